// Device value refresh of libcfdla: new matrix values on a fixed pattern
// without a host round trip.
//
// * Format scatter: the CSR FP64 values are the truth; the FP32 copy, the
//   diagonal / sliced-ELL arrays and the smoother's inverse diagonal are
//   rewritten from them through per-slot maps built once on the host.
// * The SIMPLE pressure-correction rows and right-hand side
//   (crates/cfd-2d/src/solvers/simple/pressure.rs:244-450) and the velocity /
//   pressure corrections (:487-520).
// * The AMG Galerkin refresh (amg.rs:179-214): numeric sparse products
//   replayed from per-slot plans and the dense coarsest inverse.
//
// Everything that replays a host computation does the host's floating-point
// operations in the host's order with contraction off, so results are equal
// bit for bit: each slot of a product C = A B sums its products in the
// order the host's row accumulator meets them, starting from -0.0 (the
// additive identity of IEEE round-to-nearest: -0.0 + x == x for every x,
// signed zeros included), which equals the host's "first product assigned,
// later ones added".
#include <climits>

#include "internal.hpp"
#include "fused_jacobi.hpp"
#include "simple_stencil.hpp"
#include "solver.hpp"

namespace cfdla {
namespace {

DevBuf<int32_t> g_flag;  // one device int: symmetry / singularity flags
DevBuf<double> g_max;    // max |mass imbalance| of the latest SIMPLE refresh

__global__ void k_zero_flag(int *f) { *f = 0; }

int *flag_reset() {
  g_flag.ensure(1);
  hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, stream(), g_flag.p);
  return g_flag.p;
}

int flag_read() {
  int32_t h = 0;
  CFDLA_HIP(hipMemcpyAsync(&h, g_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, stream()));
  sync_stream();
  return h;
}

__global__ __launch_bounds__(kBlock) void k_gather(int64_t nnz, const double *__restrict__ src,
                                                   const int32_t *__restrict__ perm, double *__restrict__ v64) {
  for (int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * kBlock)
    v64[k] = src[perm ? perm[k] : k];
}

// Symmetric diagonal storage: every off-diagonal entry must equal its
// transposed entry, or be zero where the pattern lacks the transpose.
__global__ __launch_bounds__(kBlock) void k_mirror_check(int64_t nnz, const double *__restrict__ v64,
                                                         const int32_t *__restrict__ mirror, int *bad) {
  for (int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * kBlock) {
    const int32_t m = mirror[k];
    if (m == -1) continue;
    const double other = m >= 0 ? v64[m] : 0.0;
    if (!(v64[k] == other)) *bad = 1;  // benign race: every writer stores 1
  }
}

__global__ __launch_bounds__(kBlock) void k_scatter(int64_t nnz, const double *__restrict__ v64, float *v32,
                                                    const int32_t *__restrict__ pos, double *fv64, float *fv32) {
  for (int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * kBlock) {
    const double v = v64[k];
    if (v32) v32[k] = (float)v;
    if (fv64) {
      const int32_t q = pos[k];
      if (q >= 0) {
        fv64[q] = v;
        if (fv32) fv32[q] = (float)v;
      }
    }
  }
}

// Smoother inverse diagonal: 1 / a_ii, 0 where |a_ii| <= 1e-15 or absent.
__global__ __launch_bounds__(kBlock) void k_sdinv(int n, const double *__restrict__ v64,
                                                  const int32_t *__restrict__ dslot, double *d64, float *d32) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const int32_t k = dslot[i];
    const double a = k >= 0 ? v64[k] : 0.0;
    const double d = fabs(a) > 1e-15 ? 1.0 / a : 0.0;
    d64[i] = d;
    if (d32) d32[i] = (float)d;
  }
}

int grid_for(int64_t units) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kGridCap, (units + kBlock - 1) / kBlock));
}

// ---- SIMPLE pressure correction ------------------------------------------

// One lane per node: interior rows (5 slots from rp[idx], columns sorted)
// and the right-hand side; boundary nodes get rhs 0.  Block maxima of
// |imbalance| (a NaN never wins the strict comparison) go to partials.
__global__ __launch_bounds__(kBlock) void k_simple_refresh(int64_t nx, int64_t ny, double dx, double dy, int negate,
                                                           const double *u, const double *v, const double *p,
                                                           const double *du, const double *dv, const double *rho,
                                                           const int32_t *__restrict__ rp, double *v64, double *rhs,
                                                           double *partials) {
  const int64_t n = nx * ny;
  double mx = 0.0;
  for (int64_t idx = blockIdx.x * (int64_t)kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock) {
    const int64_t j = idx / nx, i = idx - j * nx;
    if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1) {
      rhs[idx] = 0.0;
      continue;
    }
    const SimpleNode s = simple_node(i, j, nx, ny, dx, dy, u, v, p, du, dv, rho);
    double *row = v64 + rp[idx];
    row[0] = negate ? -s.as : s.as;
    row[1] = negate ? -s.aw : s.aw;
    row[2] = negate ? -s.ap : s.ap;
    row[3] = negate ? -s.ae : s.ae;
    row[4] = negate ? -s.an : s.an;
    rhs[idx] = negate ? -s.imbalance : s.imbalance;
    const double a = fabs(s.imbalance);
    if (a > mx) mx = a;
  }
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int t = 0; t < kBlock; ++t)
      if (sh[t] > m) m = sh[t];
    partials[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(kBlock) void k_max_final(int nparts, const double *partials, double *out) {
  __shared__ double sh[kBlock];
  double m = 0.0;
  for (int b = threadIdx.x; b < nparts; b += kBlock)
    if (partials[b] > m) m = partials[b];
  sh[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 0; t < kBlock; ++t)
      if (sh[t] > m) m = sh[t];
    *out = m;
  }
}

__global__ __launch_bounds__(kBlock) void k_simple_corrections(int64_t nx, int64_t ny, double dx, double dy,
                                                               double relax, const double *__restrict__ pp,
                                                               const double *__restrict__ du,
                                                               const double *__restrict__ dv, double *p, double *u,
                                                               double *v) {
#pragma clang fp contract(off)
  const int64_t n = nx * ny;
  const double two = 2.0;
  for (int64_t idx = blockIdx.x * (int64_t)kBlock + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * kBlock) {
    const int64_t j = idx / nx, i = idx - j * nx;
    if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1) continue;
    p[idx] = p[idx] + relax * pp[idx];
    const double dpx = (pp[idx + 1] - pp[idx - 1]) / (two * dx);
    const double dpy = (pp[idx + nx] - pp[idx - nx]) / (two * dy);
    u[idx] = u[idx] - du[idx] * dpx * dx;
    v[idx] = v[idx] - dv[idx] * dpy * dy;
  }
}

// Pressure-correction right-hand side of the cell-centred system
// (crates/cfd-2d/src/pressure_velocity/correction.rs:345-356): cell (i, j)
// of the interior is row (i-1)(ny-2) + (j-1), one less with the first cell
// pinned; rhs = -(rho/dt) (central du/dx + dv/dy), the host generator's
// operations in its order.
__global__ __launch_bounds__(kBlock) void k_pressure_rhs(int64_t nx, int64_t ny, double dx, double dy, double coeff,
                                                         const double *__restrict__ u, const double *__restrict__ v,
                                                         int pin, double *rhs) {
#pragma clang fp contract(off)
  const int64_t cells = (nx - 2) * (ny - 2);
  for (int64_t idx = blockIdx.x * (int64_t)kBlock + threadIdx.x; idx < cells; idx += (int64_t)gridDim.x * kBlock) {
    if (pin && idx == 0) continue;
    const int64_t i = idx / (ny - 2) + 1, j = idx % (ny - 2) + 1;
    const double div = (u[j * nx + i + 1] - u[j * nx + i - 1]) / (2.0 * dx) +
                       (v[(j + 1) * nx + i] - v[(j - 1) * nx + i]) / (2.0 * dy);
    rhs[pin ? idx - 1 : idx] = -coeff * div;
  }
}

// ---- Galerkin refresh ------------------------------------------------------

// Each block owns kBlock consecutive slots of C, whose products are one
// contiguous range of the plan.  The range is staged through LDS in tiles:
// the block forms the rounded products with coalesced plan loads, then every
// lane adds the products of its own slot, in plan order (the order of the
// host's row accumulator), to a sum started at -0.0.
constexpr int kProdTile = 4096;

__global__ __launch_bounds__(kBlock) void k_product_numeric(int nslots, const int32_t *__restrict__ ptr,
                                                            const int32_t *__restrict__ pa,
                                                            const int32_t *__restrict__ pb,
                                                            const double *__restrict__ av,
                                                            const double *__restrict__ bv, double *cv) {
#pragma clang fp contract(off)
  __shared__ double prod[kProdTile];
  for (int s0 = blockIdx.x * kBlock; s0 < nslots; s0 += gridDim.x * kBlock) {
    const int s = s0 + (int)threadIdx.x;
    const int q_begin = ptr[s0], q_end = ptr[min(s0 + kBlock, nslots)];
    const int my_b = s < nslots ? ptr[s] : q_end, my_e = s < nslots ? ptr[s + 1] : q_end;
    double sum = -0.0;
    for (int t0 = q_begin; t0 < q_end; t0 += kProdTile) {
      const int t1 = min(q_end, t0 + kProdTile);
      for (int q = t0 + (int)threadIdx.x; q < t1; q += kBlock) prod[q - t0] = av[pa[q]] * bv[pb[q]];
      __syncthreads();
      const int lo = max(my_b, t0), hi = min(my_e, t1);
      for (int q = lo; q < hi; ++q) sum = sum + prod[q - t0];
      __syncthreads();
    }
    if (s < nslots) cv[s] = sum;
  }
}

// Dense [A | I] (n x 2n, row-major): identity on the right, then the CSR
// entries added to the zeroed left half (0.0 + a_ij, the host's
// accumulation).
__global__ __launch_bounds__(kBlock) void k_dense_fill(int n, double *m) {
  const int64_t w = 2 * (int64_t)n;
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < (int64_t)n * w; e += (int64_t)gridDim.x * kBlock) {
    const int64_t r = e / w, c = e - r * w;
    m[e] = c == n + r ? 1.0 : 0.0;
  }
}

// One block per row; the columns of a CSR row are distinct.
__global__ __launch_bounds__(kBlock) void k_dense_scatter(int n, const int32_t *__restrict__ rp,
                                                          const int32_t *__restrict__ ci,
                                                          const double *__restrict__ v, double *m) {
  for (int r = blockIdx.x; r < n; r += gridDim.x)
    for (int k = rp[r] + (int)threadIdx.x; k < rp[r + 1]; k += kBlock)
      m[(int64_t)r * 2 * n + ci[k]] = m[(int64_t)r * 2 * n + ci[k]] + v[k];
}

// Partial pivoting as the host does it: scanning rows c, c+1, ... and
// replacing the best only on a strictly larger |m_rc| picks the lowest row of
// the largest value, and a NaN never wins -- except at row c, the initial
// best, where it makes the column singular.  On the device, row c's NaN is
// tested directly and the other candidates are reduced with "larger value,
// then lower row" (NaN entered as -1), an associative and commutative merge,
// so any reduction order finds the host's row.
struct Piv {
  double val;
  int row;
};

__device__ inline Piv piv_better(Piv a, Piv b) {
  return (b.val > a.val || (b.val == a.val && b.row < a.row)) ? b : a;
}

__device__ inline Piv piv_of(double x, int row) { return Piv{x == x ? fabs(x) : -1.0, row}; }

constexpr int kGjRows = 4;  // rows per block of the elimination step

// Block-wide merge of one candidate per thread (shuffles inside each wave,
// then the 4 wave results); the result is returned to every thread.
__device__ Piv piv_block(Piv mine) {
  __shared__ double sv[kBlock / 64];
  __shared__ int sr[kBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    mine = piv_better(mine, Piv{__shfl_xor(mine.val, o, 64), __shfl_xor(mine.row, o, 64)});
  if ((threadIdx.x & 63) == 0) {
    sv[threadIdx.x >> 6] = mine.val;
    sr[threadIdx.x >> 6] = mine.row;
  }
  __syncthreads();
  Piv best{sv[0], sr[0]};
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) best = piv_better(best, Piv{sv[w], sr[w]});
  return best;
}

// Column 0's partials of the filled matrix.
__global__ __launch_bounds__(kBlock) void k_gj_init(int n, const double *m, double *pval, int *prow) {
  const int r0 = blockIdx.x * kGjRows;
  if (threadIdx.x == 0) {
    Piv acc{-1.0, INT_MAX};
    for (int r = r0; r < min(n, r0 + kGjRows); ++r) acc = piv_better(acc, piv_of(m[(int64_t)r * 2 * n], r));
    pval[blockIdx.x] = acc.val;
    prow[blockIdx.x] = acc.row;
  }
}

// One Gauss-Jordan column step on [A | I], ping-pong src -> dst: pivot row p
// from the partials, rows c and p swapped, row c scaled by d = 1 / m_pc,
// every other row r with f = m_rc != 0 updated m_rj - f (m_pj d) for the
// columns j > c (the left-half columns up to c are dead).  Writes the next
// column's pivot partials.  A singular column (NaN at row c, best
// pivot 0 or not finite) sets *bad; every later step then returns at once.
__global__ __launch_bounds__(kBlock) void k_gj_step(int n, int c, const double *src, double *dst,
                                                    const double *pin, const int *rin, double *pout, int *rout,
                                                    int nparts, int *bad) {
#pragma clang fp contract(off)
  __shared__ int s_p;
  __shared__ double s_next[kGjRows];
  const int64_t w = 2 * (int64_t)n;
  Piv mine{-1.0, INT_MAX};
  for (int b = threadIdx.x; b < nparts; b += kBlock) mine = piv_better(mine, Piv{pin[b], rin[b]});
  const Piv best = piv_block(mine);
  if (threadIdx.x == 0) {
    const double mcc = src[(int64_t)c * w + c];
    s_p = (*bad || mcc != mcc || !(best.val > 0.0 && best.val <= 1.7976931348623157e308)) ? -1 : best.row;
  }
  __syncthreads();
  // Columns 0..c of the left half are never read after this step (pivots
  // and multipliers come from later columns), so they are not written.
  const int64_t j0 = c + 1;
  const int p = s_p;
  if (p < 0) {
    if (threadIdx.x == 0) *bad = 1;
    return;
  }
  const double d = 1.0 / src[(int64_t)p * w + c];
  const int r0 = blockIdx.x * kGjRows;
  for (int q = 0; q < kGjRows; ++q) {
    const int r = r0 + q;
    if (r >= n) break;
    const int sr = r == c ? p : (r == p ? c : r);
    const double *srow = src + (int64_t)sr * w;
    const double *prow = src + (int64_t)p * w;
    double *drow = dst + (int64_t)r * w;
    const double f = srow[c];
    for (int64_t j = j0 + threadIdx.x; j < w; j += kBlock) {
      double out;
      if (r == c) out = prow[j] * d;
      else if (f == 0.0) out = srow[j];
      else {
        const double piv_j = prow[j] * d;
        const double t = f * piv_j;
        out = srow[j] - t;
      }
      drow[j] = out;
      if (j == c + 1) s_next[q] = out;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && c + 1 < n) {
    Piv acc{-1.0, INT_MAX};
    for (int q = 0; q < kGjRows; ++q) {
      const int r = r0 + q;
      if (r >= n) break;
      if (r >= c + 1) acc = piv_better(acc, piv_of(s_next[q], r));
    }
    pout[blockIdx.x] = acc.val;
    rout[blockIdx.x] = acc.row;
  }
}

// Gauss-Jordan steps c0 .. c0 + B - 1 on [A | I] (column-major, entry
// (r, j) at j n + r) in one launch, for n <= kGjMaxRows.  Every block loads
// the B panel columns (contiguous) into LDS and
// replays the B steps on them -- pivot row, swap, scale by d = 1 / m_pc,
// eliminate with f = m_rc -- with the operations of k_gj_step (and of the
// host), applying each step to its own kGjCols columns held in registers
// (rows threadIdx.x + 256 q).  The columns left of the panel are dead, and
// the panel's own columns are dead once its steps are done, so only columns
// >= c0 + B are carried, src -> dst.  Each element meets the host's
// operations in the host's order, so the inverse is the host's bit for bit.
constexpr int kGjPanel = 8;      // steps per launch (fewer when n * B * 8 B exceeds 60 KiB)
constexpr int kGjCols = 4;       // columns per block
constexpr int kGjMaxRows = 1024; // rows held by one block (4 per thread)
constexpr int kGjPanelLds = 7680;  // doubles of the LDS panel
__global__ __launch_bounds__(kBlock) void k_gj_panel(int n, int c0, int B, const double *src, double *dst,
                                                     int *bad) {
#pragma clang fp contract(off)
  extern __shared__ double pnl[];  // n x B, column-major: entry (r, k) at k n + r
  __shared__ double u[kGjCols];    // row c of this block's columns, scaled
  __shared__ double sw[2][kGjCols];
  __shared__ int s_p;
  constexpr int RQ = kGjMaxRows / kBlock;
  if (*bad) return;  // an earlier step found the matrix singular
  const int tid = threadIdx.x;
  const int64_t w = 2 * (int64_t)n;
  const int64_t j0 = c0 + B + (int64_t)blockIdx.x * kGjCols;
  for (int e = tid; e < n * B; e += kBlock) pnl[e] = src[(int64_t)c0 * n + e];
  double x[RQ][kGjCols];
#pragma unroll
  for (int q = 0; q < RQ; ++q)
#pragma unroll
    for (int jj = 0; jj < kGjCols; ++jj) {
      const int r = tid + q * kBlock;
      x[q][jj] = r < n && j0 + jj < w ? src[(j0 + jj) * n + r] : 0.0;
    }
  __syncthreads();
  for (int k = 0; k < B; ++k) {
    const int c = c0 + k;
    Piv mine{-1.0, INT_MAX};
    for (int r = c + tid; r < n; r += kBlock) mine = piv_better(mine, piv_of(pnl[k * n + r], r));
    const Piv best = piv_block(mine);
    if (tid == 0) {
      const double mcc = pnl[k * n + c];
      s_p = (mcc != mcc || !(best.val > 0.0 && best.val <= 1.7976931348623157e308)) ? -1 : best.row;
    }
    __syncthreads();
    const int p = s_p;
    if (p < 0) {
      if (tid == 0) *bad = 1;
      return;
    }
    const double d = 1.0 / pnl[k * n + p];
    if (p != c) {
      // rows c and p trade places: the panel, then this block's columns
      if (tid < B) {
        const double a = pnl[tid * n + c], b = pnl[tid * n + p];
        pnl[tid * n + c] = b;
        pnl[tid * n + p] = a;
      }
#pragma unroll
      for (int q = 0; q < RQ; ++q)
#pragma unroll
        for (int jj = 0; jj < kGjCols; ++jj) {
          const int r = tid + q * kBlock;
          if (r == c) sw[0][jj] = x[q][jj];
          if (r == p) sw[1][jj] = x[q][jj];
        }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < RQ; ++q)
#pragma unroll
        for (int jj = 0; jj < kGjCols; ++jj) {
          const int r = tid + q * kBlock;
          if (r == c) x[q][jj] = sw[1][jj];
          if (r == p) x[q][jj] = sw[0][jj];
        }
    }
    __syncthreads();
    // row c scaled by d
    if (tid >= k && tid < B) pnl[tid * n + c] = pnl[tid * n + c] * d;
#pragma unroll
    for (int q = 0; q < RQ; ++q)
      if (tid + q * kBlock == c)
#pragma unroll
        for (int jj = 0; jj < kGjCols; ++jj) {
          x[q][jj] = x[q][jj] * d;
          u[jj] = x[q][jj];
        }
    __syncthreads();
    // every other row r with f = m_rc != 0: m_rj - f m_cj
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      const int r = tid + q * kBlock;
      if (r >= n || r == c) continue;
      const double f = pnl[k * n + r];
      if (f == 0.0) continue;
      for (int kk = k; kk < B; ++kk) {
        const double t = f * pnl[kk * n + c];
        pnl[kk * n + r] = pnl[kk * n + r] - t;
      }
#pragma unroll
      for (int jj = 0; jj < kGjCols; ++jj) {
        const double t = f * u[jj];
        x[q][jj] = x[q][jj] - t;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < RQ; ++q)
#pragma unroll
    for (int jj = 0; jj < kGjCols; ++jj) {
      const int r = tid + q * kBlock;
      if (r < n && j0 + jj < w) dst[(j0 + jj) * n + r] = x[q][jj];
    }
}

// [A | I] column-major (entry (r, j) at j n + r), for the panel path: the
// panel and each block's columns are then contiguous.
__global__ __launch_bounds__(kBlock) void k_dense_fill_cm(int n, double *m) {
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < 2 * (int64_t)n * n;
       e += (int64_t)gridDim.x * kBlock) {
    const int64_t j = e / n, r = e - j * n;
    m[e] = j == n + r ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(kBlock) void k_dense_scatter_cm(int n, const int32_t *__restrict__ rp,
                                                             const int32_t *__restrict__ ci,
                                                             const double *__restrict__ v, double *m) {
  for (int r = blockIdx.x; r < n; r += gridDim.x)
    for (int k = rp[r] + (int)threadIdx.x; k < rp[r + 1]; k += kBlock)
      m[(int64_t)ci[k] * n + r] = m[(int64_t)ci[k] * n + r] + v[k];
}

__global__ __launch_bounds__(kBlock) void k_gj_extract_cm(int n, const double *m, double *inv) {
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < (int64_t)n * n; e += (int64_t)gridDim.x * kBlock) {
    const int64_t r = e / n, c = e - r * n;
    inv[e] = m[(n + c) * (int64_t)n + r];
  }
}

__global__ __launch_bounds__(kBlock) void k_gj_extract(int n, const double *m, double *inv) {
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < (int64_t)n * n; e += (int64_t)gridDim.x * kBlock) {
    const int64_t r = e / n, c = e - r * n;
    inv[e] = m[r * 2 * n + n + c];
  }
}


// ---- variable-coefficient Poisson rows --------------------------------------
// One lane per cell: the row of cfdla_problem_varcoef (face coefficient
// 2 kP kN / (kP + kN) * inv_h2; diagonal summed over the faces -z, -y, -x,
// +x, +y, +z, then the Dirichlet faces in the same order; values written in
// column order) into v64 from rp[p], and the Dirichlet right-hand side.  FP
// contraction is off, as in the host generator, so the values are the
// host's bit for bit.
__global__ __launch_bounds__(kBlock) void k_varcoef_rows(int64_t nx, int64_t ny, int64_t nz, double inv_h2,
                                                         const double *__restrict__ k, int dir_mask, double dv0,
                                                         double dv1, double dv2, double dv3, double dv4,
                                                         double dv5, const int32_t *__restrict__ rp,
                                                         double *__restrict__ v64, double *__restrict__ rhs) {
#pragma clang fp contract(off)
  const int64_t n = nx * ny * nz, sy = nx, sz = nx * ny;
  const bool three_d = nz > 1;
  const double dv[6] = {dv0, dv1, dv2, dv3, dv4, dv5};
  for (int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x; p < n; p += (int64_t)gridDim.x * kBlock) {
    const int64_t x = p % nx, y = (p / nx) % ny, z = p / sz;
    const double kp = k[p];
    const bool has[6] = {three_d && z > 0, y > 0, x > 0, x + 1 < nx, y + 1 < ny, three_d && z + 1 < nz};
    const int64_t nb[6] = {p - sz, p - sy, p - 1, p + 1, p + sy, p + sz};
    double c[6];
    double diag = 0.0;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      c[f] = 0.0;
      if (has[f]) {
        const double kq = k[nb[f]];
        c[f] = (2.0 * kp * kq / (kp + kq)) * inv_h2;
        diag += c[f];
      }
    }
    const bool on[6] = {three_d && z == 0, y == 0, x == 0, x == nx - 1, y == ny - 1, three_d && z == nz - 1};
    const int bit[6] = {4, 2, 0, 1, 3, 5};
    double bc = 0.0;
#pragma unroll
    for (int f = 0; f < 6; ++f)
      if (on[f] && ((dir_mask >> bit[f]) & 1)) {
        const double cd = kp * inv_h2;
        diag += cd;
        bc += cd * dv[bit[f]];
      }
    if (rhs) rhs[p] = bc;
    int64_t o = rp[p];
#pragma unroll
    for (int f = 0; f < 3; ++f)
      if (has[f]) v64[o++] = -c[f];
    v64[o++] = diag;
#pragma unroll
    for (int f = 3; f < 6; ++f)
      if (has[f]) v64[o++] = -c[f];
  }
}
}  // namespace

void gather_values(int64_t nnz, const double *src, const int32_t *perm, double *v64) {
  if (nnz == 0) return;
  hipLaunchKernelGGL(k_gather, dim3(grid_for(nnz)), dim3(kBlock), 0, stream(), nnz, src, perm, v64);
}

bool symmetric_on_device(const DevCsr &d) {
  if (d.nnz == 0) return true;
  int *bad = flag_reset();
  hipLaunchKernelGGL(k_mirror_check, dim3(grid_for(d.nnz)), dim3(kBlock), 0, stream(), d.nnz, d.v64.p,
                     d.fmt_mirror.p, bad);
  return flag_read() == 0;
}

namespace {
// S and E values of a level's fused operator from A's values and the
// smoother's inverse diagonal, one lane per row (fused_jacobi.hpp).
__global__ __launch_bounds__(kBlock) void k_fused_values(int n, const int32_t *__restrict__ rp,
                                                         const int32_t *__restrict__ ci,
                                                         const double *__restrict__ av,
                                                         const double *__restrict__ dinv, double w, double *sv,
                                                         double *ev) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double di = dinv[i];
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = ci[k];
      const double a = av[k], dj = dinv[j];
      sv[k] = fused_s(i == j, w, a, di, dj);
      ev[k] = fused_e(i == j, w, a, dj);
    }
  }
}
}  // namespace

void fused_values(const DevCsr &A, double w, double *sv, double *ev) {
  if (A.n > 0)
    hipLaunchKernelGGL(k_fused_values, dim3(grid_for(A.n)), dim3(kBlock), 0, stream(), A.n, A.rp.p, A.ci.p, A.v64.p,
                       A.sdinv64.p, w, sv, ev);
}

void scatter_formats(DevCsr &d) {
  if (d.nnz > 0) {
    double *f64 = d.dia ? d.dia_v64.p : (d.sell ? d.sell_v64.p : nullptr);
    float *f32 = d.dia ? d.dia_v32.p : (d.sell ? d.sell_v32.p : nullptr);
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(d.nnz)), dim3(kBlock), 0, stream(), d.nnz, d.v64.p, d.v32.p,
                       d.fmt_pos.p, f64, f32);
  }
  if (d.sdinv64.p && d.n > 0)
    hipLaunchKernelGGL(k_sdinv, dim3(grid_for(d.n)), dim3(kBlock), 0, stream(), d.n, d.v64.p, d.diag_slot.p,
                       d.sdinv64.p, d.sdinv32.p);
}

namespace {
// One workgroup per row block at a time; the FP64 partition (both are cut
// alike when the sorted copy exists).
__global__ __launch_bounds__(kBlock) void k_sorted_values(int nrb, const int *__restrict__ rb,
                                                          const int32_t *__restrict__ sci,
                                                          const double *__restrict__ v64, double *sv64,
                                                          float *sv32) {
  for (int q = blockIdx.x; q < nrb; q += gridDim.x) {
    const int e0 = rb[2 * q + 1], e1 = rb[2 * q + 3];
    const bool regular = e1 - e0 <= rb_cap<double>();
    for (int j = e0 + (int)threadIdx.x; j < e1; j += kBlock) {
      const int k = regular ? e0 + (int)((uint32_t)sci[j] >> kRbPosShift) : j;
      const double a = v64[k];
      sv64[j] = a;
      if (sv32) sv32[j] = (float)a;
    }
  }
}
}  // namespace

void sorted_values(DevCsr &d) {
  if (d.nnz > 0 && d.nrb64 > 0)
    hipLaunchKernelGGL(k_sorted_values, dim3(std::min(d.nrb64, kMaxBlocks)), dim3(kBlock), 0, stream(), d.nrb64,
                       d.rb64.p, d.sci.p, d.v64.p, d.sv64.p, d.sv32.p);
}

void simple_refresh(int64_t nx, int64_t ny, double dx, double dy, bool negate, const double *u, const double *v,
                    const double *p, const double *du, const double *dv, const double *rho, const int32_t *rp,
                    double *v64, double *rhs, bool want_max) {
  const int grid = grid_for(nx * ny);
  RedWs &ws = ctx().ws;
  hipLaunchKernelGGL(k_simple_refresh, dim3(grid), dim3(kBlock), 0, stream(), nx, ny, dx, dy, negate ? 1 : 0, u, v,
                     p, du, dv, rho, rp, v64, rhs, ws.partials.p);
  if (!want_max) return;
  g_max.ensure(1);
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(kBlock), 0, stream(), grid, ws.partials.p, g_max.p);
}

void varcoef_refresh(int64_t nx, int64_t ny, int64_t nz, double inv_h2, const double *k, int dir_mask,
                     const double *dir_val, const int32_t *rp, double *v64, double *rhs) {
  const double z6[6] = {0, 0, 0, 0, 0, 0};
  const double *d = dir_val ? dir_val : z6;
  hipLaunchKernelGGL(k_varcoef_rows, dim3(grid_for(nx * ny * nz)), dim3(kBlock), 0, stream(), nx, ny, nz, inv_h2, k,
                     dir_mask, d[0], d[1], d[2], d[3], d[4], d[5], rp, v64, rhs);
  CFDLA_HIP(hipGetLastError());
}

double simple_max_read() {
  double h = 0.0;
  CFDLA_HIP(hipMemcpyAsync(&h, g_max.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
  sync_stream();
  return h;
}

void simple_corrections(int64_t nx, int64_t ny, double dx, double dy, double relax, const double *pp,
                        const double *du, const double *dv, double *p, double *u, double *v) {
  hipLaunchKernelGGL(k_simple_corrections, dim3(grid_for(nx * ny)), dim3(kBlock), 0, stream(), nx, ny, dx, dy,
                     relax, pp, du, dv, p, u, v);
}

void pressure_rhs(int64_t nx, int64_t ny, double dx, double dy, double coeff, const double *u, const double *v,
                  bool pin, double *rhs) {
  hipLaunchKernelGGL(k_pressure_rhs, dim3(grid_for((nx - 2) * (ny - 2))), dim3(kBlock), 0, stream(), nx, ny, dx, dy,
                     coeff, u, v, pin ? 1 : 0, rhs);
}

void product_numeric(int nslots, const int32_t *ptr, const int32_t *pa, const int32_t *pb, const double *av,
                     const double *bv, double *cv) {
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_product_numeric, dim3(std::max(1, std::min(kMaxBlocks, (nslots + kBlock - 1) / kBlock))),
                     dim3(kBlock), 0, stream(), nslots, ptr, pa, pb, av, bv, cv);
}

void dense_inverse_device(int n, const int32_t *rp, const int32_t *ci, const double *v, double *inv) {
  DevBuf<double> buf[2];
  buf[0].alloc((size_t)n * 2 * n);
  buf[1].alloc((size_t)n * 2 * n);
  const int nb = (n + kGjRows - 1) / kGjRows;
  DevBuf<double> pv[2];
  DevBuf<int> pr[2];
  for (int k = 0; k < 2; ++k) {
    pv[k].alloc((size_t)nb);
    pr[k].alloc((size_t)nb);
  }
  int *bad = flag_reset();
  if (n <= kGjMaxRows) {
    hipLaunchKernelGGL(k_dense_fill_cm, dim3(grid_for((int64_t)n * 2 * n)), dim3(kBlock), 0, stream(), n, buf[0].p);
    hipLaunchKernelGGL(k_dense_scatter_cm, dim3(std::min(n, kMaxBlocks)), dim3(kBlock), 0, stream(), n, rp, ci, v,
                       buf[0].p);
    // B steps per launch (641 launches become 81 for a 641-row level)
    const int B = std::min(kGjPanel, kGjPanelLds / n);
    int it = 0;
    for (int c0 = 0; c0 < n; c0 += B, ++it) {
      const int b = std::min(B, n - c0);
      const int64_t live = 2 * (int64_t)n - (c0 + b);
      const int blocks = (int)((live + kGjCols - 1) / kGjCols);
      hipLaunchKernelGGL(k_gj_panel, dim3(blocks), dim3(kBlock), (size_t)n * b * sizeof(double), stream(), n, c0,
                         b, buf[it & 1].p, buf[(it + 1) & 1].p, bad);
    }
    hipLaunchKernelGGL(k_gj_extract_cm, dim3(grid_for((int64_t)n * n)), dim3(kBlock), 0, stream(), n,
                       buf[it & 1].p, inv);
  } else {
    hipLaunchKernelGGL(k_dense_fill, dim3(grid_for((int64_t)n * 2 * n)), dim3(kBlock), 0, stream(), n, buf[0].p);
    hipLaunchKernelGGL(k_dense_scatter, dim3(std::min(n, kMaxBlocks)), dim3(kBlock), 0, stream(), n, rp, ci, v,
                       buf[0].p);
    hipLaunchKernelGGL(k_gj_init, dim3(nb), dim3(kBlock), 0, stream(), n, buf[0].p, pv[0].p, pr[0].p);
    for (int c = 0; c < n; ++c)
      hipLaunchKernelGGL(k_gj_step, dim3(nb), dim3(kBlock), 0, stream(), n, c, buf[c & 1].p, buf[(c + 1) & 1].p,
                         pv[c & 1].p, pr[c & 1].p, pv[(c + 1) & 1].p, pr[(c + 1) & 1].p, nb, bad);
    hipLaunchKernelGGL(k_gj_extract, dim3(grid_for((int64_t)n * n)), dim3(kBlock), 0, stream(), n, buf[n & 1].p,
                       inv);
  }
  CFDLA_REQUIRE(flag_read() == 0, CFDLA_ERR_NUMERICAL,
                "AMG coarsest operator is singular; use coarse_solver = reference sweeps");
}

}  // namespace cfdla
