// Diagonal-format (DIA) products of libcfdla for stencil matrices.
//
// One lane per row, rows dealt in 256-row tiles: on every stored diagonal a
// wavefront reads 64 consecutive values (512 B FP64 / 256 B FP32), and no
// column index is stored, so a symmetric 7-point operator moves 4 values per
// row (main diagonal + 3 upper ones) where CSR moves 7 values and 7 indices.
// The lower entries of a symmetric matrix are the upper ones of the rows
// above (a(i, i - d) = a(i - d, i)): the same diagonal array read d rows
// back, which the L2 of the XCD still holds (the tiles of an XCD are
// contiguous, and d rows are at most one grid plane).
//
// Entries are summed in ascending column order, as a sequential CSR row sum.
// Structurally absent entries (zero on the diagonal arrays) are skipped, so
// a non-finite x at an absent column does not leak into the row.
//
// Tiles are split into 8 contiguous ranges, one per XCD (block b runs on XCD
// b % 8), so the x gathers of neighbouring rows hit the L2 of the XCD that
// owns them.
#include "epilogue.hpp"
#include "internal.hpp"
#include "reduce.hpp"

namespace cfdla {
namespace {

// EPI_DIR on a stopped solve: the product is skipped, but the new-direction
// buffer still receives the current direction, so the caller's buffer swap
// keeps pointing at the latest one when the iteration resumes (after a
// residual replacement).
template <typename T>
__device__ __forceinline__ void keep_direction(int n, const T *__restrict__ p, T *__restrict__ pn) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) pn[i] = p[i];
}

template <typename T, int EPI, bool DOT, bool SYM, int ND>
__global__ __launch_bounds__(kBlock) void k_dia(int n, DiaDesc dd, const T *__restrict__ dv,
                                                const T *__restrict__ x, const T *__restrict__ b,
                                                const T *__restrict__ dinv, T omega, T *y, const T *dotw, Scal *s,
                                                int check_done, double *partials, T *y2) {
  if (check_done && s->done) {
    if constexpr (EPI == EPI_DIR) keep_direction(n, x, y2);
    return;
  }
  if constexpr (EPI == EPI_DIR) omega = (T)s->beta;
  // entries per row in column order: SYM lower k = ND-1..1, main, upper k = 1..ND-1
  constexpr int NS = SYM ? 2 * ND - 1 : ND;
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3, per = gridDim.x >> 3;
  const int tiles = (n + kBlock - 1) / kBlock;
  const int chunk = (tiles + 7) >> 3;
  const int t1 = min(tiles, xcd * chunk + chunk);
  double dacc = 0.0;
  for (int t = xcd * chunk + local; t < t1; t += per) {
    const int row = t * kBlock + (int)threadIdx.x;
    if (row >= n) continue;
    const EpiIn<T> e = epi_load<T, EPI, DOT>(row, x, b, dinv, omega, y, dotw);
    // Branch-free: every value and gather load is issued before the sums;
    // an out-of-range column re-reads row's own slot and is masked.
    int c[NS];
    bool ok[NS];
    T v[NS], xv[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const bool lower = SYM && j < ND - 1;
      const int k = lower ? ND - 1 - j : (SYM ? j - (ND - 1) : j);
      const int col = lower ? row - dd.off[k] : row + dd.off[k];
      ok[j] = col >= 0 && col < n;
      c[j] = ok[j] ? col : row;
      v[j] = dv[k * n + (lower && ok[j] ? col : row)];
    }
#pragma unroll
    for (int j = 0; j < NS; ++j) xv[j] = gather<T, EPI>(c[j], x, b, dinv, omega);
    T sum = T(0);
#pragma unroll
    for (int j = 0; j < NS; ++j) sum += (ok[j] && v[j] != T(0)) ? v[j] * xv[j] : T(0);
    epi_store<T, EPI, DOT>(e, row, sum, omega, y, dotw, dacc, y2);
  }
  if constexpr (DOT) {
    double d1[1] = {dacc};
    block_partials<1>(d1, partials);
  }
}

// ---- R rows per lane, 16-byte loads ------------------------------------------
// The same product with R consecutive rows per lane (R = 2 in FP64), so
// every diagonal, x, epilogue and y access is one 16-byte load or store per
// lane.  It applies to symmetric storage when n is a multiple of R, the first
// off-diagonal offset is 1 and the others are multiples of R (the 2D 5-point
// and 3D 7-point operators with nx % R == 0): a multiple-of-R offset keeps a
// lane's R columns one aligned vector, and the unit offset takes its R - 1
// inner neighbours from the lane's own main-diagonal vectors plus one scalar
// load at each end.
// Each row sums the same terms in the same column order as k_dia.
template <typename T, int R>
using vec_t = T __attribute__((ext_vector_type(R)));

template <typename T, int R>
__device__ __forceinline__ void ldv(const T *p, T (&o)[R]) {
  const vec_t<T, R> v = *reinterpret_cast<const vec_t<T, R> *>(p);
#pragma unroll
  for (int r = 0; r < R; ++r) o[r] = v[r];
}

// gathered operands at the R aligned columns c..c+R-1
template <typename T, int EPI, int R>
__device__ __forceinline__ void gatherv(int c, const T *__restrict__ x, const T *__restrict__ b,
                                        const T *__restrict__ dinv, T omega, T (&o)[R]) {
  if constexpr (EPI == EPI_JACOBI0) {
    T dd[R], bb[R];
    ldv<T, R>(dinv + c, dd);
    ldv<T, R>(b + c, bb);
#pragma unroll
    for (int r = 0; r < R; ++r) o[r] = omega * dd[r] * bb[r];
  } else if constexpr (EPI == EPI_DIR) {
    T pp[R], zz[R];
    ldv<T, R>(x + c, pp);
    ldv<T, R>(b + c, zz);
#pragma unroll
    for (int r = 0; r < R; ++r) o[r] = cg_dir(omega, pp[r], zz[r]);
  } else {
    ldv<T, R>(x + c, o);
  }
}

template <typename T, int EPI, bool DOT, int ND, int R>
__global__ __launch_bounds__(kBlock) void k_dia_vec(int n, DiaDesc dd, const T *__restrict__ dv,
                                                    const T *__restrict__ x, const T *__restrict__ b,
                                                    const T *__restrict__ dinv, T omega, T *y, const T *dotw,
                                                    Scal *s, int check_done, double *partials, T *y2) {
  if (check_done && s->done) {
    if constexpr (EPI == EPI_DIR) keep_direction(n, x, y2);
    return;
  }
  if constexpr (EPI == EPI_DIR) omega = (T)s->beta;
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3, per = gridDim.x >> 3;
  const int groups = n / R;
  const int tiles = (groups + kBlock - 1) / kBlock;
  const int chunk = (tiles + 7) >> 3;
  const int t1 = min(tiles, xcd * chunk + chunk);
  double dacc = 0.0;
  for (int t = xcd * chunk + local; t < t1; t += per) {
    const int g = t * kBlock + (int)threadIdx.x;
    if (g >= groups) continue;
    const int row = g * R;
    // epilogue operands, then every entry and gather load before the sums
    EpiIn<T> e[R];
    {
      T tb[R], td[R], tx[R], ty[R], tw[R];
      if constexpr (EPI == EPI_RESID || EPI == EPI_JACOBI || EPI == EPI_JACOBI0) ldv<T, R>(b + row, tb);
      if constexpr (EPI == EPI_JACOBI || EPI == EPI_JACOBI0) ldv<T, R>(dinv + row, td);
      if constexpr (EPI == EPI_JACOBI) ldv<T, R>(x + row, tx);
      if constexpr (EPI == EPI_ADD) ldv<T, R>(y + row, ty);
      if constexpr (DOT && EPI != EPI_DIR) {
        if (dotw != y) ldv<T, R>(dotw + row, tw);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        e[r] = EpiIn<T>{T(0), T(0), T(0), T(0), T(0)};
        if constexpr (EPI == EPI_RESID || EPI == EPI_JACOBI || EPI == EPI_JACOBI0) e[r].b = tb[r];
        if constexpr (EPI == EPI_JACOBI) {
          e[r].d = td[r];
          e[r].x = tx[r];
        }
        if constexpr (EPI == EPI_JACOBI0) {
          e[r].d = td[r];
          e[r].x = omega * td[r] * tb[r];
        }
        if constexpr (EPI == EPI_ADD) e[r].y = ty[r];
        if constexpr (DOT && EPI != EPI_DIR) {
          if (dotw != y) e[r].w = tw[r];
        }
      }
    }
    // main diagonal and the operand at the lane's own columns
    T vm[R], xm[R];
    ldv<T, R>(dv + row, vm);
    gatherv<T, EPI, R>(row, x, b, dinv, omega, xm);
    if constexpr (EPI == EPI_DIR) {
#pragma unroll
      for (int r = 0; r < R; ++r) e[r].x = xm[r];  // the direction at the lane's own rows
    }
    // off-diagonals: lower (column row + r - off) and upper (row + r + off)
    T vl[ND][R], xl[ND][R], vu[ND][R], xu[ND][R];
    bool okl[ND][R], oku[ND][R];
#pragma unroll
    for (int k = 1; k < ND; ++k) {
      const int off = dd.off[k];
      const T *dk = dv + (size_t)k * n;
      ldv<T, R>(dk + row, vu[k]);
      if (k == 1) {  // off == 1
        // inner neighbours come from the lane's own vectors
        const bool okf = row > 0, okb = row + R < n;
        const T vfirst = dk[okf ? row - 1 : row];
        const T xfirst = gather<T, EPI>(okf ? row - 1 : row, x, b, dinv, omega);
        const T xlast = gather<T, EPI>(okb ? row + R : row, x, b, dinv, omega);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          vl[k][r] = r == 0 ? vfirst : vu[k][r - 1];
          xl[k][r] = r == 0 ? xfirst : xm[r - 1];
          okl[k][r] = r > 0 || okf;
          xu[k][r] = r == R - 1 ? xlast : xm[r + 1];
          oku[k][r] = r < R - 1 || okb;
        }
      } else {
        // whole vectors are in or out of range (row, off and n are multiples of R)
        const bool okL = row - off >= 0, okU = row + off < n;
        ldv<T, R>(dk + (okL ? row - off : row), vl[k]);
        gatherv<T, EPI, R>(okL ? row - off : row, x, b, dinv, omega, xl[k]);
        gatherv<T, EPI, R>(okU ? row + off : row, x, b, dinv, omega, xu[k]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          okl[k][r] = okL;
          oku[k][r] = okU;
        }
      }
    }
    T out[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      // column order: lower k = ND-1..1, main, upper k = 1..ND-1
      T sum = T(0);
#pragma unroll
      for (int k = ND - 1; k >= 1; --k)
        sum += (okl[k][r] && vl[k][r] != T(0)) ? vl[k][r] * xl[k][r] : T(0);
      sum += vm[r] != T(0) ? vm[r] * xm[r] : T(0);
#pragma unroll
      for (int k = 1; k < ND; ++k)
        sum += (oku[k][r] && vu[k][r] != T(0)) ? vu[k][r] * xu[k][r] : T(0);
      out[r] = epi_out<T, EPI>(e[r], sum, omega);
      if constexpr (DOT) dacc += (double)(EPI == EPI_DIR ? e[r].x : (dotw == y ? out[r] : e[r].w)) * (double)out[r];
    }
    vec_t<T, R> ov;
#pragma unroll
    for (int r = 0; r < R; ++r) ov[r] = out[r];
    *reinterpret_cast<vec_t<T, R> *>(y + row) = ov;
    if constexpr (EPI == EPI_DIR) {
      vec_t<T, R> pv;
#pragma unroll
      for (int r = 0; r < R; ++r) pv[r] = e[r].x;
      *reinterpret_cast<vec_t<T, R> *>(y2 + row) = pv;
    }
  }
  if constexpr (DOT) {
    double d1[1] = {dacc};
    block_partials<1>(d1, partials);
  }
}

template <typename T, int EPI, bool DOT, bool SYM, int ND>
void launch_shape(const DevCsr &A, int blocks, const T *x, const T *b, const T *dinv, double omega, T *y,
                  const T *dotw, Scal *s, bool check_done, RedWs &ws, T *y2) {
  hipLaunchKernelGGL((k_dia<T, EPI, DOT, SYM, ND>), dim3(blocks), dim3(kBlock), 0, stream(), A.n, A.dia_desc,
                     A.dia_vals<T>(), x, b, dinv, (T)omega, y, dotw, s, check_done ? 1 : 0, ws.partials.p, y2);
}

template <typename T, int EPI, bool DOT, int ND, int R>
void launch_vec(const DevCsr &A, const T *x, const T *b, const T *dinv, double omega, T *y, const T *dotw,
                Scal *s, bool check_done, RedWs &ws, int blocks, T *y2) {
  hipLaunchKernelGGL((k_dia_vec<T, EPI, DOT, ND, R>), dim3(blocks), dim3(kBlock), 0, stream(), A.n, A.dia_desc,
                     A.dia_vals<T>(), x, b, dinv, (T)omega, y, dotw, s, check_done ? 1 : 0, ws.partials.p, y2);
}

bool aligned(const void *p, size_t bytes) { return (reinterpret_cast<uintptr_t>(p) & (bytes - 1)) == 0; }

// Rows per lane of the 16-byte-load product for this operator and these
// vectors, or 1 when its conditions do not hold.
// FP64 only: in FP32 the 4-row form measured no faster in the 128^3 solve
// (Jacobi sweep 11.4 against 10.8 us, y = A x 9.7 against 9.6 us; 2 rows per
// lane slower still), since at 50 MB per pass launch and ramp-up dominate.
template <typename T>
constexpr int dia_r() { return sizeof(T) == 8 ? 2 : 1; }

template <typename T>
int dia_rows_per_lane(const DevCsr &A, const void *x, const void *b, const void *dinv, const void *y,
                      const void *dotw, const void *y2) {
  constexpr int R = dia_r<T>();
  if (R == 1) return 1;
  const DiaDesc &d = A.dia_desc;
  if (!d.sym || d.nd < 2 || d.nd > 4 || A.n % R != 0) return 1;
  if (d.off[1] != 1) return 1;
  for (int k = 2; k < d.nd; ++k)
    if (d.off[k] % R != 0) return 1;
  for (const void *p : {x, b, dinv, y, dotw, y2, static_cast<const void *>(A.dia_vals<T>())})
    if (p != nullptr && !aligned(p, R * sizeof(T))) return 1;
  return R;
}

}  // namespace

// Instantiated shapes: symmetric with 2..5 stored diagonals (1D 3-point,
// 2D 5-point, 3D 7-point, 2D 9-point), general with 3, 5, 7 or 9.
bool dia_supported(const DiaDesc &d) {
  if (d.sym) return d.nd >= 2 && d.nd <= 5 && d.off[0] == 0;
  return d.nd == 3 || d.nd == 5 || d.nd == 7 || d.nd == 9;
}

template <typename T, int EPI, bool DOT>
void launch_dia(const DevCsr &A, const T *x, const T *b, const T *dinv, double omega, T *y, const T *dotw,
                Scal *s, bool check_done, int op, RedWs &ws, T *y2) {
  // one 256-row tile per block up to the partials capacity: blocks that
  // finish hand their CU to the next tile without a loop-carried wait
  // (26.2 us against 29.1 us with 4 tiles per block, FP64 y = A x, 128^3)
  const DiaDesc &d = A.dia_desc;
  constexpr int R = dia_r<T>();
  if constexpr (R > 1) {
    if (dia_rows_per_lane<T>(A, x, b, dinv, y, dotw, y2) == R) {
      const int vblocks = xcd_grid((A.n / R + kBlock - 1) / kBlock, 1, kMaxBlocks);
      switch (d.nd) {
        case 2: launch_vec<T, EPI, DOT, 2, R>(A, x, b, dinv, omega, y, dotw, s, check_done, ws, vblocks, y2); break;
        case 3: launch_vec<T, EPI, DOT, 3, R>(A, x, b, dinv, omega, y, dotw, s, check_done, ws, vblocks, y2); break;
        default: launch_vec<T, EPI, DOT, 4, R>(A, x, b, dinv, omega, y, dotw, s, check_done, ws, vblocks, y2); break;
      }
      if (DOT) finish(1, s, op, vblocks, ws, check_done);
      return;
    }
  }
  const int blocks = xcd_grid((A.n + kBlock - 1) / kBlock, 1, kMaxBlocks);
#define CFDLA_DIA(SYM, ND) launch_shape<T, EPI, DOT, SYM, ND>(A, blocks, x, b, dinv, omega, y, dotw, s, check_done, ws, y2)
  if (d.sym) {
    switch (d.nd) {
      case 2: CFDLA_DIA(true, 2); break;
      case 3: CFDLA_DIA(true, 3); break;
      case 4: CFDLA_DIA(true, 4); break;
      case 5: CFDLA_DIA(true, 5); break;
      default: throw Error(CFDLA_ERR_INVALID_ARG, "unsupported symmetric diagonal count");
    }
  } else {
    switch (d.nd) {
      case 3: CFDLA_DIA(false, 3); break;
      case 5: CFDLA_DIA(false, 5); break;
      case 7: CFDLA_DIA(false, 7); break;
      case 9: CFDLA_DIA(false, 9); break;
      default: throw Error(CFDLA_ERR_INVALID_ARG, "unsupported diagonal count");
    }
  }
#undef CFDLA_DIA
  if (DOT) finish(1, s, op, blocks, ws, check_done);
}

#define CFDLA_DIA_INST(T, E)                                                                                   \
  template void launch_dia<T, E, false>(const DevCsr &, const T *, const T *, const T *, double, T *, const T *, \
                                        Scal *, bool, int, RedWs &, T *);                                      \
  template void launch_dia<T, E, true>(const DevCsr &, const T *, const T *, const T *, double, T *, const T *,  \
                                       Scal *, bool, int, RedWs &, T *);
#define CFDLA_DIA_ALL(T)         \
  CFDLA_DIA_INST(T, EPI_Y)       \
  CFDLA_DIA_INST(T, EPI_RESID)   \
  CFDLA_DIA_INST(T, EPI_JACOBI)  \
  CFDLA_DIA_INST(T, EPI_ADD)     \
  CFDLA_DIA_INST(T, EPI_JACOBI0) \
  template void launch_dia<T, EPI_DIR, true>(const DevCsr &, const T *, const T *, const T *, double, T *, const T *, \
                                             Scal *, bool, int, RedWs &, T *);
CFDLA_DIA_ALL(double)
CFDLA_DIA_ALL(float)

}  // namespace cfdla
