// gfx950 kernels of libcfdla: CSR and sliced-ELL products with fused
// epilogues, deterministic grid reductions whose finalize kernel advances the
// Krylov scalars on the device, Krylov vector updates, Jacobi sweeps and the
// FP64 dense coarse solve.
//
// Wave64 throughout.  Reductions: wave shuffle tree, then LDS across the 4
// waves of a 256-thread block, one partial per block to global memory, and a
// single-block kernel that sums the partials in block order — the result does
// not depend on block scheduling, so runs are bitwise reproducible.
#include <cmath>

#include "epilogue.hpp"
#include "internal.hpp"
#include "reduce.hpp"

namespace cfdla {
namespace {

__device__ void finalize(Scal *s, int op) {
  switch (op) {
    case OP_INIT_RR:
      s->rr = s->red[0];
      break;
    case OP_SET_RZ:
      s->rz = s->red[0];
      break;
    case OP_CG_ALPHA: {
      double pq = s->red[0];
      s->pq = pq;
      if (!(pq > 0.0) || !isfinite(pq)) {
        s->term = CFDLA_TERM_NONPOSITIVE_CURVATURE;
        s->done = 2;
      } else {
        s->alpha = s->rz / pq;
      }
    } break;
    case OP_CG_RR: {
      double rr = s->red[0];
      s->rr = rr;
      s->iters += 1;
      if (!isfinite(rr)) {
        s->term = CFDLA_TERM_BREAKDOWN;
        s->done = 2;
      } else if (rr <= s->thr2) {
        s->done = 1;
      }
    } break;
    case OP_CG_BETA: {
      double rzn = s->red[0];
      if (rzn == 0.0 || !isfinite(rzn)) {
        s->term = CFDLA_TERM_BREAKDOWN;
        s->done = 2;
      } else {
        s->beta = s->restart_dir ? 0.0 : rzn / s->rz;
        s->restart_dir = 0;
        s->rz = rzn;
      }
    } break;
    case OP_CG_RR_RZ: {
      double rr = s->red[0], rzn = s->red[1];
      s->rr = rr;
      s->iters += 1;
      if (!isfinite(rr)) {
        s->term = CFDLA_TERM_BREAKDOWN;
        s->done = 2;
      } else if (rr <= s->thr2) {
        s->done = 1;
      } else if (rzn == 0.0 || !isfinite(rzn)) {
        s->term = CFDLA_TERM_BREAKDOWN;
        s->done = 2;
      } else {
        s->beta = rzn / s->rz;
        s->rz = rzn;
      }
    } break;
    case OP_BI_RHO: {
      double rho_new = s->red[0];
      if (rho_new == 0.0 || !isfinite(rho_new)) {
        s->term = CFDLA_TERM_BREAKDOWN;
        s->done = 2;
      } else {
        // a residual replacement far from the recurrence restarts p = r
        s->beta = s->restart_dir ? 0.0 : (rho_new / s->rho) * (s->alpha / s->omega);
        s->restart_dir = 0;
        s->rho = rho_new;
      }
    } break;
    case OP_BI_ALPHA: {
      double rv = s->red[0];
      if (rv == 0.0 || !isfinite(rv)) {
        s->term = CFDLA_TERM_BREAKDOWN;
        s->done = 2;
      } else {
        s->alpha = s->rho / rv;
      }
    } break;
    case OP_BI_SNORM: {
      double ss = s->red[0];
      s->iters += 1;
      if (!isfinite(ss)) {
        s->term = CFDLA_TERM_BREAKDOWN;
        s->done = 2;
      } else if (ss <= s->thr2) {
        s->done = 3;  // x += alpha * phat, then confirm on the host
      }
    } break;
    case OP_BI_OMEGA: {
      double ts = s->red[0], tt = s->red[1];
      s->omega = tt > 0.0 ? ts / tt : 0.0;
    } break;
    case OP_RELIABLE: {
      // FP64 residual of the accumulated FP64 solution (reliable-update CG)
      // A replaced residual far above the recursive one it replaces (the
      // FP32 operator's rounding caps what one cycle can gain) would make
      // beta = r.z / rz blow up and wreck the direction: restart it.
      double rr = s->red[0];
      const double rec = s->rr;
      s->rr = rr;
      s->rr_true = rr;
      s->nupd += 1;
      if (!isfinite(rr)) {
        s->term = CFDLA_TERM_BREAKDOWN;
        s->done = 2;
      } else {
        s->done = rr <= s->thr_final2 ? 4 : 0;
        s->restart_dir = rr > 4.0 * rec ? 1 : 0;
        s->nrestart += s->restart_dir;
      }
    } break;
    case OP_BI_RR:
    case OP_BI_RR_RHO: {
      double rr = s->red[0];
      s->rr = rr;
      if (!isfinite(rr)) {
        s->term = CFDLA_TERM_BREAKDOWN;
        s->done = 2;
      } else if (rr <= s->thr2) {
        s->done = 1;
      } else if (s->omega == 0.0 || !isfinite(s->omega)) {
        s->term = CFDLA_TERM_BREAKDOWN;
        s->done = 2;
      } else if (op == OP_BI_RR_RHO) {
        // the next iteration's OP_BI_RHO, with rho_new = rhat . r
        double rho_new = s->red[1];
        if (rho_new == 0.0 || !isfinite(rho_new)) {
          s->term = CFDLA_TERM_BREAKDOWN;
          s->done = 2;
        } else {
          s->beta = s->restart_dir ? 0.0 : (rho_new / s->rho) * (s->alpha / s->omega);
          s->restart_dir = 0;
          s->rho = rho_new;
        }
      }
    } break;
    default:
      break;
  }
}

// Reducing kernels only write their block partials; the finalize op runs in
// the single-block k_finalize launched right after them on the same stream.
template <int NV>
__device__ __forceinline__ void reduce_and_finalize(double (&v)[NV], Scal *, int, double *partials, unsigned *) {
  block_partials<NV>(v, partials);
}

template <int NV>
__global__ __launch_bounds__(kBlock) void k_finalize(Scal *s, int op, int nblocks, const double *partials,
                                                     int check_done) {
  if (check_done && s->done) return;
  double tot[NV];
  sum_partials<NV>(nblocks, partials, tot);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) s->red[i] = tot[i];
    finalize(s, op);
  }
}

}  // namespace

void finish(int nv, Scal *s, int op, int nblocks, RedWs &ws, bool check_done) {
  if (nv == 1)
    hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kBlock), 0, stream(), s, op, nblocks, ws.partials.p,
                       check_done ? 1 : 0);
  else
    hipLaunchKernelGGL((k_finalize<2>), dim3(1), dim3(kBlock), 0, stream(), s, op, nblocks, ws.partials.p,
                       check_done ? 1 : 0);
}

// Grid of a multiple of 8 blocks (one per XCD round-robin), at most cap.
int xcd_grid(int units, int units_per_block, int cap) {
  int blocks = (units + units_per_block - 1) / units_per_block;
  blocks = std::max(1, std::min(blocks, cap));
  return (blocks + 7) & ~7;
}

namespace {

// ---- row-block CSR product ----------------------------------------------------
// Four consecutive entries of the matrix streams (columns, values), loaded as
// single 8- or 16-byte loads, nontemporal (NT) for operators too large to
// stay in the caches between products: those streams are read once per
// product, so they need not displace the gathered vector.
// The arrays carry kRbPad zero entries past nnz, so a quad that starts
// inside the array never reads past its allocation.
typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned short v4u16 __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

template <bool NT, typename V>
__device__ __forceinline__ V load_vec(const V *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void load_quad(const double *p, double (&v)[4]) {
  const v2d a = load_vec<NT>(reinterpret_cast<const v2d *>(p));
  const v2d b = load_vec<NT>(reinterpret_cast<const v2d *>(p + 2));
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
template <bool NT>
__device__ __forceinline__ void load_quad(const float *p, float (&v)[4]) {
  const v4f a = load_vec<NT>(reinterpret_cast<const v4f *>(p));
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
template <bool NT>
__device__ __forceinline__ void load_quad_cols(const int *p, int (&c)[4]) {
  const v4i a = load_vec<NT>(reinterpret_cast<const v4i *>(p));
  c[0] = a.x; c[1] = a.y; c[2] = a.z; c[3] = a.w;
}
template <bool NT>
__device__ __forceinline__ void load_quad_cols(const uint16_t *p, int base, int (&c)[4]) {
  const v4u16 a = load_vec<NT>(reinterpret_cast<const v4u16 *>(p));
  c[0] = base + (int)a.x; c[1] = base + (int)a.y; c[2] = base + (int)a.z; c[3] = base + (int)a.w;
}
// Four entries into LDS at a 4-aligned entry offset (16- or 32-byte stores).
__device__ __forceinline__ void store_quad(int *p, const int (&c)[4]) {
  *reinterpret_cast<v4i *>(p) = v4i{c[0], c[1], c[2], c[3]};
}
__device__ __forceinline__ void store_quad(float *p, const float (&v)[4]) {
  *reinterpret_cast<v4f *>(p) = v4f{v[0], v[1], v[2], v[3]};
}
__device__ __forceinline__ void store_quad(double *p, const double (&v)[4]) {
  reinterpret_cast<v2d *>(p)[0] = v2d{v[0], v[1]};
  reinterpret_cast<v2d *>(p)[1] = v2d{v[2], v[3]};
}

// The host packs rows in order into blocks of at most rb_rows = kBlock rows
// and rb_cap nonzeros.  A workgroup streams its block's columns and values in
// quads from the 4-aligned entry at or below its first one (up to
// rb_cap / (4 kBlock) = 2 quads per thread, the at most 3 entries past them by
// three lanes), all loads in flight together.  Each row is then summed by tpr
// lanes (the largest power of two up to 64 with tpr * rows <= kBlock), in one
// of two gather layouts picked per operator at upload (rb_pick_layout):
//   * row walk (WALK): the quads are parked in LDS; after the barrier each
//     row's lanes read their entries back (8 or 4 at a time, all LDS reads
//     issued before the gathers) and gather x, so lanes of a wavefront hold
//     consecutive rows (tpr = 1) or consecutive entries of one row and the
//     gathers of a stencil-like operator fall on a few cache lines per
//     wavefront (y = A x on the 128^3 7-point operator: 42.1 us FP64,
//     24.3 us FP32);
//   * quad: x is gathered for each quad as it arrives (4 consecutive
//     entries per lane), products are parked in LDS and summed from there;
//     fewer gather instructions where rows are short and uneven (the
//     prolongations) or a few lanes share a long row (the first
//     restriction), but the 7-point operator's gathers spread over 3 grid
//     planes (47.3 and 29.5 us).
//   * strided (forced only, CFDLA_GATHER_STRIDED): staged as for the row
//     walk, entries dealt to threads round robin so a wavefront gathers for
//     64 consecutive entries, products written back over the staged values
//     and summed as in the quad layout (a second barrier); the first
//     restriction in FP64 took 31.7 against 36.2 us in the quad layout.
//   * sorted (CFDLA_GATHER_SORTED): the quad layout over blocks stored in
//     column order, every entry packed with its row-ordered position
//     (kRbPosShift), so a wavefront's gathers fall on adjacent columns; each
//     product is parked at its position and the rows are summed exactly as
//     in the quad layout, bit for bit the same result.
// A row longer than rb_cap is a block of its own, summed in steps of rb_cap
// entries straight from memory.
//
// Blocks are dealt to the 8 XCDs round-robin; XCD q walks the q-th
// contiguous range of row blocks, so neighbouring rows share its L2.
//
// Measured against a vector-CSR kernel (a power-of-two lane group per row
// near the mean row length, no LDS or barriers) on the 128^3 hierarchy: the
// row blocks win on every transfer operator (FP64 P 736 against 1051 us per
// solve, R 780 against 1008 us); the vector kernel won only FP32 level-2
// sweeps (377 against 437 us).
//
// C16: columns are 16-bit offsets from the block's base column (cb[q]).
// LAYOUT: kRbQuad, kRbRows (row walk), kRbStrided or kRbSorted; separate
// instantiations, so the quad layout keeps its smaller LDS footprint (8
// blocks per CU in FP64 instead of 6) and its own register allocation.
template <typename T, int EPI, bool DOT, bool C16, bool NT, int LAYOUT>
__global__ __launch_bounds__(kBlock) void k_csr_rb(int nrb, const int2 *__restrict__ rb, const int *__restrict__ rp,
                                                   const int *__restrict__ ci, const uint16_t *__restrict__ c16,
                                                   const int *__restrict__ cb, const T *__restrict__ av,
                                                   const T *__restrict__ x, const T *__restrict__ b,
                                                   const T *__restrict__ dinv, T omega, T *y, const T *dotw,
                                                   Scal *s, int check_done, double *partials, T *y2) {
  if (check_done && s->done) return;
  constexpr int CAP = rb_cap<T>();
  constexpr int Q = CAP / (4 * kBlock);  // quads per thread
  static_assert(Q * 4 * kBlock == CAP, "the row-block cap is a whole number of quads per thread");
  static_assert(rb_rows<T>() == kBlock, "one row per thread at most");
  constexpr bool STAGED = LAYOUT == kRbRows || LAYOUT == kRbStrided;
  static_assert(LAYOUT != kRbSorted || !C16, "column-sorted blocks keep packed 32-bit columns");
  __shared__ T sval[CAP + 4];  // staged values (row walk, strided), or products
  __shared__ int scol[STAGED ? CAP + 4 : 1];
  __shared__ T wsum[kBlock / 64];
  const int tid = threadIdx.x;
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3, per = gridDim.x >> 3;
  const int chunk = (nrb + 7) >> 3;
  const int q1 = min(nrb, xcd * chunk + chunk);
  double dacc = 0.0;
  int q = xcd * chunk + local;
  // block descriptors (first row, first nonzero) of this and the next block
  int2 d0 = make_int2(0, 0), d1 = make_int2(0, 0);
  int b0 = 0;
  if (q < q1) {
    d0 = rb[q];
    d1 = rb[q + 1];
    if constexpr (C16) b0 = cb[q];
  }
  // column of nonzero k of this block
  auto col = [&](int k) -> int {
    if constexpr (C16) return b0 + (int)c16[k];
    else return ci[k];
  };
  while (q < q1) {
    const int qn = q + per;
    int2 n0 = d0, n1 = d1;
    int nb = b0;
    if (qn < q1) {  // prefetched while this block is processed
      n0 = rb[qn];
      n1 = rb[qn + 1];
      if constexpr (C16) nb = cb[qn];
    }
    const int r0 = d0.x, e0 = d0.y, nr = d1.x - d0.x, ne = d1.y - d0.y;
    if (ne > CAP) {
      // a single long row, CAP entries per step: every thread loads Q quads
      // (all in flight before their gathers) and keeps a running sum
      const EpiIn<T> e = epi_load<T, EPI, DOT>(r0, x, b, dinv, omega, y, dotw);
      const int ea = e0 & ~3, e1 = e0 + ne;
      const int safe = C16 ? b0 : 0;
      T acc = T(0);
      for (int base = ea; base < e1; base += CAP) {
        int c[Q][4];
        T v[Q][4];
#pragma unroll
        for (int u = 0; u < Q; ++u) {
          const int k = base + 4 * (tid + u * kBlock);
          if (k < e1) {
            if constexpr (C16) load_quad_cols<NT>(c16 + k, b0, c[u]);
            else load_quad_cols<NT>(ci + k, c[u]);
            load_quad<NT>(av + k, v[u]);
          } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              c[u][j] = safe;
              v[u][j] = T(0);
            }
          }
        }
        // every gather of the step in flight before the sums
        T xv[Q][4];
#pragma unroll
        for (int u = 0; u < Q; ++u) {
          const int k = base + 4 * (tid + u * kBlock);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const bool ok = k + j >= e0 && k + j < e1;
            xv[u][j] = gather<T, EPI>(ok ? c[u][j] : safe, x, b, dinv, omega);
          }
        }
#pragma unroll
        for (int u = 0; u < Q; ++u) {
          const int k = base + 4 * (tid + u * kBlock);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const bool ok = k + j >= e0 && k + j < e1;
            acc += ok ? v[u][j] * xv[u][j] : T(0);
          }
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if ((tid & 63) == 0) wsum[tid >> 6] = acc;
      __syncthreads();
      if (tid == 0) {
        T sum = T(0);
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) sum += wsum[w];
        epi_store<T, EPI, DOT>(e, r0, sum, omega, y, dotw, dacc, y2);
      }
      __syncthreads();
    } else {
      // Row bounds and epilogue operands first, then every quad of the block
      // and the tail entries into LDS; after the barrier each row's lanes
      // walk its entries and gather x.
      const int ea = e0 & ~3, e1 = e0 + ne;
      int tpr = 1;
      while (tpr * 2 * nr <= kBlock && tpr < 64) tpr *= 2;
      const int lane = tid & (tpr - 1);
      const int rl = min(tid / tpr, nr - 1);
      const bool own = tid / tpr < nr;
      // both bounds loaded unconditionally (rl is clamped into the block), so
      // the select waits for nothing before the streams are issued
      const int ks = rp[r0 + rl], kr = rp[r0 + rl + 1];
      const int ke = own ? kr : ks;
      const EpiIn<T> e = epi_load<T, EPI, DOT>(r0 + rl, x, b, dinv, omega, y, dotw);
      // a column that is valid for any gather: entries outside the row or
      // the block gather it and drop the product (selected after the
      // multiply, so a stored zero times a non-finite x still gives NaN)
      const int safe = C16 ? b0 : 0;
      const T *sv = sval - ea;
      T sum = T(0);
      if constexpr (STAGED) {
        if (ne > 0) {
          // every quad and the tail loaded first (addresses clamped into the
          // block, so all loads are issued together), then stored to LDS
          int c[Q][4];
          T v[Q][4];
#pragma unroll
          for (int u = 0; u < Q; ++u) {
            const int k = ea + 4 * (tid + u * kBlock);
            const int kc = k < e1 ? k : ea;
            if constexpr (C16) load_quad_cols<NT>(c16 + kc, b0, c[u]);
            else load_quad_cols<NT>(ci + kc, c[u]);
            load_quad<NT>(av + kc, v[u]);
          }
          const int kt = ea + 4 * Q * kBlock + tid;  // entries past the quads (e1 - ea <= CAP + 3)
          const bool tail = tid < 3 && kt < e1;
          const int ktc = tail ? kt : ea;
          const int ct = col(ktc);
          const T vt = av[ktc];
#pragma unroll
          for (int u = 0; u < Q; ++u) {
            const int k = ea + 4 * (tid + u * kBlock);
            if (k < e1) {
              store_quad(&scol[k - ea], c[u]);
              store_quad(&sval[k - ea], v[u]);
            }
          }
          if (tail) {
            scol[kt - ea] = ct;
            sval[kt - ea] = vt;
          }
        }
        __syncthreads();
        const int *sc = scol - ea;
        if constexpr (LAYOUT == kRbStrided) {
          // entries dealt to threads round robin (thread tid takes
          // e0 + tid + kBlock u), so a wavefront gathers for 64 consecutive
          // entries; each product overwrites its own staged value, then rows
          // are summed from LDS as in the quad layout.  A thread without an
          // entry reads slot CAP + 3, which no block stages.
          constexpr int U = CAP / kBlock;
          int kk[U];
          T vk[U], xv[U];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int k = e0 + tid + kBlock * u;
            const bool ok = k < e1;
            kk[u] = ok ? k - ea : CAP + 3;
            const int cc = scol[kk[u]];
            vk[u] = sval[kk[u]];
            xv[u] = gather<T, EPI>(ok ? cc : safe, x, b, dinv, omega);
          }
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (kk[u] != CAP + 3) sval[kk[u]] = vk[u] * xv[u];
          __syncthreads();
          T s4[4] = {T(0), T(0), T(0), T(0)};
          int k = ks + lane;
          for (; k + 3 * tpr < ke; k += 4 * tpr) {
#pragma unroll
            for (int u = 0; u < 4; ++u) s4[u] += sv[k + u * tpr];
          }
#pragma unroll
          for (int u = 0; u < 3; ++u)
            if (k + u * tpr < ke) s4[u] += sv[k + u * tpr];
          sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
        } else if (tpr == 1) {
          // the row's entries in order, 8 gathers in flight
          for (int k = ks; k < ke; k += 8) {
            T vv[8], xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              // LDS index clamped to the row's first entry, so the reads
              // are unconditional and issue back to back
              const bool ok = k + u < ke;
              const int kk = ok ? k + u : ks;
              const int cc = sc[kk];
              const T vk = sv[kk];
              vv[u] = ok ? vk : T(0);
              xv[u] = gather<T, EPI>(ok ? cc : safe, x, b, dinv, omega);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
              if (k + u < ke) sum += vv[u] * xv[u];
          }
        } else {
          // tpr lanes per row, four running sums per lane
          T s4[4] = {T(0), T(0), T(0), T(0)};
          for (int k = ks + lane; k < ke; k += 4 * tpr) {
            T vv[4], xv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const bool ok = k + u * tpr < ke;
              const int kk = ok ? k + u * tpr : k;
              const int cc = sc[kk];
              const T vk = sv[kk];
              vv[u] = ok ? vk : T(0);
              xv[u] = gather<T, EPI>(ok ? cc : safe, x, b, dinv, omega);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (k + u * tpr < ke) s4[u] += vv[u] * xv[u];
          }
          sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
        }
      } else {
        // gathers in the quad layout, products parked in LDS; column-sorted
        // blocks park each product at its entry's row-ordered position
        if (ne > 0) {
          int c[Q][4];
          T v[Q][4], xv[Q][4];
#pragma unroll
          for (int u = 0; u < Q; ++u) {
            const int k = ea + 4 * (tid + u * kBlock);
            if (k < e1) {
              if constexpr (C16) load_quad_cols<NT>(c16 + k, b0, c[u]);
              else load_quad_cols<NT>(ci + k, c[u]);
              load_quad<NT>(av + k, v[u]);
            } else {
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                c[u][j] = safe;
                v[u][j] = T(0);
              }
            }
          }
          const int kt = ea + 4 * Q * kBlock + tid;
          const bool tail = tid < 3 && kt < e1;
          int ct = tail ? col(kt) : safe;
          const T vt = tail ? av[kt] : T(0);
          // LDS slot of each product (relative to ea)
          int slot[Q][4];
          int st = kt - ea;
#pragma unroll
          for (int u = 0; u < Q; ++u) {
            const int k = ea + 4 * (tid + u * kBlock);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              if constexpr (LAYOUT == kRbSorted) {
                slot[u][j] = (int)((uint32_t)c[u][j] >> kRbPosShift) + (e0 - ea);
                c[u][j] &= kRbColMask;
              } else {
                slot[u][j] = k + j - ea;
              }
            }
          }
          if constexpr (LAYOUT == kRbSorted) {
            st = (int)((uint32_t)ct >> kRbPosShift) + (e0 - ea);
            ct &= kRbColMask;
          }
#pragma unroll
          for (int u = 0; u < Q; ++u) {
            const int k = ea + 4 * (tid + u * kBlock);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const bool ok = k + j >= e0 && k + j < e1;
              xv[u][j] = gather<T, EPI>(ok ? c[u][j] : safe, x, b, dinv, omega);
            }
          }
          const T xt = gather<T, EPI>(ct, x, b, dinv, omega);
#pragma unroll
          for (int u = 0; u < Q; ++u) {
            const int k = ea + 4 * (tid + u * kBlock);
            if (k < e1) {
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                const bool ok = k + j >= e0 && k + j < e1;
                if constexpr (LAYOUT == kRbSorted) {
                  // entries outside the block belong to its neighbours
                  if (ok) sval[slot[u][j]] = v[u][j] * xv[u][j];
                } else {
                  sval[slot[u][j]] = ok ? v[u][j] * xv[u][j] : T(0);
                }
              }
            }
          }
          if (tail) sval[st] = vt * xt;
        }
        __syncthreads();
        // four running sums per lane, so four LDS reads are in flight
        T s4[4] = {T(0), T(0), T(0), T(0)};
        int k = ks + lane;
        for (; k + 3 * tpr < ke; k += 4 * tpr) {
#pragma unroll
          for (int u = 0; u < 4; ++u) s4[u] += sv[k + u * tpr];
        }
#pragma unroll
        for (int u = 0; u < 3; ++u)
          if (k + u * tpr < ke) s4[u] += sv[k + u * tpr];
        sum = (s4[0] + s4[1]) + (s4[2] + s4[3]);
      }
      for (int o = tpr >> 1; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
      if (own && lane == 0) epi_store<T, EPI, DOT>(e, r0 + tid / tpr, sum, omega, y, dotw, dacc, y2);
      __syncthreads();
    }
    d0 = n0;
    d1 = n1;
    b0 = nb;
    q = qn;
  }
  if constexpr (DOT) {
    double dv[1] = {dacc};
    block_partials<1>(dv, partials);
  }
}

// Which products load their streams nontemporally, measured on the 128^3
// solves (FP64 / FP32-refined ms, two runs each): never 6.54 / 4.54; always
// 6.67 / 4.33; FP64 operators of at least 32 MiB 6.53 / 4.42, of at least
// 64 MiB 6.54 / 4.55.  So FP32 products always stream nontemporally and FP64
// ones from 32 MiB up: the small FP64 operators (level-1 transfers, level 2)
// gain from keeping their lines for the next sweep, the FP32 ones do not.
#ifndef CFDLA_RB_NT_MIN_BYTES
#define CFDLA_RB_NT_MIN_BYTES (32.0 * 1024.0 * 1024.0)
#endif

template <typename T, int EPI, bool DOT, bool C16, bool NT, int LAYOUT>
void launch_csr_rb_k(const DevCsr &A, int blocks, const T *x, const T *b, const T *dinv, double omega, T *y,
                     const T *dotw, Scal *s, bool check_done, RedWs &ws, T *y2) {
  const int2 *rb = reinterpret_cast<const int2 *>(A.rb<T>());
  hipLaunchKernelGGL((k_csr_rb<T, EPI, DOT, C16, NT, LAYOUT>), dim3(blocks), dim3(kBlock), 0, stream(), A.nrb<T>(), rb,
                     A.rp.p, C16 ? nullptr : A.rb_cols<T>(), C16 ? A.cols16<T>() : nullptr,
                     C16 ? A.cbase<T>() : nullptr, A.rb_vals<T>(), x, b, dinv, (T)omega, y, dotw, s,
                     check_done ? 1 : 0, ws.partials.p, y2);
}

template <typename T, int EPI, bool DOT>
void launch_csr_rb(const DevCsr &A, const T *x, const T *b, const T *dinv, double omega, T *y, const T *dotw,
                   Scal *s, bool check_done, int op, RedWs &ws, T *y2 = nullptr) {
  const int nrb = A.nrb<T>();
  const int blocks = xcd_grid(nrb, 1, kMaxBlocks);
  const bool c16 = A.has_cols16<T>();
  const bool nt = sizeof(T) == 4 || (double)A.nnz * (double)(sizeof(T) + (c16 ? 2 : 4)) >= CFDLA_RB_NT_MIN_BYTES;
  const int layout = A.rb_layout<T>();
#define CFDLA_RB(C, N, L) \
  launch_csr_rb_k<T, EPI, DOT, C, N, L>(A, blocks, x, b, dinv, omega, y, dotw, s, check_done, ws, y2)
#define CFDLA_RB3(C, N)                                  \
  if (layout == kRbRows) CFDLA_RB(C, N, kRbRows);        \
  else if (layout == kRbStrided) CFDLA_RB(C, N, kRbStrided); \
  else CFDLA_RB(C, N, kRbQuad);
  // column-sorted blocks never have 16-bit columns
#define CFDLA_RB4(N)                                      \
  if (layout == kRbSorted) CFDLA_RB(false, N, kRbSorted); \
  else { CFDLA_RB3(false, N) }
  if (c16) {
    if (nt) { CFDLA_RB3(true, true) }
    else { CFDLA_RB3(true, false) }
  } else {
    if (nt) { CFDLA_RB4(true) }
    else { CFDLA_RB4(false) }
  }
#undef CFDLA_RB4
#undef CFDLA_RB3
#undef CFDLA_RB
  if (DOT) finish(1, s, op, blocks, ws, check_done);
}

// ---- sliced-ELL product ------------------------------------------------------
// One lane per row of a 64-row slice; KS wavefronts share a slice, each
// summing a contiguous 1/KS of its slot groups, and the partial sums meet in
// LDS.  A lane's slots come in groups of GS (1 or 4) consecutive entries, one
// load of the values and one of the offsets per group; GPS groups per
// unrolled step are all in flight before their x gathers.  (With one
// wavefront per slice the first AMG level of the 128^3 problem, 47 slots per
// row, ran a chain of 6 dependent load-then-gather steps per wavefront at
// about 4 wavefronts per SIMD: latency-bound, 19 us per FP32 sweep for
// 65 MB.)  Workgroups are dealt to the 8 XCDs round-robin (block b runs on
// XCD b % 8), so the slices are split into 8 contiguous ranges and XCD q
// walks range q: the x gathers of neighbouring rows then hit the L2 of the
// XCD that owns them.  The grid is a multiple of 8 blocks.
typedef short v4i16 __attribute__((ext_vector_type(4)));

// The streams are loaded with the default cache policy: the level matrices
// are re-read by the next sweep of the same level, and nontemporal loads,
// which keep them out of the caches, ran the 128^3 solves slower (FP64 7.00
// against 6.51 ms, FP32 4.52 against 4.33 ms).
template <int GS, typename T>
__device__ __forceinline__ void sell_group_load(const short *d, const T *v, int (&dd)[GS], T (&vv)[GS]) {
  if constexpr (GS == 4) {
    const v4i16 a = *reinterpret_cast<const v4i16 *>(d);
    dd[0] = a.x; dd[1] = a.y; dd[2] = a.z; dd[3] = a.w;
    load_quad<false>(v, vv);
  } else {
    dd[0] = d[0];
    vv[0] = v[0];
  }
}

template <typename T, int EPI, bool DOT, int GS, int GPS, int KS>
__global__ __launch_bounds__(kBlock) void k_sell(int n, int nslices, const int *__restrict__ sp,
                                                 const short *__restrict__ dl, const T *__restrict__ av,
                                                 const T *__restrict__ x, const T *__restrict__ b,
                                                 const T *__restrict__ dinv, T omega, T *y, const T *dotw, Scal *s,
                                                 int check_done, double *partials) {
  if (check_done && s->done) return;
  constexpr int WPB = kBlock / 64;
  constexpr int SPB = WPB / KS;     // slices per block and step
  constexpr int GSTRIDE = 64 * GS;  // slots between a lane's consecutive groups
  static_assert(WPB % KS == 0, "KS must divide the wavefronts of a block");
  __shared__ T part[KS > 1 ? (KS - 1) * SPB * 64 : 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sw = wave / KS, kp = wave % KS;  // slice of the step, part of its slots
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3, per = gridDim.x >> 3;
  const int chunk = (nslices + 7) >> 3;
  const int s0 = xcd * chunk;
  const int s1 = min(nslices, s0 + chunk);
  double dacc = 0.0;
  // block-uniform trip count: every wavefront reaches the barriers
  for (int sb = s0 + local * SPB; sb < s1; sb += per * SPB) {
    const int sl = sb + sw;
    const int row = sl * 64 + lane;
    const bool live = sl < s1 && row < n;
    T sum = T(0);
    EpiIn<T> e{T(0), T(0), T(0), T(0), T(0)};
    if (live) {
      const int base = sp[sl];
      const int ng = (sp[sl + 1] - base) / GSTRIDE;
      const int ga = ng * kp / KS, gb = ng * (kp + 1) / KS;
      const short *dd = dl + base + lane * GS;
      const T *vv = av + base + lane * GS;
      if (kp == 0) e = epi_load<T, EPI, DOT>(row, x, b, dinv, omega, y, dotw);
      // Branch-free: group indices clamped to the part, padding gathers
      // x[row] and is masked out of the sum, so all loads of a step are in
      // flight at once.  Entries are summed in row order within a part.
      for (int g0 = ga; g0 < gb; g0 += GPS) {
        int d[GPS][GS];
        T v[GPS][GS], xv[GPS][GS];
        bool ok[GPS][GS];
#pragma unroll
        for (int u = 0; u < GPS; ++u) {
          const int g = min(g0 + u, gb - 1);
          sell_group_load<GS>(dd + g * GSTRIDE, vv + g * GSTRIDE, d[u], v[u]);
        }
#pragma unroll
        for (int u = 0; u < GPS; ++u)
#pragma unroll
          for (int j = 0; j < GS; ++j) {
            ok[u][j] = g0 + u < gb && d[u][j] != kSellPad;
            xv[u][j] = gather<T, EPI>(row + (ok[u][j] ? d[u][j] : 0), x, b, dinv, omega);
          }
#pragma unroll
        for (int u = 0; u < GPS; ++u)
#pragma unroll
          for (int j = 0; j < GS; ++j) sum += ok[u][j] ? v[u][j] * xv[u][j] : T(0);
      }
    }
    if constexpr (KS > 1) {
      if (kp > 0) part[((kp - 1) * SPB + sw) * 64 + lane] = sum;
      __syncthreads();
      if (kp == 0)
#pragma unroll
        for (int q = 1; q < KS; ++q) sum += part[((q - 1) * SPB + sw) * 64 + lane];
      __syncthreads();  // the next step rewrites part
    }
    if (kp == 0 && live) epi_store<T, EPI, DOT>(e, row, sum, omega, y, dotw, dacc);
  }
  if constexpr (DOT) {
    double v[1] = {dacc};
    block_partials<1>(v, partials);
  }
}

template <typename T, int EPI, bool DOT>
void launch_sell(const DevCsr &A, const T *x, const T *b, const T *dinv, double omega, T *y, const T *dotw,
                 Scal *s, bool check_done, int op, RedWs &ws) {
  const int cd = check_done ? 1 : 0;
  if (A.sell_group == 4) {
    // 4 wavefronts per slice, one group in flight per step.  On the first
    // AMG level of the 128^3 problem (solves FP64 / FP32-refined, ms): one
    // wavefront per slice 6.52 / 4.38 (groups of 1: 6.52 / 4.34); two per
    // slice 6.40 / 4.26; four 6.35 / 4.23 (two groups per step) and
    // 6.34 / 4.22 (one).
    constexpr int KS = 4;
    const int blocks = xcd_grid(A.nslices, (kBlock / 64) / KS, kMaxBlocks);
    hipLaunchKernelGGL((k_sell<T, EPI, DOT, 4, 1, KS>), dim3(blocks), dim3(kBlock), 0, stream(), A.n, A.nslices,
                       A.sell_ptr.p, A.sell_delta.p, A.sell_vals<T>(), x, b, dinv, (T)omega, y, dotw, s, cd,
                       ws.partials.p);
    if (DOT) finish(1, s, op, blocks, ws, check_done);
  } else {
    // 8 entries per unrolled step: on the first AMG level of the 128^3
    // problem 16 per step measured slower in both precisions (solve 6.69
    // against 6.60 ms FP64, 4.88 against 4.83 ms FP32)
    const int blocks = xcd_grid(A.nslices, kBlock / 64);
    hipLaunchKernelGGL((k_sell<T, EPI, DOT, 1, 8, 1>), dim3(blocks), dim3(kBlock), 0, stream(), A.n, A.nslices,
                       A.sell_ptr.p, A.sell_delta.p, A.sell_vals<T>(), x, b, dinv, (T)omega, y, dotw, s, cd,
                       ws.partials.p);
    if (DOT) finish(1, s, op, blocks, ws, check_done);
  }
}

template <typename T, int EPI, bool DOT>
void launch_csr_e(const DevCsr &A, const T *x, const T *b, const T *dinv, double omega, T *y, const T *dotw,
                  Scal *s, bool check_done, int op, RedWs &ws) {
  if (A.dia && A.dia_vals<T>() != nullptr) {
    launch_dia<T, EPI, DOT>(A, x, b, dinv, omega, y, dotw, s, check_done, op, ws);
    return;
  }
  if (A.sell && A.sell_vals<T>() != nullptr) {
    launch_sell<T, EPI, DOT>(A, x, b, dinv, omega, y, dotw, s, check_done, op, ws);
    return;
  }
  launch_csr_rb<T, EPI, DOT>(A, x, b, dinv, omega, y, dotw, s, check_done, op, ws);
}

// ---- reductions and vector updates ---------------------------------------------
template <typename T, int NV>
__global__ __launch_bounds__(kBlock) void k_dot(int n, const T *__restrict__ a0, const T *__restrict__ b0,
                                                const T *__restrict__ a1, const T *__restrict__ b1, Scal *s,
                                                int check_done, int op, double *partials, unsigned *counter) {
  if (check_done && s->done) return;
  double v[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    v[0] += (double)a0[i] * (double)b0[i];
    if constexpr (NV > 1) v[1] += (double)a1[i] * (double)b1[i];
  }
  reduce_and_finalize<NV>(v, s, op, partials, counter);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_jacobi_first(int n, const T *__restrict__ dinv, T omega,
                                                         const T *__restrict__ b, T *__restrict__ x, Scal *s,
                                                         int check_done) {
  if (check_done && s->done) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] = omega * dinv[i] * b[i];
}

template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_cg_update(int n, const T *__restrict__ p, const T *__restrict__ q,
                                                      T *__restrict__ x, T *__restrict__ r,
                                                      const T *__restrict__ jdinv, T *__restrict__ z, Scal *s,
                                                      double *partials, unsigned *counter) {
  if (s->done) return;
  const T alpha = (T)s->alpha;
  double v[JAC ? 2 : 1];
  v[0] = 0.0;
  if constexpr (JAC) v[1] = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    T ri = r[i] - alpha * q[i];
    x[i] += alpha * p[i];
    r[i] = ri;
    v[0] += (double)ri * (double)ri;
    if constexpr (JAC) {
      T zi = jdinv[i] * ri;
      z[i] = zi;
      v[1] += (double)ri * (double)zi;
    }
  }
  reduce_and_finalize<JAC ? 2 : 1>(v, s, JAC ? OP_CG_RR_RZ : OP_CG_RR, partials, counter);
}

// The FP32 update of a step that replaces the residual, with the correction
// folded into the FP64 iterate in the same pass: xc + alpha p goes to x64 and
// xc restarts at zero, r -= alpha q, r.r for the finalize op.  The same
// operations, in the same forms, as k_cg_update<float, false> followed by
// k_fold_correction.
__global__ __launch_bounds__(kBlock) void k_cg_update_fold(int n, const float *__restrict__ p,
                                                           const float *__restrict__ q, float *__restrict__ xc,
                                                           double *__restrict__ x64, float *__restrict__ r, Scal *s,
                                                           double *partials, unsigned *counter) {
  if (s->done) return;
  const float alpha = (float)s->alpha;
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    float ri = r[i] - alpha * q[i];
    float ci = xc[i];
    ci += alpha * p[i];
    x64[i] += (double)ci;
    xc[i] = 0.0f;
    r[i] = ri;
    v[0] += (double)ri * (double)ri;
  }
  reduce_and_finalize<1>(v, s, OP_CG_RR, partials, counter);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_direction(int n, const T *__restrict__ z, T *__restrict__ p,
                                                         Scal *s) {
  if (s->done) return;
  const T beta = (T)s->beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = cg_dir(beta, p[i], z[i]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy(int n, const T *__restrict__ src, T *__restrict__ dst, Scal *s,
                                                 int check_done) {
  if (check_done && s->done) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) dst[i] = src[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(int n, T value, T *__restrict__ dst) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) dst[i] = value;
}

__global__ __launch_bounds__(kBlock) void k_f64_to_f32_scaled(int n, const double *__restrict__ src, double scale,
                                                              float *__restrict__ dst) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) dst[i] = (float)(src[i] * scale);
}

__global__ __launch_bounds__(kBlock) void k_axpy_f32_to_f64(int n, double a, const float *__restrict__ x,
                                                            double *__restrict__ y) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] += a * (double)x[i];
}

__global__ __launch_bounds__(kBlock) void k_fold_correction(int n, float *__restrict__ xc, double *__restrict__ x,
                                                            Scal *s, int check_done) {
  if (check_done && s->done) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] += (double)xc[i];
    xc[i] = 0.0f;
  }
}

__global__ __launch_bounds__(kBlock) void k_f64_to_f32(int n, const double *__restrict__ src,
                                                       float *__restrict__ dst, Scal *s, int check_done) {
  if (check_done && s->done) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) dst[i] = (float)src[i];
}

// x = A^-1 b with the dense inverse, one wavefront per row.  b[j] is the
// sum of `parts` consecutive entries (parts > 1: the partial sums of a
// restriction whose long rows were split into pieces).  Each block stages
// the sums in LDS while its rows' first 16 x 64 inverse entries are already
// in flight; the row sum runs in the order j = lane, lane + 64, ...
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dense_gemv(int n, const double *__restrict__ a,
                                                       const T *__restrict__ b, int parts, T *__restrict__ x,
                                                       Scal *s, int check_done) {
  if (check_done && s->done) return;
  extern __shared__ double bs[];  // n sums
  constexpr int PF = 16;
  const int lane = threadIdx.x & 63;
  const int row = (int)((blockIdx.x * (unsigned)kBlock + threadIdx.x) >> 6);
  double av[PF];
#pragma unroll
  for (int u = 0; u < PF; ++u) {
    const int j = lane + 64 * u;
    av[u] = row < n && j < n ? a[(size_t)row * n + j] : 0.0;
  }
  for (int j = threadIdx.x; j < n; j += kBlock) {
    double bj = (double)b[(size_t)j * parts];
    for (int k = 1; k < parts; ++k) bj += (double)b[(size_t)j * parts + k];
    bs[j] = bj;
  }
  __syncthreads();
  if (row >= n) return;  // whole wave leaves together
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < PF; ++u) {
    const int j = lane + 64 * u;
    if (j < n) acc += av[u] * bs[j];
  }
  for (int j = lane + 64 * PF; j < n; j += 64) acc += a[(size_t)row * n + j] * bs[j];
  acc = wave_sum(acc);
  if (lane == 0) x[row] = (T)acc;
}

// BiCGSTAB: p = r + beta (p - omega v)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_p(int n, const T *__restrict__ r, const T *__restrict__ v,
                                                   T *__restrict__ p, Scal *s) {
  if (s->done) return;
  const T beta = (T)s->beta, omega = (T)s->omega;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

// BiCGSTAB: s = r - alpha v, ||s||^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_s(int n, const T *__restrict__ r, const T *__restrict__ v,
                                                   T *__restrict__ sv, Scal *s, double *partials,
                                                   unsigned *counter) {
  if (s->done) return;
  const T alpha = (T)s->alpha;
  double acc[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    T si = r[i] - alpha * v[i];
    sv[i] = si;
    acc[0] += (double)si * (double)si;
  }
  reduce_and_finalize<1>(acc, s, OP_BI_SNORM, partials, counter);
}

// BiCGSTAB: x += alpha ph + omega sh; r = s - omega t; ||r||^2 (and, with
// RHO, rhat . r for the next iteration)
template <typename T, bool RHO>
__global__ __launch_bounds__(kBlock) void k_bicg_xr(int n, const T *__restrict__ ph, const T *__restrict__ sh,
                                                    const T *__restrict__ sv, const T *__restrict__ t,
                                                    T *__restrict__ x, T *__restrict__ r, const T *__restrict__ rhat,
                                                    Scal *s, double *partials, unsigned *counter) {
  if (s->done) return;
  const T alpha = (T)s->alpha, omega = (T)s->omega;
  double acc[RHO ? 2 : 1] = {};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] += alpha * ph[i] + omega * sh[i];
    T ri = sv[i] - omega * t[i];
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
    if constexpr (RHO) acc[1] += (double)rhat[i] * (double)ri;
  }
  reduce_and_finalize<RHO ? 2 : 1>(acc, s, RHO ? OP_BI_RR_RHO : OP_BI_RR, partials, counter);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy_alpha(int n, const T *__restrict__ x, T *__restrict__ y,
                                                       Scal *s) {
  const T a = (T)s->alpha;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) y[i] += a * x[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sor_colour(int count, const int *__restrict__ perm,
                                                       const int *__restrict__ rp, const int *__restrict__ ci,
                                                       const T *__restrict__ v, const T *__restrict__ piv,
                                                       const T *__restrict__ r, T *z, Scal *s, int check_done) {
  if (check_done && s->done) return;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < count; t += gridDim.x * kBlock) {
    const int i = perm[t];
    T acc = r[i];
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = ci[k];
      if (j != i) acc -= v[k] * z[j];
    }
    z[i] = piv[i] * acc;
  }
}

// One colour of a multicolour Gauss-Seidel sweep on A x = b, the reference's
// row update (smoothers.rs:42-82): with the diagonal d and the off-diagonal
// sum of row i (columns in stored order), x_i += omega ((b_i - sum) / d - x_i)
// unless |d| <= 1e-15.  Rows of one colour are not coupled, so updating them
// together equals updating them one after another.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gs_colour(int count, const int *__restrict__ perm,
                                                      const int *__restrict__ rp, const int *__restrict__ ci,
                                                      const T *__restrict__ v, const T *__restrict__ b, T *x, T omega,
                                                      Scal *s, int check_done) {
  if (check_done && s->done) return;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < count; t += gridDim.x * kBlock) {
    const int i = perm[t];
    T sum = T(0), diag = T(0);
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = ci[k];
      if (j == i) diag = v[k];
      else sum += v[k] * x[j];
    }
    if (fabs((double)diag) > 1e-15) {
      const T nv = (b[i] - sum) / diag;
      x[i] = x[i] + omega * (nv - x[i]);
    }
  }
}

// The first colour of a sweep from z = 0 sees only zero neighbours:
// z_i = piv_i r_i there and 0 elsewhere, one streaming pass instead of a
// fill and a sweep over the matrix rows.  (After convergence, z = 0.)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sor_first(int n, const uint8_t *__restrict__ first,
                                                      const T *__restrict__ piv, const T *__restrict__ r, T *z,
                                                      Scal *s, int check_done) {
  const bool skip = check_done && s->done;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    z[i] = (!skip && first[i]) ? piv[i] * r[i] : T(0);
}

__global__ void k_scal_reset(Scal *s, double thr2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int i = 0; i < 8; ++i) s->red[i] = 0.0;
    s->rz = s->beta = s->rr = s->pq = s->hn = s->rr_true = 0.0;
    s->rho = s->omega = s->alpha = 1.0;
    s->thr2 = s->thr_final2 = thr2;
    s->done = 0;
    s->iters = 0;
    s->kstep = 0;
    s->nupd = 0;
    s->restart_dir = 0;
    s->nrestart = 0;
    s->term = CFDLA_TERM_MAX_ITERATIONS;
  }
}

__global__ void k_scal_set_done(Scal *s, int value) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->done = value;
}

__global__ void k_scal_set_final(Scal *s, double thr_final2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->thr_final2 = thr_final2;
}

__global__ void k_scal_bicg_restart(Scal *s) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    s->rho = s->alpha = s->omega = 1.0;
    s->done = 0;
  }
}

// dJ/dg for the 2D variable-coefficient operator (one thread per cell).
__global__ __launch_bounds__(kBlock) void k_coef_grad_2d(int nx, int ny, double ih2, int dir_mask,
                                                         double dv0, double dv1, double dv2, double dv3,
                                                         const double *__restrict__ k,
                                                         const double *__restrict__ p,
                                                         const double *__restrict__ lam,
                                                         const double *__restrict__ dJdk,
                                                         double *__restrict__ dJdg) {
  const int n = nx * ny;
  for (int c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock) {
    const int x = c % nx, y = c / nx;
    const double kp = k[c], pp = p[c], lp = lam[c];
    double g = dJdk ? dJdk[c] : 0.0;
    auto face = [&](int q) {
      const double kq = k[q], s = kp + kq;
      const double dkf = 2.0 * kq * kq / (s * s);
      g -= ih2 * dkf * (lp - lam[q]) * (pp - p[q]);
    };
    if (x > 0) face(c - 1);
    if (x + 1 < nx) face(c + 1);
    if (y > 0) face(c - nx);
    if (y + 1 < ny) face(c + nx);
    if (x == 0 && (dir_mask & 1)) g -= ih2 * lp * (pp - dv0);
    if (x == nx - 1 && (dir_mask & 2)) g -= ih2 * lp * (pp - dv1);
    if (y == 0 && (dir_mask & 4)) g -= ih2 * lp * (pp - dv2);
    if (y == ny - 1 && (dir_mask & 8)) g -= ih2 * lp * (pp - dv3);
    dJdg[c] = kp * g;
  }
}

}  // namespace

// ---- launchers -----------------------------------------------------------------------
template <typename T>
void csr_apply(const DevCsr &A, int epi, const T *x, const T *b, const T *dinv, double omega, T *y, Scal *s,
               bool check_done, int dot_op, const T *dotw, RedWs &ws, T *y2) {
  const bool dot = dot_op != OP_NONE;
  if (!dotw) dotw = x;
#define CFDLA_EPI_CASE(E)                                                                         \
  case E:                                                                                         \
    if (dot) launch_csr_e<T, E, true>(A, x, b, dinv, omega, y, dotw, s, check_done, dot_op, ws);   \
    else launch_csr_e<T, E, false>(A, x, b, dinv, omega, y, dotw, s, check_done, dot_op, ws);     \
    break;
  switch (epi) {
    CFDLA_EPI_CASE(EPI_Y)
    CFDLA_EPI_CASE(EPI_RESID)
    CFDLA_EPI_CASE(EPI_JACOBI)
    CFDLA_EPI_CASE(EPI_ADD)
    CFDLA_EPI_CASE(EPI_JACOBI0)
    case EPI_RJ0:
      CFDLA_REQUIRE(!dot && y2 != nullptr && dinv != nullptr, CFDLA_ERR_INVALID_ARG,
                    "restriction epilogue takes the coarse dinv and a second output, no dot");
      launch_csr_rb<T, EPI_RJ0, false>(A, x, b, dinv, omega, y, dotw, s, check_done, dot_op, ws, y2);
      break;
    case EPI_DIR:
      CFDLA_REQUIRE(dot && y2 != nullptr && b != nullptr && fuses_direction<T>(A), CFDLA_ERR_INVALID_ARG,
                    "direction epilogue takes diagonal storage, z, a second output and a dot");
      launch_dia<T, EPI_DIR, true>(A, x, b, dinv, omega, y, dotw, s, check_done, dot_op, ws, y2);
      break;
    default:
      throw Error(CFDLA_ERR_INVALID_ARG, "unknown CSR epilogue");
  }
#undef CFDLA_EPI_CASE
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void dot2(int n, const T *a0, const T *b0, const T *a1, const T *b1, Scal *s, bool check_done, int op,
          RedWs &ws) {
  int blocks = blocks_for(n, kBlock * 4);
  if (a1)
    hipLaunchKernelGGL((k_dot<T, 2>), dim3(blocks), dim3(kBlock), 0, stream(), n, a0, b0, a1, b1, s,
                       check_done ? 1 : 0, op, ws.partials.p, ws.counter.p);
  else
    hipLaunchKernelGGL((k_dot<T, 1>), dim3(blocks), dim3(kBlock), 0, stream(), n, a0, b0, a1, b1, s,
                       check_done ? 1 : 0, op, ws.partials.p, ws.counter.p);
  finish(a1 ? 2 : 1, s, op, blocks, ws, check_done);
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void jacobi_first(int n, const T *dinv, double omega, const T *b, T *x, Scal *s, bool check_done) {
  hipLaunchKernelGGL((k_jacobi_first<T>), dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, dinv,
                     (T)omega, b, x, s, check_done ? 1 : 0);
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void cg_update(int n, const T *p, const T *q, T *x, T *r, const T *jdinv, T *z, Scal *s, RedWs &ws) {
  int blocks = blocks_for(n, kBlock * 4);
  if (jdinv)
    hipLaunchKernelGGL((k_cg_update<T, true>), dim3(blocks), dim3(kBlock), 0, stream(), n, p, q, x, r, jdinv, z,
                       s, ws.partials.p, ws.counter.p);
  else
    hipLaunchKernelGGL((k_cg_update<T, false>), dim3(blocks), dim3(kBlock), 0, stream(), n, p, q, x, r, jdinv,
                       z, s, ws.partials.p, ws.counter.p);
  finish(jdinv ? 2 : 1, s, jdinv ? OP_CG_RR_RZ : OP_CG_RR, blocks, ws, true);
  CFDLA_HIP(hipGetLastError());
}

void cg_update_fold(int n, const float *p, const float *q, float *xc, double *x64, float *r, Scal *s, RedWs &ws) {
  const int blocks = blocks_for(n, kBlock * 4);
  hipLaunchKernelGGL(k_cg_update_fold, dim3(blocks), dim3(kBlock), 0, stream(), n, p, q, xc, x64, r, s,
                     ws.partials.p, ws.counter.p);
  finish(1, s, OP_CG_RR, blocks, ws, true);
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void cg_direction(int n, const T *z, T *p, Scal *s) {
  hipLaunchKernelGGL((k_cg_direction<T>), dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, z, p, s);
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void copy_vec(int n, const T *src, T *dst, Scal *s, bool check_done) {
  hipLaunchKernelGGL((k_copy<T>), dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, src, dst, s,
                     check_done ? 1 : 0);
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_chk(int n, T *__restrict__ dst, Scal *s, int check_done) {
  if (check_done && s->done) return;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) dst[i] = T(0);
}

template <typename T>
void zero_vec(int n, T *dst, Scal *s, bool check_done) {
  if (n == 0) return;
  hipLaunchKernelGGL((k_zero_chk<T>), dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, dst, s,
                     check_done ? 1 : 0);
}
template void zero_vec<double>(int, double *, Scal *, bool);
template void zero_vec<float>(int, float *, Scal *, bool);

template <typename T>
void fill_vec(int n, T value, T *dst) {
  hipLaunchKernelGGL((k_fill<T>), dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, value, dst);
  CFDLA_HIP(hipGetLastError());
}

void f64_to_f32_scaled(int n, const double *src, double scale, float *dst) {
  hipLaunchKernelGGL(k_f64_to_f32_scaled, dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, src,
                     scale, dst);
  CFDLA_HIP(hipGetLastError());
}

void axpy_f32_to_f64(int n, double a, const float *x, double *y) {
  hipLaunchKernelGGL(k_axpy_f32_to_f64, dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, a, x, y);
  CFDLA_HIP(hipGetLastError());
}

void fold_correction(int n, float *xc, double *x, Scal *s, bool check_done) {
  hipLaunchKernelGGL(k_fold_correction, dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, xc, x, s,
                     check_done ? 1 : 0);
  CFDLA_HIP(hipGetLastError());
}

void f64_to_f32(int n, const double *src, float *dst, Scal *s, bool check_done) {
  hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, src, dst, s,
                     check_done ? 1 : 0);
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void dense_gemv(int n, const double *ainv, const T *b, T *x, Scal *s, bool check_done, int parts) {
  int blocks = (int)(((int64_t)n * 64 + kBlock - 1) / kBlock);
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((k_dense_gemv<T>), dim3(blocks), dim3(kBlock), (size_t)n * sizeof(double), stream(), n, ainv,
                     b, parts, x, s, check_done ? 1 : 0);
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void bicg_p_update(int n, const T *r, const T *v, T *p, Scal *s) {
  hipLaunchKernelGGL((k_bicg_p<T>), dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, r, v, p, s);
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void bicg_s_update(int n, const T *r, const T *v, T *sv, Scal *s, RedWs &ws) {
  const int blocks = blocks_for(n, kBlock * 4);
  hipLaunchKernelGGL((k_bicg_s<T>), dim3(blocks), dim3(kBlock), 0, stream(), n, r, v, sv, s, ws.partials.p,
                     ws.counter.p);
  finish(1, s, OP_BI_SNORM, blocks, ws, true);
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void bicg_x_r_update(int n, const T *ph, const T *sh, const T *sv, const T *t, T *x, T *r, Scal *s, RedWs &ws,
                     const T *rhat) {
  const int blocks = blocks_for(n, kBlock * 4);
  if (rhat) {
    hipLaunchKernelGGL((k_bicg_xr<T, true>), dim3(blocks), dim3(kBlock), 0, stream(), n, ph, sh, sv, t, x, r, rhat,
                       s, ws.partials.p, ws.counter.p);
    finish(2, s, OP_BI_RR_RHO, blocks, ws, true);
  } else {
    hipLaunchKernelGGL((k_bicg_xr<T, false>), dim3(blocks), dim3(kBlock), 0, stream(), n, ph, sh, sv, t, x, r,
                       nullptr, s, ws.partials.p, ws.counter.p);
    finish(1, s, OP_BI_RR, blocks, ws, true);
  }
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void axpy_scal(int n, const T *x, T *y, Scal *s, int /*which*/) {
  hipLaunchKernelGGL((k_axpy_alpha<T>), dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, x, y, s);
  CFDLA_HIP(hipGetLastError());
}

template <typename T>
void sor_colour(int count, const int *perm, const int *rp, const int *ci, const T *v, const T *piv, const T *r, T *z,
                Scal *s, bool check_done) {
  if (count == 0) return;
  hipLaunchKernelGGL((k_sor_colour<T>), dim3(blocks_for(count, kBlock * 4)), dim3(kBlock), 0, stream(), count, perm,
                     rp, ci, v, piv, r, z, s, check_done ? 1 : 0);
  CFDLA_HIP(hipGetLastError());
}
template <typename T>
void gs_colour(int count, const int *perm, const int *rp, const int *ci, const T *v, const T *b, T *x, double omega,
               Scal *s, bool check_done) {
  if (count == 0) return;
  hipLaunchKernelGGL((k_gs_colour<T>), dim3(blocks_for(count, kBlock * 4)), dim3(kBlock), 0, stream(), count, perm,
                     rp, ci, v, b, x, (T)omega, s, check_done ? 1 : 0);
  CFDLA_HIP(hipGetLastError());
}
template void gs_colour<double>(int, const int *, const int *, const int *, const double *, const double *, double *,
                                double, Scal *, bool);
template void gs_colour<float>(int, const int *, const int *, const int *, const float *, const float *, float *,
                               double, Scal *, bool);
template <typename T>
void sor_first(int n, const uint8_t *first, const T *piv, const T *r, T *z, Scal *s, bool check_done) {
  if (n == 0) return;
  hipLaunchKernelGGL((k_sor_first<T>), dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, first, piv, r,
                     z, s, check_done ? 1 : 0);
  CFDLA_HIP(hipGetLastError());
}
template void sor_first<double>(int, const uint8_t *, const double *, const double *, double *, Scal *, bool);
template void sor_first<float>(int, const uint8_t *, const float *, const float *, float *, Scal *, bool);

template void sor_colour<double>(int, const int *, const int *, const int *, const double *, const double *,
                                 const double *, double *, Scal *, bool);
template void sor_colour<float>(int, const int *, const int *, const int *, const float *, const float *,
                                const float *, float *, Scal *, bool);

void scal_reset(Scal *s, double thr2) {
  hipLaunchKernelGGL(k_scal_reset, dim3(1), dim3(64), 0, stream(), s, thr2);
  CFDLA_HIP(hipGetLastError());
}

void scal_set_done(Scal *s, int value) {
  hipLaunchKernelGGL(k_scal_set_done, dim3(1), dim3(64), 0, stream(), s, value);
  CFDLA_HIP(hipGetLastError());
}

void scal_set_final(Scal *s, double thr_final2) {
  hipLaunchKernelGGL(k_scal_set_final, dim3(1), dim3(64), 0, stream(), s, thr_final2);
  CFDLA_HIP(hipGetLastError());
}

void scal_bicg_restart(Scal *s) {
  hipLaunchKernelGGL(k_scal_bicg_restart, dim3(1), dim3(64), 0, stream(), s);
  CFDLA_HIP(hipGetLastError());
}

void coef_grad_poisson2d(int64_t nx, int64_t ny, double inv_h2, int dir_mask, const double *dir_val,
                         const double *k, const double *p, const double *lam, const double *dJdk,
                         double *dJdg) {
  hipLaunchKernelGGL(k_coef_grad_2d, dim3(blocks_for(nx * ny, kBlock * 4)), dim3(kBlock), 0, stream(), (int)nx,
                     (int)ny, inv_h2, dir_mask, dir_val[0], dir_val[1], dir_val[2], dir_val[3], k, p, lam, dJdk,
                     dJdg);
  CFDLA_HIP(hipGetLastError());
}

#define CFDLA_INSTANTIATE(T)                                                                                   \
  template void csr_apply<T>(const DevCsr &, int, const T *, const T *, const T *, double, T *, Scal *, bool, \
                             int, const T *, RedWs &, T *);                                                   \
  template void dot2<T>(int, const T *, const T *, const T *, const T *, Scal *, bool, int, RedWs &);          \
  template void jacobi_first<T>(int, const T *, double, const T *, T *, Scal *, bool);                        \
  template void cg_update<T>(int, const T *, const T *, T *, T *, const T *, T *, Scal *, RedWs &);            \
  template void cg_direction<T>(int, const T *, T *, Scal *);                                                  \
  template void copy_vec<T>(int, const T *, T *, Scal *, bool);                                                \
  template void fill_vec<T>(int, T, T *);                                                                      \
  template void dense_gemv<T>(int, const double *, const T *, T *, Scal *, bool, int);                             \
  template void bicg_p_update<T>(int, const T *, const T *, T *, Scal *);                                      \
  template void bicg_s_update<T>(int, const T *, const T *, T *, Scal *, RedWs &);                             \
  template void bicg_x_r_update<T>(int, const T *, const T *, const T *, const T *, T *, T *, Scal *, RedWs &,  \
                                   const T *);                                                               \
  template void axpy_scal<T>(int, const T *, T *, Scal *, int);

CFDLA_INSTANTIATE(double)
CFDLA_INSTANTIATE(float)

}  // namespace cfdla
