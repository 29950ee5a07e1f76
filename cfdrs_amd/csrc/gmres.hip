// Restarted, right-preconditioned GMRES(m) on the device.
//
// Reference: krylov::gmres_preconditioned[_with_workspace]
// (crates/cfd-math/src/linear_solver/krylov.rs:150-251) delegating to the
// un-vendored athena Gmres<B, M>; the restart width is rounded up the ladder
// 8/16/32/64/128/256 (krylov.rs:84-114) when restart_ladder is set.
//
// Arnoldi uses classical Gram-Schmidt applied twice (CGS2): each pass is one
// multi-dot over all basis vectors (one read of w, one read of each v_k)
// and one multi-axpy, instead of the j+1 dependent dots of modified
// Gram-Schmidt.  The Hessenberg column, Givens rotations and residual
// estimate are updated on the device by the last block of the norm
// reduction, so a whole restart cycle is enqueued without a host round trip.
// The preconditioner is linear, so the cycle's update is
// x += M^-1 (V_k y) with a single preconditioner application.
#include <algorithm>
#include <cmath>

#include "reduce.hpp"
#include "solver.hpp"

namespace cfdla {
namespace {

constexpr int kMdotBlocks = 256;

struct GmLayout {
  int m;
  double *H, *cs, *sn, *g, *y, *hc, *hp, *part;
  static size_t doubles(int m) { return (size_t)(m + 1) * m + 2 * m + (m + 1) + m + 2 * (m + 1) + (size_t)(m + 1) * kMdotBlocks; }
  GmLayout(double *base, int m_) : m(m_) {
    H = base;
    cs = H + (size_t)(m + 1) * m;
    sn = cs + m;
    g = sn + m;
    y = g + (m + 1);
    hc = y + m;
    hp = hc + (m + 1);
    part = hp + (m + 1);
  }
};

// part[k * NB + bx] = sum over the block's slice of V_k . w, k = blockIdx.y
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mdot(int n, const T *__restrict__ V, size_t ld, const T *__restrict__ w,
                                                 double *__restrict__ part, Scal *s) {
  if (s->done) return;
  const int k = blockIdx.y;
  const T *v = V + (size_t)k * ld;
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) acc += (double)v[i] * (double)w[i];
  __shared__ double sh[kBlock / 64];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) t += sh[q];
    part[(size_t)k * gridDim.x + blockIdx.x] = t;
  }
}

// hp[k] = sum_b part[k * nb + b]; hc[k] (+)= hp[k]
__global__ __launch_bounds__(kBlock) void k_mdot_fin(int nb, const double *__restrict__ part, double *__restrict__ hp,
                                                     double *__restrict__ hc, int first, Scal *s) {
  if (s->done) return;
  const int k = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) acc += part[(size_t)k * nb + b];
  __shared__ double sh[kBlock / 64];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) t += sh[q];
    hp[k] = t;
    hc[k] = first ? t : hc[k] + t;
  }
}

// w -= sum_{k<=j} hp[k] V_k; with NORM, also the block partials of ||w||^2
// of the updated w (the last Gram-Schmidt pass: the norm needs no pass of
// its own).  (Summing the multi-dot partials here instead of in k_mdot_fin,
// with 64 partials per vector, was slower: 22.1 -> 23.2 ms for the 512^2
// momentum GMRES.)
template <typename T, bool NORM>
__global__ __launch_bounds__(kBlock) void k_maxpy(int n, int jp1, const T *__restrict__ V, size_t ld,
                                                  const double *__restrict__ hp, T *__restrict__ w, Scal *s,
                                                  double *partials) {
  if (s->done) return;
  __shared__ double h[257];
  for (int k = threadIdx.x; k < jp1; k += kBlock) h[k] = hp[k];
  __syncthreads();
  double v[1] = {0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double acc = (double)w[i];
    for (int k = 0; k < jp1; ++k) acc -= h[k] * (double)V[(size_t)k * ld + i];
    const T wi = (T)acc;
    w[i] = wi;
    if (NORM) v[0] += (double)wi * (double)wi;
  }
  if (NORM) block_partials<1>(v, partials);
}

// One block: ||w||, Hessenberg column j, Givens rotation, residual estimate.
// The column and the earlier rotations are staged in LDS by the whole block,
// so thread 0's rotation chain waits on arithmetic, not on global memory.
__global__ __launch_bounds__(kBlock) void k_gm_arnoldi(int j, int nblocks, const double *partials, Scal *s,
                                                       GmLayout L) {
  if (s->done) return;
  __shared__ double col[258], cs[256], sn[256];
  for (int k = threadIdx.x; k <= j; k += kBlock) col[k] = L.hc[k];
  for (int k = threadIdx.x; k < j; k += kBlock) {
    cs[k] = L.cs[k];
    sn[k] = L.sn[k];
  }
  double tot[1];
  sum_partials<1>(nblocks, partials, tot);  // synchronizes the block
  const int m = L.m;
  __shared__ int broke;
  if (threadIdx.x == 0) {
    const double hn = sqrt(tot[0]);
    s->hn = hn;
    col[j + 1] = hn;
    for (int k = 0; k < j; ++k) {
      const double t0 = col[k], t1 = col[k + 1];
      col[k] = cs[k] * t0 + sn[k] * t1;
      col[k + 1] = -sn[k] * t0 + cs[k] * t1;
    }
    const double h0 = col[j], h1 = col[j + 1];
    const double den = hypot(h0, h1);
    broke = den == 0.0 || !isfinite(den);
    if (broke) {
      s->done = 4;  // breakdown: the cycle ends with kstep = j
    } else {
      const double c = h0 / den, sj = h1 / den;
      L.cs[j] = c;
      L.sn[j] = sj;
      col[j] = den;
      col[j + 1] = 0.0;
      const double gj = L.g[j];
      L.g[j + 1] = -sj * gj;
      L.g[j] = c * gj;
      s->kstep = j + 1;
      s->iters += 1;
      s->alpha = hn > 0.0 ? 1.0 / hn : 0.0;
      if (fabs(L.g[j + 1]) <= sqrt(s->thr2) || hn == 0.0) s->done = 1;
    }
  }
  __syncthreads();
  // the column as the host-visible Hessenberg H(0:j+1, j); on breakdown the
  // rotated entries up to j + 1, as before the rotation that failed
  for (int k = threadIdx.x; k <= j + 1; k += kBlock) L.H[(size_t)k * m + j] = col[k];
  (void)broke;
}

// v_{j+1} = w / hn (w already lives in V_{j+1})
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_scale(int n, T *__restrict__ w, Scal *s) {
  if (s->done) return;
  const double a = s->alpha;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) w[i] = (T)((double)w[i] * a);
}

// Back substitution H(0:k,0:k) y = g(0:k), k = kstep.
__global__ void k_gm_solve(Scal *s, GmLayout L) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int k = s->kstep, m = L.m;
  for (int i = k - 1; i >= 0; --i) {
    double t = L.g[i];
    for (int q = i + 1; q < k; ++q) t -= L.H[(size_t)i * m + q] * L.y[q];
    L.y[i] = t / L.H[(size_t)i * m + i];
  }
}

// u = sum_{k < kstep} y_k V_k
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_combine(int n, const T *__restrict__ V, size_t ld, Scal *s, GmLayout L,
                                                       T *__restrict__ u) {
  const int k = s->kstep;
  __shared__ double y[256];
  for (int q = threadIdx.x; q < k; q += kBlock) y[q] = L.y[q];
  __syncthreads();
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    double acc = 0.0;
    for (int q = 0; q < k; ++q) acc += y[q] * (double)V[(size_t)q * ld + i];
    u[i] = (T)acc;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy1(int n, const T *__restrict__ z, T *__restrict__ x) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] += z[i];
}

// v_0 = r / beta, g = beta e_1, new cycle.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_restart(int n, const T *__restrict__ r, double beta,
                                                       T *__restrict__ v0, Scal *s, GmLayout L) {
  const double a = 1.0 / beta;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v0[i] = (T)((double)r[i] * a);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int q = 0; q <= L.m; ++q) L.g[q] = 0.0;
    L.g[0] = beta;
    s->kstep = 0;
    s->done = 0;
  }
}

template <typename T> T *vec(int slot, size_t n);
template <> double *vec<double>(int slot, size_t n) { return ctx().dv_at(slot, n); }
template <> float *vec<float>(int slot, size_t n) { return ctx().fv_at(slot, n); }

struct GmWork {
  DevBuf<double> V64, st;
  DevBuf<float> V32;
};
GmWork &gmwork() {
  static GmWork w;
  return w;
}

template <typename T> T *basis(size_t count);
template <> double *basis<double>(size_t count) { gmwork().V64.ensure(count); return gmwork().V64.p; }
template <> float *basis<float>(size_t count) { gmwork().V32.ensure(count); return gmwork().V32.p; }

template <typename T>
double residual_norm(const DevCsr &A, const T *b, const T *x, T *r, Scal *s) {
  csr_apply<T>(A, EPI_RESID, x, b, nullptr, 0.0, r, s, false, OP_INIT_RR, r, ctx().ws);
  return std::sqrt(ctx().read(s).rr);
}

template <typename T>
SolveOut gmres(const DevCsr &A, Precond *M, const T *b, T *x, double thr, int64_t max_it, int m, bool confirm) {
  Context &c = ctx();
  RedWs &ws = c.ws;
  const int n = A.n;
  const size_t ld = (size_t)n;
  T *V = basis<T>((size_t)(m + 1) * ld);
  gmwork().st.ensure(GmLayout::doubles(m));
  GmLayout L(gmwork().st.p, m);
  T *r = vec<T>(0, n), *z = vec<T>(1, n), *u = vec<T>(2, n);
  Scal *s = c.s(0);
  SolveOut o;
  o.thr = thr;
  scal_reset(s, thr * thr);
  double beta = residual_norm<T>(A, b, x, r, s);
  o.rn0 = o.rn = beta;
  if (beta <= thr) {
    o.term = CFDLA_TERM_INITIAL_RESIDUAL;
    return o;
  }
  const int eb = blocks_for(n, kBlock * 4);
  const int nb = std::min(kMdotBlocks, blocks_for(n, kBlock * 4));
  int64_t enq = 0;
  while (true) {
    hipLaunchKernelGGL((k_gm_restart<T>), dim3(eb), dim3(kBlock), 0, stream(), n, r, beta, V, s, L);
    int steps = (int)std::min<int64_t>(m, max_it - enq);
    for (int j = 0; j < steps; ++j) {
      T *vj = V + (size_t)j * ld, *w = V + (size_t)(j + 1) * ld;
      M->apply(vj, z, s, true);
      csr_apply<T>(A, EPI_Y, z, nullptr, nullptr, 0.0, w, s, true, OP_NONE, nullptr, ws);
      for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL((k_mdot<T>), dim3(nb, j + 1), dim3(kBlock), 0, stream(), n, V, ld, w, L.part, s);
        hipLaunchKernelGGL(k_mdot_fin, dim3(j + 1), dim3(kBlock), 0, stream(), nb, L.part, L.hp, L.hc,
                           pass == 0 ? 1 : 0, s);
        if (pass == 0)
          hipLaunchKernelGGL((k_maxpy<T, false>), dim3(eb), dim3(kBlock), 0, stream(), n, j + 1, V, ld, L.hp, w, s,
                             nullptr);
        else
          hipLaunchKernelGGL((k_maxpy<T, true>), dim3(eb), dim3(kBlock), 0, stream(), n, j + 1, V, ld, L.hp, w, s,
                             ws.partials.p);
      }
      hipLaunchKernelGGL(k_gm_arnoldi, dim3(1), dim3(kBlock), 0, stream(), j, eb, ws.partials.p, s, L);
      hipLaunchKernelGGL((k_gm_scale<T>), dim3(eb), dim3(kBlock), 0, stream(), n, w, s);
    }
    enq += steps;
    CFDLA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gm_solve, dim3(1), dim3(64), 0, stream(), s, L);
    hipLaunchKernelGGL((k_gm_combine<T>), dim3(eb), dim3(kBlock), 0, stream(), n, V, ld, s, L, u);
    M->apply(u, z, s, false);
    hipLaunchKernelGGL((k_axpy1<T>), dim3(eb), dim3(kBlock), 0, stream(), n, z, x);
    CFDLA_HIP(hipGetLastError());
    Scal h = c.read(s);
    o.iters = h.iters;
    beta = residual_norm<T>(A, b, x, r, s);
    o.rn = beta;
    if (beta <= thr || (!confirm && h.done == 1)) {
      o.term = CFDLA_TERM_CONVERGED;
      return o;
    }
    if (h.done == 4 && h.kstep == 0) {
      o.term = CFDLA_TERM_BREAKDOWN;
      return o;
    }
    if (enq >= max_it) {
      o.term = CFDLA_TERM_MAX_ITERATIONS;
      return o;
    }
    if (!std::isfinite(beta)) {
      o.term = CFDLA_TERM_BREAKDOWN;
      return o;
    }
  }
}

}  // namespace

int restart_width(const cfdla_solver_config &cfg) {
  int req = cfg.restart > 0 ? cfg.restart : 30;
  if (cfg.restart_ladder) {
    if (req <= 8) return 8;
    if (req <= 16) return 16;
    if (req <= 32) return 32;
    if (req <= 64) return 64;
    if (req <= 128) return 128;
    return 256;
  }
  return std::min(req, 256);
}

SolveOut gmres_fp64(const DevCsr &A, Precond *M, const double *b, double *x, double thr,
                    const cfdla_solver_config &cfg) {
  return gmres<double>(A, M, b, x, thr, cfg.max_iter, restart_width(cfg), true);
}

SolveOut gmres_fp32_inner(const DevCsr &A, Precond *M, const float *b, float *x, double thr, int64_t max_it,
                          const cfdla_solver_config &cfg) {
  return gmres<float>(A, M, b, x, thr, max_it, restart_width(cfg), false);
}

}  // namespace cfdla
