// Batched Jacobi-PCG for thousands of independent Kirchhoff networks in one
// persistent launch.
//
// Reference per-network solve: crates/cfd-1d/src/solver/core/linear_system.rs:114-267
// (Jacobi preconditioner :270-332, convergence measured on the
// row-equilibrated residual :188-246).  The device solves the symmetric
// Laplacian itself with Jacobi-PCG (the reference runs CG on the
// row-scaled, nonsymmetric system) and stops on the same equilibrated
// residual ||D (b - A x)|| <= max(tol_abs, tol_rel ||D b||), D = 1 / row max.
//
// Scheduling: the persistent grid pulls work items from a queue (one atomic
// ticket per item, exits when the queue is drained).  A system with more
// than wave_max_rows rows is one item for the whole 256-thread block; smaller
// systems are grouped four to an item, one per wavefront, and reduce with
// shuffles only.  Each team keeps the shared vectors of its system (x, p) in
// LDS and each lane its rows' share of r, z, q and the FP32 correction in
// registers, so during the iterations global memory serves only the matrix
// and three per-row inputs.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "network.hpp"
#include "reduce.hpp"

namespace cfdla {

namespace {

constexpr int kMaxBatchRows = 2048;  // rows of one block-team system
constexpr int kWaveSlice = kMaxBatchRows / (kBlock / 64);
// A lane owns rows tl, tl + size, ... of its team's system: at most 8 (2048
// rows over 256 lanes, or 512 over 64).  Its share of r, z, q and the FP32
// correction stays in registers; only x and p, which the product gathers
// across lanes, live in LDS.
constexpr int kRowsPerLane = kMaxBatchRows / kBlock;
static_assert(kWaveSlice / 64 == kRowsPerLane, "wave and block teams own the same rows per lane");

struct BArgs {
  const int *sys_row, *rp, *ci;
  const double *v64, *b, *jd64, *eq;
  const float *v32, *jd32;
  double *x;
  BatchReport *rep;
  const int *items;
  unsigned *queue;
  int n_block_items, n_wave_items, n_wave_sys;
  double tol_abs, tol_rel, inner_tol, delta;
  int max_it, max_outer, refined, reliable;
};

struct WaveTeam {
  static constexpr int size = 64;
  __device__ static void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __device__ static double sum(double v, double *) { return wave_sum(v); }
  __device__ static void sum2(double &a, double &b, double *) {
    a = wave_sum(a);
    b = wave_sum(b);
  }
};

struct BlockTeam {
  static constexpr int size = kBlock;
  __device__ static void sync() { __syncthreads(); }
  __device__ static double sum(double v, double *sh) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
  }
  __device__ static void sum2(double &a, double &b, double *sh) {
    a = wave_sum(a);
    b = wave_sum(b);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
      sh[threadIdx.x >> 6] = a;
      sh[4 + (threadIdx.x >> 6)] = b;
    }
    __syncthreads();
    a = sh[0] + sh[1] + sh[2] + sh[3];
    b = sh[4] + sh[5] + sh[6] + sh[7];
  }
};

// The lane's rows of its team's system, read once per system: their entry
// ranges and equilibration factors (the matrix and the factors do not
// change during the solve).  A row the lane does not own gets the empty
// range at its first row's start, a valid index.
template <class Team>
struct LaneRows {
  int kb[kRowsPerLane], ke[kRowsPerLane];
  int kmax;  // the longest of the lane's rows
  double eq[kRowsPerLane];
  __device__ void load(const BArgs &A, int r0, int n, int tl) {
    kmax = 0;
    const int safe = A.rp[r0 + (tl < n ? tl : 0)];
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      const int i = tl + j * Team::size;
      kb[j] = ke[j] = safe;
      eq[j] = 0.0;
      if (i < n) {
        kb[j] = A.rp[r0 + i];
        ke[j] = A.rp[r0 + i + 1];
        eq[j] = A.eq[r0 + i];
        kmax = max(kmax, ke[j] - kb[j]);
      }
    }
  }
};

// q_j = (A v)_row over the team-local vector v in LDS, for all of the lane's
// rows at once, slot-major: step k loads the k-th entry of every row, so a
// step's loads are in flight together and a product costs the longest
// row's length in memory round trips rather than the sum of the lane's row
// lengths (row by row it was ~38 of the ~49 us per iteration of a 2,000-row
// system, the matrix being served from the Infinity Cache).  Each row still
// sums its entries in order.  (An ELL copy with a fixed, unrolled slot loop
// took 256 VGPRs, one block per CU, and ran slower: 39.7 against 28.5 ms for
// the 4,096 networks in FP64.)
template <typename T, typename V, class Team>
__device__ __forceinline__ void lane_product(const BArgs &A, const T *vals, const LaneRows<Team> &L, int r0,
                                             const V *v, T (&q)[kRowsPerLane]) {
#pragma unroll
  for (int j = 0; j < kRowsPerLane; ++j) q[j] = T(0);
  for (int k = 0; k < L.kmax; ++k) {
    T a[kRowsPerLane];
    int c[kRowsPerLane];
    bool ok[kRowsPerLane];
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      ok[j] = L.kb[j] + k < L.ke[j];
      const int e = ok[j] ? L.kb[j] + k : L.kb[j];
      a[j] = vals[e];
      c[j] = A.ci[e] - r0;
    }
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      const T t = (T)v[ok[j] ? c[j] : 0];
      if (ok[j]) q[j] += a[j] * t;
    }
  }
}

// r = b - A x over the lane's rows (x in LDS); returns sum (eq_i r_i)^2.
template <class Team>
__device__ __forceinline__ double team_resid(const BArgs &A, const LaneRows<Team> &L, int r0, int n, int tl,
                                             const double *xs, double (&r)[kRowsPerLane], double *sh) {
  double ax[kRowsPerLane];
  lane_product<double>(A, A.v64, L, r0, xs, ax);
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < kRowsPerLane; ++j) {
    const int i = tl + j * Team::size;
    r[j] = 0.0;
    if (i < n) {
      r[j] = A.b[r0 + i] - ax[j];
      const double e = L.eq[j] * r[j];
      acc += e * e;
    }
  }
  return Team::sum(acc, sh);
}

// Jacobi-PCG on A d = r, accumulating d into acc (x in LDS for FP64, the
// lane's FP32 correction for FP32), p in LDS, r/z/q in registers.  Returns
// the termination; iters counts across calls.
template <class Team, typename T, typename Acc>
__device__ __forceinline__ int team_cg(const BArgs &A, const LaneRows<Team> &L, int r0, int n, int tl, Acc &acc,
                                       T *ps, T (&r)[kRowsPerLane], const T *vals, const T *jdv, double thr2,
                                       int &iters, double *sh) {
  T z[kRowsPerLane], q[kRowsPerLane], jd[kRowsPerLane];
  double rz = 0.0;
#pragma unroll
  for (int j = 0; j < kRowsPerLane; ++j) {
    const int i = tl + j * Team::size;
    z[j] = T(0);
    jd[j] = T(0);
    if (i < n) {
      jd[j] = jdv[r0 + i];
      z[j] = jd[j] * r[j];
      ps[i] = z[j];
      rz += (double)r[j] * (double)z[j];
    }
  }
  rz = Team::sum(rz, sh);
  Team::sync();
  while (iters < A.max_it) {
    double pq = 0.0;
    lane_product<T>(A, vals, L, r0, ps, q);
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      const int i = tl + j * Team::size;
      if (i < n) pq += (double)ps[i] * (double)q[j];
    }
    pq = Team::sum(pq, sh);
    if (!(pq > 0.0) || !isfinite(pq)) return CFDLA_TERM_NONPOSITIVE_CURVATURE;
    const double alpha = rz / pq;
    double rr = 0.0, rzn = 0.0;
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      const int i = tl + j * Team::size;
      if (i < n) {
        acc.add(j, i, (T)alpha * ps[i]);
        r[j] = r[j] - (T)alpha * q[j];
        z[j] = jd[j] * r[j];
        const double e = L.eq[j] * (double)r[j];
        rr += e * e;
        rzn += (double)r[j] * (double)z[j];
      }
    }
    Team::sum2(rr, rzn, sh);
    ++iters;
    if (rr <= thr2) return CFDLA_TERM_CONVERGED;
    if (rzn == 0.0 || !isfinite(rzn)) return CFDLA_TERM_BREAKDOWN;
    const double beta = rzn / rz;
    rz = rzn;
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      const int i = tl + j * Team::size;
      if (i < n) ps[i] = z[j] + (T)beta * ps[i];
    }
    Team::sync();
  }
  return CFDLA_TERM_MAX_ITERATIONS;
}

// Accumulators of team_cg: the FP64 solution in LDS, or the lane's FP32
// correction in registers.
struct LdsAcc {
  double *x;
  __device__ void add(int, int i, double v) { x[i] += v; }
};
struct RegAcc {
  float d[kRowsPerLane];
  __device__ void add(int j, int, float v) { d[j] += v; }
};

// FP32 Jacobi-PCG with reliable FP64 residual replacement, team-local (the
// scheme of the single-system reliable-update PCG): the FP32 recurrence runs
// on the residual scaled by 1 / ||r0|| and accumulates its increments in d;
// when its equilibrated residual has dropped by delta since the latest
// replacement, or passes the threshold, d is folded into the FP64 solution
// and the residual is replaced by the FP64 one.  A replacement that comes out
// above twice the recursive residual restarts the direction and widens delta
// tenfold (up to 0.5): the cycle outran FP32 precision.
template <class Team>
__device__ __forceinline__ int team_cg_reliable(const BArgs &A, const LaneRows<Team> &L, int r0, int n, int tl,
                                                double *xs, float *ps, double (&r64)[kRowsPerLane], double rr_true,
                                                double thr2, int &iters, int &nrep, double &rr_out, double *sh) {
  const double sc = 1.0 / sqrt(rr_true);
  const double sc2 = sc * sc, thr2s = thr2 * sc2;
  double d2 = A.delta * A.delta;
  const float *vals = A.v32;
  float r[kRowsPerLane], z[kRowsPerLane], q[kRowsPerLane], d[kRowsPerLane], jd[kRowsPerLane];
#pragma unroll
  for (int j = 0; j < kRowsPerLane; ++j) {
    const int i = tl + j * Team::size;
    jd[j] = i < n ? A.jd32[r0 + i] : 0.0f;
  }
  auto load_residual = [&]() {  // r32 = sc * r64, z = D^-1 r32; returns r.z
    double rz = 0.0;
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      const int i = tl + j * Team::size;
      r[j] = (float)(r64[j] * sc);
      z[j] = i < n ? jd[j] * r[j] : 0.0f;
      rz += (double)r[j] * (double)z[j];
    }
    return Team::sum(rz, sh);
  };
  auto fold = [&]() {  // x += d / sc, d = 0, then the FP64 residual
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      const int i = tl + j * Team::size;
      if (i < n) xs[i] += (double)d[j] / sc;
      d[j] = 0.0f;
    }
    Team::sync();
    rr_out = team_resid<Team>(A, L, r0, n, tl, xs, r64, sh);
    ++nrep;
  };
#pragma unroll
  for (int j = 0; j < kRowsPerLane; ++j) d[j] = 0.0f;
  double rz = load_residual();
#pragma unroll
  for (int j = 0; j < kRowsPerLane; ++j) {
    const int i = tl + j * Team::size;
    if (i < n) ps[i] = z[j];
  }
  Team::sync();
  double ref = rr_true * sc2;
  bool restart = false;
  while (iters < A.max_it) {
    double pq = 0.0;
    lane_product<float>(A, vals, L, r0, ps, q);
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      const int i = tl + j * Team::size;
      if (i < n) pq += (double)ps[i] * (double)q[j];
    }
    pq = Team::sum(pq, sh);
    if (!(pq > 0.0) || !isfinite(pq)) {
      fold();
      return CFDLA_TERM_NONPOSITIVE_CURVATURE;
    }
    const double alpha = rz / pq;
    double rr = 0.0, rzn = 0.0;
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      const int i = tl + j * Team::size;
      if (i < n) {
        d[j] += (float)alpha * ps[i];
        r[j] = r[j] - (float)alpha * q[j];
        z[j] = jd[j] * r[j];
        const double e = L.eq[j] * (double)r[j];
        rr += e * e;
        rzn += (double)r[j] * (double)z[j];
      }
    }
    Team::sum2(rr, rzn, sh);
    ++iters;
    if (rr <= thr2s || rr <= d2 * ref) {
      Team::sync();
      fold();
      if (rr_out <= thr2) return CFDLA_TERM_CONVERGED;
      if (!isfinite(rr_out)) return CFDLA_TERM_BREAKDOWN;
      restart = rr_out * sc2 > 4.0 * rr;
      if (restart) d2 = fmin(0.25, d2 * 100.0);
      ref = rr_out * sc2;
      rzn = load_residual();
    }
    if (rzn == 0.0 || !isfinite(rzn)) {
      Team::sync();
      fold();
      return CFDLA_TERM_BREAKDOWN;
    }
    const double beta = restart ? 0.0 : rzn / rz;
    restart = false;
    rz = rzn;
#pragma unroll
    for (int j = 0; j < kRowsPerLane; ++j) {
      const int i = tl + j * Team::size;
      if (i < n) ps[i] = z[j] + (float)beta * ps[i];
    }
    Team::sync();
  }
  fold();
  return rr_out <= thr2 ? CFDLA_TERM_CONVERGED : CFDLA_TERM_MAX_ITERATIONS;
}

template <class Team>
__device__ __forceinline__ void solve_system(const BArgs &A, int sys, int tl, double *xs, double *psd, double *sh) {
  const int r0 = A.sys_row[sys], n = A.sys_row[sys + 1] - r0;
  double bb = 0.0;
  for (int i = tl; i < n; i += Team::size) {
    const int row = r0 + i;
    xs[i] = A.x[row];
    const double e = A.eq[row] * A.b[row];
    bb += e * e;
  }
  bb = Team::sum(bb, sh);
  const double thr = fmax(A.tol_abs, A.tol_rel * sqrt(bb));
  const double thr2 = thr * thr;
  Team::sync();
  LaneRows<Team> L;
  L.load(A, r0, n, tl);
  double r[kRowsPerLane];
  double rr = team_resid<Team>(A, L, r0, n, tl, xs, r, sh);
  const double rn0 = sqrt(rr);
  int iters = 0, outer = 0, term;
  if (rr <= thr2) {
    term = CFDLA_TERM_INITIAL_RESIDUAL;
  } else if (!A.refined) {
    LdsAcc acc{xs};
    while (true) {
      term = team_cg<Team, double>(A, L, r0, n, tl, acc, psd, r, A.v64, A.jd64, thr2, iters, sh);
      Team::sync();
      rr = team_resid<Team>(A, L, r0, n, tl, xs, r, sh);  // explicit FP64 residual
      if (rr <= thr2) {
        term = CFDLA_TERM_CONVERGED;
        break;
      }
      if (term != CFDLA_TERM_CONVERGED) break;
      if (iters >= A.max_it) {
        term = CFDLA_TERM_MAX_ITERATIONS;
        break;
      }
      // the explicit residual now in r restarts the recurrence
    }
  } else if (A.reliable) {
    term = team_cg_reliable<Team>(A, L, r0, n, tl, xs, reinterpret_cast<float *>(psd), r, rr, thr2, iters, outer, rr,
                                  sh);
  } else {
    float *ps = reinterpret_cast<float *>(psd);
    double rn = rn0;
    int stalls = 0;
    term = CFDLA_TERM_MAX_ITERATIONS;
    for (outer = 1; outer <= A.max_outer && iters < A.max_it; ++outer) {
      const double sc = 1.0 / rn;
      float r32[kRowsPerLane];
      RegAcc d;
#pragma unroll
      for (int j = 0; j < kRowsPerLane; ++j) {
        r32[j] = (float)(r[j] * sc);
        d.d[j] = 0.0f;
      }
      const int t = team_cg<Team, float>(A, L, r0, n, tl, d, ps, r32, A.v32, A.jd32,
                                          A.inner_tol * A.inner_tol, iters, sh);
#pragma unroll
      for (int j = 0; j < kRowsPerLane; ++j) {
        const int i = tl + j * Team::size;
        if (i < n) xs[i] += rn * (double)d.d[j];
      }
      Team::sync();
      rr = team_resid<Team>(A, L, r0, n, tl, xs, r, sh);
      const double rn2 = sqrt(rr);
      if (rr <= thr2) {
        term = CFDLA_TERM_CONVERGED;
        break;
      }
      if (!(rn2 < 0.9 * rn)) {
        if (++stalls >= 2) {
          term = (t == CFDLA_TERM_BREAKDOWN || t == CFDLA_TERM_NONPOSITIVE_CURVATURE) ? t : CFDLA_TERM_MAX_ITERATIONS;
          break;
        }
      } else {
        stalls = 0;
      }
      rn = rn2;
    }
    if (outer > A.max_outer) outer = A.max_outer;
  }
  for (int i = tl; i < n; i += Team::size) A.x[r0 + i] = xs[i];
  if (tl == 0) {
    BatchReport rep;
    rep.iterations = iters;
    rep.outer = outer;
    rep.rn0 = rn0;
    rep.rn = sqrt(rr);
    rep.thr = thr;
    rep.term = term;
    rep.pad = A.refined;  // the report's precision: 0 FP64, 1 FP32-refined
    A.rep[sys] = rep;
  }
}

__global__ __launch_bounds__(kBlock) void k_batch(BArgs A) {
  __shared__ double xs[kMaxBatchRows];
  __shared__ double ps[kMaxBatchRows];
  __shared__ double sh[8];
  __shared__ int item_sh;
  const int total = A.n_block_items + A.n_wave_items;
  while (true) {
    if (threadIdx.x == 0) item_sh = (int)atomicAdd(A.queue, 1u);
    __syncthreads();
    const int item = item_sh;
    __syncthreads();
    if (item >= total) break;  // queue drained: every block leaves
    if (item < A.n_block_items) {
      solve_system<BlockTeam>(A, A.items[item], threadIdx.x, xs, ps, sh);
    } else {
      const int w = threadIdx.x >> 6;
      const int k = (item - A.n_block_items) * (kBlock / 64) + w;
      if (k < A.n_wave_sys)
        solve_system<WaveTeam>(A, A.items[A.n_block_items + k], threadIdx.x & 63, xs + w * kWaveSlice,
                               ps + w * kWaveSlice, sh);
    }
    __syncthreads();
  }
}

__global__ void k_reset_queue(unsigned *q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *q = 0u;
}

}  // namespace

void BatchSolver::build(const NetworkInput &in) {
  nsys = in.nsys;
  assemble_networks(in, host, rhs, &plan);
  nrows = host.n;
  nedges = in.edge_off[nsys];
  isdir_h.assign(in.is_dirichlet, in.is_dirichlet + nrows);
  dirval_h.assign(in.dirichlet_value, in.dirichlet_value + nrows);
  neumann_h.clear();
  if (in.neumann) neumann_h.assign(in.neumann, in.neumann + nrows);
  device_newer = false;
  sys_row.resize((size_t)nsys + 1);
  for (int64_t s = 0; s <= nsys; ++s) sys_row[s] = (int32_t)in.node_off[s];
  for (int64_t s = 0; s < nsys; ++s)
    CFDLA_REQUIRE(in.node_off[s + 1] - in.node_off[s] <= kMaxBatchRows, CFDLA_ERR_CONFIG,
                  "batched networks are limited to 2048 nodes each; solve larger ones with cfdla_solve");
  on_device = false;
}

void BatchSolver::make_items() {
  const int wm = std::min(std::max(sched.wave_max_rows, 0), kWaveSlice);
  std::vector<int32_t> big, small;
  for (int64_t s = 0; s < nsys; ++s) (sys_row[s + 1] - sys_row[s] > wm ? big : small).push_back((int32_t)s);
  auto by_size = [&](int32_t a, int32_t b) {
    const int na = sys_row[a + 1] - sys_row[a], nb = sys_row[b + 1] - sys_row[b];
    return na != nb ? na > nb : a < b;
  };
  std::sort(big.begin(), big.end(), by_size);
  std::sort(small.begin(), small.end(), by_size);
  std::vector<int32_t> items(big);
  items.insert(items.end(), small.begin(), small.end());
  n_block_items = (int)big.size();
  n_wave_sys = (int)small.size();
  n_wave_items = (n_wave_sys + kBlock / 64 - 1) / (kBlock / 64);
  if (items.empty()) items.push_back(0);
  d_items.upload(items.data(), items.size());
  sync_stream();
}

void BatchSolver::upload() {
  std::vector<double> jd((size_t)nrows, 1.0), eq((size_t)nrows, 1.0);
  std::vector<float> jd32((size_t)nrows), v32(host.v.size());
  for (int64_t i = 0; i < nrows; ++i) {
    double d = 0.0, mx = 0.0;
    for (int64_t k = host.rp[i]; k < host.rp[i + 1]; ++k) {
      if (host.ci[k] == i) d = host.v[k];
      mx = std::max(mx, std::fabs(host.v[k]));
    }
    CFDLA_REQUIRE(std::isfinite(d), CFDLA_ERR_NUMERICAL, "network matrix has a non-finite diagonal");
    jd[i] = std::fabs(d) <= 2.220446049250313e-16 ? 1.0 : 1.0 / d;  // linear_system.rs:286-295
    eq[i] = (mx > 2.220446049250313e-16 && std::isfinite(mx)) ? 1.0 / mx : 1.0;  // linear_system.rs:161-169
    jd32[i] = (float)jd[i];
  }
  for (size_t k = 0; k < v32.size(); ++k) v32[k] = (float)host.v[k];
  std::vector<int32_t> rp32((size_t)nrows + 1);
  for (int64_t i = 0; i <= nrows; ++i) rp32[i] = (int32_t)host.rp[i];
  const size_t N = (size_t)std::max<int64_t>(1, nrows);
  d_sys_row.upload(sys_row.data(), sys_row.size());
  d_rp.upload(rp32.data(), rp32.size());
  d_ci.upload(host.ci.data(), std::max<size_t>(1, host.ci.size()));
  d_v64.upload(host.v.data(), std::max<size_t>(1, host.v.size()));
  d_v32.upload(v32.data(), std::max<size_t>(1, v32.size()));

  d_b.upload(rhs.data(), N);
  d_jd64.upload(jd.data(), N);
  d_eq.upload(eq.data(), N);
  d_jd32.upload(jd32.data(), N);
  d_rep.ensure((size_t)std::max<int64_t>(1, nsys));
  d_queue.ensure(1);
  d_slot_off.upload(plan.slot_off.data(), plan.slot_off.size());
  d_slot_src.upload(plan.slot_src.data(), std::max<size_t>(1, plan.slot_src.size()));
  d_rhs_off.upload(plan.rhs_off.data(), plan.rhs_off.size());
  d_rhs_edge.upload(plan.rhs_edge.data(), plan.rhs_edge.size());
  d_rhs_node.upload(plan.rhs_node.data(), plan.rhs_node.size());
  d_isdir.upload(isdir_h.data(), isdir_h.size());
  d_dirval.upload(dirval_h.data(), dirval_h.size());
  if (!neumann_h.empty()) d_neumann.upload(neumann_h.data(), neumann_h.size());
  d_bad.ensure(1);
  make_items();
  sync_stream();
  on_device = true;
}

namespace {
// ---- scratch of the allocation-free batched solve ---------------------------
// One 256-byte slot for the work-queue counter, then report staging for at
// most kNoAllocMaxSystems systems: 256 + 16384 x 48 bytes (768.25 KiB), held
// by every GPU the library runs on until the process ends.  Host code reaches
// it only through the address hipGetSymbolAddress returns for the current
// device (g_scratch), never through the array's name.
constexpr int64_t kNoAllocMaxSystems = 16384;
constexpr size_t kNoAllocQueueBytes = 256;
constexpr size_t kNoAllocBytes = kNoAllocQueueBytes + (size_t)kNoAllocMaxSystems * sizeof(BatchReport);
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
void *g_scratch[kScratchDevices] = {};

void check_batch_config(const cfdla_solver_config &cfg) {
  CFDLA_REQUIRE(cfg.kind == CFDLA_CG, CFDLA_ERR_CONFIG, "batched networks are solved with CG (SPD Laplacians)");
  CFDLA_REQUIRE(cfg.precision == CFDLA_FP64 || cfg.precision == CFDLA_FP32_REFINED ||
                    cfg.precision == CFDLA_PRECISION_AUTO,
                CFDLA_ERR_CONFIG,
                "unknown precision");
  CFDLA_REQUIRE(cfg.max_iter >= 1 && cfg.max_iter < INT32_MAX, CFDLA_ERR_CONFIG, "max_iter out of range");
}
}  // namespace

void noalloc_scratch_init(int device) {
  CFDLA_REQUIRE(device >= 0 && device < kScratchDevices, CFDLA_ERR_DEVICE,
                "device index beyond the allocation-free solve's table");
  if (g_scratch[device]) return;
  void *p = nullptr;
  CFDLA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  CFDLA_REQUIRE(p != nullptr, CFDLA_ERR_DEVICE, "no device address for the allocation-free solve's scratch");
  g_scratch[device] = p;
}

void BatchSolver::launch(double *x, const cfdla_solver_config &cfg, const double *rhs_dev, unsigned *queue,
                         BatchReport *rep) {
  BArgs a;
  a.sys_row = d_sys_row.p;
  a.rp = d_rp.p;
  a.ci = d_ci.p;
  a.v64 = d_v64.p;
  a.b = rhs_dev ? rhs_dev : d_b.p;
  a.jd64 = d_jd64.p;
  a.eq = d_eq.p;
  a.v32 = d_v32.p;
  a.jd32 = d_jd32.p;
  a.x = x;
  a.rep = rep;
  a.items = d_items.p;
  a.queue = queue;
  a.n_block_items = n_block_items;
  a.n_wave_items = n_wave_items;
  a.n_wave_sys = n_wave_sys;
  a.tol_abs = cfg.tol_abs;
  a.tol_rel = cfg.tol_rel;
  a.inner_tol = cfg.inner_tol > 0.0 ? cfg.inner_tol : 1e-5;
  a.max_it = (int)cfg.max_iter;
  a.max_outer = cfg.max_outer > 0 ? cfg.max_outer : 30;
  // automatic precision: FP64 (the team solves are L2- and latency-bound, and
  // their FP32 recurrences need more iterations)
  a.refined = cfg.precision == CFDLA_FP32_REFINED ? 1 : 0;
  a.reliable = cfg.reliable_interval >= 0 ? 1 : 0;  // a replacement interval does not apply per team: delta only
  a.delta = cfg.reliable_delta > 0.0 ? cfg.reliable_delta : 1e-1;
  const int grid = std::max(1, sched.blocks_per_cu) * num_cus();
  hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, stream(), queue);
  hipLaunchKernelGGL(k_batch, dim3(grid), dim3(kBlock), 0, stream(), a);
  CFDLA_HIP(hipGetLastError());
}

void BatchSolver::solve(double *x, const cfdla_solver_config &cfg, BatchReport *reports, const double *rhs_dev) {
  check_batch_config(cfg);
  if (!on_device) upload();
  if (nsys == 0) return;
  launch(x, cfg, rhs_dev, d_queue.p, d_rep.p);
  if (reports) {
    CFDLA_HIP(hipMemcpyAsync(reports, d_rep.p, (size_t)nsys * sizeof(BatchReport), hipMemcpyDeviceToHost, stream()));
  }
  sync_stream();
}

void BatchSolver::solve_noalloc(double *x, const cfdla_solver_config &cfg, BatchReport *reports,
                                const double *rhs_dev) {
  check_batch_config(cfg);
  CFDLA_REQUIRE(on_device, CFDLA_ERR_CONFIG,
                "batch is not resident on the device yet (its first cfdla_batch_solve_device uploads it)");
  CFDLA_REQUIRE(nsys <= kNoAllocMaxSystems, CFDLA_ERR_DIMENSION,
                "the allocation-free batched solve takes at most 16384 systems");
  if (nsys == 0) return;
  const int dev = library_device();
  void *p = (dev >= 0 && dev < kScratchDevices) ? g_scratch[dev] : nullptr;
  CFDLA_REQUIRE(p != nullptr, CFDLA_ERR_DEVICE, "allocation-free solve scratch not resolved on this device");
  unsigned *queue = static_cast<unsigned *>(p);
  BatchReport *rep = reinterpret_cast<BatchReport *>(static_cast<unsigned char *>(p) + kNoAllocQueueBytes);
  launch(x, cfg, rhs_dev, queue, rep);  // k_reset_queue clears the counter the previous call left
  if (reports) {
    CFDLA_HIP(hipMemcpyAsync(reports, rep, (size_t)nsys * sizeof(BatchReport), hipMemcpyDeviceToHost, stream()));
  }
  sync_stream();
}

void BatchSolver::set_values(const double *g, const NetworkInput &in) {
  NetworkInput copy = in;
  copy.g = g;
  assemble_networks(copy, host, rhs, &plan);
  device_newer = false;
  if (on_device) upload();
}

namespace {
// Device re-assembly from conductances, mirroring assemble_networks (and the
// Jacobi / equilibration factors of BatchSolver::upload) operation for
// operation: sequential sums in the recorded order, no contraction into FMA
// (the host side is built with -ffp-contract=off), so the values are the
// host's bit for bit.
__global__ __launch_bounds__(kBlock) void k_net_check(int64_t ne, const double *__restrict__ g, int *bad) {
  int local = 0;
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kBlock)
    local += (!isfinite(g[e]) || !(g[e] > 0.0)) ? 1 : 0;
  if (local) atomicAdd(bad, local);
}

__global__ void k_net_clear(int *bad) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *bad = 0;
}

__global__ __launch_bounds__(kBlock) void k_net_values(int64_t nnz, const int *__restrict__ slot_off,
                                                       const int *__restrict__ slot_src,
                                                       const double *__restrict__ g, double *__restrict__ v64,
                                                       float *__restrict__ v32) {
#pragma clang fp contract(off)
  for (int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * kBlock) {
    double v = 0.0;
    for (int t = slot_off[k]; t < slot_off[k + 1]; ++t) {
      const int src = slot_src[t];
      const double c = src == 0 ? 1.0 : (src > 0 ? g[src - 1] : -g[-src - 1]);
      v = t == slot_off[k] ? c : v + c;
    }
    v64[k] = v;
    v32[k] = (float)v;
  }
}

__global__ __launch_bounds__(kBlock) void k_net_rows(int64_t n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                     const double *__restrict__ v64, const int *__restrict__ rhs_off,
                                                     const int *__restrict__ rhs_edge,
                                                     const int *__restrict__ rhs_node, const double *__restrict__ g,
                                                     const unsigned char *__restrict__ isdir,
                                                     const double *__restrict__ dirval,
                                                     const double *__restrict__ neumann, double *__restrict__ b,
                                                     double *__restrict__ jd64, float *__restrict__ jd32,
                                                     double *__restrict__ eq) {
#pragma clang fp contract(off)
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    if (isdir[i]) {
      b[i] = dirval[i];
    } else {
      double acc = 0.0;
      for (int t = rhs_off[i]; t < rhs_off[i + 1]; ++t) acc += g[rhs_edge[t]] * dirval[rhs_node[t]];
      if (neumann) acc += neumann[i];
      b[i] = acc;
    }
    double d = 0.0, mx = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      if (ci[k] == i) d = v64[k];
      mx = fmax(mx, fabs(v64[k]));
    }
    const double j = fabs(d) <= 2.220446049250313e-16 ? 1.0 : 1.0 / d;
    jd64[i] = j;
    jd32[i] = (float)j;
    eq[i] = (mx > 2.220446049250313e-16 && isfinite(mx)) ? 1.0 / mx : 1.0;
  }
}
}  // namespace

void BatchSolver::set_values_device(const double *g) {
  if (!on_device) upload();
  if (nedges > 0) {
    hipLaunchKernelGGL(k_net_clear, dim3(1), dim3(64), 0, stream(), d_bad.p);
    hipLaunchKernelGGL(k_net_check, dim3(blocks_for(nedges, kBlock * 4)), dim3(kBlock), 0, stream(), nedges, g,
                       d_bad.p);
    int bad = 0;
    CFDLA_HIP(hipMemcpyAsync(&bad, d_bad.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
    sync_stream();
    CFDLA_REQUIRE(bad == 0, CFDLA_ERR_INVALID_ARG, "Non-positive or non-finite conductance encountered in network assembly");
  }
  const int64_t nnz = host.nnz();
  if (nnz > 0)
    hipLaunchKernelGGL(k_net_values, dim3(blocks_for(nnz, kBlock * 4)), dim3(kBlock), 0, stream(), nnz,
                       d_slot_off.p, d_slot_src.p, g, d_v64.p, d_v32.p);
  if (nrows > 0)
    hipLaunchKernelGGL(k_net_rows, dim3(blocks_for(nrows, kBlock * 4)), dim3(kBlock), 0, stream(), nrows, d_rp.p,
                       d_ci.p, d_v64.p, d_rhs_off.p, d_rhs_edge.p, d_rhs_node.p, g, d_isdir.p, d_dirval.p,
                       neumann_h.empty() ? nullptr : d_neumann.p, d_b.p, d_jd64.p, d_jd32.p, d_eq.p);
  CFDLA_HIP(hipGetLastError());
  sync_stream();
  device_newer = true;
}

void BatchSolver::download(std::vector<double> &values, std::vector<double> &rhs_out) const {
  if (!device_newer) {
    values = host.v;
    rhs_out = rhs;
    return;
  }
  values.resize(host.v.size());
  rhs_out.resize((size_t)nrows);
  if (!values.empty()) d_v64.download(values.data(), values.size());
  if (nrows > 0) d_b.download(rhs_out.data(), (size_t)nrows);
  sync_stream();
}

namespace {
// Rectangular-duct conductance on the device, the host formulas of
// network.cpp operation for operation (pi^5 is passed in from the host);
// tanh comes from the device math library, so G may differ from the host's
// in the last bits.
__device__ double po_of_dev(double eps, double pi5) {
#pragma clang fp contract(off)
  if (eps == 0.0) return 96.0;
  const double kPi = 3.14159265358979323846;
  const double t1 = (1.0 + eps) * (1.0 + eps);
  const double t2 = 1.0 - (192.0 * eps / pi5) * tanh(kPi / (2.0 * eps));
  return 96.0 / (t1 * t2);
}

__device__ double dpo_deps_dev(double eps, double pi5) {
#pragma clang fp contract(off)
  const double kPi = 3.14159265358979323846;
  const double c = 192.0 / pi5;
  const double u = kPi / (2.0 * eps);
  const double th = tanh(u);
  const double sech2 = 1.0 - th * th;
  const double T = 1.0 - c * eps * th;
  const double dT = -c * th + c * u * sech2;
  const double q = (1.0 + eps) * (1.0 + eps) * T;
  const double dq = 2.0 * (1.0 + eps) * T + (1.0 + eps) * (1.0 + eps) * dT;
  return -96.0 * dq / (q * q);
}

__global__ __launch_bounds__(kBlock) void k_rect_conductance(int64_t n, const double *__restrict__ w,
                                                             const double *__restrict__ h,
                                                             const double *__restrict__ L, double mu, double pi5,
                                                             double *__restrict__ g, double *__restrict__ dgdw) {
#pragma clang fp contract(off)
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBlock) {
    const double we = w[e], he = h[e], Le = L[e];
    // G = 1 / R, R = Po(eps) mu L / (2 A Dh^2)
    const double area = we * he;
    const double dh = 4.0 * area / (2.0 * (we + he));
    const double lo = we <= he ? we : he, hi = we >= he ? we : he;
    const double R = (po_of_dev(lo / hi, pi5) * mu * Le) / (2.0 * area * dh * dh);
    g[e] = 1.0 / R;
    if (dgdw) {
      const double s = we + he;
      const double f = 4.0 * we * we * we * he * he * he / (s * s);
      const double df = 4.0 * he * he * he * (3.0 * we * we / (s * s) - 2.0 * we * we * we / (s * s * s));
      double eps, deps;
      if (we >= he) {
        eps = he / we;
        deps = -he / (we * we);
      } else {
        eps = we / he;
        deps = 1.0 / he;
      }
      const double po = po_of_dev(eps, pi5);
      const double dpo = dpo_deps_dev(eps, pi5) * deps;
      const double Rf = po * mu * Le / (2.0 * f);
      const double dR = mu * Le / 2.0 * (dpo * f - po * df) / (f * f);
      dgdw[e] = -dR / (Rf * Rf);
    }
  }
}
}  // namespace

void rect_conductance_device(int64_t n, const double *w, const double *h, const double *L, double mu, double *g,
                             double *dgdw) {
  if (n == 0) return;
  const double pi5 = std::pow(3.14159265358979323846, 5);
  hipLaunchKernelGGL(k_rect_conductance, dim3(blocks_for(n, kBlock * 4)), dim3(kBlock), 0, stream(), n, w, h, L, mu,
                     pi5, g, dgdw);
  CFDLA_HIP(hipGetLastError());
}

namespace {
// dJ/dparam_e = (dJ/dG_e|explicit - (l_i - l_j)(p_i - p_j)) * dG_e/dparam_e,
// with l the adjoint restricted to free nodes (Dirichlet nodes carry 0).
__global__ __launch_bounds__(kBlock) void k_network_grad(int64_t ne, const int *__restrict__ ei,
                                                         const int *__restrict__ ej,
                                                         const unsigned char *__restrict__ isdir,
                                                         const double *__restrict__ p,
                                                         const double *__restrict__ lam,
                                                         const double *__restrict__ explicit_dg,
                                                         const double *__restrict__ dg_dparam,
                                                         double *__restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kBlock) {
    const int i = ei[e], j = ej[e];
    const double li = isdir[i] ? 0.0 : lam[i], lj = isdir[j] ? 0.0 : lam[j];
    double g = explicit_dg ? explicit_dg[e] : 0.0;
    g -= (li - lj) * (p[i] - p[j]);
    out[e] = dg_dparam ? g * dg_dparam[e] : g;
  }
}
}  // namespace

namespace {
// ---- edge flows from node pressures ------------------------------------------
// The reference's update_from_solution (crates/cfd-1d/src/domain/network/
// wrapper.rs:349-428): flow_e = (p_from - p_to) / (R_e + k_e |q_prev,e|),
// after it rejects an edge with R < 0, k < 0, both R and k below epsilon, or
// an effective resistance that is non-finite or not above epsilon.  The
// check pass runs first, so a rejected batch has no flow written.
__global__ __launch_bounds__(kBlock) void k_flow_check(int64_t ne, const double *__restrict__ R,
                                                       const double *__restrict__ kq, const double *__restrict__ q,
                                                       int *bad) {
  const double eps = 2.220446049250313e-16;
  int local = 0;
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kBlock) {
    const double r = R[e], k = kq ? kq[e] : 0.0;
    const double reff = kq ? r + k * fabs(q[e]) : r;
    const bool ok = r >= 0.0 && k >= 0.0 && !(fabs(r) < eps && fabs(k) < eps) && isfinite(reff) && reff > eps;
    local += ok ? 0 : 1;
  }
  if (local) atomicAdd(bad, local);
}

__global__ __launch_bounds__(kBlock) void k_flows(int64_t ne, const int32_t *__restrict__ ei,
                                                  const int32_t *__restrict__ ej, const double *__restrict__ p,
                                                  const double *__restrict__ R, const double *__restrict__ kq,
                                                  double *q) {
#pragma clang fp contract(off)
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kBlock) {
    const double reff = kq ? R[e] + kq[e] * fabs(q[e]) : R[e];
    q[e] = (p[ei[e]] - p[ej[e]]) / reff;
  }
}

// ---- outlet objective ----------------------------------------------------------
// J_s = sum over the edges e = (i, j) of network s of G_e x_e with
// x_e = m_j (p_i - p_j) + m_i (p_j - p_i): the flow into the nodes whose
// outlet mask m is set.  One block per network sums its edges in a fixed
// order (strided lanes, then a shuffle / LDS tree), so J is reproducible;
// x_e, the explicit dJ/dG_e, is written when asked for.
__global__ __launch_bounds__(kBlock) void k_outlet_edges(const int64_t *__restrict__ eoff,
                                                         const int32_t *__restrict__ ei,
                                                         const int32_t *__restrict__ ej,
                                                         const uint8_t *__restrict__ m, const double *__restrict__ p,
                                                         const double *__restrict__ g, double *__restrict__ J,
                                                         double *__restrict__ dJdg) {
  __shared__ double wsum[kBlock / 64];
  const int s = blockIdx.x;
  double acc = 0.0;
  for (int64_t e = eoff[s] + threadIdx.x; e < eoff[s + 1]; e += kBlock) {
    const int i = ei[e], j = ej[e];
    const double x = (m[j] ? p[i] - p[j] : 0.0) + (m[i] ? p[j] - p[i] : 0.0);
    if (dJdg) dJdg[e] = x;
    acc += g[e] * x;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += wsum[w];
    J[s] = t;
  }
}

// dJ/dp_i = sum over the edges at free node i (in edge order) of
// G_e (m_other - m_i); 0 on Dirichlet nodes, which the adjoint system keeps
// as identity rows.
__global__ __launch_bounds__(kBlock) void k_outlet_dp(int64_t nrows, const int64_t *__restrict__ inc_ptr,
                                                      const int32_t *__restrict__ inc_edge,
                                                      const int32_t *__restrict__ ei, const int32_t *__restrict__ ej,
                                                      const uint8_t *__restrict__ isdir,
                                                      const uint8_t *__restrict__ m, const double *__restrict__ g,
                                                      double *__restrict__ dJdp) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * kBlock) {
    double d = 0.0;
    if (!isdir[i]) {
      const double mi = m[i] ? 1.0 : 0.0;
      for (int64_t t = inc_ptr[i]; t < inc_ptr[i + 1]; ++t) {
        const int e = inc_edge[t];
        const int o = ei[e] == (int)i ? ej[e] : ei[e];
        d += g[e] * ((m[o] ? 1.0 : 0.0) - mi);
      }
    }
    dJdp[i] = d;
  }
}
}  // namespace

void network_flows(int64_t ne, const int32_t *ei, const int32_t *ej, const double *p, const double *R,
                   const double *kq, double *q, int *bad) {
  if (ne == 0) return;
  hipLaunchKernelGGL(k_net_clear, dim3(1), dim3(64), 0, stream(), bad);
  hipLaunchKernelGGL(k_flow_check, dim3(blocks_for(ne, kBlock * 4)), dim3(kBlock), 0, stream(), ne, R, kq, q, bad);
  int h = 0;
  CFDLA_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, stream()));
  sync_stream();
  CFDLA_REQUIRE(h == 0, CFDLA_ERR_CONFIG,
                std::to_string(h) + " edge(s) with a negative, zero or non-finite (effective) resistance");
  hipLaunchKernelGGL(k_flows, dim3(blocks_for(ne, kBlock * 4)), dim3(kBlock), 0, stream(), ne, ei, ej, p, R, kq, q);
  CFDLA_HIP(hipGetLastError());
}

void network_outlet(int64_t nsys, int64_t nrows, const int64_t *eoff, const int32_t *ei, const int32_t *ej,
                    const int64_t *inc_ptr, const int32_t *inc_edge, const uint8_t *isdir, const uint8_t *mask,
                    const double *p, const double *g, double *J, double *dJdp, double *dJdg) {
  if (nsys == 0) return;
  hipLaunchKernelGGL(k_outlet_edges, dim3((unsigned)nsys), dim3(kBlock), 0, stream(), eoff, ei, ej, mask, p, g, J,
                     dJdg);
  if (dJdp && nrows)
    hipLaunchKernelGGL(k_outlet_dp, dim3(blocks_for(nrows, kBlock * 4)), dim3(kBlock), 0, stream(), nrows, inc_ptr,
                       inc_edge, ei, ej, isdir, mask, g, dJdp);
  CFDLA_HIP(hipGetLastError());
}

void network_gradient(int64_t ne, const int32_t *ei, const int32_t *ej, const uint8_t *isdir, const double *p,
                      const double *lam, const double *explicit_dg, const double *dg_dparam, double *out) {
  if (ne == 0) return;
  hipLaunchKernelGGL(k_network_grad, dim3(blocks_for(ne, kBlock * 4)), dim3(kBlock), 0, stream(), ne, ei, ej, isdir,
                     p, lam, explicit_dg, dg_dparam, out);
  CFDLA_HIP(hipGetLastError());
  sync_stream();
}

}  // namespace cfdla
