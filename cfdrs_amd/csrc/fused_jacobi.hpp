// Two Jacobi sweeps from zero on an AMG level and the restriction of their
// residual as one product.  The V-cycle of the reference
// (crates/cfd-math/src/linear_solver/preconditioners/multigrid/amg.rs:423-503)
// starts every coarse level from x = 0; with two pre-smoothing sweeps
//   x1 = w D b,   x2 = x1 + w D (b - A x1)     =>  x2 = S b,
//   S = 2 w D - w^2 D A D,
//   b - A x2 = (I - w A D)^2 b                   =>  b_c = R (b - A x2) = M b,
//   M = R (I - w A D)^2 = (R E) E,   E = I - w A D,
// with D = diag(dinv), the smoother's inverse diagonal (0 where the Jacobi
// smoother skips a row).  Stacking S over M gives one operator F whose
// product with b writes the level's iterate and the coarse right-hand side
// in one launch, in place of a sweep, a residual and a restriction.  In
// exact arithmetic the result is that of the three products; in floating
// point it differs by rounding only.
//
// Per stored entry (i, j) of A with value a (every row of A holds its
// diagonal, so E and S share A's pattern):
//   E_ij = [i = j] - (w a) d_j
//   S_ij = [i = j] (2 w) d_i - ((w^2 d_i) a) d_j
// The host build (solver.cpp) and the device refresh (refresh.hip) both
// evaluate these functions with FP contraction off, so both give the same
// bits, and the products R E and (R E) E run in csr_multiply's order on
// either side.
#pragma once

#include <hip/hip_runtime.h>

namespace cfdla {

__host__ __device__ inline double fused_e(bool diag, double w, double a, double dj) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  const double t = (w * a) * dj;
  return diag ? 1.0 - t : -t;
}

__host__ __device__ inline double fused_s(bool diag, double w, double a, double di, double dj) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  const double t = (((w * w) * di) * a) * dj;
  return diag ? (2.0 * w) * di - t : -t;
}

}  // namespace cfdla
