// Fused epilogues of the sparse products of libcfdla (row-block CSR,
// sliced ELL and diagonal formats): the operands an epilogue needs are
// loaded before the row sum is known, so their loads overlap the entry
// loads instead of following the reduction.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace cfdla {

// y and dotw carry no __restrict__: the self dot products (r . r with
// y = dotw = r) alias them, and the store of y[row] must precede the load of
// dotw[row].
template <typename T>
struct EpiIn {
  T b, d, x, y, w;
};

// The CG direction p = z + beta p_old, rounded once (k_cg_direction and the
// EPI_DIR gathers evaluate it the same way).
__device__ __forceinline__ double cg_dir(double beta, double pold, double z) { return __builtin_fma(beta, pold, z); }
__device__ __forceinline__ float cg_dir(float beta, float pold, float z) { return __builtin_fmaf(beta, pold, z); }

template <typename T, int EPI, bool DOT>
__device__ __forceinline__ EpiIn<T> epi_load(int row, const T *__restrict__ x, const T *__restrict__ b,
                                             const T *__restrict__ dinv, T omega, const T *y, const T *dotw) {
  EpiIn<T> e{T(0), T(0), T(0), T(0), T(0)};
  if constexpr (EPI == EPI_RESID || EPI == EPI_JACOBI || EPI == EPI_JACOBI0) e.b = b[row];
  if constexpr (EPI == EPI_JACOBI) {
    e.d = dinv[row];
    e.x = x[row];
  }
  if constexpr (EPI == EPI_JACOBI0) {
    e.d = dinv[row];
    e.x = omega * e.d * e.b;
  }
  if constexpr (EPI == EPI_RJ0) e.d = dinv[row];
  if constexpr (EPI == EPI_DIR) e.x = cg_dir(omega, x[row], b[row]);
  if constexpr (EPI == EPI_ADD) e.y = y[row];
  if constexpr (DOT && EPI != EPI_DIR) {
    if (dotw != y) e.w = dotw[row];
  }
  return e;
}

// The value an epilogue writes for row sum `sum`.
template <typename T, int EPI>
__device__ __forceinline__ T epi_out(const EpiIn<T> &e, T sum, T omega) {
  if constexpr (EPI == EPI_Y || EPI == EPI_RJ0 || EPI == EPI_DIR) return sum;
  else if constexpr (EPI == EPI_RESID) return e.b - sum;
  else if constexpr (EPI == EPI_JACOBI || EPI == EPI_JACOBI0) return e.x + omega * e.d * (e.b - sum);
  else return e.y + sum;
}

// EPI_RJ0 also writes y2 = omega * dinv .* y (the next level's first Jacobi
// iterate from zero, evaluated as k_jacobi_first does); EPI_DIR writes the
// direction p to y2 and dots it with y.
template <typename T, int EPI, bool DOT>
__device__ __forceinline__ void epi_store(const EpiIn<T> &e, int row, T sum, T omega, T *y, const T *dotw,
                                          double &dacc, T *y2 = nullptr) {
  const T out = epi_out<T, EPI>(e, sum, omega);
  if constexpr (DOT) dacc += (double)(EPI == EPI_DIR ? e.x : (dotw == y ? out : e.w)) * (double)out;
  y[row] = out;
  if constexpr (EPI == EPI_RJ0) y2[row] = omega * e.d * out;
  if constexpr (EPI == EPI_DIR) y2[row] = e.x;
}

// The gathered operand of a product: x[c], for EPI_JACOBI0 the first Jacobi
// iterate omega * dinv[c] * b[c], evaluated as k_jacobi_first does, and for
// EPI_DIR the direction b[c] + omega x[c] (omega = beta).
template <typename T, int EPI>
__device__ __forceinline__ T gather(int c, const T *__restrict__ x, const T *__restrict__ b,
                                    const T *__restrict__ dinv, T omega) {
  if constexpr (EPI == EPI_JACOBI0) return omega * dinv[c] * b[c];
  else if constexpr (EPI == EPI_DIR) return cg_dir(omega, x[c], b[c]);
  else return x[c];
}


}  // namespace cfdla
