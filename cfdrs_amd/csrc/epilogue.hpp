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
  if constexpr (EPI == EPI_ADD) e.y = y[row];
  if constexpr (DOT) {
    if (dotw != y) e.w = dotw[row];
  }
  return e;
}

// The value an epilogue writes for row sum `sum`.
template <typename T, int EPI>
__device__ __forceinline__ T epi_out(const EpiIn<T> &e, T sum, T omega) {
  if constexpr (EPI == EPI_Y || EPI == EPI_RJ0) return sum;
  else if constexpr (EPI == EPI_RESID) return e.b - sum;
  else if constexpr (EPI == EPI_JACOBI || EPI == EPI_JACOBI0) return e.x + omega * e.d * (e.b - sum);
  else return e.y + sum;
}

// EPI_RJ0 also writes y2 = omega * dinv .* y (the next level's first Jacobi
// iterate from zero, evaluated as k_jacobi_first does).
template <typename T, int EPI, bool DOT>
__device__ __forceinline__ void epi_store(const EpiIn<T> &e, int row, T sum, T omega, T *y, const T *dotw,
                                          double &dacc, T *y2 = nullptr) {
  const T out = epi_out<T, EPI>(e, sum, omega);
  if constexpr (DOT) dacc += (double)(dotw == y ? out : e.w) * (double)out;
  y[row] = out;
  if constexpr (EPI == EPI_RJ0) y2[row] = omega * e.d * out;
}

// The gathered operand of a product: x[c], or for EPI_JACOBI0 the first
// Jacobi iterate omega * dinv[c] * b[c], evaluated as k_jacobi_first does.
template <typename T, int EPI>
__device__ __forceinline__ T gather(int c, const T *__restrict__ x, const T *__restrict__ b,
                                    const T *__restrict__ dinv, T omega) {
  if constexpr (EPI == EPI_JACOBI0) return omega * dinv[c] * b[c];
  else return x[c];
}


}  // namespace cfdla
