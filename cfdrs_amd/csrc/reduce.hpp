// Device-side reduction helpers shared by the HIP translation units of
// libcfdla.  A reducing kernel sums its values per block (wave64 shuffle
// tree, then LDS across the 4 waves of a 256-thread block) and stores one
// partial per block with a plain store.  A single-block kernel launched next
// on the same stream sums the partials in block order and advances the
// solver scalars.  No atomics, no fences: the kernel boundary orders the
// partial stores before their loads, and the fixed summation order makes
// results bitwise reproducible run to run.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace cfdla {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Every thread of the block must call this; partial i of this block lands at
// partials[i * kMaxBlocks + blockIdx.x].
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *partials) {
  __shared__ double sh[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double t = wave_sum(v[i]);
    if (lane == 0) sh[i][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kBlock / 64; ++w) t += sh[i][w];
      partials[i * kMaxBlocks + blockIdx.x] = t;
    }
  }
}

// One block of kBlock threads sums nblocks partials per value; the totals are
// valid in thread 0.
template <int NV>
__device__ __forceinline__ void sum_partials(int nblocks, const double *partials, double (&tot)[NV]) {
  __shared__ double sh[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0.0;
  // unrolled so up to 8 loads per value are in flight (same summation order)
#pragma unroll 8
  for (int b = threadIdx.x; b < nblocks; b += kBlock) {
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] += partials[i * kMaxBlocks + b];
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double t = wave_sum(acc[i]);
    if (lane == 0) sh[i][wid] = t;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) t += sh[i][w];
    tot[i] = t;
  }
}

inline int blocks_for(int64_t work_items, int per_block) {
  int64_t b = (work_items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > kGridCap) b = kGridCap;
  return (int)b;
}

}  // namespace cfdla
