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

template <typename T, int EPI, bool DOT, bool SYM, int ND>
__global__ __launch_bounds__(kBlock) void k_dia(int n, DiaDesc dd, const T *__restrict__ dv,
                                                const T *__restrict__ x, const T *__restrict__ b,
                                                const T *__restrict__ dinv, T omega, T *y, const T *dotw, Scal *s,
                                                int check_done, double *partials) {
  if (check_done && s->done) return;
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
    epi_store<T, EPI, DOT>(e, row, sum, omega, y, dotw, dacc);
  }
  if constexpr (DOT) {
    double d1[1] = {dacc};
    block_partials<1>(d1, partials);
  }
}

template <typename T, int EPI, bool DOT, bool SYM, int ND>
void launch_shape(const DevCsr &A, int blocks, const T *x, const T *b, const T *dinv, double omega, T *y,
                  const T *dotw, Scal *s, bool check_done, RedWs &ws) {
  hipLaunchKernelGGL((k_dia<T, EPI, DOT, SYM, ND>), dim3(blocks), dim3(kBlock), 0, stream(), A.n, A.dia_desc,
                     A.dia_vals<T>(), x, b, dinv, (T)omega, y, dotw, s, check_done ? 1 : 0, ws.partials.p);
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
                Scal *s, bool check_done, int op, RedWs &ws) {
  // one 256-row tile per block up to the partials capacity: blocks that
  // finish hand their CU to the next tile without a loop-carried wait
  // (26.2 us against 29.1 us with 4 tiles per block, FP64 y = A x, 128^3)
  const int blocks = xcd_grid((A.n + kBlock - 1) / kBlock, 1, kMaxBlocks);
  const DiaDesc &d = A.dia_desc;
#define CFDLA_DIA(SYM, ND) launch_shape<T, EPI, DOT, SYM, ND>(A, blocks, x, b, dinv, omega, y, dotw, s, check_done, ws)
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
                                        Scal *, bool, int, RedWs &);                                           \
  template void launch_dia<T, E, true>(const DevCsr &, const T *, const T *, const T *, double, T *, const T *,  \
                                       Scal *, bool, int, RedWs &);
#define CFDLA_DIA_ALL(T)        \
  CFDLA_DIA_INST(T, EPI_Y)      \
  CFDLA_DIA_INST(T, EPI_RESID)  \
  CFDLA_DIA_INST(T, EPI_JACOBI) \
  CFDLA_DIA_INST(T, EPI_ADD)    \
  CFDLA_DIA_INST(T, EPI_JACOBI0)
CFDLA_DIA_ALL(double)
CFDLA_DIA_ALL(float)

}  // namespace cfdla
