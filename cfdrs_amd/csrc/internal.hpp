// Shared host-side definitions of libcfdla: error handling, device buffers,
// device CSR storage, reduction scalars and kernel launcher declarations.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cfdla.h"

namespace cfdla {

struct Error : std::runtime_error {
  cfdla_status code;
  Error(cfdla_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define CFDLA_HIP(expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::cfdla::Error(e_ == hipErrorOutOfMemory ? CFDLA_ERR_OOM : CFDLA_ERR_DEVICE, \
                           std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

#define CFDLA_REQUIRE(cond, code, msg)                 \
  do {                                                 \
    if (!(cond)) throw ::cfdla::Error((code), (msg));  \
  } while (0)

// Thread-local message returned by cfdla_last_error.
void set_last_error(const std::string &msg);

// Library device/stream state.
void init_device(int dev);
void ensure_init();
int library_device();  // the device init_device selected
// Resolves, with `device` current, the device address of the allocation-free
// batched solve's static scratch array (batch.hip); init_device calls it once
// per device.
void noalloc_scratch_init(int device);
hipStream_t stream();
void set_stream(hipStream_t s);
void sync_stream();
int num_cus();

// Owning device allocation (hipMalloc / hipFree).
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    CFDLA_HIP(hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T)));
    n = count;
  }
  void ensure(size_t count) { if (n < count) alloc(count); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T *h, size_t count) {
    ensure(count);
    if (count) CFDLA_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, stream()));
  }
  // upload plus `pad` zero elements after the data
  void upload_padded(const T *h, size_t count, size_t pad) {
    ensure(count + pad);
    if (count) CFDLA_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, stream()));
    if (pad) CFDLA_HIP(hipMemsetAsync(p + count, 0, pad * sizeof(T), stream()));
  }
  void download(T *h, size_t count) const {
    if (count) CFDLA_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, stream()));
  }
};

// Host CSR with int32 column indices and int64 row pointers (setup side).
struct HostCsr {
  int64_t n = 0, m = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  std::vector<double> v;
  int64_t nnz() const { return (int64_t)ci.size(); }
};

// Device CSR: int32 row pointers and columns, FP64 values with an optional
// FP32 copy, and the smoother's inverse diagonal (0 where |a_ii| <= 1e-15,
// the skip rule of the reference Jacobi smoother).
//
// Square matrices with short rows whose column offsets |col - row| fit in
// int16 also get a sliced-ELL copy: slices of 64 consecutive rows (one
// wavefront, one lane per row), padded to the longest row of the slice with
// kSellPad offsets.  Within a slice a lane's slots are stored in groups of
// sell_group consecutive entries, group-major: slot q of lane l sits at
// slice start + (q / G) 64 G + l G + q % G.  With G = 4 every lane loads four
// values (16 B FP32, 2 x 16 B FP64) and four offsets (8 B) per group, and a
// wavefront load is one contiguous 1 KiB / 2 KiB value segment; G = 1
// (plain column-major slices) serves matrices whose rows are so short that
// rounding the slice width up to 4 would pad them past the limit.  On the
// first AMG level of the 128^3 problem (37 entries per row on average) groups
// of 4 took the FP64 Jacobi sweep from 29.2 to 28.1 us and left FP32 at
// 19.0 us; splitting each slice's groups over 4 wavefronts (kernels.hip)
// gave the rest.  (Sorting rows by length inside 1024-row
// windows cut the padding of the first AMG level of the 128^3 problem from
// 28 % to 2.7 % but made its sweeps slower, 29.2 -> 37.8 us FP64: the lanes
// of a wavefront no longer hold neighbouring rows, so its x gathers and
// epilogue accesses stop coalescing.)
//
// Every other matrix runs the row-block kernel: rows are packed in order into
// blocks of at most rb_rows<T>() rows and rb_cap<T>() nonzeros (a longer row gets
// a block of its own).  A partition per value type is stored as nrb + 1
// (first row, first nonzero) pairs.  Both types use 2048 nonzeros: FP32
// blocks of 4096 (the same LDS bytes) ran the 128^3 refined solve 4 % slower
// (5.10 against 4.88 ms), as their rows were summed by fewer threads.
constexpr int16_t kSellPad = INT16_MIN;
// Zero entries stored past nnz in the row-block streams (columns, 16-bit
// columns, values): the product loads them in aligned quads, and a quad that
// starts before nnz ends at most 3 entries past it.
constexpr size_t kRbPad = 4;
template <typename T>
constexpr int rb_cap() {
  return 2048;
}
// Rows per block: kBlock = 256, so every row is summed by at least one whole
// thread (two rows per thread in 512-row FP64 blocks ran y = A x on the
// 128^3 7-point operator in 47.8 against 42.1 us).
template <typename T>
constexpr int rb_rows() {
  return 256;
}
// Gather layouts of the row-block product (CFDLA_GATHER_QUAD / ROWS / STRIDED
// / SORTED).
constexpr int kRbQuad = 0, kRbRows = 1, kRbStrided = 2, kRbSorted = 3;
// Column-sorted blocks store each entry as col | (position in its row-ordered
// block) << kRbPosShift: columns below 2^21, positions below 2^11 (rb_cap).
constexpr int kRbPosShift = 21;
constexpr int32_t kRbColMask = (1 << kRbPosShift) - 1;
static_assert(rb_cap<double>() <= (1 << (32 - kRbPosShift)) && rb_cap<float>() <= (1 << (32 - kRbPosShift)),
              "a block position fits the packed bits");
// Diagonal (DIA) layout for square stencil matrices: every nonzero lies on
// one of nd <= kDiaMax diagonals (offsets col - row), which hold at least
// 90 % nonzeros.  Values are stored diagonal-major, dia_v[k * n + i] =
// a(i, i + off[k]) (zero where the diagonal leaves the matrix or the entry is
// absent), so lane i of a wavefront reads consecutive addresses on every
// diagonal and no column index is stored.  A symmetric matrix (equal offset
// sets on both sides, a(i, j) == a(j, i) bit for bit) stores only the main
// diagonal and the upper ones, off[0] = 0 < off[1] < ...; the lower entry
// a(i, i - d) is read as the upper one at row i - d.  General matrices store
// all diagonals, offsets ascending.
constexpr int kDiaMax = 9;
struct DiaDesc {
  int nd = 0;           // stored diagonals
  int sym = 0;          // 1: main + upper diagonals of a symmetric matrix
  int off[kDiaMax] = {};
};

struct DevCsr {
  int n = 0, m = 0;
  int64_t nnz = 0;
  int nrb64 = 0, nrb32 = 0;
  DevBuf<int32_t> rp, ci, rb64, rb32;
  // Row-block columns as 16-bit offsets from a per-block base (the block's
  // smallest column), one copy per partition (FP64 / FP32 row blocks), when
  // every block of that partition has columns spanning at most 65535: 2 bytes
  // per nonzero instead of 4.  Each partition decides on its own.
  bool c16_f64 = false, c16_f32 = false;
  // row-block gather layout per value type (kRbQuad, kRbRows, kRbStrided,
  // kRbSorted)
  int layout64 = 0, layout32 = 0;
  int gather_request = -1;  // CFDLA_GATHER_*
  DevBuf<uint16_t> c16_64, c16_32;
  DevBuf<int32_t> cbase64, cbase32;
  // kRbSorted: every block of at most rb_cap entries stored in column order
  // (ties in row order), each entry packed with its position in the
  // row-ordered block (kRbPosShift); longer rows stay in order with position
  // 0.  One copy for both partitions (they are cut alike); a device value
  // refresh finds each entry's CSR slot from its block and position.
  DevBuf<int32_t> sci;
  DevBuf<double> sv64;
  DevBuf<float> sv32;
  DevBuf<double> v64, sdinv64;
  DevBuf<float> v32, sdinv32;
  bool sell = false;
  int sell_group = 1;  // slots per lane stored together: 1 or 4
  int nslices = 0;
  int64_t sell_entries = 0;
  DevBuf<int32_t> sell_ptr;   // nslices + 1 entry offsets (multiples of 64 sell_group)
  DevBuf<int16_t> sell_delta; // col - row, kSellPad for padding
  DevBuf<double> sell_v64;
  DevBuf<float> sell_v32;
  bool dia = false;
  bool general_dia = false;  // a device update broke symmetry once: keep every diagonal from then on
  DiaDesc dia_desc;
  DevBuf<double> dia_v64;
  DevBuf<float> dia_v32;
  // Device format requested by the caller (CFDLA_STORAGE_*): -1 picks the
  // format from the matrix (diagonal, then sliced ELL, then row-block CSR);
  // a fixed one skips the others, for measuring one product kernel.
  int fmt_request = -1;
  void upload(const HostCsr &h, bool want_fp32, bool want_dinv, bool allow_sell = true);
  // Device value refresh.  v64 (CSR order) is the truth; build_maps (once,
  // from the host pattern) records for every CSR slot its position in the
  // diagonal or sliced-ELL value arrays (fmt_pos, -1: none), for symmetric
  // diagonal storage the slot of the transposed entry (fmt_mirror, -2 when
  // the pattern lacks it, -1 on the main diagonal), and each row's diagonal
  // slot (diag_slot, the last one, -1: none).  refresh_from_v64 rewrites the
  // FP32 copy, the format arrays and the smoother's inverse diagonal from
  // v64; it returns false, having changed nothing, when v64 is no longer
  // symmetric under a symmetric diagonal layout (the caller then rebuilds
  // the layout through the host with general_dia set).
  bool maps_ready = false;
  DevBuf<int32_t> fmt_pos, fmt_mirror, diag_slot;
  void build_maps(const HostCsr &h);
  bool refresh_from_v64();
  template <typename T> const T *vals() const;
  template <typename T> const T *sdinv() const;
  template <typename T> const T *sell_vals() const;
  template <typename T> const T *dia_vals() const;
  template <typename T> const int *rb() const { return sizeof(T) == 8 ? rb64.p : rb32.p; }
  template <typename T> int nrb() const { return sizeof(T) == 8 ? nrb64 : nrb32; }
  template <typename T> bool has_cols16() const { return sizeof(T) == 8 ? c16_f64 : c16_f32; }
  template <typename T> int rb_layout() const { return sizeof(T) == 8 ? layout64 : layout32; }
  template <typename T> const uint16_t *cols16() const { return sizeof(T) == 8 ? c16_64.p : c16_32.p; }
  template <typename T> const int *cbase() const { return sizeof(T) == 8 ? cbase64.p : cbase32.p; }
  // columns and values the row-block product streams
  template <typename T> const int *rb_cols() const { return rb_layout<T>() == kRbSorted ? sci.p : ci.p; }
  template <typename T> const T *rb_vals() const;
};
template <> inline const double *DevCsr::vals<double>() const { return v64.p; }
template <> inline const float *DevCsr::vals<float>() const { return v32.p; }
template <> inline const double *DevCsr::sdinv<double>() const { return sdinv64.p; }
template <> inline const float *DevCsr::sdinv<float>() const { return sdinv32.p; }
template <> inline const double *DevCsr::sell_vals<double>() const { return sell_v64.p; }
template <> inline const float *DevCsr::sell_vals<float>() const { return sell_v32.p; }
template <> inline const double *DevCsr::dia_vals<double>() const { return dia_v64.p; }
template <> inline const float *DevCsr::dia_vals<float>() const { return dia_v32.p; }
template <> inline const double *DevCsr::rb_vals<double>() const { return layout64 == kRbSorted ? sv64.p : v64.p; }
template <> inline const float *DevCsr::rb_vals<float>() const { return layout32 == kRbSorted ? sv32.p : v32.p; }


// span > 0: a block also closes before its columns would span more than
// `span` (so 16-bit column offsets fit every block whose rows each do).
void build_row_blocks(const HostCsr &h, int cap, int max_rows, std::vector<int32_t> &rb, int64_t span = 0);
// Column-sorted copy of a row-block partition (kRbSorted): packed columns
// and FP64 values; false (nothing built) when a column reaches 2^21.
bool build_sorted_blocks(const HostCsr &h, const std::vector<int32_t> &rb, std::vector<int32_t> &packed,
                         std::vector<double> &vals);
// The gather layout (kRbQuad, kRbRows or, below 2^21 columns, kRbSorted)
// whose gathers cost least
// for this partition (value size vbytes): instructions with a live lane plus
// 0.04 per distinct 128-byte line of x, over a sample of about 64 row blocks.
int rb_pick_layout(const HostCsr &h, const std::vector<int32_t> &rb, int vbytes);

// Device-resident scalars of one solve.  A reducing kernel writes block
// partials; the single-block finalize kernel launched after it sums them into
// red[] and runs a finalize op on thread 0, so the Krylov coefficients never
// travel to the host.
struct Scal {
  double red[8];
  double rz, alpha, beta, thr2, rr, pq;
  double rho, omega, rn0;
  double hn;          // GMRES: norm of the newest Arnoldi vector
  double thr_final2;  // reliable-update CG: squared FP64 stopping threshold
  double rr_true;     // reliable-update CG: ||b - A x||^2 at the latest FP64 residual replacement
  int32_t done;   // 0 running, 1 residual estimate below threshold, 2 stopped (term), 3 see BiCGSTAB,
                  // 4 FP64 residual below the threshold (reliable-update CG)
  int32_t iters;  // total iterations (Arnoldi steps for GMRES)
  int32_t term;
  int32_t kstep;  // GMRES: Arnoldi steps in the current cycle
  int32_t nupd;   // reliable-update CG: FP64 residual replacements performed
  int32_t restart_dir;  // reliable updates: the next OP_CG_BETA / OP_BI_RHO restarts the direction (beta = 0)
  int32_t nrestart;     // reliable-update CG: replacements that restarted the direction
  int32_t pad;
};

// Per-stream reduction workspace: block partials and an arrival counter that
// the last block resets, so the workspace needs no clearing between kernels.
struct RedWs {
  DevBuf<double> partials;
  DevBuf<unsigned> counter;
  void init();
};

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 8192;  // capacity of the block partials (grid of any reducing kernel)
constexpr int kGridCap = 2048;    // grid of the grid-stride vector and row-block kernels
constexpr int kMaxRed = 8;

enum FinalizeOp : int {
  OP_NONE = 0,
  OP_STORE = 1,        // red[] only
  OP_INIT_RR = 2,      // rr = red0
  OP_SET_RZ = 3,       // rz = red0
  OP_CG_ALPHA = 4,     // pq = red0; alpha = rz / pq (NONPOS if pq <= 0)
  OP_CG_RR = 5,        // rr = red0; iters++; done = 1 if rr <= thr2
  OP_CG_BETA = 6,      // rz_new = red0; beta = rz_new / rz
  OP_CG_RR_RZ = 7,     // rr = red0, rz_new = red1 (fused Jacobi update)
  OP_BI_RHO = 8,       // BiCGSTAB: rho_new = red0 -> beta (0 once after a restart_dir replacement)
  OP_BI_ALPHA = 9,     // BiCGSTAB: alpha = rho / (rhat . v)
  OP_BI_SNORM = 10,    // BiCGSTAB: ||s||^2 = red0
  OP_BI_OMEGA = 11,    // BiCGSTAB: omega = (t.s)/(t.t), red0 = t.s, red1 = t.t
  OP_BI_RR = 12,       // BiCGSTAB: rr = red0; iters++
  OP_RELIABLE = 13,    // rr_true = rr = red0 (FP64 residual); done = 4 if rr <= thr_final2, else 0;
                       // restart_dir if it exceeds 4x the recursive rr it replaces
  OP_BI_RR_RHO = 14,   // BiCGSTAB: OP_BI_RR on red0, then (still running) OP_BI_RHO on red1 = rhat . r
};

enum Epi : int { EPI_Y = 0, EPI_RESID = 1, EPI_JACOBI = 2, EPI_ADD = 3, EPI_JACOBI0 = 4, EPI_RJ0 = 5, EPI_DIR = 6 };

// ---- kernel launchers (kernels.hip) -------------------------------------
// Diagonal-format product with epilogue (dia.hip); A.dia must hold and the
// stored diagonal count must be one of the instantiated shapes (dia_supported).
bool dia_supported(const DiaDesc &d);
template <typename T, int EPI, bool DOT>
void launch_dia(const DevCsr &A, const T *x, const T *b, const T *dinv, double omega, T *y, const T *dotw,
                Scal *s, bool check_done, int op, RedWs &ws, T *y2 = nullptr);
// Single-block kernel that sums the block partials of a reducing kernel and
// runs the finalize op (kernels.hip).
void finish(int nv, Scal *s, int op, int nblocks, RedWs &ws, bool check_done);
int xcd_grid(int units, int units_per_block, int cap = kGridCap);

// Device value refresh (refresh.hip).
// v64[k] = src[perm ? perm[k] : k] for k < nnz.
void gather_values(int64_t nnz, const double *src, const int32_t *perm, double *v64);
// Symmetric diagonal layout: true when v64 is still symmetric (synchronizes).
bool symmetric_on_device(const DevCsr &d);
// FP32 copy, diagonal / sliced-ELL arrays and smoother diagonal from v64.
void scatter_formats(DevCsr &d);
// Column-sorted row-block values from v64 (kRbSorted): entry j of block q
// takes CSR slot e0 + its packed position (a row longer than rb_cap: j).
void sorted_values(DevCsr &d);
// S (into sv) and E (into ev) values of a level's fused operator from A's
// values and smoother diagonal (fused_jacobi.hpp); both follow A's slots.
void fused_values(const DevCsr &A, double w, double *sv, double *ev);
// SIMPLE pressure-correction interior rows into v64 (5 sorted slots from
// rp[idx]) and rhs; with want_max, max |imbalance| is reduced on the device
// for simple_max_read (which synchronizes).  The row kernel's block maxima
// sit in the reduction partials until the next reducing launch.
void simple_refresh(int64_t nx, int64_t ny, double dx, double dy, bool negate, const double *u, const double *v,
                    const double *p, const double *du, const double *dv, const double *rho, const int32_t *rp,
                    double *v64, double *rhs, bool want_max);
double simple_max_read();
// Rows of the variable-coefficient Poisson operator (cfdla_problem_varcoef's
// pattern, sorted columns from rp[p]) into v64 from device k, bit for bit as
// the host generator forms them; rhs (nullable) gets the Dirichlet terms.
void varcoef_refresh(int64_t nx, int64_t ny, int64_t nz, double inv_h2, const double *k, int dir_mask,
                     const double *dir_val, const int32_t *rp, double *v64, double *rhs);
void simple_corrections(int64_t nx, int64_t ny, double dx, double dy, double relax, const double *pp,
                        const double *du, const double *dv, double *p, double *u, double *v);
// Cell-centred pressure-correction right-hand side on device arrays.
void pressure_rhs(int64_t nx, int64_t ny, double dx, double dy, double coeff, const double *u, const double *v,
                  bool pin, double *rhs);
// cv[k] = sum over q in [ptr[k], ptr[k+1]) of av[pa[q]] * bv[pb[q]], in q
// order from -0.0: a sparse product replayed from its per-slot plan.
void product_numeric(int nslots, const int32_t *ptr, const int32_t *pa, const int32_t *pb, const double *av,
                     const double *bv, double *cv);
// n x n inverse of a CSR matrix by Gauss-Jordan with partial pivoting, the
// host's operation order (throws CFDLA_ERR_NUMERICAL when singular).
void dense_inverse_device(int n, const int32_t *rp, const int32_t *ci, const double *v, double *inv);

// Sparse product (diagonal, sliced-ELL or row-block CSR) with epilogue:
//   EPI_Y       y = A x
//   EPI_RESID   y = b - A x
//   EPI_JACOBI  y = x + omega * dinv .* (b - A x)
//   EPI_ADD     y = y + A x
//   EPI_JACOBI0 two Jacobi sweeps from x = 0 in one pass: with
//               x1 = omega * dinv .* b formed on the fly at every gathered
//               column (x is not read), y = x1 + omega * dinv .* (b - A x1)
//   EPI_RJ0     y = A x and y2 = omega * dinv .* y (row-block CSR only: a
//               restriction that also forms the coarse level's first Jacobi
//               iterate from zero; dinv and y2 are the coarse level's)
//   EPI_DIR     the CG direction update folded into the product (diagonal
//               storage only, fuses_direction): p = b + beta x formed on the
//               fly at every gathered column, as k_cg_direction forms it
//               (x = the previous direction, b = z, beta = s->beta; omega
//               unused); y = A p, y2 = p, and the dot reduces p . y
// When dot_op != OP_NONE, also reduces sum_i dotw_i * y_i (dotw defaults to
// x when null) and runs dot_op.  Skipped entirely when *s->done != 0 and
// check_done is set.
template <typename T>
void csr_apply(const DevCsr &A, int epi, const T *x, const T *b, const T *dinv, double omega, T *y,
               Scal *s, bool check_done, int dot_op, const T *dotw, RedWs &ws, T *y2 = nullptr);
// True when csr_apply takes EPI_DIR for A in value type T.
template <typename T>
bool fuses_direction(const DevCsr &A) {
  return A.dia && A.dia_vals<T>() != nullptr;
}

// Reductions: up to 2 dot products of (a0,b0), (a1,b1) then op.
template <typename T>
void dot2(int n, const T *a0, const T *b0, const T *a1, const T *b1, Scal *s, bool check_done, int op,
          RedWs &ws);

// x = omega * dinv .* b (first Jacobi sweep from a zero guess).
template <typename T>
void jacobi_first(int n, const T *dinv, double omega, const T *b, T *x, Scal *s, bool check_done);

// CG update: x += alpha p; r -= alpha q; then rr (and with dinv: z = dinv r,
// rz) reduced with op OP_CG_RR or OP_CG_RR_RZ.
template <typename T>
void cg_update(int n, const T *p, const T *q, T *x, T *r, const T *jdinv, T *z, Scal *s, RedWs &ws);

// p = z + beta p
template <typename T>
void cg_direction(int n, const T *z, T *p, Scal *s);

template <typename T>
void copy_vec(int n, const T *src, T *dst, Scal *s, bool check_done);
template <typename T>
void fill_vec(int n, T value, T *dst);
// dst = 0, skipped once the solve has stopped (check_done and s->done), so
// work enqueued after convergence leaves the solver's vectors alone.
template <typename T>
void zero_vec(int n, T *dst, Scal *s, bool check_done);
void f64_to_f32_scaled(int n, const double *src, double scale, float *dst);
void axpy_f32_to_f64(int n, double a, const float *x, double *y);
// Residual replacement of the reliable-update CG, skipped when check_done
// and s->done != 0:  x += (double) xc, xc = 0  /  dst = (float) src.
void fold_correction(int n, float *xc, double *x, Scal *s, bool check_done);
// cg_update<float> (no Jacobi) and fold_correction in one pass, for a step
// that replaces the residual: x64 += xc + alpha p, xc = 0, r -= alpha q, r.r.
void cg_update_fold(int n, const float *p, const float *q, float *xc, double *x64, float *r, Scal *s, RedWs &ws);
void f64_to_f32(int n, const double *src, float *dst, Scal *s, bool check_done);

// x = Ainv * b (dense row-major FP64 inverse), b/x of type T.
template <typename T>
void dense_gemv(int n, const double *ainv, const T *b, T *x, Scal *s, bool check_done, int parts = 1);

// BiCGSTAB vector kernels
template <typename T>
void bicg_p_update(int n, const T *r, const T *v, T *p, Scal *s);  // p = r + beta (p - omega v)
template <typename T>
void bicg_s_update(int n, const T *r, const T *v, T *sv, Scal *s, RedWs &ws);  // s = r - alpha v; ||s||^2
// x += alpha ph + omega sh; r = s - omega t; ||r||^2 -> OP_BI_RR.  With rhat
// non-NULL also rhat . r -> OP_BI_RR_RHO: the next iteration's rho and beta
// come out of the same finalize.
template <typename T>
void bicg_x_r_update(int n, const T *ph, const T *sh, const T *sv, const T *t, T *x, T *r, Scal *s, RedWs &ws,
                     const T *rhat = nullptr);
template <typename T>
void axpy_scal(int n, const T *x, T *y, Scal *s, int which);  // y += s->{alpha} x

// One colour of a multicolour SOR sweep: for the rows perm[0 .. count),
// z_i = piv_i * (r_i - sum_{j != i} a_ij z_j) (rows of later colours still
// hold z = 0).  Skipped when check_done and s->done != 0.
template <typename T>
void sor_colour(int count, const int *perm, const int *rp, const int *ci, const T *v, const T *piv, const T *r, T *z,
                Scal *s, bool check_done);
// The first colour from z = 0, which sees no nonzero neighbour:
// z_i = piv_i r_i where first[i], 0 elsewhere (all 0 when check_done and
// s->done != 0).
// One colour of a multicolour Gauss-Seidel sweep (rows perm[0..count)),
// x_i += omega ((b_i - sum_{j != i} a_ij x_j) / a_ii - x_i) where |a_ii| > 1e-15.
template <typename T>
void gs_colour(int count, const int *perm, const int *rp, const int *ci, const T *v, const T *b, T *x, double omega,
               Scal *s, bool check_done);
template <typename T>
void sor_first(int n, const uint8_t *first, const T *piv, const T *r, T *z, Scal *s, bool check_done);

// Scalars helpers
void scal_reset(Scal *s, double thr2);
void scal_set_done(Scal *s, int value);
void scal_bicg_restart(Scal *s);  // rho = alpha = omega = 1, done = 0
void scal_set_final(Scal *s, double thr_final2);  // reliable-update CG stopping threshold

// Coefficient gradient of the 2D variable-coefficient Poisson operator.
void coef_grad_poisson2d(int64_t nx, int64_t ny, double inv_h2, int dir_mask, const double *dir_val,
                         const double *k, const double *p, const double *lam, const double *dJdk,
                         double *dJdg);

}  // namespace cfdla
