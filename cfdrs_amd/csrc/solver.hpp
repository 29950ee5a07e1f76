// Preconditioners and Krylov drivers of libcfdla (see solver.cpp).
#pragma once

#include <memory>
#include <vector>

#include "amg_setup.hpp"
#include "internal.hpp"

namespace cfdla {

// Library context: reduction workspace, device scalars, pinned host mirror,
// and pooled solver vectors.
struct Context {
  RedWs ws;
  DevBuf<Scal> scal;  // 4 slots
  Scal *host = nullptr;
  DevBuf<double> dv[12];
  DevBuf<float> fv[12];
  double *dv_at(int i, size_t n) { dv[i].ensure(n); return dv[i].p; }
  float *fv_at(int i, size_t n) { fv[i].ensure(n); return fv[i].p; }
  Scal *s(int i) { return scal.p + i; }
  Scal read(Scal *dev);  // synchronous copy of one slot
};
Context &ctx();

struct Precond {
  virtual ~Precond() = default;
  virtual int kind() const = 0;  // 0 identity, 1 Jacobi, 2 AMG, 3 multicolour SOR
  virtual void apply(const double *r, double *z, Scal *s, bool check_done) = 0;
  virtual void apply(const float *r, float *z, Scal *s, bool check_done) = 0;
  // z = M r, then r . z reduced and dot_op run on the device scalars.
  virtual void apply_dot(int n, const double *r, double *z, Scal *s, bool check_done, int dot_op);
  virtual void apply_dot(int n, const float *r, float *z, Scal *s, bool check_done, int dot_op);
  virtual const double *jacobi_dinv64() const { return nullptr; }
  virtual const float *jacobi_dinv32() const { return nullptr; }
  virtual void ensure_fp32() {}
  // FP64 operator bytes one application streams (matrices as CSR: 12 bytes
  // per nonzero plus row pointers), for the automatic precision choice.
  virtual double stream_bytes() const { return 0.0; }
};

struct JacobiPrecond : Precond {
  int n = 0;
  DevBuf<double> d64;
  DevBuf<float> d32;
  explicit JacobiPrecond(const HostCsr &a);
  int kind() const override { return 1; }
  void apply(const double *r, double *z, Scal *s, bool check_done) override;
  void apply(const float *r, float *z, Scal *s, bool check_done) override;
  double stream_bytes() const override { return 8.0 * n; }
  const double *jacobi_dinv64() const override { return d64.p; }
  const float *jacobi_dinv32() const override { return d32.p; }
};

// Multicolour SOR as a preconditioner: z = (D / omega + L_c)^-1 r, one
// forward sweep from z = 0 in colour order, L_c the couplings to rows of
// earlier colours.  Rows are coloured greedily in breadth-first order on the
// symmetrized pattern, so no two coupled rows share a colour and each colour
// is one parallel kernel (a 5-point stencil gets the red-black checkerboard).
// Replaces the lexicographic SuccessiveOverRelaxation of the momentum solve
// (momentum/solve.rs:137-147): a sequential sweep has no parallel form.
struct ColoredSorPrecond : Precond {
  int n = 0, ncolours = 0;
  std::vector<int32_t> colour_off;  // ncolours + 1 offsets into perm
  DevBuf<int32_t> perm, rp, ci;      // rows grouped by colour; the matrix pattern
  DevBuf<uint8_t> first;             // 1 on the rows of the first colour
  DevBuf<double> v64, piv64;         // values; omega / pivot (unit pivot where the diagonal is absent or zero)
  DevBuf<float> v32, piv32;
  ColoredSorPrecond(const HostCsr &a, double omega);
  int kind() const override { return 3; }
  void apply(const double *r, double *z, Scal *s, bool check_done) override;
  void apply(const float *r, float *z, Scal *s, bool check_done) override;
};

// Greedy colouring in breadth-first order on the symmetrized pattern of a:
// colour per row, colour_off (ncolours + 1) and the rows grouped by colour
// (ascending within a colour); returns ncolours.  No two coupled rows share
// a colour; a bipartite pattern (a 5-point stencil) gets two colours.
int greedy_colouring(const HostCsr &a, std::vector<int32_t> &colour, std::vector<int32_t> &colour_off,
                     std::vector<int32_t> &order);

struct DevLevel {
  DevCsr A, P, R;
  // Gauss-Seidel smoothers: colour classes of A's rows (gs_perm grouped by
  // colour, gs_off offsets); empty for the Jacobi smoother.
  std::vector<int32_t> gs_off;
  DevBuf<int32_t> gs_perm;
  DevBuf<double> b64, x64, t64, r64;
  DevBuf<float> b32, x32, t32, r32;
  // The fused pre-smoothing and restriction of this level (fused_jacobi.hpp,
  // F = [S; M]) and the output of its product, [x_l; b_{l+1}].
  DevCsr F;
  DevBuf<double> fx64;
  DevBuf<float> fx32;
  template <typename T> T *b();
  template <typename T> T *x();
  template <typename T> T *t();
  template <typename T> T *r();
  template <typename T> T *fx();
};

// Device copy of a ProductPlan.
struct DevPlan {
  int nslots = 0;
  DevBuf<int32_t> ptr, pa, pb;
};

// Numeric plans of one level's Galerkin product A_{l+1} = (R_l A_l) P_l and
// the values of the intermediate R_l A_l.
struct GalerkinPlan {
  DevPlan ra, rap;
  DevBuf<double> ra_v;
};

// The fused operator of one level (fused_jacobi.hpp): the host F = [S; M],
// the patterns of E = I - w A D (A's), T = R E and M = T E, and for the
// device refresh the per-slot plans of the two products with E's and T's
// device values.
struct FusedOp {
  int level = -1;
  int parts = 1;  // pieces per row of M in F (the coarse solve adds them)
  HostCsr F, E, T, M;
  DevPlan t, m;
  DevBuf<double> e64, t64;
  bool plans = false;
};

// C = A B on the device (galerkin.hip), bit-identical to csr_multiply: C's
// row pointers, columns and values, and the per-slot plan of the product.
// A is nr rows (int32 CSR), B has ncols columns.
void device_product(int nr, int ncols, const int32_t *arp, const int32_t *aci, const double *av, const int32_t *brp,
                    const int32_t *bci, const double *bv, DevPlan &plan, DevBuf<int32_t> &crp,
                    DevBuf<int32_t> &cci, DevBuf<double> &cv);

struct AmgPrecond : Precond {
  cfdla_amg_config cfg{};
  HostHierarchy host;
  std::vector<DevLevel> lv;
  DevBuf<double> coarse_inv;
  int coarse_n = 0;
  double setup_time = 0.0;
  bool on_device = false;
  int gather_request = -1;  // CFDLA_GATHER_* for every level's row-block products
  bool host_stale = false;  // a device refresh left newer level values than host
  std::vector<GalerkinPlan> gplan;  // per level, made on the first device refresh
  // The level before a dense coarsest one runs its two Jacobi pre-sweeps
  // from zero, residual and restriction as one product when the smoother is
  // Jacobi with two pre-sweeps (level -1: none).
  FusedOp fused;
  bool build_fused(size_t l);
  void refresh_fused_device();
  // AlgebraicMultigrid::recompute (amg.rs:179-214) on the device from
  // level-0 values already on the device (CSR order of the fine matrix).
  void update_values_device(const double *fine_values);
  // Host level values from the device after update_values_device.
  void sync_host();
  // Builds the host hierarchy only; device copies are made on first use.
  AmgPrecond(const HostCsr &a, const cfdla_amg_config &c);
  void upload();
  void ensure_device() {
    if (!on_device) upload();
  }
  void update_values(const double *values);
  int kind() const override { return 2; }
  double stream_bytes() const override {
    double b = 0.0;
    for (const HostLevel &h : host.levels)
      b += 12.0 * (double)(h.A.nnz() + h.P.nnz() + h.R.nnz()) + 4.0 * (double)(h.A.n + h.P.n + h.R.n);
    return b;
  }
  void apply(const double *r, double *z, Scal *s, bool check_done) override {
    ensure_device();
    vcycle<double>(0, r, z, s, check_done, OP_NONE);
  }
  void apply(const float *r, float *z, Scal *s, bool check_done) override {
    ensure_device();
    vcycle<float>(0, r, z, s, check_done, OP_NONE);
  }
  // The dot product rides on the last post-smoothing sweep of level 0.
  void apply_dot(int, const double *r, double *z, Scal *s, bool check_done, int dot_op) override {
    ensure_device();
    vcycle<double>(0, r, z, s, check_done, dot_op);
  }
  void apply_dot(int, const float *r, float *z, Scal *s, bool check_done, int dot_op) override {
    ensure_device();
    vcycle<float>(0, r, z, s, check_done, dot_op);
  }
  // first_ready: the caller's restriction already wrote the level's first
  // Jacobi iterate from zero into first_sweep_target(l).
  template <typename T> void vcycle(int l, const T *b, T *x, Scal *s, bool check_done, int dot_op,
                                    bool first_ready = false);
  // The V-cycle with a Gauss-Seidel family smoother (multicolour sweeps in
  // place on x; amg.rs:382-391 maps every smoother but Jacobi and symmetric
  // Gauss-Seidel to forward Gauss-Seidel).
  template <typename T> void vcycle_gs(int l, const T *b, T *x, Scal *s, bool check_done, int dot_op);
  template <typename T> void gs_sweeps(DevLevel &L, const T *b, T *x, int sweeps, Scal *s, bool check_done);
  // Buffer that level l's V-cycle writes its first Jacobi iterate to when it
  // starts with a plain sweep from zero; nullptr when it starts otherwise.
  template <typename T> T *first_sweep_target(int l);
};

struct SolveOut {
  int64_t iters = 0, outer = 0;
  double rn0 = 0, rn = 0, thr = 0;
  int term = CFDLA_TERM_MAX_ITERATIONS;
  int precision = CFDLA_FP64;  // the arithmetic the solve ran in
};

// GMRES(m) drivers (gmres.hip).
SolveOut gmres_fp64(const DevCsr &A, Precond *M, const double *b, double *x, double thr,
                    const cfdla_solver_config &cfg);
SolveOut gmres_fp32_inner(const DevCsr &A, Precond *M, const float *b, float *x, double thr, int64_t max_it,
                          const cfdla_solver_config &cfg);
int restart_width(const cfdla_solver_config &cfg);

// Solve on device FP64 vectors (b, x), dispatching on config kind/precision.
SolveOut solve_device(const DevCsr &A, Precond *M, const double *b, double *x, const cfdla_solver_config &cfg);

}  // namespace cfdla
