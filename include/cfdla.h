/*
 * cfdla.h — C ABI of the MI355X-native sparse linear-solve library
 * (libcfdla.so): CSR SpMV, PCG / BiCGSTAB / GMRES(m), Jacobi and
 * aggregation-AMG preconditioners, FP32-inner / FP64-refined solves,
 * adjoint solves with coefficient gradients, and batched network solves.
 *
 * Plain C: opaque handles, plain pointers and sizes, status codes.  No
 * exceptions cross this boundary.  Every entry point returns a
 * cfdla_status; on failure cfdla_last_error() returns a thread-local
 * message.
 *
 * Each declaration names the reference interface it replaces (paths are
 * relative to the CFDrs repository root).
 *
 * Threading: handles are not thread-safe.  Calls on one handle must be
 * serialized by the caller (the reference serializes AMG application with a
 * Mutex, crates/cfd-math/src/linear_solver/preconditioners/multigrid/amg.rs:519-552).
 * All device work is issued on the library stream (cfdla_set_stream).
 */
#ifndef CFDLA_H
#define CFDLA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Mirror cfd_core::error::Error variants
 * (crates/cfd-core/src/error.rs:13-45): InvalidInput/InvalidConfiguration ->
 * INVALID_ARG/CONFIG, DimensionMismatch -> DIMENSION, Numerical -> NUMERICAL,
 * GpuCompute -> DEVICE.  A solve that stalls or breaks down still returns
 * CFDLA_OK (value-semantic termination, krylov.rs:375-400). */
typedef enum {
  CFDLA_OK = 0,
  CFDLA_ERR_INVALID_ARG = 1,
  CFDLA_ERR_DIMENSION = 2,
  CFDLA_ERR_CONFIG = 3,
  CFDLA_ERR_NUMERICAL = 4,
  CFDLA_ERR_DEVICE = 5,
  CFDLA_ERR_OOM = 6
} cfdla_status;

/* Termination, mirrors athena Termination as classified by krylov::interpret
 * (crates/cfd-math/src/linear_solver/krylov.rs:375-400):
 * CONVERGED / INITIAL_RESIDUAL -> SolveOutcome::Converged,
 * MAX_ITERATIONS -> Stalled, BREAKDOWN / NONPOSITIVE_CURVATURE -> BrokenDown. */
typedef enum {
  CFDLA_TERM_CONVERGED = 0,
  CFDLA_TERM_INITIAL_RESIDUAL = 1,
  CFDLA_TERM_MAX_ITERATIONS = 2,
  CFDLA_TERM_BREAKDOWN = 3,
  CFDLA_TERM_NONPOSITIVE_CURVATURE = 4
} cfdla_termination;

/* krylov::SolverKind (krylov.rs:534-560). */
typedef enum { CFDLA_CG = 0, CFDLA_BICGSTAB = 1, CFDLA_GMRES = 2 } cfdla_solver_kind;

/* Arithmetic of the solve.  FP32_REFINED: FP32 SpMV, smoother and inner
 * Krylov iterations inside FP64 defect correction (outer residual, update
 * and coarsest AMG solve in FP64).  PRECISION_AUTO picks one per solve: the
 * refined form when an iteration streams more than 256 MiB of FP64
 * operator bytes (the matrix plus the preconditioner's matrices: beyond the
 * Infinity Cache, where halving the value bytes pays), FP64 otherwise
 * (cache-resident systems, whose kernels are latency-bound, gain nothing
 * from FP32 and pay for the residual replacements); batched network solves
 * take FP64.  The report's precision field says which ran.  No counterpart
 * in the reference. */
typedef enum { CFDLA_FP64 = 0, CFDLA_FP32_REFINED = 1, CFDLA_PRECISION_AUTO = 2 } cfdla_precision;

/* Value type of device vectors passed to *_device entry points. */
typedef enum { CFDLA_F64 = 0, CFDLA_F32 = 1 } cfdla_dtype;

typedef struct cfdla_matrix_s *cfdla_matrix;
typedef struct cfdla_precond_s *cfdla_precond;
typedef struct cfdla_batch_s *cfdla_batch;

/* Solver configuration.  Replaces IterativeSolverConfig (ext leto-ops,
 * used at krylov.rs:123-138) plus the restart argument of
 * gmres_preconditioned (krylov.rs:150-160).  Stopping rule:
 * ||b - A x||_2 <= max(tol_abs, tol_rel * ||b||_2), confirmed by an explicit
 * FP64 true residual before CONVERGED is reported. */
typedef struct {
  int32_t kind;            /* cfdla_solver_kind */
  int32_t restart;         /* GMRES restart width as requested */
  int32_t restart_ladder;  /* 1: round restart up 8/16/32/64/128/256 as krylov.rs:84-114 */
  int32_t precision;       /* cfdla_precision */
  double tol_abs;
  double tol_rel;
  int64_t max_iter;        /* total (inner) iteration budget */
  double inner_tol;        /* FP32_REFINED restarted refinement: relative tolerance of each inner solve */
  int32_t max_outer;       /* FP32_REFINED restarted refinement: refinement step budget */
  int32_t check_every;     /* host convergence poll interval in iterations (0 = auto) */
  int32_t reliable_interval; /* FP32_REFINED CG and BiCGSTAB run one uninterrupted FP32 recurrence whose
                              * residual is replaced by the FP64 residual at convergence and: 0 = when it
                              * has dropped by reliable_delta since the last replacement; k > 0 (CG only)
                              * = every k iterations.  -1 = restarted refinement (inner_tol, max_outer),
                              * which GMRES always uses */
  int32_t reserved;
  double reliable_delta;   /* residual reduction between replacements for reliable_interval 0 (<= 0: 1e-1) */
} cfdla_solver_config;

/* athena SolveReport {iterations, initial_residual_norm,
 * final_residual_norm, threshold, termination} (krylov.rs:506-516). */
typedef struct {
  int64_t iterations;
  int64_t outer_iterations;
  double initial_residual_norm;
  double final_residual_norm;
  double threshold;
  int32_t termination;  /* cfdla_termination */
  int32_t precision;    /* cfdla_precision the solve ran in (FP64 or FP32_REFINED) */
} cfdla_report;

/* AMGConfig (crates/cfd-math/src/linear_solver/preconditioners/multigrid/mod.rs:219-260).
 * Enumerations: coarsening 0 = RugeStueben, 1 = Aggregation, 2 = Hybrid;
 * interpolation 0 = Classical, 1 = Direct, 2 = SmoothedDirect (prolongator
 * smoothing P = (I - smoothing_omega D^-1 A) P_direct, no reference
 * counterpart); smoother 0 = GaussSeidel, 1 = SymmetricGaussSeidel,
 * 2 = Jacobi (the device V-cycle runs Jacobi only);
 * coarse_solver 0 = reference sweeps (pre + 10 smoother sweeps, amg.rs:501),
 * 1 = FP64 dense direct solve. */
typedef struct {
  int32_t max_levels;
  int32_t min_coarse_size;
  int32_t coarsening;
  int32_t interpolation;
  int32_t smoother;
  int32_t pre_smooth_iterations;
  int32_t post_smooth_iterations;
  int32_t coarse_solver;
  double relaxation_factor;
  double strength_threshold;
  double smoothing_omega;
  int32_t max_aggregate_size;
  int32_t max_interpolation_points;
  /* 1: form the Galerkin products A_{l+1} = (R_l A_l) P_l on the device
   * (bit-identical to the host products; their plans also serve
   * cfdla_amg_update_values_device).  No reference counterpart. */
  int32_t galerkin_on_device;
  /* 1: keep the two Jacobi pre-sweeps, the residual and the restriction of
   * the level above a dense coarsest level as three products.  0 (default):
   * when the smoother is Jacobi with two pre-sweeps and the coarsest level
   * is solved directly, that level runs them as one product with
   * F = [2wD - w^2 DAD; R (I - wAD)^2] (D the smoother's inverse diagonal),
   * the same operator in exact arithmetic.  No reference counterpart. */
  int32_t separate_coarse_restriction;
} cfdla_amg_config;

#define CFDLA_MAX_LEVELS 32

/* AMGStatistics (multigrid/mod.rs:366-395). */
typedef struct {
  int32_t num_levels;
  /* the level whose pre-sweeps and restriction run as one product (see
   * separate_coarse_restriction); -1 none or not yet on the device */
  int32_t fused_level;
  int64_t level_rows[CFDLA_MAX_LEVELS];
  int64_t level_nnz[CFDLA_MAX_LEVELS];
  double operator_complexity;
  double grid_complexity;
  double setup_time;
} cfdla_amg_statistics;

/* ---- library ----------------------------------------------------------- */

/* Thread-local message of the last failing call ("" if none). */
const char *cfdla_last_error(void);
/* Library version string. */
const char *cfdla_version(void);
/* Select the HIP device and create the library stream.  Called implicitly
 * by the first call that needs the device.  Replaces ComputeBackend::Gpu
 * selection (crates/cfd-core/src/compute/backend.rs:81-92). */
cfdla_status cfdla_init(int32_t device);
/* Issue device work on an existing hipStream_t.  NULL selects the device's
 * null stream (the default stream of PyTorch); CFDLA_LIBRARY_STREAM selects
 * the library's own non-blocking stream (the initial choice). */
#define CFDLA_LIBRARY_STREAM ((void *)(intptr_t)-1)
cfdla_status cfdla_set_stream(void *hip_stream);
/* Block until all work on the library stream has finished. */
cfdla_status cfdla_synchronize(void);

/* ---- matrices: SparseMatrix = leto_ops::CsrMatrix (crates/cfd-math/src/sparse/mod.rs:53),
 *      CsrMatrix::from_parts(values, col_indices, row_ptr, nrows, ncols) --------- */

/* Copies host CSR arrays (sorted or unsorted columns, int64 indices) to the
 * device.  The caller keeps ownership of its arrays. */
cfdla_status cfdla_matrix_create(int64_t nrows, int64_t ncols, const int64_t *row_ptr,
                                 const int64_t *col_indices, const double *values,
                                 cfdla_matrix *out);
/* Same, from device int32 row_ptr / col_indices and f64 values (copied). */
cfdla_status cfdla_matrix_create_device(int64_t nrows, int64_t ncols, int64_t nnz,
                                        const int32_t *row_ptr_dev, const int32_t *col_dev,
                                        const double *values_dev, cfdla_matrix *out);
/* Replace the values of the same pattern (host array of nnz doubles). */
cfdla_status cfdla_matrix_update_values(cfdla_matrix m, const double *values);
/* Same from a device array of nnz doubles in creation order: the FP64 and
 * FP32 device copies and the diagonal / sliced-ELL arrays are rewritten on
 * the device, without a host copy (the host copy is refreshed only when a
 * host-side consumer needs it).  A symmetric diagonal layout whose new
 * values are not symmetric is rebuilt once as a general one.  Device side of
 * refilling CsrMatrix::values_mut on a fixed pattern
 * (crates/cfd-2d/src/solvers/simple/pressure.rs:244-260). */
cfdla_status cfdla_matrix_update_values_device(cfdla_matrix m, const double *values_dev);
/* Current values into a host array of nnz doubles, in creation order. */
cfdla_status cfdla_matrix_get_values(cfdla_matrix m, double *values);
cfdla_status cfdla_matrix_info(cfdla_matrix m, int64_t *nrows, int64_t *ncols, int64_t *nnz);
/* Device storage chosen for the matrix (uploads it if needed): *format is
 * CFDLA_STORAGE_DIA (stencil matrices: diagonal-major values, no column
 * indices; symmetric ones store the main and upper diagonals only;
 * *stored_entries = stored diagonals x rows, *units = stored diagonals),
 * CFDLA_STORAGE_SELL64 (64-row slices, int16 column offsets; *stored_entries
 * counts the padded slots, *units = slices) or CFDLA_STORAGE_CSR_ROWBLOCK
 * (int32 columns; *stored_entries = nnz, *units = row blocks).  No
 * counterpart in the reference, whose CSR lives in host memory; used to state
 * the bytes a product moves. */
#define CFDLA_STORAGE_CSR_ROWBLOCK 0
#define CFDLA_STORAGE_SELL64 1
#define CFDLA_STORAGE_DIA 2
cfdla_status cfdla_matrix_storage(cfdla_matrix m, int32_t *format, int64_t *stored_entries, int64_t *units);
/* Restrict the device format of m to one CFDLA_STORAGE_* value, or
 * CFDLA_STORAGE_AUTO (the default) to let the matrix pick.  A format the
 * matrix cannot take (DIA for a non-stencil matrix, SELL64 for rows too long
 * or offsets beyond int16) falls back to CSR_ROWBLOCK; cfdla_matrix_storage
 * reports the result.  No counterpart in the reference: it exists to time
 * one product kernel (for example the row-block CSR product on a stencil
 * matrix that would otherwise use diagonal storage). */
#define CFDLA_STORAGE_AUTO (-1)
/* row-block CSR that keeps 32-bit columns even where 16-bit block offsets
 * would fit (the CSR layout the reference stores, int32 indices) */
#define CFDLA_STORAGE_CSR_INT32 3
cfdla_status cfdla_matrix_set_storage(cfdla_matrix m, int32_t format);
/* Gather layout of the row-block CSR product.  CFDLA_GATHER_QUAD: every
 * thread gathers x for 4 consecutive stored entries of its block as it loads
 * them; CFDLA_GATHER_ROWS: the block's columns and values are staged in LDS
 * and each row's lanes gather x while walking the row, so lanes of a
 * wavefront gather for consecutive rows.  CFDLA_GATHER_AUTO (the default)
 * picks per matrix and value type, at upload, among QUAD, ROWS and SORTED
 * the layout whose gathers cost least over a sample of 64 row blocks
 * (instructions with a live lane plus 0.04 per distinct 128-byte line).
 * Results differ between layouts by rounding only.  No counterpart in the
 * reference; it exists to time and test each layout. */
#define CFDLA_GATHER_AUTO (-1)
#define CFDLA_GATHER_QUAD 0
#define CFDLA_GATHER_ROWS 1
/* CFDLA_GATHER_STRIDED: the block staged in LDS as for ROWS, entries dealt
 * to threads round robin (a wavefront gathers for 64 consecutive stored
 * entries), products summed per row as for QUAD.  Never picked by
 * CFDLA_GATHER_AUTO. */
#define CFDLA_GATHER_STRIDED 2
/* CFDLA_GATHER_SORTED: every block's entries stored in column order (each
 * packed with its position in the row-ordered block), so a wavefront gathers
 * for 64 stored entries of adjacent columns as it loads them; the products
 * are placed in row order in LDS and summed as for QUAD, which they equal bit
 * for bit.  Needs fewer than 2^21 columns (otherwise QUAD is used) and keeps
 * 32-bit columns. */
#define CFDLA_GATHER_SORTED 3
cfdla_status cfdla_matrix_set_gather(cfdla_matrix m, int32_t layout);
/* Layout in use for FP64 and FP32 products (uploads the matrix if needed);
 * CFDLA_GATHER_QUAD for matrices not stored as row-block CSR. */
cfdla_status cfdla_matrix_gather(cfdla_matrix m, int32_t *layout_fp64, int32_t *layout_fp32);
cfdla_status cfdla_matrix_destroy(cfdla_matrix m);

/* y = A x on host vectors.  Replaces cfd_math::sparse::spmv / try_spmv
 * (crates/cfd-math/src/sparse/operations.rs:26-80). */
cfdla_status cfdla_spmv(cfdla_matrix m, const double *x, double *y);
/* y = A x on device vectors of dtype (F64 or F32). */
cfdla_status cfdla_spmv_device(cfdla_matrix m, int32_t dtype, const void *x, void *y);
/* r = b - A x on device FP64 vectors; writes ||r||_2 to *rnorm if non-NULL. */
cfdla_status cfdla_residual_device(cfdla_matrix m, const double *b, const double *x, double *r,
                                   double *rnorm);

/* ---- preconditioners ------------------------------------------------------ */

/* Jacobi: z = D^-1 r with identity where |d_ii| <= f64::EPSILON
 * (leto_ops::JacobiPreconditioner, crates/cfd-math/src/lib.rs:181-184;
 * DiagJacobi, crates/cfd-1d/src/solver/core/linear_system.rs:270-299).
 * Non-finite diagonal -> CFDLA_ERR_NUMERICAL. */
cfdla_status cfdla_precond_create_jacobi(cfdla_matrix m, cfdla_precond *out);
/* Multicolour SOR preconditioner, z = (D / omega + L_c)^-1 r: one forward
 * sweep from z = 0 over the rows in colour order (greedy colouring of the
 * symmetrized pattern in breadth-first order; a 5-point stencil gets
 * red-black),
 * L_c the couplings to earlier colours.  Rows without a nonzero diagonal get
 * a unit pivot.  Replaces SuccessiveOverRelaxation::from_csr_with_identity_rows
 * (momentum/solve.rs:137-147), whose lexicographic sweep is sequential; the
 * sweep order, hence the iterates, differ.  *ncolours (nullable) receives the
 * colour count. */
cfdla_status cfdla_precond_create_sor(cfdla_matrix m, double omega, int32_t *ncolours, cfdla_precond *out);
/* AlgebraicMultigrid::new (amg.rs:133-152).  Host setup (coarsening,
 * interpolation, Galerkin products); device V-cycle. */
cfdla_status cfdla_precond_create_amg(cfdla_matrix m, const cfdla_amg_config *config,
                                      cfdla_precond *out);
/* z = M^-1 r (z fully overwritten, amg.rs:553-554) on host vectors.
 * Replaces Preconditioner::apply_to (amg.rs:506-558) and
 * athena Preconditioner::apply (amg.rs:560-601). */
cfdla_status cfdla_precond_apply(cfdla_precond p, const double *r, double *z);
cfdla_status cfdla_precond_apply_device(cfdla_precond p, int32_t dtype, const void *r, void *z);
/* AlgebraicMultigrid::recompute (amg.rs:179-214): new fine values, same
 * pattern; coarse operators recomputed with the existing P and R. */
cfdla_status cfdla_amg_update_values(cfdla_precond p, const double *values);
/* Gather layout (CFDLA_GATHER_*) of the row-block CSR products of every
 * level of an AMG hierarchy (its operators, transfer operators and fused
 * coarse product); the device copies are rebuilt if they exist.
 * CFDLA_GATHER_AUTO (the default) picks per operator.  No counterpart in the
 * reference; it exists to time and test each layout. */
cfdla_status cfdla_amg_set_gather(cfdla_precond p, int32_t layout);
/* The same recompute from the current device values of m (the matrix the
 * preconditioner was built from), on the device: the level-0 copy is
 * refreshed, A_{l+1} = (R_l A_l) P_l is formed by numeric products on the
 * fixed patterns in the host's summation order, then the smoother diagonals
 * and the dense coarsest inverse (Gauss-Jordan with partial pivoting, the
 * host's operation order) are recomputed.  The result is bit-identical to
 * cfdla_amg_update_values with the same values; nothing crosses PCIe except
 * a singularity flag.  Device side of the pressure path's cached-hierarchy
 * refresh (crates/cfd-2d/src/pressure_velocity/correction.rs:56-89). */
cfdla_status cfdla_amg_update_values_device(cfdla_precond p, cfdla_matrix m);
cfdla_status cfdla_amg_stats(cfdla_precond p, cfdla_amg_statistics *out);
/* Hierarchy export for structure checks.  which: 0 = A_l, 1 = R_l, 2 = P_l.
 * First call with arrays NULL to get sizes. */
cfdla_status cfdla_amg_export_level(cfdla_precond p, int32_t level, int32_t which, int64_t *nrows,
                                    int64_t *ncols, int64_t *nnz, int64_t *row_ptr,
                                    int64_t *col_indices, double *values);
/* Aggregate map of level l (fine_to_coarse_map, -1 = None). */
cfdla_status cfdla_amg_export_map(cfdla_precond p, int32_t level, int64_t *fine_to_coarse);
cfdla_status cfdla_precond_destroy(cfdla_precond p);

/* ---- solves ------------------------------------------------------------------ */

/* Solve A x = b.  Replaces krylov::{cg,bicgstab,gmres}[_preconditioned]
 * and SolverKind::{solve,solve_preconditioned} (krylov.rs:150-373,
 * 562-612).  p NULL = identity.  x carries the initial guess in and the last
 * iterate out, whatever the termination. */
cfdla_status cfdla_solve(cfdla_matrix m, cfdla_precond p, const double *b, double *x_inout,
                         const cfdla_solver_config *config, cfdla_report *report);
/* Same on device FP64 vectors. */
cfdla_status cfdla_solve_device(cfdla_matrix m, cfdla_precond p, const double *b_dev,
                                double *x_dev, const cfdla_solver_config *config,
                                cfdla_report *report);
/* Adjoint solve A^T lambda = g (device FP64), m square.  p (nullable) must
 * have been built for a matrix of m's size (else CFDLA_ERR_DIMENSION, before
 * any device work).  For a symmetric matrix p is reused as the preconditioner
 * of A^T = A; otherwise p is not used: the transpose is formed once on the
 * host, cached until m's values change, and solved with its own Jacobi
 * preconditioner.  No reference counterpart (the reference has no adjoint
 * code). */
cfdla_status cfdla_solve_transpose_device(cfdla_matrix m, cfdla_precond p, const double *g_dev,
                                          double *lambda_dev, const cfdla_solver_config *config,
                                          cfdla_report *report);

/* ---- gradients (no reference counterpart) ----------------------------------- */

/* Cell-centred variable-coefficient 5-point Poisson on nx*ny cells with
 * face coefficient 2 kP kN/(kP+kN) * inv_h2 and Dirichlet faces
 * (dir_mask bits 0..3 = -x,+x,-y,+y; values dir_val[4]) adding kP*inv_h2.
 * Given log-coefficients g (k = exp g), the forward solution p, the adjoint
 * lambda and the explicit objective derivative dJdk (per cell), writes
 * dJ/dg = k * (dJdk - lambda^T dA/dk p + lambda^T db/dk) per cell.
 * All device FP64 arrays of nx*ny. */
cfdla_status cfdla_coef_grad_poisson2d_device(int64_t nx, int64_t ny, double inv_h2,
                                              int32_t dir_mask, const double *dir_val,
                                              const double *k, const double *p,
                                              const double *lambda, const double *dJdk,
                                              double *dJdg);

/* ---- batched Kirchhoff networks ------------------------------------------------ */

/* Many independent networks solved in one persistent launch.  No reference
 * counterpart for the batching itself; each network is assembled as
 * MatrixAssembler::assemble_into (crates/cfd-1d/src/solver/core/matrix_assembly.rs:163-247)
 * and solved like LinearSystemSolver::solve_with_initial_guess
 * (crates/cfd-1d/src/solver/core/linear_system.rs:114-155): Jacobi-PCG, with
 * convergence and the reported residual norms measured on the
 * row-equilibrated residual (linear_system.rs:188-246).  Networks are
 * limited to 2048 nodes each. */
typedef struct {
  int32_t wave_max_rows; /* networks with at most this many nodes are solved by one wavefront (<= 512) */
  int32_t blocks_per_cu; /* persistent grid = blocks_per_cu * compute units */
} cfdla_batch_schedule;

/* node_offset / edge_offset: n_sys + 1 prefix offsets; edge_i / edge_j:
 * local node ids per edge; is_dirichlet / dirichlet_value / neumann (nullable):
 * per node.  Errors follow classify_boundary_conditions_into and assemble_into
 * (matrix_assembly.rs:92-198): self loops, non-positive or non-finite
 * conductance, a network without a Dirichlet node, isolated free nodes. */
cfdla_status cfdla_batch_create(int64_t n_sys, const int64_t *node_offset, const int64_t *edge_offset,
                                const int64_t *edge_i, const int64_t *edge_j, const double *conductance,
                                const uint8_t *is_dirichlet, const double *dirichlet_value,
                                const double *neumann, cfdla_batch *out);
/* Re-assemble with new conductances (same graphs), e.g. per Picard step. */
cfdla_status cfdla_batch_update_conductance(cfdla_batch b, const double *conductance);
/* Re-forms every network's assembled values, right-hand side and Jacobi /
 * equilibration factors on the device from device conductances (n_edges
 * doubles), bit for bit as cfdla_batch_update_conductance assembles them on
 * the host (matrix_assembly.rs:163-247: the recorded order of every sum is
 * replayed with no FMA contraction).  Rejects non-positive or non-finite
 * conductances before changing anything.  No counterpart in the reference,
 * which re-assembles each network on the host: it keeps design loops (new
 * widths -> conductances -> assembly -> solve) on the GPU. */
cfdla_status cfdla_batch_update_conductance_device(cfdla_batch b, const double *conductance_dev);
cfdla_status cfdla_batch_info(cfdla_batch b, int64_t *n_sys, int64_t *n_rows, int64_t *nnz);
/* Assembled block-diagonal system (global column ids) and right-hand side. */
cfdla_status cfdla_batch_export(cfdla_batch b, int64_t *row_ptr, int64_t *col_indices, double *values,
                                double *rhs);
cfdla_status cfdla_batch_set_schedule(cfdla_batch b, const cfdla_batch_schedule *schedule);
cfdla_status cfdla_batch_get_schedule(cfdla_batch b, cfdla_batch_schedule *schedule);
/* Solve every network (config.kind must be CFDLA_CG).  x: packed node
 * pressures, initial guess in, solution out; reports: n_sys entries or NULL. */
cfdla_status cfdla_batch_solve(cfdla_batch b, const cfdla_solver_config *config, double *x_inout,
                               cfdla_report *reports);
/* Same on device vectors; rhs_dev NULL uses the assembled right-hand side. */
cfdla_status cfdla_batch_solve_device(cfdla_batch b, const cfdla_solver_config *config,
                                      const double *rhs_dev, double *x_dev, cfdla_report *reports);
/* Allocation-free variant of cfdla_batch_solve_device: same arguments and
 * results (reports is a host array of n_sys entries or NULL; the call
 * synchronizes the library stream, as cfdla_batch_solve_device does).  No
 * counterpart in the reference.  It is for callers that cannot allocate
 * device memory during a call: tight serving and design loops, harnesses
 * that forbid hipMalloc.  The work-queue counter and the per-system report
 * staging live in a static device array of 256 B + 16384 x 48 B
 * (768.25 KiB) that every GPU the library runs on holds until the process
 * ends; its address is resolved once per device by cfdla_init.  Limits:
 * at most 16384 systems (CFDLA_ERR_DIMENSION beyond), and the batch must
 * already be resident on the device (CFDLA_ERR_CONFIG otherwise: its first
 * cfdla_batch_solve_device or cfdla_batch_autotune uploads it).  The counter
 * is cleared by a kernel at the start of every call.  All calls on one GPU
 * share the array: calls that overlap on one GPU (other host threads or
 * streams) race and must be serialized by the caller. */
cfdla_status cfdla_batch_solve_device_noalloc(cfdla_batch b, const cfdla_solver_config *config,
                                              const double *rhs_dev, double *x_dev, cfdla_report *reports);
/* Time each candidate schedule with one full solve; ms_out[k] gets its time
 * and *best the fastest, which is also applied to the batch. */
cfdla_status cfdla_batch_autotune(cfdla_batch b, const cfdla_solver_config *config, int32_t n_candidates,
                                  const cfdla_batch_schedule *candidates, double *ms_out,
                                  cfdla_batch_schedule *best);
/* Edge gradient dJ/dparam_e = (dJ/dG_e|explicit - (l_i - l_j)(p_i - p_j)) * dG_e/dparam_e
 * from device pressures p and adjoint l (Dirichlet nodes' l taken as 0);
 * explicit_dg and dg_dparam are device arrays per edge or NULL.  No
 * reference counterpart. */
cfdla_status cfdla_batch_edge_gradient_device(cfdla_batch b, const double *p, const double *lambda,
                                              const double *explicit_dg, const double *dg_dparam,
                                              double *out);
/* Edge flows from device node pressures p (packed, n_rows), as the
 * reference's update_from_solution forms them
 * (crates/cfd-1d/src/domain/network/wrapper.rs:349-428):
 * flow_e = (p_i - p_j) / (R_e + k_e |flow_e|), with the previous flows read
 * from flow_inout and the new ones written there; quad_coeff (k) may be NULL
 * (linear channels: flow = (p_i - p_j) / R_e).  An edge with R < 0, k < 0,
 * R and k both below machine epsilon, or a non-finite or non-positive
 * effective resistance rejects the call (CFDLA_ERR_CONFIG) before any flow
 * is written.  All arrays on the device, n_edges long. */
cfdla_status cfdla_batch_edge_flows_device(cfdla_batch b, const double *p, const double *resistance,
                                           const double *quad_coeff, double *flow_inout);
/* Outlet flux of every network, flux[s] = sum over its edges e = (i, j) of
 * G_e (m_j (p_i - p_j) + m_i (p_j - p_i)): the flow into the nodes whose
 * outlet_mask (uint8 per node) is set, summed in a fixed order.  Optional
 * outputs for adjoint gradients: dflux_dp (per node, 0 on Dirichlet nodes)
 * and dflux_dg (per edge, the explicit derivative in G_e).  Device arrays;
 * conductance holds the G_e the pressures were solved with.  No counterpart
 * in the reference (its design objectives are host loops over the flows). */
cfdla_status cfdla_batch_outlet_flux_device(cfdla_batch b, const double *p, const double *conductance,
                                            const uint8_t *outlet_mask, double *flux, double *dflux_dp,
                                            double *dflux_dg);
cfdla_status cfdla_batch_destroy(cfdla_batch b);

/* ---- problem generators (host, deterministic) -------------------------------- */

/* count planar channel networks on a grid: random spanning tree plus extra
 * grid edges (degrees 2-4), n ~ U{nmin..nmax}, widths U[50, 500] um,
 * lengths U[1, 10] mm; graph g uses SplitMix64 state seed_base + g.
 * Two-phase: with edge_i NULL only the offsets are written. */
cfdla_status cfdla_synth_networks(int64_t count, uint64_t seed_base, int64_t nmin, int64_t nmax,
                                  int64_t *node_offset, int64_t *edge_offset, int64_t *edge_i,
                                  int64_t *edge_j, double *width, double *length);
/* First-order-upwind + diffusion momentum operator for the x velocity on an
 * nx*ny node grid (crates/cfd-2d/src/physics/momentum/coefficients.rs:199-277,406-414;
 * assembly crates/cfd-2d/src/physics/momentum/solve.rs:264-308); boundary
 * nodes are identity rows; u, v node fields indexed j*nx + i.  Two-phase. */
cfdla_status cfdla_problem_momentum_u(int64_t nx, int64_t ny, double dx, double dy, double dt, double mu,
                                      double rho, const double *u, const double *v, int64_t *nnz,
                                      int64_t *row_ptr, int64_t *col_indices, double *values);
/* Pressure-correction right-hand side -rho/dt div u* at the interior cells
 * (correction.rs:345-356), rows as cfdla_problem_pressure_correction with
 * pin_reference = 1 when no Dirichlet side is present. */
cfdla_status cfdla_problem_pressure_rhs(int64_t nx, int64_t ny, double dx, double dy, double rho, double dt,
                                        const double *u, const double *v, int32_t pin_reference,
                                        double *rhs);
/* Same right-hand side from device arrays u, v into a device rhs (the
 * host generator's operations, bit for bit): with the constant-coefficient
 * matrix of cfdla_problem_pressure_correction kept on the device, a
 * pressure-correction iteration needs no host assembly. */
cfdla_status cfdla_pressure_rhs_device(int64_t nx, int64_t ny, double dx, double dy, double rho, double dt,
                                       const double *u_dev, const double *v_dev, int32_t pin_reference,
                                       double *rhs_dev);
/* Rectangular-duct conductance G = 1/R (rectangular.rs:95-157,209-245) and,
 * if dg_dw is non-NULL, dG/dw at fixed h, L, mu. */
cfdla_status cfdla_rect_conductance(int64_t n, const double *width, const double *height,
                                    const double *length, double viscosity, double *g, double *dg_dw);
/* Same on device arrays (no geometry validation: a zero width gives a
 * non-finite G, which cfdla_batch_update_conductance_device rejects).  tanh
 * comes from the device math library, so G can differ from the host's in the
 * last bits. */
cfdla_status cfdla_rect_conductance_device(int64_t n, const double *width_dev, const double *height_dev,
                                           const double *length_dev, double viscosity, double *g_dev,
                                           double *dg_dw_dev);

/* SplitMix64 uniform[lo, hi) stream: out[i] for i < n from seed. */
cfdla_status cfdla_synth_uniform(uint64_t seed, double lo, double hi, int64_t n, double *out);
/* Periodic Gaussian random field on nx*ny cells (x fastest) with covariance
 * exp(-r^2 / (2 corr_len^2)) in cells, normalised to sample mean 0 and
 * sample standard deviation 1. */
cfdla_status cfdla_synth_grf(int64_t nx, int64_t ny, double corr_len, uint64_t seed, double *out);
/* SIMPLE pressure-correction system on an nx*ny node grid, node (i, j) =
 * row j*nx + i (crates/cfd-2d/src/solvers/simple/pressure.rs:18-240).
 * Interior rows: a_nb = rho d_face (face length / spacing) with d_face the
 * mean of the nodes' momentum d_u / d_v, a_P = -sum a_nb; rhs = Rhie-Chow
 * mass imbalance.  Boundary rows: identity across a pressure side, else 1 on
 * the diagonal and -1/m on the m nodes across the touching sides (periodic
 * sides: the partner node); node 0 is an identity row when no side is a
 * pressure side.  side[4] = {west, east, south, north} of CFDLA_SIDE_*.
 * Fields are nx*ny arrays.  rhs is 0 at boundary nodes; *max_residual
 * (nullable) = max |mass imbalance|.  positive_diagonal = 1 negates every
 * interior row and its rhs entry (exactly; the solution is unchanged), so
 * all diagonals are positive: the form to precondition with AMG, whose
 * Galerkin sums over an aggregate must not mix rows of opposite sign.
 * Two-phase: with row_ptr NULL only *nnz is written and the fields may be
 * NULL. */
#define CFDLA_SIDE_PLAIN 0    /* wall, outflow, velocity inlet */
#define CFDLA_SIDE_PRESSURE 1 /* Dirichlet / pressure inlet / pressure outlet: anchors p' */
#define CFDLA_SIDE_PERIODIC 2 /* periodic with the opposite side */
cfdla_status cfdla_problem_simple_pressure(int64_t nx, int64_t ny, double dx, double dy, const int32_t *side,
                                           int32_t positive_diagonal, const double *u, const double *v,
                                           const double *p, const double *d_u, const double *d_v, const double *rho,
                                           int64_t *nnz, int64_t *row_ptr, int64_t *col_indices, double *values,
                                           double *rhs, double *max_residual);
/* Per-iteration refresh of that system on the device (the refresh branch,
 * pressure.rs:244-450): m must hold cfdla_problem_simple_pressure's pattern
 * for the same nx, ny; its interior rows are recomputed from device fields
 * (the same operations as the host assembly, bit for bit, negated with
 * positive_diagonal = 1), boundary rows keep their values, rhs_dev (nx*ny)
 * receives the imbalance and 0 at boundary nodes.  *max_residual (host,
 * nullable; non-NULL synchronizes) receives max |mass imbalance|. */
cfdla_status cfdla_simple_pressure_refresh_device(cfdla_matrix m, int64_t nx, int64_t ny, double dx, double dy,
                                                  int32_t positive_diagonal, const double *u, const double *v,
                                                  const double *p,
                                                  const double *d_u, const double *d_v, const double *rho,
                                                  double *rhs_dev, double *max_residual);
/* SimpleAlgorithm::apply_corrections (pressure.rs:487-520) on device arrays,
 * interior nodes: p += relax p', u -= d_u (dp'/dx) dx, v -= d_v (dp'/dy) dy
 * with central differences of p'. */
cfdla_status cfdla_simple_corrections_device(int64_t nx, int64_t ny, double dx, double dy, double relax,
                                             const double *p_prime, const double *d_u, const double *d_v,
                                             double *p, double *u, double *v);
/* Cell-centred pressure-correction Laplacian on nxc*nyc interior cells,
 * all fluid (crates/cfd-2d/src/pressure_velocity/correction.rs:273-343);
 * dirichlet[4] = {west, east, south, north}.  Without a Dirichlet side the
 * first cell is pinned as in correction.rs:225-246 and the system has
 * nxc*nyc - 1 rows.  Two-phase: call with arrays NULL to get nnz. */
cfdla_status cfdla_problem_pressure_correction(int64_t nxc, int64_t nyc, double dx, double dy,
                                               const int32_t *dirichlet, int64_t *nnz,
                                               int64_t *row_ptr, int64_t *col_indices,
                                               double *values);
/* Variable-coefficient cell-centred Poisson on nx*ny*nz cells (nz = 1 for
 * 2D) with harmonic-mean faces; dir_mask bits 0..5 = -x,+x,-y,+y,-z,+z;
 * rhs_bc receives the Dirichlet contributions.  Two-phase like above. */
cfdla_status cfdla_problem_varcoef(int64_t nx, int64_t ny, int64_t nz, double inv_h2,
                                   const double *k, int32_t dir_mask, const double *dir_val,
                                   int64_t *nnz, int64_t *row_ptr, int64_t *col_indices,
                                   double *values, double *rhs_bc);
/* Rewrites the values of m, which must hold cfdla_problem_varcoef's pattern
 * for this grid (checked on the first call), from device coefficients k_dev
 * (nx*ny*nz doubles), on the device and bit for bit as
 * cfdla_problem_varcoef forms them; rhs_bc_dev (device, nullable) receives
 * the Dirichlet contributions.  Derived device copies (FP32, diagonal
 * storage, smoother diagonal) follow; a preconditioner built on m is
 * brought up to date with cfdla_amg_update_values_device.  No counterpart in
 * the reference, which assembles on the host: it keeps design loops over
 * the coefficients (adjoint gradients) on the GPU. */
cfdla_status cfdla_varcoef_refresh_device(cfdla_matrix m, int64_t nx, int64_t ny, int64_t nz, double inv_h2,
                                         const double *k_dev, int32_t dir_mask, const double *dir_val,
                                         double *rhs_bc_dev);

#ifdef __cplusplus
}
#endif

#endif /* CFDLA_H */
