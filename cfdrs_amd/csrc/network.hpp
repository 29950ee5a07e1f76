// Batched Kirchhoff networks: host-side generation and assembly (network.cpp)
// and the device batch solver (batch.hip).
#pragma once

#include <memory>
#include <vector>

#include "internal.hpp"

namespace cfdla {

struct SynthNetworks {
  std::vector<int64_t> node_off, edge_off;  // count + 1 each
  std::vector<int32_t> ei, ej;              // local node ids, i < j
  std::vector<double> width, length;        // per edge [m]
};

// count graphs; graph g draws from SplitMix64 state seed_base + g; n ~ U{nmin..nmax}.
void synth_networks(int64_t count, uint64_t seed_base, int64_t nmin, int64_t nmax, SynthNetworks &out);

double rect_resistance(double w, double h, double L, double mu);
// dG/dw of G = 1 / rect_resistance(w, h, L, mu) at fixed h, L, mu.
double rect_conductance_dw(double w, double h, double L, double mu);

struct NetworkInput {
  int64_t nsys;
  const int64_t *node_off, *edge_off;
  const int64_t *ei, *ej;  // local node ids
  const double *g;
  const uint8_t *is_dirichlet;     // per node (global)
  const double *dirichlet_value;   // per node (global)
  const double *neumann;           // per node (global), nullable
};

// The order in which assemble_networks forms every value, so a device kernel
// can re-form them from new conductances bit for bit: slot k of the CSR is
// the sequential sum of its sources slot_src[slot_off[k] .. slot_off[k+1])
// (e + 1: +g_e, -(e + 1): -g_e, 0: the 1 of a Dirichlet row; e indexes the
// batch's edges); free row i's right-hand side is the sequential sum of
// g_{rhs_edge[t]} * p_D[rhs_node[t]] over t in rhs_off[i] .. rhs_off[i+1],
// plus its Neumann source.
struct AssemblyPlan {
  std::vector<int32_t> slot_off, slot_src;
  std::vector<int32_t> rhs_off, rhs_edge, rhs_node;
};

// Block-diagonal matrix (global column indices) and right-hand side; fills
// plan when given.
void assemble_networks(const NetworkInput &in, HostCsr &A, std::vector<double> &rhs, AssemblyPlan *plan = nullptr);

// G = 1 / R and dG/dw of rectangular ducts on the device (arrays of n).
void rect_conductance_device(int64_t n, const double *w, const double *h, const double *L, double mu, double *g,
                             double *dgdw);

// Edge-wise adjoint gradient (batch.hip); ei/ej global node ids, all device arrays.
void network_gradient(int64_t ne, const int32_t *ei, const int32_t *ej, const uint8_t *isdir, const double *p,
                      const double *lam, const double *explicit_dg, const double *dg_dparam, double *out);

// Edge flows (p_i - p_j) / (R + k |q_prev|) into q (k nullable: R only);
// throws CFDLA_ERR_CONFIG, writing nothing, when an edge breaks the
// reference's invariants.  bad: one device int of scratch.
void network_flows(int64_t ne, const int32_t *ei, const int32_t *ej, const double *p, const double *R,
                   const double *kq, double *q, int *bad);
// Outlet flux per network, J[s], and optionally dJ/dp (per node, 0 on
// Dirichlet nodes) and the explicit dJ/dG (per edge).  eoff: nsys + 1 edge
// offsets; inc_ptr / inc_edge: each node's incident edges in edge order.
void network_outlet(int64_t nsys, int64_t nrows, const int64_t *eoff, const int32_t *ei, const int32_t *ej,
                    const int64_t *inc_ptr, const int32_t *inc_edge, const uint8_t *isdir, const uint8_t *mask,
                    const double *p, const double *g, double *J, double *dJdp, double *dJdg);

struct BatchSchedule {
  int32_t wave_max_rows = 256;  // systems with n <= this are solved by one wavefront
  int32_t blocks_per_cu = 2;    // persistent grid = blocks_per_cu * CUs
};

struct BatchReport {
  int64_t iterations, outer;
  double rn0, rn, thr;
  int32_t term, pad;  // pad: the precision the solve ran in (cfdla_report.precision)
};

struct BatchSolver {
  int64_t nsys = 0, nrows = 0;
  HostCsr host;
  std::vector<double> rhs;
  std::vector<int32_t> sys_row;  // nsys + 1
  BatchSchedule sched;
  bool on_device = false;
  DevBuf<int32_t> d_sys_row, d_rp, d_ci, d_items;
  DevBuf<double> d_v64, d_b, d_jd64, d_eq;
  DevBuf<float> d_v32, d_jd32;
  DevBuf<BatchReport> d_rep;
  // device re-assembly from new conductances (set_values_device)
  AssemblyPlan plan;
  int64_t nedges = 0;
  std::vector<uint8_t> isdir_h;
  std::vector<double> dirval_h, neumann_h;
  DevBuf<int32_t> d_slot_off, d_slot_src, d_rhs_off, d_rhs_edge, d_rhs_node, d_bad;
  DevBuf<uint8_t> d_isdir;
  DevBuf<double> d_dirval, d_neumann;
  bool device_newer = false;  // values / rhs were re-formed on the device since the last host assembly
  DevBuf<unsigned> d_queue;
  int n_block_items = 0, n_wave_items = 0, n_wave_sys = 0;
  void build(const NetworkInput &in);
  void upload();
  void make_items();
  void set_values(const double *g, const NetworkInput &in);
  // Re-forms values, right-hand side, Jacobi and equilibration factors on the
  // device from device conductances g (n_edges), as assemble_networks does;
  // rejects non-positive or non-finite conductances before changing anything.
  void set_values_device(const double *g);
  // Host copies of the (possibly device-formed) values and right-hand side.
  void download(std::vector<double> &values, std::vector<double> &rhs_out) const;
  // x: device FP64 packed solutions (initial guess in, solution out); reports: host array of nsys or null
  void solve(double *x, const cfdla_solver_config &cfg, BatchReport *reports, const double *rhs_override = nullptr);
  // The same solve with the queue counter and report staging in the
  // library's static scratch array instead of d_queue / d_rep: allocates
  // nothing, so the batch must already be on the device, with at most
  // 16384 systems.
  void solve_noalloc(double *x, const cfdla_solver_config &cfg, BatchReport *reports,
                     const double *rhs_override = nullptr);
  void launch(double *x, const cfdla_solver_config &cfg, const double *rhs_override, unsigned *queue,
              BatchReport *rep);
};

}  // namespace cfdla
