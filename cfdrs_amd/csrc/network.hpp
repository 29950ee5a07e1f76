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

// Block-diagonal matrix (global column indices) and right-hand side.
void assemble_networks(const NetworkInput &in, HostCsr &A, std::vector<double> &rhs);

// Edge-wise adjoint gradient (batch.hip); ei/ej global node ids, all device arrays.
void network_gradient(int64_t ne, const int32_t *ei, const int32_t *ej, const uint8_t *isdir, const double *p,
                      const double *lam, const double *explicit_dg, const double *dg_dparam, double *out);

struct BatchSchedule {
  int32_t wave_max_rows = 256;  // systems with n <= this are solved by one wavefront
  int32_t blocks_per_cu = 2;    // persistent grid = blocks_per_cu * CUs
};

struct BatchReport {
  int64_t iterations, outer;
  double rn0, rn, thr;
  int32_t term, pad;
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
  DevBuf<unsigned> d_queue;
  int n_block_items = 0, n_wave_items = 0, n_wave_sys = 0;
  void build(const NetworkInput &in);
  void upload();
  void make_items();
  void set_values(const double *g, const NetworkInput &in);
  // x: device FP64 packed solutions (initial guess in, solution out); reports: host array of nsys or null
  void solve(double *x, const cfdla_solver_config &cfg, BatchReport *reports, const double *rhs_override = nullptr);
};

}  // namespace cfdla
