// Host assembly of the 2D SIMPLE systems (lid-driven cavity workload):
// the first-order-upwind momentum operator and the divergence right-hand
// side of the cell-centred pressure correction.
//
// Momentum: coefficients crates/cfd-2d/src/physics/momentum/coefficients.rs:199-277,406-414
// (diffusion mu dy/dx and mu dx/dy, upwind convection from the cell-centred
// velocity, a_P = a_E + a_W + a_N + a_S + rho dx dy / dt); assembly
// crates/cfd-2d/src/physics/momentum/solve.rs:264-308 (boundary nodes are
// identity rows, interior rows a_P on the diagonal and -a_nb off it).
// Node (i, j), i along x, is row j * nx + i; columns are written sorted.
//
// Pressure right-hand side: crates/cfd-2d/src/pressure_velocity/correction.rs:345-356,
// rhs = -rho / dt * div u* with central differences of the node velocities.
//
// SIMPLE pressure-correction system: crates/cfd-2d/src/solvers/simple/pressure.rs:18-240
// (interior node stencil in simple_stencil.hpp; boundary rows below).
#include <algorithm>
#include <cmath>

#include "../../include/cfdla.h"
#include "internal.hpp"
#include "simple_stencil.hpp"

namespace {
template <typename F>
cfdla_status run(F &&f) {
  try {
    f();
    cfdla::set_last_error("");
    return CFDLA_OK;
  } catch (const cfdla::Error &e) {
    cfdla::set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    cfdla::set_last_error(e.what());
    return CFDLA_ERR_INVALID_ARG;
  }
}
// Boundary row of the SIMPLE pressure-correction system
// (crates/cfd-2d/src/solvers/simple/pressure.rs:150-240, neighbour choice
// crates/cfd-2d/src/pressure_velocity/boundary.rs:29-66).  The touching sides
// are visited west, east, south, north: a pressure-anchor side makes the row
// an identity row; any other side contributes the node across it (the first
// interior node, or the partner node of a periodic side).  Without an anchor
// side anywhere, node 0 is an identity row.  Otherwise the row is 1 on the
// diagonal and -1/m on each of its m distinct neighbours.  Writes the sorted
// columns and values of the row; returns the entry count (1..3).
int simple_boundary_row(int64_t i, int64_t j, int64_t nx, int64_t ny, const int32_t *side, bool anchored,
                        int64_t *cols, double *vals) {
  const int64_t idx = j * nx + i;
  const bool on[4] = {i == 0, i == nx - 1, j == 0, j == ny - 1};
  const int64_t across[4] = {j * nx + 1, j * nx + nx - 2, nx + i, (ny - 2) * nx + i};
  const int64_t partner[4] = {j * nx + nx - 1, j * nx, (ny - 1) * nx + i, i};
  int64_t nb[2];
  int m = 0;
  bool identity = !anchored && idx == 0;
  for (int s = 0; s < 4; ++s) {
    if (!on[s]) continue;
    if (side[s] == CFDLA_SIDE_PRESSURE) {
      identity = true;
      break;
    }
    const int64_t c = side[s] == CFDLA_SIDE_PERIODIC ? partner[s] : across[s];
    if (m == 0 || nb[0] != c) nb[m++] = c;
  }
  if (identity || m == 0) {
    cols[0] = idx;
    vals[0] = 1.0;
    return 1;
  }
  if (m == 2 && nb[1] < nb[0]) std::swap(nb[0], nb[1]);
  const double w = 1.0 / (double)m;
  int o = 0;
  bool diag_done = false;
  for (int k = 0; k < m; ++k) {
    if (!diag_done && idx < nb[k]) {
      cols[o] = idx;
      vals[o++] = 1.0;
      diag_done = true;
    }
    cols[o] = nb[k];
    vals[o++] = -w;
  }
  if (!diag_done) {
    cols[o] = idx;
    vals[o++] = 1.0;
  }
  return o;
}

}  // namespace

extern "C" {

cfdla_status cfdla_problem_momentum_u(int64_t nx, int64_t ny, double dx, double dy, double dt, double mu, double rho,
                                      const double *u, const double *v, int64_t *nnz, int64_t *row_ptr,
                                      int64_t *col_indices, double *values) {
  return run([&] {
    CFDLA_REQUIRE(nx >= 3 && ny >= 3, CFDLA_ERR_INVALID_ARG, "momentum grid needs at least 3x3 nodes");
    CFDLA_REQUIRE(u && v && nnz, CFDLA_ERR_INVALID_ARG, "u, v and nnz must not be NULL");
    const int64_t interior = (nx - 2) * (ny - 2);
    *nnz = interior * 5 + (nx * ny - interior);
    if (!row_ptr) return;
    CFDLA_REQUIRE(col_indices && values, CFDLA_ERR_INVALID_ARG, "col_indices and values must not be NULL");
    const double dew = dy / dx, dns = dx / dy;
    int64_t o = 0;
    row_ptr[0] = 0;
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t idx = j * nx + i;
        if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1) {
          col_indices[o] = idx;
          values[o++] = 1.0;
          row_ptr[idx + 1] = o;
          continue;
        }
        double ae = mu * dew, aw = mu * dew, an = mu * dns, as = mu * dns;
        const double uu = u[idx];
        const double vv = (v[idx] + v[j * nx + i + 1]) / 2.0;
        const double fx = rho * uu * dy, fy = rho * vv * dx;
        if (uu > 0.0) aw = aw + fx;
        else ae = ae - fx;
        if (vv > 0.0) as = as + fy;
        else an = an - fy;
        const double ap = (ae + aw + an + as) + rho * (dx * dy) / dt;
        col_indices[o] = idx - nx; values[o++] = -as;
        col_indices[o] = idx - 1; values[o++] = -aw;
        col_indices[o] = idx; values[o++] = ap;
        col_indices[o] = idx + 1; values[o++] = -ae;
        col_indices[o] = idx + nx; values[o++] = -an;
        row_ptr[idx + 1] = o;
      }
  });
}

cfdla_status cfdla_problem_pressure_rhs(int64_t nx, int64_t ny, double dx, double dy, double rho, double dt,
                                        const double *u, const double *v, int32_t pin_reference, double *rhs) {
  return run([&] {
    CFDLA_REQUIRE(nx >= 3 && ny >= 3, CFDLA_ERR_INVALID_ARG, "pressure grid needs at least 3x3 nodes");
    CFDLA_REQUIRE(u && v && rhs, CFDLA_ERR_INVALID_ARG, "u, v and rhs must not be NULL");
    const double coeff = rho / dt;
    for (int64_t i = 1; i < nx - 1; ++i)
      for (int64_t j = 1; j < ny - 1; ++j) {
        const int64_t idx = (i - 1) * (ny - 2) + (j - 1);
        if (pin_reference && idx == 0) continue;
        const int64_t row = pin_reference ? idx - 1 : idx;
        const double div = (u[j * nx + i + 1] - u[j * nx + i - 1]) / (2.0 * dx) +
                           (v[(j + 1) * nx + i] - v[(j - 1) * nx + i]) / (2.0 * dy);
        rhs[row] = -coeff * div;
      }
  });
}

cfdla_status cfdla_problem_simple_pressure(int64_t nx, int64_t ny, double dx, double dy, const int32_t *side,
                                           int32_t positive_diagonal, const double *u, const double *v,
                                           const double *p, const double *d_u, const double *d_v, const double *rho,
                                           int64_t *nnz, int64_t *row_ptr, int64_t *col_indices, double *values,
                                           double *rhs, double *max_residual) {
  return run([&] {
    CFDLA_REQUIRE(nx >= 3 && ny >= 3, CFDLA_ERR_INVALID_ARG, "SIMPLE pressure grid needs at least 3x3 nodes");
    CFDLA_REQUIRE(side && nnz, CFDLA_ERR_INVALID_ARG, "side and nnz must not be NULL");
    for (int s = 0; s < 4; ++s)
      CFDLA_REQUIRE(side[s] >= CFDLA_SIDE_PLAIN && side[s] <= CFDLA_SIDE_PERIODIC, CFDLA_ERR_CONFIG,
                    "boundary side kind must be plain, pressure or periodic");
    // a periodic side needs a periodic partner (physics/momentum/boundary/consistency.rs:72-93)
    CFDLA_REQUIRE((side[0] == CFDLA_SIDE_PERIODIC) == (side[1] == CFDLA_SIDE_PERIODIC) &&
                      (side[2] == CFDLA_SIDE_PERIODIC) == (side[3] == CFDLA_SIDE_PERIODIC),
                  CFDLA_ERR_CONFIG, "periodic sides come in west-east / south-north pairs");
    const bool anchored = side[0] == CFDLA_SIDE_PRESSURE || side[1] == CFDLA_SIDE_PRESSURE ||
                          side[2] == CFDLA_SIDE_PRESSURE || side[3] == CFDLA_SIDE_PRESSURE;
    int64_t cols[3];
    double vals[3];
    int64_t count = (nx - 2) * (ny - 2) * 5;
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i)
        if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1)
          count += simple_boundary_row(i, j, nx, ny, side, anchored, cols, vals);
    *nnz = count;
    if (!row_ptr) return;
    CFDLA_REQUIRE(col_indices && values && rhs && u && v && p && d_u && d_v && rho, CFDLA_ERR_INVALID_ARG,
                  "fields, col_indices, values and rhs must not be NULL");
    double maxr = 0.0;
    int64_t o = 0;
    row_ptr[0] = 0;
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t idx = j * nx + i;
        if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1) {
          const int m = simple_boundary_row(i, j, nx, ny, side, anchored, cols, vals);
          for (int k = 0; k < m; ++k) {
            col_indices[o] = cols[k];
            values[o++] = vals[k];
          }
          rhs[idx] = 0.0;
        } else {
          const cfdla::SimpleNode s = cfdla::simple_node(i, j, nx, ny, dx, dy, u, v, p, d_u, d_v, rho);
          const int64_t c[5] = {idx - nx, idx - 1, idx, idx + 1, idx + nx};
          const double a[5] = {s.as, s.aw, s.ap, s.ae, s.an};
          for (int k = 0; k < 5; ++k) {
            col_indices[o] = c[k];
            values[o++] = positive_diagonal ? -a[k] : a[k];
          }
          rhs[idx] = positive_diagonal ? -s.imbalance : s.imbalance;
          // strict comparison: a NaN imbalance never becomes the maximum (pressure.rs:140-143)
          if (std::fabs(s.imbalance) > maxr) maxr = std::fabs(s.imbalance);
        }
        row_ptr[idx + 1] = o;
      }
    if (max_residual) *max_residual = maxr;
  });
}

}  // extern "C"
