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
#include <cmath>

#include "../../include/cfdla.h"
#include "internal.hpp"

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

}  // extern "C"
