// One interior node of the SIMPLE pressure-correction system
// (crates/cfd-2d/src/solvers/simple/pressure.rs:57-147, recomputed in the
// refresh branch :260-356).  The host assembly (problems.cpp) and the device
// refresh (simple.hip) both call simple_node, with FP contraction off on
// both sides, so the device values equal the host's bit for bit.
//
// Node (i, j), i along x, is index j * nx + i.  Face d-coefficients are the
// averages of the two nodes' momentum d's; face velocities are the linear
// interpolation plus d_face times (average of the two cell-centred pressure
// gradients - the face gradient), the Rhie-Chow correction, with one-sided
// extrapolation (2 p_nb - p_P) for the second neighbour next to the
// boundary.  Row: a_nb = rho d_face (face length / distance), a_P = -sum a_nb;
// right-hand side: the mass imbalance of the four face fluxes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cfdla {

struct SimpleNode {
  double ap, ae, aw, an, as, imbalance;
};

__host__ __device__ inline SimpleNode simple_node(int64_t i, int64_t j, int64_t nx, int64_t ny, double dx,
                                                  double dy, const double *u, const double *v, const double *p,
                                                  const double *du, const double *dv, const double *rho) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  const double half = 0.5, two = 2.0;
  const int64_t c = j * nx + i;
  const double pp = p[c], pe = p[c + 1], pw = p[c - 1], pn = p[c + nx], ps = p[c - nx];
  const double pee = i + 2 < nx ? p[c + 2] : two * pe - pp;
  const double pww = i >= 2 ? p[c - 2] : two * pw - pp;
  const double pnn = j + 2 < ny ? p[c + 2 * nx] : two * pn - pp;
  const double pss = j >= 2 ? p[c - 2 * nx] : two * ps - pp;
  const double de = (du[c] + du[c + 1]) * half, dw = (du[c - 1] + du[c]) * half;
  const double dn = (dv[c] + dv[c + nx]) * half, ds = (dv[c - nx] + dv[c]) * half;
  const double hx = two * dx, hy = two * dy;
  const double gxp = (pe - pw) / hx, gyp = (pn - ps) / hy;  // cell-centred gradients at P
  const double ufe = (u[c] + u[c + 1]) * half + de * ((gxp + (pee - pp) / hx) * half - (pe - pp) / dx);
  const double ufw = (u[c - 1] + u[c]) * half + dw * (((pp - pww) / hx + gxp) * half - (pp - pw) / dx);
  const double vfn = (v[c] + v[c + nx]) * half + dn * ((gyp + (pnn - pp) / hy) * half - (pn - pp) / dy);
  const double vfs = (v[c - nx] + v[c]) * half + ds * (((pp - pss) / hy + gyp) * half - (pp - ps) / dy);
  const double r = rho[c];
  SimpleNode o;
  o.imbalance = r * ufe * dy - r * ufw * dy + r * vfn * dx - r * vfs * dx;
  o.ae = r * de * dy / dx;
  o.aw = r * dw * dy / dx;
  o.an = r * dn * dx / dy;
  o.as = r * ds * dx / dy;
  o.ap = -(o.ae + o.aw + o.an + o.as);
  return o;
}

}  // namespace cfdla
