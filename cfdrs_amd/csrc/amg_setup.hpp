// Host-side AMG setup of libcfdla (see amg_setup.cpp).
#pragma once

#include <functional>
#include <vector>

#include "internal.hpp"

namespace cfdla {

struct CoarseSplit {
  std::vector<int32_t> coarse_points;
  std::vector<int32_t> fine_to_coarse;  // -1 = no coarse point
  HostCsr strength;
};

struct HostLevel {
  HostCsr A, P, R;  // P, R empty on the coarsest level
  std::vector<int32_t> fine_to_coarse;
};

struct HostHierarchy {
  std::vector<HostLevel> levels;
};

// Numeric plan of C = A B on C's fixed pattern: slot k of C sums
// A.v[pa[q]] * B.v[pb[q]] for q in [ptr[k], ptr[k + 1]) in that order, the
// order in which csr_multiply's row accumulator meets the products.
struct ProductPlan {
  std::vector<int32_t> ptr, pa, pb;
};
ProductPlan product_plan(const HostCsr &a, const HostCsr &b, const HostCsr &c);

HostCsr csr_transpose(const HostCsr &a);
HostCsr csr_multiply(const HostCsr &a, const HostCsr &b);
HostCsr strong_connections(const HostCsr &a, double theta);
CoarseSplit greedy_aggregates(const HostCsr &a, int max_size);
CoarseSplit ruge_stueben_split(const HostCsr &a, double theta);
CoarseSplit hybrid_split(const HostCsr &a, double theta);
CoarseSplit three_phase_aggregates(const HostCsr &a, double theta);
HostCsr piecewise_constant_p(const std::vector<int32_t> &f2c, int64_t n, int64_t nc);
HostCsr classical_p(const HostCsr &a, const std::vector<int32_t> &cpts, const HostCsr &s);
HostCsr smooth_prolongator(const HostCsr &a, const HostCsr &pt, double omega);
std::vector<double> dense_inverse(const HostCsr &a);

// The Galerkin product of level l, A_{l+1} = (R A) P; the host product when
// none is given.
using GalerkinFn = std::function<HostCsr(size_t level, const HostCsr &R, const HostCsr &A, const HostCsr &P)>;
void build_hierarchy(const HostCsr &fine, const cfdla_amg_config &cfg, HostHierarchy &h,
                     const GalerkinFn &galerkin = nullptr);
void recompute_hierarchy(HostHierarchy &h, const double *fine_values);

}  // namespace cfdla
