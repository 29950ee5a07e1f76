// Host side of the batched Kirchhoff-network solves: synthetic planar
// channel networks, rectangular-duct conductances and their width
// derivative, and the per-network Laplacian assembly.
//
// Assembly follows crates/cfd-1d/src/solver/core/matrix_assembly.rs:163-247:
// for each edge in order, free-free edges add G to both diagonals and -G to
// both couplings; an edge to a Dirichlet node adds G to the free diagonal and
// G * p_D to its right-hand side; Neumann sources are added next; Dirichlet
// rows become identity rows with right-hand side p_D.  Duplicate entries are
// summed in push order and every row is sorted by column.
//
// Resistance follows crates/cfd-1d/src/physics/resistance/models/rectangular.rs:95-157,209-245:
// R = Po(eps) mu L / (2 A Dh^2), Dh = 4 A / (2 (w + h)), eps = min/max(w, h),
// Po = 96 / ((1 + eps)^2 (1 - 192 eps / pi^5 tanh(pi / (2 eps)))).
#include "network.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace cfdla {

namespace {
constexpr double kPi = 3.14159265358979323846;

uint64_t mix(uint64_t &state) {
  // SplitMix64 (Steele, Lea & Flood 2014)
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double u01(uint64_t &state) { return (double)(mix(state) >> 11) * (1.0 / 9007199254740992.0); }

double po_of(double eps) {
  if (eps == 0.0) return 96.0;
  const double t1 = (1.0 + eps) * (1.0 + eps);
  const double t2 = 1.0 - (192.0 * eps / std::pow(kPi, 5)) * std::tanh(kPi / (2.0 * eps));
  return 96.0 / (t1 * t2);
}

// dPo/deps of the fit above.
double dpo_deps(double eps) {
  const double c = 192.0 / std::pow(kPi, 5);
  const double u = kPi / (2.0 * eps);
  const double th = std::tanh(u);
  const double sech2 = 1.0 - th * th;
  const double T = 1.0 - c * eps * th;
  const double dT = -c * th + c * u * sech2;  // d/deps of -c eps tanh(pi / (2 eps))
  const double q = (1.0 + eps) * (1.0 + eps) * T;
  const double dq = 2.0 * (1.0 + eps) * T + (1.0 + eps) * (1.0 + eps) * dT;
  return -96.0 * dq / (q * q);
}
}  // namespace

double rect_resistance(double w, double h, double L, double mu) {
  const double area = w * h;
  const double dh = 4.0 * area / (2.0 * (w + h));
  const double lo = w <= h ? w : h, hi = w >= h ? w : h;
  const double po = po_of(lo / hi);
  return (po * mu * L) / (2.0 * area * dh * dh);
}

double rect_conductance_dw(double w, double h, double L, double mu) {
  // G = 1/R, R = Po(eps) mu L / (2 f), f = A Dh^2 = 4 w^3 h^3 / (w + h)^2
  const double s = w + h;
  const double f = 4.0 * w * w * w * h * h * h / (s * s);
  const double df = 4.0 * h * h * h * (3.0 * w * w / (s * s) - 2.0 * w * w * w / (s * s * s));
  double eps, deps;
  if (w >= h) {
    eps = h / w;
    deps = -h / (w * w);
  } else {
    eps = w / h;
    deps = 1.0 / h;
  }
  const double po = po_of(eps);
  const double dpo = dpo_deps(eps) * deps;
  const double R = po * mu * L / (2.0 * f);
  const double dR = mu * L / 2.0 * (dpo * f - po * df) / (f * f);
  return -dR / (R * R);
}

void synth_networks(int64_t count, uint64_t seed_base, int64_t nmin, int64_t nmax, SynthNetworks &out) {
  CFDLA_REQUIRE(count >= 0 && nmin >= 2 && nmax >= nmin, CFDLA_ERR_INVALID_ARG, "invalid network synthesis sizes");
  out.node_off.assign((size_t)count + 1, 0);
  out.edge_off.assign((size_t)count + 1, 0);
  out.ei.clear();
  out.ej.clear();
  out.width.clear();
  out.length.clear();
  std::vector<std::pair<int32_t, int32_t>> grid, tree, extra;
  std::vector<double> key;
  std::vector<int32_t> order, parent, deg;
  for (int64_t g = 0; g < count; ++g) {
    uint64_t st = seed_base + (uint64_t)g;
    const int64_t n = std::min<int64_t>(nmax, nmin + (int64_t)(u01(st) * (double)(nmax - nmin + 1)));
    const int64_t W = (int64_t)std::ceil(std::sqrt((double)n));
    grid.clear();
    for (int64_t k = 0; k < n; ++k) {
      if ((k % W) + 1 < W && k + 1 < n) grid.emplace_back((int32_t)k, (int32_t)(k + 1));
      if (k + W < n) grid.emplace_back((int32_t)k, (int32_t)(k + W));
    }
    key.resize(grid.size());
    for (auto &x : key) x = u01(st);
    order.resize(grid.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    parent.resize((size_t)n);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int32_t a) {
      while (parent[a] != a) {
        parent[a] = parent[parent[a]];
        a = parent[a];
      }
      return a;
    };
    deg.assign((size_t)n, 0);
    tree.clear();
    extra.clear();
    std::vector<int32_t> rest;
    for (int32_t e : order) {
      int32_t a = find(grid[e].first), b = find(grid[e].second);
      if (a != b) {
        parent[a] = b;
        tree.push_back(grid[e]);
        deg[grid[e].first]++;
        deg[grid[e].second]++;
      } else {
        rest.push_back(e);
      }
    }
    // close leaves first so that degrees land in 2..4, then sprinkle loops
    for (int32_t e : rest) {
      auto [i, j] = grid[e];
      if (deg[i] == 1 || deg[j] == 1 || u01(st) < 0.1) {
        extra.push_back(grid[e]);
        deg[i]++;
        deg[j]++;
      }
    }
    out.node_off[g + 1] = out.node_off[g] + n;
    for (auto &e : tree) { out.ei.push_back(e.first); out.ej.push_back(e.second); }
    for (auto &e : extra) { out.ei.push_back(e.first); out.ej.push_back(e.second); }
    out.edge_off[g + 1] = (int64_t)out.ei.size();
    for (int64_t e = out.edge_off[g]; e < out.edge_off[g + 1]; ++e) {
      out.width.push_back(50e-6 + 450e-6 * u01(st));
      out.length.push_back(1e-3 + 9e-3 * u01(st));
    }
  }
}

void assemble_networks(const NetworkInput &in, HostCsr &A, std::vector<double> &rhs, AssemblyPlan *plan) {
  const int64_t ns = in.nsys;
  const int64_t N = in.node_off[ns];
  CFDLA_REQUIRE(N < INT32_MAX, CFDLA_ERR_INVALID_ARG, "too many network nodes");
  A.n = A.m = N;
  A.rp.assign((size_t)N + 1, 0);
  A.ci.clear();
  A.v.clear();
  rhs.assign((size_t)N, 0.0);
  // entries of a row in push order: (column, value, source code of the plan)
  struct Entry {
    int32_t col;
    double v;
    int32_t src;
  };
  std::vector<std::vector<Entry>> rows;
  std::vector<std::vector<std::pair<int32_t, int32_t>>> rhs_src;  // (edge, dirichlet node) per free row
  if (plan) {
    plan->slot_off.assign(1, 0);
    plan->slot_src.clear();
    plan->rhs_off.assign((size_t)N + 1, 0);
    plan->rhs_edge.clear();
    plan->rhs_node.clear();
  }
  for (int64_t s = 0; s < ns; ++s) {
    const int64_t r0 = in.node_off[s], n = in.node_off[s + 1] - r0;
    CFDLA_REQUIRE(n >= 1, CFDLA_ERR_INVALID_ARG, "network with no nodes");
    rows.assign((size_t)n, {});
    rhs_src.assign((size_t)n, {});
    std::vector<char> touched((size_t)n, 0);
    bool any_dirichlet = false;
    for (int64_t i = 0; i < n; ++i) any_dirichlet |= in.is_dirichlet[r0 + i] != 0;
    CFDLA_REQUIRE(any_dirichlet, CFDLA_ERR_INVALID_ARG, "At least one Dirichlet boundary condition is required");
    for (int64_t e = in.edge_off[s]; e < in.edge_off[s + 1]; ++e) {
      const int64_t i = in.ei[e], j = in.ej[e];
      const double c = in.g[e];
      CFDLA_REQUIRE(i >= 0 && i < n && j >= 0 && j < n, CFDLA_ERR_INVALID_ARG, "edge node out of range");
      CFDLA_REQUIRE(i != j, CFDLA_ERR_INVALID_ARG, "Self-loop edge detected in network assembly");
      CFDLA_REQUIRE(std::isfinite(c), CFDLA_ERR_INVALID_ARG, "Non-finite conductance encountered in network assembly");
      CFDLA_REQUIRE(c > 0.0, CFDLA_ERR_INVALID_ARG, "Non-positive conductance encountered in network assembly");
      touched[i] = touched[j] = 1;
      const bool di = in.is_dirichlet[r0 + i] != 0, dj = in.is_dirichlet[r0 + j] != 0;
      const int32_t plus = (int32_t)(e + 1), minus = -(int32_t)(e + 1);
      if (!di && !dj) {
        rows[i].push_back({(int32_t)i, c, plus});
        rows[j].push_back({(int32_t)j, c, plus});
        rows[i].push_back({(int32_t)j, -c, minus});
        rows[j].push_back({(int32_t)i, -c, minus});
      } else if (di && !dj) {
        rows[j].push_back({(int32_t)j, c, plus});
        rhs[r0 + j] += c * in.dirichlet_value[r0 + i];
        rhs_src[j].emplace_back((int32_t)e, (int32_t)(r0 + i));
      } else if (!di && dj) {
        rows[i].push_back({(int32_t)i, c, plus});
        rhs[r0 + i] += c * in.dirichlet_value[r0 + j];
        rhs_src[i].emplace_back((int32_t)e, (int32_t)(r0 + j));
      }
    }
    for (int64_t i = 0; i < n; ++i) {
      CFDLA_REQUIRE(touched[i] || in.is_dirichlet[r0 + i], CFDLA_ERR_INVALID_ARG,
                    "Isolated node without Dirichlet condition");
      if (in.neumann) rhs[r0 + i] += in.neumann[r0 + i];
    }
    for (int64_t i = 0; i < n; ++i)
      if (in.is_dirichlet[r0 + i]) {
        rows[i].push_back({(int32_t)i, 1.0, 0});
        rhs[r0 + i] = in.dirichlet_value[r0 + i];
        rhs_src[i].clear();
      }
    for (int64_t i = 0; i < n; ++i) {
      auto &row = rows[i];
      std::stable_sort(row.begin(), row.end(), [](const Entry &a, const Entry &b) { return a.col < b.col; });
      for (size_t k = 0; k < row.size(); ++k) {
        if (k > 0 && row[k].col == row[k - 1].col) {
          A.v.back() += row[k].v;
        } else {
          A.ci.push_back((int32_t)(row[k].col + r0));
          A.v.push_back(row[k].v);
          if (plan && k > 0) plan->slot_off.push_back((int32_t)plan->slot_src.size());
        }
        if (plan) plan->slot_src.push_back(row[k].src);
      }
      if (plan && !row.empty()) plan->slot_off.push_back((int32_t)plan->slot_src.size());
      A.rp[r0 + i + 1] = (int64_t)A.ci.size();
      if (plan) {
        for (auto &t : rhs_src[i]) {
          plan->rhs_edge.push_back(t.first);  // e runs over the batch's edges already
          plan->rhs_node.push_back(t.second);
        }
        plan->rhs_off[(size_t)(r0 + i + 1)] = (int32_t)plan->rhs_edge.size();
      }
    }
  }
  CFDLA_REQUIRE(A.nnz() < INT32_MAX, CFDLA_ERR_INVALID_ARG, "too many network matrix entries");
}

}  // namespace cfdla
