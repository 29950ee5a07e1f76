// Host-side AMG hierarchy setup of libcfdla: strength of connection,
// coarsening, interpolation, R = P^T and the Galerkin product R A P.
//
// The integer outputs (strong-connection pattern, aggregates, C/F split,
// operator patterns) are bit-exact restatements of the reference:
//   strength       crates/cfd-math/src/linear_solver/preconditioners/multigrid/coarsening/algorithms.rs:395-436
//   aggregation    coarsening/algorithms.rs:129-172 (theta fixed at 0.5, :138)
//   Ruge-Stueben   coarsening/algorithms.rs:34-126, assignment :200-230
//   hybrid         coarsening/algorithms.rs:175-197
//   direct P       multigrid/interpolation.rs:191-217
//   classical P    multigrid/interpolation.rs:38-188
//   R = P^T, A_c = (R A) P   multigrid/amg.rs:349-353
// Sparse products keep every structurally produced entry (explicit zeros
// included) with columns sorted, and accumulate each entry in row-of-A then
// row-of-B order.
#include "amg_setup.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <set>
#include <thread>

namespace cfdla {

namespace {

int worker_count() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 1;
  if (h > 16) h = 16;
  return (int)h;
}

// Run body(lo, hi) over [0, n) split into contiguous chunks.
void parallel_rows(int64_t n, const std::function<void(int64_t, int64_t)> &body) {
  int w = worker_count();
  if (n < 20000 || w == 1) {
    body(0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + w - 1) / w;
  for (int t = 0; t < w; ++t) {
    int64_t lo = t * chunk, hi = std::min<int64_t>(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back(body, lo, hi);
  }
  for (auto &x : th) x.join();
}

}  // namespace

HostCsr csr_transpose(const HostCsr &a) {
  HostCsr t;
  t.n = a.m;
  t.m = a.n;
  t.rp.assign((size_t)a.m + 1, 0);
  for (int32_t c : a.ci) t.rp[(size_t)c + 1]++;
  for (int64_t j = 0; j < a.m; ++j) t.rp[j + 1] += t.rp[j];
  t.ci.resize(a.ci.size());
  t.v.resize(a.v.size());
  std::vector<int64_t> next(t.rp.begin(), t.rp.end() - 1);
  for (int64_t i = 0; i < a.n; ++i)
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k) {
      int64_t p = next[a.ci[k]]++;
      t.ci[p] = (int32_t)i;
      t.v[p] = a.v[k];
    }
  return t;
}

// Two-pass (symbolic count, numeric fill) row-parallel product with a
// dense sparse accumulator per worker.
HostCsr csr_multiply(const HostCsr &a, const HostCsr &b) {
  CFDLA_REQUIRE(a.m == b.n, CFDLA_ERR_DIMENSION, "sparse product: inner dimensions differ");
  HostCsr c;
  c.n = a.n;
  c.m = b.m;
  c.rp.assign((size_t)a.n + 1, 0);
  parallel_rows(a.n, [&](int64_t lo, int64_t hi) {
    std::vector<int64_t> mark((size_t)b.m, -1);
    for (int64_t i = lo; i < hi; ++i) {
      int64_t cnt = 0;
      for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k) {
        int32_t r = a.ci[k];
        for (int64_t q = b.rp[r]; q < b.rp[r + 1]; ++q)
          if (mark[b.ci[q]] != i) { mark[b.ci[q]] = i; ++cnt; }
      }
      c.rp[(size_t)i + 1] = cnt;
    }
  });
  for (int64_t i = 0; i < a.n; ++i) c.rp[i + 1] += c.rp[i];
  CFDLA_REQUIRE(c.rp[a.n] < (int64_t)INT32_MAX, CFDLA_ERR_INVALID_ARG, "sparse product: nnz exceeds int32");
  c.ci.resize((size_t)c.rp[a.n]);
  c.v.resize((size_t)c.rp[a.n]);
  parallel_rows(a.n, [&](int64_t lo, int64_t hi) {
    std::vector<int64_t> mark((size_t)b.m, -1);
    std::vector<double> acc((size_t)b.m, 0.0);
    std::vector<int32_t> cols;
    for (int64_t i = lo; i < hi; ++i) {
      cols.clear();
      for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k) {
        int32_t r = a.ci[k];
        double aik = a.v[k];
        for (int64_t q = b.rp[r]; q < b.rp[r + 1]; ++q) {
          int32_t j = b.ci[q];
          double prod = aik * b.v[q];
          if (mark[j] != i) {
            mark[j] = i;
            acc[j] = prod;
            cols.push_back(j);
          } else {
            acc[j] += prod;
          }
        }
      }
      std::sort(cols.begin(), cols.end());
      int64_t o = c.rp[i];
      for (int32_t j : cols) {
        c.ci[o] = j;
        c.v[o] = acc[j];
        ++o;
      }
    }
  });
  return c;
}

ProductPlan product_plan(const HostCsr &a, const HostCsr &b, const HostCsr &c) {
  CFDLA_REQUIRE(a.n == c.n && a.m == b.n && b.m == c.m, CFDLA_ERR_DIMENSION, "product plan: shapes differ");
  const int64_t nslots = c.nnz();
  std::vector<int64_t> ptr((size_t)nslots + 1, 0);
  // pass 1: products per slot (the slots of a row belong to one worker)
  parallel_rows(a.n, [&](int64_t lo, int64_t hi) {
    std::vector<int64_t> slot((size_t)b.m, -1);
    for (int64_t i = lo; i < hi; ++i) {
      for (int64_t k = c.rp[i]; k < c.rp[i + 1]; ++k) slot[(size_t)c.ci[k]] = k;
      for (int64_t ka = a.rp[i]; ka < a.rp[i + 1]; ++ka) {
        const int32_t r = a.ci[ka];
        for (int64_t kb = b.rp[r]; kb < b.rp[r + 1]; ++kb) ptr[(size_t)slot[(size_t)b.ci[kb]] + 1]++;
      }
    }
  });
  for (int64_t k = 0; k < nslots; ++k) ptr[(size_t)k + 1] += ptr[(size_t)k];
  CFDLA_REQUIRE(ptr.back() < (int64_t)INT32_MAX && a.nnz() < (int64_t)INT32_MAX && b.nnz() < (int64_t)INT32_MAX,
                CFDLA_ERR_INVALID_ARG, "product plan exceeds 32-bit indices");
  ProductPlan p;
  p.pa.resize((size_t)ptr.back());
  p.pb.resize((size_t)ptr.back());
  // pass 2: the pairs, per slot in the order the row traversal meets them
  std::vector<int64_t> next(ptr.begin(), ptr.end() - 1);
  parallel_rows(a.n, [&](int64_t lo, int64_t hi) {
    std::vector<int64_t> slot((size_t)b.m, -1);
    for (int64_t i = lo; i < hi; ++i) {
      for (int64_t k = c.rp[i]; k < c.rp[i + 1]; ++k) slot[(size_t)c.ci[k]] = k;
      for (int64_t ka = a.rp[i]; ka < a.rp[i + 1]; ++ka) {
        const int32_t r = a.ci[ka];
        for (int64_t kb = b.rp[r]; kb < b.rp[r + 1]; ++kb) {
          const int64_t q = next[(size_t)slot[(size_t)b.ci[kb]]]++;
          p.pa[(size_t)q] = (int32_t)ka;
          p.pb[(size_t)q] = (int32_t)kb;
        }
      }
    }
  });
  p.ptr.assign(ptr.begin(), ptr.end());
  return p;
}

// Strong connections: |a_ij| >= theta * max_{k != i} |a_ik| (diagonal excluded).
HostCsr strong_connections(const HostCsr &a, double theta) {
  HostCsr s;
  s.n = a.n;
  s.m = a.n;
  s.rp.assign((size_t)a.n + 1, 0);
  std::vector<double> rowmax((size_t)a.n, 0.0);
  for (int64_t i = 0; i < a.n; ++i) {
    double mx = 0.0;
    int64_t cnt = 0;
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k)
      if (a.ci[k] != i) mx = std::max(mx, std::fabs(a.v[k]));
    double cut = theta * mx;
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k)
      if (a.ci[k] != i && std::fabs(a.v[k]) >= cut) ++cnt;
    rowmax[i] = cut;
    s.rp[i + 1] = s.rp[i] + cnt;
  }
  s.ci.resize((size_t)s.rp[a.n]);
  s.v.resize((size_t)s.rp[a.n]);
  for (int64_t i = 0; i < a.n; ++i) {
    int64_t o = s.rp[i];
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k) {
      double mag = std::fabs(a.v[k]);
      if (a.ci[k] != i && mag >= rowmax[i]) {
        s.ci[o] = a.ci[k];
        s.v[o] = mag;
        ++o;
      }
    }
  }
  return s;
}

CoarseSplit greedy_aggregates(const HostCsr &a, int max_size) {
  CoarseSplit out;
  out.strength = strong_connections(a, 0.5);
  const HostCsr &s = out.strength;
  out.fine_to_coarse.assign((size_t)a.n, -1);
  int32_t next_id = 0;
  for (int64_t root = 0; root < a.n; ++root) {
    if (out.fine_to_coarse[root] >= 0) continue;
    out.fine_to_coarse[root] = next_id;
    int members = 1;
    for (int64_t k = s.rp[root]; k < s.rp[root + 1] && members < max_size; ++k) {
      int32_t j = s.ci[k];
      if (out.fine_to_coarse[j] < 0) {
        out.fine_to_coarse[j] = next_id;
        ++members;
      }
    }
    out.coarse_points.push_back((int32_t)root);
    ++next_id;
  }
  return out;
}

// Three-phase aggregation (Vanek, Mandel & Brezina 1996); no reference
// counterpart.  Strength as in the reference (theta times the row's largest
// off-diagonal magnitude).  Phase 1 seeds an aggregate at every point whose
// strong neighbourhood is untouched; phase 2 attaches each remaining point to
// the phase-1 aggregate of its first strong neighbour (column order); phase
// 3 seeds aggregates from what is left with its untouched strong neighbours.
CoarseSplit three_phase_aggregates(const HostCsr &a, double theta) {
  CoarseSplit out;
  out.strength = strong_connections(a, theta);
  const HostCsr &s = out.strength;
  const int64_t n = a.n;
  std::vector<int32_t> agg((size_t)n, -1);
  int32_t next_id = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free_nbhd = true;
    for (int64_t k = s.rp[i]; k < s.rp[i + 1] && free_nbhd; ++k) free_nbhd = agg[s.ci[k]] < 0;
    if (!free_nbhd) continue;
    agg[i] = next_id;
    for (int64_t k = s.rp[i]; k < s.rp[i + 1]; ++k) agg[s.ci[k]] = next_id;
    out.coarse_points.push_back((int32_t)i);
    ++next_id;
  }
  const std::vector<int32_t> phase1 = agg;
  for (int64_t i = 0; i < n; ++i) {
    if (phase1[i] >= 0) continue;
    for (int64_t k = s.rp[i]; k < s.rp[i + 1]; ++k)
      if (phase1[s.ci[k]] >= 0) {
        agg[i] = phase1[s.ci[k]];
        break;
      }
  }
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    agg[i] = next_id;
    for (int64_t k = s.rp[i]; k < s.rp[i + 1]; ++k)
      if (agg[s.ci[k]] < 0) agg[s.ci[k]] = next_id;
    out.coarse_points.push_back((int32_t)i);
    ++next_id;
  }
  out.fine_to_coarse = std::move(agg);
  return out;
}

CoarseSplit ruge_stueben_split(const HostCsr &a, double theta) {
  CoarseSplit out;
  out.strength = strong_connections(a, theta);
  const HostCsr &s = out.strength;
  HostCsr st = csr_transpose(s);
  const int64_t n = a.n;
  out.fine_to_coarse.assign((size_t)n, -1);
  std::vector<int64_t> lambda((size_t)n);
  std::vector<char> status((size_t)n, 0);  // 0 undecided, 1 C, 2 F
  // Ordered by (largest lambda, smallest index): the same point the
  // reference's linear scan with a strict '>' selects.
  std::set<std::pair<int64_t, int64_t>> queue;
  for (int64_t i = 0; i < n; ++i) {
    lambda[i] = st.rp[i + 1] - st.rp[i];
    queue.emplace(-lambda[i], i);
  }
  while (!queue.empty()) {
    auto top = *queue.begin();
    if (-top.first <= 0) break;
    int64_t i = top.second;
    queue.erase(queue.begin());
    status[i] = 1;
    out.fine_to_coarse[i] = (int32_t)out.coarse_points.size();
    out.coarse_points.push_back((int32_t)i);
    for (int64_t k = st.rp[i]; k < st.rp[i + 1]; ++k) {
      int32_t j = st.ci[k];
      if (status[j] != 0) continue;
      status[j] = 2;
      queue.erase({-lambda[j], j});
      for (int64_t m = s.rp[j]; m < s.rp[j + 1]; ++m) {
        int32_t q = s.ci[m];
        if (status[q] == 0) {
          queue.erase({-lambda[q], q});
          lambda[q] += 1;
          queue.emplace(-lambda[q], q);
        }
      }
    }
  }
  for (int64_t i = 0; i < n; ++i)
    if (status[i] == 0) {
      status[i] = 1;
      out.fine_to_coarse[i] = (int32_t)out.coarse_points.size();
      out.coarse_points.push_back((int32_t)i);
    }
  // F points without a map take the coarse index of their strongest C neighbour.
  for (int64_t i = 0; i < n; ++i) {
    if (out.fine_to_coarse[i] >= 0) continue;
    bool found = false;
    double best = 0.0;
    int32_t pick = -1;
    for (int64_t k = s.rp[i]; k < s.rp[i + 1]; ++k) {
      int32_t j = s.ci[k];
      if (status[j] != 1) continue;
      if (!found || s.v[k] > best) {
        best = s.v[k];
        pick = out.fine_to_coarse[j];
        found = true;
      }
    }
    if (found) out.fine_to_coarse[i] = pick;
  }
  return out;
}

CoarseSplit hybrid_split(const HostCsr &a, double theta) {
  CoarseSplit rs = ruge_stueben_split(a, theta);
  int64_t assigned = 0;
  for (int32_t m : rs.fine_to_coarse) assigned += (m >= 0);
  if ((double)assigned / (double)a.n > 0.8) return rs;
  return greedy_aggregates(a, 4);
}

HostCsr piecewise_constant_p(const std::vector<int32_t> &f2c, int64_t n, int64_t nc) {
  HostCsr p;
  p.n = n;
  p.m = nc;
  p.rp.assign((size_t)n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (f2c[i] >= 0) {
      p.ci.push_back(f2c[i]);
      p.v.push_back(1.0);
    }
    p.rp[i + 1] = (int64_t)p.ci.size();
  }
  return p;
}

namespace {
double entry(const HostCsr &a, int64_t r, int32_t c) {
  auto b = a.ci.begin() + a.rp[r], e = a.ci.begin() + a.rp[r + 1];
  auto it = std::lower_bound(b, e, c);
  return (it != e && *it == c) ? a.v[(size_t)(it - a.ci.begin())] : 0.0;
}
}  // namespace

HostCsr classical_p(const HostCsr &a, const std::vector<int32_t> &cpts, const HostCsr &s) {
  const double eps = 2.220446049250313e-16;
  const int64_t n = a.n;
  std::vector<int32_t> cidx((size_t)n, -1);
  for (size_t l = 0; l < cpts.size(); ++l) cidx[cpts[l]] = (int32_t)l;
  HostCsr p;
  p.n = n;
  p.m = (int64_t)cpts.size();
  p.rp.assign((size_t)n + 1, 0);
  std::vector<int32_t> strong_c, strong_f;
  std::vector<double> denom;
  std::vector<std::pair<int32_t, double>> w;
  for (int64_t i = 0; i < n; ++i) {
    if (cidx[i] >= 0) {
      p.ci.push_back(cidx[i]);
      p.v.push_back(1.0);
      p.rp[i + 1] = (int64_t)p.ci.size();
      continue;
    }
    strong_c.clear();
    strong_f.clear();
    for (int64_t k = s.rp[i]; k < s.rp[i + 1]; ++k) {
      int32_t j = s.ci[k];
      if (cidx[j] >= 0) strong_c.push_back(j);
      else if (j != i) strong_f.push_back(j);
    }
    if (!strong_c.empty()) {
      double aii = 0.0, weak = 0.0;
      auto sb = s.ci.begin() + s.rp[i], se = s.ci.begin() + s.rp[i + 1];
      for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k) {
        int32_t j = a.ci[k];
        if (j == i) aii = a.v[k];
        else if (!std::binary_search(sb, se, j)) weak += a.v[k];
      }
      double diag = aii + weak;
      if (std::fabs(diag) > eps) {
        denom.clear();
        for (int32_t f : strong_f) {
          double d = 0.0;
          for (int32_t c : strong_c) d += entry(a, f, c);
          denom.push_back(d);
        }
        double ndi = 1.0 / -diag;
        w.clear();
        for (int32_t c : strong_c) {
          double direct = entry(a, i, c), indirect = 0.0;
          for (size_t q = 0; q < strong_f.size(); ++q)
            if (std::fabs(denom[q]) > eps) indirect += entry(a, i, strong_f[q]) * entry(a, strong_f[q], c) / denom[q];
          w.emplace_back(cidx[c], (direct + indirect) * ndi);
        }
        std::stable_sort(w.begin(), w.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        for (auto &e : w) {
          p.ci.push_back(e.first);
          p.v.push_back(e.second);
        }
      }
    }
    p.rp[i + 1] = (int64_t)p.ci.size();
  }
  return p;
}

HostCsr smooth_prolongator(const HostCsr &a, const HostCsr &pt, double omega) {
  HostCsr ap = csr_multiply(a, pt);
  HostCsr p;
  p.n = a.n;
  p.m = pt.m;
  p.rp.assign((size_t)a.n + 1, 0);
  std::vector<int64_t> mark((size_t)pt.m, -1);
  std::vector<double> acc((size_t)pt.m, 0.0);
  std::vector<int32_t> cols;
  for (int64_t i = 0; i < a.n; ++i) {
    double d = entry(a, i, (int32_t)i);
    double sc = omega * (std::fabs(d) > 1e-15 ? 1.0 / d : 0.0);
    cols.clear();
    for (int64_t k = pt.rp[i]; k < pt.rp[i + 1]; ++k) {
      int32_t j = pt.ci[k];
      if (mark[j] != i) { mark[j] = i; acc[j] = pt.v[k]; cols.push_back(j); }
      else acc[j] += pt.v[k];
    }
    for (int64_t k = ap.rp[i]; k < ap.rp[i + 1]; ++k) {
      int32_t j = ap.ci[k];
      double t = -(sc * ap.v[k]);
      if (mark[j] != i) { mark[j] = i; acc[j] = t; cols.push_back(j); }
      else acc[j] += t;
    }
    std::sort(cols.begin(), cols.end());
    for (int32_t j : cols) {
      p.ci.push_back(j);
      p.v.push_back(acc[j]);
    }
    p.rp[i + 1] = (int64_t)p.ci.size();
  }
  return p;
}

std::vector<double> dense_inverse(const HostCsr &a) {
  const int64_t n = a.n;
  std::vector<double> m((size_t)(n * n), 0.0), inv((size_t)(n * n), 0.0);
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k) m[i * n + a.ci[k]] += a.v[k];
    inv[i * n + i] = 1.0;
  }
  // Gauss-Jordan elimination with partial pivoting.
  for (int64_t c = 0; c < n; ++c) {
    int64_t piv = c;
    double best = std::fabs(m[c * n + c]);
    for (int64_t r = c + 1; r < n; ++r)
      if (std::fabs(m[r * n + c]) > best) { best = std::fabs(m[r * n + c]); piv = r; }
    CFDLA_REQUIRE(best > 0.0 && std::isfinite(best), CFDLA_ERR_NUMERICAL,
                  "AMG coarsest operator is singular; use coarse_solver = reference sweeps");
    if (piv != c)
      for (int64_t j = 0; j < n; ++j) {
        std::swap(m[c * n + j], m[piv * n + j]);
        std::swap(inv[c * n + j], inv[piv * n + j]);
      }
    double d = 1.0 / m[c * n + c];
    for (int64_t j = 0; j < n; ++j) { m[c * n + j] *= d; inv[c * n + j] *= d; }
    for (int64_t r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = m[r * n + c];
      if (f == 0.0) continue;
      for (int64_t j = 0; j < n; ++j) { m[r * n + j] -= f * m[c * n + j]; inv[r * n + j] -= f * inv[c * n + j]; }
    }
  }
  return inv;
}

void build_hierarchy(const HostCsr &fine, const cfdla_amg_config &cfg, HostHierarchy &h, const GalerkinFn &galerkin) {
  h.levels.clear();
  h.levels.emplace_back();
  h.levels[0].A = fine;
  const int max_levels = std::min<int>(cfg.max_levels, CFDLA_MAX_LEVELS);
  while ((int)h.levels.size() < max_levels && h.levels.back().A.n > cfg.min_coarse_size) {
    HostLevel &cur = h.levels.back();
    CoarseSplit split;
    if (cfg.coarsening == 1) split = greedy_aggregates(cur.A, cfg.max_aggregate_size > 0 ? cfg.max_aggregate_size : 8);
    else if (cfg.coarsening == 2) split = hybrid_split(cur.A, cfg.strength_threshold);
    else if (cfg.coarsening == 3) split = three_phase_aggregates(cur.A, cfg.strength_threshold);
    else split = ruge_stueben_split(cur.A, cfg.strength_threshold);
    const int64_t nc = (int64_t)split.coarse_points.size();
    HostCsr P;
    if (cfg.interpolation == 0) P = classical_p(cur.A, split.coarse_points, split.strength);
    else if (cfg.interpolation == 1) P = piecewise_constant_p(split.fine_to_coarse, cur.A.n, nc);
    else P = smooth_prolongator(cur.A, piecewise_constant_p(split.fine_to_coarse, cur.A.n, nc), cfg.smoothing_omega);
    HostCsr R = csr_transpose(P);
    HostCsr Ac = galerkin ? galerkin(h.levels.size() - 1, R, cur.A, P) : csr_multiply(csr_multiply(R, cur.A), P);
    cur.P = std::move(P);
    cur.R = std::move(R);
    cur.fine_to_coarse = std::move(split.fine_to_coarse);
    if (Ac.n == 0) break;
    h.levels.emplace_back();
    h.levels.back().A = std::move(Ac);
  }
}

void recompute_hierarchy(HostHierarchy &h, const double *fine_values) {
  HostCsr &f = h.levels[0].A;
  std::copy(fine_values, fine_values + f.nnz(), f.v.begin());
  for (size_t l = 0; l + 1 < h.levels.size(); ++l) {
    HostCsr ac = csr_multiply(csr_multiply(h.levels[l].R, h.levels[l].A), h.levels[l].P);
    h.levels[l + 1].A = std::move(ac);
  }
}

}  // namespace cfdla
