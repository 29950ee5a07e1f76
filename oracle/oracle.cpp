// CPU FP64 oracle for the CFDrs sparse linear-solve path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: it restates, in plain
// sequential C++, the reference's algorithms for CSR assembly, AMG setup and
// V-cycle, the smoothers, the Krylov recurrences and the 1D network solve
// cascade.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
// bench.py load it (through oracle/oracle.py).  The product library
// (cfdrs_amd/libcfdla.so) never links, loads or calls it.
//
// Every function cites the reference file:line it follows.  Paths are
// relative to the reference repository root.
//
// Parity pinning (each function is checked against something outside this
// file; see tests/test_oracle_golden.py):
//   strength / aggregation / direct P / transpose / SpGEMM: the reference's
//     known answers (coarsening/mod.rs:108-122, interpolation.rs:471-536)
//     and brute-force dense products on tiny inputs.
//   classical interpolation: interpolation.rs:490-536 (0.5/0.5) and
//     :581-643 (2/3, 1/3).
//   smoothers: smoothers.rs:418-491 exact vectors.
//   CG / BiCGSTAB / GMRES: the recurrences live in the un-vendored athena
//     crate; pinned at solution level by closed forms (diagonal systems,
//     exact termination on tiny SPD systems) and numpy dense solves.
//   network assembly: matrix_assembly.rs:335-455 known answers.
//   rectangular resistance: rectangular.rs:268-300 (Po(1) ~ 56.91).
//   adjoint gradients: central finite differences of the oracle's own
//     forward solve (tests/test_oracle_golden.py).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

namespace {

using i64 = int64_t;

struct Csr {
  i64 n = 0, m = 0;
  std::vector<i64> rp, ci;
  std::vector<double> v;
  i64 nnz() const { return (i64)ci.size(); }
};

// ---------------------------------------------------------------------------
// Builder: duplicates summed in insertion order, rows sorted by column.
// crates/cfd-math/src/sparse/builder.rs:303-324 (build) and :213-222 (sort).
// ---------------------------------------------------------------------------
struct Builder {
  i64 n, m;
  std::vector<std::vector<std::pair<i64, double>>> rows;
  Builder(i64 n_, i64 m_) : n(n_), m(m_), rows((size_t)n_) {}
  void add(i64 r, i64 c, double val) { rows[(size_t)r].emplace_back(c, val); }
  Csr build() {
    Csr a;
    a.n = n; a.m = m;
    a.rp.assign((size_t)n + 1, 0);
    for (i64 r = 0; r < n; ++r) {
      auto &row = rows[(size_t)r];
      // stable: duplicates keep insertion order; the first is taken as is
      // (or_insert) and later ones are added (and_modify)
      std::stable_sort(row.begin(), row.end(),
                       [](const std::pair<i64, double> &x, const std::pair<i64, double> &y) { return x.first < y.first; });
      for (size_t k = 0; k < row.size(); ++k) {
        if (k > 0 && row[k].first == row[k - 1].first) a.v.back() += row[k].second;
        else { a.ci.push_back(row[k].first); a.v.push_back(row[k].second); }
      }
      a.rp[(size_t)r + 1] = (i64)a.ci.size();
      std::vector<std::pair<i64, double>>().swap(row);
    }
    return a;
  }
};

double csr_value(const Csr &a, i64 r, i64 c) {
  // multigrid/mod.rs:206-216 (binary search on the sorted row)
  auto b = a.ci.begin() + a.rp[(size_t)r], e = a.ci.begin() + a.rp[(size_t)r + 1];
  auto it = std::lower_bound(b, e, c);
  if (it != e && *it == c) return a.v[(size_t)(it - a.ci.begin())];
  return 0.0;
}

// Threads for the solve-phase loops (row loops of products and sweeps,
// vector updates, dot products).  1 (the default) keeps every loop serial and
// every sum in index order; more threads reorder only the dot-product sums.
int g_threads = 1;

void spmv(const Csr &a, const double *x, double *y) {
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
  for (i64 i = 0; i < a.n; ++i) {
    double s = 0.0;
    for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) s += a.v[k] * x[a.ci[k]];
    y[i] = s;
  }
}

double dot(const std::vector<double> &a, const std::vector<double> &b) {
  double s = 0.0;
  const i64 n = (i64)a.size();
  if (g_threads > 1) {
#pragma omp parallel for schedule(static) reduction(+ : s) num_threads(g_threads)
    for (i64 i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  }
  for (i64 i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
double nrm2(const std::vector<double> &a) { return std::sqrt(dot(a, a)); }

// ---------------------------------------------------------------------------
// Transpose with sorted columns (leto CsrMatrix::transpose, used at
// multigrid/amg.rs:349 and coarsening/algorithms.rs:45).
// ---------------------------------------------------------------------------
Csr transpose(const Csr &a) {
  Csr t;
  t.n = a.m; t.m = a.n;
  t.rp.assign((size_t)a.m + 1, 0);
  for (i64 k = 0; k < a.nnz(); ++k) t.rp[a.ci[k] + 1]++;
  for (i64 j = 0; j < a.m; ++j) t.rp[j + 1] += t.rp[j];
  t.ci.resize(a.ci.size());
  t.v.resize(a.v.size());
  std::vector<i64> pos(t.rp.begin(), t.rp.end() - 1);
  for (i64 i = 0; i < a.n; ++i)
    for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) {
      i64 p = pos[a.ci[k]]++;
      t.ci[p] = i;
      t.v[p] = a.v[k];
    }
  return t;
}

// ---------------------------------------------------------------------------
// Sparse product C = A*B keeping every structurally produced entry
// (explicit zeros kept), accumulation in row-of-A then row-of-B order.
// Stands in for ext leto_ops::spgemm at multigrid/amg.rs:63-69,352-353.
// ---------------------------------------------------------------------------
Csr spgemm(const Csr &a, const Csr &b) {
  Builder bd(a.n, b.m);
  for (i64 i = 0; i < a.n; ++i)
    for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) {
      i64 r = a.ci[k];
      double aik = a.v[k];
      for (i64 q = b.rp[r]; q < b.rp[r + 1]; ++q) bd.add(i, b.ci[q], aik * b.v[q]);
    }
  return bd.build();
}

// ---------------------------------------------------------------------------
// Strength of connection, coarsening/algorithms.rs:395-436.
// ---------------------------------------------------------------------------
Csr strength(const Csr &a, double theta) {
  Csr s;
  s.n = a.n; s.m = a.n;
  s.rp.assign((size_t)a.n + 1, 0);
  for (i64 i = 0; i < a.n; ++i) {
    double mx = 0.0;
    for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k)
      if (a.ci[k] != i) { double v = std::fabs(a.v[k]); mx = (mx > v) ? mx : v; }
    for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k)
      if (a.ci[k] != i) {
        double st = std::fabs(a.v[k]);
        if (st >= theta * mx) { s.ci.push_back(a.ci[k]); s.v.push_back(st); }
      }
    s.rp[i + 1] = (i64)s.ci.size();
  }
  return s;
}

struct Coarsening {
  std::vector<i64> coarse_points;
  std::vector<i64> f2c;  // -1 = None
  Csr S;
};

// coarsening/algorithms.rs:129-172 (aggregation; theta fixed at 0.5, :138)
Coarsening aggregation(const Csr &a, i64 max_size) {
  Coarsening r;
  r.S = strength(a, 0.5);
  r.f2c.assign((size_t)a.n, -1);
  std::vector<char> aggregated((size_t)a.n, 0);
  i64 agg_id = 0;
  std::vector<i64> members;
  for (i64 i = 0; i < a.n; ++i) {
    if (aggregated[i]) continue;
    members.clear();
    members.push_back(i);
    aggregated[i] = 1;
    for (i64 k = r.S.rp[i]; k < r.S.rp[i + 1]; ++k) {
      i64 j = r.S.ci[k];
      if (!aggregated[j] && (i64)members.size() < max_size) { members.push_back(j); aggregated[j] = 1; }
    }
    r.coarse_points.push_back(members[0]);
    for (i64 p : members) r.f2c[p] = agg_id;
    ++agg_id;
  }
  return r;
}

// Three-phase aggregation (Vanek, Mandel & Brezina 1996).  No counterpart in
// the reference; the strength matrix is the reference's
// (coarsening/algorithms.rs:395-436) with a configurable theta.
//   1. i untouched and all strong neighbours untouched -> aggregate {i} + S_i
//   2. remaining i joins the phase-1 aggregate of its first strong neighbour
//   3. remaining i seeds {i} + its untouched strong neighbours
Coarsening standard_aggregation(const Csr &a, double theta) {
  Coarsening r;
  r.S = strength(a, theta);
  i64 n = a.n;
  r.f2c.assign((size_t)n, -1);
  i64 id = 0;
  for (i64 i = 0; i < n; ++i) {
    if (r.f2c[i] != -1) continue;
    bool ok = true;
    for (i64 k = r.S.rp[i]; k < r.S.rp[i + 1]; ++k)
      if (r.f2c[r.S.ci[k]] != -1) ok = false;
    if (!ok) continue;
    r.f2c[i] = id;
    for (i64 k = r.S.rp[i]; k < r.S.rp[i + 1]; ++k) r.f2c[r.S.ci[k]] = id;
    r.coarse_points.push_back(i);
    ++id;
  }
  std::vector<i64> first = r.f2c;
  for (i64 i = 0; i < n; ++i) {
    if (first[i] != -1) continue;
    for (i64 k = r.S.rp[i]; k < r.S.rp[i + 1]; ++k) {
      i64 j = r.S.ci[k];
      if (first[j] != -1) { r.f2c[i] = first[j]; break; }
    }
  }
  for (i64 i = 0; i < n; ++i) {
    if (r.f2c[i] != -1) continue;
    r.f2c[i] = id;
    for (i64 k = r.S.rp[i]; k < r.S.rp[i + 1]; ++k)
      if (r.f2c[r.S.ci[k]] == -1) r.f2c[r.S.ci[k]] = id;
    r.coarse_points.push_back(i);
    ++id;
  }
  return r;
}

// coarsening/algorithms.rs:200-230
void assign_closest(std::vector<i64> &f2c, const std::vector<int> &status, const Csr &S) {
  for (i64 i = 0; i < S.n; ++i) {
    if (f2c[i] != -1) continue;
    double best = 0.0;
    bool have = false;
    i64 best_idx = -1;
    for (i64 k = S.rp[i]; k < S.rp[i + 1]; ++k) {
      i64 j = S.ci[k];
      if (status[j] == 1) {
        double st = S.v[k];
        if (!have || st > best) { best = st; best_idx = f2c[j]; have = true; }
      }
    }
    if (have && best_idx != -1) f2c[i] = best_idx;
  }
}

// coarsening/algorithms.rs:34-126 (Ruge-Stueben first pass, O(n^2) scan)
Coarsening ruge_stueben(const Csr &a, double theta) {
  Coarsening r;
  i64 n = a.n;
  r.S = strength(a, theta);
  r.f2c.assign((size_t)n, -1);
  Csr St = transpose(r.S);
  std::vector<i64> lambda((size_t)n);
  for (i64 i = 0; i < n; ++i) lambda[i] = St.rp[i + 1] - St.rp[i];
  std::vector<int> status((size_t)n, 0);
  i64 undecided = n;
  while (undecided > 0) {
    i64 max_lambda = -1, iopt = -1;
    for (i64 k = 0; k < n; ++k)
      if (status[k] == 0 && lambda[k] > max_lambda) { max_lambda = lambda[k]; iopt = k; }
    if (iopt < 0) break;
    if (max_lambda <= 0) break;
    i64 i = iopt;
    status[i] = 1;
    --undecided;
    r.coarse_points.push_back(i);
    r.f2c[i] = (i64)r.coarse_points.size() - 1;
    for (i64 k = St.rp[i]; k < St.rp[i + 1]; ++k) {
      i64 j = St.ci[k];
      if (status[j] == 0) {
        status[j] = 2;
        --undecided;
        for (i64 m = r.S.rp[j]; m < r.S.rp[j + 1]; ++m) {
          i64 kk = r.S.ci[m];
          if (status[kk] == 0) lambda[kk] += 1;
        }
      }
    }
  }
  for (i64 i = 0; i < n; ++i)
    if (status[i] == 0) {
      status[i] = 1;
      r.coarse_points.push_back(i);
      r.f2c[i] = (i64)r.coarse_points.size() - 1;
    }
  assign_closest(r.f2c, status, r.S);
  return r;
}

// coarsening/algorithms.rs:175-197 (hybrid: RS unless <= 80% assigned)
Coarsening hybrid(const Csr &a, double theta, i64 max_size) {
  Coarsening rs = ruge_stueben(a, theta);
  i64 assigned = 0;
  for (i64 x : rs.f2c) assigned += (x != -1);
  if ((double)assigned / (double)a.n > 0.8) return rs;
  return aggregation(a, max_size);
}

// interpolation.rs:191-217
Csr direct_interp(const std::vector<i64> &f2c, i64 n, i64 nc) {
  Csr p;
  p.n = n; p.m = nc;
  p.rp.assign((size_t)n + 1, 0);
  for (i64 i = 0; i < n; ++i) {
    if (f2c[i] != -1) { p.ci.push_back(f2c[i]); p.v.push_back(1.0); }
    p.rp[i + 1] = (i64)p.ci.size();
  }
  return p;
}

// interpolation.rs:38-188 (classical Ruge-Stueben weights)
Csr classical_interp(const Csr &a, const std::vector<i64> &cpts, const Csr &S) {
  const double EPS = 2.220446049250313e-16;  // f64::EPSILON
  i64 n = a.n, nc = (i64)cpts.size();
  std::vector<i64> cmap((size_t)n, -1);
  for (i64 l = 0; l < nc; ++l)
    if (cpts[l] < n) cmap[cpts[l]] = l;
  Csr p;
  p.n = n; p.m = nc;
  p.rp.assign((size_t)n + 1, 0);
  for (i64 i = 0; i < n; ++i) {
    if (cmap[i] != -1) {
      p.ci.push_back(cmap[i]);
      p.v.push_back(1.0);
    } else {
      i64 rs = a.rp[i], re = a.rp[i + 1];
      i64 ss = S.rp[i], se = S.rp[i + 1];
      std::vector<i64> ci_set, fs_set;
      for (i64 k = ss; k < se; ++k) {
        i64 nb = S.ci[k];
        if (cmap[nb] != -1) ci_set.push_back(nb);
        else if (nb != i) fs_set.push_back(nb);
      }
      if (!ci_set.empty()) {
        double aii = 0.0, sum_weak = 0.0;
        for (i64 k = rs; k < re; ++k) {
          i64 nb = a.ci[k];
          double val = a.v[k];
          if (nb == i) aii = val;
          else if (!std::binary_search(S.ci.begin() + ss, S.ci.begin() + se, nb)) sum_weak += val;
        }
        double diagonal = aii + sum_weak;
        if (std::fabs(diagonal) > EPS) {
          std::vector<double> kden;
          for (i64 k : fs_set) {
            double d = 0.0;
            for (i64 m : ci_set) d += csr_value(a, k, m);
            kden.push_back(d);
          }
          std::vector<std::pair<i64, double>> w;
          double neg_diag_inv = 1.0 / -diagonal;
          for (i64 j : ci_set) {
            double aij = csr_value(a, i, j);
            double ind = 0.0;
            for (size_t q = 0; q < fs_set.size(); ++q) {
              double den = kden[q];
              if (std::fabs(den) > EPS) {
                double aik = csr_value(a, i, fs_set[q]);
                double akj = csr_value(a, fs_set[q], j);
                ind += aik * akj / den;
              }
            }
            w.emplace_back(cmap[j], (aij + ind) * neg_diag_inv);
          }
          std::stable_sort(w.begin(), w.end(),
                           [](const std::pair<i64, double> &x, const std::pair<i64, double> &y) {
                             return x.first < y.first;
                           });
          for (auto &e : w) { p.ci.push_back(e.first); p.v.push_back(e.second); }
        }
      }
    }
    p.rp[i + 1] = (i64)p.ci.size();
  }
  return p;
}

// Prolongator smoothing P = (I - omega D^-1 A) P_tent.  No counterpart in the
// reference (it has only the tentative P of interpolation.rs:191-217); this
// is the smoothed-aggregation variant.  Rows with |a_ii| <= 1e-15 are left
// unsmoothed (the smoother's own skip rule, smoothers.rs:188).
Csr smooth_interp(const Csr &a, const Csr &pt, double omega) {
  Csr ap = spgemm(a, pt);
  std::vector<double> dinv((size_t)a.n, 0.0);
  for (i64 i = 0; i < a.n; ++i) {
    double d = csr_value(a, i, i);
    dinv[i] = std::fabs(d) > 1e-15 ? 1.0 / d : 0.0;
  }
  Builder bd(a.n, pt.m);
  for (i64 i = 0; i < a.n; ++i) {
    for (i64 k = pt.rp[i]; k < pt.rp[i + 1]; ++k) bd.add(i, pt.ci[k], pt.v[k]);
    double s = omega * dinv[i];
    for (i64 k = ap.rp[i]; k < ap.rp[i + 1]; ++k) bd.add(i, ap.ci[k], -(s * ap.v[k]));
  }
  return bd.build();
}

// ---------------------------------------------------------------------------
// Smoothers, smoothers.rs.
// ---------------------------------------------------------------------------
enum SmootherKind { SM_GS = 0, SM_SGS = 1, SM_JACOBI = 2, SM_SOR = 3, SM_CHEBYSHEV = 4 };

void gs_forward(const Csr &a, double *x, const double *b, i64 iters, double w) {
  // smoothers.rs:42-82
  for (i64 it = 0; it < iters; ++it)
    for (i64 i = 0; i < a.n; ++i) {
      double sum = 0.0, diag = 0.0;
      for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) {
        if (a.ci[k] == i) diag = a.v[k];
        else sum += a.v[k] * x[a.ci[k]];
      }
      if (std::fabs(diag) > 1e-15) { double nv = (b[i] - sum) / diag; x[i] = x[i] + w * (nv - x[i]); }
    }
}

void sgs(const Csr &a, double *x, const double *b, i64 iters, double w) {
  // smoothers.rs:97-142: all forward sweeps, then all backward sweeps
  gs_forward(a, x, b, iters, w);
  for (i64 it = 0; it < iters; ++it)
    for (i64 i = a.n - 1; i >= 0; --i) {
      double sum = 0.0, diag = 0.0;
      for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) {
        if (a.ci[k] == i) diag = a.v[k];
        else sum += a.v[k] * x[a.ci[k]];
      }
      if (std::fabs(diag) > 1e-15) { double nv = (b[i] - sum) / diag; x[i] = x[i] + w * (nv - x[i]); }
    }
}

void jacobi(const Csr &a, double *x, const double *b, i64 iters, double w) {
  // smoothers.rs:157-199
  std::vector<double> xo((size_t)a.n);
  for (i64 it = 0; it < iters; ++it) {
    std::copy(x, x + a.n, xo.begin());
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (i64 i = 0; i < a.n; ++i) {
      double sum = 0.0, diag = 0.0;
      for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) {
        if (a.ci[k] == i) diag = a.v[k];
        else sum += a.v[k] * xo[a.ci[k]];
      }
      if (std::fabs(diag) > 1e-15) { double nv = (b[i] - sum) / diag; x[i] = xo[i] + w * (nv - xo[i]); }
    }
  }
}

// smoothers.rs:275-311 (Gershgorin bounds over the first 100 rows)
void chebyshev_bounds(const Csr &a, double *lo, double *hi) {
  double mn = 1e6, mx = 0.0;
  for (i64 i = 0; i < std::min<i64>(a.n, 100); ++i) {
    double rs = 0.0, d = 0.0;
    for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) {
      double v = std::fabs(a.v[k]);
      if (a.ci[k] == i) d = v; else rs += v;
    }
    double lower = d > rs ? d - rs : 0.1;
    double upper = d + rs;
    mn = std::min(mn, lower);
    mx = std::max(mx, upper);
  }
  *lo = std::max(mn, 0.1);
  *hi = std::max(mx, 1.0);
}

// smoothers.rs:317-372
void chebyshev(const Csr &a, double *x, const double *b, i64 iters, double emin, double emax, i64 degree) {
  double theta = (emax + emin) / 2.0, delta = (emax - emin) / 2.0;
  double sigma = std::fabs(delta) < 1e-10 ? 0.0 : theta / delta;
  double rho_old = sigma == 0.0 ? 0.0 : 1.0 / sigma;
  std::vector<double> r((size_t)a.n), ax((size_t)a.n);
  auto resid = [&]() { spmv(a, x, ax.data()); for (i64 i = 0; i < a.n; ++i) r[i] = b[i] - ax[i]; };
  resid();
  double alpha = sigma == 0.0 ? 1.0 / emax : (1.0 / (2.0 * sigma - rho_old)) / rho_old;
  for (i64 it = 0; it < std::min(iters, degree); ++it) {
    for (i64 i = 0; i < a.n; ++i) {
      double d = 0.0;
      for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) if (a.ci[k] == i) { d = a.v[k]; break; }
      if (std::fabs(d) > 1e-15) x[i] += alpha * r[i] / d;
    }
    resid();
  }
}

// ---------------------------------------------------------------------------
// Dense LU with partial pivoting (stands in for ext leto lu_decompose,
// used at cfd-1d/src/solver/core/linear_system.rs:248-267 and as the
// coarsest-level direct solve).
// ---------------------------------------------------------------------------
bool dense_lu_solve(i64 n, std::vector<double> a /*row-major copy*/, std::vector<double> &x) {
  std::vector<i64> piv((size_t)n);
  for (i64 k = 0; k < n; ++k) {
    i64 p = k;
    double best = std::fabs(a[k * n + k]);
    for (i64 i = k + 1; i < n; ++i)
      if (std::fabs(a[i * n + k]) > best) { best = std::fabs(a[i * n + k]); p = i; }
    if (best == 0.0 || !std::isfinite(best)) return false;
    if (p != k) {
      for (i64 j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
      std::swap(x[k], x[p]);
    }
    for (i64 i = k + 1; i < n; ++i) {
      double f = a[i * n + k] / a[k * n + k];
      if (f == 0.0) continue;
      for (i64 j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
      x[i] -= f * x[k];
    }
  }
  for (i64 i = n - 1; i >= 0; --i) {
    double s = x[i];
    for (i64 j = i + 1; j < n; ++j) s -= a[i * n + j] * x[j];
    x[i] = s / a[i * n + i];
  }
  for (i64 i = 0; i < n; ++i) if (!std::isfinite(x[i])) return false;
  return true;
}

std::vector<double> to_dense(const Csr &a) {
  std::vector<double> d((size_t)(a.n * a.m), 0.0);
  for (i64 i = 0; i < a.n; ++i)
    for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) d[i * a.m + a.ci[k]] += a.v[k];
  return d;
}

// ---------------------------------------------------------------------------
// AMG, multigrid/amg.rs and multigrid/mod.rs:219-260.
// ---------------------------------------------------------------------------
enum Coarsen { CO_RS = 0, CO_AGG = 1, CO_HYBRID = 2, CO_STD = 3 };
enum Interp { IN_CLASSICAL = 0, IN_DIRECT = 1, IN_SMOOTHED = 2 };
enum CoarseSolve { CS_SWEEPS = 0, CS_DIRECT = 1 };

struct AmgConfig {
  i64 max_levels = 10, min_coarse = 50;
  int coarsening = CO_RS, interp = IN_CLASSICAL, smoother = SM_GS;
  i64 pre = 2, post = 2;
  double omega = 1.0, theta = 0.25;
  int coarse_solve = CS_SWEEPS;
  double sa_omega = 2.0 / 3.0;
  i64 max_agg = 8;
};

struct Level {
  Csr A, R, P;
  bool has_transfer = false;
  std::vector<i64> f2c;  // aggregation/coarsening map of this level (-1 = None)
  std::vector<i64> cpts;
};

struct Amg {
  AmgConfig cfg;
  std::vector<Level> levels;
  std::vector<double> coarse_dense;  // dense copy of the coarsest operator
  double op_complexity = 1.0, grid_complexity = 1.0;
};

void amg_build(Amg &h, const Csr &A) {
  // amg.rs:217-288 (setup) and :290-379 (build_next_level)
  h.levels.clear();
  h.levels.push_back(Level{A, {}, {}, false, {}, {}});
  while ((i64)h.levels.size() < h.cfg.max_levels && h.levels.back().A.n > h.cfg.min_coarse) {
    Level &cur = h.levels.back();
    Coarsening co;
    switch (h.cfg.coarsening) {
      case CO_AGG: co = aggregation(cur.A, h.cfg.max_agg); break;
      case CO_HYBRID: co = hybrid(cur.A, h.cfg.theta, 4); break;
      case CO_STD: co = standard_aggregation(cur.A, h.cfg.theta); break;
      default: co = ruge_stueben(cur.A, h.cfg.theta); break;
    }
    Csr P;
    i64 nc = (i64)co.coarse_points.size();
    if (h.cfg.interp == IN_CLASSICAL) P = classical_interp(cur.A, co.coarse_points, co.S);
    else if (h.cfg.interp == IN_DIRECT) P = direct_interp(co.f2c, cur.A.n, nc);
    else P = smooth_interp(cur.A, direct_interp(co.f2c, cur.A.n, nc), h.cfg.sa_omega);
    Csr R = transpose(P);
    Csr Ac = spgemm(spgemm(R, cur.A), P);
    cur.R = std::move(R);
    cur.P = std::move(P);
    cur.has_transfer = true;
    cur.f2c = co.f2c;
    cur.cpts = co.coarse_points;
    if (Ac.n == 0) break;
    h.levels.push_back(Level{std::move(Ac), {}, {}, false, {}, {}});
  }
  h.coarse_dense = to_dense(h.levels.back().A);
  // amg.rs:394-420 (complexities; entries with |a| > 1e-15 counted)
  auto cnt = [](const Csr &a) { i64 c = 0; for (double x : a.v) c += std::fabs(x) > 1e-15; return c; };
  i64 fine = cnt(h.levels[0].A), tot = 0, vars = 0;
  for (auto &l : h.levels) { tot += cnt(l.A); vars += l.A.n; }
  h.op_complexity = (double)tot / (double)fine;
  h.grid_complexity = (double)vars / (double)h.levels[0].A.n;
}

void smooth(const AmgConfig &c, const Csr &a, double *x, const double *b, i64 iters) {
  // amg.rs:382-391: SGS, Jacobi, everything else -> Gauss-Seidel
  if (c.smoother == SM_SGS) sgs(a, x, b, iters, c.omega);
  else if (c.smoother == SM_JACOBI) jacobi(a, x, b, iters, c.omega);
  else gs_forward(a, x, b, iters, c.omega);
}

void v_cycle(const Amg &h, size_t l, const double *b, double *x) {
  // amg.rs:423-503
  const Level &lv = h.levels[l];
  i64 n = lv.A.n;
  if (l + 1 < h.levels.size()) {
    smooth(h.cfg, lv.A, x, b, h.cfg.pre);
    std::vector<double> ax((size_t)n), r((size_t)n);
    spmv(lv.A, x, ax.data());
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (i64 i = 0; i < n; ++i) r[i] = b[i] - ax[i];
    i64 nc = lv.R.n;
    std::vector<double> rc((size_t)nc), xc((size_t)nc, 0.0), e((size_t)n);
    spmv(lv.R, r.data(), rc.data());
    v_cycle(h, l + 1, rc.data(), xc.data());
    spmv(lv.P, xc.data(), e.data());
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (i64 i = 0; i < n; ++i) x[i] += e[i];
    smooth(h.cfg, lv.A, x, b, h.cfg.post);
  } else if (h.cfg.coarse_solve == CS_DIRECT) {
    std::vector<double> sol(b, b + n);
    if (dense_lu_solve(n, h.coarse_dense, sol)) std::copy(sol.begin(), sol.end(), x);
    else smooth(h.cfg, lv.A, x, b, 10);
  } else {
    // the coarsest level is pre-smoothed then 10 more sweeps (amg.rs:430-435,501)
    smooth(h.cfg, lv.A, x, b, h.cfg.pre);
    smooth(h.cfg, lv.A, x, b, 10);
  }
}

void amg_apply(const Amg &h, const double *r, double *z) {
  // amg.rs:506-558: z.fill(0) then one V-cycle
  std::fill(z, z + h.levels[0].A.n, 0.0);
  v_cycle(h, 0, r, z);
}

// ---------------------------------------------------------------------------
// Krylov recurrences.  The reference delegates these to the un-vendored
// athena crate (crates/cfd-math/src/linear_solver/krylov.rs:150-373); the
// restatement is textbook (Saad 2003, Alg. 9.1 PCG, Alg. 7.7 BiCGSTAB with
// right preconditioning, Alg. 9.5 right-preconditioned GMRES(m) with
// modified Gram-Schmidt and Givens rotations).  Termination codes mirror
// athena's Termination as classified in krylov.rs:375-400.
// ---------------------------------------------------------------------------
enum Term { T_CONVERGED = 0, T_INITIAL = 1, T_MAXIT = 2, T_BREAKDOWN = 3, T_NONPOS = 4 };

struct Report {
  i64 iterations;
  double initial_residual_norm, final_residual_norm, threshold;
  i64 termination;
};

enum PrecKind { PK_IDENTITY = 0, PK_JACOBI = 1, PK_AMG = 2 };

struct Prec {
  int kind = PK_IDENTITY;
  std::vector<double> inv_diag;
  const Amg *amg = nullptr;
  void apply(const double *r, double *z, i64 n) const {
    if (kind == PK_JACOBI) for (i64 i = 0; i < n; ++i) z[i] = r[i] * inv_diag[i];
    else if (kind == PK_AMG) amg_apply(*amg, r, z);
    else std::copy(r, r + n, z);
  }
};

// Jacobi inverse diagonal: identity where |d| <= eps
// (cfd-1d/src/solver/core/linear_system.rs:270-299; eps = f64::EPSILON).
std::vector<double> jacobi_inv_diag(const Csr &a) {
  std::vector<double> inv((size_t)a.n, 1.0);
  for (i64 i = 0; i < a.n; ++i) {
    double d = csr_value(a, i, i);
    inv[i] = std::fabs(d) <= 2.220446049250313e-16 ? 1.0 : 1.0 / d;
  }
  return inv;
}

std::vector<double> true_residual(const Csr &a, const double *b, const double *x) {
  std::vector<double> r((size_t)a.n);
  spmv(a, x, r.data());
  for (i64 i = 0; i < a.n; ++i) r[i] = b[i] - r[i];
  return r;
}

Report pcg(const Csr &a, const Prec &M, const double *b, double *x, double tol_abs, double tol_rel, i64 max_it) {
  i64 n = a.n;
  std::vector<double> bv(b, b + n);
  double thr = std::max(tol_abs, tol_rel * nrm2(bv));
  std::vector<double> r = true_residual(a, b, x), z((size_t)n), p((size_t)n), q((size_t)n);
  Report rep{0, nrm2(r), 0.0, thr, T_MAXIT};
  double rn = rep.initial_residual_norm;
  if (rn <= thr) { rep.final_residual_norm = rn; rep.termination = T_INITIAL; return rep; }
  M.apply(r.data(), z.data(), n);
  p = z;
  double rz = dot(r, z);
  for (i64 it = 1; it <= max_it; ++it) {
    spmv(a, p.data(), q.data());
    double pq = dot(p, q);
    if (!(pq > 0.0)) { rep.iterations = it - 1; rep.final_residual_norm = rn; rep.termination = T_NONPOS; return rep; }
    double alpha = rz / pq;
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (i64 i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    rn = nrm2(r);
    rep.iterations = it;
    if (rn <= thr) {
      std::vector<double> tr = true_residual(a, b, x);
      double tn = nrm2(tr);
      if (tn <= thr) { rep.final_residual_norm = tn; rep.termination = T_CONVERGED; return rep; }
      r = tr; rn = tn;  // residual replacement and restart of the direction
      M.apply(r.data(), z.data(), n);
      p = z;
      rz = dot(r, z);
      continue;
    }
    M.apply(r.data(), z.data(), n);
    double rz_new = dot(r, z);
    if (rz_new == 0.0 || !std::isfinite(rz_new)) { rep.final_residual_norm = rn; rep.termination = T_BREAKDOWN; return rep; }
    double beta = rz_new / rz;
    rz = rz_new;
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
    for (i64 i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  rep.final_residual_norm = nrm2(true_residual(a, b, x));
  rep.termination = T_MAXIT;
  return rep;
}

Report bicgstab(const Csr &a, const Prec &M, const double *b, double *x, double tol_abs, double tol_rel, i64 max_it) {
  i64 n = a.n;
  std::vector<double> bv(b, b + n);
  double thr = std::max(tol_abs, tol_rel * nrm2(bv));
  std::vector<double> r = true_residual(a, b, x);
  Report rep{0, nrm2(r), 0.0, thr, T_MAXIT};
  double rn = rep.initial_residual_norm;
  if (rn <= thr) { rep.final_residual_norm = rn; rep.termination = T_INITIAL; return rep; }
  std::vector<double> rhat = r, p((size_t)n, 0.0), v((size_t)n, 0.0), ph((size_t)n), s((size_t)n), sh((size_t)n), t((size_t)n);
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  for (i64 it = 1; it <= max_it; ++it) {
    double rho_new = dot(rhat, r);
    if (rho_new == 0.0 || !std::isfinite(rho_new)) { rep.iterations = it - 1; rep.final_residual_norm = rn; rep.termination = T_BREAKDOWN; return rep; }
    double beta = (rho_new / rho) * (alpha / omega);
    for (i64 i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    M.apply(p.data(), ph.data(), n);
    spmv(a, ph.data(), v.data());
    double rv = dot(rhat, v);
    if (rv == 0.0 || !std::isfinite(rv)) { rep.iterations = it - 1; rep.final_residual_norm = rn; rep.termination = T_BREAKDOWN; return rep; }
    alpha = rho_new / rv;
    for (i64 i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    rep.iterations = it;
    if (nrm2(s) <= thr) {
      for (i64 i = 0; i < n; ++i) x[i] += alpha * ph[i];
      std::vector<double> tr = true_residual(a, b, x);
      double tn = nrm2(tr);
      if (tn <= thr) { rep.final_residual_norm = tn; rep.termination = T_CONVERGED; return rep; }
      r = tr; rn = tn; rhat = r; rho = alpha = omega = 1.0;
      std::fill(p.begin(), p.end(), 0.0); std::fill(v.begin(), v.end(), 0.0);
      continue;
    }
    M.apply(s.data(), sh.data(), n);
    spmv(a, sh.data(), t.data());
    double tt = dot(t, t);
    omega = tt > 0.0 ? dot(t, s) / tt : 0.0;
    for (i64 i = 0; i < n; ++i) { x[i] += alpha * ph[i] + omega * sh[i]; r[i] = s[i] - omega * t[i]; }
    rho = rho_new;
    rn = nrm2(r);
    if (rn <= thr) {
      std::vector<double> tr = true_residual(a, b, x);
      double tn = nrm2(tr);
      if (tn <= thr) { rep.final_residual_norm = tn; rep.termination = T_CONVERGED; return rep; }
      r = tr; rn = tn; rhat = r; rho = alpha = omega = 1.0;
      std::fill(p.begin(), p.end(), 0.0); std::fill(v.begin(), v.end(), 0.0);
      continue;
    }
    if (omega == 0.0 || !std::isfinite(omega)) { rep.final_residual_norm = rn; rep.termination = T_BREAKDOWN; return rep; }
  }
  rep.final_residual_norm = nrm2(true_residual(a, b, x));
  rep.termination = T_MAXIT;
  return rep;
}

// Restart ladder, crates/cfd-math/src/linear_solver/krylov.rs:84-114
i64 restart_ladder(i64 requested) {
  if (requested <= 8) return 8;
  if (requested <= 16) return 16;
  if (requested <= 32) return 32;
  if (requested <= 64) return 64;
  if (requested <= 128) return 128;
  return 256;
}

Report gmres(const Csr &a, const Prec &M, const double *b, double *x, double tol_abs, double tol_rel, i64 max_it, i64 m) {
  i64 n = a.n;
  std::vector<double> bv(b, b + n);
  double thr = std::max(tol_abs, tol_rel * nrm2(bv));
  std::vector<double> r = true_residual(a, b, x);
  Report rep{0, nrm2(r), 0.0, thr, T_MAXIT};
  double beta = rep.initial_residual_norm;
  if (beta <= thr) { rep.final_residual_norm = beta; rep.termination = T_INITIAL; return rep; }
  std::vector<std::vector<double>> V((size_t)m + 1, std::vector<double>((size_t)n)), Z((size_t)m, std::vector<double>((size_t)n));
  std::vector<double> H((size_t)((m + 1) * m), 0.0), cs((size_t)m), sn((size_t)m), g((size_t)m + 1), w((size_t)n);
  i64 it = 0;
  while (it < max_it) {
    for (i64 i = 0; i < n; ++i) V[0][i] = r[i] / beta;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    i64 j = 0;
    bool breakdown = false;
    for (; j < m && it < max_it; ++j) {
      ++it;
      M.apply(V[j].data(), Z[j].data(), n);
      spmv(a, Z[j].data(), w.data());
      for (i64 k = 0; k <= j; ++k) {  // modified Gram-Schmidt
        double hk = dot(w, V[k]);
        H[k * m + j] = hk;
        for (i64 i = 0; i < n; ++i) w[i] -= hk * V[k][i];
      }
      double hn = nrm2(w);
      H[(j + 1) * m + j] = hn;
      for (i64 k = 0; k < j; ++k) {  // apply previous rotations
        double t0 = H[k * m + j], t1 = H[(k + 1) * m + j];
        H[k * m + j] = cs[k] * t0 + sn[k] * t1;
        H[(k + 1) * m + j] = -sn[k] * t0 + cs[k] * t1;
      }
      double h0 = H[j * m + j], h1 = H[(j + 1) * m + j];
      double den = std::hypot(h0, h1);
      if (den == 0.0) { breakdown = true; break; }
      cs[j] = h0 / den; sn[j] = h1 / den;
      H[j * m + j] = den; H[(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      bool happy = hn == 0.0;
      if (!happy) for (i64 i = 0; i < n; ++i) V[j + 1][i] = w[i] / hn;
      if (std::fabs(g[j + 1]) <= thr || happy) { ++j; break; }
    }
    // solve the triangular system and update x = x + Z y
    i64 kdim = j;
    std::vector<double> y((size_t)kdim);
    for (i64 i = kdim - 1; i >= 0; --i) {
      double s = g[i];
      for (i64 k = i + 1; k < kdim; ++k) s -= H[i * m + k] * y[k];
      y[i] = s / H[i * m + i];
    }
    for (i64 k = 0; k < kdim; ++k) for (i64 i = 0; i < n; ++i) x[i] += y[k] * Z[k][i];
    r = true_residual(a, b, x);
    beta = nrm2(r);
    rep.iterations = it;
    if (beta <= thr) { rep.final_residual_norm = beta; rep.termination = T_CONVERGED; return rep; }
    if (breakdown && kdim == 0) { rep.final_residual_norm = beta; rep.termination = T_BREAKDOWN; return rep; }
  }
  rep.final_residual_norm = beta;
  rep.termination = T_MAXIT;
  return rep;
}

// ---------------------------------------------------------------------------
// Matrix assemblies.
// ---------------------------------------------------------------------------

// Cell-centred 5-point pressure-correction Laplacian, all cells fluid:
// crates/cfd-2d/src/pressure_velocity/correction.rs:273-343.  nxc, nyc are
// interior cell counts (the reference's nx-2, ny-2); row index
// (i-1)*(ny-2)+(j-1) with i the x index.  dirichlet = {west, east, south, north}.
// Without any Dirichlet side the first cell is pinned as the pressure
// reference: system size n - 1, map_index drops it (correction.rs:225-246).
Csr assemble_pressure_correction(i64 nxc, i64 nyc, double dx, double dy, const int dirichlet[4]) {
  i64 nx = nxc + 2, ny = nyc + 2, n = nxc * nyc;
  bool has_dirichlet = dirichlet[0] || dirichlet[1] || dirichlet[2] || dirichlet[3];
  i64 ref = has_dirichlet ? -1 : 0;
  i64 size = has_dirichlet ? n : n - 1;
  auto map_index = [&](i64 idx) -> i64 {
    if (ref < 0) return idx;
    if (idx == ref) return -1;
    return idx < ref ? idx : idx - 1;
  };
  double dx2 = 1.0 / (dx * dx), dy2 = 1.0 / (dy * dy);
  Builder bd(size, size);
  for (i64 i = 1; i < nx - 1; ++i)
    for (i64 j = 1; j < ny - 1; ++j) {
      i64 idx = (i - 1) * (ny - 2) + (j - 1);
      if (idx == ref) continue;
      i64 row = map_index(idx);
      double ap = 0.0;
      if (i > 1) { ap += dx2; i64 c = map_index(idx - (ny - 2)); if (c >= 0) bd.add(row, c, -dx2); }
      else if (i == 1 && dirichlet[0]) ap += dx2;
      if (i < nx - 2) { ap += dx2; i64 c = map_index(idx + (ny - 2)); if (c >= 0) bd.add(row, c, -dx2); }
      else if (i == nx - 2 && dirichlet[1]) ap += dx2;
      if (j > 1) { ap += dy2; i64 c = map_index(idx - 1); if (c >= 0) bd.add(row, c, -dy2); }
      else if (j == 1 && dirichlet[2]) ap += dy2;
      if (j < ny - 2) { ap += dy2; i64 c = map_index(idx + 1); if (c >= 0) bd.add(row, c, -dy2); }
      else if (j == ny - 2 && dirichlet[3]) ap += dy2;
      bd.add(row, row, ap);
    }
  return bd.build();
}

// Variable-coefficient cell-centred Poisson (no counterpart in the reference:
// it generalises correction.rs:273-343 to face coefficients).  Cell index
// (z*ny + y)*nx + x.  Face coefficient = harmonic mean 2 kP kN / (kP + kN)
// times inv_h2.  A Dirichlet face adds kP*inv_h2 to the diagonal and
// kP*inv_h2*value to the right-hand side.  Faces visited in the order
// -z, -y, -x, +x, +y, +z; the diagonal accumulates in that order, then the
// Dirichlet faces in the same order.  dir_val[f] for f in
// {-x,+x,-y,+y,-z,+z}; dir_mask bit f marks a Dirichlet face.
Csr assemble_varcoef(i64 nx, i64 ny, i64 nz, double inv_h2, const double *k, int dir_mask,
                     const double dir_val[6], double *rhs_bc) {
  i64 n = nx * ny * nz;
  Builder bd(n, n);
  if (rhs_bc) std::fill(rhs_bc, rhs_bc + n, 0.0);
  const i64 sx = 1, sy = nx, sz = nx * ny;
  for (i64 z = 0; z < nz; ++z)
    for (i64 y = 0; y < ny; ++y)
      for (i64 x = 0; x < nx; ++x) {
        i64 p = (z * ny + y) * nx + x;
        double kp = k[p], diag = 0.0;
        auto face = [&](bool exists, i64 q) {
          if (!exists) return;
          double kq = k[q];
          double c = (2.0 * kp * kq / (kp + kq)) * inv_h2;
          bd.add(p, q, -c);
          diag += c;
        };
        face(nz > 1 && z > 0, p - sz);
        face(y > 0, p - sy);
        face(x > 0, p - sx);
        face(x + 1 < nx, p + sx);
        face(y + 1 < ny, p + sy);
        face(nz > 1 && z + 1 < nz, p + sz);
        const bool on[6] = {x == 0, x == nx - 1, y == 0, y == ny - 1, nz > 1 && z == 0, nz > 1 && z == nz - 1};
        const int order[6] = {4, 2, 0, 1, 3, 5};  // -z, -y, -x, +x, +y, +z
        for (int oi = 0; oi < 6; ++oi) {
          int f = order[oi];
          if (on[f] && (dir_mask >> f & 1)) {
            double c = kp * inv_h2;
            diag += c;
            if (rhs_bc) rhs_bc[p] += c * dir_val[f];
          }
        }
        bd.add(p, p, diag);
      }
  return bd.build();
}

// First-order upwind + diffusion momentum operator for the U component on an
// nx*ny node grid with Dirichlet boundary nodes (lid-driven cavity):
// coefficients crates/cfd-2d/src/physics/momentum/coefficients.rs:199-277,
// 406-414; assembly crates/cfd-2d/src/physics/momentum/solve.rs:264-308.
// u, v are node fields indexed j*nx + i; mu, rho constant.
Csr assemble_momentum_u(i64 nx, i64 ny, double dx, double dy, double dt, double mu, double rho,
                        const double *u, const double *v) {
  i64 n = nx * ny;
  Builder bd(n, n);
  double dew = dy / dx, dns = dx / dy;
  for (i64 j = 0; j < ny; ++j)
    for (i64 i = 0; i < nx; ++i) {
      i64 idx = j * nx + i;
      if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1) { bd.add(idx, idx, 1.0); continue; }
      double ae = mu * dew, aw = mu * dew, an = mu * dns, as = mu * dns;
      double uu = u[idx];
      double vv = (v[idx] + v[j * nx + i + 1]) / 2.0;
      double fx = rho * uu * dy, fy = rho * vv * dx;
      if (uu > 0.0) aw = aw + fx; else ae = ae - fx;
      if (vv > 0.0) as = as + fy; else an = an - fy;
      double sum = ae + aw + an + as;
      double ap = sum + rho * (dx * dy) / dt;
      bd.add(idx, idx, ap);
      bd.add(idx, idx - 1, -aw);
      bd.add(idx, idx + 1, -ae);
      bd.add(idx, idx - nx, -as);
      bd.add(idx, idx + nx, -an);
    }
  return bd.build();
}

// SIMPLE pressure-correction system on an nx*ny node grid (node fields
// indexed j*nx + i): crates/cfd-2d/src/solvers/simple/pressure.rs:18-240 (the
// first assembly; the refresh branch :244-450 computes the same values into
// the fixed pattern).  Interior rows: d-coefficient Laplacian with a_p the
// negated sum, right-hand side the Rhie-Chow mass imbalance (:57-147).
// Boundary rows (:150-240): a pressure-anchor side gives an identity row,
// otherwise the row averages its neighbours across each touching side
// (pressure_velocity/boundary.rs:29-66); without any anchor side node 0 is an
// identity row.  side = {west, east, south, north}: 0 plain (wall, outflow,
// velocity inlet), 1 pressure anchor, 2 periodic with the opposite side.
// max_residual = max |mass imbalance| over interior nodes.
Csr assemble_simple_pressure(i64 nx, i64 ny, double dx, double dy, const int side[4], const double *u,
                             const double *v, const double *p, const double *du, const double *dv,
                             const double *rho, double *rhs, double *max_residual) {
  const i64 n = nx * ny;
  const double half = 0.5, two = 2.0;
  Builder bd(n, n);
  std::fill(rhs, rhs + n, 0.0);
  double maxr = 0.0;
  for (i64 j = 1; j < ny - 1; ++j)
    for (i64 i = 1; i < nx - 1; ++i) {
      const i64 idx = j * nx + i;
      const double u_p = u[idx], u_e = u[idx + 1], u_w = u[idx - 1];
      const double v_p = v[idx], v_n = v[idx + nx], v_s = v[idx - nx];
      const double p_p = p[idx], p_e = p[idx + 1], p_w = p[idx - 1], p_n = p[idx + nx], p_s = p[idx - nx];
      // one-sided extrapolation next to the boundary (:77-96)
      const double p_ee = i + 2 < nx ? p[idx + 2] : two * p_e - p_p;
      const double p_ww = i >= 2 ? p[idx - 2] : two * p_w - p_p;
      const double p_nn = j + 2 < ny ? p[idx + 2 * nx] : two * p_n - p_p;
      const double p_ss = j >= 2 ? p[idx - 2 * nx] : two * p_s - p_p;
      const double d_e = (du[idx] + du[idx + 1]) * half;
      const double d_w = (du[idx - 1] + du[idx]) * half;
      const double d_n = (dv[idx] + dv[idx + nx]) * half;
      const double d_s = (dv[idx - nx] + dv[idx]) * half;
      // face velocities: linear interpolation plus the Rhie-Chow difference
      // between the averaged cell gradients and the face gradient (:103-117)
      const double gx = (p_e - p_w) / (two * dx);
      const double uf_e = (u_p + u_e) * half + d_e * ((gx + (p_ee - p_p) / (two * dx)) * half - (p_e - p_p) / dx);
      const double uf_w = (u_w + u_p) * half + d_w * (((p_p - p_ww) / (two * dx) + gx) * half - (p_p - p_w) / dx);
      const double gy = (p_n - p_s) / (two * dy);
      const double vf_n = (v_p + v_n) * half + d_n * ((gy + (p_nn - p_p) / (two * dy)) * half - (p_n - p_p) / dy);
      const double vf_s = (v_s + v_p) * half + d_s * (((p_p - p_ss) / (two * dy) + gy) * half - (p_p - p_s) / dy);
      const double r = rho[idx];
      const double imbalance = r * uf_e * dy - r * uf_w * dy + r * vf_n * dx - r * vf_s * dx;
      const double a_e = r * d_e * dy / dx, a_w = r * d_w * dy / dx;
      const double a_n = r * d_n * dx / dy, a_s = r * d_s * dx / dy;
      bd.add(idx, idx, -(a_e + a_w + a_n + a_s));
      bd.add(idx, idx + 1, a_e);
      bd.add(idx, idx - 1, a_w);
      bd.add(idx, idx + nx, a_n);
      bd.add(idx, idx - nx, a_s);
      rhs[idx] = imbalance;
      if (std::fabs(imbalance) > maxr) maxr = std::fabs(imbalance);
    }
  const bool anchored = side[0] == 1 || side[1] == 1 || side[2] == 1 || side[3] == 1;
  for (i64 j = 0; j < ny; ++j)
    for (i64 i = 0; i < nx; ++i) {
      if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) continue;
      const i64 idx = j * nx + i;
      // sides in the order west, east, south, north; the inner neighbour
      // across a side, or the periodic partner node
      const bool on[4] = {i == 0, i == nx - 1, j == 0, j == ny - 1};
      const i64 inner[4] = {j * nx + 1, j * nx + nx - 2, nx + i, (ny - 2) * nx + i};
      const i64 partner[4] = {j * nx + nx - 1, j * nx, (ny - 1) * nx + i, i};
      std::vector<i64> nb;
      bool identity = false;
      for (int s = 0; s < 4 && !identity; ++s) {
        if (!on[s]) continue;
        if (side[s] == 1) identity = true;
        else nb.push_back(side[s] == 2 ? partner[s] : inner[s]);
      }
      if (identity || (!anchored && idx == 0) || nb.empty()) {
        bd.add(idx, idx, 1.0);
        continue;
      }
      std::sort(nb.begin(), nb.end());
      nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
      const double w = 1.0 / (double)nb.size();
      bd.add(idx, idx, 1.0);
      for (i64 c : nb) bd.add(idx, c, -w);
    }
  *max_residual = maxr;
  return bd.build();
}

// Pressure and velocity corrections after the SIMPLE pressure solve
// (crates/cfd-2d/src/solvers/simple/pressure.rs:487-520), interior nodes:
// p += relax p', u -= d_u (dp'/dx) dx, v -= d_v (dp'/dy) dy with central
// differences of p'.
void simple_corrections(i64 nx, i64 ny, double dx, double dy, double relax, const double *pp, const double *du,
                        const double *dv, double *p, double *u, double *v) {
  const double two = 2.0;
  for (i64 j = 1; j < ny - 1; ++j)
    for (i64 i = 1; i < nx - 1; ++i) {
      const i64 idx = j * nx + i;
      p[idx] += relax * pp[idx];
      const double dpx = (pp[idx + 1] - pp[idx - 1]) / (two * dx);
      const double dpy = (pp[idx + nx] - pp[idx - nx]) / (two * dy);
      u[idx] -= du[idx] * dpx * dx;
      v[idx] -= dv[idx] * dpy * dy;
    }
}

// Kirchhoff network Laplacian with Dirichlet row replacement and column
// elimination: crates/cfd-1d/src/solver/core/matrix_assembly.rs:163-247.
// Returns false on a self-loop or a non-positive / non-finite conductance
// (:178-198).
bool assemble_network(i64 n, i64 ne, const i64 *ei, const i64 *ej, const double *g,
                      const unsigned char *is_dir, const double *dir_val, const double *neumann,
                      Csr &out, double *rhs) {
  Builder bd(n, n);
  std::fill(rhs, rhs + n, 0.0);
  for (i64 e = 0; e < ne; ++e) {
    i64 i = ei[e], j = ej[e];
    double c = g[e];
    if (i == j || !std::isfinite(c) || c <= 0.0) return false;
    bool di = is_dir[i], dj = is_dir[j];
    if (!di && !dj) { bd.add(i, i, c); bd.add(j, j, c); bd.add(i, j, -c); bd.add(j, i, -c); }
    else if (di && !dj) { bd.add(j, j, c); rhs[j] += c * dir_val[i]; }
    else if (!di && dj) { bd.add(i, i, c); rhs[i] += c * dir_val[j]; }
  }
  if (neumann) for (i64 i = 0; i < n; ++i) rhs[i] += neumann[i];
  for (i64 i = 0; i < n; ++i)
    if (is_dir[i]) { bd.add(i, i, 1.0); rhs[i] = dir_val[i]; }
  out = bd.build();
  return true;
}

// crates/cfd-1d/src/solver/core/solver_detection.rs:36-64
int detect_spd(const Csr &a) {
  for (i64 i = 0; i < a.n; ++i) {
    double diag = 0.0, off = 0.0;
    for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) {
      if (a.ci[k] == i) diag = a.v[k];
      else if (a.v[k] > 0.0) return 0;
      else off += std::fabs(a.v[k]);
    }
    bool ident = diag == 1.0 && off == 0.0;
    if ((diag < off || diag <= 0.0) && !ident) return 0;
  }
  return 1;
}

// Rectangular duct resistance:
// crates/cfd-1d/src/physics/resistance/models/rectangular.rs:95-157,209-245.
double poiseuille_number(double eps) {
  if (eps == 0.0) return 96.0;
  const double pi = 3.14159265358979323846;
  double t1 = (1.0 + eps) * (1.0 + eps);
  double t2 = 1.0 - (192.0 * eps / std::pow(pi, 5)) * std::tanh(pi / (2.0 * eps));
  return 96.0 / (t1 * t2);
}
double rect_resistance(double w, double h, double L, double mu) {
  double area = w * h;
  double dh = 4.0 * area / (2.0 * (w + h));
  double lo = w <= h ? w : h, hi = w >= h ? w : h;
  double po = poiseuille_number(lo / hi);
  return (po * mu * L) / (2.0 * area * dh * dh);
}

// 1D solve cascade, crates/cfd-1d/src/solver/core/linear_system.rs:114-267.
struct CascadeOut { int path; i64 iterations; };  // path: 0 dense, 1 Krylov, 2 dense fallback

double row_scale(const Csr &a, i64 i) {
  double mx = 0.0;
  for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) mx = std::max(mx, std::fabs(a.v[k]));
  return (mx > 2.220446049250313e-16 && std::isfinite(mx)) ? 1.0 / mx : 1.0;
}

bool meets_target(const Csr &a, const double *x, const double *b, double tol) {
  // linear_system.rs:188-246
  double rn = 0.0, bn = 0.0;
  for (i64 i = 0; i < a.n; ++i) {
    double ax = 0.0;
    for (i64 k = a.rp[i]; k < a.rp[i + 1]; ++k) ax += a.v[k] * x[a.ci[k]];
    double s = row_scale(a, i);
    double r = (ax - b[i]) * s, q = b[i] * s;
    rn += r * r; bn += q * q;
  }
  rn = std::sqrt(rn); bn = std::sqrt(bn);
  if (!std::isfinite(rn)) return false;
  return bn > 2.220446049250313e-16 ? rn / bn <= tol : rn <= tol;
}

bool dense_solve(const Csr &a, const double *b, double *x) {
  std::vector<double> sol(b, b + a.n);
  if (!dense_lu_solve(a.n, to_dense(a), sol)) return false;
  std::copy(sol.begin(), sol.end(), x);
  return true;
}

CascadeOut network_cascade(const Csr &a, const double *b, double *x, double tol, i64 max_it, int use_cg) {
  Csr sa = a;
  std::vector<double> sb(b, b + a.n);
  for (i64 i = 0; i < a.n; ++i) {
    double s = row_scale(a, i);
    for (i64 k = sa.rp[i]; k < sa.rp[i + 1]; ++k) sa.v[k] *= s;
    sb[i] *= s;
  }
  if (a.n <= 256) {
    if (dense_solve(sa, sb.data(), x) && meets_target(a, x, b, tol)) return {0, 0};
    dense_solve(a, b, x);
    return {2, 0};
  }
  Prec M;
  M.kind = PK_JACOBI;
  M.inv_diag = jacobi_inv_diag(sa);
  // IterativeSolverConfig::new(tol): absolute tolerance semantics are
  // unpinned (ext leto-ops); tol is applied relative to ||b|| here.
  Report r = use_cg ? pcg(sa, M, sb.data(), x, 0.0, tol, max_it) : bicgstab(sa, M, sb.data(), x, 0.0, tol, max_it);
  bool conv = r.termination == T_CONVERGED || r.termination == T_INITIAL;
  if (conv && meets_target(a, x, b, tol)) return {1, r.iterations};
  dense_solve(a, b, x);
  return {2, r.iterations};
}

// Adjoint gradient of a scalar J(p, k) for the 2D variable-coefficient
// operator of assemble_varcoef (nz = 1), k = exp(g):
//   dJ/dg = k * (dJ/dk|explicit + lambda^T (db/dk - dA/dk p)).
// Written as a scatter over faces (each face updates both cells), the
// reverse of the device kernel's per-cell gather.  dir_mask bits 0..3 =
// -x, +x, -y, +y.  No counterpart in the reference.
void coef_grad_2d(i64 nx, i64 ny, double inv_h2, int dir_mask, const double *dir_val, const double *k,
                  const double *p, const double *lam, const double *dJdk, double *dJdg) {
  const i64 n = nx * ny;
  std::vector<double> dk((size_t)n, 0.0);
  for (i64 c = 0; c < n; ++c) dk[c] = dJdk ? dJdk[c] : 0.0;
  auto face = [&](i64 a, i64 b) {
    // coefficient c = 2 ka kb / (ka + kb) * inv_h2 couples a and b
    double s = k[a] + k[b];
    double dc_da = 2.0 * k[b] * k[b] / (s * s) * inv_h2;
    double dc_db = 2.0 * k[a] * k[a] / (s * s) * inv_h2;
    double w = (lam[a] - lam[b]) * (p[a] - p[b]);  // lambda^T dA/dc p
    dk[a] -= dc_da * w;
    dk[b] -= dc_db * w;
  };
  for (i64 y = 0; y < ny; ++y)
    for (i64 x = 0; x < nx; ++x) {
      i64 c = y * nx + x;
      if (x + 1 < nx) face(c, c + 1);
      if (y + 1 < ny) face(c, c + nx);
    }
  for (i64 y = 0; y < ny; ++y)
    for (i64 x = 0; x < nx; ++x) {
      i64 c = y * nx + x;
      const bool on[4] = {x == 0, x == nx - 1, y == 0, y == ny - 1};
      for (int f = 0; f < 4; ++f)
        if (on[f] && (dir_mask >> f & 1)) {
          // diagonal += k inv_h2, rhs += k inv_h2 value
          dk[c] += lam[c] * inv_h2 * dir_val[f] - lam[c] * inv_h2 * p[c];
        }
    }
  for (i64 c = 0; c < n; ++c) dJdg[c] = k[c] * dk[c];
}

}  // namespace

// ===========================================================================
// C interface for ctypes (oracle/oracle.py).  All index arrays are int64.
// ===========================================================================
extern "C" {

typedef struct { int64_t iterations; double initial_residual_norm, final_residual_norm, threshold; int64_t termination; } orc_report;

void *orc_csr_new(int64_t n, int64_t m, const int64_t *rp, const int64_t *ci, const double *v) {
  Csr *a = new Csr;
  a->n = n; a->m = m;
  a->rp.assign(rp, rp + n + 1);
  a->ci.assign(ci, ci + rp[n]);
  a->v.assign(v, v + rp[n]);
  return a;
}
void orc_csr_free(void *h) { delete static_cast<Csr *>(h); }
void orc_csr_info(void *h, int64_t *n, int64_t *m, int64_t *nnz) {
  Csr *a = static_cast<Csr *>(h);
  *n = a->n; *m = a->m; *nnz = a->nnz();
}
void orc_csr_export(void *h, int64_t *rp, int64_t *ci, double *v) {
  Csr *a = static_cast<Csr *>(h);
  std::copy(a->rp.begin(), a->rp.end(), rp);
  std::copy(a->ci.begin(), a->ci.end(), ci);
  std::copy(a->v.begin(), a->v.end(), v);
}
void orc_spmv(void *h, const double *x, double *y) { spmv(*static_cast<Csr *>(h), x, y); }
void *orc_transpose(void *h) { return new Csr(transpose(*static_cast<Csr *>(h))); }
void *orc_spgemm(void *a, void *b) { return new Csr(spgemm(*static_cast<Csr *>(a), *static_cast<Csr *>(b))); }
void *orc_strength(void *h, double theta) { return new Csr(strength(*static_cast<Csr *>(h), theta)); }

// kind: 0 RS(theta), 1 aggregation(max_size), 2 hybrid(theta, 4), 3 three-phase(theta)
int64_t orc_coarsen(void *h, int kind, double theta, int64_t max_size, int64_t *f2c, int64_t *cpts) {
  Csr *a = static_cast<Csr *>(h);
  Coarsening c = kind == 1   ? aggregation(*a, max_size)
                 : kind == 2 ? hybrid(*a, theta, 4)
                 : kind == 3 ? standard_aggregation(*a, theta)
                             : ruge_stueben(*a, theta);
  std::copy(c.f2c.begin(), c.f2c.end(), f2c);
  std::copy(c.coarse_points.begin(), c.coarse_points.end(), cpts);
  return (int64_t)c.coarse_points.size();
}
void *orc_direct_interp(const int64_t *f2c, int64_t n, int64_t nc) {
  std::vector<i64> m(f2c, f2c + n);
  return new Csr(direct_interp(m, n, nc));
}
void *orc_classical_interp(void *a, const int64_t *cpts, int64_t nc, void *s) {
  std::vector<i64> c(cpts, cpts + nc);
  return new Csr(classical_interp(*static_cast<Csr *>(a), c, *static_cast<Csr *>(s)));
}
void *orc_smooth_interp(void *a, void *pt, double omega) {
  return new Csr(smooth_interp(*static_cast<Csr *>(a), *static_cast<Csr *>(pt), omega));
}

void orc_smoother(void *h, int kind, double omega, double *x, const double *b, int64_t iters) {
  Csr *a = static_cast<Csr *>(h);
  if (kind == SM_SGS) sgs(*a, x, b, iters, omega);
  else if (kind == SM_JACOBI) jacobi(*a, x, b, iters, omega);
  else if (kind == SM_CHEBYSHEV) {
    double lo, hi;
    chebyshev_bounds(*a, &lo, &hi);
    chebyshev(*a, x, b, iters, lo, hi, 3);
  } else gs_forward(*a, x, b, iters, omega);
}
void orc_chebyshev_bounds(void *h, double *lo, double *hi) { chebyshev_bounds(*static_cast<Csr *>(h), lo, hi); }

typedef struct {
  int64_t max_levels, min_coarse;
  int32_t coarsening, interp, smoother, coarse_solve;
  int64_t pre, post;
  double omega, theta, sa_omega;
  int64_t max_agg;
} orc_amg_config;

void *orc_amg_new(void *h, const orc_amg_config *c) {
  Amg *amg = new Amg;
  amg->cfg.max_levels = c->max_levels; amg->cfg.min_coarse = c->min_coarse;
  amg->cfg.coarsening = c->coarsening; amg->cfg.interp = c->interp;
  amg->cfg.smoother = c->smoother; amg->cfg.coarse_solve = c->coarse_solve;
  amg->cfg.pre = c->pre; amg->cfg.post = c->post;
  amg->cfg.omega = c->omega; amg->cfg.theta = c->theta; amg->cfg.sa_omega = c->sa_omega;
  amg->cfg.max_agg = c->max_agg;
  amg_build(*amg, *static_cast<Csr *>(h));
  return amg;
}
void orc_amg_free(void *h) { delete static_cast<Amg *>(h); }
int64_t orc_amg_num_levels(void *h) { return (int64_t)static_cast<Amg *>(h)->levels.size(); }
void orc_amg_complexity(void *h, double *op, double *grid) {
  Amg *a = static_cast<Amg *>(h);
  *op = a->op_complexity; *grid = a->grid_complexity;
}
// which: 0 = A, 1 = R, 2 = P (copies; caller frees)
void *orc_amg_level(void *h, int64_t l, int which) {
  Level &lv = static_cast<Amg *>(h)->levels[(size_t)l];
  return new Csr(which == 0 ? lv.A : which == 1 ? lv.R : lv.P);
}
void orc_amg_level_map(void *h, int64_t l, int64_t *f2c) {
  Level &lv = static_cast<Amg *>(h)->levels[(size_t)l];
  std::copy(lv.f2c.begin(), lv.f2c.end(), f2c);
}
void orc_amg_apply(void *h, const double *r, double *z) { amg_apply(*static_cast<Amg *>(h), r, z); }

// kind: 0 CG, 1 BiCGSTAB, 2 GMRES(restart); prec: 0 identity, 1 Jacobi, 2 AMG (amg handle)
void orc_solve(void *h, int kind, int prec, void *amg, const double *b, double *x, double tol_abs,
               double tol_rel, int64_t max_it, int64_t restart, orc_report *out) {
  Csr *a = static_cast<Csr *>(h);
  Prec M;
  M.kind = prec;
  if (prec == PK_JACOBI) M.inv_diag = jacobi_inv_diag(*a);
  if (prec == PK_AMG) M.amg = static_cast<Amg *>(amg);
  Report r;
  if (kind == 1) r = bicgstab(*a, M, b, x, tol_abs, tol_rel, max_it);
  else if (kind == 2) r = gmres(*a, M, b, x, tol_abs, tol_rel, max_it, restart);
  else r = pcg(*a, M, b, x, tol_abs, tol_rel, max_it);
  out->iterations = r.iterations;
  out->initial_residual_norm = r.initial_residual_norm;
  out->final_residual_norm = r.final_residual_norm;
  out->threshold = r.threshold;
  out->termination = r.termination;
}
int64_t orc_restart_ladder(int64_t requested) { return restart_ladder(requested); }
void orc_set_threads(int32_t n) { g_threads = n < 1 ? 1 : n; }

void orc_jacobi_inv_diag(void *h, double *out) {
  std::vector<double> d = jacobi_inv_diag(*static_cast<Csr *>(h));
  std::copy(d.begin(), d.end(), out);
}

int orc_dense_solve(void *h, const double *b, double *x) { return dense_solve(*static_cast<Csr *>(h), b, x) ? 1 : 0; }

void *orc_assemble_pressure_correction(int64_t nxc, int64_t nyc, double dx, double dy, const int32_t *dir4) {
  int d[4] = {dir4[0], dir4[1], dir4[2], dir4[3]};
  return new Csr(assemble_pressure_correction(nxc, nyc, dx, dy, d));
}
void *orc_assemble_varcoef(int64_t nx, int64_t ny, int64_t nz, double inv_h2, const double *k, int32_t dir_mask,
                           const double *dir_val, double *rhs_bc) {
  return new Csr(assemble_varcoef(nx, ny, nz, inv_h2, k, dir_mask, dir_val, rhs_bc));
}
void *orc_assemble_momentum_u(int64_t nx, int64_t ny, double dx, double dy, double dt, double mu, double rho,
                              const double *u, const double *v) {
  return new Csr(assemble_momentum_u(nx, ny, dx, dy, dt, mu, rho, u, v));
}
void *orc_assemble_simple_pressure(int64_t nx, int64_t ny, double dx, double dy, const int32_t *side4,
                                   const double *u, const double *v, const double *p, const double *du,
                                   const double *dv, const double *rho, double *rhs, double *max_residual) {
  const int sd[4] = {side4[0], side4[1], side4[2], side4[3]};
  return new Csr(assemble_simple_pressure(nx, ny, dx, dy, sd, u, v, p, du, dv, rho, rhs, max_residual));
}
void orc_simple_corrections(int64_t nx, int64_t ny, double dx, double dy, double relax, const double *pp,
                            const double *du, const double *dv, double *p, double *u, double *v) {
  simple_corrections(nx, ny, dx, dy, relax, pp, du, dv, p, u, v);
}
void *orc_assemble_network(int64_t n, int64_t ne, const int64_t *ei, const int64_t *ej, const double *g,
                           const unsigned char *is_dir, const double *dir_val, const double *neumann, double *rhs) {
  Csr *a = new Csr;
  if (!assemble_network(n, ne, ei, ej, g, is_dir, dir_val, neumann, *a, rhs)) { delete a; return nullptr; }
  return a;
}
int orc_detect_spd(void *h) { return detect_spd(*static_cast<Csr *>(h)); }
double orc_poiseuille_number(double eps) { return poiseuille_number(eps); }
double orc_rect_resistance(double w, double h, double L, double mu) { return rect_resistance(w, h, L, mu); }
int orc_network_cascade(void *h, const double *b, double *x, double tol, int64_t max_it, int use_cg, int64_t *iters) {
  CascadeOut o = network_cascade(*static_cast<Csr *>(h), b, x, tol, max_it, use_cg);
  *iters = o.iterations;
  return o.path;
}

}  // extern "C"

extern "C" void orc_coef_grad_2d(int64_t nx, int64_t ny, double inv_h2, int32_t dir_mask, const double *dir_val,
                                 const double *k, const double *p, const double *lam, const double *dJdk,
                                 double *dJdg) {
  coef_grad_2d(nx, ny, inv_h2, dir_mask, dir_val, k, p, lam, dJdk, dJdg);
}
