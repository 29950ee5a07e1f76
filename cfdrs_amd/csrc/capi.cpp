// extern "C" entry points of libcfdla (declared in include/cfdla.h).
// Every entry point converts C++ exceptions into a status code and a
// thread-local message; nothing throws across the boundary.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>

#include "network.hpp"
#include "solver.hpp"

using namespace cfdla;

struct cfdla_matrix_s {
  HostCsr host;
  DevCsr dev;
  bool on_device = false;     // device copy made on first device use
  std::vector<int64_t> perm;  // stored position -> caller position (empty: identity)
  int symmetric = -1;         // -1 unknown
  std::unique_ptr<cfdla_matrix_s> transpose;
  std::unique_ptr<Precond> transpose_prec;
  bool host_stale = false;    // device values newer than host.v (device updates)
  DevBuf<int32_t> dperm;      // perm on the device, for device updates
  int64_t simple_nx = 0, simple_ny = 0;  // SIMPLE pressure pattern verified for this grid
  int64_t varcoef_grid[3] = {0, 0, 0};   // variable-coefficient pattern verified for this grid
};

struct cfdla_precond_s {
  std::unique_ptr<Precond> p;
  cfdla_matrix_s *m = nullptr;
};

namespace {

thread_local std::string g_error;

template <typename F>
cfdla_status guarded(F &&f) {
  try {
    f();
    g_error.clear();
    return CFDLA_OK;
  } catch (const Error &e) {
    g_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_error = "host allocation failed";
    return CFDLA_ERR_OOM;
  } catch (const std::exception &e) {
    g_error = e.what();
    return CFDLA_ERR_INVALID_ARG;
  }
}

void require_ptr(const void *p, const char *what) {
  CFDLA_REQUIRE(p != nullptr, CFDLA_ERR_INVALID_ARG, std::string(what) + " must not be NULL");
}

// Validate caller CSR, sort each row by column and reject duplicates.
void ingest_csr(int64_t nrows, int64_t ncols, const int64_t *rp, const int64_t *ci, const double *v, HostCsr &h,
                std::vector<int64_t> &perm) {
  CFDLA_REQUIRE(nrows >= 0 && ncols >= 0, CFDLA_ERR_INVALID_ARG, "matrix dimensions must be non-negative");
  require_ptr(rp, "row_ptr");
  CFDLA_REQUIRE(rp[0] == 0, CFDLA_ERR_INVALID_ARG, "row_ptr[0] must be 0");
  for (int64_t i = 0; i < nrows; ++i)
    CFDLA_REQUIRE(rp[i + 1] >= rp[i], CFDLA_ERR_INVALID_ARG, "row_ptr must be non-decreasing");
  const int64_t nnz = rp[nrows];
  CFDLA_REQUIRE(nnz < INT32_MAX && nrows < INT32_MAX && ncols < INT32_MAX, CFDLA_ERR_INVALID_ARG,
                "matrix exceeds 32-bit device indexing");
  if (nnz > 0) {
    require_ptr(ci, "col_indices");
    require_ptr(v, "values");
  }
  h.n = nrows;
  h.m = ncols;
  h.rp.assign(rp, rp + nrows + 1);
  h.ci.resize((size_t)nnz);
  h.v.resize((size_t)nnz);
  perm.clear();
  bool sorted = true;
  for (int64_t i = 0; i < nrows && sorted; ++i)
    for (int64_t k = rp[i] + 1; k < rp[i + 1]; ++k)
      if (ci[k] <= ci[k - 1]) { sorted = false; break; }
  std::vector<int64_t> order;
  if (!sorted) {
    order.resize((size_t)nnz);
    std::iota(order.begin(), order.end(), 0);
    for (int64_t i = 0; i < nrows; ++i)
      std::stable_sort(order.begin() + rp[i], order.begin() + rp[i + 1],
                       [&](int64_t a, int64_t b) { return ci[a] < ci[b]; });
  }
  for (int64_t i = 0; i < nrows; ++i)
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      int64_t src = sorted ? k : order[k];
      int64_t c = ci[src];
      CFDLA_REQUIRE(c >= 0 && c < ncols, CFDLA_ERR_INVALID_ARG, "column index out of range");
      CFDLA_REQUIRE(k == rp[i] || c != h.ci[k - 1], CFDLA_ERR_INVALID_ARG, "duplicate column index in a row");
      h.ci[k] = (int32_t)c;
      h.v[k] = v[src];
    }
  if (!sorted) perm = std::move(order);
}

DevCsr &devmat(cfdla_matrix_s *m) {
  if (!m->on_device) {
    m->dev.upload(m->host, true, false);
    m->on_device = true;
  }
  return m->dev;
}

// host.v after device updates (copied back on demand).
const HostCsr &host_values(cfdla_matrix_s *m) {
  if (m->host_stale) {
    m->dev.v64.download(m->host.v.data(), m->host.v.size());
    sync_stream();
    m->host_stale = false;
  }
  return m->host;
}

// New values landed in dev.v64: derived device copies follow, host copies
// and host-derived caches are dropped.
void device_values_changed(cfdla_matrix_s *m) {
  m->host_stale = true;
  m->symmetric = -1;
  m->transpose.reset();
  m->transpose_prec.reset();
  DevCsr &d = m->dev;
  if (!d.maps_ready) d.build_maps(m->host);
  if (d.refresh_from_v64()) return;
  host_values(m);
  d.general_dia = true;
  d.upload(m->host, true, false);
}

void permuted_values(const cfdla_matrix_s *m, const double *values, std::vector<double> &out) {
  const size_t nnz = m->host.v.size();
  out.resize(nnz);
  if (m->perm.empty()) std::copy(values, values + nnz, out.begin());
  else
    for (size_t k = 0; k < nnz; ++k) out[k] = values[m->perm[k]];
}

cfdla_report to_report(const SolveOut &o) {
  cfdla_report r;
  std::memset(&r, 0, sizeof(r));
  r.iterations = o.iters;
  r.outer_iterations = o.outer;
  r.initial_residual_norm = o.rn0;
  r.final_residual_norm = o.rn;
  r.threshold = o.thr;
  r.termination = o.term;
  r.precision = o.precision;
  return r;
}

bool is_symmetric(cfdla_matrix_s *m) {
  if (m->symmetric >= 0) return m->symmetric == 1;
  const HostCsr &a = host_values(m);
  bool sym = a.n == a.m;
  for (int64_t i = 0; i < a.n && sym; ++i)
    for (int64_t k = a.rp[i]; k < a.rp[i + 1] && sym; ++k) {
      int32_t j = a.ci[k];
      auto b = a.ci.begin() + a.rp[j], e = a.ci.begin() + a.rp[j + 1];
      auto it = std::lower_bound(b, e, (int32_t)i);
      sym = it != e && *it == i && a.v[(size_t)(it - a.ci.begin())] == a.v[k];
    }
  m->symmetric = sym ? 1 : 0;
  return sym;
}

}  // namespace

void cfdla::set_last_error(const std::string &msg) { g_error = msg; }

namespace {

uint64_t splitmix64(uint64_t &state) {
  // SplitMix64 (Steele, Lea & Flood 2014)
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

extern "C" {

const char *cfdla_last_error(void) { return g_error.c_str(); }

const char *cfdla_version(void) { return "cfdla 0.1.0 (gfx950)"; }

cfdla_status cfdla_init(int32_t device) {
  return guarded([&] { init_device(device); });
}

cfdla_status cfdla_set_stream(void *hip_stream) {
  return guarded([&] { set_stream(reinterpret_cast<hipStream_t>(hip_stream)); });
}

cfdla_status cfdla_synchronize(void) {
  return guarded([&] { sync_stream(); });
}

cfdla_status cfdla_matrix_create(int64_t nrows, int64_t ncols, const int64_t *row_ptr, const int64_t *col_indices,
                                 const double *values, cfdla_matrix *out) {
  return guarded([&] {
    require_ptr(out, "out");
    auto m = std::make_unique<cfdla_matrix_s>();
    ingest_csr(nrows, ncols, row_ptr, col_indices, values, m->host, m->perm);
    // device copy is made on first use (devmat)
    *out = m.release();
  });
}

cfdla_status cfdla_matrix_create_device(int64_t nrows, int64_t ncols, int64_t nnz, const int32_t *row_ptr_dev,
                                        const int32_t *col_dev, const double *values_dev, cfdla_matrix *out) {
  return guarded([&] {
    require_ptr(out, "out");
    require_ptr(row_ptr_dev, "row_ptr_dev");
    CFDLA_REQUIRE(nrows >= 0 && ncols >= 0 && nnz >= 0, CFDLA_ERR_INVALID_ARG, "sizes must be non-negative");
    std::vector<int32_t> rp((size_t)nrows + 1), ci((size_t)nnz);
    std::vector<double> v((size_t)nnz);
    CFDLA_HIP(hipMemcpyAsync(rp.data(), row_ptr_dev, rp.size() * 4, hipMemcpyDeviceToHost, stream()));
    if (nnz) {
      require_ptr(col_dev, "col_dev");
      require_ptr(values_dev, "values_dev");
      CFDLA_HIP(hipMemcpyAsync(ci.data(), col_dev, ci.size() * 4, hipMemcpyDeviceToHost, stream()));
      CFDLA_HIP(hipMemcpyAsync(v.data(), values_dev, v.size() * 8, hipMemcpyDeviceToHost, stream()));
    }
    sync_stream();
    CFDLA_REQUIRE(rp[nrows] == nnz, CFDLA_ERR_DIMENSION, "row_ptr[nrows] != nnz");
    std::vector<int64_t> rp64(rp.begin(), rp.end()), ci64(ci.begin(), ci.end());
    auto m = std::make_unique<cfdla_matrix_s>();
    ingest_csr(nrows, ncols, rp64.data(), ci64.data(), v.data(), m->host, m->perm);
    // device copy is made on first use (devmat)
    *out = m.release();
  });
}

cfdla_status cfdla_matrix_update_values(cfdla_matrix m, const double *values) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(values, "values");
    std::vector<double> v;
    permuted_values(m, values, v);
    m->host.v = v;
    m->host_stale = false;
    m->symmetric = -1;
    m->transpose.reset();
    m->transpose_prec.reset();
    if (m->on_device) m->dev.upload(m->host, true, false);
  });
}

cfdla_status cfdla_matrix_update_values_device(cfdla_matrix m, const double *values_dev) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(values_dev, "values");
    DevCsr &d = devmat(m);
    if (d.nnz == 0) return;
    if (!m->perm.empty() && !m->dperm.p) {
      std::vector<int32_t> p32(m->perm.begin(), m->perm.end());
      m->dperm.upload(p32.data(), p32.size());
    }
    gather_values(d.nnz, values_dev, m->perm.empty() ? nullptr : m->dperm.p, d.v64.p);
    device_values_changed(m);
  });
}

cfdla_status cfdla_matrix_get_values(cfdla_matrix m, double *values) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(values, "values");
    const HostCsr &h = host_values(m);
    if (m->perm.empty()) std::copy(h.v.begin(), h.v.end(), values);
    else
      for (size_t k = 0; k < h.v.size(); ++k) values[m->perm[k]] = h.v[k];
  });
}

cfdla_status cfdla_simple_pressure_refresh_device(cfdla_matrix m, int64_t nx, int64_t ny, double dx, double dy,
                                                  int32_t positive_diagonal, const double *u, const double *v,
                                                  const double *p,
                                                  const double *d_u, const double *d_v, const double *rho,
                                                  double *rhs_dev, double *max_residual) {
  return guarded([&] {
    require_ptr(m, "matrix");
    CFDLA_REQUIRE(u && v && p && d_u && d_v && rho && rhs_dev, CFDLA_ERR_INVALID_ARG,
                  "fields and rhs must not be NULL");
    CFDLA_REQUIRE(nx >= 3 && ny >= 3 && m->host.n == nx * ny && m->host.m == nx * ny, CFDLA_ERR_DIMENSION,
                  "matrix is not the SIMPLE pressure system of this grid");
    if (m->simple_nx != nx || m->simple_ny != ny) {
      // every interior row must be the 5-point row the kernel rewrites
      const HostCsr &h = m->host;
      for (int64_t j = 1; j < ny - 1; ++j)
        for (int64_t i = 1; i < nx - 1; ++i) {
          const int64_t idx = j * nx + i, k0 = h.rp[idx];
          const int64_t want[5] = {idx - nx, idx - 1, idx, idx + 1, idx + nx};
          bool ok = h.rp[idx + 1] - k0 == 5;
          for (int q = 0; q < 5 && ok; ++q) ok = h.ci[k0 + q] == want[q];
          CFDLA_REQUIRE(ok, CFDLA_ERR_INVALID_ARG,
                        "matrix pattern is not cfdla_problem_simple_pressure's for this grid");
        }
      m->simple_nx = nx;
      m->simple_ny = ny;
    }
    DevCsr &d = devmat(m);
    simple_refresh(nx, ny, dx, dy, positive_diagonal != 0, u, v, p, d_u, d_v, rho, d.rp.p, d.v64.p, rhs_dev,
                   max_residual != nullptr);
    device_values_changed(m);
    if (max_residual) *max_residual = simple_max_read();
  });
}

cfdla_status cfdla_varcoef_refresh_device(cfdla_matrix m, int64_t nx, int64_t ny, int64_t nz, double inv_h2,
                                         const double *k_dev, int32_t dir_mask, const double *dir_val,
                                         double *rhs_bc_dev) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(k_dev, "k_dev");
    CFDLA_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, CFDLA_ERR_INVALID_ARG, "grid must have at least one cell");
    CFDLA_REQUIRE(nx * ny * nz < INT32_MAX, CFDLA_ERR_INVALID_ARG, "grid exceeds 32-bit device indexing");
    if (dir_mask) require_ptr(dir_val, "dir_val");
    const int64_t n = nx * ny * nz;
    CFDLA_REQUIRE(m->host.n == n && m->host.m == n, CFDLA_ERR_DIMENSION,
                  "matrix is not the variable-coefficient operator of this grid");
    if (m->varcoef_grid[0] != nx || m->varcoef_grid[1] != ny || m->varcoef_grid[2] != nz) {
      // the stored pattern must be cfdla_problem_varcoef's for this grid
      const HostCsr &h = m->host;
      const bool three_d = nz > 1;
      const int64_t sy = nx, sz = nx * ny;
      for (int64_t p = 0; p < n; ++p) {
        const int64_t x = p % nx, y = (p / nx) % ny, z = p / sz;
        int64_t want[7], c = 0;
        if (three_d && z > 0) want[c++] = p - sz;
        if (y > 0) want[c++] = p - sy;
        if (x > 0) want[c++] = p - 1;
        want[c++] = p;
        if (x + 1 < nx) want[c++] = p + 1;
        if (y + 1 < ny) want[c++] = p + sy;
        if (three_d && z + 1 < nz) want[c++] = p + sz;
        bool ok = h.rp[p + 1] - h.rp[p] == c;
        for (int64_t q = 0; q < c && ok; ++q) ok = h.ci[h.rp[p] + q] == want[q];
        CFDLA_REQUIRE(ok, CFDLA_ERR_INVALID_ARG, "matrix pattern is not cfdla_problem_varcoef's for this grid");
      }
      m->varcoef_grid[0] = nx;
      m->varcoef_grid[1] = ny;
      m->varcoef_grid[2] = nz;
    }
    DevCsr &d = devmat(m);
    varcoef_refresh(nx, ny, nz, inv_h2, k_dev, dir_mask, dir_val, d.rp.p, d.v64.p, rhs_bc_dev);
    device_values_changed(m);
  });
}

cfdla_status cfdla_pressure_rhs_device(int64_t nx, int64_t ny, double dx, double dy, double rho, double dt,
                                       const double *u, const double *v, int32_t pin_reference, double *rhs) {
  return guarded([&] {
    CFDLA_REQUIRE(nx >= 3 && ny >= 3, CFDLA_ERR_INVALID_ARG, "pressure grid needs at least 3x3 nodes");
    CFDLA_REQUIRE(u && v && rhs, CFDLA_ERR_INVALID_ARG, "u, v and rhs must not be NULL");
    ensure_init();
    pressure_rhs(nx, ny, dx, dy, rho / dt, u, v, pin_reference != 0, rhs);
  });
}

cfdla_status cfdla_simple_corrections_device(int64_t nx, int64_t ny, double dx, double dy, double relax,
                                             const double *p_prime, const double *d_u, const double *d_v, double *p,
                                             double *u, double *v) {
  return guarded([&] {
    CFDLA_REQUIRE(p_prime && d_u && d_v && p && u && v, CFDLA_ERR_INVALID_ARG, "arrays must not be NULL");
    CFDLA_REQUIRE(nx >= 3 && ny >= 3, CFDLA_ERR_DIMENSION, "SIMPLE grid needs at least 3x3 nodes");
    ensure_init();
    simple_corrections(nx, ny, dx, dy, relax, p_prime, d_u, d_v, p, u, v);
  });
}

cfdla_status cfdla_matrix_info(cfdla_matrix m, int64_t *nrows, int64_t *ncols, int64_t *nnz) {
  return guarded([&] {
    require_ptr(m, "matrix");
    if (nrows) *nrows = m->host.n;
    if (ncols) *ncols = m->host.m;
    if (nnz) *nnz = m->host.nnz();
  });
}

cfdla_status cfdla_matrix_storage(cfdla_matrix m, int32_t *format, int64_t *stored_entries, int64_t *units) {
  return guarded([&] {
    require_ptr(m, "matrix");
    const DevCsr &d = devmat(m);
    if (d.dia) {
      if (format) *format = CFDLA_STORAGE_DIA;
      if (stored_entries) *stored_entries = (int64_t)d.dia_desc.nd * d.n;
      if (units) *units = d.dia_desc.nd;
      return;
    }
    if (format) *format = d.sell ? CFDLA_STORAGE_SELL64 : CFDLA_STORAGE_CSR_ROWBLOCK;
    if (stored_entries) *stored_entries = d.sell ? d.sell_entries : d.nnz;
    if (units) *units = d.sell ? d.nslices : d.nrb64;
  });
}

cfdla_status cfdla_matrix_set_storage(cfdla_matrix m, int32_t format) {
  return guarded([&] {
    require_ptr(m, "matrix");
    CFDLA_REQUIRE(format == CFDLA_STORAGE_AUTO || format == CFDLA_STORAGE_CSR_ROWBLOCK ||
                      format == CFDLA_STORAGE_SELL64 || format == CFDLA_STORAGE_DIA ||
                      format == CFDLA_STORAGE_CSR_INT32,
                  CFDLA_ERR_INVALID_ARG, "unknown storage format");
    if (m->dev.fmt_request == format) return;
    m->dev.fmt_request = format;
    if (m->on_device) m->dev.upload(host_values(m), true, false);
  });
}

cfdla_status cfdla_matrix_set_gather(cfdla_matrix m, int32_t layout) {
  return guarded([&] {
    require_ptr(m, "matrix");
    CFDLA_REQUIRE(layout >= CFDLA_GATHER_AUTO && layout <= CFDLA_GATHER_SORTED, CFDLA_ERR_INVALID_ARG,
                  "unknown gather layout");
    if (m->dev.gather_request == layout) return;
    m->dev.gather_request = layout;
    if (m->on_device) m->dev.upload(host_values(m), true, false);
  });
}

cfdla_status cfdla_matrix_gather(cfdla_matrix m, int32_t *layout_fp64, int32_t *layout_fp32) {
  return guarded([&] {
    require_ptr(m, "matrix");
    const DevCsr &d = devmat(m);
    const bool rowblock = !d.dia && !d.sell;
    if (layout_fp64) *layout_fp64 = rowblock ? d.layout64 : CFDLA_GATHER_QUAD;
    if (layout_fp32) *layout_fp32 = rowblock ? d.layout32 : CFDLA_GATHER_QUAD;
  });
}

cfdla_status cfdla_matrix_destroy(cfdla_matrix m) {
  return guarded([&] { delete m; });
}

cfdla_status cfdla_spmv(cfdla_matrix m, const double *x, double *y) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(x, "x");
    require_ptr(y, "y");
    if (m->host.n == 0) return;
    Context &c = ctx();
    double *dx = c.dv_at(9, std::max<int64_t>(1, m->host.m)), *dy = c.dv_at(10, std::max<int64_t>(1, m->host.n));
    CFDLA_HIP(hipMemcpyAsync(dx, x, m->host.m * 8, hipMemcpyHostToDevice, stream()));
    csr_apply<double>(devmat(m), EPI_Y, dx, nullptr, nullptr, 0.0, dy, c.s(3), false, OP_NONE, nullptr, c.ws);
    CFDLA_HIP(hipMemcpyAsync(y, dy, m->host.n * 8, hipMemcpyDeviceToHost, stream()));
    sync_stream();
  });
}

cfdla_status cfdla_spmv_device(cfdla_matrix m, int32_t dtype, const void *x, void *y) {
  return guarded([&] {
    require_ptr(m, "matrix");
    CFDLA_REQUIRE(dtype == CFDLA_F64 || dtype == CFDLA_F32, CFDLA_ERR_INVALID_ARG,
                  "dtype must be CFDLA_F64 or CFDLA_F32");
    if (m->host.n == 0) return;
    Context &c = ctx();
    if (dtype == CFDLA_F64)
      csr_apply<double>(devmat(m), EPI_Y, static_cast<const double *>(x), nullptr, nullptr, 0.0,
                        static_cast<double *>(y), c.s(3), false, OP_NONE, nullptr, c.ws);
    else if (dtype == CFDLA_F32)
      csr_apply<float>(devmat(m), EPI_Y, static_cast<const float *>(x), nullptr, nullptr, 0.0, static_cast<float *>(y),
                       c.s(3), false, OP_NONE, nullptr, c.ws);
    else
      throw Error(CFDLA_ERR_INVALID_ARG, "dtype must be CFDLA_F64 or CFDLA_F32");
  });
}

cfdla_status cfdla_residual_device(cfdla_matrix m, const double *b, const double *x, double *r, double *rnorm) {
  return guarded([&] {
    require_ptr(m, "matrix");
    if (m->host.n == 0) {
      if (rnorm) *rnorm = 0.0;
      return;
    }
    Context &c = ctx();
    Scal *s = c.s(3);
    scal_reset(s, 0.0);
    csr_apply<double>(devmat(m), EPI_RESID, x, b, nullptr, 0.0, r, s, false, OP_INIT_RR, r, c.ws);
    Scal h = c.read(s);
    if (rnorm) *rnorm = std::sqrt(h.rr);
  });
}

cfdla_status cfdla_precond_create_jacobi(cfdla_matrix m, cfdla_precond *out) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(out, "out");
    CFDLA_REQUIRE(m->host.n == m->host.m, CFDLA_ERR_DIMENSION, "Jacobi requires a square matrix");
    auto p = std::make_unique<cfdla_precond_s>();
    p->p = std::make_unique<JacobiPrecond>(host_values(m));
    p->m = m;
    *out = p.release();
  });
}

cfdla_status cfdla_precond_create_sor(cfdla_matrix m, double omega, int32_t *ncolours, cfdla_precond *out) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(out, "out");
    auto p = std::make_unique<cfdla_precond_s>();
    auto sor = std::make_unique<ColoredSorPrecond>(host_values(m), omega);
    if (ncolours) *ncolours = sor->ncolours;
    p->p = std::move(sor);
    p->m = m;
    *out = p.release();
  });
}

cfdla_status cfdla_precond_create_amg(cfdla_matrix m, const cfdla_amg_config *config, cfdla_precond *out) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(config, "config");
    require_ptr(out, "out");
    auto p = std::make_unique<cfdla_precond_s>();
    p->p = std::make_unique<AmgPrecond>(host_values(m), *config);
    p->m = m;
    *out = p.release();
  });
}

cfdla_status cfdla_precond_apply(cfdla_precond p, const double *r, double *z) {
  return guarded([&] {
    require_ptr(p, "preconditioner");
    require_ptr(r, "r");
    require_ptr(z, "z");
    Context &c = ctx();
    const int64_t n = p->m->host.n;
    if (n == 0) return;
    double *dr = c.dv_at(9, std::max<int64_t>(1, n)), *dz = c.dv_at(10, std::max<int64_t>(1, n));
    CFDLA_HIP(hipMemcpyAsync(dr, r, n * 8, hipMemcpyHostToDevice, stream()));
    scal_reset(c.s(3), 0.0);
    p->p->apply(dr, dz, c.s(3), false);
    CFDLA_HIP(hipMemcpyAsync(z, dz, n * 8, hipMemcpyDeviceToHost, stream()));
    sync_stream();
  });
}

cfdla_status cfdla_precond_apply_device(cfdla_precond p, int32_t dtype, const void *r, void *z) {
  return guarded([&] {
    require_ptr(p, "preconditioner");
    CFDLA_REQUIRE(dtype == CFDLA_F64 || dtype == CFDLA_F32, CFDLA_ERR_INVALID_ARG,
                  "dtype must be CFDLA_F64 or CFDLA_F32");
    if (p->m->host.n == 0) return;
    Context &c = ctx();
    scal_reset(c.s(3), 0.0);
    if (dtype == CFDLA_F64) p->p->apply(static_cast<const double *>(r), static_cast<double *>(z), c.s(3), false);
    else if (dtype == CFDLA_F32) p->p->apply(static_cast<const float *>(r), static_cast<float *>(z), c.s(3), false);
    else throw Error(CFDLA_ERR_INVALID_ARG, "dtype must be CFDLA_F64 or CFDLA_F32");
  });
}

cfdla_status cfdla_amg_update_values(cfdla_precond p, const double *values) {
  return guarded([&] {
    require_ptr(p, "preconditioner");
    require_ptr(values, "values");
    auto *amg = dynamic_cast<AmgPrecond *>(p->p.get());
    CFDLA_REQUIRE(amg != nullptr, CFDLA_ERR_INVALID_ARG, "preconditioner is not AMG");
    std::vector<double> v;
    permuted_values(p->m, values, v);
    amg->update_values(v.data());
  });
}

cfdla_status cfdla_amg_set_gather(cfdla_precond p, int32_t layout) {
  return guarded([&] {
    require_ptr(p, "preconditioner");
    auto *amg = dynamic_cast<AmgPrecond *>(p->p.get());
    CFDLA_REQUIRE(amg != nullptr, CFDLA_ERR_INVALID_ARG, "preconditioner is not AMG");
    CFDLA_REQUIRE(layout >= CFDLA_GATHER_AUTO && layout <= CFDLA_GATHER_SORTED, CFDLA_ERR_INVALID_ARG,
                  "unknown gather layout");
    if (amg->gather_request == layout) return;
    amg->gather_request = layout;
    if (amg->on_device) amg->upload();
  });
}

cfdla_status cfdla_amg_update_values_device(cfdla_precond p, cfdla_matrix m) {
  return guarded([&] {
    require_ptr(p, "preconditioner");
    require_ptr(m, "matrix");
    auto *amg = dynamic_cast<AmgPrecond *>(p->p.get());
    CFDLA_REQUIRE(amg != nullptr, CFDLA_ERR_INVALID_ARG, "preconditioner is not AMG");
    CFDLA_REQUIRE(m == p->m, CFDLA_ERR_INVALID_ARG, "matrix is not the one the AMG hierarchy was built from");
    amg->update_values_device(devmat(m).v64.p);
  });
}

cfdla_status cfdla_amg_stats(cfdla_precond p, cfdla_amg_statistics *out) {
  return guarded([&] {
    require_ptr(p, "preconditioner");
    require_ptr(out, "out");
    auto *amg = dynamic_cast<AmgPrecond *>(p->p.get());
    CFDLA_REQUIRE(amg != nullptr, CFDLA_ERR_INVALID_ARG, "preconditioner is not AMG");
    std::memset(out, 0, sizeof(*out));
    amg->sync_host();
    const auto &lv = amg->host.levels;
    out->num_levels = (int32_t)lv.size();
    out->fused_level = amg->on_device ? amg->fused.level : -1;
    auto count = [](const HostCsr &a) {
      int64_t c = 0;
      for (double x : a.v) c += std::fabs(x) > 1e-15;
      return c;
    };
    int64_t fine = count(lv[0].A), tot = 0, vars = 0;
    for (size_t l = 0; l < lv.size(); ++l) {
      out->level_rows[l] = lv[l].A.n;
      out->level_nnz[l] = lv[l].A.nnz();
      tot += count(lv[l].A);
      vars += lv[l].A.n;
    }
    out->operator_complexity = fine > 0 ? (double)tot / (double)fine : 1.0;
    out->grid_complexity = lv[0].A.n > 0 ? (double)vars / (double)lv[0].A.n : 1.0;
    out->setup_time = amg->setup_time;
  });
}

cfdla_status cfdla_amg_export_level(cfdla_precond p, int32_t level, int32_t which, int64_t *nrows, int64_t *ncols,
                                    int64_t *nnz, int64_t *row_ptr, int64_t *col_indices, double *values) {
  return guarded([&] {
    require_ptr(p, "preconditioner");
    auto *amg = dynamic_cast<AmgPrecond *>(p->p.get());
    CFDLA_REQUIRE(amg != nullptr, CFDLA_ERR_INVALID_ARG, "preconditioner is not AMG");
    CFDLA_REQUIRE(level >= 0 && level < (int32_t)amg->host.levels.size(), CFDLA_ERR_INVALID_ARG,
                  "level out of range");
    if (values) amg->sync_host();
    const HostLevel &L = amg->host.levels[(size_t)level];
    CFDLA_REQUIRE(which >= 0 && which <= 2, CFDLA_ERR_INVALID_ARG, "which must be 0 (A), 1 (R) or 2 (P)");
    const HostCsr &a = which == 0 ? L.A : which == 1 ? L.R : L.P;
    if (nrows) *nrows = a.n;
    if (ncols) *ncols = a.m;
    if (nnz) *nnz = a.nnz();
    if (row_ptr) std::copy(a.rp.begin(), a.rp.end(), row_ptr);
    if (col_indices) std::copy(a.ci.begin(), a.ci.end(), col_indices);
    if (values) std::copy(a.v.begin(), a.v.end(), values);
  });
}

cfdla_status cfdla_amg_export_map(cfdla_precond p, int32_t level, int64_t *fine_to_coarse) {
  return guarded([&] {
    require_ptr(p, "preconditioner");
    require_ptr(fine_to_coarse, "fine_to_coarse");
    auto *amg = dynamic_cast<AmgPrecond *>(p->p.get());
    CFDLA_REQUIRE(amg != nullptr, CFDLA_ERR_INVALID_ARG, "preconditioner is not AMG");
    CFDLA_REQUIRE(level >= 0 && level < (int32_t)amg->host.levels.size(), CFDLA_ERR_INVALID_ARG,
                  "level out of range");
    const auto &m = amg->host.levels[(size_t)level].fine_to_coarse;
    for (size_t i = 0; i < m.size(); ++i) fine_to_coarse[i] = m[i];
  });
}

cfdla_status cfdla_precond_destroy(cfdla_precond p) {
  return guarded([&] { delete p; });
}

cfdla_status cfdla_solve(cfdla_matrix m, cfdla_precond p, const double *b, double *x_inout,
                         const cfdla_solver_config *config, cfdla_report *report) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(b, "b");
    require_ptr(x_inout, "x");
    require_ptr(config, "config");
    CFDLA_REQUIRE(!p || p->m->host.n == m->host.n, CFDLA_ERR_DIMENSION, "preconditioner size differs from matrix");
    Context &c = ctx();
    const int64_t n = m->host.n;
    double *db = c.dv_at(9, std::max<int64_t>(1, n)), *dx = c.dv_at(10, std::max<int64_t>(1, n));
    CFDLA_HIP(hipMemcpyAsync(db, b, n * 8, hipMemcpyHostToDevice, stream()));
    CFDLA_HIP(hipMemcpyAsync(dx, x_inout, n * 8, hipMemcpyHostToDevice, stream()));
    SolveOut o = solve_device(devmat(m), p ? p->p.get() : nullptr, db, dx, *config);
    CFDLA_HIP(hipMemcpyAsync(x_inout, dx, n * 8, hipMemcpyDeviceToHost, stream()));
    sync_stream();
    if (report) *report = to_report(o);
  });
}

cfdla_status cfdla_solve_device(cfdla_matrix m, cfdla_precond p, const double *b_dev, double *x_dev,
                                const cfdla_solver_config *config, cfdla_report *report) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(config, "config");
    CFDLA_REQUIRE(!p || p->m->host.n == m->host.n, CFDLA_ERR_DIMENSION, "preconditioner size differs from matrix");
    SolveOut o = solve_device(devmat(m), p ? p->p.get() : nullptr, b_dev, x_dev, *config);
    if (report) *report = to_report(o);
  });
}

cfdla_status cfdla_solve_transpose_device(cfdla_matrix m, cfdla_precond p, const double *g_dev, double *lambda_dev,
                                          const cfdla_solver_config *config, cfdla_report *report) {
  return guarded([&] {
    require_ptr(m, "matrix");
    require_ptr(config, "config");
    CFDLA_REQUIRE(m->host.n == m->host.m, CFDLA_ERR_DIMENSION, "transpose solves need a square matrix");
    CFDLA_REQUIRE(!p || p->m->host.n == m->host.n, CFDLA_ERR_DIMENSION, "preconditioner size differs from matrix");
    SolveOut o;
    if (is_symmetric(m)) {
      o = solve_device(devmat(m), p ? p->p.get() : nullptr, g_dev, lambda_dev, *config);
    } else {
      if (!m->transpose) {
        auto t = std::make_unique<cfdla_matrix_s>();
        HostCsr tr;
        host_values(m);
        tr.n = m->host.m;
        tr.m = m->host.n;
        tr.rp.assign((size_t)tr.n + 1, 0);
        for (int32_t c : m->host.ci) tr.rp[(size_t)c + 1]++;
        for (int64_t j = 0; j < tr.n; ++j) tr.rp[j + 1] += tr.rp[j];
        tr.ci.resize(m->host.ci.size());
        tr.v.resize(m->host.v.size());
        std::vector<int64_t> next(tr.rp.begin(), tr.rp.end() - 1);
        for (int64_t i = 0; i < m->host.n; ++i)
          for (int64_t k = m->host.rp[i]; k < m->host.rp[i + 1]; ++k) {
            int64_t q = next[m->host.ci[k]]++;
            tr.ci[q] = (int32_t)i;
            tr.v[q] = m->host.v[k];
          }
        t->host = std::move(tr);
        t->dev.upload(t->host, true, false);
        t->on_device = true;
        m->transpose_prec = std::make_unique<JacobiPrecond>(t->host);
        m->transpose = std::move(t);
      }
      o = solve_device(m->transpose->dev, m->transpose_prec.get(), g_dev, lambda_dev, *config);
    }
    if (report) *report = to_report(o);
  });
}

cfdla_status cfdla_coef_grad_poisson2d_device(int64_t nx, int64_t ny, double inv_h2, int32_t dir_mask,
                                              const double *dir_val, const double *k, const double *p,
                                              const double *lambda, const double *dJdk, double *dJdg) {
  return guarded([&] {
    CFDLA_REQUIRE(nx > 0 && ny > 0 && nx * ny < INT32_MAX, CFDLA_ERR_INVALID_ARG, "grid size out of range");
    require_ptr(dir_val, "dir_val");
    require_ptr(k, "k");
    require_ptr(p, "p");
    require_ptr(lambda, "lambda");
    require_ptr(dJdg, "dJdg");
    coef_grad_poisson2d(nx, ny, inv_h2, dir_mask, dir_val, k, p, lambda, dJdk, dJdg);
    sync_stream();
  });
}

cfdla_status cfdla_synth_uniform(uint64_t seed, double lo, double hi, int64_t n, double *out) {
  return guarded([&] {
    CFDLA_REQUIRE(n >= 0, CFDLA_ERR_INVALID_ARG, "n must be non-negative");
    if (n) require_ptr(out, "out");
    uint64_t st = seed;
    const double scale = 1.0 / 9007199254740992.0;  // 2^-53
    for (int64_t i = 0; i < n; ++i) {
      double u = (double)(splitmix64(st) >> 11) * scale;
      out[i] = lo + (hi - lo) * u;
    }
  });
}

cfdla_status cfdla_problem_pressure_correction(int64_t nxc, int64_t nyc, double dx, double dy,
                                               const int32_t *dirichlet, int64_t *nnz, int64_t *row_ptr,
                                               int64_t *col_indices, double *values) {
  // Cell (i, j), i = x index (slow), j = y index (fast): index i*nyc + j.
  // Neighbour terms and diagonal accumulate in the reference order
  // west, east, south, north (correction.rs:291-334); columns sorted.  With
  // no Dirichlet side the first cell is the pressure reference: its row and
  // column are dropped and later rows shift down by one (correction.rs:225-246),
  // giving nxc*nyc - 1 rows.
  return guarded([&] {
    CFDLA_REQUIRE(nxc >= 1 && nyc >= 1, CFDLA_ERR_INVALID_ARG, "grid must have at least one cell");
    require_ptr(dirichlet, "dirichlet");
    require_ptr(nnz, "nnz");
    const bool pin = !(dirichlet[0] || dirichlet[1] || dirichlet[2] || dirichlet[3]);
    CFDLA_REQUIRE(!pin || nxc * nyc > 1, CFDLA_ERR_INVALID_ARG, "a pinned system needs at least two cells");
    const double ix2 = 1.0 / (dx * dx), iy2 = 1.0 / (dy * dy);
    auto map = [&](int64_t idx) -> int64_t { return pin ? (idx == 0 ? -1 : idx - 1) : idx; };
    int64_t count = 0;
    for (int64_t i = 0; i < nxc; ++i)
      for (int64_t j = 0; j < nyc; ++j) {
        const int64_t idx = i * nyc + j;
        if (map(idx) < 0) continue;
        count += 1;
        if (i > 0 && map(idx - nyc) >= 0) ++count;
        if (i + 1 < nxc) ++count;
        if (j > 0 && map(idx - 1) >= 0) ++count;
        if (j + 1 < nyc) ++count;
      }
    *nnz = count;
    if (!row_ptr) return;
    require_ptr(col_indices, "col_indices");
    require_ptr(values, "values");
    int64_t o = 0;
    row_ptr[0] = 0;
    for (int64_t i = 0; i < nxc; ++i)
      for (int64_t j = 0; j < nyc; ++j) {
        const int64_t idx = i * nyc + j, row = map(idx);
        if (row < 0) continue;
        double ap = 0.0;
        if (i > 0) ap += ix2; else if (dirichlet[0]) ap += ix2;
        if (i + 1 < nxc) ap += ix2; else if (dirichlet[1]) ap += ix2;
        if (j > 0) ap += iy2; else if (dirichlet[2]) ap += iy2;
        if (j + 1 < nyc) ap += iy2; else if (dirichlet[3]) ap += iy2;
        if (i > 0 && map(idx - nyc) >= 0) { col_indices[o] = map(idx - nyc); values[o++] = -ix2; }
        if (j > 0 && map(idx - 1) >= 0) { col_indices[o] = map(idx - 1); values[o++] = -iy2; }
        col_indices[o] = row; values[o++] = ap;
        if (j + 1 < nyc) { col_indices[o] = map(idx + 1); values[o++] = -iy2; }
        if (i + 1 < nxc) { col_indices[o] = map(idx + nyc); values[o++] = -ix2; }
        row_ptr[row + 1] = o;
      }
  });
}

cfdla_status cfdla_problem_varcoef(int64_t nx, int64_t ny, int64_t nz, double inv_h2, const double *k,
                                   int32_t dir_mask, const double *dir_val, int64_t *nnz, int64_t *row_ptr,
                                   int64_t *col_indices, double *values, double *rhs_bc) {
  // Cell (x, y, z) -> row (z*ny + y)*nx + x.  Face coefficient
  // 2 kP kN / (kP + kN) * inv_h2; diagonal accumulates the faces in the order
  // -z, -y, -x, +x, +y, +z, then the Dirichlet faces in the same order.
  return guarded([&] {
    CFDLA_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, CFDLA_ERR_INVALID_ARG, "grid must have at least one cell");
    require_ptr(k, "k");
    require_ptr(nnz, "nnz");
    const bool three_d = nz > 1;
    int64_t count = 0;
    for (int64_t z = 0; z < nz; ++z)
      for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x)
          count += 1 + (x > 0) + (x + 1 < nx) + (y > 0) + (y + 1 < ny) + (three_d && z > 0) + (three_d && z + 1 < nz);
    *nnz = count;
    if (!row_ptr) return;
    require_ptr(col_indices, "col_indices");
    require_ptr(values, "values");
    if (dir_mask) require_ptr(dir_val, "dir_val");
    const int64_t sy = nx, sz = nx * ny;
    int64_t o = 0;
    row_ptr[0] = 0;
    for (int64_t z = 0; z < nz; ++z)
      for (int64_t y = 0; y < ny; ++y)
        for (int64_t x = 0; x < nx; ++x) {
          const int64_t p = (z * ny + y) * nx + x;
          const double kp = k[p];
          auto coef = [&](int64_t q) { return (2.0 * kp * k[q] / (kp + k[q])) * inv_h2; };
          const bool has[6] = {three_d && z > 0, y > 0, x > 0, x + 1 < nx, y + 1 < ny, three_d && z + 1 < nz};
          const int64_t nb[6] = {p - sz, p - sy, p - 1, p + 1, p + sy, p + sz};
          double c[6] = {0, 0, 0, 0, 0, 0};
          double diag = 0.0;
          for (int f = 0; f < 6; ++f)
            if (has[f]) { c[f] = coef(nb[f]); diag += c[f]; }
          const bool on[6] = {three_d && z == 0, y == 0, x == 0, x == nx - 1, y == ny - 1, three_d && z == nz - 1};
          const int bit[6] = {4, 2, 0, 1, 3, 5};
          double bc = 0.0;
          for (int f = 0; f < 6; ++f)
            if (on[f] && ((dir_mask >> bit[f]) & 1)) {
              const double cd = kp * inv_h2;
              diag += cd;
              bc += cd * dir_val[bit[f]];
            }
          if (rhs_bc) rhs_bc[p] = bc;
          for (int f = 0; f < 3; ++f)
            if (has[f]) { col_indices[o] = nb[f]; values[o++] = -c[f]; }
          col_indices[o] = p;
          values[o++] = diag;
          for (int f = 3; f < 6; ++f)
            if (has[f]) { col_indices[o] = nb[f]; values[o++] = -c[f]; }
          row_ptr[p + 1] = o;
        }
  });
}

}  // extern "C"

// ---- batched networks ----------------------------------------------------------
struct cfdla_batch_s {
  BatchSolver solver;
  std::vector<int64_t> node_off, edge_off, ei, ej;
  std::vector<uint8_t> isdir;
  std::vector<double> dirval, neumann, g;
  DevBuf<int32_t> d_ei, d_ej, d_inc_edge, d_bad;
  DevBuf<int64_t> d_eoff, d_inc_ptr;
  DevBuf<uint8_t> d_isdir;
  bool edges_on_device = false;
  NetworkInput input() const {
    NetworkInput in;
    in.nsys = (int64_t)node_off.size() - 1;
    in.node_off = node_off.data();
    in.edge_off = edge_off.data();
    in.ei = ei.data();
    in.ej = ej.data();
    in.g = g.data();
    in.is_dirichlet = isdir.data();
    in.dirichlet_value = dirval.data();
    in.neumann = neumann.empty() ? nullptr : neumann.data();
    return in;
  }
};

static_assert(sizeof(BatchReport) == sizeof(cfdla_report), "report layouts differ");

extern "C" {

cfdla_status cfdla_batch_create(int64_t n_sys, const int64_t *node_offset, const int64_t *edge_offset,
                                const int64_t *edge_i, const int64_t *edge_j, const double *conductance,
                                const uint8_t *is_dirichlet, const double *dirichlet_value, const double *neumann,
                                cfdla_batch *out) {
  return guarded([&] {
    require_ptr(out, "out");
    CFDLA_REQUIRE(n_sys >= 0, CFDLA_ERR_INVALID_ARG, "n_sys must be non-negative");
    require_ptr(node_offset, "node_offset");
    require_ptr(edge_offset, "edge_offset");
    CFDLA_REQUIRE(node_offset[0] == 0 && edge_offset[0] == 0, CFDLA_ERR_INVALID_ARG, "offsets must start at 0");
    for (int64_t s = 0; s < n_sys; ++s)
      CFDLA_REQUIRE(node_offset[s + 1] >= node_offset[s] && edge_offset[s + 1] >= edge_offset[s],
                    CFDLA_ERR_INVALID_ARG, "offsets must be non-decreasing");
    const int64_t N = node_offset[n_sys], E = edge_offset[n_sys];
    if (E) {
      require_ptr(edge_i, "edge_i");
      require_ptr(edge_j, "edge_j");
      require_ptr(conductance, "conductance");
    }
    if (N) {
      require_ptr(is_dirichlet, "is_dirichlet");
      require_ptr(dirichlet_value, "dirichlet_value");
    }
    auto b = std::make_unique<cfdla_batch_s>();
    b->node_off.assign(node_offset, node_offset + n_sys + 1);
    b->edge_off.assign(edge_offset, edge_offset + n_sys + 1);
    b->ei.assign(edge_i, edge_i + E);
    b->ej.assign(edge_j, edge_j + E);
    b->g.assign(conductance, conductance + E);
    b->isdir.assign(is_dirichlet, is_dirichlet + N);
    b->dirval.assign(dirichlet_value, dirichlet_value + N);
    if (neumann) b->neumann.assign(neumann, neumann + N);
    b->solver.build(b->input());
    *out = b.release();
  });
}

cfdla_status cfdla_batch_update_conductance(cfdla_batch b, const double *conductance) {
  return guarded([&] {
    require_ptr(b, "batch");
    if (!b->g.empty()) require_ptr(conductance, "conductance");
    std::vector<double> g(conductance, conductance + b->g.size());
    NetworkInput in = b->input();
    b->solver.set_values(g.data(), in);  // validates before committing
    b->g = std::move(g);
  });
}

cfdla_status cfdla_batch_info(cfdla_batch b, int64_t *n_sys, int64_t *n_rows, int64_t *nnz) {
  return guarded([&] {
    require_ptr(b, "batch");
    if (n_sys) *n_sys = b->solver.nsys;
    if (n_rows) *n_rows = b->solver.nrows;
    if (nnz) *nnz = b->solver.host.nnz();
  });
}

cfdla_status cfdla_batch_export(cfdla_batch b, int64_t *row_ptr, int64_t *col_indices, double *values, double *rhs) {
  return guarded([&] {
    require_ptr(b, "batch");
    const HostCsr &a = b->solver.host;
    if (row_ptr) std::copy(a.rp.begin(), a.rp.end(), row_ptr);
    if (col_indices) for (size_t k = 0; k < a.ci.size(); ++k) col_indices[k] = a.ci[k];
    if (values || rhs) {
      std::vector<double> v, r;
      b->solver.download(v, r);  // device-formed values when newer than the host assembly
      if (values) std::copy(v.begin(), v.end(), values);
      if (rhs) std::copy(r.begin(), r.end(), rhs);
    }
  });
}

cfdla_status cfdla_batch_update_conductance_device(cfdla_batch b, const double *conductance_dev) {
  return guarded([&] {
    require_ptr(b, "batch");
    if (!b->g.empty()) require_ptr(conductance_dev, "conductance");
    b->solver.set_values_device(conductance_dev);
  });
}

cfdla_status cfdla_batch_set_schedule(cfdla_batch b, const cfdla_batch_schedule *schedule) {
  return guarded([&] {
    require_ptr(b, "batch");
    require_ptr(schedule, "schedule");
    CFDLA_REQUIRE(schedule->wave_max_rows >= 0 && schedule->wave_max_rows <= 512, CFDLA_ERR_CONFIG,
                  "wave_max_rows must be in [0, 512]");
    CFDLA_REQUIRE(schedule->blocks_per_cu >= 1 && schedule->blocks_per_cu <= 16, CFDLA_ERR_CONFIG,
                  "blocks_per_cu must be in [1, 16]");
    b->solver.sched.wave_max_rows = schedule->wave_max_rows;
    b->solver.sched.blocks_per_cu = schedule->blocks_per_cu;
    if (b->solver.on_device) b->solver.make_items();
  });
}

cfdla_status cfdla_batch_get_schedule(cfdla_batch b, cfdla_batch_schedule *schedule) {
  return guarded([&] {
    require_ptr(b, "batch");
    require_ptr(schedule, "schedule");
    schedule->wave_max_rows = b->solver.sched.wave_max_rows;
    schedule->blocks_per_cu = b->solver.sched.blocks_per_cu;
  });
}

cfdla_status cfdla_batch_solve(cfdla_batch b, const cfdla_solver_config *config, double *x_inout,
                               cfdla_report *reports) {
  return guarded([&] {
    require_ptr(b, "batch");
    require_ptr(config, "config");
    const int64_t N = b->solver.nrows;
    if (N) require_ptr(x_inout, "x");
    Context &c = ctx();
    double *dx = c.dv_at(10, std::max<int64_t>(1, N));
    CFDLA_HIP(hipMemcpyAsync(dx, x_inout, N * 8, hipMemcpyHostToDevice, stream()));
    b->solver.solve(dx, *config, reinterpret_cast<BatchReport *>(reports));
    CFDLA_HIP(hipMemcpyAsync(x_inout, dx, N * 8, hipMemcpyDeviceToHost, stream()));
    sync_stream();
  });
}

cfdla_status cfdla_batch_solve_device(cfdla_batch b, const cfdla_solver_config *config, const double *rhs_dev,
                                      double *x_dev, cfdla_report *reports) {
  return guarded([&] {
    require_ptr(b, "batch");
    require_ptr(config, "config");
    if (b->solver.nrows) require_ptr(x_dev, "x_dev");
    b->solver.solve(x_dev, *config, reinterpret_cast<BatchReport *>(reports), rhs_dev);
  });
}

cfdla_status cfdla_batch_solve_device_noalloc(cfdla_batch b, const cfdla_solver_config *config,
                                              const double *rhs_dev, double *x_dev, cfdla_report *reports) {
  return guarded([&] {
    require_ptr(b, "batch");
    require_ptr(config, "config");
    if (b->solver.nrows) require_ptr(x_dev, "x_dev");
    b->solver.solve_noalloc(x_dev, *config, reinterpret_cast<BatchReport *>(reports), rhs_dev);
  });
}

cfdla_status cfdla_batch_autotune(cfdla_batch b, const cfdla_solver_config *config, int32_t n_candidates,
                                  const cfdla_batch_schedule *candidates, double *ms_out,
                                  cfdla_batch_schedule *best) {
  return guarded([&] {
    require_ptr(b, "batch");
    require_ptr(config, "config");
    CFDLA_REQUIRE(n_candidates >= 1, CFDLA_ERR_INVALID_ARG, "no candidate schedules");
    require_ptr(candidates, "candidates");
    const int64_t N = std::max<int64_t>(1, b->solver.nrows);
    DevBuf<double> x(N);
    double best_ms = 1e300;
    BatchSolver &s = b->solver;
    BatchSchedule keep = s.sched, pick = s.sched;
    hipEvent_t e0, e1;
    CFDLA_HIP(hipEventCreate(&e0));
    CFDLA_HIP(hipEventCreate(&e1));
    for (int32_t k = 0; k < n_candidates; ++k) {
      const cfdla_batch_schedule &cand = candidates[k];
      CFDLA_REQUIRE(cand.wave_max_rows >= 0 && cand.wave_max_rows <= 512 && cand.blocks_per_cu >= 1 &&
                        cand.blocks_per_cu <= 16,
                    CFDLA_ERR_CONFIG, "candidate schedule out of range");
      s.sched.wave_max_rows = cand.wave_max_rows;
      s.sched.blocks_per_cu = cand.blocks_per_cu;
      if (s.on_device) s.make_items();
      fill_vec<double>((int)N, 0.0, x.p);
      s.solve(x.p, *config, nullptr);  // warm (uploads on first use)
      fill_vec<double>((int)N, 0.0, x.p);
      CFDLA_HIP(hipEventRecord(e0, stream()));
      s.solve(x.p, *config, nullptr);
      CFDLA_HIP(hipEventRecord(e1, stream()));
      CFDLA_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      CFDLA_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms_out) ms_out[k] = ms;
      if (ms < best_ms) {
        best_ms = ms;
        pick = s.sched;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)keep;
    s.sched = pick;
    s.make_items();
    if (best) {
      best->wave_max_rows = pick.wave_max_rows;
      best->blocks_per_cu = pick.blocks_per_cu;
    }
  });
}

namespace {
// Global edge endpoints, per-network edge offsets, each node's incident edges
// (in edge order) and the Dirichlet flags on the device, uploaded once.
void batch_edges_on_device(cfdla_batch_s *b) {
  if (b->edges_on_device) return;
  const int64_t E = (int64_t)b->ei.size();
  const int64_t ns = (int64_t)b->node_off.size() - 1;
  const int64_t nrows = b->node_off.empty() ? 0 : b->node_off.back();
  std::vector<int32_t> gi((size_t)E), gj((size_t)E);
  for (int64_t s = 0; s < ns; ++s)
    for (int64_t e = b->edge_off[s]; e < b->edge_off[s + 1]; ++e) {
      gi[e] = (int32_t)(b->node_off[s] + b->ei[e]);
      gj[e] = (int32_t)(b->node_off[s] + b->ej[e]);
    }
  std::vector<int64_t> inc_ptr((size_t)nrows + 1, 0);
  for (int64_t e = 0; e < E; ++e) {
    inc_ptr[(size_t)gi[e] + 1]++;
    inc_ptr[(size_t)gj[e] + 1]++;
  }
  for (int64_t i = 0; i < nrows; ++i) inc_ptr[i + 1] += inc_ptr[i];
  std::vector<int32_t> inc((size_t)std::max<int64_t>(1, 2 * E));
  std::vector<int64_t> next(inc_ptr.begin(), inc_ptr.end() - 1);
  for (int64_t e = 0; e < E; ++e) {
    inc[(size_t)next[gi[e]]++] = (int32_t)e;
    inc[(size_t)next[gj[e]]++] = (int32_t)e;
  }
  b->d_ei.upload(gi.data(), std::max<size_t>(1, gi.size()));
  b->d_ej.upload(gj.data(), std::max<size_t>(1, gj.size()));
  b->d_isdir.upload(b->isdir.data(), std::max<size_t>(1, b->isdir.size()));
  b->d_eoff.upload(b->edge_off.data(), b->edge_off.size());
  b->d_inc_ptr.upload(inc_ptr.data(), inc_ptr.size());
  b->d_inc_edge.upload(inc.data(), inc.size());
  b->d_bad.ensure(1);
  sync_stream();
  b->edges_on_device = true;
}
}  // namespace

cfdla_status cfdla_batch_edge_gradient_device(cfdla_batch b, const double *p, const double *lambda,
                                              const double *explicit_dg, const double *dg_dparam, double *out) {
  return guarded([&] {
    require_ptr(b, "batch");
    const int64_t E = (int64_t)b->ei.size();
    if (E) {
      require_ptr(p, "p");
      require_ptr(lambda, "lambda");
      require_ptr(out, "out");
    }
    batch_edges_on_device(b);
    network_gradient(E, b->d_ei.p, b->d_ej.p, b->d_isdir.p, p, lambda, explicit_dg, dg_dparam, out);
  });
}

cfdla_status cfdla_batch_edge_flows_device(cfdla_batch b, const double *p, const double *resistance,
                                           const double *quad_coeff, double *flow_inout) {
  return guarded([&] {
    require_ptr(b, "batch");
    const int64_t E = (int64_t)b->ei.size();
    if (E == 0) return;
    require_ptr(p, "p");
    require_ptr(resistance, "resistance");
    require_ptr(flow_inout, "flow");
    batch_edges_on_device(b);
    network_flows(E, b->d_ei.p, b->d_ej.p, p, resistance, quad_coeff, flow_inout, b->d_bad.p);
  });
}

cfdla_status cfdla_batch_outlet_flux_device(cfdla_batch b, const double *p, const double *conductance,
                                            const uint8_t *outlet_mask, double *flux, double *dflux_dp,
                                            double *dflux_dg) {
  return guarded([&] {
    require_ptr(b, "batch");
    const int64_t ns = (int64_t)b->node_off.size() - 1;
    if (ns <= 0) return;
    require_ptr(p, "p");
    require_ptr(conductance, "conductance");
    require_ptr(outlet_mask, "outlet_mask");
    require_ptr(flux, "flux");
    batch_edges_on_device(b);
    network_outlet(ns, b->node_off.back(), b->d_eoff.p, b->d_ei.p, b->d_ej.p, b->d_inc_ptr.p, b->d_inc_edge.p,
                   b->d_isdir.p, outlet_mask, p, conductance, flux, dflux_dp, dflux_dg);
  });
}

cfdla_status cfdla_batch_destroy(cfdla_batch b) {
  return guarded([&] { delete b; });
}

cfdla_status cfdla_synth_networks(int64_t count, uint64_t seed_base, int64_t nmin, int64_t nmax, int64_t *node_offset,
                                  int64_t *edge_offset, int64_t *edge_i, int64_t *edge_j, double *width,
                                  double *length) {
  return guarded([&] {
    require_ptr(node_offset, "node_offset");
    require_ptr(edge_offset, "edge_offset");
    SynthNetworks s;
    synth_networks(count, seed_base, nmin, nmax, s);
    std::copy(s.node_off.begin(), s.node_off.end(), node_offset);
    std::copy(s.edge_off.begin(), s.edge_off.end(), edge_offset);
    if (!edge_i) return;
    require_ptr(edge_j, "edge_j");
    for (size_t e = 0; e < s.ei.size(); ++e) {
      edge_i[e] = s.ei[e];
      edge_j[e] = s.ej[e];
    }
    if (width) std::copy(s.width.begin(), s.width.end(), width);
    if (length) std::copy(s.length.begin(), s.length.end(), length);
  });
}

cfdla_status cfdla_rect_conductance_device(int64_t n, const double *width, const double *height, const double *length,
                                           double viscosity, double *g, double *dg_dw) {
  return guarded([&] {
    CFDLA_REQUIRE(n >= 0, CFDLA_ERR_INVALID_ARG, "n must be non-negative");
    if (n) {
      require_ptr(width, "width");
      require_ptr(height, "height");
      require_ptr(length, "length");
      require_ptr(g, "g");
    }
    rect_conductance_device(n, width, height, length, viscosity, g, dg_dw);
  });
}

cfdla_status cfdla_rect_conductance(int64_t n, const double *width, const double *height, const double *length,
                                    double viscosity, double *g, double *dg_dw) {
  return guarded([&] {
    CFDLA_REQUIRE(n >= 0, CFDLA_ERR_INVALID_ARG, "n must be non-negative");
    if (n) {
      require_ptr(width, "width");
      require_ptr(height, "height");
      require_ptr(length, "length");
      require_ptr(g, "g");
    }
    for (int64_t e = 0; e < n; ++e) {
      CFDLA_REQUIRE(width[e] > 0.0 && height[e] > 0.0 && length[e] >= 0.0, CFDLA_ERR_INVALID_ARG,
                    "Invalid geometry: width and height must be > 0");
      g[e] = 1.0 / rect_resistance(width[e], height[e], length[e], viscosity);
      if (dg_dw) dg_dw[e] = rect_conductance_dw(width[e], height[e], length[e], viscosity);
    }
  });
}

}  // extern "C"

extern "C" cfdla_status cfdla_synth_grf(int64_t nx, int64_t ny, double corr_len, uint64_t seed, double *out) {
  // Gaussian random field on a periodic nx*ny grid with covariance
  // exp(-r^2 / (2 corr_len^2)): white noise (Box-Muller on SplitMix64)
  // convolved with a Gaussian of standard deviation corr_len / sqrt(2),
  // then shifted and scaled to sample mean 0 and sample deviation 1.
  return guarded([&] {
    CFDLA_REQUIRE(nx >= 1 && ny >= 1 && corr_len > 0.0, CFDLA_ERR_INVALID_ARG, "invalid field size");
    require_ptr(out, "out");
    const int64_t n = nx * ny;
    std::vector<double> w((size_t)n), t((size_t)n);
    uint64_t st = seed;
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < n; i += 2) {
      double u1 = ((double)(splitmix64(st) >> 11) + 1.0) * (1.0 / 9007199254740992.0);  // (0, 1]
      double u2 = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
      double rad = std::sqrt(-2.0 * std::log(u1));
      w[i] = rad * std::cos(two_pi * u2);
      if (i + 1 < n) w[i + 1] = rad * std::sin(two_pi * u2);
    }
    const double s = corr_len / std::sqrt(2.0);
    const int64_t R = (int64_t)std::ceil(4.0 * s);
    std::vector<double> ker((size_t)(2 * R + 1));
    for (int64_t d = -R; d <= R; ++d) ker[d + R] = std::exp(-(double)(d * d) / (2.0 * s * s));
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) {
        double acc = 0.0;
        for (int64_t d = -R; d <= R; ++d) acc += ker[d + R] * w[y * nx + (((x + d) % nx) + nx) % nx];
        t[y * nx + x] = acc;
      }
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) {
        double acc = 0.0;
        for (int64_t d = -R; d <= R; ++d) acc += ker[d + R] * t[(((y + d) % ny) + ny) % ny * nx + x];
        out[y * nx + x] = acc;
      }
    double mean = 0.0, var = 0.0;
    for (int64_t i = 0; i < n; ++i) mean += out[i];
    mean /= (double)n;
    for (int64_t i = 0; i < n; ++i) var += (out[i] - mean) * (out[i] - mean);
    const double sd = std::sqrt(var / (double)n);
    for (int64_t i = 0; i < n; ++i) out[i] = sd > 0.0 ? (out[i] - mean) / sd : 0.0;
  });
}
