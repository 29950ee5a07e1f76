// Device runtime, preconditioners and Krylov drivers of libcfdla.
//
// The Krylov recurrences are driven from the host but keep every scalar on
// the device: a reducing kernel stores one partial per block, and a one-block
// finalize kernel launched after it sums them in block order and updates
// alpha/beta/rho/omega and the convergence flag in a Scal record.  The host only enqueues kernels and polls the flag one step
// behind (a pinned copy of Scal plus an event per step), so the GPU never
// waits for the host, and kernels of steps enqueued after convergence exit
// immediately.
#include "solver.hpp"
#include "fused_jacobi.hpp"

#include <algorithm>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

namespace cfdla {

// ---------------------------------------------------------------------------
// Device state
// ---------------------------------------------------------------------------
namespace {
struct DeviceState {
  bool ready = false;
  int device = 0;
  int cus = 0;
  hipStream_t own = nullptr;
  hipStream_t current = nullptr;
  std::unique_ptr<Context> context;
  hipEvent_t ev[2] = {nullptr, nullptr};
  Scal *poll = nullptr;  // pinned, 2 slots
};
DeviceState &dstate() {
  static DeviceState s;
  return s;
}
}  // namespace

void init_device(int dev) {
  DeviceState &d = dstate();
  if (d.ready) {
    CFDLA_REQUIRE(dev == d.device, CFDLA_ERR_INVALID_ARG, "library already initialised on another device");
    return;
  }
  int count = 0;
  CFDLA_HIP(hipGetDeviceCount(&count));
  CFDLA_REQUIRE(count > 0, CFDLA_ERR_DEVICE, "no HIP device visible");
  CFDLA_REQUIRE(dev >= 0 && dev < count, CFDLA_ERR_INVALID_ARG, "device index out of range");
  CFDLA_HIP(hipSetDevice(dev));
  noalloc_scratch_init(dev);
  hipDeviceProp_t prop;
  CFDLA_HIP(hipGetDeviceProperties(&prop, dev));
  d.cus = prop.multiProcessorCount;
  CFDLA_HIP(hipStreamCreateWithFlags(&d.own, hipStreamNonBlocking));
  d.current = d.own;
  d.device = dev;
  CFDLA_HIP(hipEventCreateWithFlags(&d.ev[0], hipEventDisableTiming));
  CFDLA_HIP(hipEventCreateWithFlags(&d.ev[1], hipEventDisableTiming));
  CFDLA_HIP(hipHostMalloc(reinterpret_cast<void **>(&d.poll), 2 * sizeof(Scal), hipHostMallocDefault));
  d.ready = true;
  d.context = std::make_unique<Context>();
  d.context->ws.init();
  d.context->scal.alloc(4);
  CFDLA_HIP(hipHostMalloc(reinterpret_cast<void **>(&d.context->host), sizeof(Scal), hipHostMallocDefault));
}

void ensure_init() {
  if (!dstate().ready) init_device(0);
}

int library_device() {
  ensure_init();
  return dstate().device;
}

hipStream_t stream() {
  ensure_init();
  return dstate().current;
}

void set_stream(hipStream_t s) {
  ensure_init();
  dstate().current = (s == reinterpret_cast<hipStream_t>(CFDLA_LIBRARY_STREAM)) ? dstate().own : s;
}

int num_cus() {
  ensure_init();
  return dstate().cus;
}

void sync_stream() { CFDLA_HIP(hipStreamSynchronize(stream())); }

Context &ctx() {
  ensure_init();
  return *dstate().context;
}

Scal Context::read(Scal *dev) {
  CFDLA_HIP(hipMemcpyAsync(host, dev, sizeof(Scal), hipMemcpyDeviceToHost, stream()));
  CFDLA_HIP(hipStreamSynchronize(stream()));
  return *host;
}

void RedWs::init() {
  partials.alloc((size_t)kMaxRed * kMaxBlocks);
  counter.alloc(1);
  CFDLA_HIP(hipMemsetAsync(counter.p, 0, sizeof(unsigned), stream()));
  CFDLA_HIP(hipStreamSynchronize(stream()));
}

// Lagged poll: after each enqueued step, copy Scal to a pinned slot and
// record an event; check the previous step's slot.
struct Poller {
  int k = 0;
  Scal *dev;
  explicit Poller(Scal *s) : dev(s) {}
  void post() {
    DeviceState &d = dstate();
    CFDLA_HIP(hipMemcpyAsync(d.poll + (k & 1), dev, sizeof(Scal), hipMemcpyDeviceToHost, stream()));
    CFDLA_HIP(hipEventRecord(d.ev[k & 1], stream()));
    ++k;
  }
  // State as of the step before the last one posted (nullptr if none).
  const Scal *lagged() {
    if (k < 2) return nullptr;
    DeviceState &d = dstate();
    CFDLA_HIP(hipEventSynchronize(d.ev[(k - 2) & 1]));
    return d.poll + ((k - 2) & 1);
  }
  const Scal *latest() {
    DeviceState &d = dstate();
    CFDLA_HIP(hipEventSynchronize(d.ev[(k - 1) & 1]));
    return d.poll + ((k - 1) & 1);
  }
};

// ---------------------------------------------------------------------------
// Device CSR
// ---------------------------------------------------------------------------
void build_row_blocks(const HostCsr &h, int cap, int max_rows, std::vector<int32_t> &rb, int64_t span) {
  rb.assign({0, 0});
  // a row that alone spans more than `span` rules 16-bit offsets out for the
  // whole partition: then blocks are not cut for the span at all
  for (int64_t r = 0; r < h.n && span > 0; ++r)
    if (h.rp[r + 1] > h.rp[r]) {
      const auto mm = std::minmax_element(h.ci.begin() + h.rp[r], h.ci.begin() + h.rp[r + 1]);
      if ((int64_t)*mm.second - (int64_t)*mm.first > span) span = 0;
    }
  int64_t rows = 0, nz = 0;
  int64_t lo = INT64_MAX, hi = INT64_MIN;  // column range of the open block
  for (int64_t r = 0; r < h.n; ++r) {
    const int64_t len = h.rp[r + 1] - h.rp[r];
    int64_t rlo = INT64_MAX, rhi = INT64_MIN;
    if (span > 0)
      for (int64_t k = h.rp[r]; k < h.rp[r + 1]; ++k) {
        rlo = std::min<int64_t>(rlo, h.ci[k]);
        rhi = std::max<int64_t>(rhi, h.ci[k]);
      }
    const bool wide = span > 0 && len > 0 && nz > 0 && std::max(hi, rhi) - std::min(lo, rlo) > span;
    if (rows > 0 && (nz + len > cap || rows == max_rows || wide)) {
      rb.push_back((int32_t)r);
      rb.push_back((int32_t)h.rp[r]);
      rows = nz = 0;
      lo = INT64_MAX;
      hi = INT64_MIN;
    }
    rows += 1;
    nz += len;
    if (len > 0) {
      lo = std::min(lo, rlo);
      hi = std::max(hi, rhi);
    }
  }
  if (rows > 0) {
    rb.push_back((int32_t)h.n);
    rb.push_back((int32_t)h.rp[h.n]);
  }
}

bool build_sorted_blocks(const HostCsr &h, const std::vector<int32_t> &rb, std::vector<int32_t> &packed,
                         std::vector<double> &vals) {
  if (h.m > (int64_t)kRbColMask + 1) return false;
  const int64_t nnz = h.nnz(), nrb = (int64_t)rb.size() / 2 - 1;
  packed.resize((size_t)nnz);
  vals.resize((size_t)nnz);
  // blocks are independent: contiguous ranges of them go to up to 16 host
  // threads (sorting the 24M entries of the 128^3 hierarchy's transfer and
  // level-2 operators on one thread took most of the hierarchy's upload)
  auto work = [&](int64_t q0, int64_t q1) {
    for (int64_t q = q0; q < q1; ++q) {
      const int64_t e0 = rb[2 * q + 1], e1 = rb[2 * q + 3];
      const bool regular = e1 - e0 <= rb_cap<double>();
      // the packed word itself is the sort key: column, then position, so
      // ties keep row order (a long row keeps its order, with position 0)
      for (int64_t k = e0; k < e1; ++k)
        packed[(size_t)k] =
            (int32_t)((uint32_t)h.ci[(size_t)k] | (regular ? (uint32_t)(k - e0) << kRbPosShift : 0u));
      if (regular) {
        const auto key = [](int32_t w) {
          return ((uint32_t)w & (uint32_t)kRbColMask) << 11 | (uint32_t)w >> kRbPosShift;
        };
        std::sort(packed.begin() + e0, packed.begin() + e1, [&](int32_t a, int32_t b) { return key(a) < key(b); });
      }
      for (int64_t j = e0; j < e1; ++j)
        vals[(size_t)j] = h.v[(size_t)(regular ? e0 + ((uint32_t)packed[(size_t)j] >> kRbPosShift) : j)];
    }
  };
  const int64_t nt = std::min<int64_t>({16, std::max(1u, std::thread::hardware_concurrency()), nnz / (1 << 16) + 1});
  std::vector<std::thread> pool;
  for (int64_t t = 1; t < nt; ++t) pool.emplace_back(work, nrb * t / nt, nrb * (t + 1) / nt);
  work(0, nrb / nt);
  for (std::thread &th : pool) th.join();
  return true;
}

int rb_pick_layout(const HostCsr &h, const std::vector<int32_t> &rb, int vbytes) {
  // The layouts of k_csr_rb's regular path, wavefront by wavefront:
  //   quad:    gather instruction (u, j) of wave w covers entry
  //            ea + 4 (64 w + lane + 256 u) + j of the block;
  //   rows:    lane t of wave w walks row t / tpr; its p-th gather is entry
  //            ks + p (tpr = 1) or ks + t % tpr + p tpr (tpr > 1);
  //   sorted:  as quad, over the block's columns in ascending order (only
  //            where every column is below 2^21).
  // (The strided layout is not chosen here: on this count it looks best for
  // R0 and the prolongations, but its second barrier and pass of products
  // through LDS, which the count leaves out, made it slower on P0 (+7-12 %)
  // and on R0 in FP32 (+3 %); it won on R0 in FP64 (36.2 -> 31.7 us).  It is
  // available through CFDLA_GATHER_STRIDED.)
  // An instruction counts when some lane has a live entry; its lines are the
  // distinct 128-byte lines of x those lanes read.  Measured on the 128^3
  // hierarchy (per launch, FP64 / FP32 us), 0.04 per line orders the three
  // layouts for every operator: the 7-point operator in rows (42.7 / 24.2;
  // sorted 42.8 / 25.3, quad 47.3 / 29.3), R0, P0, R1, P1 and the level-2
  // operator sorted (R0 28.4 / 16.5 against quad 31.8 / 18.8, P1 6.5 / 6.1
  // against 7.3 / 6.9; the others 2-10 % ahead of the better of quad and
  // rows).  The count leaves out what sorted blocks give up, 16-bit columns
  // and a contiguous LDS store; they won regardless.
  constexpr double kLineCost = 0.04;
  const bool sortable = h.m <= (int64_t)kRbColMask + 1;
  const int64_t nrb = (int64_t)rb.size() / 2 - 1;
  if (nrb <= 0) return kRbQuad;
  const int64_t step = std::max<int64_t>(1, nrb / 64);
  const int cap = rb_cap<double>();
  double instr[4] = {0.0, 0.0, 0.0, 0.0}, nl[4] = {0.0, 0.0, 0.0, 0.0};
  std::vector<int64_t> lines, sorted_cols;
  lines.reserve(64);
  auto count = [&](int layout) {
    if (lines.empty()) return;
    std::sort(lines.begin(), lines.end());
    instr[layout] += 1.0;
    nl[layout] += (double)(std::unique(lines.begin(), lines.end()) - lines.begin());
    lines.clear();
  };
  auto line = [&](int64_t k) { return ((int64_t)h.ci[k] * vbytes) >> 7; };
  for (int64_t q = 0; q < nrb; q += step) {
    const int64_t r0 = rb[2 * q], e0 = rb[2 * q + 1], r1 = rb[2 * q + 2], e1 = rb[2 * q + 3];
    const int64_t nr = r1 - r0, ea = e0 & ~int64_t(3);
    if (e1 - e0 > cap) continue;  // a long row: the kernel's other path
    for (int w = 0; w < kBlock / 64; ++w)
      for (int u = 0; u < cap / (4 * kBlock); ++u)
        for (int j = 0; j < 4; ++j) {
          for (int l = 0; l < 64; ++l) {
            const int64_t k = ea + 4 * (64 * w + l + kBlock * u) + j;
            if (k >= e0 && k < e1) lines.push_back(line(k));
          }
          count(kRbQuad);
        }
    if (sortable) {
      sorted_cols.assign(h.ci.begin() + e0, h.ci.begin() + e1);
      std::sort(sorted_cols.begin(), sorted_cols.end());
      for (int w = 0; w < kBlock / 64; ++w)
        for (int u = 0; u < cap / (4 * kBlock); ++u)
          for (int j = 0; j < 4; ++j) {
            for (int l = 0; l < 64; ++l) {
              const int64_t k = ea + 4 * (64 * w + l + kBlock * u) + j;
              if (k >= e0 && k < e1) lines.push_back((sorted_cols[(size_t)(k - e0)] * vbytes) >> 7);
            }
            count(kRbSorted);
          }
    }
    int64_t tpr = 1;
    while (tpr * 2 * nr <= kBlock && tpr < 64) tpr *= 2;
    for (int w = 0; w < kBlock / 64; ++w) {
      int64_t ks[64], len[64], maxp = 0;
      for (int l = 0; l < 64; ++l) {
        const int64_t t = 64 * w + l, row = t / tpr, sub = t % tpr;
        ks[l] = 0;
        len[l] = 0;
        if (row >= nr) continue;
        const int64_t a = h.rp[r0 + row], b = h.rp[r0 + row + 1];
        ks[l] = a + sub;
        len[l] = b - a > sub ? (b - a - sub + tpr - 1) / tpr : 0;
        maxp = std::max(maxp, len[l]);
      }
      for (int64_t p = 0; p < maxp; ++p) {
        for (int l = 0; l < 64; ++l)
          if (p < len[l]) lines.push_back(line(ks[l] + p * tpr));
        count(kRbRows);
      }
    }
  }
  auto cost = [&](int layout) { return instr[layout] + kLineCost * nl[layout]; };
  int best = cost(kRbRows) < cost(kRbQuad) ? kRbRows : kRbQuad;
  if (sortable && cost(kRbSorted) < cost(best)) best = kRbSorted;
  return best;
}

namespace {
// 16-bit column offsets for a row-block partition: base[q] = the smallest
// column of block q, c16[k] = ci[k] - base[q].  False when some block has
// max - min > 65535 over its columns.
bool build_cols16(const HostCsr &h, const std::vector<int32_t> &rb, std::vector<uint16_t> &c16,
                  std::vector<int32_t> &base) {
  const size_t nrb = rb.size() / 2 - 1;
  c16.assign(h.ci.size(), 0);
  base.assign(nrb, 0);
  for (size_t q = 0; q < nrb; ++q) {
    const int64_t e0 = rb[2 * q + 1], e1 = rb[2 * q + 3];
    if (e1 <= e0) continue;
    const auto mm = std::minmax_element(h.ci.begin() + e0, h.ci.begin() + e1);
    if ((int64_t)*mm.second - (int64_t)*mm.first > 65535) return false;
    base[q] = *mm.first;
    for (int64_t k = e0; k < e1; ++k) c16[(size_t)k] = (uint16_t)(h.ci[(size_t)k] - *mm.first);
  }
  return true;
}

void smoother_dinv(const HostCsr &h, std::vector<double> &d) {
  d.assign((size_t)h.n, 0.0);
  for (int64_t i = 0; i < h.n; ++i)
    for (int64_t k = h.rp[i]; k < h.rp[i + 1]; ++k)
      if (h.ci[k] == i) {
        double a = h.v[k];
        d[i] = std::fabs(a) > 1e-15 ? 1.0 / a : 0.0;
      }
}
}  // namespace

namespace {
// Slice start offsets of the sliced-ELL layout with groups of gs slots: slice
// s holds 64 lanes x (its longest row rounded up to gs) slots.  Returns the
// total slot count, or -1 past int32.
int64_t sell_slices(const HostCsr &h, int gs, std::vector<int32_t> &ptr) {
  const int64_t nslices = (h.n + 63) / 64;
  ptr.assign((size_t)nslices + 1, 0);
  int64_t total = 0;
  for (int64_t s = 0; s < nslices; ++s) {
    int64_t w = 0;
    for (int64_t r = s * 64; r < std::min<int64_t>(h.n, s * 64 + 64); ++r) w = std::max<int64_t>(w, h.rp[r + 1] - h.rp[r]);
    total += (w + gs - 1) / gs * gs * 64;
    if (total >= INT32_MAX) return -1;
    ptr[(size_t)s + 1] = (int32_t)total;
  }
  return total;
}

// Position of entry q of the row in lane l of a slice starting at `start`.
inline int64_t sell_pos(int64_t start, int64_t q, int64_t l, int gs) {
  return start + q / gs * 64 * gs + l * gs + q % gs;
}

// Sliced-ELL layout for square matrices of at least 4096 rows when every
// offset fits in int16, rows average at most 64 nonzeros and the padding
// stays under 35 % (7-point stencils: ~0.2 %; the first smoothed-aggregation
// Galerkin level of a 3D Poisson problem: ~28 %).  Groups of 4 slots when
// they stay under that limit, else groups of 1; gs returns the choice.
bool build_sell(const HostCsr &h, std::vector<int32_t> &ptr, std::vector<int16_t> &delta, std::vector<double> &val,
                int &gs) {
  if (h.n != h.m || h.n < 4096) return false;
  if ((double)h.nnz() > 64.0 * (double)h.n) return false;
  for (int64_t r = 0; r < h.n; ++r)
    for (int64_t k = h.rp[r]; k < h.rp[r + 1]; ++k) {
      const int64_t d = (int64_t)h.ci[k] - r;
      if (d < -32767 || d > 32767) return false;
    }
  const double limit = 1.35 * (double)h.nnz() + 64.0 * 16.0;
  int64_t total = -1;
  for (int g : {4, 1}) {
    total = sell_slices(h, g, ptr);
    if (total >= 0 && (double)total <= limit) {
      gs = g;
      break;
    }
    total = -1;
  }
  if (total < 0) return false;
  delta.assign((size_t)total, kSellPad);
  val.assign((size_t)total, 0.0);
  for (int64_t s = 0; s + 1 < (int64_t)ptr.size(); ++s)
    for (int64_t l = 0; l < 64; ++l) {
      const int64_t r = s * 64 + l;
      if (r >= h.n) break;
      const int64_t k0 = h.rp[r];
      for (int64_t k = k0; k < h.rp[r + 1]; ++k) {
        const size_t pos = (size_t)sell_pos(ptr[(size_t)s], k - k0, l, gs);
        delta[pos] = (int16_t)((int64_t)h.ci[k] - r);
        val[pos] = h.v[k];
      }
    }
  return true;
}
}  // namespace

namespace {
// Diagonal layout (see DiaDesc) for square matrices of at least 4096 rows
// whose nonzeros lie on few diagonals filled to at least 90 %.  Returns the
// stored diagonals diagonal-major in val (nd x n).  A matrix that stores an
// explicit zero is left to the indexed formats: the diagonal kernels skip
// zero values (the slots of absent entries), while a CSR product multiplies
// a stored zero, which turns a non-finite x into NaN.
bool build_dia(const HostCsr &h, DiaDesc &d, std::vector<double> &val, bool allow_sym) {
  if (h.n != h.m || h.n < 4096 || h.nnz() == 0) return false;
  for (const double a : h.v)
    if (a == 0.0) return false;
  const int64_t n = h.n;
  std::vector<int64_t> offs;
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = h.rp[i]; k < h.rp[i + 1]; ++k) {
      const int64_t o = (int64_t)h.ci[k] - i;
      if (std::find(offs.begin(), offs.end(), o) == offs.end()) {
        offs.push_back(o);
        if (offs.size() > (size_t)2 * kDiaMax) return false;
      }
    }
  std::sort(offs.begin(), offs.end());
  const int64_t nall = (int64_t)offs.size();
  if ((double)h.nnz() < 0.9 * (double)n * (double)nall) return false;
  if (nall * n >= INT32_MAX) return false;
  std::vector<double> full((size_t)(nall * n), 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = h.rp[i]; k < h.rp[i + 1]; ++k) {
      const int64_t kd = std::lower_bound(offs.begin(), offs.end(), (int64_t)h.ci[k] - i) - offs.begin();
      full[(size_t)(kd * n + i)] = h.v[k];
    }
  auto index_of = [&](int64_t o) -> int64_t {
    auto it = std::lower_bound(offs.begin(), offs.end(), o);
    return it != offs.end() && *it == o ? it - offs.begin() : -1;
  };
  bool sym = allow_sym;
  for (int64_t kd = 0; kd < nall && sym; ++kd) {
    const int64_t o = offs[(size_t)kd];
    if (o <= 0) continue;
    const int64_t km = index_of(-o);
    if (km < 0) {
      sym = false;
      break;
    }
    for (int64_t i = 0; i + o < n; ++i)
      if (!(full[(size_t)(kd * n + i)] == full[(size_t)(km * n + i + o)])) {
        sym = false;
        break;
      }
  }
  std::vector<int64_t> keep;
  if (sym) {
    keep.push_back(index_of(0));  // -1: no main diagonal, stored as zeros
    for (int64_t kd = 0; kd < nall; ++kd)
      if (offs[(size_t)kd] > 0) keep.push_back(kd);
  } else {
    for (int64_t kd = 0; kd < nall; ++kd) keep.push_back(kd);
  }
  if ((int64_t)keep.size() > kDiaMax) return false;
  d = DiaDesc{};
  d.nd = (int)keep.size();
  d.sym = sym ? 1 : 0;
  for (int j = 0; j < d.nd; ++j) d.off[j] = keep[(size_t)j] < 0 ? 0 : (int)offs[(size_t)keep[(size_t)j]];
  if (!dia_supported(d)) return false;
  val.assign((size_t)(d.nd * n), 0.0);
  for (int j = 0; j < d.nd; ++j)
    if (keep[(size_t)j] >= 0)
      std::copy(full.begin() + keep[(size_t)j] * n, full.begin() + (keep[(size_t)j] + 1) * n, val.begin() + j * n);
  return true;
}
}  // namespace

void DevCsr::upload(const HostCsr &h, bool want_fp32, bool want_dinv, bool allow_sell) {
  CFDLA_REQUIRE(h.n < INT32_MAX && h.m < INT32_MAX && h.nnz() < INT32_MAX, CFDLA_ERR_INVALID_ARG,
                "matrix too large for 32-bit device indices");
  n = (int)h.n;
  m = (int)h.m;
  nnz = h.nnz();
  std::vector<int32_t> rp32((size_t)h.n + 1);
  for (int64_t i = 0; i <= h.n; ++i) rp32[i] = (int32_t)h.rp[i];
  rp.upload(rp32.data(), rp32.size());
  std::vector<int32_t> blocks;
  // (Smaller caps for small matrices, to reach 8 blocks per CU, measured
  // slower on every AMG level of the 128^3 problem: 7.6 -> 10.4 us per level-2
  // sweep; the per-block descriptor loads and barriers dominate.)
  std::vector<uint16_t> c16v;
  std::vector<int32_t> cb;
  // with 16-bit columns allowed, a block also closes before its columns
  // would span more than 65535 (the first restriction of the 128^3
  // hierarchy spans up to three planes of 16384 per row)
  const int64_t span = fmt_request != CFDLA_STORAGE_CSR_INT32 ? 65535 : 0;
  build_row_blocks(h, rb_cap<double>(), rb_rows<double>(), blocks, span);
  nrb64 = (int)blocks.size() / 2 - 1;
  rb64.upload(blocks.data(), blocks.size());
  const std::vector<int32_t> blocks64 = blocks;
  const bool allow16 = span > 0;
  c16_f64 = allow16 && build_cols16(h, blocks, c16v, cb);
  if (c16_f64) {
    c16_64.upload_padded(c16v.data(), c16v.size(), kRbPad);
    cbase64.upload(cb.data(), cb.size());
    sync_stream();
  } else {
    c16_64.release();
    cbase64.release();
  }
  c16_f32 = false;
  if (want_fp32) {
    build_row_blocks(h, rb_cap<float>(), rb_rows<float>(), blocks, span);
    nrb32 = (int)blocks.size() / 2 - 1;
    rb32.upload(blocks.data(), blocks.size());
    c16_f32 = allow16 && build_cols16(h, blocks, c16v, cb);
    if (c16_f32) {
      c16_32.upload_padded(c16v.data(), c16v.size(), kRbPad);
      cbase32.upload(cb.data(), cb.size());
      sync_stream();
    }
  }
  if (!c16_f32) {
    c16_32.release();
    cbase32.release();
  }
  ci.upload_padded(h.ci.data(), h.ci.size(), kRbPad);
  v64.upload_padded(h.v.data(), h.v.size(), kRbPad);
  std::vector<float> f;
  if (want_fp32) {
    f.resize(h.v.size());
    for (size_t k = 0; k < f.size(); ++k) f[k] = (float)h.v[k];
    v32.upload_padded(f.data(), f.size(), kRbPad);
  }
  std::vector<double> d;
  std::vector<float> df;
  if (want_dinv) {
    smoother_dinv(h, d);
    sdinv64.upload(d.data(), d.size());
    df.resize(d.size());
    for (size_t k = 0; k < d.size(); ++k) df[k] = (float)d[k];
    sdinv32.upload(df.data(), df.size());
  }
  std::vector<double> dvals;
  const bool want_dia = fmt_request == -1 || fmt_request == CFDLA_STORAGE_DIA;
  const bool want_sell = fmt_request == -1 || fmt_request == CFDLA_STORAGE_SELL64;
  dia = allow_sell && want_dia && build_dia(h, dia_desc, dvals, !general_dia);
  maps_ready = false;
  if (dia) {
    dia_v64.upload(dvals.data(), dvals.size());
    if (want_fp32) {
      std::vector<float> df32(dvals.size());
      for (size_t k = 0; k < df32.size(); ++k) df32[k] = (float)dvals[k];
      dia_v32.upload(df32.data(), df32.size());
      sync_stream();
    }
  }
  if (!dia) {
    dia_v64.release();
    dia_v32.release();
  }
  std::vector<int32_t> sp;
  std::vector<int16_t> sd;
  std::vector<double> sv;
  sell = allow_sell && want_sell && !dia && build_sell(h, sp, sd, sv, sell_group);
  if (!sell) {
    sell_ptr.release();
    sell_delta.release();
    sell_v64.release();
    sell_v32.release();
  }
  if (sell) {
    nslices = (int)(sp.size() - 1);
    sell_entries = sp.back();
    sell_ptr.upload(sp.data(), sp.size());
    sell_delta.upload(sd.data(), sd.size());
    sell_v64.upload(sv.data(), sv.size());
    if (want_fp32) {
      std::vector<float> sf(sv.size());
      for (size_t k = 0; k < sf.size(); ++k) sf[k] = (float)sv[k];
      sell_v32.upload(sf.data(), sf.size());
      sync_stream();
    }
  }
  // gather layout of the row-block product (blocks holds the FP32 partition
  // when there is one); evaluated only for matrices stored as row blocks
  static_assert(CFDLA_GATHER_QUAD == kRbQuad && CFDLA_GATHER_ROWS == kRbRows && CFDLA_GATHER_STRIDED == kRbStrided &&
                    CFDLA_GATHER_SORTED == kRbSorted,
                "gather layout codes");
  auto pick = [&](const std::vector<int32_t> &part, int vbytes) {
    if (gather_request != CFDLA_GATHER_AUTO) return gather_request;
    return !dia && !sell ? rb_pick_layout(h, part, vbytes) : kRbQuad;
  };
  layout64 = pick(blocks64, 8);
  layout32 = want_fp32 ? pick(blocks, 4) : kRbQuad;
  // the column-sorted copy: row-block storage only, one copy for both
  // partitions (cut alike), 32-bit packed columns in place of 16-bit ones
  std::vector<int32_t> packed;
  std::vector<double> svals;
  const bool want_sorted = !dia && !sell && (layout64 == kRbSorted || layout32 == kRbSorted) &&
                           (!want_fp32 || blocks == blocks64) && build_sorted_blocks(h, blocks64, packed, svals);
  if (!want_sorted) {
    if (layout64 == kRbSorted) layout64 = kRbQuad;
    if (layout32 == kRbSorted) layout32 = kRbQuad;
    sci.release();
    sv64.release();
    sv32.release();
  } else {
    sci.upload_padded(packed.data(), packed.size(), kRbPad);
    sv64.upload_padded(svals.data(), svals.size(), kRbPad);
    if (layout32 == kRbSorted) {
      std::vector<float> sf(svals.size());
      for (size_t k = 0; k < sf.size(); ++k) sf[k] = (float)svals[k];
      sv32.upload_padded(sf.data(), sf.size(), kRbPad);
      sync_stream();  // sf is freed at the end of this block
    } else {
      sv32.release();
    }
    if (layout64 == kRbSorted && c16_f64) {
      c16_f64 = false;
      c16_64.release();
      cbase64.release();
    }
    if (layout32 == kRbSorted && c16_f32) {
      c16_f32 = false;
      c16_32.release();
      cbase32.release();
    }
  }
  sync_stream();
}

void DevCsr::build_maps(const HostCsr &h) {
  std::vector<int32_t> pos((size_t)h.nnz(), -1), mirror, dslot((size_t)h.n, -1);
  if (dia && dia_desc.sym) mirror.assign((size_t)h.nnz(), -1);
  for (int64_t i = 0; i < h.n; ++i)
    for (int64_t k = h.rp[i]; k < h.rp[i + 1]; ++k) {
      const int64_t c = h.ci[k];
      if (c == i) dslot[(size_t)i] = (int32_t)k;
      if (dia) {
        const int64_t o = c - i;
        if (dia_desc.sym && o != 0) {
          // the transposed entry (c, i); sorted rows
          auto b = h.ci.begin() + h.rp[c], e = h.ci.begin() + h.rp[c + 1];
          auto it = std::lower_bound(b, e, (int32_t)i);
          mirror[(size_t)k] = it != e && *it == i ? (int32_t)(it - h.ci.begin()) : -2;
          if (o < 0) continue;  // lower entries are read through their upper mirror
        }
        for (int d = 0; d < dia_desc.nd; ++d)
          if (dia_desc.off[d] == o) pos[(size_t)k] = (int32_t)(d * h.n + i);
      }
    }
  if (sell) {
    // the layout of build_sell
    std::vector<int32_t> sp;
    sell_slices(h, sell_group, sp);
    for (int64_t r = 0; r < h.n; ++r)
      for (int64_t k = h.rp[r]; k < h.rp[r + 1]; ++k)
        pos[(size_t)k] = (int32_t)sell_pos(sp[(size_t)(r / 64)], k - h.rp[r], r % 64, sell_group);
  }
  fmt_pos.upload(pos.data(), pos.size());
  if (!mirror.empty()) fmt_mirror.upload(mirror.data(), mirror.size());
  diag_slot.upload(dslot.data(), dslot.size());
  sync_stream();
  maps_ready = true;
}

bool DevCsr::refresh_from_v64() {
  CFDLA_REQUIRE(maps_ready, CFDLA_ERR_INVALID_ARG, "device refresh before its maps were built");
  if (dia && dia_desc.sym && !symmetric_on_device(*this)) return false;
  scatter_formats(*this);
  if (sci.p) sorted_values(*this);
  return true;
}

// ---------------------------------------------------------------------------
// Preconditioners
// ---------------------------------------------------------------------------
void Precond::apply_dot(int n, const double *r, double *z, Scal *s, bool check_done, int dot_op) {
  apply(r, z, s, check_done);
  dot2<double>(n, r, z, nullptr, nullptr, s, check_done, dot_op, ctx().ws);
}
void Precond::apply_dot(int n, const float *r, float *z, Scal *s, bool check_done, int dot_op) {
  apply(r, z, s, check_done);
  dot2<float>(n, r, z, nullptr, nullptr, s, check_done, dot_op, ctx().ws);
}

JacobiPrecond::JacobiPrecond(const HostCsr &a) {
  n = (int)a.n;
  std::vector<double> inv((size_t)a.n, 1.0);
  for (int64_t i = 0; i < a.n; ++i) {
    double d = 0.0;
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k)
      if (a.ci[k] == i) d += a.v[k];
    CFDLA_REQUIRE(std::isfinite(d), CFDLA_ERR_NUMERICAL, "Jacobi preconditioner: non-finite diagonal");
    inv[i] = std::fabs(d) <= 2.220446049250313e-16 ? 1.0 : 1.0 / d;
  }
  std::vector<float> f(inv.size());
  for (size_t i = 0; i < inv.size(); ++i) f[i] = (float)inv[i];
  d64.upload(inv.data(), inv.size());
  d32.upload(f.data(), f.size());
  sync_stream();
}

// z = dinv .* r via the fused CG kernel is the hot path; the standalone apply
// uses a 1x1 "Jacobi sweep" from zero: z = 1 * dinv .* r.
void JacobiPrecond::apply(const double *r, double *z, Scal *s, bool chk) { jacobi_first<double>(n, d64.p, 1.0, r, z, s, chk); }
void JacobiPrecond::apply(const float *r, float *z, Scal *s, bool chk) { jacobi_first<float>(n, d32.p, 1.0, r, z, s, chk); }

int greedy_colouring(const HostCsr &a, std::vector<int32_t> &colour, std::vector<int32_t> &colour_off,
                     std::vector<int32_t> &order) {
  const int n = (int)a.n;
  int ncolours = 0;
  // symmetrized adjacency (couplings in either direction)
  std::vector<int64_t> deg((size_t)n + 1, 0);
  for (int64_t i = 0; i < a.n; ++i)
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k)
      if (a.ci[k] != i) {
        deg[(size_t)i + 1]++;
        deg[(size_t)a.ci[k] + 1]++;
      }
  for (int i = 0; i < n; ++i) deg[(size_t)i + 1] += deg[(size_t)i];
  std::vector<int32_t> adj((size_t)deg[(size_t)n]);
  std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
  for (int64_t i = 0; i < a.n; ++i)
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k)
      if (a.ci[k] != i) {
        adj[(size_t)fill[(size_t)i]++] = a.ci[k];
        adj[(size_t)fill[(size_t)a.ci[k]]++] = (int32_t)i;
      }
  // Greedy colouring in breadth-first order (from the lowest unvisited row of
  // each component): on a bipartite pattern -- any 5-point stencil, identity
  // boundary rows included -- every row's coloured neighbours sit one level
  // up, so it yields the two colours of red-black ordering.
  std::vector<int32_t> bfs;  // visit order; doubles as the queue
  bfs.reserve((size_t)n);
  {
    std::vector<char> queued((size_t)n, 0);
    for (int root = 0; root < n; ++root) {
      if (queued[(size_t)root]) continue;
      queued[(size_t)root] = 1;
      bfs.push_back(root);
      for (size_t head = bfs.size() - 1; head < bfs.size(); ++head) {
        const int32_t i = bfs[head];
        for (int64_t k = deg[(size_t)i]; k < deg[(size_t)i + 1]; ++k) {
          const int32_t j = adj[(size_t)k];
          if (queued[(size_t)j]) continue;
          queued[(size_t)j] = 1;
          bfs.push_back(j);
        }
      }
    }
  }
  colour.assign((size_t)n, -1);
  std::vector<int32_t> seen;  // seen[c] == i: colour c is taken by a neighbour of i
  for (const int i : bfs) {
    for (int64_t k = deg[(size_t)i]; k < deg[(size_t)i + 1]; ++k) {
      const int32_t c = colour[(size_t)adj[(size_t)k]];
      if (c >= 0) {
        if ((size_t)c >= seen.size()) seen.resize((size_t)c + 1, -1);
        seen[(size_t)c] = i;
      }
    }
    int32_t c = 0;
    while ((size_t)c < seen.size() && seen[(size_t)c] == i) ++c;
    colour[(size_t)i] = c;
    ncolours = std::max(ncolours, c + 1);
  }
  colour_off.assign((size_t)ncolours + 1, 0);
  for (int i = 0; i < n; ++i) colour_off[(size_t)colour[(size_t)i] + 1]++;
  for (int c = 0; c < ncolours; ++c) colour_off[(size_t)c + 1] += colour_off[(size_t)c];
  order.assign((size_t)n, 0);
  std::vector<int32_t> pos(colour_off.begin(), colour_off.end() - 1);
  for (int i = 0; i < n; ++i) order[(size_t)pos[(size_t)colour[(size_t)i]]++] = i;
  return ncolours;
}

ColoredSorPrecond::ColoredSorPrecond(const HostCsr &a, double omega) {
  CFDLA_REQUIRE(a.n == a.m, CFDLA_ERR_DIMENSION, "SOR requires a square matrix");
  CFDLA_REQUIRE(std::isfinite(omega) && omega > 0.0 && omega < 2.0, CFDLA_ERR_CONFIG,
                "SOR relaxation factor must lie in (0, 2)");
  n = (int)a.n;
  std::vector<int32_t> colour, order;
  ncolours = greedy_colouring(a, colour, colour_off, order);
  // pivots omega / a_ii; a row without a (nonzero) diagonal takes a unit
  // pivot, as SuccessiveOverRelaxation::from_csr_with_identity_rows does
  std::vector<double> piv((size_t)n, omega);
  for (int64_t i = 0; i < a.n; ++i)
    for (int64_t k = a.rp[i]; k < a.rp[i + 1]; ++k)
      if (a.ci[k] == i && a.v[k] != 0.0) piv[(size_t)i] = omega / a.v[k];
  for (double p : piv) CFDLA_REQUIRE(std::isfinite(p), CFDLA_ERR_NUMERICAL, "SOR: non-finite pivot");
  std::vector<int32_t> rp32((size_t)n + 1);
  for (int i = 0; i <= n; ++i) rp32[(size_t)i] = (int32_t)a.rp[(size_t)i];
  std::vector<float> v32h(a.v.size()), piv32h(piv.size());
  for (size_t k = 0; k < a.v.size(); ++k) v32h[k] = (float)a.v[k];
  for (size_t k = 0; k < piv.size(); ++k) piv32h[k] = (float)piv[k];
  std::vector<uint8_t> first_h((size_t)n);
  for (int i = 0; i < n; ++i) first_h[(size_t)i] = colour[(size_t)i] == 0 ? 1 : 0;
  first.upload(first_h.data(), first_h.size());
  perm.upload(order.data(), order.size());
  rp.upload(rp32.data(), rp32.size());
  ci.upload(a.ci.data(), a.ci.size());
  v64.upload(a.v.data(), a.v.size());
  v32.upload(v32h.data(), v32h.size());
  piv64.upload(piv.data(), piv.size());
  piv32.upload(piv32h.data(), piv32h.size());
  sync_stream();
}

void ColoredSorPrecond::apply(const double *r, double *z, Scal *s, bool chk) {
  sor_first<double>(n, first.p, piv64.p, r, z, s, chk);
  for (int c = 1; c < ncolours; ++c)
    sor_colour<double>(colour_off[(size_t)c + 1] - colour_off[(size_t)c], perm.p + colour_off[(size_t)c], rp.p, ci.p,
                       v64.p, piv64.p, r, z, s, chk);
}

void ColoredSorPrecond::apply(const float *r, float *z, Scal *s, bool chk) {
  sor_first<float>(n, first.p, piv32.p, r, z, s, chk);
  for (int c = 1; c < ncolours; ++c)
    sor_colour<float>(colour_off[(size_t)c + 1] - colour_off[(size_t)c], perm.p + colour_off[(size_t)c], rp.p, ci.p,
                      v32.p, piv32.p, r, z, s, chk);
}

struct IdentityPrecond : Precond {
  int n;
  explicit IdentityPrecond(int n_) : n(n_) {}
  int kind() const override { return 0; }
  void apply(const double *r, double *z, Scal *s, bool chk) override { copy_vec<double>(n, r, z, s, chk); }
  void apply(const float *r, float *z, Scal *s, bool chk) override { copy_vec<float>(n, r, z, s, chk); }
};

template <> double *DevLevel::b<double>() { return b64.p; }
template <> float *DevLevel::b<float>() { return b32.p; }
template <> double *DevLevel::x<double>() { return x64.p; }
template <> float *DevLevel::x<float>() { return x32.p; }
template <> double *DevLevel::t<double>() { return t64.p; }
template <> float *DevLevel::t<float>() { return t32.p; }
template <> double *DevLevel::r<double>() { return r64.p; }
template <> float *DevLevel::r<float>() { return r32.p; }
template <> double *DevLevel::fx<double>() { return fx64.p; }
template <> float *DevLevel::fx<float>() { return fx32.p; }

AmgPrecond::AmgPrecond(const HostCsr &a, const cfdla_amg_config &c) : cfg(c) {
  CFDLA_REQUIRE(a.n == a.m, CFDLA_ERR_DIMENSION, "AMG requires a square matrix");
  CFDLA_REQUIRE(a.n >= 1, CFDLA_ERR_INVALID_ARG, "AMG requires at least one row");
  CFDLA_REQUIRE(c.max_levels >= 1, CFDLA_ERR_CONFIG, "AMG max_levels must be >= 1");
  CFDLA_REQUIRE(c.pre_smooth_iterations >= 0 && c.post_smooth_iterations >= 0, CFDLA_ERR_CONFIG,
                "AMG smoothing counts must be >= 0");
  CFDLA_REQUIRE(c.coarsening >= 0 && c.coarsening <= 3, CFDLA_ERR_CONFIG, "AMG coarsening must be 0, 1, 2 or 3");
  CFDLA_REQUIRE(c.interpolation >= 0 && c.interpolation <= 2, CFDLA_ERR_CONFIG,
                "AMG interpolation must be 0, 1 or 2");
  CFDLA_REQUIRE(c.smoother >= 0 && c.smoother <= 4, CFDLA_ERR_CONFIG, "AMG smoother must be 0, 1, 2, 3 or 4");
  CFDLA_REQUIRE(std::isfinite(c.relaxation_factor) && c.relaxation_factor > 0.0, CFDLA_ERR_CONFIG,
                "AMG relaxation factor must be positive");
  CFDLA_REQUIRE(c.coarse_solver == 0 || c.coarse_solver == 1, CFDLA_ERR_CONFIG, "AMG coarse_solver must be 0 or 1");
  auto t0 = std::chrono::steady_clock::now();
  GalerkinFn galerkin;
  if (cfg.galerkin_on_device) {
    ensure_init();
    galerkin = [this](size_t l, const HostCsr &R, const HostCsr &A, const HostCsr &P) {
      struct Dev {
        DevBuf<int32_t> rp, ci;
        DevBuf<double> v;
        explicit Dev(const HostCsr &h) {
          CFDLA_REQUIRE(h.nnz() < INT32_MAX && h.n < INT32_MAX, CFDLA_ERR_INVALID_ARG,
                        "matrix too large for 32-bit device indices");
          std::vector<int32_t> rp32(h.rp.begin(), h.rp.end());
          rp.upload(rp32.data(), rp32.size());
          ci.upload(h.ci.data(), h.ci.size());
          v.upload(h.v.data(), h.v.size());
          sync_stream();
        }
      };
      const Dev r(R), a(A), p(P);
      if (gplan.size() <= l) gplan.resize(l + 1);
      GalerkinPlan &g = gplan[l];
      DevBuf<int32_t> ra_rp, ra_ci, c_rp, c_ci;
      DevBuf<double> c_v;
      device_product((int)R.n, (int)A.m, r.rp.p, r.ci.p, r.v.p, a.rp.p, a.ci.p, a.v.p, g.ra, ra_rp, ra_ci, g.ra_v);
      device_product((int)R.n, (int)P.m, ra_rp.p, ra_ci.p, g.ra_v.p, p.rp.p, p.ci.p, p.v.p, g.rap, c_rp, c_ci, c_v);
      HostCsr ac;
      ac.n = R.n;
      ac.m = P.m;
      std::vector<int32_t> rp32((size_t)ac.n + 1);
      c_rp.download(rp32.data(), rp32.size());
      sync_stream();
      ac.rp.assign(rp32.begin(), rp32.end());
      const size_t nnz = (size_t)ac.rp.back();
      ac.ci.resize(nnz);
      ac.v.resize(nnz);
      c_ci.download(ac.ci.data(), nnz);
      c_v.download(ac.v.data(), nnz);
      sync_stream();
      return ac;
    };
  }
  build_hierarchy(a, cfg, host, galerkin);
  // plans of products whose level was kept; a final empty product is dropped
  if (galerkin) gplan.resize(host.levels.size() - 1);
  setup_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void AmgPrecond::upload() {
  sync_host();
  const size_t L = host.levels.size();
  lv.clear();
  lv.resize(L);
  for (size_t l = 0; l < L; ++l) {
    HostLevel &h = host.levels[l];
    DevLevel &d = lv[l];
    d.A.gather_request = d.P.gather_request = d.R.gather_request = gather_request;
    d.A.upload(h.A, true, true);
    if (l + 1 < L) {
      d.P.upload(h.P, true, false);
      d.R.upload(h.R, true, false);
    }
    if (cfg.smoother != 2) {
      std::vector<int32_t> colour, order;
      greedy_colouring(h.A, colour, d.gs_off, order);
      d.gs_perm.upload(order.data(), std::max<size_t>(1, order.size()));
    }
    size_t n = (size_t)h.A.n;
    if (l > 0) {
      d.b64.alloc(n); d.x64.alloc(n); d.b32.alloc(n); d.x32.alloc(n);
    }
    d.t64.alloc(n); d.r64.alloc(n); d.t32.alloc(n); d.r32.alloc(n);
  }
  fused = FusedOp{};
  if (L >= 3 && cfg.smoother == 2 && cfg.coarse_solver == 1 && cfg.pre_smooth_iterations == 2 &&
      !cfg.separate_coarse_restriction)
    build_fused(L - 2);
  coarse_n = (int)host.levels.back().A.n;
  if (cfg.coarse_solver == 1) {
    CFDLA_REQUIRE(coarse_n <= 8192, CFDLA_ERR_CONFIG,
                  "AMG coarsest level too large for the dense direct solve; raise max_levels");
    std::vector<double> inv = dense_inverse(host.levels.back().A);
    coarse_inv.upload(inv.data(), inv.size());
  }
  sync_stream();
  on_device = true;
}

void AmgPrecond::update_values(const double *values) {
  recompute_hierarchy(host, values);
  host_stale = false;
  if (on_device) upload();
}

namespace {
// Rewrites the device formats of a level from its v64; a symmetric diagonal
// layout whose values stopped being symmetric is rebuilt (once) as general.
void refresh_level(DevCsr &d, HostCsr &h, bool want_dinv) {
  if (!d.maps_ready) d.build_maps(h);
  if (d.refresh_from_v64()) return;
  d.v64.download(h.v.data(), h.v.size());
  sync_stream();
  d.general_dia = true;
  d.upload(h, true, want_dinv);
}

void put_plan(DevPlan &d, const ProductPlan &p) {
  d.nslots = (int)p.ptr.size() - 1;
  d.ptr.upload(p.ptr.data(), p.ptr.size());
  d.pa.upload(p.pa.data(), p.pa.size());
  d.pb.upload(p.pb.data(), p.pb.size());
  sync_stream();
}
}  // namespace

bool AmgPrecond::build_fused(size_t l) {
  const HostCsr &A = host.levels[l].A, &R = host.levels[l].R;
  for (int64_t i = 0; i < A.n; ++i) {
    bool diag = false;
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) diag = diag || A.ci[k] == i;
    if (!diag) return false;
  }
  std::vector<double> d;
  smoother_dinv(A, d);
  const double w = cfg.relaxation_factor;
  FusedOp f;
  f.E = A;
  HostCsr S = A;
  for (int64_t i = 0; i < A.n; ++i)
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
      const int64_t j = A.ci[k];
      f.E.v[k] = fused_e(i == j, w, A.v[k], d[j]);
      S.v[k] = fused_s(i == j, w, A.v[k], d[i], d[j]);
    }
  f.T = csr_multiply(R, f.E);
  f.M = csr_multiply(f.T, f.E);
  // F stands in for a sweep and a residual over A and the restriction R;
  // keep the three products when it would stream more than twice their bytes
  if ((double)(S.nnz() + f.M.nnz()) > 2.0 * (2.0 * (double)A.nnz() + (double)R.nnz())) return false;
  // M's rows (thousands of entries on the level above a 641-row coarsest
  // level) are cut into `parts` pieces of at most one row block's capacity,
  // so they pack into ordinary row blocks instead of running as single-
  // block long rows (31 against 13 us FP64 at 128^3); the dense coarse solve
  // adds each row's pieces.  Pieces are consecutive slices of the row, so
  // F's values are S's followed by M's unchanged.
  const int64_t piece = rb_cap<double>();
  int64_t parts = 1;
  for (int64_t j = 0; j < f.M.n; ++j) parts = std::max(parts, (f.M.rp[j + 1] - f.M.rp[j] + piece - 1) / piece);
  HostCsr &F = f.F;
  F.n = A.n + f.M.n * parts;
  F.m = A.m;
  F.rp = S.rp;
  F.rp.reserve((size_t)F.n + 1);
  for (int64_t j = 0; j < f.M.n; ++j)
    for (int64_t k = 1; k <= parts; ++k)
      F.rp.push_back(S.nnz() + std::min(f.M.rp[j + 1], f.M.rp[j] + k * piece));
  F.ci = S.ci;
  F.ci.insert(F.ci.end(), f.M.ci.begin(), f.M.ci.end());
  F.v = S.v;
  F.v.insert(F.v.end(), f.M.v.begin(), f.M.v.end());
  DevLevel &dl = lv[l];
  dl.F.gather_request = gather_request;
  dl.F.upload(F, true, false, false);
  dl.fx64.alloc((size_t)F.n);
  dl.fx32.alloc((size_t)F.n);
  f.level = (int)l;
  f.parts = (int)parts;
  fused = std::move(f);
  return true;
}

void AmgPrecond::refresh_fused_device() {
  const size_t l = (size_t)fused.level;
  DevLevel &d = lv[l];
  FusedOp &f = fused;
  if (!f.plans) {
    put_plan(f.t, product_plan(host.levels[l].R, f.E, f.T));
    put_plan(f.m, product_plan(f.T, f.E, f.M));
    f.e64.alloc((size_t)f.E.nnz());
    f.t64.alloc((size_t)f.T.nnz());
    f.plans = true;
  }
  // S into F's first rows, E beside; then T = R E and M = T E into F's last rows
  fused_values(d.A, cfg.relaxation_factor, d.F.v64.p, f.e64.p);
  product_numeric(f.t.nslots, f.t.ptr.p, f.t.pa.p, f.t.pb.p, d.R.v64.p, f.e64.p, f.t64.p);
  product_numeric(f.m.nslots, f.m.ptr.p, f.m.pa.p, f.m.pb.p, f.t64.p, f.e64.p, d.F.v64.p + f.E.nnz());
  refresh_level(d.F, f.F, false);
}

void AmgPrecond::update_values_device(const double *fine_values) {
  ensure_device();
  const size_t L = lv.size();
  DevCsr &a0 = lv[0].A;
  if (a0.nnz > 0 && a0.v64.p != fine_values)
    CFDLA_HIP(hipMemcpyAsync(a0.v64.p, fine_values, (size_t)a0.nnz * sizeof(double), hipMemcpyDeviceToDevice,
                             stream()));
  refresh_level(a0, host.levels[0].A, true);
  if (gplan.size() != L - 1) {
    // per-slot product lists of the host's two products per level (host
    // work once per hierarchy; the intermediate R_l A_l pattern is the
    // host product's)
    gplan.clear();
    gplan.resize(L - 1);
    for (size_t l = 0; l + 1 < L; ++l) {
      const HostLevel &h = host.levels[l];
      const HostCsr ra_pat = csr_multiply(h.R, h.A);
      put_plan(gplan[l].ra, product_plan(h.R, h.A, ra_pat));
      put_plan(gplan[l].rap, product_plan(ra_pat, h.P, host.levels[l + 1].A));
      gplan[l].ra_v.alloc((size_t)ra_pat.nnz());
    }
  }
  for (size_t l = 0; l + 1 < L; ++l) {
    DevLevel &f = lv[l], &c = lv[l + 1];
    GalerkinPlan &g = gplan[l];
    product_numeric(g.ra.nslots, g.ra.ptr.p, g.ra.pa.p, g.ra.pb.p, f.R.v64.p, f.A.v64.p, g.ra_v.p);
    product_numeric(g.rap.nslots, g.rap.ptr.p, g.rap.pa.p, g.rap.pb.p, g.ra_v.p, f.P.v64.p, c.A.v64.p);
    refresh_level(c.A, host.levels[l + 1].A, true);
  }
  if (fused.level >= 0) refresh_fused_device();
  if (cfg.coarse_solver == 1) {
    const DevCsr &cA = lv.back().A;
    dense_inverse_device(coarse_n, cA.rp.p, cA.ci.p, cA.v64.p, coarse_inv.p);
  }
  host_stale = true;
}

void AmgPrecond::sync_host() {
  if (!host_stale) return;
  for (size_t l = 0; l < lv.size(); ++l) lv[l].A.v64.download(host.levels[l].A.v.data(), host.levels[l].A.v.size());
  sync_stream();
  host_stale = false;
}

template <typename T>
T *AmgPrecond::first_sweep_target(int l) {
  DevLevel &L = lv[(size_t)l];
  const bool coarsest = (size_t)l + 1 == lv.size();
  if (coarsest && cfg.coarse_solver == 1) return nullptr;
  if (l == fused.level) return nullptr;  // the fused product starts from b alone
  const int down = coarsest ? cfg.pre_smooth_iterations + 10 : cfg.pre_smooth_iterations;
  const int up = coarsest ? 0 : cfg.post_smooth_iterations;
  if (down < 1 || (down >= 2 && L.A.nnz <= 8 * (int64_t)L.A.n)) return nullptr;
  const int flips = down - 1 + up;
  return flips % 2 == 0 ? L.x<T>() : L.t<T>();
}

template <typename T>
void AmgPrecond::gs_sweeps(DevLevel &L, const T *b, T *x, int sweeps, Scal *s, bool chk) {
  const int nc = (int)L.gs_off.size() - 1;
  const double w = cfg.relaxation_factor;
  auto colour = [&](int c) {
    gs_colour<T>(L.gs_off[(size_t)c + 1] - L.gs_off[(size_t)c], L.gs_perm.p + L.gs_off[(size_t)c], L.A.rp.p, L.A.ci.p,
                 L.A.vals<T>(), b, x, w, s, chk);
  };
  for (int k = 0; k < sweeps; ++k)
    for (int c = 0; c < nc; ++c) colour(c);
  if (cfg.smoother == 1)  // symmetric: the backward sweeps follow (smoothers.rs:97-137)
    for (int k = 0; k < sweeps; ++k)
      for (int c = nc - 1; c >= 0; --c) colour(c);
}

template <typename T>
void AmgPrecond::vcycle_gs(int l, const T *b, T *x, Scal *s, bool chk, int dot_op) {
  DevLevel &L = lv[(size_t)l];
  RedWs &ws = ctx().ws;
  const int n = L.A.n;
  const bool coarsest = (size_t)l + 1 == lv.size();
  if (coarsest && cfg.coarse_solver == 1) {
    dense_gemv<T>(n, coarse_inv.p, b, x, s, chk);
  } else {
    // amg.rs:423-503 from the zero guess (amg.rs:553): pre-smoothing, and on
    // the coarsest level 10 more sweeps (amg.rs:501)
    zero_vec<T>(n, x, s, chk);
    gs_sweeps<T>(L, b, x, cfg.pre_smooth_iterations, s, chk);
    if (coarsest) {
      gs_sweeps<T>(L, b, x, 10, s, chk);
    } else {
      DevLevel &C = lv[(size_t)l + 1];
      csr_apply<T>(L.A, EPI_RESID, x, b, nullptr, 0.0, L.r<T>(), s, chk, OP_NONE, nullptr, ws);
      csr_apply<T>(L.R, EPI_Y, L.r<T>(), nullptr, nullptr, 0.0, C.b<T>(), s, chk, OP_NONE, nullptr, ws);
      vcycle_gs<T>(l + 1, C.b<T>(), C.x<T>(), s, chk, OP_NONE);
      csr_apply<T>(L.P, EPI_ADD, C.x<T>(), nullptr, nullptr, 0.0, x, s, chk, OP_NONE, nullptr, ws);
      gs_sweeps<T>(L, b, x, cfg.post_smooth_iterations, s, chk);
    }
  }
  if (dot_op != OP_NONE) dot2<T>(n, b, x, nullptr, nullptr, s, chk, dot_op, ws);
}

template <typename T>
void AmgPrecond::vcycle(int l, const T *b, T *x, Scal *s, bool chk, int dot_op, bool first_ready) {
  if (cfg.smoother != 2) {
    vcycle_gs<T>(l, b, x, s, chk, dot_op);
    return;
  }
  DevLevel &L = lv[(size_t)l];
  RedWs &ws = ctx().ws;
  const int n = L.A.n;
  const bool coarsest = (size_t)l + 1 == lv.size();
  const double w = cfg.relaxation_factor;
  const T *dinv = L.A.sdinv<T>();
  if (coarsest && cfg.coarse_solver == 1) {
    dense_gemv<T>(n, coarse_inv.p, b, x, s, chk);
    if (dot_op != OP_NONE) dot2<T>(n, b, x, nullptr, nullptr, s, chk, dot_op, ctx().ws);
    return;
  }
  // amg.rs:423-503 with Jacobi: pre-smoothing from the zero guess, then on
  // the coarsest level 10 more sweeps (amg.rs:501).
  const int down = coarsest ? cfg.pre_smooth_iterations + 10 : cfg.pre_smooth_iterations;
  const int up = coarsest ? 0 : cfg.post_smooth_iterations;
  if (l == fused.level) {
    // the two pre-sweeps from zero, the residual and the restriction as one
    // product: [x; pieces of b_c] = F b (fused_jacobi.hpp), then the dense
    // coarse solve (adding the pieces), the prolongation and the
    // post-sweeps, the last one into x
    DevLevel &C = lv[(size_t)l + 1];
    T *fx = L.fx<T>();
    csr_apply<T>(L.F, EPI_Y, b, nullptr, nullptr, 0.0, fx, s, chk, OP_NONE, nullptr, ws);
    dense_gemv<T>(C.A.n, coarse_inv.p, fx + n, C.x<T>(), s, chk, fused.parts);
    csr_apply<T>(L.P, EPI_ADD, C.x<T>(), nullptr, nullptr, 0.0, fx, s, chk, OP_NONE, nullptr, ws);
    const T *src = fx;
    for (int j = 1; j <= up; ++j) {
      T *dst = (up - j) % 2 == 0 ? x : L.t<T>();
      const int op = j == up ? dot_op : OP_NONE;
      csr_apply<T>(L.A, EPI_JACOBI, src, b, dinv, w, dst, s, chk, op, op != OP_NONE ? b : nullptr, ws);
      src = dst;
    }
    if (up == 0) copy_vec<T>(n, fx, x, s, chk);
    if (up == 0 && dot_op != OP_NONE) dot2<T>(n, b, x, nullptr, nullptr, s, chk, dot_op, ws);
    return;
  }
  T *buf[2] = {x, L.t<T>()};
  // number of buffer flips after the first sweep; the iterate must end in x
  const int flips = (down > 0 ? down - 1 : 0) + up;
  int cur = (flips % 2 == 0) ? 0 : 1;
  int k = 1;
  // Fusing the zero-guess sweep into the next one gathers dinv and b at
  // every column instead of x: a win on stencil rows (<= 8 entries), a loss
  // on the long rows of Galerkin levels (measured 44.7 us fused against
  // 26.2 + ~3 us unfused on the first coarse level of the 128^3 problem).
  if (down >= 2 && L.A.nnz <= 8 * (int64_t)n) {
    // the zero-guess sweep and the next one in a single pass
    csr_apply<T>(L.A, EPI_JACOBI0, b, b, dinv, w, buf[cur ^ 1], s, chk, OP_NONE, nullptr, ws);
    cur ^= 1;
    k = 2;
  } else if (down >= 1) {
    if (!first_ready) jacobi_first<T>(n, dinv, w, b, buf[cur], s, chk);
  } else {
    zero_vec<T>(n, buf[cur], s, chk);
  }
  for (; k < down; ++k) {
    csr_apply<T>(L.A, EPI_JACOBI, buf[cur], b, dinv, w, buf[cur ^ 1], s, chk, OP_NONE, nullptr, ws);
    cur ^= 1;
  }
  if (!coarsest) {
    DevLevel &C = lv[(size_t)l + 1];
    csr_apply<T>(L.A, EPI_RESID, buf[cur], b, nullptr, 0.0, L.r<T>(), s, chk, OP_NONE, nullptr, ws);
    // b_c = R r; when the coarse level starts with a plain sweep from zero,
    // the restriction also writes that first iterate, omega D_c^-1 b_c
    T *first = first_sweep_target<T>(l + 1);
    if (first)
      csr_apply<T>(L.R, EPI_RJ0, L.r<T>(), nullptr, C.A.sdinv<T>(), w, C.b<T>(), s, chk, OP_NONE, nullptr, ws,
                   first);
    else
      csr_apply<T>(L.R, EPI_Y, L.r<T>(), nullptr, nullptr, 0.0, C.b<T>(), s, chk, OP_NONE, nullptr, ws);
    vcycle<T>(l + 1, C.b<T>(), C.x<T>(), s, chk, OP_NONE, first != nullptr);
    csr_apply<T>(L.P, EPI_ADD, C.x<T>(), nullptr, nullptr, 0.0, buf[cur], s, chk, OP_NONE, nullptr, ws);
    for (int j = 0; j < up; ++j) {
      // the last sweep writes x (the flip count fixes the parity) and, when
      // asked, also reduces b . x and runs dot_op
      const bool last = j + 1 == up && (cur ^ 1) == 0;
      const int op = last ? dot_op : OP_NONE;
      csr_apply<T>(L.A, EPI_JACOBI, buf[cur], b, dinv, w, buf[cur ^ 1], s, chk, op, op != OP_NONE ? b : nullptr,
                   ws);
      if (last) dot_op = OP_NONE;
      cur ^= 1;
    }
  }
  if (cur != 0) copy_vec<T>(n, buf[cur], x, s, chk);
  if (dot_op != OP_NONE) dot2<T>(n, b, x, nullptr, nullptr, s, chk, dot_op, ws);
}

template void AmgPrecond::vcycle<double>(int, const double *, double *, Scal *, bool, int, bool);
template void AmgPrecond::vcycle_gs<double>(int, const double *, double *, Scal *, bool, int);
template void AmgPrecond::vcycle_gs<float>(int, const float *, float *, Scal *, bool, int);
template void AmgPrecond::vcycle<float>(int, const float *, float *, Scal *, bool, int, bool);

// ---------------------------------------------------------------------------
// Krylov drivers
// ---------------------------------------------------------------------------
namespace {

template <typename T> T *vec(int slot, size_t n);
template <> double *vec<double>(int slot, size_t n) { return ctx().dv_at(slot, n); }
template <> float *vec<float>(int slot, size_t n) { return ctx().fv_at(slot, n); }

template <typename T> const T *jdinv(const Precond *M);
template <> const double *jdinv<double>(const Precond *M) { return M->jacobi_dinv64(); }
template <> const float *jdinv<float>(const Precond *M) { return M->jacobi_dinv32(); }

// From the lagged poll: does the chunk enqueued last probably reach the
// threshold?  Near convergence the residual falls about geometrically, so
// rr_next ~ rr (rr / rr_prev).  The host then waits for that chunk's state
// instead of enqueuing another chunk whose kernels would only exit early.
struct ConvergenceGate {
  double prev = -1.0;
  bool expect(double rr, double thr2) {
    const bool hit = prev > 0.0 && rr > 0.0 && rr < prev && rr * (rr / prev) <= thr2;
    prev = rr;
    return hit;
  }
};

int auto_chunk(const Precond *M, int n) {
  if (M->kind() == 2) return 1;
  return n > 1000000 ? 2 : 8;
}

// r = b - A x, returns ||r||_2 (synchronous).
template <typename T>
double true_residual(const DevCsr &A, const T *b, const T *x, T *r, Scal *s) {
  csr_apply<T>(A, EPI_RESID, x, b, nullptr, 0.0, r, s, false, OP_INIT_RR, r, ctx().ws);
  Scal h = ctx().read(s);
  return std::sqrt(h.rr);
}

// Preconditioned CG (Saad 2003, Alg. 9.1).  confirm: check the explicit
// residual before reporting CONVERGED and restart from it otherwise.
template <typename T>
SolveOut pcg(const DevCsr &A, Precond *M, const T *b, T *x, double thr, int64_t max_it, int chunk, bool confirm) {
  Context &c = ctx();
  RedWs &ws = c.ws;
  const int n = A.n;
  T *r = vec<T>(0, n), *z = vec<T>(1, n), *p = vec<T>(2, n), *q = vec<T>(3, n);
  // Diagonal storage folds p = z + beta p into the next product (EPI_DIR),
  // which writes the new direction into the other buffer of a pair.
  const bool fuse = fuses_direction<T>(A);
  T *pn = fuse ? vec<T>(4, n) : nullptr;
  bool fresh = true;  // p was just set from z: the next product takes it as is
  Scal *s = c.s(0);
  SolveOut o;
  o.thr = thr;
  scal_reset(s, thr * thr);
  o.rn0 = o.rn = true_residual<T>(A, b, x, r, s);
  if (o.rn0 <= thr) {
    o.term = CFDLA_TERM_INITIAL_RESIDUAL;
    return o;
  }
  const T *jd = jdinv<T>(M);
  auto start_direction = [&]() {
    M->apply_dot(n, r, z, s, false, OP_SET_RZ);
    copy_vec<T>(n, z, p, s, false);
    fresh = true;
  };
  start_direction();
  if (chunk <= 0) chunk = auto_chunk(M, n);
  int64_t enq = 0;
  Poller poll(s);
  ConvergenceGate gate;
  bool finished = false;
  while (!finished) {
    const Scal *h = nullptr;
    if (enq < max_it) {
      int64_t steps = std::min<int64_t>(chunk, max_it - enq);
      for (int64_t k = 0; k < steps; ++k) {
        if (fuse && !fresh) {
          csr_apply<T>(A, EPI_DIR, p, z, nullptr, 0.0, q, s, true, OP_CG_ALPHA, nullptr, ws, pn);
          std::swap(p, pn);
        } else {
          csr_apply<T>(A, EPI_Y, p, nullptr, nullptr, 0.0, q, s, true, OP_CG_ALPHA, p, ws);
        }
        fresh = false;
        if (jd) {
          cg_update<T>(n, p, q, x, r, jd, z, s, ws);
        } else {
          cg_update<T>(n, p, q, x, r, nullptr, nullptr, s, ws);
          M->apply_dot(n, r, z, s, true, OP_CG_BETA);
        }
        if (!fuse) cg_direction<T>(n, z, p, s);
      }
      enq += steps;
      poll.post();
      h = poll.lagged();
    } else {
      h = poll.latest();
    }
    if (!h) continue;
    if (h->done == 0 && enq < max_it) {
      if (!gate.expect(h->rr, thr * thr)) continue;
      h = poll.latest();
      if (h->done == 0) continue;
    }
    sync_stream();
    Scal now = c.read(s);
    o.iters = now.iters;
    if (now.done == 2) {
      o.term = now.term;
      finished = true;
    } else if (now.done == 1) {
      if (!confirm) {
        o.term = CFDLA_TERM_CONVERGED;
        o.rn = std::sqrt(now.rr);
        return o;
      }
      double tn = true_residual<T>(A, b, x, r, s);
      if (tn <= thr) {
        o.term = CFDLA_TERM_CONVERGED;
        o.rn = tn;
        return o;
      }
      scal_set_done(s, 0);  // residual replacement: restart from the explicit residual
      start_direction();
      poll = Poller(s);
    } else {
      finished = true;  // budget exhausted
    }
  }
  o.rn = true_residual<T>(A, b, x, r, s);
  return o;
}

// FP32-inner PCG with reliable updates (residual replacement): the FP32
// recurrence runs uninterrupted -- no restart, so the Krylov space and its
// convergence rate survive -- while its solution increments accumulate in an
// FP32 correction xc that is folded into the FP64 solution, and the FP32
// residual is replaced by the FP64 residual b - A x (Sleijpen & van der
// Vorst 1996; van der Vorst & Ye 2000; Clark et al., Comput. Phys. Commun.
// 181 (2010) 1517, mixed-precision CG with reliable updates):
//   * interval > 0: every `interval` iterations, between the residual
//     update of an iteration and its preconditioner application, enqueued
//     with the iteration (skipped once the solve has stopped);
//   * delta > 0: in the next enqueued iteration once the lagged host poll
//     shows the recursive residual at or below delta times the latest FP64
//     residual (the reduction criterion of Clark et al.), delta widened
//     tenfold (up to 0.5) after each replacement that restarted the direction;
//   * when the FP32 recursive residual passes the FP64 threshold (done = 1,
//     seen by the lagged host poll), to confirm convergence.
// A replacement leaves rz and p in place: z = M r_new, beta = r_new.z / rz,
// p = z + beta p continue the iteration -- unless the FP64 residual comes out
// above twice the recursive one it replaces, when beta = 0 restarts the
// direction (the FP32 operator's rounding bounds the true-residual gain of
// one cycle near eps32 * cond(A); past that the recursive residual is
// fiction and beta would blow up).  The FP64 residual decides
// convergence (OP_RELIABLE: done = 4), so CONVERGED is always backed by an
// explicit FP64 residual.
SolveOut pcg_reliable(const DevCsr &A, Precond *M, const double *b, double *x, double thr, int64_t max_it,
                      int interval, double delta, int chunk) {
  CFDLA_REQUIRE(A.v32.p != nullptr, CFDLA_ERR_CONFIG, "matrix has no FP32 copy");
  Context &c = ctx();
  RedWs &ws = c.ws;
  const int n = A.n;
  double *r64 = c.dv_at(8, n);
  float *r = c.fv_at(0, n), *z = c.fv_at(1, n), *p = c.fv_at(2, n), *q = c.fv_at(3, n), *xc = c.fv_at(4, n);
  // Diagonal storage folds p = z + beta p into the next product (EPI_DIR)
  const bool fuse = fuses_direction<float>(A);
  float *pn = fuse ? c.fv_at(5, n) : nullptr;
  bool fresh = true;  // p was just set from z: the next product takes it as is
  Scal *s = c.s(0);
  SolveOut o;
  o.thr = thr;
  scal_reset(s, thr * thr);
  o.rn0 = o.rn = true_residual<double>(A, b, x, r64, s);
  if (o.rn0 <= thr) {
    o.term = CFDLA_TERM_INITIAL_RESIDUAL;
    return o;
  }
  f64_to_f32(n, r64, r, s, false);
  fill_vec<float>(n, 0.0f, xc);
  const float *jd = jdinv<float>(M);
  // z = M r (+ rz) and p = z
  M->apply_dot(n, r, z, s, false, OP_SET_RZ);
  copy_vec<float>(n, z, p, s, false);
  if (chunk <= 0) chunk = auto_chunk(M, n);
  // FP64 residual replacement; check_done skips it once the solve stopped.
  // folded: the step's update has already moved the correction into x.
  auto replace = [&](bool check_done, bool folded = false) {
    if (!folded) fold_correction(n, xc, x, s, check_done);
    csr_apply<double>(A, EPI_RESID, x, b, nullptr, 0.0, r64, s, check_done, OP_RELIABLE, r64, ws);
    f64_to_f32(n, r64, r, s, check_done);
  };
  // the preconditioner half of an iteration: z = M r, beta, p = z + beta p
  auto direction = [&]() {
    if (jd) {
      jacobi_first<float>(n, jd, 1.0, r, z, s, true);
      dot2<float>(n, r, z, nullptr, nullptr, s, true, OP_CG_BETA, ws);
    } else {
      M->apply_dot(n, r, z, s, true, OP_CG_BETA);
    }
    if (!fuse) cg_direction<float>(n, z, p, s);
  };
  int64_t enq = 0;
  int32_t replacements = 0;  // enqueued, including those the device skips after stopping
  bool want = false;         // a reduction by delta was seen: replace in the next step
  Poller poll(s);
  ConvergenceGate gate;
  for (;;) {
    const Scal *h = nullptr;
    if (enq < max_it) {
      const int64_t steps = std::min<int64_t>(chunk, max_it - enq);
      for (int64_t k = 0; k < steps; ++k) {
        if (fuse && !fresh) {
          csr_apply<float>(A, EPI_DIR, p, z, nullptr, 0.0, q, s, true, OP_CG_ALPHA, nullptr, ws, pn);
          std::swap(p, pn);
        } else {
          csr_apply<float>(A, EPI_Y, p, nullptr, nullptr, 0.0, q, s, true, OP_CG_ALPHA, p, ws);
        }
        fresh = false;
        const bool rep = want || (interval > 0 && (enq + k + 1) % interval == 0);
        if (rep) ++replacements;
        want = false;
        if (jd && !rep) {
          // xc += alpha p, r -= alpha q, z = D^-1 r, rr and rz -> beta in one pass
          cg_update<float>(n, p, q, xc, r, jd, z, s, ws);
          if (!fuse) cg_direction<float>(n, z, p, s);
        } else if (rep) {
          // xc + alpha p into x (xc = 0), r -= alpha q, rr; then the FP64
          // residual.  A step whose rr stops the solve skips the residual;
          // the fold it already did leaves the final fold nothing to add.
          cg_update_fold(n, p, q, xc, x, r, s, ws);
          replace(true, true);
          direction();
        } else {
          cg_update<float>(n, p, q, xc, r, nullptr, nullptr, s, ws);  // xc += alpha p, r -= alpha q, rr
          direction();
        }
      }
      enq += steps;
      poll.post();
      h = poll.lagged();
    } else {
      h = poll.latest();
    }
    if (!h) continue;
    if (h->done == 0 && enq < max_it) {
      // Reduction trigger (Clark et al. 2010): once every replacement enqueued
      // so far shows in the polled state, replace again when the recursive
      // residual has dropped by delta since the latest FP64 residual.
      // Each replacement that had to restart the direction shows a cycle
      // longer than FP32 precision carries: widen delta tenfold (at most 0.5).
      const double ref = h->nupd > 0 ? h->rr_true : o.rn0 * o.rn0;
      const double de = std::min(0.5, delta * std::pow(10.0, (double)h->nrestart));
      if (delta > 0.0 && h->nupd >= replacements && h->rr <= de * de * ref) want = true;
      if (!gate.expect(h->rr, thr * thr)) continue;
      h = poll.latest();
      if (h->done == 0) continue;
    }
    if (h->done == 1 && enq < max_it) {
      // the FP32 residual passed the threshold: confirm with the FP64 one,
      // read it before enqueuing more work, and continue the iteration from
      // it if it is not there yet
      ++replacements;
      replace(false);
      const Scal conf = c.read(s);
      if (conf.done != 4) {
        direction();
        poll = Poller(s);
        gate = ConvergenceGate();
        poll.post();  // the replaced state, so latest() is defined once the budget is spent
        continue;
      }
    }
    sync_stream();
    Scal now = c.read(s);
    o.iters = now.iters;
    if (now.done == 4) {
      o.term = CFDLA_TERM_CONVERGED;
      o.outer = now.nupd;
      o.rn = std::sqrt(now.rr_true);
      return o;
    }
    // stopped (breakdown, curvature) or out of budget: fold what the FP32
    // recurrence has and report the FP64 residual
    fold_correction(n, xc, x, s, false);
    o.rn = true_residual<double>(A, b, x, r64, s);
    o.outer = now.nupd + 1;
    if (now.done == 2) o.term = now.term;
    else o.term = o.rn <= thr ? CFDLA_TERM_CONVERGED : CFDLA_TERM_MAX_ITERATIONS;
    return o;
  }
}

// Right-preconditioned BiCGSTAB (Saad 2003, Alg. 7.7).
template <typename T>
SolveOut bicgstab(const DevCsr &A, Precond *M, const T *b, T *x, double thr, int64_t max_it, int chunk,
                  bool confirm) {
  Context &c = ctx();
  RedWs &ws = c.ws;
  const int n = A.n;
  T *r = vec<T>(0, n), *rhat = vec<T>(1, n), *p = vec<T>(2, n), *v = vec<T>(3, n);
  T *ph = vec<T>(4, n), *sv = vec<T>(5, n), *sh = vec<T>(6, n), *t = vec<T>(7, n);
  Scal *s = c.s(0);
  SolveOut o;
  o.thr = thr;
  scal_reset(s, thr * thr);
  o.rn0 = o.rn = true_residual<T>(A, b, x, r, s);
  if (o.rn0 <= thr) {
    o.term = CFDLA_TERM_INITIAL_RESIDUAL;
    return o;
  }
  // rho = rhat . r is formed by the x/r update of the previous iteration
  // (OP_BI_RR_RHO), except right after a (re)start
  bool need_rho = true;
  auto restart = [&]() {
    copy_vec<T>(n, r, rhat, s, false);
    fill_vec<T>(n, T(0), p);
    fill_vec<T>(n, T(0), v);
    scal_bicg_restart(s);
    need_rho = true;
  };
  restart();
  if (chunk <= 0) chunk = auto_chunk(M, n);
  int64_t enq = 0;
  Poller poll(s);
  bool finished = false;
  while (!finished) {
    const Scal *h = nullptr;
    if (enq < max_it) {
      int64_t steps = std::min<int64_t>(chunk, max_it - enq);
      for (int64_t k = 0; k < steps; ++k) {
        if (need_rho) dot2<T>(n, rhat, r, nullptr, nullptr, s, true, OP_BI_RHO, ws);
        need_rho = false;
        bicg_p_update<T>(n, r, v, p, s);
        M->apply(p, ph, s, true);
        csr_apply<T>(A, EPI_Y, ph, nullptr, nullptr, 0.0, v, s, true, OP_BI_ALPHA, rhat, ws);
        bicg_s_update<T>(n, r, v, sv, s, ws);
        M->apply(sv, sh, s, true);
        csr_apply<T>(A, EPI_Y, sh, nullptr, nullptr, 0.0, t, s, true, OP_NONE, nullptr, ws);
        dot2<T>(n, t, sv, t, t, s, true, OP_BI_OMEGA, ws);
        bicg_x_r_update<T>(n, ph, sh, sv, t, x, r, s, ws, rhat);
      }
      enq += steps;
      poll.post();
      h = poll.lagged();
    } else {
      h = poll.latest();
    }
    if (!h) continue;
    if (h->done == 0 && enq < max_it) continue;
    sync_stream();
    Scal now = c.read(s);
    o.iters = now.iters;
    if (now.done == 2) {
      o.term = now.term;
      finished = true;
    } else if (now.done == 1 || now.done == 3) {
      if (now.done == 3) axpy_scal<T>(n, ph, x, s, 0);  // x += alpha * phat
      double tn = true_residual<T>(A, b, x, r, s);
      if (!confirm || tn <= thr) {
        o.term = CFDLA_TERM_CONVERGED;
        o.rn = tn;
        return o;
      }
      scal_set_done(s, 0);
      restart();
      poll = Poller(s);
    } else {
      finished = true;
    }
  }
  o.rn = true_residual<T>(A, b, x, r, s);
  return o;
}

// FP32-inner BiCGSTAB with reliable FP64 residual replacement (Sleijpen &
// van der Vorst 1996, reliable updated residuals in hybrid Bi-CG methods):
// the FP32 recurrence accumulates its solution increments in an FP32
// correction xc and keeps running; its residual is replaced by the FP64
// residual b - A (x + xc) when the lagged host poll shows it has dropped by
// delta since the last replacement (OP_RELIABLE: done = 4 when that residual
// is below the threshold), and when the recurrence reports convergence, to
// confirm it.  A confirmation that fails restarts BiCGSTAB from the FP64
// residual (rhat = r); a replacement far above the recursive residual
// restarts the direction (p = r).
SolveOut bicgstab_reliable(const DevCsr &A, Precond *M, const double *b, double *x, double thr, int64_t max_it,
                           double delta, int chunk) {
  CFDLA_REQUIRE(A.v32.p != nullptr, CFDLA_ERR_CONFIG, "matrix has no FP32 copy");
  Context &c = ctx();
  RedWs &ws = c.ws;
  const int n = A.n;
  double *r64 = c.dv_at(8, n);
  float *r = c.fv_at(0, n), *rhat = c.fv_at(1, n), *p = c.fv_at(2, n), *v = c.fv_at(3, n);
  float *ph = c.fv_at(4, n), *sv = c.fv_at(5, n), *sh = c.fv_at(6, n), *t = c.fv_at(7, n), *xc = c.fv_at(8, n);
  Scal *s = c.s(0);
  SolveOut o;
  o.thr = thr;
  scal_reset(s, thr * thr);
  o.rn0 = o.rn = true_residual<double>(A, b, x, r64, s);
  if (o.rn0 <= thr) {
    o.term = CFDLA_TERM_INITIAL_RESIDUAL;
    return o;
  }
  f64_to_f32(n, r64, r, s, false);
  fill_vec<float>(n, 0.0f, xc);
  auto restart = [&]() {
    copy_vec<float>(n, r, rhat, s, false);
    fill_vec<float>(n, 0.0f, p);
    fill_vec<float>(n, 0.0f, v);
    scal_bicg_restart(s);
  };
  auto replace = [&](bool check_done) {
    fold_correction(n, xc, x, s, check_done);
    csr_apply<double>(A, EPI_RESID, x, b, nullptr, 0.0, r64, s, check_done, OP_RELIABLE, r64, ws);
    f64_to_f32(n, r64, r, s, check_done);
  };
  restart();
  if (chunk <= 0) chunk = auto_chunk(M, n);
  int64_t enq = 0;
  int32_t replacements = 0;
  bool want = false;
  Poller poll(s);
  for (;;) {
    const Scal *h = nullptr;
    if (enq < max_it) {
      const int64_t steps = std::min<int64_t>(chunk, max_it - enq);
      for (int64_t k = 0; k < steps; ++k) {
        if (want) {
          ++replacements;
          replace(true);
          want = false;
        }
        dot2<float>(n, rhat, r, nullptr, nullptr, s, true, OP_BI_RHO, ws);
        bicg_p_update<float>(n, r, v, p, s);
        M->apply(p, ph, s, true);
        csr_apply<float>(A, EPI_Y, ph, nullptr, nullptr, 0.0, v, s, true, OP_BI_ALPHA, rhat, ws);
        bicg_s_update<float>(n, r, v, sv, s, ws);
        M->apply(sv, sh, s, true);
        csr_apply<float>(A, EPI_Y, sh, nullptr, nullptr, 0.0, t, s, true, OP_NONE, nullptr, ws);
        dot2<float>(n, t, sv, t, t, s, true, OP_BI_OMEGA, ws);
        bicg_x_r_update<float>(n, ph, sh, sv, t, xc, r, s, ws);
      }
      enq += steps;
      poll.post();
      h = poll.lagged();
    } else {
      h = poll.latest();
    }
    if (!h) continue;
    if (h->done == 0 && enq < max_it) {
      const double ref = h->nupd > 0 ? h->rr_true : o.rn0 * o.rn0;
      const double de = std::min(0.5, delta * std::pow(10.0, (double)h->nrestart));
      if (delta > 0.0 && h->nupd >= replacements && h->rr > 0.0 && h->rr <= de * de * ref) want = true;
      continue;
    }
    sync_stream();
    Scal now = c.read(s);
    o.iters = now.iters;
    if ((now.done == 1 || now.done == 3) && enq < max_it) {
      // the recurrence reports convergence: confirm in FP64, restart if not
      if (now.done == 3) axpy_scal<float>(n, ph, xc, s, 0);  // xc += alpha phat
      ++replacements;
      replace(false);
      now = c.read(s);
      if (now.done == 4) {
        o.term = CFDLA_TERM_CONVERGED;
        o.outer = now.nupd;
        o.rn = std::sqrt(now.rr_true);
        return o;
      }
      if (now.done == 2) {
        o.term = now.term;
        o.outer = now.nupd;
        o.rn = std::sqrt(now.rr_true);
        return o;
      }
      restart();
      poll = Poller(s);
      continue;
    }
    if (now.done == 4) {
      o.term = CFDLA_TERM_CONVERGED;
      o.outer = now.nupd;
      o.rn = std::sqrt(now.rr_true);
      return o;
    }
    // stopped (breakdown) or out of budget: fold and report the FP64 residual
    if (now.done == 3) axpy_scal<float>(n, ph, xc, s, 0);
    fold_correction(n, xc, x, s, false);
    o.rn = true_residual<double>(A, b, x, r64, s);
    o.outer = now.nupd + 1;
    if (now.done == 2) o.term = now.term;
    else o.term = o.rn <= thr ? CFDLA_TERM_CONVERGED : CFDLA_TERM_MAX_ITERATIONS;
    return o;
  }
}

SolveOut run_fp64(const DevCsr &A, Precond *M, const double *b, double *x, double thr, const cfdla_solver_config &cfg) {
  switch (cfg.kind) {
    case CFDLA_CG: return pcg<double>(A, M, b, x, thr, cfg.max_iter, cfg.check_every, true);
    case CFDLA_BICGSTAB: return bicgstab<double>(A, M, b, x, thr, cfg.max_iter, cfg.check_every, true);
    default: return gmres_fp64(A, M, b, x, thr, cfg);
  }
}

// FP32-inner / FP64-refined defect correction.
SolveOut run_refined(const DevCsr &A, Precond *M, const double *b, double *x, double thr,
                     const cfdla_solver_config &cfg) {
  CFDLA_REQUIRE(A.v32.p != nullptr, CFDLA_ERR_CONFIG, "matrix has no FP32 copy");
  const double delta = cfg.reliable_delta > 0.0 ? cfg.reliable_delta : 1e-1;
  if (cfg.kind == CFDLA_CG && cfg.reliable_interval >= 0)
    return pcg_reliable(A, M, b, x, thr, cfg.max_iter, cfg.reliable_interval,
                        cfg.reliable_interval == 0 ? delta : 0.0, cfg.check_every);
  if (cfg.kind == CFDLA_BICGSTAB && cfg.reliable_interval >= 0)
    return bicgstab_reliable(A, M, b, x, thr, cfg.max_iter, delta, cfg.check_every);
  Context &c = ctx();
  const int n = A.n;
  double *r64 = c.dv_at(8, n);
  float *r32 = c.fv_at(8, n), *d32 = c.fv_at(9, n);
  Scal *so = c.s(1);
  SolveOut o;
  o.thr = thr;
  scal_reset(so, thr * thr);
  double rn = true_residual<double>(A, b, x, r64, so);
  o.rn0 = o.rn = rn;
  if (rn <= thr) {
    o.term = CFDLA_TERM_INITIAL_RESIDUAL;
    return o;
  }
  const double inner_tol = cfg.inner_tol > 0.0 ? cfg.inner_tol : 1e-5;
  const int max_outer = cfg.max_outer > 0 ? cfg.max_outer : 30;
  int stalls = 0;
  for (int outer = 1; outer <= max_outer; ++outer) {
    int64_t budget = cfg.max_iter - o.iters;
    if (budget <= 0) break;
    f64_to_f32_scaled(n, r64, 1.0 / rn, r32);
    fill_vec<float>(n, 0.0f, d32);
    SolveOut in;
    if (cfg.kind == CFDLA_BICGSTAB) in = bicgstab<float>(A, M, r32, d32, inner_tol, budget, cfg.check_every, false);
    else if (cfg.kind == CFDLA_CG) in = pcg<float>(A, M, r32, d32, inner_tol, budget, cfg.check_every, false);
    else in = gmres_fp32_inner(A, M, r32, d32, inner_tol, budget, cfg);
    o.iters += in.iters;
    axpy_f32_to_f64(n, rn, d32, x);
    double rn_new = true_residual<double>(A, b, x, r64, so);
    o.outer = outer;
    o.rn = rn_new;
    if (rn_new <= thr) {
      o.term = CFDLA_TERM_CONVERGED;
      return o;
    }
    if (!(rn_new < 0.9 * rn)) {
      if (++stalls >= 2 || in.iters == 0) {
        o.term = (in.term == CFDLA_TERM_BREAKDOWN || in.term == CFDLA_TERM_NONPOSITIVE_CURVATURE)
                     ? in.term
                     : CFDLA_TERM_MAX_ITERATIONS;
        return o;
      }
    } else {
      stalls = 0;
    }
    rn = rn_new;
  }
  o.term = CFDLA_TERM_MAX_ITERATIONS;
  return o;
}

}  // namespace

SolveOut solve_device(const DevCsr &A, Precond *M, const double *b, double *x, const cfdla_solver_config &cfg) {
  CFDLA_REQUIRE(A.n == A.m, CFDLA_ERR_DIMENSION, "solve requires a square matrix");
  CFDLA_REQUIRE(cfg.kind >= 0 && cfg.kind <= 2, CFDLA_ERR_CONFIG, "unknown solver kind");
  CFDLA_REQUIRE(cfg.precision == CFDLA_FP64 || cfg.precision == CFDLA_FP32_REFINED ||
                    cfg.precision == CFDLA_PRECISION_AUTO,
                CFDLA_ERR_CONFIG, "unknown precision");
  CFDLA_REQUIRE(cfg.max_iter >= 1, CFDLA_ERR_CONFIG, "max_iter must be >= 1");
  CFDLA_REQUIRE(cfg.reliable_interval >= -1, CFDLA_ERR_CONFIG, "reliable_interval must be >= -1");
  CFDLA_REQUIRE(std::isfinite(cfg.tol_abs) && std::isfinite(cfg.tol_rel) && cfg.tol_abs >= 0.0 && cfg.tol_rel >= 0.0,
                CFDLA_ERR_CONFIG, "tolerances must be finite and non-negative");
  if (A.n == 0) {  // nothing to solve: ||b|| = 0 meets any threshold
    SolveOut o;
    o.thr = cfg.tol_abs;
    o.term = CFDLA_TERM_INITIAL_RESIDUAL;
    return o;
  }
  std::unique_ptr<IdentityPrecond> ident;
  if (!M) {
    ident = std::make_unique<IdentityPrecond>(A.n);
    M = ident.get();
  }
  // ||b|| for the relative threshold
  Context &c = ctx();
  Scal *s = c.s(2);
  scal_reset(s, 0.0);
  dot2<double>(A.n, b, b, nullptr, nullptr, s, false, OP_STORE, c.ws);
  double bn = std::sqrt(c.read(s).red[0]);
  CFDLA_REQUIRE(std::isfinite(bn), CFDLA_ERR_NUMERICAL, "right-hand side is not finite");
  double thr = std::max(cfg.tol_abs, cfg.tol_rel * bn);
  int precision = cfg.precision;
  if (precision == CFDLA_PRECISION_AUTO) {
    // FP32 halves the value bytes; it pays once an iteration streams more
    // than the 256 MiB Infinity Cache holds (matrix and preconditioner).
    const double bytes = 12.0 * (double)A.nnz + 4.0 * (double)A.n + M->stream_bytes();
    precision = bytes > 256.0 * 1024.0 * 1024.0 ? CFDLA_FP32_REFINED : CFDLA_FP64;
  }
  SolveOut o = precision == CFDLA_FP32_REFINED ? run_refined(A, M, b, x, thr, cfg) : run_fp64(A, M, b, x, thr, cfg);
  o.precision = precision;
  sync_stream();
  return o;
}

}  // namespace cfdla
