// Sparse product C = A B on the device for the AMG setup's Galerkin products
// (A_c = (R A) P, amg.rs:349-353), bit-identical to the host product
// csr_multiply (amg_setup.cpp).
//
// Expand-sort-compress: every product a_ik b_kj becomes a key (row i,
// column j, traversal index q) with the entry pair (k_a, k_b) as payload,
// where q counts the products of row i in the host's traversal order (the
// entries of A's row, and for each the entries of B's row).  One global
// radix sort orders the keys by row, then column, then q.  A run of equal
// (row, column) is one entry of C; its products, in q order, are the order
// in which the host's row accumulator met them.  The sorted payload is
// therefore also the per-slot plan that the value refresh replays
// (refresh.hip, k_product_numeric), and C's values are formed by that same
// kernel.
//
// The sort is hipCUB's device radix sort: this is setup work, not the
// solve's hot path.
#include <hipcub/hipcub.hpp>

#include "internal.hpp"
#include "solver.hpp"

namespace cfdla {
namespace {

int bits_for(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b) != 0) ++b;
  return std::max(b, 1);
}

int grid_of(int64_t units) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kGridCap, (units + kBlock - 1) / kBlock));
}

// Products per row of A B, and the largest.
__global__ __launch_bounds__(kBlock) void k_count(int nr, const int32_t *__restrict__ arp,
                                                  const int32_t *__restrict__ aci, const int32_t *__restrict__ brp,
                                                  int64_t *cnt) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nr; i += gridDim.x * kBlock) {
    int64_t c = 0;
    for (int k = arp[i]; k < arp[i + 1]; ++k) c += brp[aci[k] + 1] - brp[aci[k]];
    cnt[i] = c;
  }
}

// Keys (row << (cb + sb)) | (col << sb) | q and payload (k_a << 32) | k_b,
// one lane per row of A, in the host's traversal order.
__global__ __launch_bounds__(kBlock) void k_expand(int nr, const int32_t *__restrict__ arp,
                                                   const int32_t *__restrict__ aci, const int32_t *__restrict__ brp,
                                                   const int32_t *__restrict__ bci, const int64_t *__restrict__ offs,
                                                   int cb, int sb, uint64_t *keys, uint64_t *pay) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nr; i += gridDim.x * kBlock) {
    int64_t o = offs[i];
    uint64_t q = 0;
    const uint64_t rowkey = (uint64_t)i << (cb + sb);
    for (int ka = arp[i]; ka < arp[i + 1]; ++ka) {
      const int r = aci[ka];
      for (int kb = brp[r]; kb < brp[r + 1]; ++kb) {
        keys[o] = rowkey | ((uint64_t)bci[kb] << sb) | q;
        pay[o] = ((uint64_t)(uint32_t)ka << 32) | (uint32_t)kb;
        ++o;
        ++q;
      }
    }
  }
}

// 1 where a sorted key starts a new (row, column) run.
__global__ __launch_bounds__(kBlock) void k_run_flags(int64_t total, const uint64_t *__restrict__ keys, int sb,
                                                      int32_t *flag) {
  for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < total; q += (int64_t)gridDim.x * kBlock)
    flag[q] = (q == 0 || (keys[q] >> sb) != (keys[q - 1] >> sb)) ? 1 : 0;
}

// Per run start: the slot's first product, its column, and one count for
// its row; every product: the unpacked plan pair.
__global__ __launch_bounds__(kBlock) void k_compress(int64_t total, const uint64_t *__restrict__ keys,
                                                     const uint64_t *__restrict__ pay,
                                                     const int32_t *__restrict__ flag,
                                                     const int32_t *__restrict__ sid, int cb, int sb,
                                                     int32_t *slot_ptr, int32_t *cci, int32_t *rowcnt, int32_t *pa,
                                                     int32_t *pb) {
  const uint64_t cmask = (cb >= 64) ? ~0ull : ((1ull << cb) - 1);
  for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < total; q += (int64_t)gridDim.x * kBlock) {
    pa[q] = (int32_t)(pay[q] >> 32);
    pb[q] = (int32_t)(pay[q] & 0xffffffffull);
    if (flag[q]) {
      const int32_t s = sid[q] - 1;  // inclusive scan of the flags
      slot_ptr[s] = (int32_t)q;
      cci[s] = (int32_t)((keys[q] >> sb) & cmask);
      atomicAdd(&rowcnt[keys[q] >> (cb + sb)], 1);  // integer counts: order-independent
    }
  }
}

__global__ void k_set_last(int32_t *slot_ptr, int nslots, int32_t total) { slot_ptr[nslots] = total; }

template <typename T>
void exclusive_scan(const T *in, T *out, int64_t n) {
  size_t tb = 0;
  CFDLA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, (int)n, stream()));
  DevBuf<char> tmp(std::max<size_t>(tb, 1));
  CFDLA_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, in, out, (int)n, stream()));
  sync_stream();
}

}  // namespace

void device_product(int nr, int ncols, const int32_t *arp, const int32_t *aci, const double *av, const int32_t *brp,
                    const int32_t *bci, const double *bv, DevPlan &plan, DevBuf<int32_t> &crp,
                    DevBuf<int32_t> &cci, DevBuf<double> &cv) {
  crp.ensure((size_t)nr + 1);
  if (nr == 0) {
    plan.nslots = 0;
    return;
  }
  // products per row -> offsets (one extra slot holds the total)
  DevBuf<int64_t> cnt((size_t)nr + 1), offs((size_t)nr + 1);
  CFDLA_HIP(hipMemsetAsync(cnt.p + nr, 0, sizeof(int64_t), stream()));
  hipLaunchKernelGGL(k_count, dim3(grid_of(nr)), dim3(kBlock), 0, stream(), nr, arp, aci, brp, cnt.p);
  exclusive_scan<int64_t>(cnt.p, offs.p, (int64_t)nr + 1);
  std::vector<int64_t> hc((size_t)nr);
  cnt.download(hc.data(), hc.size());
  int64_t total = 0;
  CFDLA_HIP(hipMemcpyAsync(&total, offs.p + nr, sizeof(int64_t), hipMemcpyDeviceToHost, stream()));
  sync_stream();
  int64_t maxrow = 0;
  for (int64_t c : hc) maxrow = std::max(maxrow, c);
  CFDLA_REQUIRE(total < (int64_t)INT32_MAX, CFDLA_ERR_INVALID_ARG, "device product: more than 2^31 products");
  const int rb = bits_for((uint64_t)nr), cb = bits_for((uint64_t)ncols), sb = bits_for((uint64_t)maxrow);
  CFDLA_REQUIRE(rb + cb + sb <= 64, CFDLA_ERR_INVALID_ARG, "device product: key range exceeds 64 bits");
  plan.pa.ensure((size_t)std::max<int64_t>(total, 1));
  plan.pb.ensure((size_t)std::max<int64_t>(total, 1));
  if (total == 0) {
    plan.nslots = 0;
    CFDLA_HIP(hipMemsetAsync(crp.p, 0, ((size_t)nr + 1) * sizeof(int32_t), stream()));
    plan.ptr.ensure(1);
    CFDLA_HIP(hipMemsetAsync(plan.ptr.p, 0, sizeof(int32_t), stream()));
    sync_stream();
    return;
  }
  // expand, sort by (row, column, traversal index)
  DevBuf<uint64_t> keys((size_t)total), pay((size_t)total), keys2((size_t)total), pay2((size_t)total);
  hipLaunchKernelGGL(k_expand, dim3(grid_of(nr)), dim3(kBlock), 0, stream(), nr, arp, aci, brp, bci, offs.p, cb, sb,
                     keys.p, pay.p);
  {
    size_t tb = 0;
    CFDLA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.p, keys2.p, pay.p, pay2.p, (int)total, 0,
                                                 rb + cb + sb, stream()));
    DevBuf<char> tmp(std::max<size_t>(tb, 1));
    CFDLA_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keys.p, keys2.p, pay.p, pay2.p, (int)total, 0,
                                                 rb + cb + sb, stream()));
    sync_stream();
  }
  keys.release();
  pay.release();
  // runs of equal (row, column): slot ids by an inclusive scan of the flags
  DevBuf<int32_t> flag((size_t)total), sid((size_t)total);
  hipLaunchKernelGGL(k_run_flags, dim3(grid_of(total)), dim3(kBlock), 0, stream(), total, keys2.p, sb, flag.p);
  {
    size_t tb = 0;
    CFDLA_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, flag.p, sid.p, (int)total, stream()));
    DevBuf<char> tmp(std::max<size_t>(tb, 1));
    CFDLA_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tb, flag.p, sid.p, (int)total, stream()));
    sync_stream();
  }
  int32_t nslots = 0;
  CFDLA_HIP(hipMemcpyAsync(&nslots, sid.p + (total - 1), sizeof(int32_t), hipMemcpyDeviceToHost, stream()));
  sync_stream();
  plan.nslots = nslots;
  plan.ptr.ensure((size_t)nslots + 1);
  cci.ensure((size_t)std::max(nslots, 1));
  cv.ensure((size_t)std::max(nslots, 1));
  DevBuf<int32_t> rowcnt((size_t)nr + 1);
  CFDLA_HIP(hipMemsetAsync(rowcnt.p, 0, ((size_t)nr + 1) * sizeof(int32_t), stream()));
  hipLaunchKernelGGL(k_compress, dim3(grid_of(total)), dim3(kBlock), 0, stream(), total, keys2.p, pay2.p, flag.p,
                     sid.p, cb, sb, plan.ptr.p, cci.p, rowcnt.p, plan.pa.p, plan.pb.p);
  hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, stream(), plan.ptr.p, nslots, (int32_t)total);
  exclusive_scan<int32_t>(rowcnt.p, crp.p, (int64_t)nr + 1);
  product_numeric(nslots, plan.ptr.p, plan.pa.p, plan.pb.p, av, bv, cv.p);
  sync_stream();
}

}  // namespace cfdla
