// Order of the long rows.
//
// The long-row kernels take rows from one list with a global counter, so
// rows that are close in the list run at the same time on every XCD.  In
// power-law products a few B rows feed most of the products.  On R-MAT
// scale 20 the 64 longest rows of B hold 4.7 MB and supply 39% of the long
// rows' products, and each is gathered by thousands of long rows.  In row
// order the rows that gather a given hub row are far apart in time.  Between
// two of its uses an XCD's 4 MB L2 then streams several times its size of
// other B rows, so even the hub rows miss.
//
// Sorting the long rows by which of B's HUBS longest rows they gather gives
// those reuses a short distance.  The sort key is a bitmask, the longest row
// in the top bit, sorted in descending order.  All rows that gather the
// longest B row come first, so 64 of them run on every XCD at once.  Within
// that group, the rows that also gather the second longest come first, and
// so on.  Only the order changes: a row's list slot still indexes its tile
// counts and bitmaps, and the row's result does not depend on when it runs.
#include <hipcub/hipcub.hpp>

#include <type_traits>

#include "device.hpp"
#include "spgemm.hpp"

namespace rt {
using namespace dev;

namespace {

// 16, 32 and 64 hubs measured the same on R-MAT scale 20 (long-row numeric
// kernel 65.2-65.8 ms): the first few hubs carry the reuse.
#ifndef RT_HUBS
#define RT_HUBS 32
#endif
constexpr int HUBS = RT_HUBS;
using HubKey = std::conditional_t<(HUBS > 32), u64, u32>;
static_assert(HUBS <= 64, "hub keys are at most 64 bits");

// Top HUBS of a list of keys, largest first: workgroup g sorts keys
// [1024 g, 1024 g + 1024) in LDS (bitonic, descending) and writes its first
// HUBS to out[HUBS g ...].  Rounds of this shrink any list 1024/HUBS-fold
// until one workgroup holds it.  A key is (row length << 32) | ~row, so
// ties go to the lower row and a zero key marks nothing.  With rp set, the
// keys are made from B's row pointers (n = rows of B); otherwise they are
// read from in.
constexpr int TOPK_WG = 1024;

__global__ __launch_bounds__(TOPK_WG) void k_topk(i64 n, const i64* __restrict__ rp,
                                                  const u64* __restrict__ in,
                                                  u64* __restrict__ out) {
  __shared__ u64 s_k[TOPK_WG];
  const int t = threadIdx.x;
  const i64 i = (i64)blockIdx.x * TOPK_WG + t;
  u64 k = 0;
  if (i < n) {
    if (rp) {
      const i64 len = min<i64>(rp[i + 1] - rp[i], 0xffffffffLL);
      k = len > 0 ? ((u64)len << 32) | (u32)~(u32)i : 0;
    } else {
      k = in[i];
    }
  }
  s_k[t] = k;
  __syncthreads();
  for (int size = 2; size <= TOPK_WG; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int j = t ^ stride;
      if (j > t) {
        const u64 a = s_k[t], b = s_k[j];
        // Descending overall: blocks with (t & size) == 0 sort descending.
        const bool desc = (t & size) == 0;
        if (desc ? a < b : a > b) {
          s_k[t] = b;
          s_k[j] = a;
        }
      }
      __syncthreads();
    }
  }
  if (t < HUBS) out[(i64)blockIdx.x * HUBS + t] = s_k[t];
}

// rank[k] = h for hub h (row k of B), 0xff for every other row.
__global__ __launch_bounds__(256) void k_hub_ranks(i64 nb, const u64* __restrict__ top,
                                                   unsigned char* __restrict__ rank) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= nb) return;
  unsigned char r = 0xff;
#pragma unroll
  for (int h = 0; h < HUBS; ++h)
    if (top[h] && (i32)~(u32)top[h] == i) r = (unsigned char)h;
  rank[i] = r;
}

// One workgroup per long row: bit (HUBS - 1 - h) of the row's key is set
// when its A entries name hub h.
__global__ __launch_bounds__(256) void k_hub_keys(const i32* __restrict__ rows,
                                                  const i64* __restrict__ arp,
                                                  const i32* __restrict__ acol,
                                                  const unsigned char* __restrict__ rank,
                                                  HubKey* __restrict__ keys) {
  __shared__ HubKey s_or[256 / WAVE];
  const i32 row = rows[blockIdx.x];
  const i64 e = arp[row + 1];
  HubKey key = 0;
  for (i64 j = arp[row] + threadIdx.x; j < e; j += 256) {
    const u32 h = rank[acol[j]];
    if (h < HUBS) key |= (HubKey)1 << (HUBS - 1 - h);
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    key |= (u32)__shfl_xor((int)(u32)key, o, WAVE);
    if constexpr (sizeof(HubKey) == 8)
      key |= (HubKey)(u32)__shfl_xor((int)(u32)(key >> 32), o, WAVE) << 32;
  }
  if ((threadIdx.x & (WAVE - 1)) == 0) s_or[threadIdx.x / WAVE] = key;
  __syncthreads();
  if (threadIdx.x == 0) keys[blockIdx.x] = s_or[0] | s_or[1] | s_or[2] | s_or[3];
}

}  // namespace

void order_long_rows(const Csr& A, const Csr& B, i32* rows, i32 n, hipStream_t s) {
  if (n < 2 || B.rows < 1) return;
  ProfScope ps("order_long_rows", s);
  // The HUBS longest rows of B, longest first.
  i64 len = B.rows;
  i64 blocks = (len + TOPK_WG - 1) / TOPK_WG;
  DBuf<u64> cand(2 * (size_t)blocks * HUBS);
  u64* src = nullptr;
  u64* dst = cand.p;
  u64* other = cand.p + (size_t)blocks * HUBS;
  hipLaunchKernelGGL(k_topk, dim3((unsigned)blocks), dim3(TOPK_WG), 0, s, len, B.rp, nullptr,
                     dst);
  while (blocks > 1) {
    len = blocks * HUBS;
    blocks = (len + TOPK_WG - 1) / TOPK_WG;
    src = dst;
    dst = other;
    other = src;
    hipLaunchKernelGGL(k_topk, dim3((unsigned)blocks), dim3(TOPK_WG), 0, s, len, nullptr, src,
                       dst);
  }
  DBuf<unsigned char> rank((size_t)B.rows);
  hipLaunchKernelGGL(k_hub_ranks, dim3((unsigned)((B.rows + 255) / 256)), dim3(256), 0, s, B.rows,
                     dst, rank.p);
  RT_HIP(hipGetLastError());
  // Keys of the long rows, then the rows sorted by key, descending (stable).
  DBuf<HubKey> keys(2 * (size_t)n);
  DBuf<i32> vals(2 * (size_t)n);
  hipcub::DoubleBuffer<HubKey> kk(keys.p, keys.p + n);
  hipcub::DoubleBuffer<i32> kv(vals.p, vals.p + n);
  size_t tb = 0;
  RT_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb, kk, kv, n, 0, HUBS, s));
  DBuf<unsigned char> tmp(tb);
  RT_HIP(hipMemcpyAsync(kv.Current(), rows, (size_t)n * sizeof(i32), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_hub_keys, dim3((unsigned)n), dim3(256), 0, s, kv.Current(), A.rp, A.col,
                     rank.p, kk.Current());
  RT_HIP(hipGetLastError());
  RT_HIP(hipcub::DeviceRadixSort::SortPairsDescending(tmp.p, tb, kk, kv, n, 0, HUBS, s));
  RT_HIP(hipMemcpyAsync(rows, kv.Current(), (size_t)n * sizeof(i32), hipMemcpyDeviceToDevice, s));
  RT_HIP(hipStreamSynchronize(s));  // the scratch buffers are released at scope exit
}

}  // namespace rt
