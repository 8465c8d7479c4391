// Exact two-phase SpGEMM for gfx950: row analysis, row binning, LDS-hash and
// LDS-bitmap symbolic kernels, LDS-hash and dense-LDS-column-tile numeric
// kernels, exclusive scan, and the host pipeline spgemm_plain().
//
// Reference mapping.  The reference computes C = A*B by tracing one ray per
// nonzero A(i,k) (optixSpMSpM.cu:113-201) whose any-hit program adds
// A(i,k)*B(k,j) into a dense C for every sphere B(k,j) the ray meets
// (optixSpMSpM.cu:233-260).  Its GPU comparator is cuSPARSE's two-phase
// SpGEMM (cuSparse/src/main.cpp:464-539).  Here the "hits" of a ray are the
// entries of B's row k, read through B's row pointer, and the any-hit
// accumulate becomes an insert into an LDS table owned by output row i.
#include <algorithm>
#include <climits>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "learned.hpp"
#include "spgemm.hpp"

namespace rt {
using namespace dev;

constexpr int NBIN = 6;
constexpr int TILE_W = 16384;                   // columns per dense numeric tile
// Threads of the long-row numeric kernels: one presence word per thread of a
// tile, and small enough LDS (< 80 KB) for two workgroups per CU.
constexpr int HWG = TILE_W / 32;

struct Thr {
  i64 t[5];
};

// A list of rows for a grid-stride kernel: either (list, n) known on the
// host, or bin `bin` of a device-side binning (count dcount[bin], rows from
// list[doff[bin]]), so that launches need no host round trip.
struct RowSet {
  const i32* list;
  i32 n;
  const i32* dcount;
  const i32* doff;
  int bin;
  __device__ __forceinline__ i32 count() const { return dcount ? dcount[bin] : n; }
  __device__ __forceinline__ const i32* rows() const { return dcount ? list + doff[bin] : list; }
};

// Fills (kernels rather than memsets, so that graph capture can record them).
__global__ __launch_bounds__(256) void k_fill_i64(i64* __restrict__ p, i64 n, i64 v) {
  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256) p[i] = v;
}
__global__ __launch_bounds__(256) void k_fill_i32(i32* __restrict__ p, i64 n, i32 v) {
  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256) p[i] = v;
}

// B's entries as (column, value bits) pairs, so that a numeric kernel
// gathers a product's operands with one 8-byte load instead of two 4-byte
// loads (half the vector-memory instructions and address work per product).
__global__ __launch_bounds__(256) void k_pair_entries(i64 n, const i32* __restrict__ col,
                                                      const float* __restrict__ val,
                                                      uint2* __restrict__ cv) {
  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256)
    cv[i] = make_uint2((u32)col[i], __float_as_uint(val[i]));
}

// Where a numeric kernel gathers B's entries from: separate column and value
// arrays, or (column, value) pairs.
// PER is the count of consecutive entries a lane gathers per block load.
struct SplitEntries {
  static constexpr int PER = 1;
  const i32* __restrict__ col;
  const float* __restrict__ val;
  template <class OffT>
  __device__ __forceinline__ void load(OffT i, i32& c, float& v) const {
    c = col[i];
    v = val[i];
  }
  // Entry base + j, with base uniform (a scalar base address) and j the
  // lane's offset in units of LANE_UNIT entries (j = lane * LANE_UNIT).
  static constexpr int LANE_UNIT = 1;
  template <class OffT>
  __device__ __forceinline__ void load_at(OffT base, int j, i32& c, float& v) const {
    c = (col + base)[j];
    v = (val + base)[j];
  }
};
struct PairedEntries {
  static constexpr int PER = 1;
  const uint2* __restrict__ cv;
  template <class OffT>
  __device__ __forceinline__ void load(OffT i, i32& c, float& v) const {
    const uint2 x = cv[i];
    c = (i32)x.x;
    v = __uint_as_float(x.y);
  }
  // The lane's offset comes in bytes (LANE_UNIT 8), as for ColumnEntries.
  static constexpr int LANE_UNIT = 8;
  template <class OffT>
  __device__ __forceinline__ void load_at(OffT base, int jb, i32& c, float& v) const {
    const uint2 x = *reinterpret_cast<const uint2*>(reinterpret_cast<const char*>(cv + base) + jb);
    c = (i32)x.x;
    v = __uint_as_float(x.y);
  }
};

// Column indices only (symbolic kernels): the value is left unset.
struct ColumnEntries {
  static constexpr int PER = 1;
  const i32* __restrict__ col;
  template <class OffT>
  __device__ __forceinline__ void load(OffT i, i32& c, float&) const {
    c = col[i];
  }
  // The lane's offset comes in bytes (LANE_UNIT 4): selecting it between 0
  // and a precomputed lane * 4 saves a shift per gather.
  static constexpr int LANE_UNIT = 4;
  template <class OffT>
  __device__ __forceinline__ void load_at(OffT base, int jb, i32& c, float&) const {
    c = *reinterpret_cast<const i32*>(reinterpret_cast<const char*>(col + base) + jb);
  }
};

// Column indices, two consecutive entries per lane in the block walk: one
// 8-byte load at an even entry index.  The array must hold one readable
// entry past the last one a run can end on (a pair whose second half lies
// past a run's end is loaded whole and that half dropped).
struct ColumnPairs {
  static constexpr int PER = 2;
  const i32* __restrict__ col;  // 8-byte aligned
  template <class OffT>
  __device__ __forceinline__ void load(OffT i, i32& c, float&) const {
    c = col[i];
  }
  // Entries base, base + 1 (base even and uniform) + the lane's byte offset
  // jb (lane * 8).
  static constexpr int LANE_UNIT = 8;
  template <class OffT>
  __device__ __forceinline__ void load2_at(OffT base, int jb, i32& c0, i32& c1, float&,
                                           float&) const {
    const uint2 x =
        *reinterpret_cast<const uint2*>(reinterpret_cast<const char*>(col + base) + jb);
    c0 = (i32)x.x;
    c1 = (i32)x.y;
  }
};

// B's column indices copied with one spare entry past the end (the
// ColumnPairs source reads a pair whole).  With maxp set (captured stages,
// where the long-row count stays on the device), nothing is copied when no
// row has more than th products: no row can be a long row.
__global__ __launch_bounds__(256) void k_copy_cols(i64 n, const i32* __restrict__ col,
                                                   i32* __restrict__ out,
                                                   const unsigned long long* __restrict__ maxp,
                                                   i64 th) {
  if (maxp && (i64)*maxp <= th) return;
  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i <= n; i += (i64)gridDim.x * 256)
    out[i] = i < n ? col[i] : 0;
}

// Zeroes the row-slot counter of a one-row-per-workgroup kernel.
void zero_slot(i32* next, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, next, (i64)1, 0);
}

// ---------------------------------------------------------------------------
// Row analysis: P_i = sum over A(i,k) of nnz(B(k,:)).  16 lanes per row,
// rows grid-stride; rows of more than RP_LONG entries are queued for a
// workgroup each.  totals[0] += sum P_i, totals[1] = max P_i, with one
// atomic of each kind per workgroup (same-address atomics serialise).
constexpr i64 RP_LONG = 1024;  // longer A rows go to the workgroup-per-row kernel

template <bool BATCH>
__global__ __launch_bounds__(256) void k_row_products(
    i64 nrows, const i64* __restrict__ arp, const i32* __restrict__ acol,
    const i64* __restrict__ brp, const i64* __restrict__ bbase, i64* __restrict__ prod,
    unsigned long long* __restrict__ totals, i32* __restrict__ longq) {
  constexpr int G = 16;
  __shared__ i64 s_scr[4];
  const int l = threadIdx.x & (G - 1);
  i64 tsum = 0, tmax = 0;
  for (i64 r0 = (i64)blockIdx.x * (256 / G); r0 < nrows; r0 += (i64)gridDim.x * (256 / G)) {
    const i64 row = r0 + threadIdx.x / G;
    i64 sum = 0;
    bool mine = row < nrows;
    if (mine) {
      const i64 a0 = arp[row], a1 = arp[row + 1];
      if (a1 - a0 > RP_LONG) {
        mine = false;  // queued for k_row_products_long
        if (l == 0) longq[1 + atomicAdd(&longq[0], 1)] = (i32)row;
      } else {
        const i64 bb = BATCH ? bbase[row] : 0;
        for (i64 e = a0 + l; e < a1; e += G) {
          const i64 k = (i64)acol[e] + bb;
          sum += brp[k + 1] - brp[k];
        }
      }
    }
    sum = seg_reduce_sum<G>(sum);
    if (mine && l == 0) prod[row] = sum;
    if (l == 0) {
      tsum += sum;
      tmax = max(tmax, sum);
    }
  }
  // The workgroup's (sum, max) go to its own pair of partials, reduced by
  // k_row_totals: thousands of workgroups' atomics on one cache line cost
  // more than the analysis of a small product.
  __shared__ i64 s_mx[256 / WAVE];
  const i64 mx = seg_reduce_max<WAVE>(tmax);
  if ((threadIdx.x & (WAVE - 1)) == 0) s_mx[threadIdx.x / WAVE] = mx;
  const i64 ws = block_reduce_sum<256>(tsum, s_scr);
  if (threadIdx.x == 0) {
    i64 m = 0;
#pragma unroll
    for (int w = 0; w < 256 / WAVE; ++w) m = max(m, s_mx[w]);
    totals[2 + 2 * blockIdx.x] = (unsigned long long)ws;
    totals[3 + 2 * blockIdx.x] = (unsigned long long)m;
  }
}

// totals[0] += the n workgroups' sums, totals[1] = max(totals[1], their
// maxima), from the partials at totals[2 ..].  One workgroup.
__global__ __launch_bounds__(256) void k_row_totals(i32 n, unsigned long long* __restrict__ totals) {
  __shared__ i64 s_scr[256 / WAVE];
  __shared__ i64 s_mx[256 / WAVE];
  i64 sum = 0, mx = 0;
  for (i32 i = threadIdx.x; i < n; i += 256) {
    sum += (i64)totals[2 + 2 * i];
    mx = max(mx, (i64)totals[3 + 2 * i]);
  }
  mx = seg_reduce_max<WAVE>(mx);
  if ((threadIdx.x & (WAVE - 1)) == 0) s_mx[threadIdx.x / WAVE] = mx;
  sum = block_reduce_sum<256>(sum, s_scr);
  if (threadIdx.x == 0) {
    i64 m = 0;
#pragma unroll
    for (int w = 0; w < 256 / WAVE; ++w) m = max(m, s_mx[w]);
    if (sum > 0) atomicAdd(&totals[0], (unsigned long long)sum);
    if (m > 0) atomicMax(&totals[1], (unsigned long long)m);
  }
}

// The rows of more than RP_LONG A entries (longq[0] of them, listed from
// longq[1]): one workgroup per row, four independent entries per thread in
// flight, so the longest rows no longer serialise on 16 lanes.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_row_products_long(
    const i32* __restrict__ longq, const i64* __restrict__ arp, const i32* __restrict__ acol,
    const i64* __restrict__ brp, const i64* __restrict__ bbase, i64* __restrict__ prod,
    unsigned long long* __restrict__ totals) {
  constexpr int U = 4;
  __shared__ i64 s_scr[4];
  const i32 n = longq[0];
  for (i32 i = blockIdx.x; i < n; i += gridDim.x) {
    const i32 row = longq[1 + i];
    const i64 a0 = arp[row], a1 = arp[row + 1];
    const i64 bb = BATCH ? bbase[row] : 0;
    i64 sum = 0;
    for (i64 e0 = a0 + threadIdx.x; e0 < a1; e0 += 256 * U) {
      i64 k[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const i64 e = e0 + u * 256;
        k[u] = e < a1 ? (i64)acol[e] + bb : -1;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (k[u] >= 0) sum += brp[k[u] + 1] - brp[k[u]];
    }
    sum = block_reduce_sum<256>(sum, s_scr);
    if (threadIdx.x == 0) {
      prod[row] = sum;
      atomicAdd(&totals[0], (unsigned long long)sum);
      atomicMax(&totals[1], (unsigned long long)sum);
    }
  }
}

// Grid of the row-analysis kernel: enough workgroups to fill the device.
inline unsigned row_products_grid(i64 nrows) {
  const i64 need = (nrows * 16 + 255) / 256;
  return (unsigned)std::max<i64>(1, std::min<i64>(need, (i64)device().cus * 16));
}

// Entries of the long-row queue of the row analysis (count + rows).
inline size_t row_products_queue(i64 annz) { return (size_t)(annz / RP_LONG + 2); }

// Entries of the row analysis' totals: (sum, max) and a pair of partials per
// workgroup of k_row_products.
inline size_t row_totals_len() { return 2 + 2 * (size_t)device().cus * 16; }

// Row analysis: prod[i] = P_i; totals (row_totals_len() entries, the first
// two zeroed by the caller) += (sum, max).  longq holds
// row_products_queue(A.nnz) entries.  Kernels only (graph capture records
// them).
template <bool BATCH>
void launch_row_products(i64 m, const i64* arp, const i32* acol, const i64* brp,
                         const i64* bbase, i64* prod, unsigned long long* totals, i32* longq,
                         hipStream_t s) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, longq, (i64)1, 0);
  const unsigned grid = row_products_grid(m);
  hipLaunchKernelGGL(k_row_products<BATCH>, dim3(grid), dim3(256), 0, s, m, arp, acol, brp, bbase,
                     prod, totals, longq);
  hipLaunchKernelGGL(k_row_totals, dim3(1), dim3(256), 0, s, (i32)grid, totals);
  hipLaunchKernelGGL(k_row_products_long<BATCH>, dim3((unsigned)device().cus * 2), dim3(256), 0,
                     s, longq, arp, acol, brp, bbase, prod, totals);
}

// ---------------------------------------------------------------------------
// Row binning.  key = min(x[i], cap); key 0 -> no bin (empty output row).
__device__ __forceinline__ int bin_of(i64 k, const Thr& th) {
  if (k <= 0) return -1;
#pragma unroll
  for (int b = 0; b < 5; ++b)
    if (k <= th.t[b]) return b;
  return 5;
}

// Row binning in three launches with no atomics: k_bin_count writes every
// workgroup's bin counts over its rows, k_bin_scan (one workgroup) turns
// them into each workgroup's first slot in every bin (plus the bins' counts
// and offsets), and k_bin_fill writes each workgroup's rows there, tile by
// tile in row order, so the lists are deterministic.  Both row kernels run
// a capped grid (stride_grid); workgroup w takes the contiguous run of whole
// 256-row tiles from bin_run(n) * w.  (One workgroup per 256 rows with a
// global atomic per bin and workgroup made the counters the bottleneck:
// 1.08 ms per kernel for 24M rows; one atomic per bin and workgroup of a
// capped grid still cost 0.14 ms per kernel, the atomics of all workgroups
// arriving together on one cache line.)
__device__ __forceinline__ void bin_run(i64 n, i64& r0, i64& r1) {
  const i64 per = (((n + gridDim.x - 1) / gridDim.x) + 255) & ~(i64)255;
  r0 = min(n, (i64)blockIdx.x * per);
  r1 = min(n, r0 + per);
}

__global__ __launch_bounds__(256) void k_bin_count(i64 n, const i64* __restrict__ x, i64 cap,
                                                   Thr th, i32* __restrict__ wgc) {
  __shared__ i32 s_w[256 / WAVE][NBIN];
  i64 r0, r1;
  bin_run(n, r0, r1);
  i32 c[NBIN];
#pragma unroll
  for (int b = 0; b < NBIN; ++b) c[b] = 0;
  for (i64 i = r0 + threadIdx.x; i < r1; i += 256) {
    const int b = bin_of(min(x[i], cap), th);
#pragma unroll
    for (int q = 0; q < NBIN; ++q) c[q] += b == q;
  }
  const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int q = 0; q < NBIN; ++q) {
    i32 v = c[q];
#pragma unroll
    for (int d = WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d);
    if (lane == 0) s_w[w][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < NBIN) {
    i32 t = 0;
#pragma unroll
    for (int v = 0; v < 256 / WAVE; ++v) t += s_w[v][threadIdx.x];
    wgc[(i64)blockIdx.x * NBIN + threadIdx.x] = t;
  }
}

// One workgroup: wgc[w][b] (nwg workgroups' counts) becomes workgroup w's
// first slot in the whole list for bin b; cnt[b] = bin b's rows, off[b] =
// its first slot (off[NBIN] = all rows), cursor (if set) = off.
__global__ __launch_bounds__(256) void k_bin_scan(i32 nwg, i32* __restrict__ wgc,
                                                  i32* __restrict__ cnt, i32* __restrict__ off,
                                                  i32* __restrict__ cursor) {
  __shared__ i32 s_scr[256 / WAVE];
  const int per = (nwg + 255) / 256;
  const int w0 = min(nwg, (int)threadIdx.x * per), w1 = min(nwg, w0 + per);
  i32 run = 0;  // rows of the bins before b
  for (int b = 0; b < NBIN; ++b) {
    i32 mine = 0;
    for (int w = w0; w < w1; ++w) mine += wgc[(i64)w * NBIN + b];
    i32 tot;
    const i32 inc = block_scan_incl<256>(mine, s_scr, &tot);
    i32 at = run + inc - mine;
    for (int w = w0; w < w1; ++w) {
      const i32 v = wgc[(i64)w * NBIN + b];
      wgc[(i64)w * NBIN + b] = at;
      at += v;
    }
    if (threadIdx.x == 0) {
      cnt[b] = tot;
      off[b] = run;
      if (cursor) cursor[b] = run;
    }
    run += tot;
    __syncthreads();  // s_scr is reused by the next bin's scan
  }
  if (threadIdx.x == 0) off[NBIN] = run;
}

// The workgroup's rows, tile by tile in row order, from its first slots in
// wgc (ranks within a wave from ballots).
__global__ __launch_bounds__(256) void k_bin_fill(i64 n, const i64* __restrict__ x, i64 cap,
                                                  Thr th, const i32* __restrict__ wgc,
                                                  i32* __restrict__ list) {
  __shared__ i32 s_w[256 / WAVE][NBIN];
  __shared__ i32 s_base[NBIN];
  if (threadIdx.x < NBIN) s_base[threadIdx.x] = wgc[(i64)blockIdx.x * NBIN + threadIdx.x];
  __syncthreads();
  i64 r0, r1;
  bin_run(n, r0, r1);
  const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const u64 below = (1ull << lane) - 1ull;
  for (i64 t0 = r0; t0 < r1; t0 += 256) {  // uniform
    const i64 i = t0 + threadIdx.x;
    const int b = i < r1 ? bin_of(min(x[i], cap), th) : -1;
    int rank = 0;
#pragma unroll
    for (int q = 0; q < NBIN; ++q) {
      const u64 m = __ballot(b == q);
      if (b == q) rank = (int)__popcll(m & below);
      if (lane == 0) s_w[w][q] = (i32)__popcll(m);
    }
    __syncthreads();
    if (b >= 0) {
      int pos = s_base[b] + rank;
      for (int v = 0; v < w; ++v) pos += s_w[v][b];
      list[pos] = (i32)i;
    }
    __syncthreads();  // every thread has read s_base and s_w
    if (threadIdx.x < NBIN) {
      int add = 0;
#pragma unroll
      for (int v = 0; v < 256 / WAVE; ++v) add += s_w[v][threadIdx.x];
      s_base[threadIdx.x] += add;
    }
    __syncthreads();
  }
}

// A grid-stride launch over n items: at most 8 workgroups of 256 per CU.
unsigned stride_grid(i64 n) {
  return (unsigned)std::max<i64>(1, std::min<i64>((n + 255) / 256, (i64)device().cus * 8));
}

i64 bin_scratch_ints() { return (i64)device().cus * 8 * NBIN + 2 * NBIN + 1; }

// ---------------------------------------------------------------------------
// Product p of the current chunk -> entry index, from the previous entry
// `lo`: one LDS read when the entry still covers p (long B rows), a binary
// search over the rest otherwise (short B rows).
__device__ __forceinline__ int advance_entry(const i32* __restrict__ ex, int lo, int nvalid,
                                             int p) {
  if (lo + 1 < nvalid && ex[lo + 1] <= p) {
    int a = lo + 1, b = nvalid - 1;
    while (a < b) {
      const int mid = (a + b + 1) >> 1;
      if (ex[mid] <= p) a = mid; else b = mid - 1;
    }
    lo = a;
  }
  return lo;
}

// Product enumeration for the long-row kernels.  A chunk of up to WG
// entries A(i,k) -- entry l held by thread l as (run length, first B index,
// A value) -- has products 0..tot-1.  One block scan gives every entry both
// its first product and its first 64-product block, and the chunk is walked
// in one of two ways, chosen per chunk:
//
// * Blocks, when whole 64-product blocks leave at most 2/5 of the lanes
//   idle (long runs: the usual case in long rows).  Each run is cut
//   into blocks of 64 consecutive products of one entry, and wave w takes
//   an equal contiguous range of blocks.  For a batch of 64 blocks, lane j
//   finds the owner of block j by binary search over the entries' first
//   blocks and computes the block's first B index, length and A value; the
//   wave then walks the batch ROWS blocks at a time, lane l taking product
//   l of each block, with the block's fields read from lane j (v_readlane):
//   no per-product entry lookup, and the B loads use a scalar base.  A
//   source with PER = 2 (ColumnPairs) has blocks of 128 products starting
//   at even B indices, and lane l loads products 2l and 2l + 1 of a block
//   with one 8-byte load (half the loads and block bookkeeping per
//   product); an entry starting at an odd index masks its first slot, and
//   such blocks are taken when they keep at least half the lanes busy.
// * Owner marks, otherwise (short runs).  Wave w takes an equal contiguous
//   slice of the products and walks it in sub-batches of 64*ROWS (lane l:
//   products 64r + l).  For a sub-batch the wave marks, in its own zeroed
//   u16 array, the start of every nonempty entry that begins inside it (e+1
//   at ex[e]); each product's entry is the running maximum of the marks (a
//   DPP max-scan over the wave), seeded with the entry that covers the
//   sub-batch's first product.
//
// Either way a lane issues its ROWS B loads before using any (loads under a
// branch would make the compiler wait for each in turn) and hands the ROWS
// products to f at once, f(c, p, ok) with arrays of ROWS: c the B columns,
// p the products A*B (0 when VAL is false), ok whether the lane holds a
// product in that row.  Where it does not, c is the column of another
// product of the chunk (so within the columns the chunk's runs cover), and
// f should still use it unconditionally: a load whose value is used only
// under `ok` is sunk into the branch by the compiler and then waited for
// alone, which costs the round a second memory latency.  No wave waits
// for another inside the chunk; four barriers per chunk (two in the scan).
// OffT is the B index type (int32 when nnz(B) < 2^31).  A chunk's products
// must number fewer than 2^30.
template <class OffT>
struct ChunkLds {
  OffT* off;   // [WG] B index of the entry's product p is off[e] + p (marks) or off[e] + 64 b + lane (blocks)
  float* av;   // [WG] A value of the entry
  i32* ex;     // [WG+1] first product (marks) or first block (blocks) of the entry
  u16* mark;   // [WG/64][64*ROWS] per-wave zeroed owner marks (16-byte aligned);
               // block mode keeps each entry's product end (64 x first block
               // + run length) in its first WG*4 bytes and zeroes them after
  u64* scr;    // [WG/64] scan scratch
  // With int32 offsets and values, off and av share one 8-byte word per
  // entry (A value bits << 32 | offset), read with one LDS instruction:
  // off then points at WG u64 words and av is unused.
  static constexpr bool packs(bool val) { return val && sizeof(OffT) == 4; }
};

// v read from lane j (j uniform); the result is uniform.
template <class T>
__device__ __forceinline__ T lane_value(T v, int j) {
  if constexpr (sizeof(T) == 8) {
    const u64 u = (u64)v;
    const u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)u, j);
    const u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(u >> 32), j);
    return (T)(((u64)hi << 32) | lo);
  } else {
    return (T)__builtin_amdgcn_readlane((int)v, j);
  }
}
__device__ __forceinline__ float lane_value(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// Returns the chunk's product count (uniform).
template <int WG, int ROWS, bool VAL, class OffT, class Src, class F>
__device__ __forceinline__ int chunk_products(int len, OffT start, float av, int nvalid,
                                              const Src& src, const ChunkLds<OffT>& s, F&& f) {
  static_assert(WG % WAVE == 0 && WG <= 65535, "marks are u16 entry indices + 1");
  static_assert(WAVE % ROWS == 0 && ROWS >= 2, "blocks are walked ROWS at a time; the entry "
                                               "ends fit the mark array");
  constexpr int NW = WG / WAVE;
  constexpr bool PACK = ChunkLds<OffT>::packs(VAL);
  // Entries per lane and block gather (a block holds BW products).  With
  // PER 2 a block starts at an even B index, so an entry that starts at an
  // odd one leads with one masked product (`lead`).
  constexpr int PER = Src::PER;
  constexpr int BW = WAVE * PER;
  static_assert(PER == 1 || PER == 2, "one or two entries per lane");
  static_assert(ROWS % PER == 0, "whole blocks per batch");
  const int l = threadIdx.x;
  const int lane = l & (WAVE - 1);
  const int w = __builtin_amdgcn_readfirstlane(l / WAVE);
  const int lead = PER == 2 && len > 0 ? (int)(start & 1) : 0;
  const int nb = (len + lead + BW - 1) / BW;
  u64 both;
  const u64 inc = block_scan_incl<WG>(((u64)(u32)len << 32) | (u32)nb, s.scr, &both);
  const int tot = (int)(both >> 32), tb = (int)(u32)both;
  if (tot == 0) {
    __syncthreads();  // every wave has read s.scr before the next scan writes it
    return 0;
  }
  // Blocks when their lanes are at least 3/5 busy (1/2 with pairs, whose
  // block walk costs about half as many instructions per slot): a busy block
  // slot costs about 20 VALU instructions in the numeric kernels, a product
  // of the owner-mark walk about 36.
  const bool blocks = PER == 1 ? (i64)tot * 5 >= (i64)tb * (3 * BW) : (i64)tot * 2 >= (i64)tb * BW;
  if (blocks) {
    const int bex = (int)(u32)inc - nb;
    i32* s_end = reinterpret_cast<i32*>(s.mark);
    s.ex[l] = bex;
    s_end[l] = bex * BW + lead + len;
    // B index of the entry's block 0 (even with PER 2, its low bit then
    // carrying `lead`).
    const OffT offp = start - (OffT)bex * BW;
    if constexpr (PACK) {
      reinterpret_cast<u64*>(s.off)[l] = ((u64)__float_as_uint(av) << 32) | (u32)offp;
    } else {
      s.off[l] = offp;
      if constexpr (VAL) s.av[l] = av;
    }
    __syncthreads();
    const int per = (tb + NW - 1) / NW;
    const int pb = w * per, pe = min(tb, pb + per);
    const int lane_b = lane * Src::LANE_UNIT;  // the lane's offset in a block, as src takes it
    int lo = 0;  // owner of the batch's first block, or an entry before it
    for (int b0 = pb; b0 < pe; b0 += WAVE) {
      const int nblk = min(WAVE, pe - b0);
      const int b = b0 + min(lane, nblk - 1);
      int e = lo, hi = nvalid - 1;  // last entry whose first block is <= b
      while (e < hi) {
        const int mid = (e + hi + 1) >> 1;
        if (s.ex[mid] <= b) e = mid; else hi = mid - 1;
      }
      lo = __builtin_amdgcn_readlane(e, WAVE - 1);
      OffT bs;
      float ae = 0.f;
      if constexpr (PACK) {
        const u64 pk = reinterpret_cast<const u64*>(s.off)[e];
        bs = (OffT)(i32)(u32)pk;
        ae = __uint_as_float((u32)(pk >> 32));
      } else {
        bs = s.off[e];
        if constexpr (VAL) ae = s.av[e];
      }
      int first = 0;  // PER 2: the block is its entry's first and leads
      if constexpr (PER == 2) {
        first = (int)(bs & 1);
        bs -= (OffT)first;
        if (s.ex[e] != b) first = 0;
      }
      bs += (OffT)b * BW;                                    // the block's first B index
      const int rem = lane < nblk ? s_end[e] - b * BW : 0;  // its slots (> BW: all)
      const int remx = PER == 2 ? 2 * rem + first : rem;
      for (int j0 = 0; j0 < nblk; j0 += ROWS / PER) {
        i32 c[ROWS];
        float v[ROWS], p[ROWS];
        bool ok[ROWS];
        if constexpr (PER == 1) {
#pragma unroll
          for (int r = 0; r < ROWS; ++r) {
            ok[r] = lane < __builtin_amdgcn_readlane(rem, j0 + r);
            src.load_at(lane_value(bs, j0 + r), ok[r] ? lane_b : 0, c[r], v[r]);
          }
        } else {
          // Lane l takes slots 2l and 2l + 1 of the block: entries 2r and
          // 2r + 1 of the lane's arrays are neighbouring products.
#pragma unroll
          for (int r = 0; r < ROWS / 2; ++r) {
            const int x = __builtin_amdgcn_readlane(remx, j0 + r);
            ok[2 * r] = 2 * lane < (x >> 1) && (lane != 0 || (x & 1) == 0);
            ok[2 * r + 1] = 2 * lane + 1 < (x >> 1);
            src.load2_at(lane_value(bs, j0 + r), 2 * lane < (x >> 1) ? lane_b : 0, c[2 * r],
                         c[2 * r + 1], v[2 * r], v[2 * r + 1]);
          }
        }
        // All ROWS gathers leave before any result is used.  Without this
        // fence the scheduler moves the first product's use up between the
        // loads, and the wait for it holds back the other seven.
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < ROWS; ++r) p[r] = VAL ? lane_value(ae, j0 + r / PER) * v[r] : 0.f;
        f(c, p, ok);
      }
    }
  } else {
    constexpr int SB = ROWS * WAVE;  // products per sub-batch
    const int ex = (int)(inc >> 32) - len;
    s.ex[l] = ex;
    if (l == 0) s.ex[nvalid] = tot;
    if constexpr (PACK) {
      reinterpret_cast<u64*>(s.off)[l] = ((u64)__float_as_uint(av) << 32) | (u32)(start - (OffT)ex);
    } else {
      s.off[l] = start - (OffT)ex;
      if constexpr (VAL) s.av[l] = av;
    }
    __syncthreads();
    u16* wm = s.mark + w * SB;
    const int per = ((tot + NW - 1) / NW + WAVE - 1) & ~(WAVE - 1);
    const int pb = w * per, pe = min(tot, pb + per);
    if (pb < pe) {  // wave-uniform
      int lo = 0, hi = nvalid - 1;  // last entry with ex <= pb: the owner of pb
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s.ex[mid] <= pb) lo = mid; else hi = mid - 1;
      }
      int carry = lo + 1;  // owner of the sub-batch's first product, plus one
      for (int p0 = pb; p0 < pe; p0 += SB) {
        const int p1 = min(pe, p0 + SB);
        // Mark the nonempty entries that start in (p0, p1).
        for (int c0 = carry; ; c0 += WAVE) {
          const int c = c0 + lane;
          const int xc = c < nvalid ? s.ex[c] : INT_MAX;
          if (xc < p1 && s.ex[c + 1] > xc) wm[xc - p0] = (u16)(c + 1);
          if (!__any(xc < p1)) break;
        }
        wave_sync();
        OffT idx[ROWS];
        float a[ROWS];
        bool ok[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          const int q = r * WAVE + lane;
          int m = wm[q];
          m = max(wave_max_scan(m), carry);
          carry = __builtin_amdgcn_readlane(m, WAVE - 1);
          const int e = m - 1;
          ok[r] = p0 + q < p1;
          // Lanes past the sub-batch hold its last product's owner: point
          // them at that product.
          const int pq = min(p0 + q, p1 - 1);
          if constexpr (PACK) {
            const u64 pk = reinterpret_cast<const u64*>(s.off)[e];
            idx[r] = (OffT)(i32)(u32)pk + (OffT)pq;
            a[r] = __uint_as_float((u32)(pk >> 32));
          } else {
            idx[r] = s.off[e] + (OffT)pq;
            a[r] = 0.f;
            if constexpr (VAL) a[r] = s.av[e];
          }
        }
        // Clear the sub-batch's marks, ROWS per lane in one store (the fence
        // keeps the mark reads above before it).
        wave_sync();
        if constexpr (ROWS == 8) {
          reinterpret_cast<uint4*>(wm)[lane] = make_uint4(0u, 0u, 0u, 0u);
        } else if constexpr (ROWS == 4) {
          reinterpret_cast<uint2*>(wm)[lane] = make_uint2(0u, 0u);
        } else {
#pragma unroll
          for (int r = 0; r < ROWS; ++r) wm[r * WAVE + lane] = 0;
        }
        wave_sync();  // marks cleared before the next sub-batch writes them
        i32 c[ROWS];
        float v[ROWS], p[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) src.load(idx[r], c[r], v[r]);
        __builtin_amdgcn_sched_barrier(0);  // as above: every gather in flight first
#pragma unroll
        for (int r = 0; r < ROWS; ++r) p[r] = VAL ? a[r] * v[r] : 0.f;
        f(c, p, ok);
      }
    }
  }
  __syncthreads();  // the chunk's entries are read before the next chunk writes them
  if (blocks) reinterpret_cast<i32*>(s.mark)[l] = 0;  // marks start the next chunk zeroed
  return tot;
}

// Batched LDS FP32 accumulation of ROWS products per lane: the first
// compare-and-swap of every row is issued before any result is checked, so
// their LDS round trips overlap; a row whose swap lost a race retries alone.
// old[r] is the word's value read before (any earlier value is correct, only
// slower).
template <int ROWS>
__device__ __forceinline__ void lds_add_f32_rows(float* base, const int (&x)[ROWS],
                                                 const float (&p)[ROWS], const bool (&ok)[ROWS],
                                                 const u32 (&old)[ROWS]) {
  u32* q = reinterpret_cast<u32*>(base);
  u32 got[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    got[r] = old[r];
    if (ok[r]) got[r] = atomicCAS(&q[x[r]], old[r], __float_as_uint(__uint_as_float(old[r]) + p[r]));
  }
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    if (ok[r] && got[r] != old[r]) {
      u32 cur = got[r], assumed;
      do {
        assumed = cur;
        cur = atomicCAS(&q[x[r]], assumed, __float_as_uint(__uint_as_float(assumed) + p[r]));
      } while (cur != assumed);
    }
  }
}

// As lds_add_f32_rows, with the words given as byte offsets from base.
template <int ROWS>
__device__ __forceinline__ void lds_add_f32_rows_at(unsigned char* base, const u32 (&at)[ROWS],
                                                    const float (&p)[ROWS], const bool (&ok)[ROWS],
                                                    const u32 (&old)[ROWS]) {
  u32 got[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    got[r] = old[r];
    if (ok[r])
      got[r] = atomicCAS(reinterpret_cast<u32*>(base + at[r]), old[r],
                         __float_as_uint(__uint_as_float(old[r]) + p[r]));
  }
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    if (ok[r] && got[r] != old[r]) {
      u32 cur = got[r], assumed;
      do {
        assumed = cur;
        cur = atomicCAS(reinterpret_cast<u32*>(base + at[r]), assumed,
                        __float_as_uint(__uint_as_float(assumed) + p[r]));
      } while (cur != assumed);
    }
  }
}

// Sorts the n packed entries buf[0..n) of a whole-workgroup row (column
// << 32 | value bits; columns unique, all in [cmn, cmn + span)) by column,
// in place: least-significant-digit radix passes of 7 bits over
// column - cmn (three passes for a span of 2^20 columns).  A pass loads
// entry p = i * WG + t into thread t (i < E = ceil(n / WG) <= EMAX), ranks
// it among the same-digit entries of its wave from seven ballots, has the
// digit's first lane record the wave's count for block (i, wave), and one
// thread per digit turns the counts into exclusive prefixes in (i, wave)
// order; the digits' totals are scanned into bucket starts, so every entry's
// new position keeps the order of the pass before (stable), and the entries
// are written back.  cells holds EMAX x (WG / 64) x 128 u16 counts, start
// 128 ints.  Four barriers per pass (two in the scan) and none between
// passes' loads and the previous pass's stores beyond those.  Replaces a
// bitonic sort of 55-66 barrier-separated stages for rows of 1024-4096
// entries.
template <int WG, int EMAX>
__device__ __forceinline__ void radix_sort_row(u64* buf, int n, i32 cmn, u32 span, u16* cells,
                                               i32* start, i32* scr) {
  constexpr int NW = WG / WAVE;
  constexpr int DB = 7, ND = 1 << DB;
  static_assert(WG >= ND, "one thread per digit in the scan");
  const int t = threadIdx.x;
  const int w = t / WAVE;
  const int E = (n + WG - 1) / WG;  // uniform, <= EMAX
  const int bits = n <= 1 || span <= 1u ? 0 : 32 - __clz((int)(span - 1u));
  u32* cells32 = reinterpret_cast<u32*>(cells);
  for (int shift = 0; shift < bits; shift += DB) {
    for (int q = t; q < E * NW * (ND / 2); q += WG) cells32[q] = 0u;
    u64 x[EMAX];
    int dr[EMAX];  // digit | rank in the wave << DB
#pragma unroll
    for (int i = 0; i < EMAX; ++i) {
      const int p = i * WG + t;
      x[i] = i < E && p < n ? buf[p] : ~0ull;
    }
    __syncthreads();  // cells zeroed; every entry loaded before any is stored
#pragma unroll
    for (int i = 0; i < EMAX; ++i) {
      dr[i] = 0;
      if (i < E) {  // uniform
        const bool v = i * WG + t < n;
        const int d = v ? (int)((((u32)(x[i] >> 32) - (u32)cmn) >> shift) & (ND - 1)) : 0;
        u64 peers = __ballot(v);
#pragma unroll
        for (int b = 0; b < DB; ++b) {
          const u64 bb = __ballot(v && ((d >> b) & 1));
          peers &= ((d >> b) & 1) ? bb : ~bb;
        }
        const int r = (int)__builtin_amdgcn_mbcnt_hi((u32)(peers >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((u32)peers, 0u));
        dr[i] = d | r << DB;
        if (v && r == 0) cells[(i * NW + w) * ND + d] = (u16)__popcll(peers);
      }
    }
    __syncthreads();
    int tot = 0;
    if (t < ND) {
      for (int c = 0; c < E * NW; ++c) {
        u16* q = &cells[c * ND + t];
        const int v = *q;
        *q = (u16)tot;
        tot += v;
      }
    }
    int all;
    const int incl = block_scan_incl<WG>(tot, scr, &all);
    if (t < ND) start[t] = incl - tot;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < EMAX; ++i) {
      if (i < E && i * WG + t < n) {
        const int d = dr[i] & (ND - 1);
        const int pos = start[d] + cells[(i * NW + w) * ND + d] + (dr[i] >> DB);
        buf[min(pos, n - 1)] = x[i];
      }
    }
    __syncthreads();
  }
}

// LDS-hash row kernel.  A group of G lanes owns one output row and a table of
// T slots (linear probing, load factor <= 1/2 by binning).  The products of
// the row are enumerated flat: a chunk of up to G entries A(i,k) is loaded,
// their B-row lengths are prefix-summed, and lane l takes products l, l+G, ...
// of the chunk, locating its A entry by binary search over the prefix sums.
// Consecutive lanes therefore read consecutive entries of one B row.
//   MODE_SYM  : count distinct columns -> a.nnz[row].
//   MODE_NUM  : accumulate values, order the row by column (a bitmap over
//               its span, or a sort: radix for whole-workgroup rows, bitonic
//               within a wave) and write it at a.crp[row] of (a.ccol, a.cval).
//   MODE_LEARN: as MODE_NUM, but the row's size is a predicted capacity
//               a.cap[row] (the table has T >= 2*cap slots) and the row is
//               written at a.crp[row] of a temporary buffer; a row whose
//               distinct columns exceed the capacity is abandoned and flagged
//               in a.ovf[row] for the exact rerun.  a.nnz[row] gets the count.
constexpr int MODE_SYM = 0, MODE_NUM = 1, MODE_LEARN = 2;

// One product into a row's LDS hash table (T slots, linear probing), with
// the key of its first probe slot h already read (kk).  MODE_SYM counts
// new columns in mycnt; MODE_NUM adds v into the column's value; MODE_LEARN
// also counts, gives up after a probe run of PLIM slots and flags the row
// in *ovf instead (so that no lane can spin on a crowded table; the row's
// distinct count is checked against its predicted capacity afterwards).
// The flag is only written here: a row that overflowed keeps inserting its
// remaining products (each bounded by PLIM probes) and is rerun exactly, so
// no insert waits on a read of the flag.
template <int MODE, int T>
__device__ __forceinline__ void hash_insert(i32* keys, float* vals, i32 c, u32 h, i32 kk, float v,
                                            int& mycnt, i32* ovf) {
  if constexpr (MODE == MODE_LEARN) {
    constexpr int PLIM = T < 64 ? T : 64;
    bool found = false;
    for (int probe = 0; probe < PLIM; ++probe) {
      if (kk == c) { found = true; break; }
      if (kk == EMPTY_KEY) {
        kk = atomicCAS(&keys[h], EMPTY_KEY, c);
        if (kk == EMPTY_KEY) { ++mycnt; found = true; break; }
        if (kk == c) { found = true; break; }
      }
      h = (h + 1) & (T - 1);
      kk = keys[h];
    }
    if (found) lds_add_f32(&vals[h], v); else *ovf = 1;
  } else {
    while (true) {
      if (kk == c) break;
      if (kk == EMPTY_KEY) {
        kk = atomicCAS(&keys[h], EMPTY_KEY, c);
        if (kk == EMPTY_KEY) {
          if constexpr (MODE == MODE_SYM) ++mycnt;
          break;
        }
        if (kk == c) break;
      }
      h = (h + 1) & (T - 1);
      kk = keys[h];
    }
    if constexpr (MODE != MODE_SYM) lds_add_f32(&vals[h], v);
  }
}

struct HashArgs {
  const i32* list;
  i32 nlist;
  const i64* arp;
  const i32* acol;
  const float* aval;
  const i64* brp;
  const i32* bcol;
  const float* bval;
  const i64* bbase;
  i64* nnz;
  const i64* crp;
  i32* ccol;
  float* cval;
  const i64* cap;
  i32* ovf;
  // When dcount is set the kernel is persistent: its rows are the bin `bin`
  // of a device-side binning (count dcount[bin], first row list[doff[bin]]),
  // and its workgroups take runs of TICKET_RUN row groups in turn from the
  // counter ticket[bin] (zeroed with the binning), so that one that drew
  // cheap rows takes more.
  const i32* dcount;
  const i32* doff;
  int bin;
  i32* ticket;
  // B's column count (0: unknown).  Symbolic rows count their columns in a
  // bitmap instead of a hash table when every column fits the table's bits.
  i64 bcols;
};
constexpr int TICKET_RUN = 32;

// Minimum and maximum over a group of G lanes (G <= 64), or over the whole
// workgroup (G == WG, through scr[WG/64]); every lane receives both.
template <int G, int WG>
__device__ __forceinline__ void group_minmax(i32& mn, i32& mx, i32* scr) {
  constexpr int W = G <= WAVE ? G : WAVE;
#pragma unroll
  for (int d = W >> 1; d > 0; d >>= 1) {
    mn = min(mn, __shfl_xor(mn, d, W));
    mx = max(mx, __shfl_xor(mx, d, W));
  }
  if constexpr (G > WAVE) {
    constexpr int NW = WG / WAVE;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
    __syncthreads();
    if (lane == 0) scr[w] = mn;
    __syncthreads();
    for (int i = 0; i < NW; ++i) mn = min(mn, scr[i]);
    __syncthreads();
    if (lane == 0) scr[w] = mx;
    __syncthreads();
    for (int i = 0; i < NW; ++i) mx = max(mx, scr[i]);
    __syncthreads();
  }
}

template <int G, int T, int WG, int MODE, bool BATCH>
// A 512-thread workgroup keeps two resident per CU: four waves per SIMD
// (the second bound), so at most 128 VGPRs; so do the other whole-workgroup
// rows (their radix sort would otherwise take the 2048-slot bin past 128
// and cost it a wave per SIMD).  Groups within a wave holding
// at most 8 table slots per lane keep seven waves per SIMD (at most 72
// VGPRs): the learned mode's extra state would otherwise cost it a wave per
// SIMD against the exact numeric mode (measured 10% on its 512-slot bin).
__global__ __launch_bounds__(WG, WG >= 512 || G == WG ? 4 : (G <= WAVE && T / G <= 8 ? 7 : 1))
void k_hash(HashArgs a) {
  static_assert(G <= WAVE || G == WG, "a multi-wave group must be the whole workgroup");
  constexpr bool NUMERIC = MODE != MODE_SYM;
  constexpr int R = WG / G;
  constexpr int LT = Log2<T>::value;
  constexpr int SPL = T / G;
  __shared__ __align__(16) i32 s_keys[R * T];
  __shared__ float s_vals[NUMERIC ? R * T : 1];
  // Per-entry chunk data: B start (8 B), first product (4 B), A value (4 B).
  __shared__ __align__(16) unsigned char s_ent[WG * 16];
  i64* s_bs = reinterpret_cast<i64*>(s_ent);
  i32* s_ex = reinterpret_cast<i32*>(s_ent + 8 * WG);
  float* s_av = reinterpret_cast<float*>(s_ent + 12 * WG);
  __shared__ i32 s_scr[WG / WAVE];
  __shared__ i32 s_ovf[MODE == MODE_LEARN ? R : 1];
  // Column bitmap of a numeric row: the output order when the row's columns
  // span fewer than 32*T.  Once every lane holds its slots in registers the
  // table is free: the bitmap takes the key array (T words), the words'
  // prefixes of set bits the value array (u16), and the row is staged in
  // column order in the key array again.
  constexpr int BMW = T;
  constexpr int WPL = BMW / G;  // bitmap words per lane
  static_assert(BMW % G == 0, "bitmap words split evenly over the group");
  static_assert(T <= 65535, "u16 word prefixes (a row has fewer than T entries)");
  // Whole-workgroup rows enumerate their products with chunk_products (owner
  // marks or 64-product blocks, HR products per lane in flight).
  constexpr bool WHOLE = G == WG;
  constexpr int HR = MODE == MODE_LEARN ? 2 : 4;  // learned rows: fewer registers
  __shared__ i32 s_cex[WHOLE ? WG + 1 : 1];
  __shared__ __align__(16) u16 s_cmark[WHOLE ? HR * WG : 8];
  __shared__ u64 s_cscr[WHOLE ? WG / WAVE : 1];
  const i32 nlist = a.dcount ? a.dcount[a.bin] : a.nlist;
  const i32* __restrict__ list = a.dcount ? a.list + a.doff[a.bin] : a.list;
  const i64* __restrict__ arp = a.arp;
  const i32* __restrict__ acol = a.acol;
  const float* __restrict__ aval = a.aval;
  const i64* __restrict__ brp = a.brp;
  const i32* __restrict__ bcol = a.bcol;
  const float* __restrict__ bval = a.bval;

  const int g = threadIdx.x / G;
  const int l = threadIdx.x & (G - 1);
  // Symbolic mode over B columns below 32*T: the table's T words are a
  // bitmap of the row's columns (one LDS OR per product, no probing) and the
  // count is its population.  Uniform over the launch.
  const bool bmp = MODE == MODE_SYM && a.bcols > 0 && a.bcols <= (i64)32 * T;
  // The same condition for exact numeric whole-workgroup rows of up to 2048
  // slots: the row's columns are ranked through a bitmap and its values
  // accumulated densely in rank order (two walks over the products, no hash
  // table).  Larger tables keep the hash: the second path would cost their
  // kernels register spills.
  constexpr bool BMN = MODE == MODE_NUM && G == WG && T <= 2048;
  const bool bmn = BMN && a.bcols > 0 && a.bcols <= (i64)32 * T;
  i32* keys = s_keys + g * T;
  float* vals = NUMERIC ? s_vals + g * T : nullptr;
  const int gb = g * G;
  const i64 ngroups = ((i64)nlist + R - 1) / R;
  if constexpr (WHOLE) {
    for (int d = threadIdx.x; d < HR * WG; d += WG) s_cmark[d] = 0;  // zeroed between chunks
    __syncthreads();
  }
  // Row groups are taken grid-stride (a single pass when the grid covers the
  // bin), or, for a persistent launch, in runs from the bin's ticket counter
  // (one atomic per run: one per row group serialises on the counter).  The
  // loop is uniform over the workgroup; in block mode the tables are shared
  // by the whole workgroup, hence the barrier between groups.
  __shared__ i32 s_ticket;
  i64 run_end = 0;
  auto first = [&]() -> i64 {
    if (!a.ticket) return (i64)blockIdx.x;
    if (threadIdx.x == 0) s_ticket = atomicAdd(&a.ticket[a.bin], TICKET_RUN);
    __syncthreads();
    const i64 g0 = s_ticket;
    run_end = g0 + TICKET_RUN;
    return g0;
  };
  auto next = [&](i64 gi) -> i64 {
    if (!a.ticket) return gi + gridDim.x;
    if (gi + 1 < run_end) return gi + 1;
    __syncthreads();  // every thread has read the previous ticket
    if (threadIdx.x == 0) s_ticket = atomicAdd(&a.ticket[a.bin], TICKET_RUN);
    __syncthreads();
    const i64 g0 = s_ticket;
    run_end = g0 + TICKET_RUN;
    return g0;
  };
  for (i64 gi = first(); gi < ngroups; gi = next(gi)) {
  [&]() {
  const i64 ridx = gi * R + g;
  const bool active = ridx < nlist;

  // The row's header loads are issued first; the table is cleared while
  // they are in flight.
  const i32 row = active ? list[ridx] : 0;
  i64 a0 = 0, a1 = 0, bb = 0, cap = 0;
  if (active) {
    a0 = arp[row];
    a1 = arp[row + 1];
    if constexpr (BATCH) bb = a.bbase[row];
    if constexpr (MODE == MODE_LEARN) cap = a.cap[row];
  }
  for (int s = l; s < T; s += G) {
    keys[s] = bmp || bmn ? 0 : EMPTY_KEY;
    if constexpr (NUMERIC) vals[s] = 0.f;
  }
  if constexpr (MODE == MODE_LEARN) {
    if (l == 0) s_ovf[g] = 0;
  }
  group_sync<G>();
  int mycnt = 0;
  if constexpr (BMN) {
    if (bmn) {
      // LDS: the key array is the row's column bitmap (later its columns in
      // rank order); the value array holds each bitmap word's prefix count
      // (u16, first half) and the row's sums in rank order (second half; a
      // numeric row has at most T/2 entries).
      u32* bm = reinterpret_cast<u32*>(keys);
      u16* wpre = reinterpret_cast<u16*>(vals);
      float* acc = vals + T / 2;
      const ChunkLds<i64> cl{s_bs, s_av, s_cex, s_cmark, s_cscr};
      auto walk = [&](auto val, auto&& f) {
        constexpr bool V = decltype(val)::value;
        for (i64 e0 = a0; e0 < a1; e0 += WG) {  // uniform: one row per workgroup
          const i64 e = e0 + l;
          int len = 0;
          i64 bs = 0;
          float av = 0.f;
          if (e < a1) {
            const i64 k = (i64)acol[e] + bb;
            bs = brp[k];
            len = (int)(brp[k + 1] - bs);
            if constexpr (V) av = aval[e];
          }
          const int nvalid = (int)min((i64)WG, a1 - e0);
          if constexpr (V)
            chunk_products<WG, HR, true>(len, bs, av, nvalid, SplitEntries{bcol, bval}, cl, f);
          else
            chunk_products<WG, HR, false>(len, bs, 0.f, nvalid, ColumnEntries{bcol}, cl, f);
        }
      };
      walk(std::false_type{}, [&](const auto& c, const auto&, const auto& ok) {
#pragma unroll
        for (int r = 0; r < HR; ++r)
          if (ok[r]) lds_or_u32(bm + (((u32)c[r] >> 5) & (T - 1)), 1u << ((u32)c[r] & 31u));
      });
      __syncthreads();
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < WPL; ++j) cnt += __popc(bm[l * WPL + j]);
      int N;
      const int inc = block_scan_incl<WG>(cnt, s_scr, &N);
      {
        int run = inc - cnt;
#pragma unroll
        for (int j = 0; j < WPL; ++j) {
          wpre[l * WPL + j] = (u16)run;
          run += __popc(bm[l * WPL + j]);
        }
      }
      for (int q = l; q < N; q += WG) acc[q] = 0.f;
      __syncthreads();
      walk(std::true_type{}, [&](const auto& c, const auto& p, const auto& ok) {
#pragma unroll
        for (int r = 0; r < HR; ++r) {
          if (ok[r]) {
            const u32 x = (u32)c[r], wi = (x >> 5) & (T - 1);
            const int rk = wpre[wi] + __popc(bm[wi] & ((1u << (x & 31u)) - 1u));
            lds_add_f32(&acc[rk], p[r]);
          }
        }
      });
      __syncthreads();
      // Columns in rank order: every lane expands its own words.
      u32 w[WPL];
#pragma unroll
      for (int j = 0; j < WPL; ++j) w[j] = bm[l * WPL + j];
      __syncthreads();  // every word is read before the columns overwrite the bitmap
      {
        int run = inc - cnt;
#pragma unroll
        for (int j = 0; j < WPL; ++j) {
          u32 r = w[j];
          while (r) {
            const int b = __ffs(r) - 1;
            r &= r - 1u;
            bm[run++] = (u32)((l * WPL + j) * 32 + b);
          }
        }
      }
      __syncthreads();
      if (active) {
        const i64 off = a.crp[row];
        for (int q = l; q < N; q += WG) {
          a.ccol[off + q] = (i32)bm[q];
          a.cval[off + q] = acc[q];
        }
      }
      return;
    }
  }
  if constexpr (WHOLE) {
    // Batched insert of HR products per lane: the first probe of every row
    // is read before any is used.
    static_assert(HR == 2 || HR == 4, "named first probes");
    auto ins = [&](const auto& c, const auto& p, const auto& ok) {
      if constexpr (MODE == MODE_SYM) {
        if (bmp) {
#pragma unroll
          for (int r = 0; r < HR; ++r)
            if (ok[r])
              lds_or_u32(reinterpret_cast<u32*>(keys) + (((u32)c[r] >> 5) & (T - 1)),
                         1u << ((u32)c[r] & 31u));
          return;
        }
      }
      i32* ovf = MODE == MODE_LEARN ? &s_ovf[g] : nullptr;
      const u32 h0 = hash_col<LT>(c[0]), h1 = hash_col<LT>(c[1]);
      if constexpr (HR == 4) {
        const u32 h2 = hash_col<LT>(c[2]), h3 = hash_col<LT>(c[3]);
        const i32 q0 = keys[h0], q1 = keys[h1], q2 = keys[h2], q3 = keys[h3];
        if (ok[0]) hash_insert<MODE, T>(keys, vals, c[0], h0, q0, p[0], mycnt, ovf);
        if (ok[1]) hash_insert<MODE, T>(keys, vals, c[1], h1, q1, p[1], mycnt, ovf);
        if (ok[2]) hash_insert<MODE, T>(keys, vals, c[2], h2, q2, p[2], mycnt, ovf);
        if (ok[3]) hash_insert<MODE, T>(keys, vals, c[3], h3, q3, p[3], mycnt, ovf);
      } else {
        const i32 q0 = keys[h0], q1 = keys[h1];
        if (ok[0]) hash_insert<MODE, T>(keys, vals, c[0], h0, q0, p[0], mycnt, ovf);
        if (ok[1]) hash_insert<MODE, T>(keys, vals, c[1], h1, q1, p[1], mycnt, ovf);
      }
    };
    const ChunkLds<i64> cl{s_bs, s_av, s_cex, s_cmark, s_cscr};
    for (i64 e0 = a0; e0 < a1; e0 += WG) {  // uniform: one row per workgroup
      const i64 e = e0 + l;
      int len = 0;
      i64 bs = 0;
      float av = 0.f;
      if (e < a1) {
        const i64 k = (i64)acol[e] + bb;
        bs = brp[k];
        len = (int)(brp[k + 1] - bs);
        if constexpr (NUMERIC) av = aval[e];
      }
      const int nvalid = (int)min((i64)WG, a1 - e0);
      if constexpr (NUMERIC)
        chunk_products<WG, HR, true>(len, bs, av, nvalid, SplitEntries{bcol, bval}, cl, ins);
      else
        chunk_products<WG, HR, false>(len, bs, 0.f, nvalid, ColumnEntries{bcol}, cl, ins);
    }
  } else
  for (i64 e0 = a0; e0 < a1; e0 += G) {
    const i64 e = e0 + l;
    int len = 0;
    i64 bs = 0;
    float av = 0.f;
    if (e < a1) {
      const i64 k = (i64)acol[e] + bb;
      bs = brp[k];
      len = (int)(brp[k + 1] - bs);
      if constexpr (NUMERIC) av = aval[e];
    }
    int inc, total;
    if constexpr (G <= WAVE) {
      inc = seg_scan_incl<G>(len);
      total = __shfl(inc, G - 1, G);
    } else {
      inc = block_scan_incl<WG>(len, s_scr, &total);
    }
    s_ex[gb + l] = inc - len;
    s_bs[gb + l] = bs;
    if constexpr (NUMERIC) s_av[gb + l] = av;
    group_sync<G>();
    const int nvalid = (int)min((i64)G, a1 - e0);
    int lo = 0;
    // HU products per lane per iteration: their B loads are issued together
    // (unconditionally; lanes past the end read entry 0), then inserted.
    constexpr int HU = 4;
    for (int p0 = l; p0 < total; p0 += G * HU) {
      i64 idx[HU];
      float avu[HU];
      bool okp[HU];
#pragma unroll
      for (int u = 0; u < HU; ++u) {
        const int p = p0 + u * G;
        okp[u] = p < total;
        idx[u] = 0;
        avu[u] = 0.f;
        if (okp[u]) {
          lo = advance_entry(s_ex + gb, lo, nvalid, p);
          idx[u] = s_bs[gb + lo] + (p - s_ex[gb + lo]);
          if constexpr (NUMERIC) avu[u] = s_av[gb + lo];
        }
      }
      i32 cu[HU];
      float bvu[HU];
#pragma unroll
      for (int u = 0; u < HU; ++u) {
        cu[u] = bcol[idx[u]];
        if constexpr (NUMERIC) bvu[u] = bval[idx[u]];
      }
      if constexpr (MODE == MODE_SYM) {
        if (bmp) {
#pragma unroll
          for (int u = 0; u < HU; ++u)
            if (okp[u])
              lds_or_u32(reinterpret_cast<u32*>(keys) + (((u32)cu[u] >> 5) & (T - 1)),
                         1u << ((u32)cu[u] & 31u));
          continue;
        }
      }
#pragma unroll
      for (int u = 0; u < HU; ++u) {
        if (!okp[u]) continue;
        const i32 c = cu[u];
        const u32 h = hash_col<LT>(c);
        hash_insert<MODE, T>(keys, vals, c, h, keys[h], NUMERIC ? avu[u] * bvu[u] : 0.f, mycnt,
                             MODE == MODE_LEARN ? &s_ovf[g] : nullptr);
      }
    }
    group_sync<G>();
  }

  if constexpr (!NUMERIC) {
    if (bmp) {
      group_sync<G>();  // every bit is set
      mycnt = 0;
#pragma unroll
      for (int m = 0; m < SPL; ++m) mycnt += __popc((u32)keys[l + m * G]);
    }
    int tot;
    if constexpr (G <= WAVE) tot = seg_reduce_sum<G>(mycnt);
    else tot = block_reduce_sum<WG>(mycnt, s_scr);
    if (active && l == 0) a.nnz[row] = tot;
  } else {
    if constexpr (MODE == MODE_LEARN) {
      // s_ovf is uniform over the group (and over the workgroup when G == WG).
      int distinct;
      if constexpr (G <= WAVE) distinct = seg_reduce_sum<G>(mycnt);
      else distinct = block_reduce_sum<WG>(mycnt, s_scr);
      if (s_ovf[g] || distinct > cap) {
        if (active && l == 0) {
          a.ovf[row] = 1;
          a.nnz[row] = 0;
        }
        return;
      }
    }
    // Every lane takes its slots (packed column << 32 | value bits; ~0 when
    // empty) and the row's column range.  When the range is narrower than
    // the bitmap, the row's columns are set in the bitmap, each word's
    // prefix of set bits is stored, and every entry finds its rank in the
    // row (word prefix + set bits below it in the word): the row is staged
    // in column order and written with coalesced stores.  No sort, no second
    // probe.  Wider rows are compacted and sorted.
    u64 ent[SPL];
    i32 cmn = INT_MAX, cmx = -1;
    int mc = 0;
#pragma unroll
    for (int m = 0; m < SPL; ++m) {
      const int s = l + m * G;
      const i32 kk = keys[s];
      ent[m] = ~0ull;
      if (kk != EMPTY_KEY) {
        ent[m] = ((u64)(u32)kk << 32) | __float_as_uint(vals[s]);
        cmn = min(cmn, kk);
        cmx = max(cmx, kk);
        ++mc;
      }
    }
    group_minmax<G, WG>(cmn, cmx, s_scr);
    group_sync<G>();  // every slot is read before the table is reused
    if (cmx >= 0 && (i64)cmx - cmn < (i64)BMW * 32) {  // uniform over the group
      u32* bm = reinterpret_cast<u32*>(keys);
      u16* wpre = reinterpret_cast<u16*>(vals);
      for (int d = l; d < BMW; d += G) bm[d] = 0u;
      group_sync<G>();
#pragma unroll
      for (int m = 0; m < SPL; ++m) {
        if (ent[m] != ~0ull) {
          const int x = (i32)(ent[m] >> 32) - cmn;
          lds_or_u32(&bm[x >> 5], 1u << (x & 31));
        }
      }
      group_sync<G>();
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < WPL; ++j) cnt += __popc(bm[l * WPL + j]);
      int inc, N;
      if constexpr (G <= WAVE) {
        inc = seg_scan_incl<G>(cnt);
        N = __shfl(inc, G - 1, G);
      } else {
        inc = block_scan_incl<WG>(cnt, s_scr, &N);
      }
      {
        int run = inc - cnt;
#pragma unroll
        for (int j = 0; j < WPL; ++j) {
          wpre[l * WPL + j] = (u16)run;
          run += __popc(bm[l * WPL + j]);
        }
      }
      group_sync<G>();
      int rk[SPL];
#pragma unroll
      for (int m = 0; m < SPL; ++m) {
        rk[m] = -1;
        if (ent[m] != ~0ull) {
          const int x = (i32)(ent[m] >> 32) - cmn;
          rk[m] = wpre[x >> 5] + __popc(bm[x >> 5] & ((1u << (x & 31)) - 1u));
        }
      }
      group_sync<G>();  // every rank is known before the staging overwrites the bitmap
      u64* buf = reinterpret_cast<u64*>(keys);  // T/2 packed entries; N <= T/2
#pragma unroll
      for (int m = 0; m < SPL; ++m)
        if (rk[m] >= 0) buf[rk[m]] = ent[m];
      group_sync<G>();
      if (active) {
        const i64 off = a.crp[row];
        for (int q = l; q < N; q += G) {
          const u64 v = buf[q];
          a.ccol[off + q] = (i32)(v >> 32);
          a.cval[off + q] = __uint_as_float((u32)v);
        }
        if constexpr (MODE == MODE_LEARN) {
          if (l == 0) a.nnz[row] = N;
        }
      }
      return;
    }
    // Otherwise compact the entries ...
    int inc, N;
    if constexpr (G <= WAVE) {
      inc = seg_scan_incl<G>(mc);
      N = __shfl(inc, G - 1, G);
    } else {
      inc = block_scan_incl<WG>(mc, s_scr, &N);
    }
    int pos = inc - mc;
    u64* buf = reinterpret_cast<u64*>(keys);  // T/2 packed entries; N <= T/2
#pragma unroll
    for (int m = 0; m < SPL; ++m)
      if (ent[m] != ~0ull) buf[pos++] = ent[m];
    if constexpr (WHOLE) {
      // ... radix-sorted by column (whole-workgroup rows; the value array is
      // free for the digit counts) ...
      static_assert(T / (2 * WG) * (WG / WAVE) * 256 + 512 <= 4 * T, "radix counts fit the values");
      group_sync<G>();
      radix_sort_row<WG, T / (2 * WG)>(
          buf, N, cmn, (u32)(cmx - cmn) + 1u, reinterpret_cast<u16*>(vals),
          reinterpret_cast<i32*>(reinterpret_cast<unsigned char*>(vals) + 2 * T), s_scr);
    } else {
      int S = 1;
      while (S < N) S <<= 1;
      for (int q = N + l; q < S; q += G) buf[q] = ~0ull;
      group_sync<G>();
      // ... bitonic sort ascending (columns are unique, so the order is total) ...
      for (int k = 2; k <= S; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int p = l; p < (S >> 1); p += G) {
            const int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
            const int ixj = i + j;
            const u64 x = buf[i], y = buf[ixj];
            const bool up = (i & k) == 0;
            if ((x > y) == up) {
              buf[i] = y;
              buf[ixj] = x;
            }
          }
          group_sync<G>();
        }
      }
    }
    // ... and write the row.
    if (active) {
      const i64 off = a.crp[row];
      for (int q = l; q < N; q += G) {
        const u64 v = buf[q];
        a.ccol[off + q] = (i32)(v >> 32);
        a.cval[off + q] = __uint_as_float((u32)v);
      }
      if constexpr (MODE == MODE_LEARN) {
        if (l == 0) a.nnz[row] = N;
      }
    }
  }
  }();
  group_sync<G>();
  }
}

// ---------------------------------------------------------------------------
// Dynamic row scheduling for the one-row-per-workgroup kernels: a workgroup
// takes its next list position from a device counter (zeroed before the
// launch), so one that drew cheap rows takes more of them.  The grid is one
// resident workgroup per CU.  Contains two barriers.
__device__ __forceinline__ i64 next_slot(i32* next, i32* s_slot) {
  __syncthreads();  // every thread has read the previous slot
  if (threadIdx.x == 0) *s_slot = atomicAdd(next, 1);
  __syncthreads();
  return *s_slot;
}

// LDS-bitmap symbolic kernel for long rows: one workgroup of WG threads per
// row (resident: 1024 threads per CU), one bit per output column over a
// window of WG/16 tiles (2^19 columns and 64 KB of LDS with 512 threads, half
// that with 256); wider matrices take several windows, and the slab pointers of B restrict every window to its
// own products, so each product is gathered once.  With PAIRS, bcol is a
// copy of B's columns with a spare entry past the end (k_copy_cols) and the
// block walk gathers two neighbouring columns per lane.  Records for every tile of
// the row its entry count and its count of nonzero bitmap words, packed as
// entries | words << 16 (the long-row numeric kernels' step plan and output
// offsets), and, when rowbits is set, the row's bitmap in compressed form:
// rowmask gets per tile 16 masks of nonzero words (one per 32-word group),
// rowbits the nonzero words themselves, packed in column order.  About a
// quarter of an R-MAT long row's bitmap words are nonzero.
// The pair walk runs in workgroups of SYM_PAIR_WG threads, four resident per
// CU (16-tile windows of 2^18 columns, 40 KB of LDS each): twice the rows in
// flight per CU hide more of each window's setup and epilogue latency than
// the 512-thread workgroups' wider windows save (128 threads, eight per CU:
// slower again).
constexpr int SYM_PAIR_WG = 256;

template <bool BATCH, class OffT, int ROWS, bool PAIRS, int WG = HWG>
__global__ __launch_bounds__(WG, 4) void k_sym_bitmap(
    RowSet rs, const i64* __restrict__ arp,
    const i32* __restrict__ acol, const i64* __restrict__ brp,
    const i32* __restrict__ bcol, const i64* __restrict__ bbase, const i32* __restrict__ slab,
    i64 ncols, i32 nt, i32* __restrict__ tilecnt, i64* __restrict__ out_nnz,
    u32* __restrict__ rowbits, u64* __restrict__ rowmask, i32* __restrict__ next) {
  // A window of one 32-word group per thread: WT tiles.
  constexpr int WT = WG / (TILE_W / 32 / 32);
  constexpr int SYM_WIN_TILES = WT;
  constexpr int SYM_WIN = WT * TILE_W;
  constexpr int WW = SYM_WIN / 32;  // bitmap words per window
  constexpr int GPT = TILE_W / 32 / 32;  // 32-word groups per tile
  static_assert(WW / 32 == WG, "one 32-word group of the window per thread");
  static_assert((WW & (WW - 1)) == 0, "windows of a power of two words");
  constexpr u32 WWB = __builtin_ctz(WW);  // bits of a word index in the window
  const i32 nlist = rs.count();
  const i32* __restrict__ list = rs.rows();
  // The window's bitmap, with one spare word after every 32-word group: word
  // w sits at w + w / 32, so group g starts on bank g mod 32 and the
  // threads reading their own groups in the epilogue do not collide.
  constexpr int GW = 33;  // stored words per group
  // (rounded up to whole 16-byte words, which the clear writes)
  __shared__ __align__(16) u32 s_bits[(WW / 32 * GW + 3) & ~3];
  __shared__ OffT s_off[WG];
  __shared__ i32 s_ex[WG + 1];
  __shared__ __align__(16) u16 s_mark[ROWS * WG];
  __shared__ u64 s_scr[WG / WAVE];
  __shared__ i64 s_scr64[WG / WAVE];
  const ChunkLds<OffT> cl{s_off, nullptr, s_ex, s_mark, s_scr};
  using Src = std::conditional_t<PAIRS, ColumnPairs, ColumnEntries>;
  const Src src{bcol};
  const int l = threadIdx.x;
  for (int d = l; d < ROWS * WG; d += WG) s_mark[d] = 0;
  for (int d = l; d < (int)(sizeof(s_bits) / 16); d += WG)
    reinterpret_cast<uint4*>(s_bits)[d] = make_uint4(0u, 0u, 0u, 0u);
  __shared__ i32 s_slot;
  for (i64 it = next_slot(next, &s_slot); it < nlist; it = next_slot(next, &s_slot)) {
    const i32 row = list[it];
    const i64 a0 = arp[row], a1 = arp[row + 1];
    const i64 bb = BATCH ? bbase[row] : 0;
    i64 total = 0;
    int wcarry = 0;  // nonzero words of the row stored so far
    for (int tw0 = 0; tw0 < nt; tw0 += SYM_WIN_TILES) {
      const int tw1 = min(nt, tw0 + SYM_WIN_TILES);
      const i64 win = (i64)tw0 * TILE_W;
      const i64 wcols = min((i64)SYM_WIN, ncols - win);
      const int nd = (int)((wcols + 31) / 32);
      // s_bits is all zero here: cleared at the start and left zero by
      // every window's epilogue.
      auto set_bits = [&](const auto& c, const auto&, const auto& ok) {
        if constexpr (PAIRS) {
          // Entries 2r and 2r+1 of a lane are neighbouring products of one
          // B row in the block walk (often in one word when the row is
          // dense): when their words match, one atomic sets both bits.
#pragma unroll
          for (int r = 0; r < ROWS; r += 2) {
            // Windows start at multiples of SYM_WIN, so a product's word in
            // the window is a bit field of its column (the slab bounds keep
            // the columns of busy slots in the window) and its bit is the
            // column's low 5 bits.
            const u32 w0 = __builtin_amdgcn_ubfe((u32)c[r], 5u, WWB);
            const u32 w1 = __builtin_amdgcn_ubfe((u32)c[r + 1], 5u, WWB);
            u32 b0 = ok[r] ? 1u << ((u32)c[r] & 31u) : 0u;
            u32 b1 = ok[r + 1] ? 1u << ((u32)c[r + 1] & 31u) : 0u;
            if (w0 == w1) {
              b0 |= b1;
              b1 = 0u;
            }
            // Byte offsets (w + w / 32) * 4: one shift and one add-shift.
            if (b0)
              lds_or_u32(reinterpret_cast<u32*>(reinterpret_cast<unsigned char*>(s_bits) +
                                                ((w0 + (w0 >> 5)) << 2)), b0);
            if (b1)
              lds_or_u32(reinterpret_cast<u32*>(reinterpret_cast<unsigned char*>(s_bits) +
                                                ((w1 + (w1 >> 5)) << 2)), b1);
          }
        } else {
          // Lanes 2k and 2k+1 hold neighbouring products (consecutive
          // entries of a B row, often in one word when the row is dense):
          // when their words match, the even lane sets both bits and the
          // odd lane none, which halves the same-address atomics a run of
          // one word costs.  (Merging quads as well, through a second
          // exchange, measured no faster: the added VALU work cost what the
          // atomics saved.)
          const bool odd = threadIdx.x & 1;
#pragma unroll
          for (int r = 0; r < ROWS; ++r) {
            const int x = (int)((i64)c[r] - win);  // in [0, wcols) by the slab bounds
            const int w = x >> 5;
            u32 bit = ok[r] ? 1u << (x & 31) : 0u;
            const int wp = __builtin_amdgcn_mov_dpp(w, 0xB1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
            const u32 bp = (u32)__builtin_amdgcn_mov_dpp((int)bit, 0xB1, 0xf, 0xf, false);
            if (wp == w) bit = odd ? 0u : bit | bp;
            if (bit) lds_or_u32(&s_bits[w + (w >> 5)], bit);
          }
        }
      };
      for (i64 e0 = a0; e0 < a1; e0 += WG) {
        const i64 e = e0 + l;
        int len = 0;
        OffT start = 0;
        if (e < a1) {
          const i64 k = (i64)acol[e] + bb;
          const i32* sk = slab + k * (nt + 1);
          const i32 s0 = sk[tw0];
          start = (OffT)(brp[k] + s0);
          len = sk[tw1] - s0;
        }
        const int nvalid = (int)min((i64)WG, a1 - e0);
        chunk_products<WG, ROWS, false>(len, start, 0.f, nvalid, src, cl, set_bits);
      }
      const int ntw = tw1 - tw0;
      i64 wsum = 0;  // the window's entries (rowbits path)
      // Thread l owns the window's 32-word group l (words past nd are stale
      // and count as zero): its nonzero-word mask and entry count.  The 16
      // groups of a tile sit on 16 consecutive lanes.
      u32 gw[32];
      u32 gmask = 0;
      int gent = 0;
      {
        const int lim = nd - l * 32;
        u32* q = s_bits + l * GW;
#pragma unroll
        for (int j = 0; j < 32; ++j) gw[j] = q[j];
        // The group's words go back to zero (the packing below writes only
        // below wtot, and the stores zero what they take from there).
#pragma unroll
        for (int j = 0; j < 32; ++j) q[j] = 0u;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          if (j >= lim) gw[j] = 0u;
          gmask |= (gw[j] != 0u ? 1u : 0u) << j;
          gent += __popc(gw[j]);
        }
      }
      const int gcnt = __popc(gmask);
      // Tile counts: entries | nonzero words << 16 (entries <= 16384).
      const int tent = seg_reduce_sum<GPT>(gent), twd = seg_reduce_sum<GPT>(gcnt);
      const int tl = l / GPT;
      if ((l & (GPT - 1)) == 0 && tl < ntw) tilecnt[it * nt + tw0 + tl] = tent | (twd << 16);
      if (rowbits) {
        // Keep the row's column set for the numeric pass, compressed: per
        // tile the 16 group masks, and the nonzero words packed in column
        // order (a tile's words start after those of the row's earlier
        // tiles).  The words are packed in place in s_bits (every thread
        // holds its own; the scan's barriers order the reads before the
        // writes) and written with coalesced stores.  Both are read once by
        // the numeric pass: non-temporal stores keep them from evicting B's
        // rows from L2.
        // One scan gives the word positions and (high half) the window's
        // entry count.
        u64 both;
        const int gpos =
            (int)(u32)block_scan_incl<WG>(((u64)(u32)gent << 32) | (u32)gcnt, s_scr, &both) -
            gcnt;
        const int wtot = (int)(u32)both;
        wsum = (i64)(both >> 32);
        if (tl < ntw)
          __builtin_nontemporal_store((u64)gmask | ((u64)(u32)gent << 32),
                                      &rowmask[it * ((i64)nt * GPT) + (i64)tw0 * GPT + l]);
        int k = gpos;
#pragma unroll
        for (int j = 0; j < 32; ++j)
          if (gw[j] != 0u) s_bits[k++] = gw[j];
        __syncthreads();
        u32* dst = rowbits + it * ((i64)nt * (TILE_W / 32)) + wcarry;
        // Lanes shifted to 128-byte line boundaries (non-temporal stores are
        // not merged in L2).
        const int q0 = l - (int)(reinterpret_cast<uintptr_t>(dst) / 4 & 31);
        for (int q = q0; q < wtot; q += WG)
          if (q >= 0) {
            __builtin_nontemporal_store(s_bits[q], &dst[q]);
            s_bits[q] = 0u;
          }
        wcarry += wtot;
      }
      if (rowbits) {
        total += wsum;
        __syncthreads();  // s_bits is read before the next window clears it
      } else {
        // (ends with a barrier: s_bits is read before the next window clears it)
        total += block_reduce_sum<WG>((i64)gent, s_scr64);
      }
    }
    if (l == 0) out_nnz[row] = total;
    __syncthreads();
  }
}

// Per B row, the offset (relative to the row start) of the first entry in
// each TILE_W-column tile: slab[k*(nt+1)+t], t = 0..nt.  Grid-stride; with
// maxp set, nothing is written when no row has more than th products (as
// k_copy_cols).
__global__ __launch_bounds__(256) void k_slab_ptrs(i64 brows, i32 nt, const i64* __restrict__ brp,
                                                   const i32* __restrict__ bcol,
                                                   i32* __restrict__ slab,
                                                   const unsigned long long* __restrict__ maxp,
                                                   i64 th) {
  if (maxp && (i64)*maxp <= th) return;
  const i64 total = brows * (i64)(nt + 1);
  for (i64 gid = (i64)blockIdx.x * 256 + threadIdx.x; gid < total;
       gid += (i64)gridDim.x * 256) {
  const i64 k = gid / (nt + 1);
  const int t = (int)(gid - k * (nt + 1));
  const i64 r0 = brp[k], r1 = brp[k + 1];
  i64 res;
  if (t == 0) {
    res = 0;
  } else if (t == nt) {
    res = r1 - r0;
  } else {
    const i64 target = (i64)t * TILE_W;
    i64 lo = r0, hi = r1;
    while (lo < hi) {
      const i64 mid = (lo + hi) >> 1;
      if (bcol[mid] < target) lo = mid + 1; else hi = mid;
    }
    res = lo - r0;
  }
  slab[gid] = (i32)res;
  }
}

// Dense LDS column-tile numeric kernel for long rows.  One workgroup per row
// (grid-stride over the rows the LDS-bitmap symbolic kernel handled, keeping
// those whose key exceeds `th`).  The row's 32768-column tiles are processed
// in order: a tile's 32768 fp32 accumulators and presence bits live in LDS,
// the row's products restricted to the tile are read through the B slab
// pointers, and the tile is emitted in column order by a scan over the
// presence bits (no sort), at a running output offset.  The A row (up to
// HCH*1024 entries) is read once and kept in registers across tiles, the
// next tile's slab boundaries are prefetched while the current tile runs,
// and tiles the symbolic pass found empty are skipped without a barrier.
constexpr int HCH = 4;
constexpr int TC_LDS = 256;  // tile counts staged in LDS (ncols <= 8M)

// Accumulates the products of one chunk of A entries (len/start/a given per
// thread) into the LDS tile and its presence bits; returns the chunk's
// product count (uniform).
template <int WG>
__device__ __forceinline__ int tile_chunk(int len, i64 start, float av, i64 cbase,
                                          const i32* __restrict__ bcol,
                                          const float* __restrict__ bval, float* s_vals,
                                          u32* s_bits, const ChunkLds<i64>& cl, int nvalid) {
  return chunk_products<WG, 4, true>(len, start, av, nvalid, SplitEntries{bcol, bval}, cl,
                                     [&](const auto& c, const auto& p, const auto& ok) {
    int x[4];
    u32 old[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) x[r] = (int)(c[r] - cbase);  // in the tile for every lane
#pragma unroll
    for (int r = 0; r < 4; ++r) old[r] = __float_as_uint(s_vals[x[r]]);
    lds_add_f32_rows<4>(s_vals, x, p, ok, old);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const u32 bit = ok[r] ? 1u << (x[r] & 31) : 0u;
      if (bit) lds_or_u32(&s_bits[x[r] >> 5], bit);
    }
  });
}

// Position of the r-th (0-based) set bit of b; requires r < popcount(b).
__device__ __forceinline__ int select_bit(u32 b, int r) {
  int pos = 0, c;
  c = __popc(b & 0xFFFFu);
  if (r >= c) { r -= c; b >>= 16; pos += 16; }
  c = __popc(b & 0xFFu);
  if (r >= c) { r -= c; b >>= 8; pos += 8; }
  c = __popc(b & 0xFu);
  if (r >= c) { r -= c; b >>= 4; pos += 4; }
  c = __popc(b & 0x3u);
  if (r >= c) { r -= c; b >>= 2; pos += 2; }
  c = (int)(b & 1u);
  if (r >= c) pos += 1;
  return pos;
}

// Emits the finished LDS tile at `out` in column order and re-zeroes it.
// Output position q is written by lane q mod WG (coalesced): its word comes
// from a binary search over the per-word popcount prefix, its bit from a
// select inside the word.  Returns the tile's entry count, or -1 (and
// writes nothing) when that count exceeds `limit`.
template <int WG>
__device__ __forceinline__ int tile_emit(i64 out, i64 cbase, i64 limit, float* s_vals,
                                         u32* s_bits, i32* s_pfx, i32* s_scr,
                                         i32* __restrict__ ccol, float* __restrict__ cval) {
  constexpr int NW = TILE_W / 32;
  static_assert(NW == WG, "one presence word per thread");
  const int l = threadIdx.x;
  const u32 bits = s_bits[l];
  const int pc = __popc(bits);
  int tot;
  const int inc = block_scan_incl<WG>(pc, s_scr, &tot);
  s_pfx[l] = inc - pc;
  __syncthreads();
  const bool fits = (i64)tot <= limit;
  if (fits) {
    int w = 0;
    for (int q = l; q < tot; q += WG) {
      int hi = NW - 1;
      while (w < hi) {
        const int mid = (w + hi + 1) >> 1;
        if (s_pfx[mid] <= q) w = mid; else hi = mid - 1;
      }
      const int cl = w * 32 + select_bit(s_bits[w], q - s_pfx[w]);
      ccol[out + q] = (i32)(cbase + cl);
      cval[out + q] = s_vals[cl];
    }
  }
  __syncthreads();
  u32 b = bits;
  while (b) {
    const int bi = __ffs(b) - 1;
    b &= b - 1;
    s_vals[l * 32 + bi] = 0.f;
  }
  s_bits[l] = 0;
  __syncthreads();
  return fits ? tot : -1;
}

// LEARN = false: exact rows (key > th) of the symbolic bitmap list, output at
//   crp[row] of C, empty tiles skipped by the symbolic tile counts.
// LEARN = true: rows sized by the predictor, output at crp[row] (= the
//   temporary offsets) with capacity cap[row]; nnz[row] receives the count,
//   or ovf[row] = 1 when the row outgrows its capacity (the rest of the row
//   is then abandoned).  No tile counts: a tile whose chunk scans total zero
//   products is skipped.
template <bool BATCH, bool LEARN>
__global__ __launch_bounds__(HWG, 4) void k_num_heavy(
    RowSet rs, i32 nt, const i32* __restrict__ tilecnt,
    const i64* __restrict__ key, i64 th, const i64* __restrict__ arp,
    const i32* __restrict__ acol, const float* __restrict__ aval,
    const i64* __restrict__ brp, const i32* __restrict__ bcol,
    const float* __restrict__ bval, const i64* __restrict__ bbase,
    const i32* __restrict__ slab, const i64* __restrict__ crp, i32* __restrict__ ccol,
    float* __restrict__ cval, const i64* __restrict__ cap, i64* __restrict__ nnz,
    i32* __restrict__ ovf, i32* __restrict__ next) {
  constexpr int WG = HWG;
  const i32 nlist = rs.count();
  const i32* __restrict__ list = rs.rows();
  __shared__ float s_vals[TILE_W];
  __shared__ u32 s_bits[TILE_W / 32];
  __shared__ i64 s_off[WG];
  __shared__ i32 s_ex[WG + 1];
  __shared__ float s_av[WG];
  __shared__ __align__(16) u16 s_mark[4 * WG];
  __shared__ i32 s_scr[WG / WAVE];
  __shared__ u64 s_cscr[WG / WAVE];
  __shared__ i32 s_tc[TC_LDS];
  // The emit's word prefix reuses s_off (rewritten at every chunk start).
  i32* s_pfx = reinterpret_cast<i32*>(s_off);
  const ChunkLds<i64> cl{s_off, s_av, s_ex, s_mark, s_cscr};
  const int l = threadIdx.x;
  for (int d = l; d < TILE_W; d += WG) s_vals[d] = 0.f;
  for (int d = l; d < 4 * WG; d += WG) s_mark[d] = 0;
  s_bits[l] = 0;
  __syncthreads();
  __shared__ i32 s_slot;
  for (i64 h = next_slot(next, &s_slot); h < nlist; h = next_slot(next, &s_slot)) {
    const i32 row = list[h];
    if (key[row] <= th) continue;
    const i32* tc = LEARN ? nullptr : tilecnt + h * nt;
    if constexpr (!LEARN) {
      for (int t = l; t < min(nt, TC_LDS); t += WG) s_tc[t] = tc[t];
    }
    const i64 a0 = arp[row], a1 = arp[row + 1], nA = a1 - a0;
    const i64 bb = BATCH ? bbase[row] : 0;
    const i64 base = crp[row];
    const i64 capacity = LEARN ? cap[row] : LLONG_MAX;
    i64 out = base;
    bool over = false;
    __syncthreads();
    if (nA <= (i64)HCH * WG) {
      const int nch = (int)((nA + WG - 1) / WG);
      bool ok[HCH];
      float av[HCH];
      i64 b0[HCH];
      const i32* sp[HCH];
      i32 cur[HCH], nxt[HCH];
#pragma unroll
      for (int c = 0; c < HCH; ++c) {
        const i64 e = a0 + (i64)c * WG + l;
        ok[c] = e < a1;
        av[c] = 0.f;
        b0[c] = 0;
        sp[c] = slab;
        cur[c] = nxt[c] = 0;
        if (ok[c]) {
          const i64 k = (i64)acol[e] + bb;
          av[c] = aval[e];
          b0[c] = brp[k];
          sp[c] = slab + k * (nt + 1);
          nxt[c] = sp[c][1];
        }
      }
      for (int t = 0; t < nt && !over; ++t) {
        const int cnt = LEARN ? 1 : (t < TC_LDS ? s_tc[t] : tc[t]);
        i32 pre[HCH];
#pragma unroll
        for (int c = 0; c < HCH; ++c) pre[c] = (ok[c] && t + 2 <= nt) ? sp[c][t + 2] : 0;
        if (cnt > 0) {
          const i64 cbase = (i64)t * TILE_W;
          int any = 0;
#pragma unroll
          for (int c = 0; c < HCH; ++c) {
            if (c < nch) {
              const int nvalid = (int)min((i64)WG, nA - (i64)c * WG);
              any += tile_chunk<WG>(ok[c] ? nxt[c] - cur[c] : 0, b0[c] + cur[c], av[c], cbase,
                                    bcol, bval, s_vals, s_bits, cl, nvalid);
            }
          }
          if (any > 0) {
            const int n = tile_emit<WG>(out, cbase, capacity - (out - base), s_vals, s_bits,
                                        s_pfx, s_scr, ccol, cval);
            if (n < 0) over = true; else out += n;
          }
        }
#pragma unroll
        for (int c = 0; c < HCH; ++c) {
          cur[c] = nxt[c];
          nxt[c] = pre[c];
        }
      }
    } else {
      // Very long A rows: slab boundaries are read per tile and chunk.
      for (int t = 0; t < nt && !over; ++t) {
        const int cnt = LEARN ? 1 : (t < TC_LDS ? s_tc[t] : tc[t]);
        if (cnt == 0) continue;
        const i64 cbase = (i64)t * TILE_W;
        int any = 0;
        for (i64 e0 = a0; e0 < a1; e0 += WG) {
          const i64 e = e0 + l;
          int len = 0;
          i64 start = 0;
          float a = 0.f;
          if (e < a1) {
            const i64 k = (i64)acol[e] + bb;
            const i32* sk = slab + k * (nt + 1);
            const i32 s0 = sk[t], s1 = sk[t + 1];
            start = brp[k] + s0;
            len = s1 - s0;
            a = aval[e];
          }
          any += tile_chunk<WG>(len, start, a, cbase, bcol, bval, s_vals, s_bits, cl,
                                (int)min((i64)WG, a1 - e0));
        }
        if (any > 0) {
          const int n = tile_emit<WG>(out, cbase, capacity - (out - base), s_vals, s_bits, s_pfx,
                                      s_scr, ccol, cval);
          if (n < 0) over = true; else out += n;
        }
      }
    }
    if constexpr (LEARN) {
      if (l == 0) {
        nnz[row] = over ? 0 : out - base;
        if (over) ovf[row] = 1;
      }
    }
    __syncthreads();  // s_tc is reloaded for the next row
  }
}

// ---------------------------------------------------------------------------
// Segmented long-row numeric kernel (exact method, row bitmaps available).
//
// The symbolic pass stored each long row's column set as a compressed bitmap
// (group masks and packed nonzero words), and per tile its entry count and
// nonzero-word count.  With them the row is
// processed in few steps.  A "compressed" step covers up to SEG_MAXT
// consecutive tiles (2^20 columns) whose accumulators, staged words and
// group table fit the 64 KB accumulator region: only the window's nonzero
// bitmap words are staged in LDS (entry prefix, word index, bits), with a
// table of its 32-word groups (the slot of the group's first nonzero word
// and a mask of its nonzero words).  A product's rank is two LDS reads away
// (group -> word), so the accumulators sit in column order.  A first tile
// whose compressed form alone exceeds that budget is one "dense" step: 16384
// LDS accumulators indexed by column.  (Choosing dense steps more often --
// for tiles of more than 1K to 8K entries -- measured slower on R-MAT s20:
// the row then takes more, narrower steps, and per-step work dominates.)
// Output needs no search and no sort.
//
// Rows of up to HC*WG entries keep their entries in registers for the whole
// row: B row start, slab row, A value, and a cursor into the B row that
// advances from step to step (a step's B range of entry k ends at
// slab[k][t1], and the next step starts there, because the tiles skipped in
// between hold no product of the row).  The next step's range ends are
// loaded while the current step runs.  Longer rows re-read their entries per
// step and chunk.
constexpr int SEG_MAXT = 64;      // tiles per compressed step: one per lane of a planning wave
constexpr int SEG_DENSE = TILE_W;  // a first tile with more entries is a dense step
constexpr int HSEG_LDS = TILE_W * 4;
constexpr int SEG_BUDGET = HSEG_LDS - 8;             // compressed-step bytes (alignment slack)
constexpr int SEG_GPT = TILE_W / 32 / 32;            // 32-word groups per tile
constexpr int SEG_TILE_GT = SEG_GPT * 8;             // group-table bytes per tile
// Compressed-step layout of the accumulator region (bytes), for n entries in
// w nonzero words over 16 groups per tile: accumulators [0, 4n); staged
// words from OW = 4n rounded up to 8 (u64: word index << 48 | byte offset
// of the word's first accumulator << 32 | bits); the group table from
// OW + 8w (u64: byte offset of the group's first staged word << 32 |
// nonzero-word mask).  Offsets count from the region's start, so a product
// finds its accumulator with two LDS reads and no subtraction of the
// step's first column.  A step's tiles cost 4c + 8w + 128 bytes each.
static_assert(SEG_MAXT * TILE_W / 32 <= (1 << 15) && HSEG_LDS / 4 < (1 << 17),
              "staged word fields");

// Tile counts as the symbolic pass packs them: entries | nonzero words << 16.
__device__ __forceinline__ int tile_entries(int v) { return v & 0xffff; }
__device__ __forceinline__ int tile_words(int v) { return v >> 16; }

struct SegStep {
  int t, t1;   // tiles [t, t1)
  int n, w;    // entries; nonzero bitmap words
  bool dense;
};

__device__ __forceinline__ int tile_count(const i32* s_tc, const i32* tc, int nt, int j) {
  return j < nt ? (j < TC_LDS ? s_tc[j] : tc[j]) : 0;
}

// Plans the step that starts at the first nonempty tile at or after t
// (t = nt when none is left).  Lane j of every wave looks at tile t + j: wave
// scans of the entry and word counts and a ballot find the longest window
// within SEG_MAXT tiles and SEG_BUDGET bytes; trailing empty tiles are
// dropped.  The results are uniform; every lane of the wave must be active.
__device__ __forceinline__ SegStep seg_plan(const i32* s_tc, const i32* tc, int nt, int t) {
  static_assert(SEG_MAXT == WAVE, "one tile per lane");
  const int lane = threadIdx.x & (WAVE - 1);
  SegStep p{nt, nt, 0, 0, false};
  int v;
  for (;; t += WAVE) {
    if (t >= nt) return p;
    v = tile_count(s_tc, tc, nt, t + lane);
    const u64 ne = __ballot(v != 0);
    if (ne) {
      const int f = __ffsll((long long)ne) - 1;
      if (f > 0) {
        t += f;
        v = tile_count(s_tc, tc, nt, t + lane);
      }
      break;
    }
  }
  p.t = t;
  const int c = tile_entries(v), wd = tile_words(v);
  const int c0 = __builtin_amdgcn_readfirstlane(c), w0 = __builtin_amdgcn_readfirstlane(wd);
  p.dense = c0 > SEG_DENSE || 4 * c0 + 8 * w0 + SEG_TILE_GT > SEG_BUDGET;
  if (p.dense) {
    p.t1 = t + 1;
    p.n = c0;
    p.w = w0;
    return p;
  }
  const int sc = seg_scan_incl<WAVE>(c);
  const int sw = seg_scan_incl<WAVE>(wd);
  const int bytes = 4 * sc + 8 * sw + SEG_TILE_GT * (lane + 1);
  const u64 bad = __ballot(t + lane >= nt || bytes > SEG_BUDGET);
  const u64 keep = bad ? (1ull << (__ffsll((long long)bad) - 1)) - 1 : ~0ull;  // tile 0 fits
  const u64 ne = __ballot(c > 0) & keep;
  const int last = 63 - __clzll((long long)ne);
  p.t1 = t + last + 1;
  p.n = __builtin_amdgcn_readfirstlane(__shfl(sc, last));
  p.w = __builtin_amdgcn_readfirstlane(__shfl(sw, last));
  return p;
}

// Stages compressed step [t, t1) (layout above) from the row's compressed
// bitmap: pw = the step's nw packed nonzero words, mk = its tiles' group
// records (16 per tile: nonzero-word mask | entry count << 32).  The words
// are copied, coalesced, to s_r (the accumulator region's first nw words:
// nw <= n), with every load of the copy and of the records issued before
// the first is used; then thread l takes the window's 32-word group g = l
// (then l + WG): the 16 groups of a tile sit on 16 consecutive lanes, whose
// scans of (nonzero words, entries), added to the tile's own prefixes (lane
// j of every wave computes tile j's from the tile counts), give each group's
// first slot and entry prefix, and the thread stages its group's words
// eight per LDS round trip.  cwd_off is s_cwd's byte offset in the region.
// Two barriers; s_r is left dirty.  The copied words and records are read
// once: non-temporal loads keep them from evicting B's rows from L2.
template <int WG>
__device__ __forceinline__ void stage_packed(const u32* __restrict__ pw,
                                             const u64* __restrict__ mk, const i32* s_tc,
                                             const i32* tc, int nt, int t, int t1, int nw,
                                             u32* s_r, u64* s_cwd, u64* s_gt, u32 cwd_off) {
  static_assert(SEG_MAXT * SEG_GPT <= 2 * WG, "two groups per thread");
  constexpr int CU = 4;  // copy loads in flight per thread
  constexpr int WB = 8;  // group words per LDS round trip
  const int l = threadIdx.x;
  const int lane = l & (WAVE - 1);
  const int ngroups = (t1 - t) * SEG_GPT;
  const u64 m0 = __builtin_nontemporal_load(&mk[min(l, ngroups - 1)]);
  for (int q0 = 0; q0 < nw; q0 += CU * WG) {  // uniform; unconditional (clamped) loads
    u32 r[CU];
#pragma unroll
    for (int k = 0; k < CU; ++k) r[k] = __builtin_nontemporal_load(&pw[min(q0 + k * WG + l, nw - 1)]);
#pragma unroll
    for (int k = 0; k < CU; ++k)
      if (q0 + k * WG + l < nw) s_r[q0 + k * WG + l] = r[k];
  }
  const int v = lane < t1 - t ? tile_count(s_tc, tc, nt, t + lane) : 0;
  const int c = tile_entries(v), wd = tile_words(v);
  const int pc = seg_scan_incl<WAVE>(c) - c, pw0 = seg_scan_incl<WAVE>(wd) - wd;
  __syncthreads();
  for (int g0 = 0; g0 < ngroups; g0 += WG) {  // uniform
    const int g = g0 + l;
    const u64 rec = g < ngroups ? (g0 ? __builtin_nontemporal_load(&mk[g]) : m0) : 0ull;
    const u32 mask = (u32)rec;
    const int cnt = __popc(mask), ent = (int)(rec >> 32);
    const int tj = (g / SEG_GPT) & (WAVE - 1);  // the group's tile in the step (< 64)
    const int slot0 = __shfl(pw0, tj) + seg_scan_incl<SEG_GPT>(cnt) - cnt;
    u32 epre = (u32)(__shfl(pc, tj) + seg_scan_incl<SEG_GPT>(ent) - ent);
    if (g < ngroups) {
      s_gt[g] = ((u64)(cwd_off + 8u * (u32)slot0) << 32) | mask;
      u32 mm = mask;
      for (int q0 = 0; q0 < cnt; q0 += WB) {
        u32 wv[WB];
#pragma unroll
        for (int k = 0; k < WB; ++k) wv[k] = s_r[slot0 + min(q0 + k, cnt - 1)];
#pragma unroll
        for (int k = 0; k < WB; ++k) {
          if (q0 + k < cnt) {
            const int j = __ffs(mm) - 1;
            mm &= mm - 1;
            s_cwd[slot0 + q0 + k] = ((u64)((u32)(g * 32 + j) << 16 | 4u * epre) << 32) | wv[k];
            epre += __popc(wv[k]);
          }
        }
      }
    }
  }
  __syncthreads();
}

// Writes a compressed step (accumulators acc[0..n) in rank order) at
// out..out+n-1 of C with coalesced stores: the values first; then the
// threads turn the accumulators of the staged words (slot l, l + WG, ...)
// into column offsets in place; these are written and the array re-zeroed.
// C is written once and never read here: non-temporal stores keep it from
// evicting B's rows from L2.  Such stores are not merged in L2, so the lanes
// are shifted to line boundaries (32 entries): every wave then writes whole
// 128-byte lines, and only a step's two end lines are partial.  Three
// barriers.
template <int WG>
__device__ __forceinline__ void emit_ranked(const u64* s_cwd, int nwords, int n, i64 out,
                                            i64 cbase, float* acc, i32* __restrict__ ccol,
                                            float* __restrict__ cval) {
  constexpr int U = 4;  // LDS reads issued before their stores
  const int l = threadIdx.x;
  const int q0 = l - (int)(out & 31);
  for (int qb = q0; qb < n; qb += U * WG) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = qb + u * WG;
      v[u] = q >= 0 && q < n ? acc[q] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = qb + u * WG;
      if (q >= 0 && q < n) __builtin_nontemporal_store(v[u], &cval[out + q]);
    }
  }
  __syncthreads();
  // Column offsets of the staged words' entries, two set bits per round
  // (a wave runs as many rounds as its densest word needs).
  i32* colx = reinterpret_cast<i32*>(acc);
  constexpr int UW = 4;  // staged words read per LDS round trip
  for (int s0 = l; s0 < nwords; s0 += UW * WG) {
    u64 ew[UW];
#pragma unroll
    for (int u = 0; u < UW; ++u) ew[u] = s0 + u * WG < nwords ? s_cwd[s0 + u * WG] : 0ull;
#pragma unroll
    for (int u = 0; u < UW; ++u) {
      const u64 e = ew[u];
      u32 bits = (u32)e;
      int q = (int)((e >> 34) & 0x3fff);
      const int x0 = (int)(e >> 48) * 32;
      while (bits) {
        const int b0 = __ffs(bits) - 1;
        bits &= bits - 1;
        colx[q] = x0 + b0;
        if (bits) {
          colx[q + 1] = x0 + (__ffs(bits) - 1);
          bits &= bits - 1;
        }
        q += 2;
      }
    }
  }
  __syncthreads();
  for (int qb = q0; qb < n; qb += U * WG) {
    i32 c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = qb + u * WG;
      c[u] = q >= 0 && q < n ? colx[q] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = qb + u * WG;
      if (q >= 0 && q < n) {
        __builtin_nontemporal_store((i32)(cbase + c[u]), &ccol[out + q]);
        colx[q] = 0;  // also 0.0f: the accumulators start the next step at zero
      }
    }
  }
  __syncthreads();
}

// Two workgroups of HWG threads per CU: 16 waves, four per SIMD (<= 128 VGPRs).
template <bool BATCH, class OffT, int HC, int ROWS>
__global__ __launch_bounds__(HWG, 4) void k_num_heavy_seg(
    RowSet rs, i32 nt, const i32* __restrict__ tilecnt, const u32* __restrict__ rowbits,
    const u64* __restrict__ rowmask, const i64* __restrict__ key, i64 th,
    const i64* __restrict__ arp, const i32* __restrict__ acol, const float* __restrict__ aval,
    const i64* __restrict__ brp, const uint2* __restrict__ bcv, const i64* __restrict__ bbase,
    const i32* __restrict__ slab, const i64* __restrict__ crp, i32* __restrict__ ccol,
    float* __restrict__ cval, i32* __restrict__ next) {
  constexpr int WG = HWG;
  static_assert(TILE_W / 32 == WG, "one word of a dense tile per thread");
  __shared__ __align__(16) unsigned char s_raw[HSEG_LDS];
  // chunk entries: WG offsets and WG values, or WG packed words (int32 offsets)
  __shared__ __align__(16) unsigned char s_ent[WG * (sizeof(OffT) + 4)];
  OffT* s_off = reinterpret_cast<OffT*>(s_ent);
  float* s_av = reinterpret_cast<float*>(s_ent + WG * sizeof(OffT));
  __shared__ i32 s_ex[WG + 1];
  __shared__ __align__(16) u16 s_mark[ROWS * WG];
  __shared__ i32 s_scr[WG / WAVE];
  __shared__ u64 s_cscr[WG / WAVE];
  __shared__ i32 s_tc[TC_LDS];
  float* s_acc = reinterpret_cast<float*>(s_raw);
  const ChunkLds<OffT> cl{s_off, s_av, s_ex, s_mark, s_cscr};
  const PairedEntries src{bcv};
  const i32 nlist = rs.count();
  const i32* __restrict__ list = rs.rows();
  const int l = threadIdx.x;
  const i64 row_words = (i64)nt * (TILE_W / 32);
  for (int d = l; d < TILE_W; d += WG) s_acc[d] = 0.f;
  for (int d = l; d < ROWS * WG; d += WG) s_mark[d] = 0;
  // s_acc[dz0, dz1) may hold staged words and group entries; the rest is 0.
  int dz0 = 0, dz1 = 0;
  __syncthreads();
  __shared__ i32 s_slot;
  for (i64 h = next_slot(next, &s_slot); h < nlist; h = next_slot(next, &s_slot)) {
    const i32 row = list[h];
    if (key[row] <= th) continue;
    const i32* tc = tilecnt + h * nt;
    for (int t = l; t < min(nt, TC_LDS); t += WG) s_tc[t] = tc[t];
    const i64 a0 = arp[row], a1 = arp[row + 1], nA = a1 - a0;
    const i64 bb = BATCH ? bbase[row] : 0;
    const u32* bits = rowbits + h * row_words;                 // packed nonzero words
    const u64* masks = rowmask + h * ((i64)nt * SEG_GPT);     // 16 group records per tile
    i64 wbase = 0;  // packed words of the row's tiles before the current step
    i64 out = crp[row];
    const bool inreg = nA <= (i64)HC * WG;
    const int nch = (int)((nA + WG - 1) / WG);
    bool ok[HC];
    float av[HC];
    OffT b0[HC];
    const i32* sp[HC];
    i32 cur[HC], nxt[HC];
    if (inreg) {
#pragma unroll
      for (int c = 0; c < HC; ++c) {
        const i64 e = a0 + (i64)c * WG + l;
        ok[c] = e < a1;
        av[c] = 0.f;
        b0[c] = 0;
        sp[c] = slab;
        cur[c] = nxt[c] = 0;
        if (ok[c]) {
          const i64 k = (i64)acol[e] + bb;
          av[c] = aval[e];
          b0[c] = (OffT)brp[k];
          sp[c] = slab + k * (nt + 1);
        }
      }
    }
    __syncthreads();  // s_tc
    SegStep P = seg_plan(s_tc, tc, nt, 0);
    if (P.t < nt) {
      if (inreg) {
#pragma unroll
        for (int c = 0; c < HC; ++c) nxt[c] = ok[c] ? sp[c][P.t1] : 0;
      }
    }
    while (P.t < nt) {
      // The next step: its extent, and its range ends for the register
      // entries (in flight while this step runs).
      const SegStep Q = seg_plan(s_tc, tc, nt, P.t1);
      i32 pre[HC];
#pragma unroll
      for (int c = 0; c < HC; ++c) pre[c] = 0;
      if (Q.t < nt) {
        if (inreg) {
#pragma unroll
          for (int c = 0; c < HC; ++c) pre[c] = ok[c] ? sp[c][Q.t1] : 0;
        }
      }
      const int t = P.t, t1 = P.t1, n = P.n;
      const i64 cbase = (i64)t * TILE_W;
      u64* s_cwd = reinterpret_cast<u64*>(s_raw + ((4 * n + 7) & ~7));
      u64* s_gt = s_cwd + P.w;
      u32 dword = 0;  // dense: this thread's word of the tile
      if (P.dense) {
        if (dz1 > dz0) {
          for (int d = dz0 + l; d < dz1; d += WG) s_acc[d] = 0.f;
          dz0 = dz1 = 0;
        }
        // This thread's word of the tile, from the compressed bitmap: the
        // tile's group masks (in s_ex, free between chunks) locate it among
        // the packed words.
        if (l < SEG_GPT) s_ex[l] = (i32)(u32)masks[(i64)t * SEG_GPT + l];
        __syncthreads();
        const int g = l >> 5, j = l & 31;
        int pre = 0;
        for (int q = 0; q < g; ++q) pre += __popc((u32)s_ex[q]);
        const u32 m = (u32)s_ex[g];
        dword = (m >> j) & 1u ? bits[wbase + pre + __popc(m & ((1u << j) - 1u))] : 0u;
      } else {
        stage_packed<WG>(bits + wbase, masks + (i64)t * SEG_GPT, s_tc, tc, nt, t, t1, P.w,
                         reinterpret_cast<u32*>(s_raw), s_cwd, s_gt, (u32)((4 * n + 7) & ~7));
        // Accumulators [0, n) start at zero: clear the copied words and the
        // earlier staged data they overlap (ordered before the products by
        // the chunk's scan barriers).
        for (int d = l; d < P.w; d += WG) s_acc[d] = 0.f;
        for (int d = max(dz0, P.w) + l; d < min(dz1, n); d += WG) s_acc[d] = 0.f;
        const int end = (int)((reinterpret_cast<unsigned char*>(s_gt + (t1 - t) * SEG_GPT) -
                               s_raw) / 4);
        dz0 = n;  // the output re-zeroes [0, n)
        dz1 = max(dz1, end);
      }
      // Batched accumulation (every LDS stage of the ROWS products is
      // issued before the next stage waits): dense steps index the
      // accumulators by column; compressed steps by rank, two LDS reads
      // away (group -> staged word).
      auto add_dense = [&](const auto& c, const auto& p, const auto& ok) {
        int x[ROWS];
        u32 old[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) x[r] = (int)(c[r] - cbase);  // in the tile for every lane
#pragma unroll
        for (int r = 0; r < ROWS; ++r) old[r] = __float_as_uint(s_acc[x[r]]);
        lds_add_f32_rows<ROWS>(s_acc, x, p, ok, old);
      };
      // Byte offset of the step's group table, less that of the step's
      // first column's group (so that column c's group entry is at
      // gt_base + 8 * (c >> 10); u32 arithmetic wraps).
      const u32 gt_base = (u32)((4 * n + 7) & ~7) + 8u * (u32)P.w - 8u * (u32)(cbase >> 10);
      auto add_rank = [&](const auto& c, const auto& p, const auto& ok) {
        // In halves of H rows (register pressure).  The step's tiles start
        // at a multiple of 1024 columns, so a column's bits below 10 place
        // it in its group's words and its word without subtracting cbase;
        // bfe takes its field width from the low 5 bits of the operand.
        constexpr int H = ROWS / 2;
#pragma unroll
        for (int h0 = 0; h0 < ROWS; h0 += H) {
          u32 at[H], old[H];
          u64 gg[H], e[H];
          float ph[H];
          bool okh[H];
#pragma unroll
          for (int r = 0; r < H; ++r) {
            okh[r] = ok[h0 + r];
            ph[r] = p[h0 + r];
          }
#pragma unroll
          for (int r = 0; r < H; ++r)  // first staged word's offset << 32 | nonzero-word mask
            gg[r] = *reinterpret_cast<const u64*>(s_raw + (gt_base + 8u * ((u32)c[h0 + r] >> 10)));
#pragma unroll
          for (int r = 0; r < H; ++r) {
            const u32 below = __builtin_amdgcn_ubfe((u32)gg[r], 0u, (u32)c[h0 + r] >> 5);
            // word index << 48 | first accumulator's offset << 32 | bits
            e[r] = *reinterpret_cast<const u64*>(s_raw + ((u32)(gg[r] >> 32) + 8u * __popc(below)));
          }
#pragma unroll
          for (int r = 0; r < H; ++r) {
            const u32 below = __builtin_amdgcn_ubfe((u32)e[r], 0u, (u32)c[h0 + r]);
            at[r] = ((u32)(e[r] >> 32) & 0xffffu) + 4u * __popc(below);
          }
#pragma unroll
          for (int r = 0; r < H; ++r) old[r] = *reinterpret_cast<const u32*>(s_raw + at[r]);
          lds_add_f32_rows_at<H>(s_raw, at, ph, okh, old);
        }
      };
      if (inreg) {
        for (int c = 0; c < nch; ++c) {
          // c is uniform; the unrolled register arrays are selected by it
          int len = 0;
          OffT start = 0;
          float a = 0.f;
#pragma unroll
          for (int cc = 0; cc < HC; ++cc)
            if (cc == c && ok[cc]) {
              len = nxt[cc] - cur[cc];
              start = b0[cc] + (OffT)cur[cc];
              a = av[cc];
            }
          const int nvalid = (int)min((i64)WG, nA - (i64)c * WG);
          if (P.dense)
            chunk_products<WG, ROWS, true>(len, start, a, nvalid, src, cl, add_dense);
          else
            chunk_products<WG, ROWS, true>(len, start, a, nvalid, src, cl, add_rank);
        }
#pragma unroll
        for (int c = 0; c < HC; ++c) {
          cur[c] = nxt[c];
          nxt[c] = pre[c];
        }
      } else {
        for (i64 e0 = a0; e0 < a1; e0 += WG) {
          const i64 e = e0 + l;
          int len = 0;
          OffT start = 0;
          float a = 0.f;
          if (e < a1) {
            const i64 k = (i64)acol[e] + bb;
            const i32* sk = slab + k * (nt + 1);
            const i32 s0 = sk[t], s1 = sk[t1];
            start = (OffT)(brp[k] + s0);
            len = s1 - s0;
            a = aval[e];
          }
          const int nvalid = (int)min((i64)WG, a1 - e0);
          if (P.dense)
            chunk_products<WG, ROWS, true>(len, start, a, nvalid, src, cl, add_dense);
          else
            chunk_products<WG, ROWS, true>(len, start, a, nvalid, src, cl, add_rank);
        }
      }
      // Output.  Dense tile: thread l writes the entries of its own word
      // (words of a dense tile carry similar counts) at the word's prefix.
      if (P.dense) {
        const int pcb = (int)__popc(dword);
        int tot;
        const int inc = block_scan_incl<WG>(pcb, s_scr, &tot);
        i64 q = out + (inc - pcb);
        u32 r = dword;
        while (r) {
          const int x = l * 32 + (__ffs(r) - 1);
          r &= r - 1;
          // Ordinary stores: each lane writes its own short run, and L2
          // merges the runs into whole lines (non-temporal stores would
          // reach memory piecemeal).
          ccol[q] = (i32)(cbase + x);
          cval[q] = s_acc[x];
          s_acc[x] = 0.f;
          ++q;
        }
        __syncthreads();  // s_scr is reused by the next scan
      } else {
        emit_ranked<WG>(s_cwd, P.w, n, out, cbase, s_acc, ccol, cval);
      }
      out += n;
      wbase += P.w;
      P = Q;
    }
    __syncthreads();  // s_tc is reloaded for the next row
  }
}

// ---------------------------------------------------------------------------
// Exclusive scan of n int64 values into out[0..n] (out[n] = total).
constexpr int SCAN_TILE = 4096;

__global__ __launch_bounds__(1024) void k_scan_reduce(const i64* __restrict__ in, i64 n,
                                                      i64* __restrict__ bsum) {
  __shared__ i64 s_scr[16];
  const i64 base = (i64)blockIdx.x * SCAN_TILE;
  i64 s = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const i64 i = base + threadIdx.x + j * 1024;
    if (i < n) s += in[i];
  }
  s = block_reduce_sum<1024>(s, s_scr);
  if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void k_scan_top(i64* __restrict__ bsum, i64 nb,
                                                   i64* __restrict__ out_total) {
  __shared__ i64 s_scr[16];
  i64 carry = 0;
  for (i64 b0 = 0; b0 < nb; b0 += 1024) {
    const i64 i = b0 + threadIdx.x;
    const i64 v = i < nb ? bsum[i] : 0;
    i64 tot;
    const i64 inc = block_scan_incl<1024>(v, s_scr, &tot);
    if (i < nb) bsum[i] = carry + inc - v;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(1024) void k_scan_apply(const i64* __restrict__ in, i64 n,
                                                     const i64* __restrict__ bsum,
                                                     i64* __restrict__ out) {
  __shared__ i64 s_scr[16];
  const i64 base = (i64)blockIdx.x * SCAN_TILE + (i64)threadIdx.x * 4;
  i64 v[4];
  i64 s = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    v[j] = (base + j < n) ? in[base + j] : 0;
    s += v[j];
  }
  i64 tot;
  const i64 inc = block_scan_incl<1024>(s, s_scr, &tot);
  i64 run = bsum[blockIdx.x] + inc - s;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
}

i64 scan_partials(i64 n) { return std::max<i64>(1, (n + SCAN_TILE - 1) / SCAN_TILE); }

void exclusive_scan_i64(const i64* in, i64* out, i64 n, hipStream_t s, i64* partials) {
  const i64 nb = scan_partials(n);
  DBuf<i64> own;
  if (!partials) {
    own.alloc(nb);
    partials = own.p;
  }
  {
    ProfScope ps("scan_reduce", s);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(1024), 0, s, in, n, partials);
  }
  {
    ProfScope ps("scan_top", s);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, partials, nb, out + n);
  }
  if (n > 0) {
    ProfScope ps("scan_apply", s);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(1024), 0, s, in, n, partials, out);
  }
  RT_HIP(hipGetLastError());
  // Partials of our own are released to the pool here; the pool only hands
  // them out again after this synchronisation.  Caller partials need none
  // (and a captured stream allows none).
  if (own.p) RT_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Segment fill: out[r] = values[b] for the segment b containing row r.
__global__ __launch_bounds__(256) void k_fill_segments(i32 count, const i64* __restrict__ start,
                                                       const i64* __restrict__ values,
                                                       i64* __restrict__ out, i64 total) {
  const i64 r = (i64)blockIdx.x * 256 + threadIdx.x;
  if (r >= total) return;
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= r) lo = mid; else hi = mid - 1;
  }
  out[r] = values[lo];
}

void fill_segments(i32 count, const i64* seg_rows, const i64* values, i64* out_dev,
                   i64 total_rows, hipStream_t s) {
  std::vector<i64> st(count + 1, 0);
  for (int i = 0; i < count; ++i) st[i + 1] = st[i] + seg_rows[i];
  DBuf<i64> dst(count + 1), dv(count);
  RT_HIP(hipMemcpyAsync(dst.p, st.data(), (count + 1) * 8, hipMemcpyHostToDevice, s));
  RT_HIP(hipMemcpyAsync(dv.p, values, count * 8, hipMemcpyHostToDevice, s));
  if (total_rows > 0) {
    ProfScope ps("fill_segments", s);
    hipLaunchKernelGGL(k_fill_segments, dim3((unsigned)((total_rows + 255) / 256)), dim3(256), 0,
                       s, count, dst.p, dv.p, out_dev, total_rows);
  }
  RT_HIP(hipGetLastError());
  RT_HIP(hipStreamSynchronize(s));
}

// Structural validation of a device CSR: row_ptr runs from 0 to nnz and is
// monotone; columns lie in [0, limit) (limit = row_limit[r] when given) and
// increase strictly within each row.  flag[0] |= 1 on any violation.
__global__ __launch_bounds__(256) void k_validate(i64 rows, i64 nnz, const i64* __restrict__ rp,
                                                  const i32* __restrict__ col, i64 col_limit,
                                                  const i64* __restrict__ row_limit,
                                                  int* __restrict__ flag) {
  constexpr int G = 16;
  const i64 r = ((i64)blockIdx.x * 256 + threadIdx.x) / G;
  const int l = threadIdx.x & (G - 1);
  if (r >= rows) return;
  const i64 a = rp[r], b = rp[r + 1];
  bool bad = a > b || a < 0 || b > nnz || (r == 0 && a != 0) || (r == rows - 1 && b != nnz);
  if (!bad) {
    const i64 lim = row_limit ? row_limit[r] : col_limit;
    for (i64 e = a + l; e < b; e += G) {
      const i32 c = col[e];
      if (c < 0 || c >= lim || (e > a && col[e - 1] >= c)) bad = true;
    }
  }
  if (bad) atomicOr(flag, 1);
}

void validate_csr_dev(const Csr& M, i64 col_limit, const i64* row_limit, hipStream_t s,
                      const char* what, int* flag) {
  DBuf<int> own;
  if (!flag) {
    own.alloc(1);
    flag = own.p;
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, flag, (i64)1, 0);
  if (M.rows > 0) {
    ProfScope ps("validate", s);
    hipLaunchKernelGGL(k_validate, dim3((unsigned)((M.rows * 16 + 255) / 256)), dim3(256), 0, s,
                       M.rows, M.nnz, M.rp, M.col, col_limit, row_limit, flag);
    RT_HIP(hipGetLastError());
  } else {
    RT_CHECK(M.rows == 0, RT_E_INVALID, "negative rows");
  }
  const int f = read_scalar(flag, s);
  RT_CHECK(f == 0, RT_E_INVALID,
           std::string(what) + ": malformed CSR (row_ptr must run monotonically from 0 to nnz; "
           "columns must be in range, sorted and unique within each row)");
}

// Fingerprint of a CSR result, to compare two runs of one product: wrapping
// sums of a 64-bit mix of (index, row pointer) and of (position, column)
// -- equal exactly when the structures are equal, up to hash collisions --
// and two double sums of the values, sum |v| and sum v * w(position) with a
// pseudo-random weight in [-1/2, 1/2), which differ between runs only by
// rounding of the FP32 accumulation order.
__device__ __forceinline__ u64 mix64(u64 x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

__global__ __launch_bounds__(256) void k_fingerprint(i64 rows, i64 nnz, const i64* __restrict__ rp,
                                                     const i32* __restrict__ col,
                                                     const float* __restrict__ val,
                                                     unsigned long long* __restrict__ hsum,
                                                     double* __restrict__ dsum) {
  __shared__ u64 s_u[4];
  __shared__ double s_d[4];
  u64 hr = 0, hc = 0;
  double va = 0, vd = 0;
  const i64 stride = (i64)gridDim.x * 256;
  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i <= rows; i += stride)
    hr += mix64(((u64)i << 40) ^ (u64)rp[i]);
  for (i64 e = (i64)blockIdx.x * 256 + threadIdx.x; e < nnz; e += stride) {
    const u64 m = mix64((u64)e);
    hc += mix64(m ^ (u64)(u32)col[e]);
    const double v = val[e];
    va += v < 0 ? -v : v;
    vd += v * ((double)(m >> 11) * 0x1p-53 - 0.5);
  }
  hr = block_reduce_sum<256>(hr, s_u);
  hc = block_reduce_sum<256>(hc, s_u);
  va = block_reduce_sum<256>(va, s_d);
  vd = block_reduce_sum<256>(vd, s_d);
  if (threadIdx.x == 0) {
    atomicAdd(&hsum[0], (unsigned long long)hr);
    atomicAdd(&hsum[1], (unsigned long long)hc);
    atomicAdd(&dsum[0], va);
    atomicAdd(&dsum[1], vd);
  }
}

Fingerprint fingerprint(const CsrOut& C, hipStream_t s) {
  DBuf<unsigned long long> h(2);
  DBuf<double> d(2);
  RT_HIP(hipMemsetAsync(h.p, 0, 2 * sizeof(unsigned long long), s));
  RT_HIP(hipMemsetAsync(d.p, 0, 2 * sizeof(double), s));
  const i64 n = std::max<i64>(C.rows + 1, C.nnz);
  const unsigned grid = (unsigned)std::max<i64>(1, std::min<i64>((n + 255) / 256,
                                                                  (i64)device().cus * 8));
  hipLaunchKernelGGL(k_fingerprint, dim3(grid), dim3(256), 0, s, C.rows, C.nnz, C.rp, C.col, C.val,
                     h.p, d.p);
  RT_HIP(hipGetLastError());
  Fingerprint f;
  unsigned long long hh[2];
  RT_HIP(hipMemcpyAsync(hh, h.p, sizeof(hh), hipMemcpyDeviceToHost, s));
  RT_HIP(hipMemcpyAsync(f.vsum, d.p, sizeof(f.vsum), hipMemcpyDeviceToHost, s));
  RT_HIP(hipStreamSynchronize(s));
  f.hrp = hh[0];
  f.hcol = hh[1];
  f.nnz = C.nnz;
  return f;
}

// ---------------------------------------------------------------------------
// Host pipeline pieces (shared by the plain and learned methods).
namespace {

template <int G, int T, int WG, int MODE, bool BATCH>
void launch_hash(const char* name, HashArgs a, const i32* list, i32 n, hipStream_t s) {
  if (n <= 0) return;
  constexpr int R = WG / G;
  a.list = list;
  a.nlist = n;
  const dim3 grid((unsigned)((n + R - 1) / R));
  ProfScope ps(name, s);
  hipLaunchKernelGGL((k_hash<G, T, WG, MODE, BATCH>), grid, dim3(WG), 0, s, a);
  RT_HIP(hipGetLastError());
}

// The five hash bins, each in the workgroup shape the tuning picks:
// lanes per row G, table slots T, workgroup WG (G == WG: one row per
// workgroup).  Default shapes, then the alternates (see hash_group).
template <int MODE, bool BATCH, class L>
void for_hash_bins(const int shape[5], L&& launch) {
  if (shape[0]) launch.template go<16, 32, 256, MODE, BATCH>(0);
  else launch.template go<8, 32, 256, MODE, BATCH>(0);
  if (shape[1]) launch.template go<32, 128, 256, MODE, BATCH>(1);
  else launch.template go<16, 128, 256, MODE, BATCH>(1);
  if (shape[2]) launch.template go<32, 512, 256, MODE, BATCH>(2);
  else launch.template go<64, 512, 256, MODE, BATCH>(2);
  if (shape[3]) launch.template go<256, 2048, 256, MODE, BATCH>(3);
  else launch.template go<128, 2048, 128, MODE, BATCH>(3);
  if (shape[4] == 2) launch.template go<256, 4096, 256, MODE, BATCH>(4);
  else if (shape[4]) launch.template go<1024, 8192, 1024, MODE, BATCH>(4);
  else launch.template go<512, 8192, 512, MODE, BATCH>(4);
}

// Host-listed bins: bin b's rows are list[off[b] .. off[b] + cnt[b]).
struct ListedBins {
  const HashArgs& a;
  const i32* list;
  const int* cnt;
  const int* off;
  const char* const* names;
  hipStream_t s;
  template <int G, int T, int WG, int M, bool BT>
  void go(int b) {
    launch_hash<G, T, WG, M, BT>(names[T == 4096 ? 5 : b], a, list + off[b], cnt[b], s);
  }
};

template <int MODE, bool BATCH>
void launch_hash_bins(const HashArgs& a, const i32* list, const int* cnt, const int* off,
                      const int shape[5], hipStream_t s) {
  static const char* names[3][6] = {
      {"sym_hash_t32", "sym_hash_t128", "sym_hash_t512", "sym_hash_t2048", "sym_hash_t8192",
       "sym_hash_t4096"},
      {"num_hash_t32", "num_hash_t128", "num_hash_t512", "num_hash_t2048", "num_hash_t8192",
       "num_hash_t4096"},
      {"learn_hash_t32", "learn_hash_t128", "learn_hash_t512", "learn_hash_t2048",
       "learn_hash_t8192", "learn_hash_t4096"}};
  ListedBins l{a, list, cnt, off, names[MODE], s};
  for_hash_bins<MODE, BATCH>(shape, l);
}

// int32 B offsets (and eight products per lane in flight) when B's entries
// are indexable by int32.  colpad: B's columns with a spare entry past the
// end (k_copy_cols), or nullptr; with it the block walk gathers two columns
// per lane.
template <bool BATCH>
void launch_sym_bitmap(unsigned grid, const RowSet& rs, const Csr& A, const Csr& B,
                       const i64* bbase, const i32* slab, i32 nt, i32* tilecnt, i64* nnz,
                       u32* rowbits, u64* rowmask, i32* next, hipStream_t s,
                       const i32* colpad = nullptr) {
  if (colpad && B.nnz < ((i64)1 << 31))
    hipLaunchKernelGGL((k_sym_bitmap<BATCH, i32, 8, true, SYM_PAIR_WG>), dim3(grid * 2),
                       dim3(SYM_PAIR_WG), 0, s, rs,
                       A.rp, A.col, B.rp, colpad, bbase, slab, B.cols, nt, tilecnt, nnz, rowbits,
                       rowmask, next);
  else if (B.nnz < ((i64)1 << 31))
    hipLaunchKernelGGL((k_sym_bitmap<BATCH, i32, 8, false>), dim3(grid), dim3(HWG), 0, s, rs,
                       A.rp, A.col, B.rp, B.col, bbase, slab, B.cols, nt, tilecnt, nnz, rowbits,
                       rowmask, next);
  else
    hipLaunchKernelGGL((k_sym_bitmap<BATCH, i64, 4, false>), dim3(grid), dim3(HWG), 0, s, rs,
                       A.rp, A.col, B.rp, B.col, bbase, slab, B.cols, nt, tilecnt, nnz, rowbits,
                       rowmask, next);
}

// B's columns with one spare entry, for the two-column block walk.
void copy_cols(const Csr& B, DBuf<i32>& out, hipStream_t s) {
  out.alloc((size_t)B.nnz + 1);
  const unsigned g = (unsigned)std::min<i64>((B.nnz + 256) / 256, (i64)device().cus * 16);
  hipLaunchKernelGGL(k_copy_cols, dim3(g), dim3(256), 0, s, B.nnz, B.col, out.p, nullptr, (i64)0);
  RT_HIP(hipGetLastError());
}

HashArgs make_args(const Csr& A, const Csr& B, const i64* bbase) {
  HashArgs a{};
  a.ticket = nullptr;
  a.arp = A.rp;
  a.acol = A.col;
  a.aval = A.val;
  a.brp = B.rp;
  a.bcol = B.col;
  a.bval = B.val;
  a.bbase = bbase;
  a.bcols = B.cols;
  return a;
}

template <bool BATCH>
void symbolic_impl(const Csr& A, const Csr& B, const i64* bbase, const i64* key, i64* nnz,
                   const Tuning& tu, hipStream_t s, HeavyRows& heavy) {
  const i64 m = A.rows;
  heavy.list.alloc(std::max<i64>(m, 1));
  int cnt[NBIN], off[NBIN + 1];
  bin_rows(key, m, B.cols, tu.sym, heavy.list.p, cnt, off, s);
  HashArgs a = make_args(A, B, bbase);
  a.nnz = nnz;
  launch_hash_bins<MODE_SYM, BATCH>(a, heavy.list.p, cnt, off, tu.shape, s);
  heavy.off = off[5];
  heavy.n = cnt[5];
  heavy.nt = (i32)std::max<i64>(1, (B.cols + TILE_W - 1) / TILE_W);
  if (heavy.n > 0) {
    heavy.tilecnt.alloc((size_t)heavy.n * heavy.nt);
    // Keep the long rows' column bitmaps for the segmented numeric kernel
    // when they fit comfortably in free device memory.
    // (Worst case: every word nonzero.)
    const size_t bm_bytes = (size_t)heavy.n * heavy.nt * (TILE_W / 8);
    const size_t mask_bytes = (size_t)heavy.n * heavy.nt * (TILE_W / 32 / 32) * sizeof(u64);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    if (bm_bytes + mask_bytes <= free_b / 3 + pool_cached_bytes() / 3) {
      heavy.rowbits.alloc(bm_bytes / sizeof(u32));
      heavy.rowmask.alloc(mask_bytes / sizeof(u64));
    }
    // Tile boundaries of every B row (kept for the numeric pass).
    heavy.slab.alloc((size_t)B.rows * (heavy.nt + 1));
    if (B.rows > 0) {
      const i64 nthreads = B.rows * (i64)(heavy.nt + 1);
      ProfScope ps("slab_ptrs", s);
      hipLaunchKernelGGL(k_slab_ptrs, dim3(stride_grid(nthreads)), dim3(256), 0, s, B.rows,
                         heavy.nt, B.rp, B.col, heavy.slab.p, nullptr, (i64)0);
      RT_HIP(hipGetLastError());
    }
    // Rows that gather the same long B rows run together (their L2 reuse).
    // Only where B (8 bytes an entry) outgrows the eight 4 MB L2s together:
    // below that its rows stay cached in any order and the ordering's few
    // launches cost more than they save.  Batches skip it: their B rows
    // belong to different products.
    if (!BATCH && heavy.n >= 1024 && B.nnz * 8 > ((i64)32 << 20))
      order_long_rows(A, B, heavy.list.p + heavy.off, heavy.n, s);
    const unsigned grid = (unsigned)std::min<i64>(heavy.n, (i64)device().cus * 2);
    heavy.next.alloc(1);
    zero_slot(heavy.next.p, s);
    if (B.nnz < ((i64)1 << 31)) {
      ProfScope ps("copy_cols", s);
      copy_cols(B, heavy.colpad, s);
    }
    ProfScope ps("sym_bitmap", s);
    const RowSet rs{heavy.list.p + heavy.off, heavy.n, nullptr, nullptr, 0};
    launch_sym_bitmap<BATCH>(grid, rs, A, B, bbase, heavy.slab.p, heavy.nt, heavy.tilecnt.p, nnz,
                             heavy.rowbits.p, heavy.rowmask.p, heavy.next.p, s, heavy.colpad.p);
    RT_HIP(hipGetLastError());
  }
}

template <bool BATCH>
void numeric_impl(const Csr& A, const Csr& B, const i64* bbase, const i64* key,
                  const HeavyRows& heavy, CsrOut& C, const Tuning& tu, hipStream_t s) {
  const i64 m = A.rows;
  DBuf<i32> list(std::max<i64>(m, 1));
  int cnt[NBIN], off[NBIN + 1];
  bin_rows(key, m, LLONG_MAX, tu.num, list.p, cnt, off, s);
  if (cnt[5] > 0) {
    RT_CHECK(heavy.n > 0 && heavy.slab.p, RT_E_INVALID,
             "tuning: numeric tile rows without symbolic tile counts");
    const i32* const slab = heavy.slab.p;
    const unsigned grid = (unsigned)std::min<i64>(heavy.n, (i64)device().cus * 2);
    DBuf<i32> next(1);
    zero_slot(next.p, s);
    DBuf<uint2> pairs;
    if (heavy.rowbits.p && B.nnz > 0) {
      pairs.alloc((size_t)B.nnz);
      ProfScope ps("pair_entries", s);
      const unsigned pg = (unsigned)std::min<i64>((B.nnz + 255) / 256, (i64)device().cus * 16);
      hipLaunchKernelGGL(k_pair_entries, dim3(pg), dim3(256), 0, s, B.nnz, B.col, B.val, pairs.p);
      RT_HIP(hipGetLastError());
    }
    {
      ProfScope ps("num_heavy", s);
      const RowSet rs{heavy.list.p + heavy.off, heavy.n, nullptr, nullptr, 0};
      // int32 B offsets (and eight products per lane in flight) when B's
      // entries are indexable by int32.
      if (heavy.rowbits.p && B.nnz < ((i64)1 << 31))
        hipLaunchKernelGGL((k_num_heavy_seg<BATCH, i32, 1, 8>), dim3(grid), dim3(HWG), 0, s, rs,
                           heavy.nt, heavy.tilecnt.p, heavy.rowbits.p, heavy.rowmask.p, key,
                           tu.num[4], A.rp, A.col,
                           A.val, B.rp, pairs.p, bbase, slab, C.rp, C.col, C.val, next.p);
      else if (heavy.rowbits.p)
        hipLaunchKernelGGL((k_num_heavy_seg<BATCH, i64, 2, 4>), dim3(grid), dim3(HWG), 0, s, rs,
                           heavy.nt, heavy.tilecnt.p, heavy.rowbits.p, heavy.rowmask.p, key,
                           tu.num[4], A.rp, A.col,
                           A.val, B.rp, pairs.p, bbase, slab, C.rp, C.col, C.val, next.p);
      else
        hipLaunchKernelGGL((k_num_heavy<BATCH, false>), dim3(grid), dim3(HWG), 0, s, rs,
                           heavy.nt, heavy.tilecnt.p, key, tu.num[4], A.rp, A.col, A.val, B.rp,
                           B.col, B.val, bbase, slab, C.rp, C.col, C.val, nullptr, nullptr,
                           nullptr, next.p);
      RT_HIP(hipGetLastError());
    }
    HashArgs a = make_args(A, B, bbase);
    a.crp = C.rp;
    a.ccol = C.col;
    a.cval = C.val;
    launch_hash_bins<MODE_NUM, BATCH>(a, list.p, cnt, off, tu.shape, s);
    RT_HIP(hipStreamSynchronize(s));  // slab is released at scope exit
  } else {
    HashArgs a = make_args(A, B, bbase);
    a.crp = C.rp;
    a.ccol = C.col;
    a.cval = C.val;
    launch_hash_bins<MODE_NUM, BATCH>(a, list.p, cnt, off, tu.shape, s);
    RT_HIP(hipStreamSynchronize(s));
  }
}

template <bool BATCH>
void learned_impl(const Csr& A, const Csr& B, const i64* bbase, const i64* key, const i64* cap,
                  const i64* tmp_off, i32* tcol, float* tval, i64* nnz, i32* ovf,
                  const Tuning& tu, hipStream_t s) {
  const i64 m = A.rows;
  DBuf<i32> list(std::max<i64>(m, 1));
  int cnt[NBIN], off[NBIN + 1];
  bin_rows(key, m, LLONG_MAX, tu.num, list.p, cnt, off, s);
  RT_CHECK(cnt[5] == 0, RT_E_INVALID, "learned rows must fit the LDS hash bins");
  HashArgs a = make_args(A, B, bbase);
  a.crp = tmp_off;
  a.ccol = tcol;
  a.cval = tval;
  a.cap = cap;
  a.ovf = ovf;
  a.nnz = nnz;
  launch_hash_bins<MODE_LEARN, BATCH>(a, list.p, cnt, off, tu.shape, s);
  RT_HIP(hipStreamSynchronize(s));
}

}  // namespace

void bin_rows(const i64* x, i64 n, i64 cap, const i64 th_in[5], i32* list, int cnt[6],
              int off[7], hipStream_t s, i32* counters) {
  Thr th;
  for (int b = 0; b < 5; ++b) th.t[b] = th_in[b];
  if (n <= 0) {
    for (int b = 0; b < NBIN; ++b) cnt[b] = 0;
    for (int b = 0; b <= NBIN; ++b) off[b] = 0;
    return;
  }
  DBuf<i32> own;
  if (!counters) {
    own.alloc((size_t)bin_scratch_ints());
    counters = own.p;
  }
  // counters: [cnt NBIN | off NBIN + 1 | per-workgroup slots]
  i32* wgc = counters + 2 * NBIN + 1;
  const unsigned grid = stride_grid(n);
  {
    ProfScope ps("bin_count", s);
    hipLaunchKernelGGL(k_bin_count, dim3(grid), dim3(256), 0, s, n, x, cap, th, wgc);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(256), 0, s, (i32)grid, wgc, counters,
                       counters + NBIN, nullptr);
  }
  RT_HIP(hipGetLastError());
  {
    ProfScope ps("bin_fill", s);
    hipLaunchKernelGGL(k_bin_fill, dim3(grid), dim3(256), 0, s, n, x, cap, th, wgc, list);
  }
  RT_HIP(hipGetLastError());
  int host[2 * NBIN + 1];
  RT_HIP(hipMemcpyAsync(host, counters, sizeof(host), hipMemcpyDeviceToHost, s));
  RT_HIP(hipStreamSynchronize(s));
  for (int b = 0; b < NBIN; ++b) cnt[b] = host[b];
  for (int b = 0; b <= NBIN; ++b) off[b] = host[NBIN + b];
}

void row_products(const Csr& A, const Csr& B, const i64* bbase, i64* prod, hipStream_t s,
                  RunStats* st) {
  const i64 m = A.rows;
  DBuf<unsigned long long> tot(row_totals_len());
  RT_HIP(hipMemsetAsync(tot.p, 0, 2 * sizeof(unsigned long long), s));
  if (m > 0) {
    ProfScope ps("row_products", s);
    DBuf<i32> longq(row_products_queue(A.nnz));
    if (bbase)
      launch_row_products<true>(m, A.rp, A.col, B.rp, bbase, prod, tot.p, longq.p, s);
    else
      launch_row_products<false>(m, A.rp, A.col, B.rp, nullptr, prod, tot.p, longq.p, s);
    RT_HIP(hipGetLastError());
    RT_HIP(hipStreamSynchronize(s));  // longq is released at scope exit
  }
  RT_HIP(hipGetLastError());
  unsigned long long h[2];
  RT_HIP(hipMemcpyAsync(h, tot.p, sizeof(h), hipMemcpyDeviceToHost, s));
  RT_HIP(hipStreamSynchronize(s));
  if (st) {
    st->products = (i64)h[0];
    st->max_row_products = (i64)h[1];
  }
}

void symbolic_pass(const Csr& A, const Csr& B, const i64* bbase, const i64* key, i64* nnz,
                   const Tuning& tu, hipStream_t s, HeavyRows& heavy) {
  if (bbase) symbolic_impl<true>(A, B, bbase, key, nnz, tu, s, heavy);
  else symbolic_impl<false>(A, B, nullptr, key, nnz, tu, s, heavy);
}

void numeric_pass(const Csr& A, const Csr& B, const i64* bbase, const i64* key,
                  const HeavyRows& heavy, CsrOut& C, const Tuning& tu, hipStream_t s) {
  if (bbase) numeric_impl<true>(A, B, bbase, key, heavy, C, tu, s);
  else numeric_impl<false>(A, B, nullptr, key, heavy, C, tu, s);
}

void learned_numeric_pass(const Csr& A, const Csr& B, const i64* bbase, const i64* key,
                          const i64* cap, const i64* tmp_off, i32* tcol, float* tval, i64* nnz,
                          i32* ovf, const Tuning& tu, hipStream_t s) {
  if (bbase) learned_impl<true>(A, B, bbase, key, cap, tmp_off, tcol, tval, nnz, ovf, tu, s);
  else learned_impl<false>(A, B, nullptr, key, cap, tmp_off, tcol, tval, nnz, ovf, tu, s);
}

void learned_heavy_pass(const Csr& A, const Csr& B, const i64* bbase, const i64* key,
                        const i64* cap, const i64* tmp_off, i32* tcol, float* tval, i64* nnz,
                        i32* ovf, hipStream_t s) {
  const i64 m = A.rows;
  DBuf<i32> list(std::max<i64>(m, 1));
  int cnt[NBIN], off[NBIN + 1];
  const i64 zero[5] = {0, 0, 0, 0, 0};  // every row with key > 0 lands in bin 5
  bin_rows(key, m, LLONG_MAX, zero, list.p, cnt, off, s);
  if (cnt[5] == 0) return;
  const i32 nt = (i32)std::max<i64>(1, (B.cols + TILE_W - 1) / TILE_W);
  DBuf<i32> slab((size_t)B.rows * (nt + 1));
  {
    const i64 nthreads = B.rows * (i64)(nt + 1);
    ProfScope ps("slab_ptrs", s);
    hipLaunchKernelGGL(k_slab_ptrs, dim3(stride_grid(nthreads)), dim3(256), 0, s, B.rows, nt,
                       B.rp, B.col, slab.p, nullptr, (i64)0);
    RT_HIP(hipGetLastError());
  }
  const unsigned grid = (unsigned)std::min<i64>(cnt[5], (i64)device().cus * 2);
  DBuf<i32> next(1);
  zero_slot(next.p, s);
  {
    ProfScope ps("learn_heavy", s);
    if (bbase)
      hipLaunchKernelGGL((k_num_heavy<true, true>), dim3(grid), dim3(HWG), 0, s,
                         RowSet{list.p + off[5], cnt[5], nullptr, nullptr, 0}, nt, nullptr, key,
                         (i64)0, A.rp, A.col, A.val, B.rp, B.col, B.val, bbase, slab.p, tmp_off,
                         tcol, tval, cap, nnz, ovf, next.p);
    else
      hipLaunchKernelGGL((k_num_heavy<false, true>), dim3(grid), dim3(HWG), 0, s,
                         RowSet{list.p + off[5], cnt[5], nullptr, nullptr, 0}, nt, nullptr, key,
                         (i64)0, A.rp, A.col, A.val, B.rp, B.col, B.val, nullptr, slab.p, tmp_off,
                         tcol, tval, cap, nnz, ovf, next.p);
    RT_HIP(hipGetLastError());
  }
  RT_HIP(hipStreamSynchronize(s));  // list and slab are released at scope exit
}

void alloc_c(CsrOut& C, const i64* nnz, i64 m, i64 cols, hipStream_t s) {
  C.rows = m;
  C.cols = cols;
  C.rp = static_cast<i64*>(dmalloc((m + 1) * sizeof(i64)));
  exclusive_scan_i64(nnz, C.rp, m, s);
  C.nnz = read_scalar(C.rp + m, s);
  C.col = static_cast<i32*>(dmalloc(std::max<i64>(C.nnz, 1) * sizeof(i32)));
  C.val = static_cast<float*>(dmalloc(std::max<i64>(C.nnz, 1) * sizeof(float)));
}

void spgemm_plain(const Csr& A, const Csr& B, const Batch& batch, CsrOut& C,
                  const Tuning& tune, hipStream_t s, RunStats* st) {
  const i64 m = A.rows;
  RT_CHECK(m < (i64)INT32_MAX, RT_E_LIMIT, "A has too many rows for 32-bit row lists");
  DBuf<i64> prod(std::max<i64>(m, 1)), nnz(std::max<i64>(m, 1));
  row_products(A, B, batch.b_base, prod.p, s, st);
  RT_HIP(hipMemsetAsync(nnz.p, 0, std::max<i64>(m, 1) * sizeof(i64), s));
  HeavyRows heavy;
  symbolic_pass(A, B, batch.b_base, prod.p, nnz.p, tune, s, heavy);
  alloc_c(C, nnz.p, m, B.cols, s);
  numeric_pass(A, B, batch.b_base, nnz.p, heavy, C, tune, s);
}

}  // namespace rt

// ---------------------------------------------------------------------------
// Graph-captured exact method.
//
// The same kernels, launched so that a stage needs no host round trip: row
// bins are counted, offset and filled on the device, every row kernel is
// persistent (fixed grid, its row count read from device memory), clears
// are fill kernels, and all scratch is allocated before capture.  Stage A
// (products, binning, symbolic, scan) and stage B (binning by nnz, long-row
// and LDS-hash numeric) are each captured into a hipGraph and launched; the
// only host read between them is C's size, to allocate C.
namespace rt {
using namespace dev;

namespace {

unsigned fill_grid(i64 n) {
  return (unsigned)std::max<i64>(1, std::min<i64>((n + 255) / 256, (i64)device().cus * 8));
}

struct DevBins {
  DBuf<i32> list, cnt, off, cursor, ticket, wgc;
  explicit DevBins(i64 m)
      : list(std::max<i64>(m, 1)), cnt(NBIN), off(NBIN + 1), cursor(NBIN), ticket(NBIN),
        wgc((size_t)device().cus * 8 * NBIN) {}
};

void bin_rows_dev(const i64* x, i64 n, i64 cap, const i64 th_in[5], DevBins& b, hipStream_t s) {
  Thr th;
  for (int i = 0; i < 5; ++i) th.t[i] = th_in[i];
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, b.ticket.p, (i64)NBIN, 0);
  if (n > 0) {
    const unsigned grid = stride_grid(n);
    hipLaunchKernelGGL(k_bin_count, dim3(grid), dim3(256), 0, s, n, x, cap, th, b.wgc.p);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(256), 0, s, (i32)grid, b.wgc.p, b.cnt.p,
                       b.off.p, b.cursor.p);
    hipLaunchKernelGGL(k_bin_fill, dim3(grid), dim3(256), 0, s, n, x, cap, th, b.wgc.p, b.list.p);
  } else {
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(256), 0, s, 0, b.wgc.p, b.cnt.p, b.off.p,
                       b.cursor.p);
  }
  RT_HIP(hipGetLastError());
}

template <int G, int T, int WG, int MODE, bool BATCH>
void launch_hash_dev(HashArgs a, const DevBins& b, int bin, i64 max_rows, hipStream_t s) {
  constexpr int R = WG / G;
  a.list = b.list.p;
  a.nlist = 0;
  a.dcount = b.cnt.p;
  a.doff = b.off.p;
  a.bin = bin;
  a.ticket = b.ticket.p;
  const i64 groups = std::max<i64>(1, (max_rows + R - 1) / R);
  // Exactly the workgroups the device holds at once, so the grid-stride
  // loop runs in a single round on every CU.
  static int per_cu = 0;
  if (per_cu == 0) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_hash<G, T, WG, MODE, BATCH>, WG, 0) !=
            hipSuccess ||
        n < 1)
      n = 1;
    per_cu = n;
  }
  const unsigned grid = (unsigned)std::min<i64>(groups, (i64)device().cus * per_cu);
  hipLaunchKernelGGL((k_hash<G, T, WG, MODE, BATCH>), dim3(grid), dim3(WG), 0, s, a);
  RT_HIP(hipGetLastError());
}

// Device-binned rows (persistent kernels, counts read on the device).
struct DeviceBins {
  const HashArgs& a;
  const DevBins& b;
  i64 max_rows;
  hipStream_t s;
  template <int G, int T, int WG, int M, bool BT>
  void go(int bin) {
    launch_hash_dev<G, T, WG, M, BT>(a, b, bin, max_rows, s);
  }
};

template <int MODE, bool BATCH>
void launch_hash_bins_dev(const HashArgs& a, const DevBins& b, i64 max_rows, const int shape[5],
                          hipStream_t s) {
  DeviceBins l{a, b, max_rows, s};
  for_hash_bins<MODE, BATCH>(shape, l);
}

// Instantiated graphs of this host thread, by stage tag.  A stage captured
// again with the same tag is written into the cached executable graph with
// hipGraphExecUpdate (new buffer addresses and launch sizes) instead of
// being instantiated anew; when the update is refused (another topology) the
// graph is instantiated and replaces the cached one.  Kept until the
// process ends (destroying them at exit could outlive the runtime).
std::map<std::string, hipGraphExec_t>& graph_cache() {
  thread_local std::map<std::string, hipGraphExec_t>* cache =
      new std::map<std::string, hipGraphExec_t>();
  return *cache;
}

// Captures the launches `body` issues on `s` into a graph, runs it once (as
// the cached executable graph of `tag`) and returns the device time of the
// run.  Kernel timing is off inside the capture.
template <class F>
double run_captured(hipStream_t s, const std::string& tag, F&& body) {
  hipGraph_t g = nullptr;
  const bool prof = Prof::enabled;
  Prof::enabled = false;
  RT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    Prof::enabled = prof;
    hipGraph_t tmp = nullptr;
    (void)hipStreamEndCapture(s, &tmp);
    if (tmp) (void)hipGraphDestroy(tmp);
    throw;
  }
  Prof::enabled = prof;
  RT_HIP(hipStreamEndCapture(s, &g));
  hipGraphExec_t& ge = graph_cache()[tag];
  if (ge) {
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    if (hipGraphExecUpdate(ge, g, &err_node, &res) != hipSuccess ||
        res != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(ge);
      ge = nullptr;
    }
  }
  if (!ge) {
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      ge = nullptr;
      (void)hipGraphDestroy(g);
      RT_HIP(e);
    }
  }
  (void)hipGraphDestroy(g);
  hipEvent_t a, b;
  RT_HIP(hipEventCreate(&a));
  RT_HIP(hipEventCreate(&b));
  RT_HIP(hipEventRecord(a, s));
  RT_HIP(hipGraphLaunch(ge, s));
  RT_HIP(hipEventRecord(b, s));
  RT_HIP(hipEventSynchronize(b));
  float ms = 0;
  RT_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

// Cache tag of a stage: its name, the operand shape class and the tuning
// (which fixes the kernels launched).
std::string stage_tag(const char* stage, bool batch, i64 m, i32 nt, const Tuning& tu) {
  std::string t = std::string(stage) + (batch ? "/B/" : "/S/") + std::to_string(m) + "/" +
                  std::to_string(nt);
  for (int b = 0; b < 5; ++b)
    t += "/" + std::to_string(tu.sym[b]) + "," + std::to_string(tu.num[b]) + "," +
         std::to_string(tu.shape[b]);
  return t + "/" + std::to_string(tu.learned_heavy) + "," + std::to_string(tu.learned_bins);
}

template <bool BATCH>
void plain_graph_impl(const Csr& A, const Csr& B, const i64* bbase, CsrOut& C, const Tuning& tu,
                      hipStream_t s, RunStats* st) {
  const i64 m = A.rows, M = std::max<i64>(m, 1);
  const i32 nt = (i32)std::max<i64>(1, (B.cols + TILE_W - 1) / TILE_W);
  RT_CHECK(m < (i64)INT32_MAX, RT_E_LIMIT, "A has too many rows for 32-bit row lists");
  RT_CHECK((double)M * nt * 4.0 < 8.0 * (1ull << 30), RT_E_LIMIT,
           "graph mode: tile-count scratch too large; use the direct launch path");
  DBuf<i64> prod(M), nnz(M);
  DBuf<unsigned long long> tot(row_totals_len());
  DBuf<i32> longq(row_products_queue(A.nnz));
  DevBins sb(m), nb(m);
  DBuf<i32> tilecnt((size_t)M * nt), slab((size_t)std::max<i64>(B.rows, 1) * (nt + 1));
  DBuf<i32> colpad;  // B's columns for the symbolic long-row kernel's pair walk
  if (B.nnz < ((i64)1 << 31)) colpad.alloc((size_t)B.nnz + 1);
  const i64 nbs = std::max<i64>(1, (m + SCAN_TILE - 1) / SCAN_TILE);
  DBuf<i64> bsum(nbs);
  C.rows = m;
  C.cols = B.cols;
  C.rp = static_cast<i64*>(dmalloc((m + 1) * sizeof(i64)));
  HashArgs base = make_args(A, B, bbase);
  const unsigned heavy_grid = (unsigned)device().cus;
  DBuf<i32> next(2);  // row-slot counters of the two long-row kernels
  // No long row (symbolic or numeric) when no row has more products.
  const i64 long_th = std::min(tu.sym[4], tu.num[4]);
  double ms_a = 0, ms_b = 0;
  try {
    ms_a = run_captured(s, stage_tag("plain_a", BATCH, m, nt, tu), [&] {
      hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(256), 0, s, reinterpret_cast<i64*>(tot.p),
                         (i64)2, (i64)0);
      launch_row_products<BATCH>(m, A.rp, A.col, B.rp, bbase, prod.p, tot.p, longq.p, s);
      hipLaunchKernelGGL(k_fill_i64, dim3(fill_grid(M)), dim3(256), 0, s, nnz.p, M, (i64)0);
      bin_rows_dev(prod.p, m, B.cols, tu.sym, sb, s);
      HashArgs a = base;
      a.nnz = nnz.p;
      launch_hash_bins_dev<MODE_SYM, BATCH>(a, sb, m, tu.shape, s);
      if (B.rows > 0)
        hipLaunchKernelGGL(k_slab_ptrs, dim3(stride_grid(B.rows * (i64)(nt + 1))), dim3(256), 0,
                           s, B.rows, nt, B.rp, B.col, slab.p, tot.p + 1, long_th);
      const RowSet heavy{sb.list.p, 0, sb.cnt.p, sb.off.p, 5};
      zero_slot(next.p, s);
      if (colpad.p)
        hipLaunchKernelGGL(k_copy_cols, dim3(fill_grid(B.nnz + 1)), dim3(256), 0, s, B.nnz, B.col,
                           colpad.p, tot.p + 1, long_th);
      launch_sym_bitmap<BATCH>(heavy_grid * 2, heavy, A, B, bbase, slab.p, nt, tilecnt.p, nnz.p,
                               nullptr, nullptr, next.p, s, colpad.p);
      hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nbs), dim3(1024), 0, s, nnz.p, m, bsum.p);
      hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, bsum.p, nbs, C.rp + m);
      if (m > 0)
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nbs), dim3(1024), 0, s, nnz.p, m, bsum.p,
                           C.rp);
      RT_HIP(hipGetLastError());
    });
    C.nnz = read_scalar(C.rp + m, s);
    C.col = static_cast<i32*>(dmalloc(std::max<i64>(C.nnz, 1) * sizeof(i32)));
    C.val = static_cast<float*>(dmalloc(std::max<i64>(C.nnz, 1) * sizeof(float)));
    ms_b = run_captured(s, stage_tag("plain_b", BATCH, m, nt, tu), [&] {
      bin_rows_dev(nnz.p, m, LLONG_MAX, tu.num, nb, s);
      const RowSet heavy{sb.list.p, 0, sb.cnt.p, sb.off.p, 5};
      zero_slot(next.p + 1, s);
      hipLaunchKernelGGL((k_num_heavy<BATCH, false>), dim3(heavy_grid * 2), dim3(HWG), 0, s, heavy,
                         nt, tilecnt.p, nnz.p, tu.num[4], A.rp, A.col, A.val, B.rp, B.col, B.val,
                         bbase, slab.p, C.rp, C.col, C.val, nullptr, nullptr, nullptr, next.p + 1);
      HashArgs a = base;
      a.crp = C.rp;
      a.ccol = C.col;
      a.cval = C.val;
      launch_hash_bins_dev<MODE_NUM, BATCH>(a, nb, m, tu.shape, s);
      RT_HIP(hipGetLastError());
    });
  } catch (...) {
    (void)hipStreamSynchronize(s);
    throw;
  }
  unsigned long long h[2];
  RT_HIP(hipMemcpyAsync(h, tot.p, sizeof(h), hipMemcpyDeviceToHost, s));
  RT_HIP(hipStreamSynchronize(s));
  if (st) {
    st->products = (i64)h[0];
    st->max_row_products = (i64)h[1];
  }
  if (Prof::enabled) {
    Prof::record("graph_stage_a", ms_a);
    Prof::record("graph_stage_b", ms_b);
  }
}

// The learned method in three captured stages (two host reads between them:
// the temporary buffer's size, then C's).
//   A: products, features, predictions, capacities, keys, scan of the
//      capacities (temporary offsets), B's tile boundaries.
//   B: predicted rows through the learned hash kernels (and, with
//      learned_heavy, long rows through the learned tile kernel) into the
//      temporary buffer; overflowed rows join the exact set; the exact rows'
//      symbolic pass (hash bins, bitmap kernel); scan of nnz into C's row
//      pointer.
//   C: predicted rows copied into C; the exact rows' numeric pass.
template <bool BATCH>
void learned_graph_impl(const Csr& A, const Csr& B, const i64* bbase, CsrOut& C,
                        const Tuning& tu, hipStream_t s, RunStats* st) {
  const i64 m = A.rows, M = std::max<i64>(m, 1);
  const i32 nt = (i32)std::max<i64>(1, (B.cols + TILE_W - 1) / TILE_W);
  RT_CHECK(m < (i64)INT32_MAX, RT_E_LIMIT, "A has too many rows for 32-bit row lists");
  RT_CHECK((double)M * nt * 4.0 < 8.0 * (1ull << 30), RT_E_LIMIT,
           "graph mode: tile-count scratch too large; use the direct launch path");
  std::vector<float> hp;
  float margin = 0;
  learned_weights(hp, margin);
  DBuf<float> params(hp.size());
  RT_HIP(hipMemcpyAsync(params.p, hp.data(), hp.size() * sizeof(float), hipMemcpyHostToDevice, s));
  DBuf<i64> prod(M), nnz(M), cap(M), key_light(M), key_heavy(M), key_exact(M), key_num(M),
      tmp_off(m + 1);
  DBuf<i32> ovf(M);
  DBuf<float> feat(M * RT_NFEAT), pred(M);
  DBuf<unsigned long long> tot(row_totals_len()), counts(2);
  DBuf<i32> longq(row_products_queue(A.nnz));
  DevBins lb(m), hb(m), sb(m), nb(m);
  DBuf<i32> tilecnt((size_t)M * nt), slab((size_t)std::max<i64>(B.rows, 1) * (nt + 1));
  DBuf<i32> colpad;  // B's columns for the symbolic long-row kernel's pair walk
  if (B.nnz < ((i64)1 << 31)) colpad.alloc((size_t)B.nnz + 1);
  DBuf<i64> part(scan_partials(m));
  DBuf<i32> next(3);  // row-slot counters: learned tile, symbolic bitmap, numeric tile
  const i64 long_th = std::min(tu.sym[4], tu.num[4]);
  const unsigned heavy_grid = (unsigned)device().cus * 2;
  HashArgs base = make_args(A, B, bbase);
  C.rows = m;
  C.cols = B.cols;
  C.rp = static_cast<i64*>(dmalloc((m + 1) * sizeof(i64)));
  DBuf<i32> tcol;
  DBuf<float> tval;
  double ms[3] = {0, 0, 0};
  try {
    ms[0] = run_captured(s, stage_tag("learned_a", BATCH, m, nt, tu), [&] {
      hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(256), 0, s, reinterpret_cast<i64*>(tot.p),
                         (i64)2, (i64)0);
      hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(256), 0, s, reinterpret_cast<i64*>(counts.p),
                         (i64)2, (i64)0);
      launch_row_products<BATCH>(m, A.rp, A.col, B.rp, bbase, prod.p, tot.p, longq.p, s);
      hipLaunchKernelGGL(k_fill_i64, dim3(fill_grid(M)), dim3(256), 0, s, nnz.p, M, (i64)0);
      learned_predict_launch(A, B, bbase, prod.p, params.p, feat.p, pred.p,
                             tu.learned_heavy ? LLONG_MAX : long_th, predicted_floor(tu), s);
      learned_capacity_launch(m, prod.p, B.cols, pred.p, margin, tu, cap.p, key_light.p,
                              key_heavy.p, key_exact.p, ovf.p, s);
      exclusive_scan_i64(cap.p, tmp_off.p, m, s, part.p);
      if (B.rows > 0)
        hipLaunchKernelGGL(k_slab_ptrs, dim3(stride_grid(B.rows * (i64)(nt + 1))), dim3(256), 0,
                           s, B.rows, nt, B.rp, B.col, slab.p, tot.p + 1, long_th);
      RT_HIP(hipGetLastError());
    });
    const i64 tmp_total = read_scalar(tmp_off.p + m, s);
    tcol.alloc(std::max<i64>(tmp_total, 1));
    tval.alloc(std::max<i64>(tmp_total, 1));
    ms[1] = run_captured(s, stage_tag("learned_b", BATCH, m, nt, tu), [&] {
      HashArgs a = base;
      a.crp = tmp_off.p;
      a.ccol = tcol.p;
      a.cval = tval.p;
      a.cap = cap.p;
      a.ovf = ovf.p;
      a.nnz = nnz.p;
      bin_rows_dev(key_light.p, m, LLONG_MAX, tu.num, lb, s);
      launch_hash_bins_dev<MODE_LEARN, BATCH>(a, lb, m, tu.shape, s);
      if (tu.learned_heavy) {
        const i64 zero[5] = {0, 0, 0, 0, 0};  // every row with a key lands in bin 5
        bin_rows_dev(key_heavy.p, m, LLONG_MAX, zero, hb, s);
        zero_slot(next.p, s);
        hipLaunchKernelGGL((k_num_heavy<BATCH, true>), dim3(heavy_grid), dim3(HWG), 0, s,
                           RowSet{hb.list.p, 0, hb.cnt.p, hb.off.p, 5}, nt, nullptr, key_heavy.p,
                           (i64)0, A.rp, A.col, A.val, B.rp, B.col, B.val, bbase, slab.p,
                           tmp_off.p, tcol.p, tval.p, cap.p, nnz.p, ovf.p, next.p);
      }
      learned_merge_launch(m, prod.p, cap.p, ovf.p, key_exact.p, counts.p, s);
      bin_rows_dev(key_exact.p, m, B.cols, tu.sym, sb, s);
      HashArgs y = base;
      y.nnz = nnz.p;
      launch_hash_bins_dev<MODE_SYM, BATCH>(y, sb, m, tu.shape, s);
      zero_slot(next.p + 1, s);
      if (colpad.p)
        hipLaunchKernelGGL(k_copy_cols, dim3(fill_grid(B.nnz + 1)), dim3(256), 0, s, B.nnz, B.col,
                           colpad.p, tot.p + 1, long_th);
      launch_sym_bitmap<BATCH>(heavy_grid, RowSet{sb.list.p, 0, sb.cnt.p, sb.off.p, 5}, A, B,
                               bbase, slab.p, nt, tilecnt.p, nnz.p, nullptr, nullptr, next.p + 1,
                               s, colpad.p);
      exclusive_scan_i64(nnz.p, C.rp, m, s, part.p);
      RT_HIP(hipGetLastError());
    });
    C.nnz = read_scalar(C.rp + m, s);
    C.col = static_cast<i32*>(dmalloc(std::max<i64>(C.nnz, 1) * sizeof(i32)));
    C.val = static_cast<float*>(dmalloc(std::max<i64>(C.nnz, 1) * sizeof(float)));
    ms[2] = run_captured(s, stage_tag("learned_c", BATCH, m, nt, tu), [&] {
      learned_finish_launch(m, cap.p, ovf.p, tmp_off.p, tcol.p, tval.p, key_exact.p, nnz.p,
                            key_num.p, C, s);
      bin_rows_dev(key_num.p, m, LLONG_MAX, tu.num, nb, s);
      zero_slot(next.p + 2, s);
      hipLaunchKernelGGL((k_num_heavy<BATCH, false>), dim3(heavy_grid), dim3(HWG), 0, s,
                         RowSet{sb.list.p, 0, sb.cnt.p, sb.off.p, 5}, nt, tilecnt.p, key_num.p,
                         tu.num[4], A.rp, A.col, A.val, B.rp, B.col, B.val, bbase, slab.p, C.rp,
                         C.col, C.val, nullptr, nullptr, nullptr, next.p + 2);
      HashArgs a = base;
      a.crp = C.rp;
      a.ccol = C.col;
      a.cval = C.val;
      launch_hash_bins_dev<MODE_NUM, BATCH>(a, nb, m, tu.shape, s);
      RT_HIP(hipGetLastError());
    });
  } catch (...) {
    (void)hipStreamSynchronize(s);
    throw;
  }
  unsigned long long h[4];
  RT_HIP(hipMemcpyAsync(h, tot.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  RT_HIP(hipMemcpyAsync(h + 2, counts.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  RT_HIP(hipStreamSynchronize(s));
  if (st) {
    st->products = (i64)h[0];
    st->max_row_products = (i64)h[1];
    st->predicted_rows = (i64)h[2];
    st->overflow_rows = (i64)h[3];
  }
  if (Prof::enabled) {
    Prof::record("graph_learned_a", ms[0]);
    Prof::record("graph_learned_b", ms[1]);
    Prof::record("graph_learned_c", ms[2]);
  }
}

}  // namespace

void spgemm_learned_graph(const Csr& A, const Csr& B, const Batch& batch, CsrOut& C,
                          const Tuning& tune, hipStream_t s, RunStats* st) {
  if (tune.learned_bins == 0 && !tune.learned_heavy)  // no row predicted: the exact method
    spgemm_plain_graph(A, B, batch, C, tune, s, st);
  else if (batch.b_base)
    learned_graph_impl<true>(A, B, batch.b_base, C, tune, s, st);
  else
    learned_graph_impl<false>(A, B, nullptr, C, tune, s, st);
}

void spgemm_plain_graph(const Csr& A, const Csr& B, const Batch& batch, CsrOut& C,
                        const Tuning& tune, hipStream_t s, RunStats* st) {
  if (batch.b_base)
    plain_graph_impl<true>(A, B, batch.b_base, C, tune, s, st);
  else
    plain_graph_impl<false>(A, B, nullptr, C, tune, s, st);
}

}  // namespace rt

// ---------------------------------------------------------------------------
// The exact two-phase product without any device allocation
// (rt_spmspm_noalloc), for callers that cannot allocate: C goes into the
// caller's arrays, and every scratch buffer of the pipeline -- per-row
// products and counts, the two row-bin lists and their counters, the scan
// partials, the validation flag, B's tile boundaries, the long rows' tile
// counts and the row-slot counters -- is carved from `scratch`, one
// file-scope device array of NOALLOC_SCRATCH_BYTES.  The kernels are those of
// the direct exact method (LDS-hash bins, LDS-bitmap symbolic and dense-tile
// numeric long-row kernels), launched with host-read bin counts.
namespace rt {

// In host code this name is only a stand-in for the device symbol: kernels
// get the address hipGetSymbolAddress returns (scratch_base).
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

// Device id -> that device's copy of `scratch`, looked up once per device
// (with the device current) on the first call there.
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_addr;

unsigned char* scratch_base() {
  int dev_id = 0;
  RT_HIP(hipGetDevice(&dev_id));
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_addr.find(dev_id);
  if (it != g_scratch_addr.end()) return it->second;
  void* p = nullptr;
  RT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  RT_CHECK(p != nullptr, RT_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch_addr[dev_id] = static_cast<unsigned char*>(p);
  return static_cast<unsigned char*>(p);
}

// Bump allocation of 256-byte aligned pieces of the scratch array.
struct ScratchCarve {
  unsigned char* base;
  i64 used = 0;
  i64 need(i64 bytes) const { return ((used + 255) & ~(i64)255) + bytes; }
  template <class T>
  T* take(i64 count) {
    const i64 off = (used + 255) & ~(i64)255;
    const i64 bytes = std::max<i64>(count, 1) * (i64)sizeof(T);
    RT_CHECK(off + bytes <= NOALLOC_SCRATCH_BYTES, RT_E_LIMIT,
             "rt_spmspm_noalloc: this product needs more than the " +
                 std::to_string(NOALLOC_SCRATCH_BYTES >> 20) +
                 " MiB of scratch the entry point has (at least " +
                 std::to_string((off + bytes + (1 << 20) - 1) >> 20) +
                 " MiB); use rt_spmspm_ex");
    used = off + bytes;
    return reinterpret_cast<T*>(base + off);
  }
};

// Calls on one device share its scratch: they are serialised here (each
// call is synchronous, so the next one starts after the previous finished).
std::mutex g_noalloc_mu;

}  // namespace

void spgemm_noalloc(const Csr& A, const Csr& B, i64* crp, i32* ccol, float* cval, i64 capacity,
                    i64* c_nnz, bool trusted, hipStream_t s) {
  std::lock_guard<std::mutex> serial(g_noalloc_mu);
  if (c_nnz) *c_nnz = -1;
  const i64 m = A.rows, M = std::max<i64>(m, 1);
  RT_CHECK(m < (i64)INT32_MAX, RT_E_LIMIT, "A has too many rows for 32-bit row lists");
  const i32 nt = (i32)std::max<i64>(1, (B.cols + TILE_W - 1) / TILE_W);
  const Tuning tu;
  ScratchCarve sc{scratch_base()};
  // Everything but the long rows' pieces is taken (and its size checked)
  // before the first launch.
  int* flag = sc.take<int>(1);
  unsigned long long* tot = sc.take<unsigned long long>((i64)row_totals_len());
  i32* longq = sc.take<i32>((i64)row_products_queue(A.nnz));
  i64* prod = sc.take<i64>(M);
  i64* nnz = sc.take<i64>(M);
  i32* slist = sc.take<i32>(M);
  i32* nlist = sc.take<i32>(M);
  i32* counters = sc.take<i32>(bin_scratch_ints());
  i64* partials = sc.take<i64>(scan_partials(m));
  i32* next = sc.take<i32>(2);
  if (!trusted) {
    validate_csr_dev(A, A.cols, nullptr, s, "A", flag);
    validate_csr_dev(B, B.cols, nullptr, s, "B", flag);
  }
  // Symbolic phase.
  hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(256), 0, s, reinterpret_cast<i64*>(tot), (i64)2,
                     (i64)0);
  launch_row_products<false>(m, A.rp, A.col, B.rp, nullptr, prod, tot, longq, s);
  hipLaunchKernelGGL(k_fill_i64, dim3(fill_grid(M)), dim3(256), 0, s, nnz, M, (i64)0);
  RT_HIP(hipGetLastError());
  int scnt[NBIN], soff[NBIN + 1];
  bin_rows(prod, m, B.cols, tu.sym, slist, scnt, soff, s, counters);
  HashArgs a = make_args(A, B, nullptr);
  a.nnz = nnz;
  launch_hash_bins<MODE_SYM, false>(a, slist, scnt, soff, tu.shape, s);
  const i32 nlong = scnt[5];
  i32* slab = nullptr;
  i32* tilecnt = nullptr;
  if (nlong > 0) {
    slab = sc.take<i32>(std::max<i64>(B.rows, 1) * (nt + 1));
    tilecnt = sc.take<i32>((i64)nlong * nt);
    if (B.rows > 0)
      hipLaunchKernelGGL(k_slab_ptrs, dim3(stride_grid(B.rows * (i64)(nt + 1))), dim3(256), 0, s,
                         B.rows, nt, B.rp, B.col, slab, nullptr, (i64)0);
    zero_slot(next, s);
    const RowSet rs{slist + soff[5], nlong, nullptr, nullptr, 0};
    launch_sym_bitmap<false>((unsigned)std::min<i64>(nlong, (i64)device().cus * 2), rs, A, B,
                             nullptr, slab, nt, tilecnt, nnz, nullptr, nullptr, next, s);
    RT_HIP(hipGetLastError());
  }
  exclusive_scan_i64(nnz, crp, m, s, partials);
  const i64 total = read_scalar(crp + m, s);
  if (c_nnz) *c_nnz = total;
  RT_CHECK(total <= capacity, RT_E_LIMIT,
           "rt_spmspm_noalloc: C has " + std::to_string(total) + " entries, the output arrays hold " +
               std::to_string(capacity) + " (C's row pointer is filled)");
  RT_CHECK(total == 0 || (ccol && cval), RT_E_INVALID, "rt_spmspm_noalloc: NULL C arrays");
  // Numeric phase.
  int ncnt[NBIN], noff[NBIN + 1];
  bin_rows(nnz, m, LLONG_MAX, tu.num, nlist, ncnt, noff, s, counters);
  if (ncnt[5] > 0) {
    zero_slot(next + 1, s);
    const RowSet rs{slist + soff[5], nlong, nullptr, nullptr, 0};
    hipLaunchKernelGGL((k_num_heavy<false, false>),
                       dim3((unsigned)std::min<i64>(nlong, (i64)device().cus * 2)), dim3(HWG), 0,
                       s, rs, nt, tilecnt, nnz, tu.num[4], A.rp, A.col, A.val, B.rp, B.col, B.val,
                       nullptr, slab, crp, ccol, cval, nullptr, nullptr, nullptr, next + 1);
    RT_HIP(hipGetLastError());
  }
  a = make_args(A, B, nullptr);
  a.crp = crp;
  a.ccol = ccol;
  a.cval = cval;
  launch_hash_bins<MODE_NUM, false>(a, nlist, ncnt, noff, tu.shape, s);
  RT_HIP(hipGetLastError());
  RT_HIP(hipStreamSynchronize(s));
}

}  // namespace rt
