// Device-side helpers shared by the SpGEMM kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rt {
namespace dev {

using i64 = int64_t;
using i32 = int32_t;
using u16 = uint16_t;
using u32 = uint32_t;
using u64 = uint64_t;

constexpr int WAVE = 64;
constexpr i32 EMPTY_KEY = -1;

// Make LDS writes of this wave visible to its other lanes and stop the
// compiler from moving LDS accesses across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A "group" is the set of G consecutive lanes that own one output row.
// G <= 64: the group lives inside one wave and syncs with wave_sync.
// G  > 64: the group is the whole workgroup (one row per workgroup).
template <int G>
__device__ __forceinline__ void group_sync() {
  if constexpr (G <= WAVE)
    wave_sync();
  else
    __syncthreads();
}

// Inclusive scan over segments of W lanes (W <= 64, power of two).
template <int W, class T>
__device__ __forceinline__ T seg_scan_incl(T v) {
  const int l = threadIdx.x & (W - 1);
#pragma unroll
  for (int d = 1; d < W; d <<= 1) {
    T o = __shfl_up(v, d, W);
    if (l >= d) v += o;
  }
  return v;
}

template <int W, class T>
__device__ __forceinline__ T seg_reduce_sum(T v) {
#pragma unroll
  for (int d = W >> 1; d > 0; d >>= 1) v += __shfl_xor(v, d, W);
  return v;
}

template <int W, class T>
__device__ __forceinline__ T seg_reduce_max(T v) {
#pragma unroll
  for (int d = W >> 1; d > 0; d >>= 1) {
    T o = __shfl_xor(v, d, W);
    v = o > v ? o : v;
  }
  return v;
}

// Inclusive max-scan over the 64 lanes of a wave for values >= 0, by DPP:
// row_shr 1/2/4/8 scan each row of 16 lanes (lanes shifted in from outside
// the row read 0), then row_bcast 15 / 31 carry row 0 into row 1, row 2 into
// row 3, and rows 0-1 into rows 2-3.  Every lane of the wave must be active.
__device__ __forceinline__ int wave_max_scan(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return v;
}

// Inclusive scan over the whole workgroup of WG threads.  `scratch` holds at
// least WG/64 elements; *total receives the sum.  Contains two barriers; the
// caller must barrier before reusing `scratch`.
template <int WG, class T>
__device__ __forceinline__ T block_scan_incl(T v, T* scratch, T* total) {
  constexpr int NW = WG / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  v = seg_scan_incl<WAVE>(v);
  if (lane == WAVE - 1) scratch[w] = v;
  __syncthreads();
  if (w == 0) {
    T s = lane < NW ? scratch[lane] : T(0);
    s = seg_scan_incl<(NW < 2 ? 2 : NW)>(s);
    if (lane < NW) scratch[lane] = s;
  }
  __syncthreads();
  if (w > 0) v += scratch[w - 1];
  *total = scratch[NW - 1];
  return v;
}

template <int WG, class T>
__device__ __forceinline__ T block_reduce_sum(T v, T* scratch) {
  constexpr int NW = WG / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  v = seg_reduce_sum<WAVE>(v);
  if (lane == 0) scratch[w] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) s += scratch[i];
  __syncthreads();
  return s;
}

// Multiplicative hash of a column index into [0, 2^LOG2T).
template <int LOG2T>
__device__ __forceinline__ u32 hash_col(i32 c) {
  return ((u32)c * 0x9E3779B1u) >> (32 - LOG2T);
}

template <int N>
struct Log2 {
  static constexpr int value = 1 + Log2<N / 2>::value;
};
template <>
struct Log2<1> {
  static constexpr int value = 0;
};

// Relaxed LDS / global atomics whose results are unused.
//
// FP32 add into LDS by a compare-and-swap loop on the word's bits.  On
// gfx950 the native LDS float atomic (ds_add_f32) sustains about 177 CU
// cycles per wave instruction, while integer LDS atomics (compare-and-swap
// included) run at 5-14: measured with 64 lanes on distinct addresses.  The
// loop keeps the atomic's IEEE single-precision semantics (one rounded add
// per product, in arbitrary order); lanes of a wave that collide on one
// word retry, and every round completes at least one of them.
__device__ __forceinline__ void lds_add_f32(float* p, float v) {
  u32* q = reinterpret_cast<u32*>(p);
  u32 old = *q, assumed;
  do {
    assumed = old;
    old = atomicCAS(q, assumed, __float_as_uint(__uint_as_float(assumed) + v));
  } while (old != assumed);
}
__device__ __forceinline__ void lds_or_u32(u32* p, u32 v) {
  __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void global_add_f32(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace dev
}  // namespace rt
