// Faithful "ray launch" SpGEMM: one work-item per nonzero of A, as the
// reference launches one ray per A nonzero (optixLaunch of width nnz(A),
// optixSpMSpM.cpp:607-608; __raygen__rg, optixSpMSpM.cu:113-201).  The ray's
// hits are the spheres of B's row k (Sphere.cpp:72-75 places B(k,j) at (k,j,0)
// and the ray runs along x = k, optixSpMSpM.cu:143-145); here they are read
// through B's row pointer.  Every hit performs the any-hit accumulate
// C(i,j) += A(i,k)*B(k,j) with one global fp32 atomic add
// (optixSpMSpM.cu:245-259), exactly once per product.
//
// Output buffers:
//   dense: fp32 C[rows*cols] plus a byte touch map (the reference's dense
//          result buffer, optixSpMSpM.cpp:164-171; the touch map keeps the
//          structure of cancelled entries, which a dense buffer alone loses);
//          converted to CSR afterwards.
//   slot:  C's CSR structure comes from the exact symbolic+numeric pass; each
//          hit binary-searches its column in row i and adds into that slot.
#include <climits>

#include "device.hpp"
#include "spgemm.hpp"

namespace rt {
using namespace dev;

// Row of every A nonzero (CSR -> COO expansion), by binary search.
__global__ __launch_bounds__(256) void k_csr_rows(i64 nnz, i64 nrows, const i64* __restrict__ rp,
                                                  i32* __restrict__ rowidx) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= nnz) return;
  i64 lo = 0, hi = nrows - 1;
  while (lo < hi) {
    const i64 mid = (lo + hi + 1) >> 1;
    if (rp[mid] <= e) lo = mid; else hi = mid - 1;
  }
  rowidx[e] = (i32)lo;
}

// ABL: RT_ABL_* (full / no_atomic / traverse / no_hit / empty).
// OUT: 1 dense, 2 slot.
template <int ABL, int OUT>
__global__ __launch_bounds__(256) void k_ray(
    i64 nnz, const i32* __restrict__ rowidx, const i32* __restrict__ acol,
    const float* __restrict__ aval, const i64* __restrict__ brp, const i32* __restrict__ bcol,
    const float* __restrict__ bval, i64 ncols, float* __restrict__ dense,
    unsigned char* __restrict__ touched, const i64* __restrict__ crp,
    const i32* __restrict__ ccol, float* __restrict__ cval, float* __restrict__ abl_buf) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if constexpr (ABL == RT_ABL_EMPTY) {
    return;  // NOTHING: launch only (optixSpMSpM.cu:106-111)
  } else {
    if (e >= nnz) return;
    const i32 i = rowidx[e];
    const i64 k = acol[e];
    const float a = aval[e];
    if constexpr (ABL == RT_ABL_NO_HIT) {
      // NOINT (optixSpMSpM.cu:151-189): a ray that hits nothing followed by a
      // loop of nnz/5 steps whose result is never stored; the loop has no
      // effect and is removed by the compiler, in the reference as here.
      float dummy = 0.f;
      for (i64 q = 0; q < nnz; q += 5) dummy = a * 3.33f * (float)q;
      (void)dummy;
      return;
    } else {
      const i64 q0 = brp[k], q1 = brp[k + 1];
      float probe = 0.f;
      for (i64 q = q0; q < q1; ++q) {
        const i32 j = bcol[q];
        if constexpr (ABL == RT_ABL_NO_ATOMIC) {
          // ARCHSUP (optixSpMSpM.cu:222-227): racy non-atomic increment of a
          // 256-entry buffer indexed by the primitive id.
          abl_buf[q & 0xff] += 1.f;
        } else if constexpr (ABL == RT_ABL_TRAVERSE) {
          // NOMEM (optixSpMSpM.cu:228-229): the hit is found and ignored.
          probe += (float)j;
        } else {
          const float v = a * bval[q];
          if constexpr (OUT == RT_RAY_DENSE) {
            const i64 idx = (i64)i * ncols + j;
            global_add_f32(dense + idx, v);
            touched[idx] = 1;
          } else {
            i64 lo = crp[i], hi = crp[i + 1] - 1;
            while (lo < hi) {
              const i64 mid = (lo + hi) >> 1;
              if (ccol[mid] < j) lo = mid + 1; else hi = mid;
            }
            global_add_f32(cval + lo, v);
          }
        }
      }
      if constexpr (ABL == RT_ABL_TRAVERSE) {
        if (probe == -1.f) abl_buf[0] = probe;  // keeps the walk alive; never true
      }
    }
  }
}

// Dense touch map -> CSR: count per row, then fill in column order.
__global__ __launch_bounds__(256) void k_dense_count(i64 rows, i64 cols,
                                                     const unsigned char* __restrict__ touched,
                                                     i64* __restrict__ cnt) {
  __shared__ i64 s_scr[4];
  const i64 r = blockIdx.x;
  const unsigned char* t = touched + r * cols;
  i64 c = 0;
  for (i64 j = threadIdx.x; j < cols; j += 256) c += t[j];
  c = block_reduce_sum<256>(c, s_scr);
  if (threadIdx.x == 0) cnt[r] = c;
}

__global__ __launch_bounds__(256) void k_dense_fill(i64 rows, i64 cols,
                                                    const unsigned char* __restrict__ touched,
                                                    const float* __restrict__ dense,
                                                    const i64* __restrict__ crp,
                                                    i32* __restrict__ ccol,
                                                    float* __restrict__ cval) {
  __shared__ i64 s_scr[4];
  const i64 r = blockIdx.x;
  const unsigned char* t = touched + r * cols;
  const float* d = dense + r * cols;
  const i64 per = (cols + 255) / 256;
  const i64 j0 = threadIdx.x * per, j1 = min(cols, j0 + per);
  i64 c = 0;
  for (i64 j = j0; j < j1; ++j) c += t[j];
  i64 tot;
  const i64 inc = block_scan_incl<256>(c, s_scr, &tot);
  i64 pos = crp[r] + inc - c;
  for (i64 j = j0; j < j1; ++j)
    if (t[j]) {
      ccol[pos] = (i32)j;
      cval[pos] = d[j];
      ++pos;
    }
}

namespace {
template <int ABL, int OUT>
void launch_ray(i64 nnz, const i32* rowidx, const Csr& A, const Csr& B, float* dense,
                unsigned char* touched, const CsrOut& C, float* abl_buf, hipStream_t s) {
  const i64 grid = std::max<i64>(1, (nnz + 255) / 256);
  ProfScope ps("ray_launch", s);
  hipLaunchKernelGGL((k_ray<ABL, OUT>), dim3((unsigned)grid), dim3(256), 0, s, nnz, rowidx, A.col,
                     A.val, B.rp, B.col, B.val, B.cols, dense, touched, C.rp, C.col, C.val,
                     abl_buf);
  RT_HIP(hipGetLastError());
}

template <int OUT>
void dispatch_ray(int abl, i64 nnz, const i32* rowidx, const Csr& A, const Csr& B, float* dense,
                  unsigned char* touched, const CsrOut& C, float* abl_buf, hipStream_t s) {
  switch (abl) {
    case RT_ABL_FULL: launch_ray<RT_ABL_FULL, OUT>(nnz, rowidx, A, B, dense, touched, C, abl_buf, s); break;
    case RT_ABL_NO_ATOMIC: launch_ray<RT_ABL_NO_ATOMIC, OUT>(nnz, rowidx, A, B, dense, touched, C, abl_buf, s); break;
    case RT_ABL_TRAVERSE: launch_ray<RT_ABL_TRAVERSE, OUT>(nnz, rowidx, A, B, dense, touched, C, abl_buf, s); break;
    case RT_ABL_NO_HIT: launch_ray<RT_ABL_NO_HIT, OUT>(nnz, rowidx, A, B, dense, touched, C, abl_buf, s); break;
    case RT_ABL_EMPTY: launch_ray<RT_ABL_EMPTY, OUT>(nnz, rowidx, A, B, dense, touched, C, abl_buf, s); break;
    default: fail(RT_E_INVALID, "unknown ablation " + std::to_string(abl));
  }
}
}  // namespace

void spgemm_ray(const Csr& A, const Csr& B, CsrOut& C, int ablation, int output, hipStream_t s,
                RunStats* st) {
  const i64 m = A.rows, n = B.cols;
  RT_CHECK(m < (i64)INT32_MAX, RT_E_LIMIT, "A has too many rows");
  const double dense_bytes = (double)m * (double)n * 5.0;
  if (output == RT_RAY_AUTO) output = dense_bytes <= 4.0 * (1ull << 30) ? RT_RAY_DENSE : RT_RAY_SLOT;
  RT_CHECK(output == RT_RAY_DENSE || output == RT_RAY_SLOT, RT_E_INVALID, "bad ray output mode");

  DBuf<i32> rowidx(std::max<i64>(A.nnz, 1));
  if (A.nnz > 0) {
    ProfScope ps("csr_rows", s);
    hipLaunchKernelGGL(k_csr_rows, dim3((unsigned)((A.nnz + 255) / 256)), dim3(256), 0, s, A.nnz,
                       m, A.rp, rowidx.p);
    RT_HIP(hipGetLastError());
  }
  DBuf<float> abl(256);
  RT_HIP(hipMemsetAsync(abl.p, 0, 256 * sizeof(float), s));

  if (output == RT_RAY_DENSE) {
    C.rows = m;
    C.cols = n;
    DBuf<float> dense(std::max<i64>(m * n, 1));
    DBuf<unsigned char> touched(std::max<i64>(m * n, 1));
    RT_HIP(hipMemsetAsync(dense.p, 0, std::max<i64>(m * n, 1) * sizeof(float), s));
    RT_HIP(hipMemsetAsync(touched.p, 0, std::max<i64>(m * n, 1), s));
    if (st) st->products = -1;  // not computed on this path
    dispatch_ray<RT_RAY_DENSE>(ablation, A.nnz, rowidx.p, A, B, dense.p, touched.p, C, abl.p, s);
    // Structure from the touch map (the ablations touch nothing: they return
    // an empty C, as the reference's ablation builds leave C zero).
    DBuf<i64> cnt(std::max<i64>(m, 1));
    if (m > 0 && n > 0) {
      ProfScope ps("dense_count", s);
      hipLaunchKernelGGL(k_dense_count, dim3((unsigned)m), dim3(256), 0, s, m, n, touched.p,
                         cnt.p);
      RT_HIP(hipGetLastError());
    } else {
      RT_HIP(hipMemsetAsync(cnt.p, 0, std::max<i64>(m, 1) * sizeof(i64), s));
    }
    C.rp = static_cast<i64*>(dmalloc((m + 1) * sizeof(i64)));
    exclusive_scan_i64(cnt.p, C.rp, m, s);
    C.nnz = read_scalar(C.rp + m, s);
    C.col = static_cast<i32*>(dmalloc(std::max<i64>(C.nnz, 1) * sizeof(i32)));
    C.val = static_cast<float*>(dmalloc(std::max<i64>(C.nnz, 1) * sizeof(float)));
    if (m > 0 && n > 0) {
      ProfScope ps("dense_fill", s);
      hipLaunchKernelGGL(k_dense_fill, dim3((unsigned)m), dim3(256), 0, s, m, n, touched.p,
                         dense.p, C.rp, C.col, C.val);
      RT_HIP(hipGetLastError());
    }
    RT_HIP(hipStreamSynchronize(s));
  } else {
    // Structure from the exact two-phase pass, then the ray kernel fills it.
    spgemm_plain(A, B, Batch{}, C, Tuning{}, s, st);
    RT_HIP(hipMemsetAsync(C.val, 0, std::max<i64>(C.nnz, 1) * sizeof(float), s));
    dispatch_ray<RT_RAY_SLOT>(ablation, A.nnz, rowidx.p, A, B, nullptr, nullptr, C, abl.p, s);
    RT_HIP(hipStreamSynchronize(s));
  }
}

}  // namespace rt

// ---------------------------------------------------------------------------
// Streaming device copy (16-byte vectors, grid-stride; byte tail handled by
// the first thread).  Used by rt_device_copy.
namespace rt {
__global__ __launch_bounds__(256) void k_device_copy(uint4* __restrict__ dst,
                                                     const uint4* __restrict__ src, i64 n16,
                                                     unsigned char* __restrict__ dtail,
                                                     const unsigned char* __restrict__ stail,
                                                     int ntail) {
  const i64 stride = (i64)gridDim.x * 256;
  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int t = 0; t < ntail; ++t) dtail[t] = stail[t];
}

double device_copy(void* dst, const void* src, i64 bytes, hipStream_t s) {
  RT_CHECK(((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0, RT_E_INVALID,
           "device_copy needs 16-byte aligned pointers");
  const i64 n16 = bytes / 16;
  const int ntail = (int)(bytes - n16 * 16);
  const unsigned grid = (unsigned)std::max<i64>(1, std::min<i64>((n16 + 255) / 256,
                                                                  (i64)device().cus * 16));
  hipEvent_t a, b;
  RT_HIP(hipEventCreate(&a));
  RT_HIP(hipEventCreate(&b));
  RT_HIP(hipEventRecord(a, s));
  hipLaunchKernelGGL(k_device_copy, dim3(grid), dim3(256), 0, s, static_cast<uint4*>(dst),
                     static_cast<const uint4*>(src), n16,
                     static_cast<unsigned char*>(dst) + n16 * 16,
                     static_cast<const unsigned char*>(src) + n16 * 16, ntail);
  RT_HIP(hipGetLastError());
  RT_HIP(hipEventRecord(b, s));
  RT_HIP(hipEventSynchronize(b));
  float ms = 0;
  RT_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}
}  // namespace rt
