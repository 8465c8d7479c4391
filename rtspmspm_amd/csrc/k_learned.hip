// Learned row sizing for the numeric pass.
//
// The two-phase method (cuSPARSE's workEstimation/estimateMemory/compute in
// the reference's comparator, cuSparse/src/main.cpp:464-539; the exact
// symbolic pass here) spends one full pass over the products only to learn
// each output row's size.  This path predicts that size instead:
//   1. per-row features (A-row nnz, product count, longest and mean B row of
//      the row's columns, occupancy of a 64-bucket hash of a sample of the
//      row's product columns, ...),
//   2. a small MLP (8 -> 32 -> 32 -> 1, ReLU) predicts log2(1 + nnz) of the
//      row; capacity = 2^(prediction + margin) - 1, clamped to [1, U],
//   3. short rows run through the numeric LDS-hash kernels with tables sized
//      from the capacity, long rows through the dense-tile kernel; both write
//      the row into a temporary buffer at capacity-sized offsets, and a row
//      that finds more distinct columns than its capacity is abandoned and
//      flagged,
//   4. flagged rows go through the exact symbolic + numeric path; the others
//      are copied into C.
// C is identical to the exact method's: the structure of every row is
// computed exactly either way, only the sizing of the tables differs.
//
// The network is trained on (features, exact nnz) pairs that the exact
// symbolic pass emits, with the forward/backward/Adam kernels below.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "device.hpp"
#include "hostcsr.hpp"
#include "learned.hpp"

namespace rt {
using namespace dev;

// ---- network layout ---------------------------------------------------------
constexpr int NF = RT_NFEAT;  // 8 inputs
constexpr int H = 32;         // two hidden layers of 32
constexpr int OW1 = 0, OB1 = OW1 + H * NF, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H, NPARAM = OB3 + 1;  // 1377
// The output is a residual on log2(1 + min(P, cols)) (feature 6 * 32): the
// network learns how far below the upper bound the row's nnz falls.
constexpr float BASE_SCALE = 32.f;

// ---- features -----------------------------------------------------------------
template <bool BATCH>
__global__ __launch_bounds__(256) void k_features(i64 m, const i64* __restrict__ arp,
                                                  const i32* __restrict__ acol,
                                                  const i64* __restrict__ brp,
                                                  const i32* __restrict__ bcol,
                                                  const i64* __restrict__ bbase, i64 ncols,
                                                  const i64* __restrict__ prod, i64 skip_lo,
                                                  i64 skip_u, float* __restrict__ feat) {
  constexpr int G = 16;
  const i64 row = ((i64)blockIdx.x * 256 + threadIdx.x) / G;
  const int l = threadIdx.x & (G - 1);
  if (row >= m) return;  // whole group leaves together
  const i64 u_row = min(prod[row], ncols);
  if (u_row > skip_u || u_row <= skip_lo) {  // a row the caller sends to the exact path unseen
    if (l < NF) feat[row * NF + l] = 0.f;
    return;
  }
  const i64 a0 = arp[row], a1 = arp[row + 1], nA = a1 - a0;
  const i64 bb = BATCH ? bbase[row] : 0;
  i64 maxlen = 0;
  for (i64 e = a0 + l; e < a1; e += G) {
    const i64 k = (i64)acol[e] + bb;
    maxlen = max(maxlen, brp[k + 1] - brp[k]);
  }
  maxlen = seg_reduce_max<G>(maxlen);
  // Sample: up to 16 entries of the A row spread evenly, and up to 8
  // columns spread evenly over each one's B row; hash into 64 buckets.
  u32 mlo = 0, mhi = 0;
  int ns = 0;
  if (l < min((i64)G, nA)) {
    const i64 e = a0 + (nA <= G ? l : ((i64)l * nA) / G);
    const i64 k = (i64)acol[e] + bb;
    const i64 q0 = brp[k], L = brp[k + 1] - q0;
    ns = (int)min((i64)8, L);
    for (int t = 0; t < ns; ++t) {
      const i32 c = bcol[q0 + (t * L) / ns];
      const u32 bkt = ((u32)c * 0x9E3779B1u) >> 26;
      if (bkt < 32) mlo |= 1u << bkt; else mhi |= 1u << (bkt - 32);
    }
  }
#pragma unroll
  for (int d = G >> 1; d > 0; d >>= 1) {
    mlo |= __shfl_xor(mlo, d, G);
    mhi |= __shfl_xor(mhi, d, G);
  }
  ns = seg_reduce_sum<G>(ns);
  if (l < NF) {
    const double P = (double)prod[row];
    const double U = fmin(P, (double)ncols);
    float f;
    switch (l) {
      case 0: f = (float)(log2(1.0 + (double)nA) / 16.0); break;
      case 1: f = (float)(log2(1.0 + P) / 32.0); break;
      case 2: f = (float)(log2(1.0 + (double)maxlen) / 24.0); break;
      case 3: f = (float)(log2(1.0 + P / fmax(1.0, (double)nA)) / 24.0); break;
      case 4: f = (float)(__popc(mlo) + __popc(mhi)) / 64.f; break;
      case 5: f = (float)(log2(1.0 + (double)ns) / 8.0); break;
      case 6: f = (float)(log2(1.0 + U) / 32.0); break;
      default: f = (float)(log2(1.0 + (double)ncols) / 32.0); break;
    }
    feat[row * NF + l] = f;
  }
}

// ---- network kernels -------------------------------------------------------------
__device__ __forceinline__ float mlp_fwd(const float* __restrict__ p, const float (&x)[NF],
                                         float (&h1)[H], float (&h2)[H]) {
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float s = p[OB1 + j];
#pragma unroll
    for (int i = 0; i < NF; ++i) s = fmaf(p[OW1 + j * NF + i], x[i], s);
    h1[j] = fmaxf(s, 0.f);
  }
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float s = p[OB2 + j];
#pragma unroll
    for (int i = 0; i < H; ++i) s = fmaf(p[OW2 + j * H + i], h1[i], s);
    h2[j] = fmaxf(s, 0.f);
  }
  float o = p[OB3];
#pragma unroll
  for (int j = 0; j < H; ++j) o = fmaf(p[OW3 + j], h2[j], o);
  return o + BASE_SCALE * x[6];
}

// One forward + backward over samples idx[0..n) (idx may be null: 0..n).
// Loss = mean (y_hat - y)^2 with mean over `inv_n`; accumulates dLoss/dparam
// into grad and the loss into *loss.  Per-sample activations are staged in
// LDS and the parameter gradients reduced over the workgroup's samples as
// small matrix products (one parameter per thread at a time).
constexpr int TWG = 128;
constexpr int LDX = NF + 1, LDH = H + 1;  // padded rows against bank conflicts

__global__ __launch_bounds__(TWG) void k_mlp_grad(const float* __restrict__ params,
                                                  const float* __restrict__ feat,
                                                  const float* __restrict__ target,
                                                  const i32* __restrict__ idx, i64 n,
                                                  float inv_n, float* __restrict__ grad,
                                                  float* __restrict__ loss, int want_grad) {
  __shared__ float sp[NPARAM];
  __shared__ float sX[TWG * LDX];
  __shared__ float sH1[TWG * LDH], sD1[TWG * LDH], sH2[TWG * LDH], sD2[TWG * LDH];
  __shared__ float sG[TWG];
  __shared__ float s_red[TWG / WAVE];
  const int t = threadIdx.x;
  for (int q = t; q < NPARAM; q += TWG) sp[q] = params[q];
  __syncthreads();
  const i64 s = (i64)blockIdx.x * TWG + t;
  const bool valid = s < n;
  float x[NF], h1[H], h2[H];
  float y = 0.f;
  if (valid) {
    const i64 r = idx ? (i64)idx[s] : s;
#pragma unroll
    for (int i = 0; i < NF; ++i) x[i] = feat[r * NF + i];
    y = target[r];
  } else {
#pragma unroll
    for (int i = 0; i < NF; ++i) x[i] = 0.f;
  }
  const float o = mlp_fwd(sp, x, h1, h2);
  const float diff = valid ? (o - y) : 0.f;
  const float g = 2.f * diff * inv_n;
  float lsum = block_reduce_sum<TWG>(diff * diff * inv_n, s_red);
  if (t == 0) atomicAdd(loss, lsum);
  if (!want_grad) return;
  float d2[H];
#pragma unroll
  for (int j = 0; j < H; ++j) d2[j] = h2[j] > 0.f ? g * sp[OW3 + j] : 0.f;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < H; ++j) acc = fmaf(d2[j], sp[OW2 + j * H + i], acc);
    sD1[t * LDH + i] = h1[i] > 0.f ? acc : 0.f;
    sH1[t * LDH + i] = h1[i];
    sH2[t * LDH + i] = h2[i];
    sD2[t * LDH + i] = d2[i];
  }
#pragma unroll
  for (int i = 0; i < NF; ++i) sX[t * LDX + i] = x[i];
  sG[t] = g;
  __syncthreads();
  for (int q = t; q < NPARAM; q += TWG) {
    float acc = 0.f;
    if (q < OB1) {
      const int j = (q - OW1) / NF, i = (q - OW1) % NF;
      for (int r = 0; r < TWG; ++r) acc = fmaf(sD1[r * LDH + j], sX[r * LDX + i], acc);
    } else if (q < OW2) {
      const int j = q - OB1;
      for (int r = 0; r < TWG; ++r) acc += sD1[r * LDH + j];
    } else if (q < OB2) {
      const int j = (q - OW2) / H, i = (q - OW2) % H;
      for (int r = 0; r < TWG; ++r) acc = fmaf(sD2[r * LDH + j], sH1[r * LDH + i], acc);
    } else if (q < OW3) {
      const int j = q - OB2;
      for (int r = 0; r < TWG; ++r) acc += sD2[r * LDH + j];
    } else if (q < OB3) {
      const int j = q - OW3;
      for (int r = 0; r < TWG; ++r) acc = fmaf(sG[r], sH2[r * LDH + j], acc);
    } else {
      for (int r = 0; r < TWG; ++r) acc += sG[r];
    }
    atomicAdd(&grad[q], acc);
  }
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m1, float* __restrict__ m2,
                                              int n, float lr, float b1, float b2, float eps,
                                              float bc1, float bc2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
}

__global__ __launch_bounds__(256) void k_mlp_predict(const float* __restrict__ params,
                                                     const float* __restrict__ feat, i64 n,
                                                     float* __restrict__ out) {
  __shared__ float sp[NPARAM];
  for (int q = threadIdx.x; q < NPARAM; q += 256) sp[q] = params[q];
  __syncthreads();
  const i64 s = (i64)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  float x[NF], h1[H], h2[H];
#pragma unroll
  for (int i = 0; i < NF; ++i) x[i] = feat[s * NF + i];
  out[s] = mlp_fwd(sp, x, h1, h2);
}

// ---- capacities, key selection, copy ---------------------------------------
// cap = clamp(ceil(2^(pred + margin) - 1), 1, U) for every row with products,
// then raised to the upper bound of its row bin (bt[b], the table a row of
// that bin gets holds twice that), still at most U: the slack is free, so a
// row overflows only when it outgrows the table it is given.  Rows with
// U <= th go to the LDS-hash learned kernels (key_light); longer rows go to
// the learned dense-tile kernel (key_heavy) when learned_heavy is set, else
// to the exact path (key_exact = P, no capacity).
struct BinTop {
  i64 t[5];
};

__global__ __launch_bounds__(256) void k_capacity(i64 m, const i64* __restrict__ prod, i64 ncols,
                                                  const float* __restrict__ pred, float margin,
                                                  BinTop bt, i64 th, int learned_heavy,
                                                  int bins_mask, i64* __restrict__ cap,
                                                  i64* __restrict__ key_light,
                                                  i64* __restrict__ key_heavy,
                                                  i64* __restrict__ key_exact,
                                                  i32* __restrict__ ovf) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const i64 P = prod[i];
  const i64 U = min(P, ncols);
  i64 c = 0;
  int bin = 5;
  if (P > 0) {
    const float e = fminf(pred[i] + margin, 40.f);
    c = (i64)ceilf(exp2f(e) - 1.f);
    c = max((i64)1, min(c, U));
#pragma unroll
    for (int b = 0; b < 5; ++b)
      if (c <= bt.t[b]) {
        c = min(bt.t[b], U);
        bin = b;
        break;
      }
  }
  i64 ke = 0;
  // Long rows on the exact path, and rows predicted into a bin the tuning
  // leaves to the exact path.
  if (P > 0 && ((U > th && !learned_heavy) || (U <= th && (bin > 4 || !((bins_mask >> bin) & 1))))) {
    ke = P;
    c = 0;
  }
  cap[i] = c;
  key_light[i] = U <= th ? c : 0;
  key_heavy[i] = U > th ? c : 0;
  key_exact[i] = ke;
  ovf[i] = 0;
}

// Overflowed rows join the exact set; counts predicted and overflowed rows.
__global__ __launch_bounds__(256) void k_merge_overflow(i64 m, const i64* __restrict__ prod,
                                                        const i64* __restrict__ cap,
                                                        const i32* __restrict__ ovf,
                                                        i64* __restrict__ key_exact,
                                                        unsigned long long* __restrict__ counts) {
  __shared__ i64 s_scr[4];
  i64 pr = 0, ov = 0;
  // Grid-stride over a capped grid: each workgroup's two counts cost one
  // atomic each on a shared cache line.
  for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < m; i += (i64)gridDim.x * 256) {
    const bool o = ovf[i] != 0;
    pr += cap[i] > 0;
    ov += o;
    if (o) key_exact[i] = prod[i];
  }
  pr = block_reduce_sum<256>(pr, s_scr);
  ov = block_reduce_sum<256>(ov, s_scr);
  if (threadIdx.x == 0) {
    if (pr) atomicAdd(&counts[0], (unsigned long long)pr);
    if (ov) atomicAdd(&counts[1], (unsigned long long)ov);
  }
}

// Copy the rows the learned pass completed from the temporary buffer into C:
// CR lanes per row (a predicted row has at most a hash table's worth of
// entries, mostly few), two coalesced loads per array in flight per lane.
constexpr int CR = 16;
__global__ __launch_bounds__(256) void k_copy_rows(i64 m, const i64* __restrict__ cap,
                                                   const i32* __restrict__ ovf,
                                                   const i64* __restrict__ tmp_off,
                                                   const i32* __restrict__ tcol,
                                                   const float* __restrict__ tval,
                                                   const i64* __restrict__ crp,
                                                   i32* __restrict__ ccol,
                                                   float* __restrict__ cval) {
  constexpr int U = 2;
  const i64 row = ((i64)blockIdx.x * 256 + threadIdx.x) / CR;
  const int l = threadIdx.x & (CR - 1);
  if (row >= m || cap[row] == 0 || ovf[row]) return;
  const i64 src = tmp_off[row], dst = crp[row], n = crp[row + 1] - dst;
  for (i64 q0 = l; q0 < n; q0 += U * CR) {
    i32 c[U];
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const i64 q = q0 + u * CR;
      if (q < n) {
        c[u] = tcol[src + q];
        v[u] = tval[src + q];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const i64 q = q0 + u * CR;
      if (q < n) {
        ccol[dst + q] = c[u];
        cval[dst + q] = v[u];
      }
    }
  }
}

// Numeric keys: exact rows use their exact nnz, the others are skipped.
__global__ __launch_bounds__(256) void k_numeric_keys(i64 m, const i64* __restrict__ key_exact,
                                                      const i64* __restrict__ nnz,
                                                      i64* __restrict__ key_num) {
  const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i < m) key_num[i] = key_exact[i] > 0 ? nnz[i] : 0;
}

// ---- predictor state ----------------------------------------------------------
namespace {
struct MlpState {
  std::mutex mu;
  std::vector<float> params;
  float margin = 0.f;
  bool loaded = false;
  bool tried_file = false;
};
MlpState& mlp() {
  static MlpState* s = new MlpState();
  return *s;
}

std::string lib_dir() {
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&mlp), &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    const size_t k = p.rfind('/');
    return k == std::string::npos ? std::string(".") : p.substr(0, k);
  }
  return ".";
}

// File format: "RTMLP001", int32 count, float margin, float params[count].
bool load_weights_file(const std::string& path, std::vector<float>& p, float& margin) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  char magic[8];
  int32_t n = 0;
  f.read(magic, 8);
  f.read(reinterpret_cast<char*>(&n), 4);
  f.read(reinterpret_cast<char*>(&margin), 4);
  if (!f || std::memcmp(magic, "RTMLP001", 8) != 0 || n != NPARAM) return false;
  p.resize(n);
  f.read(reinterpret_cast<char*>(p.data()), n * 4);
  return (bool)f;
}

// Current weights (loading rtspmspm_amd/weights/rownnz_mlp.bin on first use).
void current_weights(std::vector<float>& p, float& margin) {
  MlpState& S = mlp();
  std::lock_guard<std::mutex> g(S.mu);
  if (!S.loaded && !S.tried_file) {
    S.tried_file = true;
    std::vector<float> q;
    float mg = 0;
    if (load_weights_file(lib_dir() + "/weights/rownnz_mlp.bin", q, mg)) {
      S.params = q;
      S.margin = mg;
      S.loaded = true;
    }
  }
  RT_CHECK(S.loaded, RT_E_INVALID,
           "learned mode needs predictor weights: rt_mlp_set_weights or weights/rownnz_mlp.bin");
  p = S.params;
  margin = S.margin;
}

template <class T>
void upload(DBuf<T>& d, const T* h, size_t n, hipStream_t s) {
  d.alloc(std::max<size_t>(n, 1));
  if (n) RT_HIP(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}

// Rows whose product bound min(P_i, cols) exceeds skip_u or is at most
// skip_lo get zero features (their prediction is not used: such rows run the
// exact path).
void features_dev(const Csr& A, const Csr& B, const i64* bbase, const i64* prod, float* feat,
                  hipStream_t s, i64 skip_u = LLONG_MAX, i64 skip_lo = -1) {
  if (A.rows == 0) return;
  const dim3 grid((unsigned)((A.rows * 16 + 255) / 256));
  ProfScope ps("row_features", s);
  if (bbase)
    hipLaunchKernelGGL(k_features<true>, grid, dim3(256), 0, s, A.rows, A.rp, A.col, B.rp, B.col,
                       bbase, B.cols, prod, skip_lo, skip_u, feat);
  else
    hipLaunchKernelGGL(k_features<false>, grid, dim3(256), 0, s, A.rows, A.rp, A.col, B.rp, B.col,
                       nullptr, B.cols, prod, skip_lo, skip_u, feat);
  RT_HIP(hipGetLastError());
}

void predict_dev(const float* params, const float* feat, i64 n, float* out, hipStream_t s) {
  if (n == 0) return;
  ProfScope ps("mlp_predict", s);
  hipLaunchKernelGGL(k_mlp_predict, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, params,
                     feat, n, out);
  RT_HIP(hipGetLastError());
}

void grad_dev(const float* params, const float* feat, const float* target, const i32* idx, i64 n,
              float inv_n, float* grad, float* loss, int want_grad, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)((n + TWG - 1) / TWG)), dim3(TWG), 0, s, params,
                     feat, target, idx, n, inv_n, grad, loss, want_grad);
  RT_HIP(hipGetLastError());
}

}  // namespace

i64 predicted_floor(const Tuning& tu) {
  int b0 = 0;
  while (b0 < 5 && !((tu.learned_bins >> b0) & 1)) ++b0;
  return b0 == 0 ? -1 : tu.num[b0 - 1];
}

void spgemm_learned(const Csr& A, const Csr& B, const Batch& batch, CsrOut& C,
                    const Tuning& tu, hipStream_t s, RunStats* st) {
  const i64 m = A.rows;
  RT_CHECK(m < (i64)INT32_MAX, RT_E_LIMIT, "A has too many rows for 32-bit row lists");
  if (tu.learned_bins == 0 && !tu.learned_heavy) {
    // The tuning predicts no row: every row takes the exact path.
    if (batch.b_base) spgemm_plain_graph(A, B, batch, C, tu, s, st);
    else spgemm_plain(A, B, batch, C, tu, s, st);
    return;
  }
  std::vector<float> hp;
  float margin = 0;
  current_weights(hp, margin);
  const i64* bbase = batch.b_base;
  const i64 M = std::max<i64>(m, 1);
  DBuf<float> params;
  upload(params, hp.data(), hp.size(), s);
  DBuf<i64> prod(M), cap(M), key_light(M), key_heavy(M), key_exact(M), nnz(M), tmp_off(m + 1),
      key_num(M);
  DBuf<i32> ovf(M);
  DBuf<float> feat(M * NF), pred(M);
  RunStats local;
  RunStats& S = st ? *st : local;
  row_products(A, B, bbase, prod.p, s, &S);
  RT_HIP(hipMemsetAsync(nnz.p, 0, M * sizeof(i64), s));
  if (!tu.learned_heavy && std::min(S.max_row_products, B.cols) <= predicted_floor(tu)) {
    // No row's bound reaches a predicted bin: the exact path, on these
    // row products, without the learned stages.
    HeavyRows heavy;
    symbolic_pass(A, B, bbase, prod.p, nnz.p, tu, s, heavy);
    alloc_c(C, nnz.p, m, B.cols, s);
    numeric_pass(A, B, bbase, nnz.p, heavy, C, tu, s);
    return;
  }
  // Long rows go to the exact path unpredicted unless the learned long-row
  // kernel is chosen (the bound k_capacity applies).
  const i64 long_th = std::min(tu.sym[4], tu.num[4]);
  features_dev(A, B, bbase, prod.p, feat.p, s, tu.learned_heavy ? LLONG_MAX : long_th,
               predicted_floor(tu));
  predict_dev(params.p, feat.p, m, pred.p, s);
  const dim3 g256((unsigned)((M + 255) / 256));
  {
    ProfScope ps("capacity", s);
    BinTop bt;
    for (int b = 0; b < 5; ++b) bt.t[b] = tu.num[b];
    hipLaunchKernelGGL(k_capacity, g256, dim3(256), 0, s, m, prod.p, B.cols, pred.p, margin, bt,
                       std::min(tu.sym[4], tu.num[4]), tu.learned_heavy, tu.learned_bins, cap.p,
                       key_light.p,
                       key_heavy.p,
                       key_exact.p, ovf.p);
    RT_HIP(hipGetLastError());
  }
  exclusive_scan_i64(cap.p, tmp_off.p, m, s);
  const i64 tmp_total = read_scalar(tmp_off.p + m, s);
  DBuf<i32> tcol(std::max<i64>(tmp_total, 1));
  DBuf<float> tval(std::max<i64>(tmp_total, 1));
  learned_numeric_pass(A, B, bbase, key_light.p, cap.p, tmp_off.p, tcol.p, tval.p, nnz.p, ovf.p,
                       tu, s);
  learned_heavy_pass(A, B, bbase, key_heavy.p, cap.p, tmp_off.p, tcol.p, tval.p, nnz.p, ovf.p, s);
  DBuf<unsigned long long> counts(2);
  RT_HIP(hipMemsetAsync(counts.p, 0, 2 * sizeof(unsigned long long), s));
  {
    ProfScope ps("merge_overflow", s);
    hipLaunchKernelGGL(k_merge_overflow, g256, dim3(256), 0, s, m, prod.p, cap.p, ovf.p,
                       key_exact.p, counts.p);
    RT_HIP(hipGetLastError());
  }
  unsigned long long hc[2];
  RT_HIP(hipMemcpyAsync(hc, counts.p, sizeof(hc), hipMemcpyDeviceToHost, s));
  HeavyRows heavy;
  symbolic_pass(A, B, bbase, key_exact.p, nnz.p, tu, s, heavy);
  S.predicted_rows = (i64)hc[0];
  S.overflow_rows = (i64)hc[1];
  alloc_c(C, nnz.p, m, B.cols, s);
  if (m > 0) {
    {
      ProfScope ps("copy_rows", s);
      hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)((m * CR + 255) / 256)), dim3(256), 0, s, m,
                         cap.p, ovf.p, tmp_off.p, tcol.p, tval.p, C.rp, C.col, C.val);
      RT_HIP(hipGetLastError());
    }
    ProfScope ps("numeric_keys", s);
    hipLaunchKernelGGL(k_numeric_keys, g256, dim3(256), 0, s, m, key_exact.p, nnz.p, key_num.p);
    RT_HIP(hipGetLastError());
  }
  numeric_pass(A, B, bbase, key_num.p, heavy, C, tu, s);
}

void row_features(const Csr& A, const Csr& B, const Batch& batch, float* feat, const i64* prod,
                  hipStream_t s) {
  features_dev(A, B, batch.b_base, prod, feat, s);
}

// ---- launch wrappers for the graph-captured learned path ----------------------
// Kernels only: no allocation, no host synchronisation, no memset, so that a
// stream capture records them (spgemm_learned_graph, k_spgemm.hip).
void learned_weights(std::vector<float>& params, float& margin) { current_weights(params, margin); }

void learned_predict_launch(const Csr& A, const Csr& B, const i64* bbase, const i64* prod,
                            const float* params, float* feat, float* pred, i64 skip_u,
                            i64 skip_lo, hipStream_t s) {
  features_dev(A, B, bbase, prod, feat, s, skip_u, skip_lo);
  predict_dev(params, feat, A.rows, pred, s);
}

void learned_capacity_launch(i64 m, const i64* prod, i64 ncols, const float* pred, float margin,
                             const Tuning& tu, i64* cap, i64* key_light, i64* key_heavy,
                             i64* key_exact, i32* ovf, hipStream_t s) {
  if (m <= 0) return;
  BinTop bt;
  for (int b = 0; b < 5; ++b) bt.t[b] = tu.num[b];
  hipLaunchKernelGGL(k_capacity, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, prod,
                     ncols, pred, margin, bt, std::min(tu.sym[4], tu.num[4]), tu.learned_heavy,
                     tu.learned_bins, cap, key_light, key_heavy, key_exact, ovf);
  RT_HIP(hipGetLastError());
}

void learned_merge_launch(i64 m, const i64* prod, const i64* cap, const i32* ovf, i64* key_exact,
                          unsigned long long* counts, hipStream_t s) {
  if (m <= 0) return;
  const unsigned grid =
      (unsigned)std::max<i64>(1, std::min<i64>((m + 255) / 256, (i64)device().cus * 8));
  hipLaunchKernelGGL(k_merge_overflow, dim3(grid), dim3(256), 0, s, m, prod, cap, ovf, key_exact,
                     counts);
  RT_HIP(hipGetLastError());
}

void learned_finish_launch(i64 m, const i64* cap, const i32* ovf, const i64* tmp_off,
                           const i32* tcol, const float* tval, const i64* key_exact,
                           const i64* nnz, i64* key_num, const CsrOut& C, hipStream_t s) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)((m * CR + 255) / 256)), dim3(256), 0, s, m, cap,
                     ovf, tmp_off, tcol, tval, C.rp, C.col, C.val);
  hipLaunchKernelGGL(k_numeric_keys, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m,
                     key_exact, nnz, key_num);
  RT_HIP(hipGetLastError());
}

// ---- training ------------------------------------------------------------------
namespace {
std::vector<float> init_params(uint64_t seed) {
  std::mt19937_64 g(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  std::vector<float> p(NPARAM, 0.f);
  for (int q = OW1; q < OB1; ++q) p[q] = nd(g) * std::sqrt(2.f / NF);
  for (int q = OW2; q < OB2; ++q) p[q] = nd(g) * std::sqrt(2.f / H);
  for (int q = OW3; q < OB3; ++q) p[q] = nd(g) * std::sqrt(1.f / H) * 0.1f;
  return p;
}
}  // namespace

}  // namespace rt

using namespace rt;

namespace {
template <class F>
int guard(F&& f) {
  try {
    set_last_error("");
    f();
    return RT_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return RT_E_INVALID;
  }
}
}  // namespace

extern "C" {

int64_t rt_mlp_param_count(void) { return NPARAM; }

int rt_mlp_set_weights(const float* params, int64_t n, float margin) {
  return guard([&] {
    RT_CHECK(params && n == NPARAM, RT_E_INVALID, "expected " + std::to_string(NPARAM) + " parameters");
    RT_CHECK(std::isfinite(margin), RT_E_INVALID, "margin must be finite");
    for (int64_t i = 0; i < n; ++i) RT_CHECK(std::isfinite(params[i]), RT_E_INVALID, "non-finite weight");
    MlpState& S = mlp();
    std::lock_guard<std::mutex> g(S.mu);
    S.params.assign(params, params + n);
    S.margin = margin;
    S.loaded = true;
  });
}

int rt_mlp_get_weights(float* params, int64_t n, float* margin) {
  return guard([&] {
    RT_CHECK(params && n == NPARAM, RT_E_INVALID, "expected " + std::to_string(NPARAM) + " parameters");
    std::vector<float> p;
    float mg;
    current_weights(p, mg);
    std::copy(p.begin(), p.end(), params);
    if (margin) *margin = mg;
  });
}

int rt_row_features(const rt_csr* A, const rt_csr* B, float* feat, int64_t* row_nnz) {
  return guard([&] {
    RT_CHECK(A && B && feat, RT_E_INVALID, "NULL argument");
    check_csr_header(A, "A");
    check_csr_header(B, "B");
    RT_CHECK(A->mem == RT_MEM_HOST && B->mem == RT_MEM_HOST, RT_E_INVALID, "host operands expected");
    RT_CHECK(A->cols == B->rows, RT_E_SHAPE, "shape mismatch");
    validate_host(A, "A");  // before any upload: a malformed CSR never reaches a kernel
    validate_host(B, "B");
    device();
    hipStream_t s = lib_stream();
    const i64 m = A->rows, M = std::max<i64>(m, 1);
    DBuf<i64> arp, brp;
    DBuf<i32> acol, bcol;
    DBuf<float> aval, bval;
    upload(arp, A->row_ptr, m + 1, s);
    upload(acol, A->col, A->nnz, s);
    upload(aval, A->val, A->nnz, s);
    upload(brp, B->row_ptr, B->rows + 1, s);
    upload(bcol, B->col, B->nnz, s);
    upload(bval, B->val, B->nnz, s);
    Csr a{m, A->cols, A->nnz, arp.p, acol.p, aval.p};
    Csr b{B->rows, B->cols, B->nnz, brp.p, bcol.p, bval.p};
    DBuf<i64> prod(M), nnz(M);
    DBuf<float> f(M * NF);
    row_products(a, b, nullptr, prod.p, s, nullptr);
    features_dev(a, b, nullptr, prod.p, f.p, s);
    if (m) RT_HIP(hipMemcpyAsync(feat, f.p, m * NF * sizeof(float), hipMemcpyDeviceToHost, s));
    if (row_nnz) {
      RT_HIP(hipMemsetAsync(nnz.p, 0, M * sizeof(i64), s));
      HeavyRows heavy;
      symbolic_pass(a, b, nullptr, prod.p, nnz.p, Tuning{}, s, heavy);
      if (m) RT_HIP(hipMemcpyAsync(row_nnz, nnz.p, m * sizeof(i64), hipMemcpyDeviceToHost, s));
    }
    RT_HIP(hipStreamSynchronize(s));
  });
}

int rt_mlp_loss_grad(const float* params, const float* feat, const float* target, int64_t n,
                     float* loss, float* grad) {
  return guard([&] {
    RT_CHECK(params && feat && target && n > 0, RT_E_INVALID, "bad arguments");
    device();
    hipStream_t s = lib_stream();
    DBuf<float> p, f, t, g(NPARAM), l(1);
    upload(p, params, NPARAM, s);
    upload(f, feat, (size_t)n * NF, s);
    upload(t, target, (size_t)n, s);
    RT_HIP(hipMemsetAsync(g.p, 0, NPARAM * sizeof(float), s));
    RT_HIP(hipMemsetAsync(l.p, 0, sizeof(float), s));
    grad_dev(p.p, f.p, t.p, nullptr, n, 1.f / (float)n, g.p, l.p, 1, s);
    if (grad) RT_HIP(hipMemcpyAsync(grad, g.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, s));
    float hl = 0;
    RT_HIP(hipMemcpyAsync(&hl, l.p, sizeof(float), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (loss) *loss = hl;
  });
}

int rt_mlp_predict(const float* params, const float* feat, int64_t n, float* pred) {
  return guard([&] {
    RT_CHECK(feat && pred && n >= 0, RT_E_INVALID, "bad arguments");
    device();
    hipStream_t s = lib_stream();
    std::vector<float> hp;
    float mg;
    if (params) hp.assign(params, params + NPARAM); else current_weights(hp, mg);
    DBuf<float> p, f, o(std::max<int64_t>(n, 1));
    upload(p, hp.data(), NPARAM, s);
    upload(f, feat, (size_t)n * NF, s);
    predict_dev(p.p, f.p, n, o.p, s);
    if (n) RT_HIP(hipMemcpyAsync(pred, o.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
  });
}

// Adam over shuffled mini-batches; the trained weights become the library's
// predictor, with margin = the 99th percentile of (target - prediction) on
// the training set (so about 1% of rows are expected to overflow).
int rt_mlp_train(const float* feat, const float* target, int64_t n, int32_t epochs,
                 int32_t batch, float lr, uint64_t seed, float* loss) {
  return guard([&] {
    RT_CHECK(feat && target && n > 0 && epochs > 0 && batch > 0 && lr > 0, RT_E_INVALID,
             "bad training arguments");
    device();
    hipStream_t s = lib_stream();
    std::vector<float> p0 = init_params(seed);
    DBuf<float> p, f, t, g(NPARAM), m1(NPARAM), m2(NPARAM), l(1);
    upload(p, p0.data(), NPARAM, s);
    upload(f, feat, (size_t)n * NF, s);
    upload(t, target, (size_t)n, s);
    RT_HIP(hipMemsetAsync(m1.p, 0, NPARAM * sizeof(float), s));
    RT_HIP(hipMemsetAsync(m2.p, 0, NPARAM * sizeof(float), s));
    std::vector<i32> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    DBuf<i32> didx(n);
    std::mt19937_64 rng(seed ^ 0xA5A5A5A5ull);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
      std::shuffle(perm.begin(), perm.end(), rng);
      RT_HIP(hipMemcpyAsync(didx.p, perm.data(), n * sizeof(i32), hipMemcpyHostToDevice, s));
      for (int64_t b0 = 0; b0 < n; b0 += batch) {
        const int64_t nb = std::min<int64_t>(batch, n - b0);
        RT_HIP(hipMemsetAsync(g.p, 0, NPARAM * sizeof(float), s));
        RT_HIP(hipMemsetAsync(l.p, 0, sizeof(float), s));
        grad_dev(p.p, f.p, t.p, didx.p + b0, nb, 1.f / (float)nb, g.p, l.p, 1, s);
        ++step;
        const float bc1 = 1.f - std::pow(b1, (float)step), bc2 = 1.f - std::pow(b2, (float)step);
        hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, s, p.p, g.p, m1.p,
                           m2.p, NPARAM, lr, b1, b2, eps, bc1, bc2);
        RT_HIP(hipGetLastError());
      }
      RT_HIP(hipStreamSynchronize(s));  // perm is re-uploaded next epoch
    }
    // Final loss over the whole set, predictions and the overflow margin.
    RT_HIP(hipMemsetAsync(l.p, 0, sizeof(float), s));
    grad_dev(p.p, f.p, t.p, nullptr, n, 1.f / (float)n, g.p, l.p, 0, s);
    DBuf<float> pr(n);
    predict_dev(p.p, f.p, n, pr.p, s);
    std::vector<float> hp(NPARAM), hpred(n);
    float hl = 0;
    RT_HIP(hipMemcpyAsync(hp.data(), p.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(hpred.data(), pr.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(&hl, l.p, sizeof(float), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    std::vector<float> res(n);
    for (int64_t i = 0; i < n; ++i) res[i] = target[i] - hpred[i];
    const int64_t q = std::min<int64_t>(n - 1, (int64_t)(0.99 * (double)n));
    std::nth_element(res.begin(), res.begin() + q, res.end());
    const float margin = std::max(0.f, res[q]);
    if (loss) *loss = hl;
    MlpState& S = mlp();
    std::lock_guard<std::mutex> lk(S.mu);
    S.params = hp;
    S.margin = margin;
    S.loaded = true;
  });
}

}  // extern "C"
