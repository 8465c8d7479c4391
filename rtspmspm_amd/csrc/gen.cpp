// Seeded synthetic input generators (host, OpenMP).  Every row (or R-MAT
// edge) draws from its own counter-derived random stream, so the matrices do
// not depend on the number of threads.  These replace the reference's
// SuiteSparse downloads and top-left sampling (scripts/download_dataset.sh,
// scripts/matrixSampling.py:26-41).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "hostcsr.hpp"

namespace rt {

namespace {

inline u64 splitmix(u64 x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Rng {
  u64 s;
  Rng(u64 seed, u64 stream, u64 idx) : s(splitmix(splitmix(seed * 0x632BE59BD9B4E019ull + stream) ^ splitmix(idx + 0x1234567ull))) {}
  u64 next() { return splitmix(s += 0x9E3779B97F4A7C15ull); }
  double u01() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }  // (0,1)
  float sym() { return (float)(((next() >> 40) + 0.5) * (2.0 / 16777216.0) - 1.0); }  // U(-1,1)
  double normal() {
    const double u1 = u01(), u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  i64 below(i64 n) { return (i64)(u01() * (double)n) % n; }
  i64 poisson(double lam) {  // Knuth; lam <= ~60
    const double L = std::exp(-lam);
    double p = 1.0;
    i64 k = 0;
    do {
      ++k;
      p *= u01();
    } while (p > L);
    return k - 1;
  }
};

// k distinct uniform columns of [0, n), sorted, appended to out.
void distinct_cols(Rng& r, i64 n, i64 k, i32* out) {
  if (k <= 0) return;
  if (k >= n) {
    for (i64 j = 0; j < n; ++j) out[j] = (i32)j;
    return;
  }
  if (k * 2 > n) {  // pick the excluded set instead
    std::vector<i32> ex;
    ex.reserve(n - k);
    std::vector<char> mark(n, 0);
    i64 got = 0;
    while (got < n - k) {
      const i64 c = r.below(n);
      if (!mark[c]) { mark[c] = 1; ++got; }
    }
    i64 w = 0;
    for (i64 j = 0; j < n; ++j)
      if (!mark[j]) out[w++] = (i32)j;
    return;
  }
  std::vector<i32> v;
  v.reserve(k + 8);
  while ((i64)v.size() < k) {
    const i64 need = k - (i64)v.size();
    for (i64 t = 0; t < need; ++t) v.push_back((i32)r.below(n));
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  std::memcpy(out, v.data(), k * sizeof(i32));
}

// Build a CSR whose row lengths are given; fill(row, rng, cols, vals).
template <class Fill>
void build_rows(i64 rows, i64 cols, const std::vector<i64>& len, HostCsr& out, Fill fill) {
  out.alloc(rows, cols, 0);
  out.rp[0] = 0;
  for (i64 i = 0; i < rows; ++i) out.rp[i + 1] = out.rp[i] + len[i];
  out.set_nnz(out.rp[rows]);
#pragma omp parallel for schedule(dynamic, 256)
  for (i64 i = 0; i < rows; ++i) fill(i, out.col + out.rp[i], out.val + out.rp[i], len[i]);
}

}  // namespace

void gen_uniform_k(i64 n, i64 k, u64 seed, HostCsr& out) {
  RT_CHECK(n > 0 && k >= 0 && k <= n, RT_E_INVALID, "gen_uniform_k: need 0 <= k <= n, n > 0");
  std::vector<i64> len(n, k);
  build_rows(n, n, len, out, [&](i64 i, i32* c, float* v, i64 L) {
    Rng r(seed, 0, (u64)i);
    distinct_cols(r, n, L, c);
    for (i64 t = 0; t < L; ++t) v[t] = r.sym();
  });
}

void gen_fixed_normal(i64 rows, i64 cols, i64 k, u64 seed, u64 stream, HostCsr& out) {
  RT_CHECK(rows >= 0 && cols > 0 && k >= 0 && k <= cols, RT_E_INVALID, "gen_fixed_normal: bad sizes");
  std::vector<i64> len(rows, k);
  build_rows(rows, cols, len, out, [&](i64 i, i32* c, float* v, i64 L) {
    Rng r(seed, stream * 2 + 1, (u64)i);
    distinct_cols(r, cols, L, c);
    for (i64 t = 0; t < L; ++t) v[t] = (float)r.normal();
  });
}

void gen_lognormal_rows(i64 rows, i64 cols, double mu, double sigma, i64 lo, i64 hi, u64 seed,
                        u64 stream, HostCsr& out) {
  RT_CHECK(rows >= 0 && cols > 0 && lo >= 0 && hi >= lo && hi <= cols, RT_E_INVALID,
           "gen_lognormal_rows: bad sizes");
  std::vector<i64> len(rows);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < rows; ++i) {
    Rng r(seed, stream * 2, (u64)i);
    const double x = std::exp(mu + sigma * r.normal());
    len[i] = std::min(hi, std::max(lo, (i64)std::llround(x)));
  }
  build_rows(rows, cols, len, out, [&](i64 i, i32* c, float* v, i64 L) {
    Rng r(seed, stream * 2 + 1, (u64)i);
    distinct_cols(r, cols, L, c);
    for (i64 t = 0; t < L; ++t) v[t] = (float)r.normal();
  });
}

void gen_poisson(i64 n, double lambda, u64 seed, u64 stream, HostCsr& out) {
  RT_CHECK(n > 0 && lambda >= 0 && lambda < 700, RT_E_INVALID, "gen_poisson: bad sizes");
  std::vector<i64> len(n);
  for (i64 i = 0; i < n; ++i) {
    Rng r(seed, stream * 2, (u64)i);
    len[i] = std::min<i64>(n, r.poisson(lambda));
  }
  build_rows(n, n, len, out, [&](i64 i, i32* c, float* v, i64 L) {
    Rng r(seed, stream * 2 + 1, (u64)i);
    distinct_cols(r, n, L, c);
    for (i64 t = 0; t < L; ++t) v[t] = r.sym();
  });
}

void gen_rmat(int scale, int edge_factor, double a, double b, double c, u64 seed, HostCsr& out) {
  RT_CHECK(scale >= 1 && scale <= 30 && edge_factor >= 1, RT_E_INVALID, "gen_rmat: bad sizes");
  RT_CHECK(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, RT_E_INVALID, "gen_rmat: bad probabilities");
  const i64 n = (i64)1 << scale;
  const i64 E = n * edge_factor;
  std::vector<u64> key(E);
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (i64 e = 0; e < E; ++e) {
    Rng r(seed, 0, (u64)e);
    u64 row = 0, col = 0;
    for (int lev = 0; lev < scale; ++lev) {
      const double u = r.u01();
      const u64 rb = u >= ab;                      // quadrants c, d: lower half
      const u64 cb = (u >= a && u < ab) || u >= abc;  // quadrants b, d: right half
      row = (row << 1) | rb;
      col = (col << 1) | cb;
    }
    key[e] = (row << 32) | col;
  }
  // Counting sort by row, then sort + merge duplicates inside each row.
  std::vector<i64> cnt(n + 1, 0);
  for (i64 e = 0; e < E; ++e) ++cnt[(key[e] >> 32) + 1];
  for (i64 i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<i32> cols(E);
  {
    std::vector<i64> cur(cnt.begin(), cnt.end() - 1);
    for (i64 e = 0; e < E; ++e) cols[cur[key[e] >> 32]++] = (i32)(key[e] & 0xffffffffu);
  }
  key.clear();
  key.shrink_to_fit();
  std::vector<i64> len(n);
#pragma omp parallel for schedule(dynamic, 1024)
  for (i64 i = 0; i < n; ++i) {
    i32* b0 = cols.data() + cnt[i];
    i32* b1 = cols.data() + cnt[i + 1];
    std::sort(b0, b1);
    len[i] = std::unique(b0, b1) - b0;
  }
  build_rows(n, n, len, out, [&](i64 i, i32* cc, float* v, i64 L) {
    std::memcpy(cc, cols.data() + cnt[i], L * sizeof(i32));
    Rng r(seed, 1, (u64)i);
    for (i64 t = 0; t < L; ++t) v[t] = r.sym();
  });
}

void gen_stencil27(i64 n, u64 seed, HostCsr& out) {
  RT_CHECK(n >= 1 && n <= 1024, RT_E_INVALID, "gen_stencil27: bad grid");
  const i64 N = n * n * n;
  std::vector<i64> len(N);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < N; ++i) {
    const i64 x = i % n, y = (i / n) % n, z = i / (n * n);
    const i64 sx = 1 + (x > 0) + (x < n - 1), sy = 1 + (y > 0) + (y < n - 1), sz = 1 + (z > 0) + (z < n - 1);
    len[i] = sx * sy * sz;
  }
  build_rows(N, N, len, out, [&](i64 i, i32* c, float* v, i64) {
    const i64 x = i % n, y = (i / n) % n, z = i / (n * n);
    Rng r(seed, 0, (u64)i);
    i64 w = 0;
    for (i64 dz = -1; dz <= 1; ++dz)
      for (i64 dy = -1; dy <= 1; ++dy)
        for (i64 dx = -1; dx <= 1; ++dx) {
          const i64 xx = x + dx, yy = y + dy, zz = z + dz;
          if (xx < 0 || yy < 0 || zz < 0 || xx >= n || yy >= n || zz >= n) continue;
          c[w] = (i32)((zz * n + yy) * n + xx);
          v[w] = (dx == 0 && dy == 0 && dz == 0) ? 26.0f : (float)(-1.0 + 0.01 * (double)r.sym());
          ++w;
        }
  });
}

void gen_aggregation(i64 n, HostCsr& P, HostCsr& R) {
  RT_CHECK(n >= 2 && n % 2 == 0 && n <= 1024, RT_E_INVALID, "gen_aggregation: n must be even");
  const i64 N = n * n * n, h = n / 2, M = h * h * h;
  std::vector<i64> one(N, 1), eight(M, 8);
  build_rows(N, M, one, P, [&](i64 i, i32* c, float* v, i64) {
    const i64 x = i % n, y = (i / n) % n, z = i / (n * n);
    c[0] = (i32)(((z / 2) * h + y / 2) * h + x / 2);
    v[0] = 1.0f;
  });
  build_rows(M, N, eight, R, [&](i64 J, i32* c, float* v, i64) {
    const i64 X = J % h, Y = (J / h) % h, Z = J / (h * h);
    i64 w = 0;
    for (i64 dz = 0; dz < 2; ++dz)
      for (i64 dy = 0; dy < 2; ++dy)
        for (i64 dx = 0; dx < 2; ++dx) {
          c[w] = (i32)(((2 * Z + dz) * n + (2 * Y + dy)) * n + (2 * X + dx));
          v[w] = 1.0f;
          ++w;
        }
  });
}

void gen_batch_poisson(i32 count, i64 nmin, i64 nmax, double lmin, double lmax, u64 seed,
                       i64* a_rows, i64* b_rows, i64* b_cols, HostCsr& A, HostCsr& B) {
  RT_CHECK(count > 0 && nmin >= 1 && nmax >= nmin && lmin >= 0 && lmax >= lmin, RT_E_INVALID,
           "gen_batch_poisson: bad arguments");
  std::vector<double> lam(count);
  for (i32 b = 0; b < count; ++b) {
    Rng r(seed, 7, (u64)b);
    const double ln = std::log((double)nmin) + r.u01() * (std::log((double)nmax) - std::log((double)nmin));
    a_rows[b] = b_rows[b] = b_cols[b] = std::min(nmax, std::max(nmin, (i64)std::llround(std::exp(ln))));
    lam[b] = lmin + r.u01() * (lmax - lmin);
  }
  // Row lengths of every pair, then one stacked CSR per operand.
  for (int op = 0; op < 2; ++op) {
    HostCsr& M = op == 0 ? A : B;
    i64 rows = 0, maxc = 0;
    std::vector<i64> start(count + 1, 0);
    for (i32 b = 0; b < count; ++b) {
      start[b + 1] = start[b] + (op == 0 ? a_rows[b] : b_rows[b]);
      maxc = std::max(maxc, b_cols[b]);
    }
    rows = start[count];
    std::vector<i64> len(rows);
#pragma omp parallel for schedule(dynamic, 4)
    for (i32 b = 0; b < count; ++b)
      for (i64 i = 0; i < start[b + 1] - start[b]; ++i) {
        Rng r(seed, (u64)(b * 4 + op * 2) + 16, (u64)i);
        len[start[b] + i] = std::min<i64>(b_cols[b], r.poisson(lam[b]));
      }
    std::vector<i32> pair_of(rows);
    for (i32 b = 0; b < count; ++b)
      for (i64 i = start[b]; i < start[b + 1]; ++i) pair_of[i] = b;
    build_rows(rows, maxc, len, M, [&](i64 gi, i32* c, float* v, i64 L) {
      const i32 b = pair_of[gi];
      const i64 i = gi - start[b];
      Rng r(seed, (u64)(b * 4 + op * 2) + 17, (u64)i);
      distinct_cols(r, b_cols[b], L, c);
      for (i64 t = 0; t < L; ++t) v[t] = r.sym();
    });
  }
}

}  // namespace rt
