/*
 * asan_check.c -- runs every function of oracle.c under AddressSanitizer +
 * UndefinedBehaviorSanitizer (`make asan`; TEST INFRASTRUCTURE, like the
 * oracle itself).  Small random CSR operands with empty rows, an empty B
 * row and explicit zeros; each result is compared with a dense product.
 * Exit status 0 when every check holds.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;
typedef int32_t i32;

void orc_count(i64, i64, const i64*, const i32*, const i64*, const i32*, i64*);
void orc_fill_f64(i64, i64, const i64*, const i32*, const float*, const i64*, const i32*,
                  const float*, const i64*, i32*, double*, double*);
void orc_fill_f32(i64, i64, const i64*, const i32*, const float*, const i64*, const i32*,
                  const float*, const i64*, i32*, float*, int);
void orc_stream_f32(i64, i64, i64, const i64*, const i32*, const float*, const i64*,
                    const i32*, const float*, int, double*);
void orc_rows_f64(i64, const i64*, i64, const i64*, const i32*, const float*, const i64*,
                  const i32*, const float*, i64*, i32*, double*, double*);

static unsigned long long g_state = 88172645463325252ull;
static double urand(void) {
  g_state ^= g_state << 13;
  g_state ^= g_state >> 7;
  g_state ^= g_state << 17;
  return (double)(g_state >> 11) * (1.0 / 9007199254740992.0);
}

/* Random CSR from a dense pattern with density d; stored zeros kept. */
static void make(i64 r, i64 c, double d, i64** rp, i32** col, float** val, double* dense) {
  *rp = calloc(r + 1, sizeof(i64));
  *col = malloc((r * c + 1) * sizeof(i32));
  *val = malloc((r * c + 1) * sizeof(float));
  i64 n = 0;
  for (i64 i = 0; i < r; ++i) {
    for (i64 j = 0; j < c; ++j) {
      dense[i * c + j] = 0;
      if (i % 7 == 3 || urand() >= d) continue; /* every 7th row empty */
      const float v = urand() < 0.05 ? 0.0f : (float)(2 * urand() - 1);
      (*col)[n] = (i32)j;
      (*val)[n] = v;
      dense[i * c + j] = v;
      ++n;
    }
    (*rp)[i + 1] = n;
  }
}

int main(void) {
  const i64 m = 60, k = 45, n = 70;
  double *da = malloc(m * k * sizeof(double)), *db = malloc(k * n * sizeof(double));
  i64 *arp, *brp;
  i32 *acol, *bcol;
  float *aval, *bval;
  make(m, k, 0.15, &arp, &acol, &aval, da);
  make(k, n, 0.2, &brp, &bcol, &bval, db);
  int bad = 0;
  i64* cnt = calloc(m, sizeof(i64));
  orc_count(m, n, arp, acol, brp, bcol, cnt);
  i64* rp = calloc(m + 1, sizeof(i64));
  for (i64 i = 0; i < m; ++i) rp[i + 1] = rp[i] + cnt[i];
  const i64 nnz = rp[m];
  i32* col = malloc((nnz + 1) * sizeof(i32));
  i32* col32 = malloc((nnz + 1) * sizeof(i32));
  double* v64 = malloc((nnz + 1) * sizeof(double));
  double* ab = malloc((nnz + 1) * sizeof(double));
  float* v32 = malloc((nnz + 1) * sizeof(float));
  orc_fill_f64(m, n, arp, acol, aval, brp, bcol, bval, rp, col, v64, ab);
  orc_fill_f32(m, n, arp, acol, aval, brp, bcol, bval, rp, col32, v32, 2);
  for (i64 i = 0; i < m; ++i) {
    for (i64 e = rp[i]; e < rp[i + 1]; ++e) {
      double want = 0;
      for (i64 q = 0; q < k; ++q) want += da[i * k + q] * db[q * n + col[e]];
      if (fabs(v64[e] - want) > 1e-12 * (ab[e] + 1e-30)) bad |= 1;
      if (col32[e] != col[e] || fabs(v32[e] - want) > 1e-5 * (ab[e] + 1e-30)) bad |= 2;
      if (e > rp[i] && col[e] <= col[e - 1]) bad |= 4;
    }
  }
  double out[2];
  orc_stream_f32(0, m, n, arp, acol, aval, brp, bcol, bval, 2, out);
  if ((i64)out[0] != nnz) bad |= 8;
  i64 sel[3] = {0, 3, m - 1};
  i64 off[4] = {0, 0, 0, 0};
  orc_rows_f64(3, sel, n, arp, acol, aval, brp, bcol, bval, off, NULL, NULL, NULL);
  for (int s = 0; s < 3; ++s) off[s + 1] += off[s];
  i32* rcol = malloc((off[3] + 1) * sizeof(i32));
  double* rval = malloc((off[3] + 1) * sizeof(double));
  double* rab = malloc((off[3] + 1) * sizeof(double));
  orc_rows_f64(3, sel, n, arp, acol, aval, brp, bcol, bval, off, rcol, rval, rab);
  for (int s = 0; s < 3; ++s) {
    const i64 i = sel[s];
    if (off[s + 1] - off[s] != cnt[i]) bad |= 16;
    for (i64 t = 0; t < cnt[i] && !(bad & 16); ++t)
      if (rcol[off[s] + t] != col[rp[i] + t] || rval[off[s] + t] != v64[rp[i] + t]) bad |= 32;
  }
  free(da); free(db); free(arp); free(brp); free(acol); free(bcol); free(aval); free(bval);
  free(cnt); free(rp); free(col); free(col32); free(v64); free(ab); free(v32);
  free(rcol); free(rval); free(rab);
  if (bad) {
    fprintf(stderr, "oracle asan check failed (mask %d)\n", bad);
    return 1;
  }
  printf("oracle asan checks: ok (nnz %lld)\n", (long long)nnz);
  return 0;
}
