/*
 * oracle.c -- CPU restatement of the reference's SpGEMM, for TESTING ONLY.
 *
 * This file is test infrastructure.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker or as
 * the timed CPU baseline; the library (rtspmspm_amd) never calls it.
 *
 * What it restates.  The reference computes C = A*B by launching one ray per
 * nonzero A(i,k) (optixSpMSpM.cu:113-201, launch width nnz(A) at
 * optixSpMSpM.cpp:607-608).  The ray travels along x = k
 * (optixSpMSpM.cu:143-145) and so meets exactly the spheres placed at
 * (k, j, 0) for the nonzeros B(k,j) (Sphere.cpp:72-75); each hit runs the
 * any-hit program, which adds a * B(k,j) into C[i*numCol + j]
 * (optixSpMSpM.cu:245-259).  Summed over all rays that is
 *     C(i,j) = sum_k A(i,k) * B(k,j)
 * over the stored entries, i.e. Gustavson's row-by-row product.  The
 * structure of C is the union pattern of the products (the cuSPARSE
 * comparator's structure comes from its symbolic phase,
 * cuSparse/src/main.cpp:464-539), so a sum that cancels to 0 is a stored
 * entry; columns are sorted ascending within a row.
 *
 * Functions:
 *   orc_count          exact nnz of every row of C (structure only)
 *   orc_fill_f64       FP64 values plus sum |a*b| per entry (the accuracy bar)
 *   orc_fill_f32       FP32 values accumulated in A-row order, then B-row
 *                      order: the order one sequential any-hit would add them
 *   orc_stream_f32     FP32 product of a row range computed and discarded
 *                      (CPU baseline for outputs larger than host memory);
 *                      returns nnz and a checksum
 *   orc_rows_f64       FP64 result of selected rows only (large configs)
 * All index arithmetic is int64; OpenMP over rows with dynamic scheduling.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;
typedef int32_t i32;

static int cmp_i32(const void* a, const void* b) {
  const i32 x = *(const i32*)a, y = *(const i32*)b;
  return (x > y) - (x < y);
}

/* Distinct columns of row i of A*B, sorted, into cols (returns count).
 * mark[ncols] holds the stamp of the last row that touched each column. */
static i64 row_pattern(i64 i, i64 stamp, const i64* arp, const i32* acol, const i64* brp,
                       const i32* bcol, i64* mark, i32* cols) {
  i64 n = 0;
  for (i64 e = arp[i]; e < arp[i + 1]; ++e) {
    const i64 k = acol[e];
    for (i64 q = brp[k]; q < brp[k + 1]; ++q) {
      const i32 j = bcol[q];
      if (mark[j] != stamp) {
        mark[j] = stamp;
        cols[n++] = j;
      }
    }
  }
  qsort(cols, (size_t)n, sizeof(i32), cmp_i32);
  return n;
}

void orc_count(i64 rows, i64 ncols, const i64* arp, const i32* acol, const i64* brp,
               const i32* bcol, i64* row_nnz) {
#pragma omp parallel
  {
    i64* mark = (i64*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(i64));
    i32* cols = (i32*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(i32));
    for (i64 j = 0; j < ncols; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 64)
    for (i64 i = 0; i < rows; ++i) row_nnz[i] = row_pattern(i, i, arp, acol, brp, bcol, mark, cols);
    free(mark);
    free(cols);
  }
}

/* crp: row pointers from orc_count (exclusive scan). */
void orc_fill_f64(i64 rows, i64 ncols, const i64* arp, const i32* acol, const float* aval,
                  const i64* brp, const i32* bcol, const float* bval, const i64* crp,
                  i32* ccol, double* cval, double* cabs) {
#pragma omp parallel
  {
    i64* mark = (i64*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(i64));
    i32* cols = (i32*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(i32));
    double* acc = (double*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(double));
    double* aab = (double*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(double));
    for (i64 j = 0; j < ncols; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 64)
    for (i64 i = 0; i < rows; ++i) {
      const i64 n = row_pattern(i, i, arp, acol, brp, bcol, mark, cols);
      for (i64 e = arp[i]; e < arp[i + 1]; ++e) {
        const i64 k = acol[e];
        const double a = aval[e];
        for (i64 q = brp[k]; q < brp[k + 1]; ++q) {
          const double p = a * (double)bval[q];
          acc[bcol[q]] += p;
          aab[bcol[q]] += p < 0 ? -p : p;
        }
      }
      for (i64 t = 0; t < n; ++t) {
        const i32 j = cols[t];
        ccol[crp[i] + t] = j;
        cval[crp[i] + t] = acc[j];
        cabs[crp[i] + t] = aab[j];
        acc[j] = 0.0;
        aab[j] = 0.0;
      }
    }
    free(mark);
    free(cols);
    free(acc);
    free(aab);
  }
}

void orc_fill_f32(i64 rows, i64 ncols, const i64* arp, const i32* acol, const float* aval,
                  const i64* brp, const i32* bcol, const float* bval, const i64* crp,
                  i32* ccol, float* cval, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    i64* mark = (i64*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(i64));
    i32* cols = (i32*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(i32));
    float* acc = (float*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(float));
    for (i64 j = 0; j < ncols; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 64)
    for (i64 i = 0; i < rows; ++i) {
      const i64 n = row_pattern(i, i, arp, acol, brp, bcol, mark, cols);
      for (i64 e = arp[i]; e < arp[i + 1]; ++e) {
        const i64 k = acol[e];
        const float a = aval[e];
        for (i64 q = brp[k]; q < brp[k + 1]; ++q) acc[bcol[q]] += a * bval[q];
      }
      for (i64 t = 0; t < n; ++t) {
        const i32 j = cols[t];
        ccol[crp[i] + t] = j;
        cval[crp[i] + t] = acc[j];
        acc[j] = 0.f;
      }
    }
    free(mark);
    free(cols);
    free(acc);
  }
}

/* FP32 Gustavson over rows [r0, r1), emitting each row (sorted) into a
 * per-thread buffer that is then discarded.  out[0] = nnz, out[1] = checksum
 * (sum over entries of (j+1) * value, in double).  Same semantics as
 * orc_fill_f32; used as the timed CPU baseline when C exceeds host memory. */
void orc_stream_f32(i64 r0, i64 r1, i64 ncols, const i64* arp, const i32* acol,
                    const float* aval, const i64* brp, const i32* bcol, const float* bval,
                    int threads, double* out) {
  double nnz = 0, chk = 0;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel reduction(+ : nnz, chk)
  {
    i64* mark = (i64*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(i64));
    i32* cols = (i32*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(i32));
    float* acc = (float*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(float));
    float* outv = (float*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(float));
    for (i64 j = 0; j < ncols; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 16)
    for (i64 i = r0; i < r1; ++i) {
      const i64 n = row_pattern(i, i, arp, acol, brp, bcol, mark, cols);
      for (i64 e = arp[i]; e < arp[i + 1]; ++e) {
        const i64 k = acol[e];
        const float a = aval[e];
        for (i64 q = brp[k]; q < brp[k + 1]; ++q) acc[bcol[q]] += a * bval[q];
      }
      for (i64 t = 0; t < n; ++t) {
        outv[t] = acc[cols[t]];
        acc[cols[t]] = 0.f;
        chk += (double)(cols[t] + 1) * (double)outv[t];
      }
      nnz += (double)n;
    }
    free(mark);
    free(cols);
    free(acc);
    free(outv);
  }
  out[0] = nnz;
  out[1] = chk;
}

/* FP64 result of the rows listed in sel[nsel]: for row sel[s] the entries
 * go to [off[s], off[s+1]) of ccol/cval/cabs; off[] comes from a first call
 * with ccol == NULL, which writes the per-row counts into off[s+1]. */
void orc_rows_f64(i64 nsel, const i64* sel, i64 ncols, const i64* arp, const i32* acol,
                  const float* aval, const i64* brp, const i32* bcol, const float* bval,
                  i64* off, i32* ccol, double* cval, double* cabs) {
#pragma omp parallel
  {
    i64* mark = (i64*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(i64));
    i32* cols = (i32*)malloc((size_t)(ncols > 0 ? ncols : 1) * sizeof(i32));
    double* acc = (double*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(double));
    double* aab = (double*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(double));
    for (i64 j = 0; j < ncols; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 1)
    for (i64 s = 0; s < nsel; ++s) {
      const i64 i = sel[s];
      const i64 n = row_pattern(i, s, arp, acol, brp, bcol, mark, cols);
      if (!ccol) {
        off[s + 1] = n;
        continue;
      }
      for (i64 e = arp[i]; e < arp[i + 1]; ++e) {
        const i64 k = acol[e];
        const double a = aval[e];
        for (i64 q = brp[k]; q < brp[k + 1]; ++q) {
          const double p = a * (double)bval[q];
          acc[bcol[q]] += p;
          aab[bcol[q]] += p < 0 ? -p : p;
        }
      }
      for (i64 t = 0; t < n; ++t) {
        const i32 j = cols[t];
        ccol[off[s] + t] = j;
        cval[off[s] + t] = acc[j];
        cabs[off[s] + t] = aab[j];
        acc[j] = 0.0;
        aab[j] = 0.0;
      }
    }
    free(mark);
    free(cols);
    free(acc);
    free(aab);
  }
}
