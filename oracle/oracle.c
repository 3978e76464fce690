/*
 * oracle.c — CPU restatement of CoLa's single-partition SpMM, used ONLY as a
 * checker: by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  Nothing in the product (cola_amd/) links, loads or calls
 * this file.  Paths cited are relative to the CoLa source tree.
 *
 * Functions and what pins each one (see tests/test_oracle.py):
 *   oracle_spmm_f64        C64 = A[lb:ub,:]·B in FP64 (weighted form of
 *                          SAG_reference_host, include/utils.cuh:44-69).
 *                          Pinned by the reference's own known answer
 *                          (B = 1, values = 1 => C[i,:] = deg(i)), the GCN
 *                          closed form and brute-force dense products.
 *   oracle_ppl0_f32        FP32 restatement of ppl0_batch_SpMM_kernel
 *                          (include/dist_spmm_ppl_batch.cuh:76-200): virtual
 *                          local rows accumulated then added, virtual remote
 *                          columns scattered, reverse order when peid & 2.
 *                          Pinned against oracle_spmm_f64 on random inputs
 *                          and by the known answer.
 *   oracle_nnz_split       include/dist_spmm_ppl_batch.cuh:29-56.  Pinned by
 *                          its invariants (pieces per row = ceil(deg/gran),
 *                          gran a multiple of 32, ptr covering [0,nnz]) and
 *                          hand-worked cases.
 *   oracle_sag_unweighted  SAG_reference_host (utils.cuh:44-69), FP32,
 *   oracle_check_result    check_result's 1e-2 relative test (utils.cuh:71-100).
 *   oracle_bf16_from_f32, oracle_fp8e4m3_from_f32, oracle_fp8e4m3_to_f32:
 *                          round-to-nearest-even conversions written from the
 *                          bfloat16 and OCP FP8 E4M3 format definitions.
 *                          Pinned by exhaustive decode of all 256 FP8 codes
 *                          against the format's closed form and by
 *                          round-trip / midpoint cases.
 *   oracle_spmm_omp_f32    OpenMP FP32 CSR SpMM over the partition's rows
 *                          (local + remote in one row loop): the CPU baseline.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* C64[i, j] = sum_k val[k] * B[col[k], j] over rows lb..ub-1 of the global
 * CSR.  B is n x N FP32 row-major (ld = ldB), C64 is (ub-lb) x N. */
void oracle_spmm_f64(const int64_t *row_ptr, const int32_t *col_idx,
                     const float *val, int64_t lb, int64_t ub, const float *B,
                     int64_t ldB, int N, double *C64) {
  for (int64_t i = lb; i < ub; ++i) {
    double *c = C64 + (i - lb) * (int64_t)N;
    for (int j = 0; j < N; ++j) c[j] = 0.0;
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
      const double a = val ? (double)val[k] : 1.0;
      const float *b = B + (int64_t)col_idx[k] * ldB;
      for (int j = 0; j < N; ++j) c[j] += a * (double)b[j];
    }
  }
}

/* Reference semantics of one partition, FP32, accumulating into C (P x dim).
 * Local virtual row v: res = sum val_l*B_p[col]; C[rowIdx_l[v]] += res.
 * Remote virtual column u (order reversed when peid & 2): the B row of global
 * column colIdx_r[u] is read from B_global, and every entry k of the column
 * adds val_r[k] * row into C[rowIdx_r[k]]. */
void oracle_ppl0_f32(float *C, const float *B_global, int dim, int nodesPerPE,
                     int peid, const int32_t *rowPtr_l, const int32_t *colIdx_l,
                     const int32_t *rowIdx_l, const float *val_l, int L,
                     const int32_t *colPtr_r, const int32_t *rowIdx_r,
                     const int32_t *colIdx_r, const float *val_r, int R) {
  const float *Bp = B_global + (int64_t)peid * nodesPerPE * dim;
  float *res = (float *)malloc(sizeof(float) * (size_t)(dim > 0 ? dim : 1));
  for (int v = 0; v < L; ++v) {
    for (int j = 0; j < dim; ++j) res[j] = 0.f;
    for (int k = rowPtr_l[v]; k < rowPtr_l[v + 1]; ++k) {
      const float a = val_l[k];
      const float *b = Bp + (int64_t)colIdx_l[k] * dim;
      for (int j = 0; j < dim; ++j) res[j] += b[j] * a;
    }
    float *c = C + (int64_t)rowIdx_l[v] * dim;
    for (int j = 0; j < dim; ++j) c[j] += res[j];
  }
  for (int t = 0; t < R; ++t) {
    const int u = (peid & 2) == 0 ? t : R - t - 1;
    const float *b = B_global + (int64_t)colIdx_r[u] * dim;
    for (int k = colPtr_r[u]; k < colPtr_r[u + 1]; ++k) {
      const float a = val_r[k];
      float *c = C + (int64_t)rowIdx_r[k] * dim;
      for (int j = 0; j < dim; ++j) c[j] += b[j] * a;
    }
  }
  free(res);
}

/* nnz_split: returns the number of virtual rows; fills target/ptr when they
 * are non-NULL (capacity n_virtual and n_virtual+1).  *gran receives the
 * granularity; nrows == 0 produces nothing. */
int64_t oracle_nnz_split(const int32_t *rowIdx, int64_t nrows,
                         const int32_t *rowPtr, int64_t nnz, int32_t *target,
                         int32_t *ptr, int32_t *gran) {
  if (nrows == 0) {
    *gran = 0;
    return 0;
  }
  const int64_t avg = nnz / nrows;
  const int64_t g = ((avg + 31) / 32) * 32;
  *gran = (int32_t)g;
  if (g == 0) return -1;
  int64_t w = 0;
  if (ptr) ptr[0] = 0;
  for (int64_t i = 0; i < nrows; ++i) {
    int64_t lb = rowPtr[i];
    const int64_t hb = rowPtr[i + 1];
    for (; lb < hb; lb += g) {
      if (target) target[w] = rowIdx[i];
      ++w;
      if (ptr) ptr[w] = (int32_t)(lb + g > hb ? hb : lb + g);
    }
  }
  return w;
}

/* Unweighted host reference over global rows [lb, ub) (FP32). */
void oracle_sag_unweighted(int64_t lb, int64_t ub, int N, const int64_t *row_ptr,
                           const int32_t *col_idx, const float *B, float *C_ref) {
  const int64_t M = ub - lb;
  memset(C_ref, 0, sizeof(float) * (size_t)(M * N));
  for (int64_t i = 0; i < M; ++i)
    for (int64_t p = row_ptr[lb + i]; p < row_ptr[lb + i + 1]; ++p) {
      const int64_t k = col_idx[p];
      for (int64_t j = 0; j < N; ++j) C_ref[i * N + j] += B[k * N + j];
    }
}

/* Number of elements with |c - c_ref| > 1e-2 * |c_ref| (0 = "passed"). */
int64_t oracle_check_result(int64_t M, int N, const float *C, const float *C_ref) {
  int64_t bad = 0;
  for (int64_t i = 0; i < M * N; ++i)
    if (fabs((double)C[i] - (double)C_ref[i]) > 1e-2 * fabs((double)C_ref[i])) ++bad;
  return bad;
}

static uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* bfloat16 = upper 16 bits of the FP32 pattern after round-to-nearest-even;
 * NaN keeps its sign and becomes the quiet NaN 0x7fc0. */
uint16_t oracle_bf16_from_f32(float f) {
  const uint32_t u = f2u(f);
  if (isnan(f)) return (uint16_t)(((u >> 16) & 0x8000u) | 0x7fc0u);
  const uint32_t lower = u & 0xffffu, upper = u >> 16;
  uint32_t r = upper;
  if (lower > 0x8000u || (lower == 0x8000u && (upper & 1u))) r = upper + 1;
  return (uint16_t)r;
}

/* E4M3 (OCP "fn" variant): value of a code from the format definition. */
float oracle_fp8e4m3_to_f32(uint8_t c) {
  const int s = c >> 7, e = (c >> 3) & 15, m = c & 7;
  float mag;
  if (e == 15 && m == 7) return NAN;
  if (e == 0) mag = ldexpf((float)m, -9);              /* m/8 * 2^(1-7) */
  else mag = ldexpf(1.0f + (float)m / 8.0f, e - 7);
  return s ? -mag : mag;
}

/* Nearest E4M3 code (ties to even mantissa), saturating to +-448; found by
 * search over the 127 non-negative finite codes, independent of any bit
 * trick. */
uint8_t oracle_fp8e4m3_from_f32(float f) {
  const uint8_t sign = signbit(f) ? 0x80 : 0;
  const double a = fabs((double)f);
  if (isnan(f) || a >= 448.0) return sign | 0x7e;
  int best = 0;
  double bd = a;
  for (int c = 1; c <= 0x7e; ++c) {
    const double v = (double)oracle_fp8e4m3_to_f32((uint8_t)c);
    const double d = fabs(v - a);
    if (d < bd || (d == bd && (c & 1) == 0 && (best & 1) == 1)) {
      bd = d;
      best = c;
    }
  }
  return sign | (uint8_t)best;
}

/* Row scale used by the FP8 storage form: max|x| / 448 (1 for a zero row). */
float oracle_fp8_row_scale(const float *x, int n) {
  float m = 0.f;
  for (int i = 0; i < n; ++i) m = fmaxf(m, fabsf(x[i]));
  return m > 0.f ? m / 448.0f : 1.0f;
}

/* OpenMP FP32 SpMM over global rows [lb, ub): C = A[lb:ub,:]·B (overwrite).
 * Returns the thread count used. */
/* One output row of the CPU baseline: c = sum_k val[k] * B[col[k], :].
 * Built for AVX-512, AVX2+FMA and plain x86-64, picked at load time for the
 * host CPU (the library is built once and runs on whichever host the GPU
 * sits in), so the dim-wide inner loop runs in the widest vectors the host
 * has. */
__attribute__((target_clones("avx512f", "avx2", "default")))
static void baseline_row_f32(const int32_t *restrict col_idx, const float *restrict val, int64_t k0,
                             int64_t k1, const float *restrict B, int64_t ldB, int N, float *restrict c) {
  for (int j = 0; j < N; ++j) c[j] = 0.f;
  for (int64_t k = k0; k < k1; ++k) {
    const float a = val[k];
    const float *restrict b = B + (int64_t)col_idx[k] * ldB;
    for (int j = 0; j < N; ++j) c[j] += a * b[j];
  }
}

int oracle_spmm_omp_f32(const int64_t *row_ptr, const int32_t *col_idx,
                        const float *val, int64_t lb, int64_t ub, const float *B,
                        int64_t ldB, int N, float *C) {
  int threads = 1;
#ifdef _OPENMP
#pragma omp parallel
  {
#pragma omp single
    threads = omp_get_num_threads();
  }
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t i = lb; i < ub; ++i)
    baseline_row_f32(col_idx, val, row_ptr[i], row_ptr[i + 1], B, ldB, N, C + (i - lb) * (int64_t)N);
  return threads;
}

/* Reference f32 -> bf16 -> f32 round trip helper for tests. */
float oracle_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }
