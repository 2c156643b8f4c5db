/* mp_oracle.c — CPU restatement of the reference's hot-path arithmetic.  TEST INFRASTRUCTURE ONLY:
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker; never linked into
 * or called by the product library.
 *
 * Build: make -C oracle   (gcc, -ffp-contract=off: every fused multiply-add below is an explicit fma/fmaf)
 *
 * Functions and the reference code they restate (paths relative to the reference repository):
 *   or_resid_fp64  k_resid            src/mp_kernels.cuh:67-78   R = B - A X, T = |A||X| + |B|, FP64 FMA, sequential j
 *   or_resid_dd    k_resid_dd         src/mp_kernels.cuh:83-87   TwoProd via FMA + TwoSum accumulation, final renormalization
 *   or_colstats    k_colstats         src/mp_kernels.cuh:92-96   max|R|, max|X|, max R/T (0/0 -> 0, x/0 -> inf), sum R^2
 *   or_getrf       custom_blocked_lu  src/mp_core.cuh:129-150    blocked right-looking GEPP, FP32 storage, panel getf2,
 *                  laswp (src/mp_kernels.cuh:29-31), unit-lower TRSM (blocked by 32 with inverted diagonal
 *                  blocks), operands rounded to the trailing format
 *                  (src/mp_kernels.cuh:14-18), FP32 accumulation of the trailing update
 *   or_getrs       apply_inverse      src/mp_core.cuh:182-190    FP64 -> FP32 cast, FP32 substitution, cast back, scale
 *   or_round_*     __float2half / the FP16 conversion of src/mp_kernels.cuh:16; BF16 and OCP FP8-e4m3 follow the
 *                  same round-to-nearest-even rule from the format definitions.
 *   or_pack        k_f32_to_f16_2d    src/mp_kernels.cuh:14-18   one operand's conversion with its overflow / subnormal
 *                  counters, after the per-32-index-tile power-of-two scale the trailing update of or_getrf applies
 * Operation order inside the panel and the TRSM is the one the HIP kernels use (documented per function), so
 * exact-arithmetic inputs give bitwise-identical factors.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* ------------------------------------------------------------------ residuals */
void or_resid_fp64(int64_t n, int nrhs, int a_bits, const void* Av, int64_t lda, const double* X, int64_t ldx,
                   const double* B, int64_t ldb, double* R, double* T, int64_t ldr) {
  for (int c = 0; c < nrhs; c++) {
    for (int64_t i = 0; i < n; i++) {
      double s = 0.0, t = 0.0;
      for (int64_t j = 0; j < n; j++) {
        const double a = (a_bits == 32) ? (double)((const float*)Av)[i + j * lda] : ((const double*)Av)[i + j * lda];
        const double x = X[j + c * ldx];
        s = fma(a, x, s);
        t = fma(fabs(a), fabs(x), t);
      }
      const double bv = B ? B[i + c * ldb] : 0.0;
      R[i + c * ldr] = bv - s;
      if (T) T[i + c * ldr] = t + fabs(bv);
    }
  }
}

void or_resid_dd(int64_t n, int nrhs, int a_bits, const void* Av, int64_t lda, const double* X, int64_t ldx,
                 const double* B, int64_t ldb, double* Rhi, double* Rlo, int64_t ldr) {
  for (int c = 0; c < nrhs; c++) {
    for (int64_t i = 0; i < n; i++) {
      double sh = B ? B[i + c * ldb] : 0.0, sl = 0.0;
      for (int64_t j = 0; j < n; j++) {
        const double a = (a_bits == 32) ? (double)((const float*)Av)[i + j * lda] : ((const double*)Av)[i + j * lda];
        const double x = X[j + c * ldx];
        const double p = -a * x;
        const double pe = fma(-a, x, -p);
        const double s = sh + p;
        const double bb = s - sh;
        const double e = (sh - (s - bb)) + (p - bb);
        sh = s; sl += e + pe;
      }
      const double s = sh + sl;
      sl = sl - (s - sh);
      Rhi[i + c * ldr] = s; Rlo[i + c * ldr] = sl;
    }
  }
}

/* out[4c + k]: 0 max|R|, 1 max|X|, 2 max |R_i|/T_i, 3 sum R^2; worst[c] = smallest row index of the max ratio.
 * NaN propagates into the maxima (the reference's per-thread loop can lose a NaN; the restatement keeps it). */
void or_colstats(int64_t n, int nrhs, const double* R, const double* X, const double* T, int64_t ld, double* out, int32_t* worst) {
  for (int c = 0; c < nrhs; c++) {
    double m0 = 0, m1 = 0, m2 = 0, m3 = 0; int64_t wi = -1;
    for (int64_t i = 0; i < n; i++) {
      const double r = fabs(R[i + c * ld]), x = fabs(X[i + c * ld]), t = T[i + c * ld];
      double q;
      if (r == 0.0) q = 0.0; else if (t == 0.0) q = INFINITY; else q = r / t;
      if (m0 == m0 && (r != r || r > m0)) m0 = r;
      if (m1 == m1 && (x != x || x > m1)) m1 = x;
      if (wi < 0 || (m2 == m2 && (q != q || q > m2))) { m2 = q; wi = i; }
      m3 = fma(r, r, m3);
    }
    out[4 * c] = m0; out[4 * c + 1] = m1; out[4 * c + 2] = m2; out[4 * c + 3] = m3;
    worst[c] = (int32_t)wi;
  }
}

/* ------------------------------------------------------------------ operand rounding (round to nearest even) */
float or_round_fp16(float x) {
  /* IEEE binary16: 11-bit significand, emin = -14, subnormal step 2^-24, max 65504, overflow to inf */
  if (x != x) return x;
  const float a = fabsf(x);
  if (a >= 65520.0f) return copysignf(INFINITY, x);
  if (a < 6.103515625e-05f) {                 /* subnormal grid k * 2^-24 */
    const float q = rintf(a * 16777216.0f);
    return copysignf(q / 16777216.0f, x);
  }
  int e; frexpf(a, &e);                        /* a in [2^(e-1), 2^e) */
  const float ulp = ldexpf(1.0f, e - 11);
  return copysignf(rintf(a / ulp) * ulp, x);
}
float or_round_bf16(float x) {
  if (x != x) return x;
  uint32_t u; memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r; memcpy(&r, &u, 4);
  return r;
}
/* OCP FP8 E4M3: bias 7, 3-bit mantissa, min normal 2^-6, subnormal step 2^-9, max 448 (saturating), NaN only */
float or_round_e4m3(float x) {
  if (x != x) return x;
  float a = fabsf(x);
  if (a > 448.0f) a = 448.0f;
  if (a < 0.015625f) return copysignf(rintf(a * 512.0f) / 512.0f, x);
  int e; frexpf(a, &e);
  const float ulp = ldexpf(1.0f, e - 4);
  float r = rintf(a / ulp) * ulp;
  if (r > 448.0f) r = 448.0f;
  return copysignf(r, x);
}
/* formats: 0 FP16, 1 BF16, 2 FP8 e4m3, 3 FP32, 4 two-term FP16 (hi = fp16(x), lo = fp16(x - hi)) */
static float round_fmt(int fmt, float x) {
  switch (fmt) { case 0: case 4: return or_round_fp16(x); case 1: return or_round_bf16(x); case 2: return or_round_e4m3(x); default: return x; }
}
static int top_exp(int fmt) { return (fmt == 0 || fmt == 4) ? 14 : (fmt == 1 ? 14 : (fmt == 2 ? 7 : 0)); }
static float fmt_max(int fmt) { return (fmt == 0 || fmt == 4) ? 65504.0f : (fmt == 1 ? 3.3895313892515355e+38f : (fmt == 2 ? 448.0f : FLT_MAX)); }
static float fmt_minnorm(int fmt) { return (fmt == 0 || fmt == 4) ? 6.103515625e-05f : (fmt == 2 ? 0.015625f : FLT_MIN); }

/* power-of-two exponent of a 32-index tile: tile max lands in [2^TOP, 2^(TOP+1)) */
static int tile_exp(int fmt, float tmax, int tile_scale) {
  if (!tile_scale || fmt == 3 || !(tmax > 0.0f) || !(tmax <= FLT_MAX)) return 0;
  int e; frexpf(tmax, &e);
  int s = top_exp(fmt) - (e - 1);
  if (s < -120) s = -120; if (s > 120) s = 120;
  return s;
}

/* One operand's conversion: S is nidx x K, element (i, q) at S[i * K + q].  Tiles of 32 indices are scaled by
 * 2^exps[t] (tile_exp of the tile's largest |x|, NaN counting as the largest) and rounded to fmt; q_out / qlo_out
 * (same layout) receive the rounded values as FP32 (qlo: the low FP16 term of format 4, else 0); counts[0] /
 * counts[1] are incremented by the scaled values above the format's maximum / in its subnormal range, as the
 * reference's pack counts them (src/mp_kernels.cuh:17) and as or_getrf does per trailing update. */
void or_pack(int fmt, int tile_scale, int64_t nidx, int K, const float* S, float* q_out, float* qlo_out, int* exps,
             unsigned long long* counts) {
  for (int64_t t = 0; t * 32 < nidx; t++) {
    float m = 0.0f;
    for (int64_t i = t * 32; i < t * 32 + 32 && i < nidx; i++)
      for (int q = 0; q < K; q++) { const float v = fabsf(S[i * K + q]); if (v != v || v > m) m = v; }
    exps[t] = tile_exp(fmt, m, tile_scale);
    const float sc = ldexpf(1.0f, exps[t]);
    for (int64_t i = t * 32; i < t * 32 + 32 && i < nidx; i++)
      for (int q = 0; q < K; q++) {
        const float v = S[i * K + q] * sc;
        const float av = fabsf(v);
        if (av > fmt_max(fmt)) counts[0]++; else if (av > 0.0f && av < fmt_minnorm(fmt)) counts[1]++;
        const float hv = round_fmt(fmt, v);
        q_out[i * K + q] = hv;
        qlo_out[i * K + q] = (fmt == 4) ? or_round_fp16(v - hv) : 0.0f;
      }
  }
}

/* ------------------------------------------------------------------ blocked LU
 * A: n x n FP32 column-major (lda), factored in place.  ipiv: 1-based LAPACK pivots.  perm: composed permutation.
 * Panel step j (global column k+j): pivot p = first row of max |a| among rows >= k+j (NaN counts as +inf);
 * swap rows inside the panel; l = a * (1/piv) when |piv| >= FLT_MIN else a / piv; a_ic = fmaf(-l, u_jc, a_ic).
 * Then the kb interchanges are applied to all other columns and to perm, U12 = L11^{-1} U12 blocked by 32
 * rows (each 32x32 diagonal block inverted by column-wise substitution and applied with fmaf in increasing
 * column order, then the rank-32 update of the rows below), and A22 -= L21 U12 with both operands rounded to the
 * format after the per-32-row (L21) / per-32-column (U12) power-of-two scaling; the products are summed in
 * double, rounded once to FP32 (acc) and applied as fmaf(-acc, 2^-(sL+sU), a).
 * counts[0] / counts[1] = scaled operands above the format maximum / nonzero below its smallest normal. */
int or_getrf(int64_t n, int nb, float* A, int64_t lda, int64_t* ipiv, int32_t* perm, int fmt, int tile_scale,
             unsigned long long* counts) {
  int info = 0;
  for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
  if (counts) { counts[0] = 0; counts[1] = 0; }
  float* qL = NULL; float* qU = NULL; float* qLl = NULL; float* qUl = NULL; int* eL = NULL; int* eU = NULL;
  for (int64_t k = 0; k < n; k += nb) {
    const int kb = (int)((n - k < nb) ? (n - k) : nb);
    for (int j = 0; j < kb; j++) {
      const int64_t col = k + j;
      int64_t p = -1; float best = -1.0f;
      for (int64_t i = col; i < n; i++) {
        float v = fabsf(A[i + col * lda]);
        if (v != v) v = INFINITY;
        if (p < 0 || v > best) { best = v; p = i; }
      }
      if (p < 0) p = col;
      ipiv[col] = p + 1;
      if (p != col)
        for (int c = 0; c < kb; c++) { const float t = A[col + (k + c) * lda]; A[col + (k + c) * lda] = A[p + (k + c) * lda]; A[p + (k + c) * lda] = t; }
      const float piv = A[col + col * lda];
      if (piv == 0.0f && info == 0) info = (int)(col + 1);
      if (piv != 0.0f) {
        const float rp = 1.0f / piv;
        for (int64_t i = col + 1; i < n; i++) {
          const float x = A[i + col * lda];
          const float l = (fabsf(piv) >= FLT_MIN) ? x * rp : x / piv;
          A[i + col * lda] = l;
          for (int c = j + 1; c < kb; c++) A[i + (k + c) * lda] = fmaf(-l, A[col + (k + c) * lda], A[i + (k + c) * lda]);
        }
      }
    }
    /* interchanges on the other columns and on perm (sequential LAPACK semantics) */
    for (int j = 0; j < kb; j++) {
      const int64_t a = k + j, b = ipiv[k + j] - 1;
      if (a == b) continue;
      for (int64_t c = 0; c < n; c++) {
        if (c >= k && c < k + kb) continue;
        const float t = A[a + c * lda]; A[a + c * lda] = A[b + c * lda]; A[b + c * lda] = t;
      }
      const int32_t t = perm[a]; perm[a] = perm[b]; perm[b] = t;
    }
    const int64_t m2 = n - k - kb;
    if (m2 <= 0) continue;
    /* U12 <- L11^{-1} U12, blocked by 32 rows: T_b^{-1} (column-wise substitution) applied to the block,
       then the rank-32 update of the rows below (the HIP kernels' operation order) */
    for (int rb = 0; rb < kb; rb += 32) {
      const int rib = (kb - rb < 32) ? (kb - rb) : 32;
      float Ti[32][32];
      for (int c = 0; c < 32; c++) {
        float x[32];
        for (int r = 0; r < 32; r++) x[r] = (r == c) ? 1.0f : 0.0f;
        for (int r = 1; r < 32; r++) {
          float acc = x[r];
          for (int q = c; q < r; q++) {
            const float T = (r < rib && q < rib) ? A[(k + rb + r) + (k + rb + q) * lda] : 0.0f;
            acc = fmaf(-T, x[q], acc);
          }
          if (r > c) x[r] = acc;
        }
        for (int r = 0; r < 32; r++) Ti[r][c] = (r < rib && c < rib) ? x[r] : (r == c ? 1.0f : 0.0f);
      }
      for (int64_t c = k + kb; c < n; c++) {
        float nv[32];
        for (int r = 0; r < rib; r++) {
          float acc = A[(k + rb + r) + c * lda];
          for (int q = 0; q < r; q++) acc = fmaf(Ti[r][q], A[(k + rb + q) + c * lda], acc);
          nv[r] = acc;
        }
        for (int r = 0; r < rib; r++) A[(k + rb + r) + c * lda] = nv[r];
        for (int r = rb + 32; r < kb; r++) {
          float acc = A[(k + r) + c * lda];
          for (int q = 0; q < 32; q++) {
            const float L = (q < rib) ? A[(k + r) + (k + rb + q) * lda] : 0.0f;
            acc = fmaf(-L, A[(k + rb + q) + c * lda], acc);
          }
          A[(k + r) + c * lda] = acc;
        }
      }
    }
    /* operand rounding with per-tile power-of-two scales */
    const int64_t nt = (m2 + 31) / 32;
    qL = (float*)realloc(qL, sizeof(float) * (size_t)m2 * kb);
    qU = (float*)realloc(qU, sizeof(float) * (size_t)m2 * kb);
    qLl = (float*)realloc(qLl, sizeof(float) * (size_t)m2 * kb);
    qUl = (float*)realloc(qUl, sizeof(float) * (size_t)m2 * kb);
    eL = (int*)realloc(eL, sizeof(int) * (size_t)nt);
    eU = (int*)realloc(eU, sizeof(int) * (size_t)nt);
    for (int64_t t = 0; t < nt; t++) {
      float mL = 0.0f, mU = 0.0f;
      for (int64_t ii = t * 32; ii < t * 32 + 32 && ii < m2; ii++)
        for (int q = 0; q < kb; q++) {
          const float vl = fabsf(A[(k + kb + ii) + (k + q) * lda]);
          const float vu = fabsf(A[(k + q) + (k + kb + ii) * lda]);
          if (vl != vl || vl > mL) mL = vl;
          if (vu != vu || vu > mU) mU = vu;
        }
      eL[t] = tile_exp(fmt, mL, tile_scale);
      eU[t] = tile_exp(fmt, mU, tile_scale);
      const float sL = ldexpf(1.0f, eL[t]), sU = ldexpf(1.0f, eU[t]);
      for (int64_t ii = t * 32; ii < t * 32 + 32 && ii < m2; ii++)
        for (int q = 0; q < kb; q++) {
          const float vl = A[(k + kb + ii) + (k + q) * lda] * sL;
          const float vu = A[(k + q) + (k + kb + ii) * lda] * sU;
          if (counts) {
            const float al = fabsf(vl), au = fabsf(vu);
            if (al > fmt_max(fmt)) counts[0]++; else if (al > 0.0f && al < fmt_minnorm(fmt)) counts[1]++;
            if (au > fmt_max(fmt)) counts[0]++; else if (au > 0.0f && au < fmt_minnorm(fmt)) counts[1]++;
          }
          qL[ii * kb + q] = round_fmt(fmt, vl);
          qU[ii * kb + q] = round_fmt(fmt, vu);
          qLl[ii * kb + q] = (fmt == 4) ? or_round_fp16(vl - qL[ii * kb + q]) : 0.0f;
          qUl[ii * kb + q] = (fmt == 4) ? or_round_fp16(vu - qU[ii * kb + q]) : 0.0f;
        }
    }
    for (int64_t c = 0; c < m2; c++)
      for (int64_t r = 0; r < m2; r++) {
        double s = 0.0;
        for (int q = 0; q < kb; q++) {
          s += (double)qL[r * kb + q] * (double)qU[c * kb + q];
          if (fmt == 4) s += (double)qL[r * kb + q] * (double)qUl[c * kb + q] + (double)qLl[r * kb + q] * (double)qU[c * kb + q];
        }
        const float acc = (float)s;
        const float sc = ldexpf(1.0f, -(eL[r / 32] + eU[c / 32]));
        float* cp = &A[(k + kb + r) + (k + kb + c) * lda];
        *cp = fmaf(-acc, sc, *cp);
      }
  }
  free(qL); free(qU); free(qLl); free(qUl); free(eL); free(eU);
  return info;
}

/* d = scale * U^{-1} L^{-1} P r with FP32 arithmetic (FMA, increasing / decreasing column order) */
void or_getrs(int64_t n, int nrhs, const float* LU, int64_t lda, const int32_t* perm, const double* R, int64_t ldr,
              double* D, int64_t ldd, double scale) {
  float* y = (float*)malloc(sizeof(float) * (size_t)n);
  for (int c = 0; c < nrhs; c++) {
    for (int64_t i = 0; i < n; i++) y[i] = (float)R[perm[i] + c * ldr];
    for (int64_t i = 0; i < n; i++) {
      float s = y[i];
      for (int64_t j = 0; j < i; j++) s = fmaf(-LU[i + j * lda], y[j], s);
      y[i] = s;
    }
    for (int64_t i = n - 1; i >= 0; i--) {
      float s = y[i];
      for (int64_t j = n - 1; j > i; j--) s = fmaf(-LU[i + j * lda], y[j], s);
      y[i] = s / LU[i + i * lda];
    }
    for (int64_t i = 0; i < n; i++) D[i + c * ldd] = scale * (double)y[i];
  }
  free(y);
}
