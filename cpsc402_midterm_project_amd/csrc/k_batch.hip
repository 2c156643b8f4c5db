// k_batch.hip — many independent small dense systems (n <= 512, up to 16 right-hand sides each) solved to FP64
// accuracy by one persistent launch.
//
// The reference solves one system per process (custom_blocked_lu + Engine::refine_stage, src/mp_core.cuh:129-150,
// 281-296); this kernel runs the same method on every system of a batch: FP32 factorization with partial
// pivoting (blocked by 32, FP32 panel, unit-lower TRSM through the inverted diagonal block, trailing update with
// products summed in FP64 and rounded once to FP32, the reference's cublasSgemm format), then stationary
// refinement with FP64 residuals and FP32 corrections until the normwise backward error reaches tau.
//
// One workgroup owns one system at a time and takes the next one from an atomic queue in decreasing order of
// n^3 (longest-processing-time first), so uneven sizes balance over the CUs without a second launch.  Everything a
// system needs stays inside its workgroup:
//   * the factor lives in a workspace slice as row-major FP32 in ORIGIN row order: rows never move, the panel
//     tracks each row's LAPACK position (rowat[position] = origin row), so no interchange pass exists; at the
//     end the factor is written column-major in LAPACK row order (LUc) for the substitutions;
//   * the panel keeps two rows per thread in registers; one barrier per column: every wave publishes its best
//     candidate (|value|, position) and that row, and all threads pick the winner among the four;
//   * TRSM and the trailing update run thread-per-column on the row-major factor (coalesced row accesses), the
//     U12 column of a thread in registers and the L21 rows staged in LDS as FP64;
//   * the triangular solves are wave-level (rows lane + 64 s, the solved value broadcast by readlane), one wave
//     per right-hand side;
//   * residuals are thread-per-row sequential FMAs over the FP64 input.
// Operation order per element equals the oracle's (FMA chains in the same index order, divisions correctly
// rounded), so each system's iterates are bitwise those of the same method run on the CPU.
#include "mp_kernels.hpp"

#include <cfloat>
#include <numeric>

namespace mpk {

namespace {

constexpr int BT = 256;
constexpr int BNMAX = 512;
constexpr int BRHS = 16;

struct BSys {
  const double* A;        // n x n, lda n
  const double* B;        // n x nrhs, ld n
  double* X;              // n x nrhs, ld n
  int64_t w_off;          // factor workspace offset in floats: row-major n x ldr, then column-major n x n
  int64_t r_off;          // residual scratch offset in doubles (n x nrhs)
  int32_t n, nrhs, idx, ldr;
};

struct BRes { int32_t terminal, outer, info, pad; double eta, omega; };

enum { BT_CONVERGED = 0, BT_NONFINITE = 1, BT_MAX_OUTER = 2, BT_STAGNATION = 3, BT_SINGULAR = 4 };

__device__ __forceinline__ double nanmax(double m, double v) { return (m != m) ? m : ((v != v || v > m) ? v : m); }

__device__ __forceinline__ double shfl_xor_d(double v, int m) {
  const int lo = __shfl_xor(__double2loint(v), m), hi = __shfl_xor(__double2hiint(v), m);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
  const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)v, m), hi = (unsigned)__shfl_xor((int)(unsigned)(v >> 32), m);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ double wave_nanmax(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = nanmax(v, shfl_xor_d(v, m));
  return v;
}

struct Lds {
  int rowat[BNMAX];                 // origin row at each LAPACK position
  float slot[2][4][32];             // per-column candidate rows of the four waves (double-buffered)
  unsigned long long key[2][4];
  float L11[32][33];
  float Ti[32][33];
  double buf[2080];                 // L21 chunk (64 x 32 FP64) / transpose tile (64 x 65 FP32) / X chunk
  double red[4][4][4];              // [wave][rhs][stat]
  double stat[BRHS][4];             // per right-hand side: rinf, xinf, omega, eta
  double binf[BRHS];
  float ys[BRHS][BNMAX];            // right-hand sides / solutions of the FP32 substitutions
  double anorm, amax;
  int misc[8];
};

// ------------------------------------------------------------------ block reductions (all threads call)
__device__ double block_nanmax(Lds& S, double v, int slot) {
  const int t = threadIdx.x, w = t >> 6;
  v = wave_nanmax(v);
  if ((t & 63) == 0) S.red[w][0][slot] = v;
  __syncthreads();
  double r = S.red[0][0][slot];
  for (int i = 1; i < 4; i++) r = nanmax(r, S.red[i][0][slot]);
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------ setup: max|A|, row sums, FP32 row-major copy
__device__ void setup(Lds& S, const double* __restrict__ A, int n, int ldr, float* __restrict__ LU, double& g) {
  const int t = threadIdx.x;
  double am = 0.0, rs = 0.0;
  for (int i = t; i < n; i += BT) {
    double s = 0.0;
#pragma unroll 16
    for (int j = 0; j < n; j++) {
      const double v = fabs(A[i + (int64_t)j * n]);
      am = nanmax(am, v);
      s += v;
    }
    rs = nanmax(rs, s);
  }
  const double amax = block_nanmax(S, am, 0);
  const double anorm = block_nanmax(S, rs, 1);
  // global scale from the maximum of the FP32 copy, fl32(amax) (rounding is monotonic); entries are rounded to
  // FP32 first and then scaled in FP32, as the convert-then-scale of src/mp_core.cuh:161,167
  g = 1.0;
  const double amax32 = (double)(float)amax;
  if (amax32 > 0.0 && amax32 <= DBL_MAX) { int e; frexp(amax32, &e); g = ldexp(1.0, -e); }
  if (t == 0) { S.anorm = anorm; S.amax = amax; }
  for (int i = t; i < n; i += BT) S.rowat[i] = i;
  // FP64 column-major -> FP32 row-major through a 64 x 65 LDS tile (coalesced on both sides)
  float* tile = (float*)S.buf;
  for (int rb = 0; rb < n; rb += 64)
    for (int cb = 0; cb < ldr; cb += 64) {
      const int rl = t & 63;
      for (int m = 0; m < 16; m++) {
        const int cl = (t >> 6) + 4 * m, r = rb + rl, c = cb + cl;
        float v = 0.0f;
        if (r < n && c < n) v = (float)A[r + (int64_t)c * n] * (float)g;
        tile[rl * 65 + cl] = v;
      }
      __syncthreads();
      const int cl = t & 63;
      for (int m = 0; m < 16; m++) {
        const int rl2 = (t >> 6) + 4 * m, r = rb + rl2, c = cb + cl;
        if (r < n && c < ldr) LU[(int64_t)r * ldr + c] = tile[rl2 * 65 + cl];
      }
      __syncthreads();
    }
}

// ------------------------------------------------------------------ panel: rows at positions [k, n), columns k..k+kb-1
__device__ __forceinline__ void load32(float (&a)[32], const float* p) {
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const float4 v = ((const float4*)p)[q];
    a[4 * q] = v.x; a[4 * q + 1] = v.y; a[4 * q + 2] = v.z; a[4 * q + 3] = v.w;
  }
}
__device__ __forceinline__ void store32(float* p, const float (&a)[32]) {
#pragma unroll
  for (int q = 0; q < 8; q++) ((float4*)p)[q] = make_float4(a[4 * q], a[4 * q + 1], a[4 * q + 2], a[4 * q + 3]);
}

__device__ __forceinline__ unsigned long long cand_key(float v, int pos) {
  v = fabsf(v);
  if (v != v) v = INFINITY;                  // NaN counts as the largest magnitude
  return ((unsigned long long)__float_as_uint(v) << 32) | (unsigned)(0xFFFF - pos);
}

// Partial-pivot LU of the panel with LAPACK's pivot choice (first position of the largest |a|).  Rows stay
// in their origin slots; pos0/pos1 follow the LAPACK positions.  Returns the first zero pivot (1-based) or 0.
__device__ int panel(Lds& S, float* __restrict__ LU, int n, int ldr, int k, int kb, int info) {
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  float a0[32], a1[32];
  int pos0 = k + t, pos1 = k + t + BT;
  const int org0 = (pos0 < n) ? S.rowat[pos0] : -1;
  const int org1 = (pos1 < n) ? S.rowat[pos1] : -1;
  // rows past n load row 0 (any valid row) and never take part
  load32(a0, LU + (int64_t)max(org0, 0) * ldr + k);
  load32(a1, LU + (int64_t)max(org1, 0) * ldr + k);
#pragma unroll
  for (int j = 0; j < 32; j++) {
    if (j < kb) {
    const int col = k + j, par = j & 1;
    const unsigned long long k0 = (org0 >= 0 && pos0 >= col) ? cand_key(a0[j], pos0) : 0ull;
    const unsigned long long k1 = (org1 >= 0 && pos1 >= col) ? cand_key(a1[j], pos1) : 0ull;
    unsigned long long wk = k0 > k1 ? k0 : k1;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) { const unsigned long long o = shfl_xor_u64(wk, m); wk = o > wk ? o : wk; }
    if (lane == 0) S.key[par][w] = wk;
    if (wk != 0ull && wk == k0) {
#pragma unroll
      for (int q = 0; q < 8; q++) ((float4*)S.slot[par][w])[q] = make_float4(a0[4 * q], a0[4 * q + 1], a0[4 * q + 2], a0[4 * q + 3]);
    } else if (wk != 0ull && wk == k1) {
#pragma unroll
      for (int q = 0; q < 8; q++) ((float4*)S.slot[par][w])[q] = make_float4(a1[4 * q], a1[4 * q + 1], a1[4 * q + 2], a1[4 * q + 3]);
    }
    __syncthreads();
    unsigned long long best = S.key[par][0]; int wb = 0;
#pragma unroll
    for (int i = 1; i < 4; i++) { const unsigned long long o = S.key[par][i]; if (o > best) { best = o; wb = i; } }
    const int p = 0xFFFF - (int)(unsigned)(best & 0xFFFFFFFFull);
    float u[32];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const float4 v = ((const float4*)S.slot[par][wb])[q];
      u[4 * q] = v.x; u[4 * q + 1] = v.y; u[4 * q + 2] = v.z; u[4 * q + 3] = v.w;
    }
    // LAPACK interchange of positions col and p
    if (pos0 == p) pos0 = col; else if (pos0 == col) pos0 = p;
    if (pos1 == p) pos1 = col; else if (pos1 == col) pos1 = p;
    const float piv = u[j];
    if (piv == 0.0f && info == 0) info = col + 1;
    if (piv != 0.0f) {
      const float rp = 1.0f / piv;
      const bool big = fabsf(piv) >= FLT_MIN;
      if (org0 >= 0 && pos0 > col) {
        const float l = big ? a0[j] * rp : a0[j] / piv;
        a0[j] = l;
#pragma unroll
        for (int c = 0; c < 32; c++) if (c > j) a0[c] = fmaf(-l, u[c], a0[c]);
      }
      if (org1 >= 0 && pos1 > col) {
        const float l = big ? a1[j] * rp : a1[j] / piv;
        a1[j] = l;
#pragma unroll
        for (int c = 0; c < 32; c++) if (c > j) a1[c] = fmaf(-l, u[c], a1[c]);
      }
    }
    }
  }
  if (org0 >= 0) { store32(LU + (int64_t)org0 * ldr + k, a0); S.rowat[pos0] = org0; }
  if (org1 >= 0) { store32(LU + (int64_t)org1 * ldr + k, a1); S.rowat[pos1] = org1; }
  if (org0 >= 0 && pos0 < k + kb) {
#pragma unroll
    for (int q = 0; q < 32; q++) S.L11[pos0 - k][q] = a0[q];
  }
  if (org1 >= 0 && pos1 < k + kb) {
#pragma unroll
    for (int q = 0; q < 32; q++) S.L11[pos1 - k][q] = a1[q];
  }
  __syncthreads();
  return info;
}

// Row offset (in floats) of the origin row at a LAPACK position that is the same for the whole wave: read once
// into a scalar register, so addresses are a scalar base plus the lane's column.
__device__ __forceinline__ int urow_off(const Lds& S, int pos, int ldr) {
  return __builtin_amdgcn_readfirstlane(S.rowat[pos]) * ldr;
}

// ------------------------------------------------------------------ TRSM + trailing update for columns [k+32, n)
// T^{-1} of the unit-lower 32 x 32 block by column-wise substitution (thread c forms column c).
__device__ void invert_l11(Lds& S) {
  const int c = threadIdx.x;
  if (c < 32) {
    float x[32];
#pragma unroll
    for (int r = 0; r < 32; r++) x[r] = (r == c) ? 1.0f : 0.0f;
#pragma unroll
    for (int r = 1; r < 32; r++) {
      float acc = x[r];
#pragma unroll
      for (int q = 0; q < 32; q++)
        if (q < r) acc = (q >= c) ? fmaf(-S.L11[r][q], x[q], acc) : acc;
      if (r > c) x[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < 32; r++) S.Ti[r][c] = x[r];
  }
  __syncthreads();
}

// U12 column c: u_r = a_r + sum_{q<r} Ti[r][q] a_q (FMA in increasing q), written back and returned.
__device__ __forceinline__ void trsm_col(Lds& S, float* __restrict__ LU, int ldr, int k, int c, float (&ud)[32], bool store) {
  float a[32];
#pragma unroll
  for (int r = 0; r < 32; r++) a[r] = LU[urow_off(S, k + r, ldr) + c];
#pragma unroll
  for (int r = 0; r < 32; r++) {
    float acc = a[r];
#pragma unroll
    for (int q = 0; q < 32; q++)
      if (q < r) acc = fmaf(S.Ti[r][q], a[q], acc);
    if (store) LU[urow_off(S, k + r, ldr) + c] = acc;
    ud[r] = acc;
    // one row's T^{-1} reads at a time (unbounded hoisting of the 496 LDS reads exhausts the registers)
    asm volatile("" ::: "memory");
  }
}

// t4 / t2: trailing sizes m2 at or below which the rows are split over 4 / 2 thread groups (per n bucket, from
// the launch's split table; the split changes which thread computes an element, never its operation order)
__device__ void update(Lds& S, float* __restrict__ LU, int n, int ldr, int k, int t4, int t2) {
  const int t = threadIdx.x;
  double* Lc = S.buf;                                   // [64 rows][32] FP64
  const int m2 = n - k - 32;
  // G thread groups of BT/G threads: each group takes every column of the pass and a share of the rows, so
  // narrow trailing matrices still keep all threads busy (each group forms its columns' U12 itself)
  const int G = (m2 <= t4) ? 4 : (m2 <= t2 ? 2 : 1);
  const int TG = BT / G, g = t / TG, tl = t % TG;
#pragma unroll 1
  for (int cb = k + 32; cb < n; cb += TG) {
    const int c = cb + tl;
    const bool has = c < n;
    double ud[32];
    {
      // every group reads the column and forms U12; group 0 writes it back once all groups have read it
      float uf[32];
      if (has) trsm_col(S, LU, ldr, k, c, uf, false);
      __syncthreads();
      if (has && g == 0) {
#pragma unroll
        for (int r = 0; r < 32; r++) LU[urow_off(S, k + r, ldr) + c] = uf[r];
      }
#pragma unroll
      for (int q = 0; q < 32; q++) ud[q] = has ? (double)uf[q] : 0.0;
    }
#pragma unroll 1
    for (int rc = k + 32; rc < n; rc += 64) {
      const int nr = min(64, n - rc);
      {
        const int r = t >> 2, q0 = (t & 3) * 8;
        if (r < nr) {
          const float* src = LU + (int64_t)S.rowat[rc + r] * ldr + k + q0;
          const float4 v0 = ((const float4*)src)[0], v1 = ((const float4*)src)[1];
          double* d = Lc + r * 32 + q0;
          d[0] = v0.x; d[1] = v0.y; d[2] = v0.z; d[3] = v0.w; d[4] = v1.x; d[5] = v1.y; d[6] = v1.z; d[7] = v1.w;
        }
      }
      __syncthreads();
      if (has) {
        // RG rows at a time, their C values loaded before the RG independent FMA chains run.  Rows past the
        // chunk are clamped to its last row: they recompute that row's value from the same operands and store
        // it again, so the group needs no branch (a conditional store lets the compiler sink each chain into
        // its branch and hoist every LDS read above them, which exhausts the registers)
        constexpr int RG = 8;
#pragma unroll 1
        for (int r0 = g * RG; r0 < nr; r0 += G * RG) {
          int ro[RG], rl[RG]; float v[RG]; double acc[RG];
#pragma unroll
          for (int i = 0; i < RG; i++) {
            rl[i] = min(r0 + i, nr - 1);
            ro[i] = urow_off(S, rc + rl[i], ldr);
            v[i] = LU[ro[i] + c];
            acc[i] = 0.0;
          }
#pragma unroll
          for (int q = 0; q < 32; q++)
#pragma unroll
            for (int i = 0; i < RG; i++) acc[i] = fma(Lc[rl[i] * 32 + q], ud[q], acc[i]);
#pragma unroll
          for (int i = 0; i < RG; i++) LU[ro[i] + c] = fmaf(-(float)acc[i], 1.0f, v[i]);
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ LAPACK-order column-major copy of the factor
__device__ void to_colmajor(Lds& S, const float* __restrict__ LU, int n, int ldr, float* __restrict__ LUc) {
  const int t = threadIdx.x;
  float* tile = (float*)S.buf;
  for (int ib = 0; ib < n; ib += 64)
    for (int jb = 0; jb < n; jb += 64) {
      const int jl = t & 63;
      for (int m = 0; m < 16; m++) {
        const int il = (t >> 6) + 4 * m, i = ib + il, j = jb + jl;
        if (i < n && j < n) tile[il * 65 + jl] = LU[urow_off(S, i, ldr) + j];   // i is the same across the wave
      }
      __syncthreads();
      const int il = t & 63;
      for (int m = 0; m < 16; m++) {
        const int jl2 = (t >> 6) + 4 * m, i = ib + il, j = jb + jl2;
        if (i < n && j < n) LUc[i + (int64_t)j * n] = tile[il * 65 + jl2];
      }
      __syncthreads();
    }
}

// ------------------------------------------------------------------ substitutions
// X[:, c] (=|+=) g * U^{-1} L^{-1} (float) src[perm, c], blocked by 64 columns.  The diagonal block is solved
// by one wave per right-hand side (row jb + lane in the lane, its 64 factor values in registers, each solved
// value broadcast by readlane); the rows beyond the block then take the block's 64 terms in one parallel pass
// (thread per row).  Every row receives its terms in increasing (forward) / decreasing (backward) column order,
// so the result equals the sequential row-oriented substitution.
__device__ void solve_all(Lds& S, const float* __restrict__ LUc, int n, int nrhs, const double* __restrict__ src,
                          double* __restrict__ X, double g, bool first) {
#pragma clang fp contract(off)
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  for (int c = 0; c < nrhs; c++)
    for (int i = t; i < n; i += BT) S.ys[c][i] = (float)src[S.rowat[i] + (int64_t)c * n];
  __syncthreads();
  // forward, unit lower
  for (int jb = 0; jb < n; jb += 64) {
    const int nbk = min(64, n - jb);
    if (w < nrhs) {
      const int i = jb + lane, ic = min(i, n - 1);
      float L[64];
#pragma unroll
      for (int jj = 0; jj < 64; jj++) L[jj] = LUc[ic + (int64_t)min(jb + jj, n - 1) * n];
      for (int c = w; c < nrhs; c += 4) {
        float y = S.ys[c][ic];
#pragma unroll
        for (int jj = 0; jj < 64; jj++) {
          if (jj < nbk) {
            const float yj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(y), jj));
            if (lane > jj) y = fmaf(-L[jj], yj, y);
          }
        }
        if (i < n) S.ys[c][i] = y;
      }
    }
    __syncthreads();
    for (int i = jb + 64 + t; i < n; i += BT)
      for (int c = 0; c < nrhs; c++) {
        float acc = S.ys[c][i];
#pragma unroll 16
        for (int jj = 0; jj < 64; jj++) acc = fmaf(-LUc[i + (int64_t)(jb + jj) * n], S.ys[c][jb + jj], acc);
        S.ys[c][i] = acc;
      }
    __syncthreads();
  }
  // backward, upper with division by the diagonal
  for (int jb = ((n - 1) / 64) * 64; jb >= 0; jb -= 64) {
    const int nbk = min(64, n - jb);
    if (w < nrhs) {
      const int i = jb + lane, ic = min(i, n - 1);
      float U[64];
#pragma unroll
      for (int jj = 0; jj < 64; jj++) U[jj] = LUc[ic + (int64_t)min(jb + jj, n - 1) * n];
      for (int c = w; c < nrhs; c += 4) {
        float y = S.ys[c][ic];
#pragma unroll
        for (int jj = 63; jj >= 0; jj--) {
          if (jj < nbk) {
            if (lane == jj) y = y / U[jj];
            const float yj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(y), jj));
            if (lane < jj) y = fmaf(-U[jj], yj, y);
          }
        }
        if (i < n) S.ys[c][i] = y;
      }
    }
    __syncthreads();
    for (int i = t; i < jb; i += BT)
      for (int c = 0; c < nrhs; c++) {
        float acc = S.ys[c][i];
        for (int jj = nbk - 1; jj >= 0; jj--) acc = fmaf(-LUc[i + (int64_t)(jb + jj) * n], S.ys[c][jb + jj], acc);
        S.ys[c][i] = acc;
      }
    __syncthreads();
  }
  for (int c = 0; c < nrhs; c++)
    for (int i = t; i < n; i += BT) {
      double* xp = X + i + (int64_t)c * n;
      const double d = g * (double)S.ys[c][i];
      *xp = first ? d : *xp + d;
    }
  __syncthreads();
}

// ------------------------------------------------------------------ residual R = B - A X, T = |A||X| + |B| and metrics
// Thread per row (rows t, t + 256), sequential FMA over j; X staged through LDS 64 rows at a time.
template <int NC>
__device__ void resid_chunk(Lds& S, const double* __restrict__ A, const double* __restrict__ B, const double* __restrict__ X,
                            double* __restrict__ R, int n, int c0, bool first) {
#pragma clang fp contract(off)
  const int t = threadIdx.x, w = t >> 6;
  double s[2][NC], tt[2][NC];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int c = 0; c < NC; c++) { s[h][c] = 0.0; tt[h][c] = 0.0; }
  const int i0 = t, i1 = t + BT;
  double* xs = S.buf;                                   // [64][4]
  for (int jb = 0; jb < n; jb += 64) {
    {
      const int jl = t & 63, c = t >> 6, j = jb + jl;
      xs[jl * 4 + c] = (j < n && c < NC) ? X[j + (int64_t)(c0 + c) * n] : 0.0;
    }
    __syncthreads();
    const int je = min(64, n - jb);
    if (i0 < n) {
#pragma unroll 8
      for (int jl = 0; jl < je; jl++) {
        const int64_t off = (int64_t)(jb + jl) * n;
        const double a0 = A[i0 + off];
        const double a1v = A[min(i1, n - 1) + off];
        const double a1 = (i1 < n) ? a1v : 0.0;
#pragma unroll
        for (int c = 0; c < NC; c++) {
          const double xv = xs[jl * 4 + c];
          s[0][c] = fma(a0, xv, s[0][c]); tt[0][c] = fma(fabs(a0), fabs(xv), tt[0][c]);
          s[1][c] = fma(a1, xv, s[1][c]); tt[1][c] = fma(fabs(a1), fabs(xv), tt[1][c]);
        }
      }
    }
    __syncthreads();
  }
  // per-row quantities and NaN-propagating maxima per right-hand side
  double m[NC][3], bmax[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) { m[c][0] = 0.0; m[c][1] = 0.0; m[c][2] = 0.0; bmax[c] = 0.0; }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int i = h ? i1 : i0;
    if (i < n) {
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const int64_t o = i + (int64_t)(c0 + c) * n;
        const double bv = B[o];
        const double r = bv - s[h][c], tv = tt[h][c] + fabs(bv);
        R[o] = r;
        const double ar = fabs(r);
        const double q = (ar == 0.0) ? 0.0 : (tv == 0.0 ? INFINITY : ar / tv);
        m[c][0] = nanmax(m[c][0], ar);
        m[c][1] = nanmax(m[c][1], fabs(X[o]));
        m[c][2] = nanmax(m[c][2], q);
        bmax[c] = nanmax(bmax[c], fabs(bv));
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const double r0 = wave_nanmax(m[c][0]), r1 = wave_nanmax(m[c][1]), r2 = wave_nanmax(m[c][2]), r3 = wave_nanmax(bmax[c]);
    if ((t & 63) == 0) { S.red[w][c][0] = r0; S.red[w][c][1] = r1; S.red[w][c][2] = r2; S.red[w][c][3] = r3; }
  }
  __syncthreads();
  if (t < NC) {
    double v[4];
    for (int q = 0; q < 4; q++) { v[q] = S.red[0][t][q]; for (int i = 1; i < 4; i++) v[q] = nanmax(v[q], S.red[i][t][q]); }
    const int c = c0 + t;
    if (first) S.binf[c] = v[3];
    const double den = S.anorm * v[1] + S.binf[c];
    const double eta = (v[0] == 0.0) ? 0.0 : (den == 0.0 ? INFINITY : v[0] / den);
    S.stat[c][0] = v[0]; S.stat[c][1] = v[1]; S.stat[c][2] = v[2]; S.stat[c][3] = eta;
  }
  __syncthreads();
}

__device__ void residual(Lds& S, const double* A, const double* B, const double* X, double* R, int n, int nrhs, bool first) {
  for (int c0 = 0; c0 < nrhs; c0 += 4) {
    switch (min(4, nrhs - c0)) {
      case 1: resid_chunk<1>(S, A, B, X, R, n, c0, first); break;
      case 2: resid_chunk<2>(S, A, B, X, R, n, c0, first); break;
      case 3: resid_chunk<3>(S, A, B, X, R, n, c0, first); break;
      default: resid_chunk<4>(S, A, B, X, R, n, c0, first); break;
    }
  }
}

__global__ void k_batch_zero(int32_t* p) { if (threadIdx.x == 0) p[0] = 0; }

struct BSplit { int t4[4], t2[4]; };   // row-split thresholds per n bucket (n - 1) / 128

__global__ __launch_bounds__(BT, 2) void k_batch(const BSys* __restrict__ sys, int count, int32_t* __restrict__ queue,
                                                 float* __restrict__ Wall, double* __restrict__ Rall,
                                                 BRes* __restrict__ res, double tau, int max_outer, int eta_only, BSplit split) {
  __shared__ Lds S;
  const int t = threadIdx.x;
  for (;;) {
    if (t == 0) S.misc[0] = atomicAdd(queue, 1);
    __syncthreads();
    // uniform by construction; readfirstlane tells the compiler, so the descriptor and every pointer derived
    // from it live in scalar registers
    const int q = __builtin_amdgcn_readfirstlane(S.misc[0]);
    __syncthreads();
    if (q >= count) return;
    const BSys d = sys[q];
    const int n = d.n, nrhs = d.nrhs, ldr = d.ldr;
    const double* A = d.A;
    const double* B = d.B;
    double* X = d.X;
    float* LU = Wall + d.w_off;
    float* LUc = LU + (int64_t)n * ldr;
    double* R = Rall + d.r_off;
    const int bucket = min(3, (n - 1) / 128);
    const int t4 = split.t4[bucket], t2 = split.t2[bucket];
    double g;
    setup(S, A, n, ldr, LU, g);
    int info = 0;
    for (int k = 0; k < n; k += 32) {
      const int kb = min(32, n - k);
      info = panel(S, LU, n, ldr, k, kb, info);
      if (n - k - kb > 0) {
        invert_l11(S);
        update(S, LU, n, ldr, k, t4, t2);
      }
    }
    BRes out; out.pad = 0; out.info = info; out.outer = 0;
    if (info != 0) {
      for (int c = 0; c < nrhs; c++)
        for (int i = t; i < n; i += BT) X[i + (int64_t)c * n] = 0.0;
      out.terminal = BT_SINGULAR; out.eta = -1.0; out.omega = -1.0;
    } else {
      to_colmajor(S, LU, n, ldr, LUc);
      solve_all(S, LUc, n, nrhs, B, X, g, true);
      double prev = INFINITY; int stall = 0, outer = 0, term;
      for (bool first = true;; first = false) {
        residual(S, A, B, X, R, n, nrhs, first);
        double eta = S.stat[0][3], omega = S.stat[0][2]; bool finite = true;
        for (int c = 0; c < nrhs; c++) {
          if (c) { eta = nanmax(eta, S.stat[c][3]); omega = nanmax(omega, S.stat[c][2]); }
          finite = finite && (S.stat[c][0] <= DBL_MAX) && (S.stat[c][1] <= DBL_MAX);
        }
        out.eta = eta; out.omega = omega;
        if (finite && eta <= tau && (eta_only || omega <= tau)) { term = BT_CONVERGED; break; }
        if (!finite) { term = BT_NONFINITE; break; }
        if (outer >= max_outer) { term = BT_MAX_OUTER; break; }
        stall = (eta > 0.9 * prev) ? stall + 1 : 0;
        if (stall >= 3) { term = BT_STAGNATION; break; }
        prev = eta;
        solve_all(S, LUc, n, nrhs, R, X, g, false);
        outer++;
      }
      out.terminal = term; out.outer = outer;
    }
    if (t == 0) res[d.idx] = out;
    __syncthreads();
  }
}

}  // namespace

int batch_max_n() { return BNMAX; }
int batch_max_rhs() { return BRHS; }

void batch_solve(int count, const int32_t* n, const int32_t* nrhs, const double* const* A, const double* const* B,
                 double* const* X, const BatchParams& p, Workspace& W, BatchResult* out, double* device_s) {
  MP_REQUIRE(count >= 0, "count must be non-negative");
  if (device_s) *device_s = 0.0;
  if (count == 0) return;
  MP_REQUIRE(n && nrhs && A && B && X && out, "NULL size/pointer/result array");
  MP_REQUIRE(p.wg_per_cu >= 1 && p.wg_per_cu <= 8, "wg_per_cu must be in [1, 8]");
  for (int i = 0; i < count; i++) {
    MP_REQUIRE(n[i] >= 1 && n[i] <= BNMAX, "every n must be in [1, 512] (larger systems go through mp_solve)");
    MP_REQUIRE(nrhs[i] >= 1 && nrhs[i] <= BRHS, "every nrhs must be in [1, 16]");
    MP_REQUIRE(A[i] && B[i] && X[i], "NULL device pointer of a system");
  }
  // longest first: n^3 dominates the work of a system
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
  std::vector<BSys> tab(count);
  int64_t woff = 0, roff = 0;
  for (int q = 0; q < count; q++) {
    const int i = order[q];
    BSys& s = tab[q];
    s.n = n[i]; s.nrhs = nrhs[i]; s.idx = i; s.ldr = (n[i] + 31) / 32 * 32;
    s.A = A[i]; s.B = B[i]; s.X = X[i];
    s.w_off = woff; woff += (int64_t)s.n * s.ldr + (int64_t)s.n * s.n;
    woff = (woff + 63) / 64 * 64;                       // 256-byte aligned slices
    s.r_off = roff; roff += (int64_t)s.n * s.nrhs;
  }
  DevTracker& tr = *W.tracker;
  W.bsys.ensure(tr, (size_t)count * sizeof(BSys), "batch table");
  W.bwork.ensure(tr, (size_t)woff, "batch factors");
  W.bresid.ensure(tr, (size_t)roff, "batch residuals");
  W.bres.ensure(tr, (size_t)count * sizeof(BRes), "batch results");
  W.bqueue.ensure(tr, 1, "batch queue");
  HIP_CK(hipMemcpyAsync(W.bsys.p, tab.data(), (size_t)count * sizeof(BSys), hipMemcpyHostToDevice, W.stream));
  const int grid = std::min(count, W.num_cus * p.wg_per_cu);
  hipEvent_t e0, e1;
  HIP_CK(hipEventCreate(&e0)); HIP_CK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(64), 0, W.stream, W.bqueue.p);
  LAUNCH_CK();
  HIP_CK(hipEventRecord(e0, W.stream));
  BSplit split;
  for (int b = 0; b < 4; b++) { split.t4[b] = p.split4[b]; split.t2[b] = p.split2[b]; }
  hipLaunchKernelGGL(k_batch, dim3(grid), dim3(BT), 0, W.stream, (const BSys*)W.bsys.p, count, W.bqueue.p,
                     W.bwork.p, W.bresid.p, (BRes*)W.bres.p, p.tau, p.max_outer, p.eta_only ? 1 : 0, split);
  LAUNCH_CK();
  HIP_CK(hipEventRecord(e1, W.stream));
  std::vector<BRes> hr(count);
  HIP_CK(hipMemcpyAsync(hr.data(), W.bres.p, (size_t)count * sizeof(BRes), hipMemcpyDeviceToHost, W.stream));
  HIP_CK(hipStreamSynchronize(W.stream));
  float ms = 0; HIP_CK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CK(hipEventDestroy(e0)); HIP_CK(hipEventDestroy(e1));
  if (device_s) *device_s = ms * 1e-3;
  for (int i = 0; i < count; i++) {
    out[i].terminal = hr[i].terminal; out[i].outer_iters = hr[i].outer; out[i].info = hr[i].info;
    out[i].eta_max = hr[i].eta; out[i].omega_max = hr[i].omega;
  }
}

// ------------------------------------------------------------------ autotuner of the launch shape
// The knobs that change only which thread or workgroup does a piece of work (never an element's operation
// order, so every candidate returns the same bits): resident workgroups per CU, and per n bucket the trailing
// sizes at which a system's update splits its rows over 2 or 4 thread groups.  The panel width (32) and the
// workgroup (4 waves) are fixed by the kernel's structure: the 32-wide inverted diagonal block and U12 column
// live in registers, and the panel's argmax is a 4-wave exchange.
// Per n bucket, a sample of the batch (at least twice the CU count of systems, spread over the bucket's sizes,
// so the device is as full as in the real launch) is solved once per split candidate and the fastest split wins.
// Occupancy (resident workgroups per CU) is timed on the whole batch: a sample too small to fill the device
// cannot tell one workgroup per CU from two.  Last, the tuned shape is kept only if the whole batch runs
// faster with it than with the default shape.  The choice is cached per batch signature (the systems per
// bucket on a log2 scale and the largest right-hand-side count).
static std::string batch_key(int count, const int32_t* n, const int32_t* nrhs) {
  int cnt[4] = {0, 0, 0, 0}, rmax = 0;
  for (int i = 0; i < count; i++) { cnt[std::min(3, (n[i] - 1) / 128)]++; rmax = std::max(rmax, (int)nrhs[i]); }
  std::ostringstream o;
  for (int b = 0; b < 4; b++) o << (cnt[b] ? 1 + (int)std::log2((double)cnt[b]) : 0) << "/";
  o << (rmax <= 1 ? 1 : (rmax <= 4 ? 4 : 16));
  return o.str();
}

void batch_tune(int count, const int32_t* n, const int32_t* nrhs, const double* const* A, const double* const* B,
                double* const* X, BatchParams& p, Workspace& W, std::string& detail, bool tune_wg) {
  MP_REQUIRE(count > 0 && n && nrhs && A && B && X, "batch_tune: empty batch or NULL array");
  for (int i = 0; i < count; i++) {
    MP_REQUIRE(n[i] >= 1 && n[i] <= BNMAX, "every n must be in [1, 512] (larger systems go through mp_solve)");
    MP_REQUIRE(nrhs[i] >= 1 && nrhs[i] <= BRHS, "every nrhs must be in [1, 16]");
  }
  const std::string key = batch_key(count, n, nrhs);
  auto it = W.batch_tuned.find(key);
  if (it != W.batch_tuned.end()) {
    const BatchParams& c = it->second;
    for (int b = 0; b < 4; b++) { p.split4[b] = c.split4[b]; p.split2[b] = c.split2[b]; }
    if (tune_wg) p.wg_per_cu = c.wg_per_cu;
    detail = "{\"source\":\"cache\",\"key\":\"" + key + "\"}";
    return;
  }
  const double t0 = wall_now();
  std::vector<std::vector<int>> by(4);
  for (int i = 0; i < count; i++) by[std::min(3, (n[i] - 1) / 128)].push_back(i);
  std::vector<int> all;
  std::vector<BatchResult> scratch(count);
  auto run = [&](const std::vector<int>& ids, const BatchParams& q) {
    std::vector<int32_t> nn, rr; std::vector<const double*> aa, bb; std::vector<double*> xx;
    for (int i : ids) { nn.push_back(n[i]); rr.push_back(nrhs[i]); aa.push_back(A[i]); bb.push_back(B[i]); xx.push_back(X[i]); }
    double s = 0;
    batch_solve((int)ids.size(), nn.data(), rr.data(), aa.data(), bb.data(), xx.data(), q, W, scratch.data(), &s);
    return s;
  };
  static const int cand[][2] = {{0, 0}, {32, 64}, {64, 128}, {128, 256}, {256, 512}};
  std::ostringstream js;
  js << "{\"source\":\"tuned\",\"key\":\"" << key << "\",\"buckets\":{";
  bool first_b = true;
  for (int b = 0; b < 4; b++) {
    if (by[b].empty()) continue;
    std::vector<int> ids = by[b];
    std::sort(ids.begin(), ids.end(), [&](int x, int y) { return n[x] < n[y]; });
    std::vector<int> smp;
    const size_t want = std::min<size_t>(std::max<size_t>(96, (size_t)2 * W.num_cus), ids.size());
    for (size_t s = 0; s < want; s++) smp.push_back(ids[s * ids.size() / want]);
    all.insert(all.end(), smp.begin(), smp.end());
    BatchParams q = p; q.wg_per_cu = 2;
    run(smp, q);                                   // warm-up: workspace growth and first-launch costs
    double best = 1e300; int bi = 2;
    js << (first_b ? "" : ",") << "\"" << 128 * b + 1 << "-" << 128 * (b + 1) << "\":{";
    first_b = false;
    for (int c = 0; c < 5; c++) {
      q.split4[b] = cand[c][0]; q.split2[b] = cand[c][1];
      const double s = run(smp, q);
      js << (c ? "," : "") << "\"" << cand[c][0] << "/" << cand[c][1] << "\":" << s * 1e3;
      if (s < best) { best = s; bi = c; }
    }
    js << "}";
    p.split4[b] = cand[bi][0]; p.split2[b] = cand[bi][1];
  }
  std::vector<int> every(count);
  for (int i = 0; i < count; i++) every[i] = i;
  js << "},\"wg_per_cu\":{";
  double tuned_s = -1;
  if (tune_wg) {
    double best = 1e300; int bw = p.wg_per_cu;
    for (int w = 1; w <= 3; w++) {
      BatchParams q = p; q.wg_per_cu = w;
      const double s = run(every, q);
      js << (w > 1 ? "," : "") << "\"" << w << "\":" << s * 1e3;
      if (s < best) { best = s; bw = w; }
    }
    p.wg_per_cu = bw;
    tuned_s = best;
  } else {
    tuned_s = run(every, p);
  }
  // the default shape on the whole batch: the tuned one must beat it
  BatchParams d0;
  d0.tau = p.tau; d0.max_outer = p.max_outer; d0.eta_only = p.eta_only;
  if (!tune_wg) d0.wg_per_cu = p.wg_per_cu;
  const double default_s = run(every, d0);
  const bool keep = tuned_s < default_s;
  if (!keep) {
    for (int b = 0; b < 4; b++) { p.split4[b] = d0.split4[b]; p.split2[b] = d0.split2[b]; }
    if (tune_wg) p.wg_per_cu = d0.wg_per_cu;
  }
  js << "},\"tuned_ms\":" << tuned_s * 1e3 << ",\"default_ms\":" << default_s * 1e3 << ",\"kept\":\""
     << (keep ? "tuned" : "default") << "\",\"sample_systems\":" << all.size() << ",\"tune_s\":" << wall_now() - t0 << "}";
  detail = js.str();
  W.batch_tuned[key] = p;
}

}  // namespace mpk
