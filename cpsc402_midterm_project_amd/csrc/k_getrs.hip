// k_getrs.hip — d = scale * U^{-1} L^{-1} P r with FP32 LU factors, FP32 arithmetic, FP64 input and output.
//
// Replaces the reference's factor application apply_inverse (src/mp_core.cuh:182-190: FP64->FP32 cast,
// cusolverDnXgetrs on FP32 factors, cast back, power-of-two unscaling) with HIP kernels of its own.  The
// permutation gather, the FP64->FP32 cast, the scaling and the FP32->FP64 cast are fused into them.
//
// Blocked by 128 rows.  The 128 x 128 diagonal blocks of L (unit lower) and U are inverted once per
// factorization (k_trinv, one workgroup per block: FP32 substitution on 16 x 16 diagonal blocks, then 2 x 2 block
// products up to 128), so the
// diagonal step of the solve is a parallel matrix-vector product, and so are the chain matrices
// M_i = T_i^{-1} B_i (B_i the off-diagonal block next to the diagonal, k_chainmat): block i of the solution
// is x_i = c_i - M_i x_prev with c_i = T_i^{-1} (r_i - sum of the other off-diagonal blocks' products), and
// c_i does not depend on x_prev.  k_trsv2: one workgroup of 512 threads per block row.  It streams the
// off-diagonal blocks of its block row in the order their solution blocks are produced, with the next
// block's loads in flight while the current one is applied, and forms c_i; the solution blocks are
// exchanged as tagged 64-bit words (FP32 value + launch tag), so a reader needs no flag and no fence.  Only
// one 128 x 128 matrix-vector product (M_i, preloaded into registers) sits on the dependency chain between
// consecutive block rows.  Lower-numbered workgroups never wait for higher-numbered ones, so in-order
// dispatch guarantees progress.
#include "mp_internal.hpp"
#include "mp_kernels.hpp"

namespace mpk {

constexpr long GS_SPIN_MAX = 1L << 22;
// Phase markers: empty in the library build; a probe build may define them to record timestamps.
#ifndef GS_MARK
#define GS_MARK(phase)
#endif
#ifndef GS_LOAD
#define GS_LOAD(p) (*(p))
#endif
constexpr int GB = 128;            // rows per block

__device__ __forceinline__ void gs_put(uint64_t* p, uint32_t tag, float v) {
  __hip_atomic_store(p, ((uint64_t)tag << 32) | (uint64_t)__float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t gs_load(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// dist = how many chain steps before this workgroup needs the block (0: it is on the dependency chain).
// Every workgroup waiting at the frontier polls the same few cache lines, so the poll interval grows with
// dist, but only slowly: a workgroup that oversleeps has to apply several blocks in a row once it wakes
// (about 0.5 us each), and falling behind the chain that way costs more than the extra polling.
__device__ __forceinline__ void gs_sleep(int dist) {
  if (dist <= 1) __builtin_amdgcn_s_sleep(1);
  else if (dist < 8) __builtin_amdgcn_s_sleep(2);
  else __builtin_amdgcn_s_sleep(8);
}
// The word loaded ahead (`first`) is checked before the polling loop: the loop re-loads, and at its head the
// compiler can only wait for every outstanding load, which would also wait for the prefetch in flight.
__device__ __forceinline__ float gs_wait(const uint64_t* p, uint64_t first, uint32_t tag, volatile lds_i32* s_abort, int dist,
                                        long spin) {
  if ((uint32_t)(first >> 32) == tag) return __uint_as_float((uint32_t)first);
  for (long it = 0; it < spin; it++) {
    if (*s_abort) return 0.0f;
    gs_sleep(dist);
    const uint64_t w = gs_load(p);
    if ((uint32_t)(w >> 32) == tag) return __uint_as_float((uint32_t)w);
  }
  *s_abort = 1;
  return 0.0f;
}

// Inverse of the 128 x 128 diagonal block b of L (unit lower, LOWER) or U (upper).  Output column-major per
// block: Tinv[b * 128 * 128 + col * 128 + row]; rows / columns beyond n are the identity.
// Blocked, so the serial chain is short: the eight 16 x 16 diagonal blocks are inverted by substitution (thread c
// takes column c % 16 of block c / 16, each sum accumulated in increasing p), then three levels combine pairs of
// s x s inverses into 2s x 2s ones (s = 16, 32, 64):
//   lower [A 0; B C]^{-1} = [A^{-1} 0; -C^{-1} (B A^{-1})  C^{-1}],
//   upper [A B; 0 C]^{-1} = [A^{-1}  -A^{-1} (B C^{-1}); 0 C^{-1}],
// every product summed in increasing k over its nonzero range.
template <bool LOWER>
__device__ __forceinline__ void trinv_block(const float* __restrict__ LU, int64_t lda, int64_t n, float* __restrict__ Tinv,
                                            int b, float* sm) {
  float* T = sm;                      // [128][129] row-major T[q * 129 + p]
  float* X = sm + GB * (GB + 1);      // [128][128] X[p * 128 + c], the inverse being built
  float* Wt = X + GB * GB;            // [64 * 64] products of one level
  const int t = threadIdx.x;
  const int64_t o = (int64_t)b * GB;
  const int nb = (int)min((int64_t)GB, n - o);
  {
    // thread q reads row q of every column p (coalesced over q), 32 columns in flight at a time; rows and
    // columns past n read a clamped element and store the identity
    const int q = t;
    const float* src = LU + o + min(q, nb - 1);
    for (int pb = 0; pb < GB; pb += 32) {
      float v[32];
#pragma unroll
      for (int i = 0; i < 32; i++) v[i] = src[(o + min(pb + i, nb - 1)) * lda];
#pragma unroll
      for (int i = 0; i < 32; i++) {
        const int p = pb + i;
        T[q * (GB + 1) + p] = (q < nb && p < nb) ? v[i] : ((q == p) ? 1.0f : 0.0f);
      }
    }
    for (int idx = t; idx < GB * GB; idx += GB) X[idx] = 0.0f;
  }
  __syncthreads();
  {
    // 16 x 16 leaves: block lb = t / 16, column c = t % 16
    const int lb = t >> 4, c = t & 15, ob = lb * 16;
    float x[16];
    if (LOWER) {
#pragma unroll
      for (int q = 0; q < 16; q++) x[q] = (q == c) ? 1.0f : 0.0f;
#pragma unroll
      for (int q = 1; q < 16; q++) {
        float s = 0.0f;
#pragma unroll
        for (int p = 0; p < q; p++) {
          const float ns = fmaf(T[(ob + q) * (GB + 1) + ob + p], x[p], s);
          s = (p >= c) ? ns : s;
        }
        x[q] = (q > c) ? -s : x[q];
      }
    } else {
#pragma unroll
      for (int q = 15; q >= 0; q--) {
        float s = (q == c) ? 1.0f : 0.0f;
#pragma unroll
        for (int p = q + 1; p < 16; p++) {
          const float ns = fmaf(-T[(ob + q) * (GB + 1) + ob + p], x[p], s);
          s = (p <= c) ? ns : s;
        }
        x[q] = (q <= c) ? s / T[(ob + q) * (GB + 1) + ob + q] : 0.0f;
      }
    }
#pragma unroll
    for (int q = 0; q < 16; q++) X[(ob + q) * GB + ob + c] = x[q];
  }
  __syncthreads();
  for (int s = 16; s < GB; s *= 2) {
    const int npair = GB / (2 * s), ss = s * s;
    // W = B A^{-1} (lower, B = T[o2 + s.., o2..]) or B C^{-1} (upper, B = T[o2.., o2 + s..]), pair by pair
    for (int e = t; e < npair * ss; e += GB) {
      const int pr = e / ss, i = (e % ss) / s, j = e % s, o2 = pr * 2 * s;
      float acc = 0.0f;
      if (LOWER) {
        for (int k = j; k < s; k++) acc = fmaf(T[(o2 + s + i) * (GB + 1) + o2 + k], X[(o2 + k) * GB + o2 + j], acc);
      } else {
        for (int k = 0; k <= j; k++) acc = fmaf(T[(o2 + i) * (GB + 1) + o2 + s + k], X[(o2 + s + k) * GB + o2 + s + j], acc);
      }
      Wt[pr * ss + i * s + j] = acc;
    }
    __syncthreads();
    // the off-diagonal block: lower -C^{-1} W into X[o2 + s.., o2..], upper -A^{-1} W into X[o2.., o2 + s..]
    for (int e = t; e < npair * ss; e += GB) {
      const int pr = e / ss, i = (e % ss) / s, j = e % s, o2 = pr * 2 * s;
      float acc = 0.0f;
      if (LOWER) {
        for (int k = 0; k <= i; k++) acc = fmaf(X[(o2 + s + i) * GB + o2 + s + k], Wt[pr * ss + k * s + j], acc);
        X[(o2 + s + i) * GB + o2 + j] = -acc;
      } else {
        for (int k = i; k < s; k++) acc = fmaf(X[(o2 + i) * GB + o2 + k], Wt[pr * ss + k * s + j], acc);
        X[(o2 + i) * GB + o2 + s + j] = -acc;
      }
    }
    __syncthreads();
  }
  for (int idx = t; idx < GB * GB; idx += GB) {
    const int col = idx >> 7, row = idx & (GB - 1);
    Tinv[(size_t)b * GB * GB + idx] = X[row * GB + col];
  }
}
// L's block inverses (workgroups 0 .. nbk-1) and U's (nbk .. 2 nbk-1) in one launch: each is one workgroup's
// serial work, so the two triangles take the time of one
__global__ __launch_bounds__(128) void k_trinv(const float* __restrict__ LU, int64_t lda, int64_t n, float* __restrict__ TinvL,
                                               float* __restrict__ TinvU, int nbk) {
  extern __shared__ float sm[];
  const int bx = (int)blockIdx.x;
  if (bx < nbk) trinv_block<true>(LU, lda, n, TinvL, bx, sm);
  else trinv_block<false>(LU, lda, n, TinvU, bx - nbk, sm);
}

// LOWER: y = Linv_i (P r)_i - sum_{j<i} L_ij y_j, output Y (FP32).   UPPER: x_i = Uinv_i (y_i - sum_{j>i} U_ij x_j),
// output D = scale * x (FP64).  Thread t: row r = t % 128 of the block, quarter qq = t / 128 of the columns.
// Chain matrix of block i: LOWER M_i = Linv_i L_{i,i-1} (i >= 1), UPPER M_i = Uinv_i U_{i,i+1} (i <= nbk-2);
// zero for the first block in production order.  Output column-major per block like Tinv; rows and columns
// beyond n are zero.  Thread (tr, tc) forms an 8 x 8 tile, each element summed over k in increasing order.
template <bool LOWER>
__global__ __launch_bounds__(256) void k_chainmat(const float* __restrict__ LU, int64_t lda, int64_t n,
                                                  const float* __restrict__ Tinv, float* __restrict__ M) {
  extern __shared__ float sm[];
  float* Ts = sm;                     // [128][129]: Ts[row * 129 + k] = Tinv_i[row][k]
  float* Bs = sm + GB * (GB + 1);     // [128][129]: Bs[k * 129 + col] = B_i[k][col]
  const int nbk = (int)((n + GB - 1) / GB);
  const int i = blockIdx.x, t = threadIdx.x;
  const int j = LOWER ? i - 1 : i + 1;
  float* Mo = M + (size_t)i * GB * GB;
  if (j < 0 || j >= nbk) {
    for (int idx = t; idx < GB * GB; idx += 256) Mo[idx] = 0.0f;
    return;
  }
  const int64_t ro = (int64_t)i * GB, co = (int64_t)j * GB;
  {
    // thread (row, column parity): 16 columns of both operands in flight at a time, coalesced over rows;
    // elements past n read a clamped address and store zero
    const int row = t & (GB - 1), ch = t >> 7;
    const float* tsrc = Tinv + (size_t)i * GB * GB + row;
    const float* bsrc = LU + min(ro + row, n - 1);
    for (int cb = 0; cb < GB; cb += 32) {
      float tv[16], bv[16];
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int col = cb + 2 * e + ch;
        tv[e] = tsrc[(size_t)col * GB];
        bv[e] = bsrc[min(co + col, n - 1) * lda];
      }
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int col = cb + 2 * e + ch;
        Ts[row * (GB + 1) + col] = tv[e];
        Bs[row * (GB + 1) + col] = (ro + row < n && co + col < n) ? bv[e] : 0.0f;
      }
    }
  }
  __syncthreads();
  const int tr = t >> 4, tc = t & 15;
  float acc[8][8];
#pragma unroll
  for (int q = 0; q < 8; q++)
#pragma unroll
    for (int p = 0; p < 8; p++) acc[q][p] = 0.0f;
  for (int k = 0; k < GB; k++) {
    float av[8], bv[8];
#pragma unroll
    for (int q = 0; q < 8; q++) av[q] = Ts[(tr * 8 + q) * (GB + 1) + k];
#pragma unroll
    for (int p = 0; p < 8; p++) bv[p] = Bs[k * (GB + 1) + tc * 8 + p];
#pragma unroll
    for (int q = 0; q < 8; q++)
#pragma unroll
      for (int p = 0; p < 8; p++) acc[q][p] = fmaf(av[q], bv[p], acc[q][p]);
  }
#pragma unroll
  for (int p = 0; p < 8; p++)
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int row = tr * 8 + q, col = tc * 8 + p;
      Mo[(size_t)col * GB + row] = (ro + row < n) ? acc[q][p] : 0.0f;
    }
}

template <int NR, bool LOWER>
__global__ __launch_bounds__(512) void k_trsv2(const float* __restrict__ LU, int64_t lda, int64_t n,
                                               const float* __restrict__ Tinv, const float* __restrict__ Mch,
                                               const int32_t* __restrict__ perm,
                                               const double* __restrict__ Rin, int64_t ldr,
                                               float* __restrict__ Y, int64_t ldy,
                                               double* __restrict__ D, int64_t ldd, double scale,
                                               uint64_t* __restrict__ ybuf, uint32_t tag, uint32_t wtag, int32_t* __restrict__ err, long spin,
                                               const int* __restrict__ act) {
  // a step of a Krylov cycle that has already ended: every workgroup returns (none waits for another)
  if (act && *act == 0) return;
  __shared__ float ys[3][NR][GB];
  __shared__ float part[4][NR][GB];
  __shared__ int s_abort;
  const int nbk = (int)((n + GB - 1) / GB);
  const int t = threadIdx.x, r = t & (GB - 1), qq = t >> 7;
  const int bi = LOWER ? (int)blockIdx.x : (nbk - 1 - (int)blockIdx.x);
  const int64_t row = (int64_t)bi * GB + r;
  const bool rvalid = row < n;
  if (t == 0) s_abort = 0;
  const int ndep = LOWER ? bi : (nbk - 1 - bi);     // blocks this row depends on
  const int nbulk = ndep > 0 ? ndep - 1 : 0;         // all but the immediately preceding one
  auto blk = [&](int s) { return LOWER ? s : (nbk - 1 - s); };   // s-th dependency in production order
  // Unconditional loads from clamped addresses, consumed only when the block is applied (a guarded or
  // masked load is waited for on the spot, which would serialise the 32 loads).  Clamped values are
  // harmless: a column beyond n meets a solution entry published as exactly 0, and a row beyond n is never
  // published (its inverse-block row and column are the identity padding).
  const int64_t rowc = rvalid ? row : n - 1;
  auto ldL = [&](int bj, float (&v)[32]) {
    const float* base = LU + rowc;
#pragma unroll
    for (int k = 0; k < 32; k++) {
      const int64_t col = (int64_t)bj * GB + qq * 32 + k;
      v[k] = GS_LOAD(base + (col < n ? col : n - 1) * lda);
    }
  };
  // preloads: inverse block row, chain matrix row, the right-hand side
  float Tv[32], Mt[32], rhs[NR];
#pragma unroll
  for (int k = 0; k < 32; k++) Tv[k] = Tinv[(size_t)bi * GB * GB + (size_t)(qq * 32 + k) * GB + r];
#pragma unroll
  for (int k = 0; k < 32; k++) Mt[k] = Mch[(size_t)bi * GB * GB + (size_t)(qq * 32 + k) * GB + r];
#pragma unroll
  for (int c = 0; c < NR; c++) {
    rhs[c] = 0.0f;
    if (qq == 0 && rvalid) rhs[c] = LOWER ? (float)Rin[(int64_t)c * ldr + perm[row]] : Y[(int64_t)c * ldy + row];
  }
  float acc[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) acc[c] = 0.0f;
  __syncthreads();

  // bulk: dependencies 0 .. nbulk-1, loads of step s+1 in flight while step s is applied.  Two register
  // buffers alternate (no copies, which would wait for the loads just issued) and every prefetch is issued
  // unconditionally (a clamped step past the end reloads block 0 and is ignored), so the number of loads
  // behind each consumed value is fixed and the waits stay partial.
  float La[32], Lb[32];
  uint64_t ya[NR], yb[NR];
  auto ldY = [&](int s2, uint64_t (&y)[NR]) {
    const int bj = blk(s2 < nbulk ? s2 : 0);
#pragma unroll
    for (int c = 0; c < NR; c++) y[c] = gs_load(ybuf + ((int64_t)bj * NR + c) * GB + (t & (GB - 1)));
  };
  auto step = [&](int s2, const float (&L)[32], const uint64_t (&y)[NR]) {
    const int bj = blk(s2);
    if (t < GB) {
#pragma unroll
      for (int c = 0; c < NR; c++) ys[s2 % 3][c][t] = gs_wait(ybuf + ((int64_t)bj * NR + c) * GB + t, y[c], wtag, LDS_FLAG(&s_abort), ndep - 1 - s2, spin);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k++)
#pragma unroll
      for (int c = 0; c < NR; c++) acc[c] = fmaf(L[k], ys[s2 % 3][c][qq * 32 + k], acc[c]);
  };
  if (nbulk > 0) { ldL(blk(0), La); ldY(0, ya); }
  for (int s = 0; s < nbulk; s += 2) {
    ldL(blk(s + 1 < nbulk ? s + 1 : 0), Lb); ldY(s + 1, yb);
    step(s, La, ya);
    if (s + 1 >= nbulk) break;
    ldL(blk(s + 2 < nbulk ? s + 2 : 0), La); ldY(s + 2, ya);
    step(s + 1, Lb, yb);
  }
  GS_MARK(0);
  // off the chain: b = rhs - (bulk sums), c = Tinv b
  const int sb = (nbulk + 1) % 3, sc = nbulk % 3;     // slots not read by the last two bulk steps
#pragma unroll
  for (int c = 0; c < NR; c++) part[qq][c][r] = acc[c];
  __syncthreads();
  if (qq == 0) {
#pragma unroll
    for (int c = 0; c < NR; c++)
      ys[sb][c][r] = rhs[c] - ((part[0][c][r] + part[1][c][r]) + (part[2][c][r] + part[3][c][r]));
  }
  __syncthreads();
  float px[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) px[c] = 0.0f;
#pragma unroll
  for (int k = 0; k < 32; k++)
#pragma unroll
    for (int c = 0; c < NR; c++) px[c] = fmaf(Tv[k], ys[sb][c][qq * 32 + k], px[c]);
  // part was last read before the previous barrier
#pragma unroll
  for (int c = 0; c < NR; c++) part[qq][c][r] = px[c];
  __syncthreads();
  float xv[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) xv[c] = (part[0][c][r] + part[1][c][r]) + (part[2][c][r] + part[3][c][r]);
  // the chain: x = c - M x_prev
  if (ndep > 0) {
    const int bj = blk(ndep - 1);
    if (t < GB) {
#pragma unroll
      for (int c = 0; c < NR; c++) {
        const uint64_t* p = ybuf + ((int64_t)bj * NR + c) * GB + t;
        ys[sc][c][t] = gs_wait(p, gs_load(p), wtag, LDS_FLAG(&s_abort), 0, spin);
      }
    }
    __syncthreads();                                  // also orders the reads of part above before the writes below
    GS_MARK(1);
    float d[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) d[c] = 0.0f;
#pragma unroll
    for (int k = 0; k < 32; k++)
#pragma unroll
      for (int c = 0; c < NR; c++) d[c] = fmaf(Mt[k], ys[sc][c][qq * 32 + k], d[c]);
#pragma unroll
    for (int c = 0; c < NR; c++) part[qq][c][r] = d[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NR; c++) xv[c] -= (part[0][c][r] + part[1][c][r]) + (part[2][c][r] + part[3][c][r]);
  }
  if (qq == 0) {
#pragma unroll
    for (int c = 0; c < NR; c++) {
      const float x = xv[c];
      gs_put(ybuf + ((int64_t)bi * NR + c) * GB + r, tag, rvalid ? x : 0.0f);
      if (rvalid) {
        if (LOWER) Y[(int64_t)c * ldy + row] = x;
        else D[(int64_t)c * ldd + row] = scale * (double)x;
      }
    }
  }
  GS_MARK(2);
  if (t == 0 && s_abort) err[0] = 1;
}

// ------------------------------------------------------------------ 256-row blocks, four workgroups per block row
// The chain of the substitution has one step per block row and each step costs a cross-CU exchange under the
// load of the streaming workgroups (about 1.8 us), so 256-row blocks halve the chain of 128-row blocks.  A
// block row's 256 rows are split over four workgroups of 64 rows, so each still streams one 64-row x 256-column
// slice (64 KB) of a previous block per chain step.  T_i^{-1} of the 256 x 256 diagonal block mixes the four
// slices: after the bulk sums, each workgroup publishes its 64-row slice of s_i = r_i - sum_{j < i-1} L_ij x_j
// (tagged words), reads the slices its rows of T_i^{-1} need (lower: the slices above, upper: below) and
// forms c_i; the chain step is x_i = c_i - M_i x_{i-1} with M_i = T_i^{-1} L_{i,i-1} (upper: U_{i,i+1}).
// Workgroup b of the launch: block row b / 4 in production order, slice 0..3 in dependency order (upper:
// from the bottom), so no workgroup waits for a higher-numbered one.
constexpr int GB3 = 256;           // rows per block
constexpr int SR3 = 64;            // rows per workgroup

// acc (thread (tr, tc)'s 8 x 8 tile of a 128 x 128 product) += As * Bs, both staged in LDS with row stride 129:
// As[row * 129 + k], Bs[k * 129 + col]; each element summed over k in increasing order.
__device__ __forceinline__ void mm128_lds(const float* As, const float* Bs, int tr, int tc, float (&acc)[8][8]) {
  for (int k = 0; k < GB; k++) {
    float av[8], bv[8];
#pragma unroll
    for (int q = 0; q < 8; q++) av[q] = As[(tr * 8 + q) * (GB + 1) + k];
#pragma unroll
    for (int p = 0; p < 8; p++) bv[p] = Bs[k * (GB + 1) + tc * 8 + p];
#pragma unroll
    for (int q = 0; q < 8; q++)
#pragma unroll
      for (int p = 0; p < 8; p++) acc[q][p] = fmaf(av[q], bv[p], acc[q][p]);
  }
}
// Stage a 128 x 128 block into LDS (row stride 129) from a loader f(row, col) -> value: thread (row, column
// parity) keeps 16 loads in flight.
template <class F>
__device__ __forceinline__ void stage128(float* S, int t, F&& f) {
  const int row = t & (GB - 1), ch = t >> 7;
  for (int cb = 0; cb < GB; cb += 32) {
    float v[16];
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = f(row, cb + 2 * e + ch);
#pragma unroll
    for (int e = 0; e < 16; e++) S[row * (GB + 1) + cb + 2 * e + ch] = v[e];
  }
}

// T_i^{-1} of 256-row block i (column-major 256 x 256 per block) from the 128-row inverses T128 (k_trinv) and the
// off-diagonal 128 x 128 part B of the block: lower [A 0; B C]^{-1} = [A^-1 0; -C^-1 B A^-1  C^-1],
// upper [A B; 0 C]^{-1} = [A^-1  -A^-1 B C^-1; 0 C^-1].  A half block past n is the identity (B = 0).
template <bool LOWER>
__device__ __forceinline__ void tinv256_block(const float* __restrict__ LU, int64_t lda, int64_t n,
                                              const float* __restrict__ T128, float* __restrict__ T256, int i, float* sm) {
  float* S0 = sm;                     // [128][129]
  float* S1 = sm + GB * (GB + 1);     // [128][129]
  const int t = threadIdx.x;
  const int nbk128 = (int)((n + GB - 1) / GB);
  const int b0 = 2 * i, b1 = 2 * i + 1;
  const bool has1 = b1 < nbk128;
  float* out = T256 + (size_t)i * GB3 * GB3;
  // the two diagonal 128-blocks and the zero block
  for (int idx = t; idx < GB * GB; idx += 256) {
    const int col = idx >> 7, row = idx & (GB - 1);
    const float a = T128[(size_t)b0 * GB * GB + idx];
    const float c = has1 ? T128[(size_t)b1 * GB * GB + idx] : (row == col ? 1.0f : 0.0f);
    out[(size_t)col * GB3 + row] = a;
    out[(size_t)(col + GB) * GB3 + row + GB] = c;
    if (LOWER) out[(size_t)(col + GB) * GB3 + row] = 0.0f; else out[(size_t)col * GB3 + row + GB] = 0.0f;
  }
  const int tr = t >> 4, tc = t & 15;
  float acc[8][8];
#pragma unroll
  for (int q = 0; q < 8; q++)
#pragma unroll
    for (int p = 0; p < 8; p++) acc[q][p] = 0.0f;
  if (has1) {
    const int64_t o = (int64_t)i * GB3;
    // B: lower rows o+128.., cols o..; upper rows o.., cols o+128..  (entries past n are zero)
    auto ldB = [&](int row, int col) -> float {
      const int64_t gr = LOWER ? o + GB + row : o + row, gc = LOWER ? o + col : o + GB + col;
      const float v = LU[min(gc, n - 1) * lda + min(gr, n - 1)];
      return (gr < n && gc < n) ? v : 0.0f;
    };
    auto ldT = [&](int b) { return [=](int row, int col) -> float { return T128[(size_t)b * GB * GB + (size_t)col * GB + row]; }; };
    // first product P = B A^-1 (lower) or A^-1 B (upper)
    if (LOWER) { stage128(S0, t, ldB); stage128(S1, t, ldT(b0)); }
    else { stage128(S0, t, ldT(b0)); stage128(S1, t, ldB); }
    __syncthreads();
    mm128_lds(S0, S1, tr, tc, acc);
    __syncthreads();
    // P into S1 (lower: it is the right operand of C^-1 P; upper: the left operand of P C^-1)
    float* P = LOWER ? S1 : S0;
#pragma unroll
    for (int q = 0; q < 8; q++)
#pragma unroll
      for (int p = 0; p < 8; p++) {
        // lower: S1 holds P as Bs[k * 129 + col]; upper: S0 holds P as As[row * 129 + k]
        P[(tr * 8 + q) * (GB + 1) + tc * 8 + p] = acc[q][p];
        acc[q][p] = 0.0f;
      }
    if (LOWER) stage128(S0, t, ldT(b1)); else stage128(S1, t, ldT(b1));
    __syncthreads();
    mm128_lds(S0, S1, tr, tc, acc);
  }
#pragma unroll
  for (int q = 0; q < 8; q++)
#pragma unroll
    for (int p = 0; p < 8; p++) {
      const int row = tr * 8 + q, col = tc * 8 + p;
      const float v = has1 ? -acc[q][p] : 0.0f;
      if (LOWER) out[(size_t)col * GB3 + row + GB] = v;
      else out[(size_t)(col + GB) * GB3 + row] = v;
    }
}
// L's 256-row inverses (workgroups 0 .. nbk3-1) and U's in one launch
__global__ __launch_bounds__(256) void k_tinv256(const float* __restrict__ LU, int64_t lda, int64_t n,
                                                 const float* __restrict__ T128L, const float* __restrict__ T128U,
                                                 float* __restrict__ T256L, float* __restrict__ T256U, int nbk3) {
  extern __shared__ float sm[];
  const int bx = (int)blockIdx.x;
  if (bx < nbk3) tinv256_block<true>(LU, lda, n, T128L, T256L, bx, sm);
  else tinv256_block<false>(LU, lda, n, T128U, T256U, bx - nbk3, sm);
}

// M_i = T_i^{-1} B_i for 256-row blocks (lower: B_i = L_{i,i-1}, upper: U_{i,i+1}; zero for the first block in
// production order).  One workgroup per 128 x 128 output quarter (sub-row sr, sub-column sc): the sum over the
// two inner quarters in increasing order.  Rows or columns past n are zero.
template <bool LOWER>
__device__ __forceinline__ void chain256_quarter(const float* __restrict__ LU, int64_t lda, int64_t n,
                                                 const float* __restrict__ T256, float* __restrict__ M256, int bx, float* sm) {
  float* S0 = sm;
  float* S1 = sm + GB * (GB + 1);
  const int nbk = (int)((n + GB3 - 1) / GB3);
  const int i = bx >> 2, sr = (bx >> 1) & 1, sc = bx & 1;
  const int t = threadIdx.x, tr = t >> 4, tc = t & 15;
  const int j = LOWER ? i - 1 : i + 1;
  float* out = M256 + (size_t)i * GB3 * GB3;
  float acc[8][8];
#pragma unroll
  for (int q = 0; q < 8; q++)
#pragma unroll
    for (int p = 0; p < 8; p++) acc[q][p] = 0.0f;
  if (j >= 0 && j < nbk) {
    const int64_t ro = (int64_t)i * GB3, co = (int64_t)j * GB3;
    for (int kq = 0; kq < 2; kq++) {
      // T_i^{-1} quarter (sr, kq) is zero above (lower) / below (upper) the diagonal quarters
      if (LOWER ? (kq > sr) : (kq < sr)) continue;
      stage128(S0, t, [&](int row, int col) -> float {
        return T256[(size_t)i * GB3 * GB3 + (size_t)(kq * GB + col) * GB3 + sr * GB + row];
      });
      stage128(S1, t, [&](int row, int col) -> float {
        const int64_t gr = ro + kq * GB + row, gc = co + sc * GB + col;
        const float v = LU[min(gc, n - 1) * lda + min(gr, n - 1)];
        return (gr < n && gc < n) ? v : 0.0f;
      });
      __syncthreads();
      mm128_lds(S0, S1, tr, tc, acc);
      __syncthreads();
    }
  }
#pragma unroll
  for (int q = 0; q < 8; q++)
#pragma unroll
    for (int p = 0; p < 8; p++) {
      const int row = sr * GB + tr * 8 + q, col = sc * GB + tc * 8 + p;
      const bool valid = (int64_t)i * GB3 + row < n;
      out[(size_t)col * GB3 + row] = valid ? acc[q][p] : 0.0f;
    }
}
// L's chain matrices (workgroups 0 .. 4 nbk3 - 1) and U's in one launch
__global__ __launch_bounds__(256) void k_chain256(const float* __restrict__ LU, int64_t lda, int64_t n,
                                                  const float* __restrict__ T256L, const float* __restrict__ T256U,
                                                  float* __restrict__ M256L, float* __restrict__ M256U, int nbk3) {
  extern __shared__ float sm[];
  const int bx = (int)blockIdx.x;
  if (bx < 4 * nbk3) chain256_quarter<true>(LU, lda, n, T256L, M256L, bx, sm);
  else chain256_quarter<false>(LU, lda, n, T256U, M256U, bx - 4 * nbk3, sm);
}

// The product A z that follows an upper solve, formed in the same launch (MvArgs, MV kernels): workgroups past
// the substitution's own take 512 rows and one column chunk of k_resid_partial's split (chunks nearest the
// bottom first, since the upper solve produces its blocks from the bottom).  Each waits until the substitution
// has published every solution word of its chunk, stages z = scale * x in LDS (the value the substitution
// stores in D) and forms the chunk's partial sums exactly as k_resid_partial does in product mode (same loads,
// same fma order), so the partials, and the reduction after them, are the same bits as for the separate
// product.  Substitution workgroups never wait for these, so in-order dispatch still guarantees progress.
constexpr int MV_ROWS = 512;       // rows per product workgroup
constexpr int MV_MAXC = 2048;      // largest column chunk staged in LDS
// wsrc (lower solves, one right-hand side): the right-hand side is v = wsrc / hn[0] (k_scale_copy's division),
// formed by the thread that reads it and also stored into vout.
struct MvArgs {
  const void* A = nullptr; int64_t lda = 0; int64_t cpc = 0; int nch = 0; double* P = nullptr;
  const double* wsrc = nullptr; const double* hn = nullptr; double* vout = nullptr;
};

template <typename TA>
__device__ __forceinline__ void mv_chunk(const TA* __restrict__ A, int64_t lda, int64_t n, int64_t cpc, int nch, int id,
                                         const uint64_t* __restrict__ xw, uint32_t wtag, double scale,
                                         double* __restrict__ P, double* zs, volatile lds_i32* s_abort, long spin) {
  const int t = threadIdx.x;
  const int nrb = (int)((n + MV_ROWS - 1) / MV_ROWS);
  const int64_t chunk = nch - 1 - id / nrb;
  const int64_t i = (int64_t)(id % nrb) * MV_ROWS + t;
  const int64_t j0 = chunk * cpc;
  const int64_t jl = (j0 + cpc <= n) ? cpc : (n - j0);
  // one sentinel per 64-row slice of each block the chunk touches (the last row of the slice, published by the
  // same store instruction as the rest), polled by one lane each with long sleeps
  const int b0 = (int)(j0 / GB3), b1 = (int)((j0 + jl - 1) / GB3);
  const int nsent = (b1 - b0 + 1) * (GB3 / SR3);
  if (t < nsent) {
    const uint64_t* p = xw + (int64_t)(b0 + t / (GB3 / SR3)) * GB3 + SR3 * (t % (GB3 / SR3)) + SR3 - 1;
    (void)gs_wait(p, gs_load(p), wtag, s_abort, 8, spin);
  }
  __syncthreads();
  for (int64_t q = t; q < jl; q += MV_ROWS) {
    const uint64_t* p = xw + (j0 + q);                    // one right-hand side: word of column j is xw[j]
    zs[q] = scale * (double)gs_wait(p, gs_load(p), wtag, s_abort, 0, spin);
  }
  __syncthreads();
  if (i >= n || jl <= 0) return;
  double s = 0.0, tt = 0.0;
  const TA* ap = A + j0 * lda + i;
  int64_t jj = 0;
  for (; jj + 8 <= jl; jj += 8) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; u++) a[u] = (double)__builtin_nontemporal_load(ap + (jj + u) * lda);
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const double aa = fabs(a[u]);
      const double x = zs[jj + u];
      s = fma(a[u], x, s);
      tt = fma(aa, fabs(x), tt);
    }
  }
  for (; jj < jl; jj++) {
    const double av = (double)ap[jj * lda];
    const double x = zs[jj];
    s = fma(av, x, s);
    tt = fma(fabs(av), fabs(x), tt);
  }
  double* base = P + (chunk * 2) * n + i;
  base[0] = s;
  base[n] = tt;
}

template <int NR, bool LOWER, typename TA = double, bool MV = false>
__global__ __launch_bounds__(512) void k_trsv3(const float* __restrict__ LU, int64_t lda, int64_t n,
                                               const float* __restrict__ Tinv, const float* __restrict__ Mch,
                                               const int32_t* __restrict__ perm,
                                               const double* __restrict__ Rin, int64_t ldr,
                                               float* __restrict__ Y, int64_t ldy,
                                               double* __restrict__ D, int64_t ldd, double scale,
                                               uint64_t* __restrict__ ybuf, uint32_t tag, uint32_t wtag, int32_t* __restrict__ err, long spin,
                                               const int* __restrict__ act, MvArgs mv) {
  if (act && *act == 0) return;
  __shared__ int s_abort;
  if constexpr (MV) {
    static_assert(NR == 1 && !LOWER, "the product follows a one-column upper solve");
    const int nbk0 = (int)((n + GB3 - 1) / GB3);
    if ((int)blockIdx.x >= 4 * nbk0) {
      __shared__ double zs[MV_MAXC];
      if (threadIdx.x == 0) s_abort = 0;
      __syncthreads();
      mv_chunk<TA>((const TA*)mv.A, mv.lda, n, mv.cpc, mv.nch, (int)blockIdx.x - 4 * nbk0, ybuf, wtag, scale, mv.P, zs,
                   LDS_FLAG(&s_abort), spin);
      __syncthreads();
      if (threadIdx.x == 0 && s_abort) err[0] = 1;
      return;
    }
  }
  __shared__ float ys[3][NR][GB3];
  __shared__ float part[8][NR][SR3];
  __shared__ float sv[NR][GB3];
  const int nbk = (int)((n + GB3 - 1) / GB3);
  const int t = threadIdx.x, r = t & (SR3 - 1), qq = t >> 6;
  const int pi = (int)blockIdx.x >> 2, wp = (int)blockIdx.x & 3;
  const int bi = LOWER ? pi : (nbk - 1 - pi);
  const int w = LOWER ? wp : 3 - wp;
  const int64_t row = (int64_t)bi * GB3 + SR3 * w + r;
  const bool rvalid = row < n;
  if (t == 0) s_abort = 0;
  const int ndep = LOWER ? bi : (nbk - 1 - bi);
  const int nbulk = ndep > 0 ? ndep - 1 : 0;
  auto blk = [&](int s) { return LOWER ? s : (nbk - 1 - s); };
  uint64_t* xw = ybuf;                                             // x words: [(block * NR + c) * 256 + row in block]
  uint64_t* sw = ybuf + (int64_t)nbk * NR * GB3;                   // s slices, same layout
  const int64_t rowc = rvalid ? row : n - 1;
  auto ldL = [&](int bj, float (&v)[32]) {
    const float* base = LU + rowc;
#pragma unroll
    for (int k = 0; k < 32; k++) {
      const int64_t col = (int64_t)bj * GB3 + qq * 32 + k;
      v[k] = GS_LOAD(base + (col < n ? col : n - 1) * lda);
    }
  };
  float Tv[32], Mt[32], rhs[NR];
#pragma unroll
  for (int k = 0; k < 32; k++) Tv[k] = Tinv[(size_t)bi * GB3 * GB3 + (size_t)(qq * 32 + k) * GB3 + SR3 * w + r];
#pragma unroll
  for (int k = 0; k < 32; k++) Mt[k] = Mch[(size_t)bi * GB3 * GB3 + (size_t)(qq * 32 + k) * GB3 + SR3 * w + r];
#pragma unroll
  for (int c = 0; c < NR; c++) {
    rhs[c] = 0.0f;
    if (qq == 0 && rvalid) {
      if (LOWER && NR == 1 && mv.wsrc) {
        const int64_t pr = perm[row];
        const double v = mv.wsrc[pr] / mv.hn[0];
        mv.vout[pr] = v;
        rhs[c] = (float)v;
      } else {
        rhs[c] = LOWER ? (float)Rin[(int64_t)c * ldr + perm[row]] : Y[(int64_t)c * ldy + row];
      }
    }
  }
  float acc[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) acc[c] = 0.0f;
  __syncthreads();

  // bulk, as in k_trsv2: dependencies 0 .. nbulk-1 in production order, the next block's loads in flight
  float La[32], Lb[32];
  uint64_t ya[NR], yb[NR];
  auto ldY = [&](int s2, uint64_t (&y)[NR]) {
    const int bj = blk(s2 < nbulk ? s2 : 0);
#pragma unroll
    for (int c = 0; c < NR; c++) y[c] = gs_load(xw + ((int64_t)bj * NR + c) * GB3 + (t & (GB3 - 1)));
  };
  auto step = [&](int s2, const float (&L)[32], const uint64_t (&y)[NR]) {
    const int bj = blk(s2);
    if (t < GB3) {
#pragma unroll
      for (int c = 0; c < NR; c++) ys[s2 % 3][c][t] = gs_wait(xw + ((int64_t)bj * NR + c) * GB3 + t, y[c], wtag, LDS_FLAG(&s_abort), ndep - 1 - s2, spin);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k++)
#pragma unroll
      for (int c = 0; c < NR; c++) acc[c] = fmaf(L[k], ys[s2 % 3][c][qq * 32 + k], acc[c]);
  };
  if (nbulk > 0) { ldL(blk(0), La); ldY(0, ya); }
  for (int s = 0; s < nbulk; s += 2) {
    ldL(blk(s + 1 < nbulk ? s + 1 : 0), Lb); ldY(s + 1, yb);
    step(s, La, ya);
    if (s + 1 >= nbulk) break;
    ldL(blk(s + 2 < nbulk ? s + 2 : 0), La); ldY(s + 2, ya);
    step(s + 1, Lb, yb);
  }
  GS_MARK(0);
  // this workgroup's slice of s_i = r_i - (bulk sums): published, and kept in sv
#pragma unroll
  for (int c = 0; c < NR; c++) part[qq][c][r] = acc[c];
  __syncthreads();
  if (qq == 0) {
#pragma unroll
    for (int c = 0; c < NR; c++) {
      const float s = rhs[c] - (((part[0][c][r] + part[1][c][r]) + (part[2][c][r] + part[3][c][r])) +
                                ((part[4][c][r] + part[5][c][r]) + (part[6][c][r] + part[7][c][r])));
      sv[c][SR3 * w + r] = s;
      gs_put(sw + ((int64_t)bi * NR + c) * GB3 + SR3 * w + r, tag, rvalid ? s : 0.0f);
    }
  }
  // the slices this workgroup's rows of T_i^{-1} need from the other three; the rest of sv is zero
  if (t < GB3) {
    const int ws = t >> 6;
    if (ws != w) {
      const bool need = LOWER ? (ws < w) : (ws > w);
#pragma unroll
      for (int c = 0; c < NR; c++) {
        float v = 0.0f;
        if (need) {
          const uint64_t* p = sw + ((int64_t)bi * NR + c) * GB3 + t;
          v = gs_wait(p, gs_load(p), wtag, LDS_FLAG(&s_abort), 0, spin);
        }
        sv[c][t] = v;
      }
    }
  }
  __syncthreads();
  float px[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) px[c] = 0.0f;
#pragma unroll
  for (int k = 0; k < 32; k++)
#pragma unroll
    for (int c = 0; c < NR; c++) px[c] = fmaf(Tv[k], sv[c][qq * 32 + k], px[c]);
#pragma unroll
  for (int c = 0; c < NR; c++) part[qq][c][r] = px[c];    // part's previous reads precede the last barrier
  __syncthreads();
  float xv[NR];
#pragma unroll
  for (int c = 0; c < NR; c++)
    xv[c] = ((part[0][c][r] + part[1][c][r]) + (part[2][c][r] + part[3][c][r])) +
            ((part[4][c][r] + part[5][c][r]) + (part[6][c][r] + part[7][c][r]));
  // the chain: x = c - M x_prev
  if (ndep > 0) {
    const int bj = blk(ndep - 1);
    const int sc = nbulk % 3;                              // a slot the last bulk step did not read
    if (t < GB3) {
#pragma unroll
      for (int c = 0; c < NR; c++) {
        const uint64_t* p = xw + ((int64_t)bj * NR + c) * GB3 + t;
        ys[sc][c][t] = gs_wait(p, gs_load(p), wtag, LDS_FLAG(&s_abort), 0, spin);
      }
    }
    __syncthreads();                                       // also orders the reads of part above before the writes below
    GS_MARK(1);
    float d[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) d[c] = 0.0f;
#pragma unroll
    for (int k = 0; k < 32; k++)
#pragma unroll
      for (int c = 0; c < NR; c++) d[c] = fmaf(Mt[k], ys[sc][c][qq * 32 + k], d[c]);
#pragma unroll
    for (int c = 0; c < NR; c++) part[qq][c][r] = d[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NR; c++)
      xv[c] -= ((part[0][c][r] + part[1][c][r]) + (part[2][c][r] + part[3][c][r])) +
               ((part[4][c][r] + part[5][c][r]) + (part[6][c][r] + part[7][c][r]));
  }
  if (qq == 0) {
#pragma unroll
    for (int c = 0; c < NR; c++) {
      const float x = xv[c];
      gs_put(xw + ((int64_t)bi * NR + c) * GB3 + SR3 * w + r, tag, rvalid ? x : 0.0f);
      if (rvalid) {
        if (LOWER) Y[(int64_t)c * ldy + row] = x;
        else D[(int64_t)c * ldd + row] = scale * (double)x;
      }
    }
  }
  GS_MARK(2);
  if (t == 0 && s_abort) err[0] = 1;
}

template <int NR>
static void trsv_pair(int64_t n, const float* LU, int64_t lda, const int32_t* perm, const double* R, int64_t ldr,
                      double* D, int64_t ldd, double scale, Workspace& W, int32_t* err, const int* act) {
  const int nbk = (int)cdiv(n, GB);
  const uint32_t tl = W.take_tags(1), tu = W.take_tags(1);
  const float* Ti = W.tinv.p;
  const float* Tu = W.tinv.p + (size_t)nbk * GB * GB;
  const float* Mi = W.tchain.p;
  const float* Mu = W.tchain.p + (size_t)nbk * GB * GB;
  // the forced-timeout mode waits for a tag take_tags never hands out (it stays below 0xFFFFFFF0)
  const long spin = W.gs_force ? 1L : GS_SPIN_MAX;
  const uint32_t wl = W.gs_force ? 0xFFFFFFFFu : tl, wu = W.gs_force ? 0xFFFFFFFFu : tu;
  hipEvent_t ev = W.prof_begin();
  hipLaunchKernelGGL((k_trsv2<NR, true>), dim3(nbk), dim3(512), 0, W.stream, LU, lda, n, Ti, Mi, perm, R, ldr,
                     W.rhs32.p, n, (double*)nullptr, (int64_t)0, 1.0, W.ybuf.p, tl, wl, err, spin, act);
  LAUNCH_CK();
  hipLaunchKernelGGL((k_trsv2<NR, false>), dim3(nbk), dim3(512), 0, W.stream, LU, lda, n, Tu, Mu, perm, (const double*)nullptr, (int64_t)0,
                     W.rhs32.p, n, D, ldd, scale, W.ybuf.p, tu, wu, err, spin, act);
  LAUNCH_CK();
  W.prof_end(ev, "getrs", 4.0 * (double)n * n + 24.0 * n * NR, 2.0 * (double)n * n * NR);
}

template <int NR>
static void trsv_pair3(int64_t n, const float* LU, int64_t lda, const int32_t* perm, const double* R, int64_t ldr,
                       double* D, int64_t ldd, double scale, Workspace& W, int32_t* err, const int* act) {
  const int nbk = (int)cdiv(n, GB3);
  const uint32_t tl = W.take_tags(1), tu = W.take_tags(1);
  const float* Ti = W.tinv3.p;
  const float* Tu = W.tinv3.p + (size_t)nbk * GB3 * GB3;
  const float* Mi = W.tchain3.p;
  const float* Mu = W.tchain3.p + (size_t)nbk * GB3 * GB3;
  const long spin = W.gs_force ? 1L : GS_SPIN_MAX;
  const uint32_t wl = W.gs_force ? 0xFFFFFFFFu : tl, wu = W.gs_force ? 0xFFFFFFFFu : tu;
  hipEvent_t ev = W.prof_begin();
  hipLaunchKernelGGL((k_trsv3<NR, true>), dim3(4 * nbk), dim3(512), 0, W.stream, LU, lda, n, Ti, Mi, perm, R, ldr,
                     W.rhs32.p, n, (double*)nullptr, (int64_t)0, 1.0, W.ybuf.p, tl, wl, err, spin, act, MvArgs{});
  LAUNCH_CK();
  hipLaunchKernelGGL((k_trsv3<NR, false>), dim3(4 * nbk), dim3(512), 0, W.stream, LU, lda, n, Tu, Mu, perm, (const double*)nullptr, (int64_t)0,
                     W.rhs32.p, n, D, ldd, scale, W.ybuf.p, tu, wu, err, spin, act, MvArgs{});
  LAUNCH_CK();
  W.prof_end(ev, "getrs", 4.0 * (double)n * n + 24.0 * n * NR, 2.0 * (double)n * n * NR);
}

// One right-hand side, then w = A d by the upper solve's extra workgroups and the product reduction.
static void trsv_pair3_product(int64_t n, const float* LU, int64_t lda, const int32_t* perm, const double* R, int64_t ldr,
                               double* D, double scale, Workspace& W, int32_t* err, const int* act, ProductReq& pr,
                               int64_t cpc, int64_t nch) {
  const int nbk = (int)cdiv(n, GB3);
  const uint32_t tl = W.take_tags(1), tu = W.take_tags(1);
  const float* Ti = W.tinv3.p;
  const float* Tu = W.tinv3.p + (size_t)nbk * GB3 * GB3;
  const float* Mi = W.tchain3.p;
  const float* Mu = W.tchain3.p + (size_t)nbk * GB3 * GB3;
  const long spin = W.gs_force ? 1L : GS_SPIN_MAX;
  const uint32_t wl = W.gs_force ? 0xFFFFFFFFu : tl, wu = W.gs_force ? 0xFFFFFFFFu : tu;
  W.part.ensure(*W.tracker, (size_t)nch * 2 * n, "resid_partials");
  MvArgs mv; mv.A = pr.A; mv.lda = n; mv.cpc = cpc; mv.nch = (int)nch; mv.P = W.part.p;
  const unsigned grid = (unsigned)(4 * nbk + cdiv(n, MV_ROWS) * nch);
  hipEvent_t ev = W.prof_begin();
  MvArgs lv;
  if (pr.scale_src) { lv.wsrc = pr.scale_src; lv.hn = pr.scale_div; lv.vout = (double*)R; }
  hipLaunchKernelGGL((k_trsv3<1, true>), dim3(4 * nbk), dim3(512), 0, W.stream, LU, lda, n, Ti, Mi, perm, R, ldr,
                     W.rhs32.p, n, (double*)nullptr, (int64_t)0, 1.0, W.ybuf.p, tl, wl, err, spin, act, lv);
  LAUNCH_CK();
  if (pr.a32)
    hipLaunchKernelGGL((k_trsv3<1, false, float, true>), dim3(grid), dim3(512), 0, W.stream, LU, lda, n, Tu, Mu, perm,
                       (const double*)nullptr, (int64_t)0, W.rhs32.p, n, D, n, scale, W.ybuf.p, tu, wu, err, spin, act, mv);
  else
    hipLaunchKernelGGL((k_trsv3<1, false, double, true>), dim3(grid), dim3(512), 0, W.stream, LU, lda, n, Tu, Mu, perm,
                       (const double*)nullptr, (int64_t)0, W.rhs32.p, n, D, n, scale, W.ybuf.p, tu, wu, err, spin, act, mv);
  LAUNCH_CK();
  // algorithmic bytes: the two triangles, then A once
  W.prof_end(ev, "getrs_product", 4.0 * (double)n * n + (pr.a32 ? 4.0 : 8.0) * (double)n * n + 32.0 * n,
             2.0 * (double)n * n + 2.0 * (double)n * n);
  if (pr.partials_only) { pr.part = W.part.p; pr.nch = nch; }
  else product_reduce(W.part.p, n, nch, D, pr.w, W, act);
}

bool getrs(int64_t n, int nrhs, const float* LU, int64_t lda, const int32_t* perm,
           const double* R, int64_t ldr, double* D, int64_t ldd, double scale, Workspace& W, bool reuse_inverse,
           const int* active, ProductReq* product) {
  MP_REQUIRE(n > 0 && nrhs > 0, "getrs: n, nrhs > 0");
  MP_REQUIRE(lda >= n && ldr >= n && ldd >= n, "getrs: leading dimensions >= n");
  const int64_t nbk = cdiv(n, GB);
  // 256-row blocks (four workgroups per block row) from n = 512 on; MPSOLVE_GETRS=2 keeps 128-row blocks
  const bool v3 = n >= 2 * GB3 && !(getenv("MPSOLVE_GETRS") && atoi(getenv("MPSOLVE_GETRS")) == 2);
  const int64_t nbk3 = cdiv(n, GB3);
  // exchange words: 128-row solve [block][rhs][row]; 256-row solve x and s slices, two of each per block row
  const size_t yw = std::max<size_t>((size_t)nbk * 2 * GB, (size_t)nbk3 * 2 * 2 * GB3);
  if (W.ybuf.ensure(*W.tracker, yw, "getrs_exchange"))
    HIP_CK(hipMemsetAsync(W.ybuf.p, 0, W.ybuf.n * 8, W.stream));
  if (v3) {
    W.tinv3.ensure(*W.tracker, (size_t)2 * nbk3 * GB3 * GB3, "getrs_block_inverses_256");
    W.tchain3.ensure(*W.tracker, (size_t)2 * nbk3 * GB3 * GB3, "getrs_chain_matrices_256");
  }
  W.rhs32.ensure(*W.tracker, (size_t)n * 4, "getrs_y32");
  W.dinfo.ensure(*W.tracker, 4, "getrf_info");
  W.tinv.ensure(*W.tracker, (size_t)2 * nbk * GB * GB, "getrs_block_inverses");
  if (!v3) W.tchain.ensure(*W.tracker, (size_t)2 * nbk * GB * GB, "getrs_chain_matrices");
  // A test forces the timeout path from the host: MPSOLVE_GS_FAIL = k (k > 0: the next k calls after the
  // variable takes a new value; "all": every call) makes the waits of the substitutions look for a tag that is
  // never published, with a bound of one poll.  Every dependent block then gives up (publishing zeros and
  // setting the timeout flag) and every workgroup exits normally.
  W.gs_force = false;
  if (const char* e = getenv("MPSOLVE_GS_FAIL")) {
    if (W.gs_fail_env != e) { W.gs_fail_env = e; W.gs_fail_calls = (strcmp(e, "all") == 0) ? -1 : std::max(0L, atol(e)); }
    if (W.gs_fail_calls != 0) { W.gs_force = true; if (W.gs_fail_calls > 0) W.gs_fail_calls--; }
  } else {
    W.gs_fail_env.clear();
  }
  if (!reuse_inverse || W.tinv_lu != LU || W.tinv_n != n || W.tinv_lda != lda || W.tinv_gen != W.factor_gen || W.tinv_v3 != v3) {
    static bool attr = false;
    const size_t lds = (size_t)(GB * (GB + 1) + GB * GB + 64 * 64) * 4;
    const size_t lds_cm = (size_t)2 * GB * (GB + 1) * 4;
    if (!attr) {
      HIP_CK(hipFuncSetAttribute((const void*)k_trinv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      HIP_CK(hipFuncSetAttribute((const void*)k_chainmat<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cm));
      HIP_CK(hipFuncSetAttribute((const void*)k_chainmat<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cm));
      HIP_CK(hipFuncSetAttribute((const void*)k_tinv256, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cm));
      HIP_CK(hipFuncSetAttribute((const void*)k_chain256, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cm));
      attr = true;
    }
    hipEvent_t ev = W.prof_begin();
    hipLaunchKernelGGL(k_trinv, dim3((unsigned)(2 * nbk)), dim3(GB), lds, W.stream, LU, lda, n, W.tinv.p,
                       W.tinv.p + (size_t)nbk * GB * GB, (int)nbk);
    LAUNCH_CK();
    if (v3) {
      // 256-row inverses from the 128-row ones, then the 256-row chain matrices
      const float* Tl = W.tinv.p; const float* Tu = W.tinv.p + (size_t)nbk * GB * GB;
      float* T3l = W.tinv3.p; float* T3u = W.tinv3.p + (size_t)nbk3 * GB3 * GB3;
      hipLaunchKernelGGL(k_tinv256, dim3((unsigned)(2 * nbk3)), dim3(256), lds_cm, W.stream, LU, lda, n, Tl, Tu, T3l, T3u, (int)nbk3);
      LAUNCH_CK();
      hipLaunchKernelGGL(k_chain256, dim3((unsigned)(8 * nbk3)), dim3(256), lds_cm, W.stream, LU, lda, n, T3l, T3u, W.tchain3.p,
                         W.tchain3.p + (size_t)nbk3 * GB3 * GB3, (int)nbk3);
      LAUNCH_CK();
    } else {
      hipLaunchKernelGGL(k_chainmat<true>, dim3((unsigned)nbk), dim3(256), lds_cm, W.stream, LU, lda, n, W.tinv.p, W.tchain.p);
      LAUNCH_CK();
      hipLaunchKernelGGL(k_chainmat<false>, dim3((unsigned)nbk), dim3(256), lds_cm, W.stream, LU, lda, n,
                         W.tinv.p + (size_t)nbk * GB * GB, W.tchain.p + (size_t)nbk * GB * GB);
      LAUNCH_CK();
    }
    W.prof_end(ev, "getrs_block_inverse", 16.0 * (double)nbk * GB * GB, 2.0 * (double)nbk * GB * GB * GB * (1.0 / 3.0 + 2.0));
    W.tinv_lu = LU; W.tinv_n = n; W.tinv_lda = lda; W.tinv_gen = W.factor_gen; W.tinv_v3 = v3;
  }
  if (product && v3 && nrhs == 1 && ldd == n) {
    int64_t cpc, nch;
    product_chunks(n, cpc, nch);
    if (cpc <= MV_MAXC) {
      trsv_pair3_product(n, LU, lda, perm, R, ldr, D, scale, W, W.dinfo.p + 2, active, *product, cpc, nch);
      return true;
    }
  }
  int c = 0;
  while (c < nrhs) {
    const int left = nrhs - c;
    const double* Rc = R + (int64_t)c * ldr;
    double* Dc = D + (int64_t)c * ldd;
    int32_t* err = W.dinfo.p + 2;
    if (v3) {
      if (left >= 2) { trsv_pair3<2>(n, LU, lda, perm, Rc, ldr, Dc, ldd, scale, W, err, active); c += 2; }
      else { trsv_pair3<1>(n, LU, lda, perm, Rc, ldr, Dc, ldd, scale, W, err, active); c += 1; }
    } else {
      if (left >= 2) { trsv_pair<2>(n, LU, lda, perm, Rc, ldr, Dc, ldd, scale, W, err, active); c += 2; }
      else { trsv_pair<1>(n, LU, lda, perm, Rc, ldr, Dc, ldd, scale, W, err, active); c += 1; }
    }
  }
  return false;
}

}  // namespace mpk
