// k_getrs.hip — d = scale * U^{-1} L^{-1} P r with FP32 LU factors, FP32 arithmetic, FP64 input and output.
//
// Replaces the reference's factor application apply_inverse (src/mp_core.cuh:182-190: FP64->FP32 cast,
// cusolverDnXgetrs on FP32 factors, cast back, power-of-two unscaling) with HIP kernels of its own.  The
// permutation gather, the FP64->FP32 cast, the scaling and the FP32->FP64 cast are fused into them.
//
// Blocked by 128 rows.  The 128 x 128 diagonal blocks of L (unit lower) and U are inverted once per
// factorization (k_trinv, one workgroup per block, FP32 substitution on the columns of the identity), so the
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
// Thread c computes column c: lower  x_q = -sum_{p=c}^{q-1} L_qp x_p (x_c = 1),
//                             upper  x_q = (delta_qc - sum_{p=q+1}^{c} U_qp x_p) / U_qq, q = c .. 0,
// each sum accumulated in increasing p.
template <bool LOWER>
__global__ __launch_bounds__(128) void k_trinv(const float* __restrict__ LU, int64_t lda, int64_t n, float* __restrict__ Tinv) {
  extern __shared__ float sm[];
  float* T = sm;                      // [128][129] row-major T[q * 129 + p]
  float* X = sm + GB * (GB + 1);      // [128][128] X[p * 128 + c]
  const int b = blockIdx.x, c = threadIdx.x;
  const int64_t o = (int64_t)b * GB;
  const int nb = (int)min((int64_t)GB, n - o);
  {
    // thread c reads row q = c of every column p (coalesced over q), 32 columns in flight at a time; rows and
    // columns past n read a clamped element and store the identity
    const int q = c;
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
  }
  __syncthreads();
  if (LOWER) {
    X[c * GB + c] = 1.0f;
    for (int q = 0; q < GB; q++) if (q < c) X[q * GB + c] = 0.0f;
    for (int q = 1; q < GB; q++) {
      float s = 0.0f;
      // unrolled so the LDS reads of later terms are in flight while the fmaf chain advances (same order)
#pragma unroll 8
      for (int p = 0; p < q; p++) {
        const float ns = fmaf(T[q * (GB + 1) + p], X[p * GB + c], s);
        s = (p >= c) ? ns : s;
      }
      if (q > c) X[q * GB + c] = -s;
    }
  } else {
    for (int q = GB - 1; q >= 0; q--) {
      float s = (q == c) ? 1.0f : 0.0f;
#pragma unroll 8
      for (int p = q + 1; p < GB; p++) {
        const float ns = fmaf(-T[q * (GB + 1) + p], X[p * GB + c], s);
        s = (p <= c) ? ns : s;
      }
      X[q * GB + c] = (q <= c) ? s / T[q * (GB + 1) + q] : 0.0f;
    }
  }
  __syncthreads();
  for (int idx = c; idx < GB * GB; idx += GB) {
    const int col = idx >> 7, row = idx & (GB - 1);
    Tinv[(size_t)b * GB * GB + idx] = X[row * GB + col];
  }
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

void getrs(int64_t n, int nrhs, const float* LU, int64_t lda, const int32_t* perm,
           const double* R, int64_t ldr, double* D, int64_t ldd, double scale, Workspace& W, bool reuse_inverse,
           const int* active) {
  MP_REQUIRE(n > 0 && nrhs > 0, "getrs: n, nrhs > 0");
  MP_REQUIRE(lda >= n && ldr >= n && ldd >= n, "getrs: leading dimensions >= n");
  const int64_t nbk = cdiv(n, GB);
  if (W.ybuf.ensure(*W.tracker, (size_t)nbk * 2 * GB, "getrs_exchange"))
    HIP_CK(hipMemsetAsync(W.ybuf.p, 0, W.ybuf.n * 8, W.stream));
  W.rhs32.ensure(*W.tracker, (size_t)n * 4, "getrs_y32");
  W.dinfo.ensure(*W.tracker, 4, "getrf_info");
  W.tinv.ensure(*W.tracker, (size_t)2 * nbk * GB * GB, "getrs_block_inverses");
  W.tchain.ensure(*W.tracker, (size_t)2 * nbk * GB * GB, "getrs_chain_matrices");
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
  if (!reuse_inverse || W.tinv_lu != LU || W.tinv_n != n || W.tinv_lda != lda || W.tinv_gen != W.factor_gen) {
    static bool attr = false;
    const size_t lds = (size_t)(GB * (GB + 1) + GB * GB) * 4;
    const size_t lds_cm = (size_t)2 * GB * (GB + 1) * 4;
    if (!attr) {
      HIP_CK(hipFuncSetAttribute((const void*)k_trinv<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      HIP_CK(hipFuncSetAttribute((const void*)k_trinv<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      HIP_CK(hipFuncSetAttribute((const void*)k_chainmat<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cm));
      HIP_CK(hipFuncSetAttribute((const void*)k_chainmat<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cm));
      attr = true;
    }
    hipEvent_t ev = W.prof_begin();
    hipLaunchKernelGGL(k_trinv<true>, dim3((unsigned)nbk), dim3(GB), lds, W.stream, LU, lda, n, W.tinv.p);
    LAUNCH_CK();
    hipLaunchKernelGGL(k_trinv<false>, dim3((unsigned)nbk), dim3(GB), lds, W.stream, LU, lda, n, W.tinv.p + (size_t)nbk * GB * GB);
    LAUNCH_CK();
    hipLaunchKernelGGL(k_chainmat<true>, dim3((unsigned)nbk), dim3(256), lds_cm, W.stream, LU, lda, n, W.tinv.p, W.tchain.p);
    LAUNCH_CK();
    hipLaunchKernelGGL(k_chainmat<false>, dim3((unsigned)nbk), dim3(256), lds_cm, W.stream, LU, lda, n,
                       W.tinv.p + (size_t)nbk * GB * GB, W.tchain.p + (size_t)nbk * GB * GB);
    LAUNCH_CK();
    W.prof_end(ev, "getrs_block_inverse", 16.0 * (double)nbk * GB * GB, 2.0 * (double)nbk * GB * GB * GB * (1.0 / 3.0 + 2.0));
    W.tinv_lu = LU; W.tinv_n = n; W.tinv_lda = lda; W.tinv_gen = W.factor_gen;
  }
  int c = 0;
  while (c < nrhs) {
    const int left = nrhs - c;
    const double* Rc = R + (int64_t)c * ldr;
    double* Dc = D + (int64_t)c * ldd;
    int32_t* err = W.dinfo.p + 2;
    if (left >= 2) { trsv_pair<2>(n, LU, lda, perm, Rc, ldr, Dc, ldd, scale, W, err, active); c += 2; }
    else { trsv_pair<1>(n, LU, lda, perm, Rc, ldr, Dc, ldd, scale, W, err, active); c += 1; }
  }
}

}  // namespace mpk
