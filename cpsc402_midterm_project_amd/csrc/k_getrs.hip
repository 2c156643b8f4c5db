// k_getrs.hip — d = scale * U^{-1} L^{-1} P r with FP32 LU factors, FP32 arithmetic, FP64 input and output.
//
// Replaces the reference's factor application apply_inverse (src/mp_core.cuh:182-190: FP64->FP32 cast,
// cusolverDnXgetrs on FP32 factors, cast back, power-of-two unscaling) with two persistent HIP kernels.
// The permutation gather, the FP64->FP32 cast, the scaling and the FP32->FP64 cast are fused into them.
//
// Layout: one workgroup per 64-row block of the triangle.  Workgroup i streams its block row of L (or U),
// 64 rows x 16 columns per wave per block, and waits for the solution block j it needs by polling tagged
// 64-bit words (FP32 value + launch tag), so there is no flag/fence round trip.  The off-diagonal block is
// loaded before the wait, so only the y_j poll, a 16-FMA update, one LDS reduction and the 64-step
// diagonal substitution sit on the dependency chain between consecutive blocks.  Lower-numbered
// workgroups never wait for higher-numbered ones, so in-order dispatch guarantees forward progress.
#include "mp_internal.hpp"
#include "mp_kernels.hpp"

namespace mpk {

constexpr long GS_SPIN_MAX = 1L << 22;

__device__ __forceinline__ void gs_put(uint64_t* p, uint32_t tag, float v) {
  __hip_atomic_store(p, ((uint64_t)tag << 32) | (uint64_t)__float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// urgent = the waiting block is the immediate successor (it is on the dependency chain); the others back off
// so that the words being produced are not hammered by every waiting workgroup at once.
__device__ __forceinline__ float gs_get(const uint64_t* p, uint32_t tag, volatile int* s_abort, bool urgent) {
  for (long it = 0; it < GS_SPIN_MAX; it++) {
    const uint64_t w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(w >> 32) == tag) return __uint_as_float((uint32_t)w);
    if (*s_abort) return 0.0f;
    if (urgent) __builtin_amdgcn_s_sleep(1); else __builtin_amdgcn_s_sleep(16);
  }
  *s_abort = 1;
  return 0.0f;
}

// LOWER = true: unit-lower forward substitution on P r (input FP64 via perm gather), output Y (FP32).
// LOWER = false: upper backward substitution on Y (FP32), output D = scale * x (FP64).
template <int NR, bool LOWER>
__global__ __launch_bounds__(256) void k_trsv(const float* __restrict__ LU, int64_t lda, int64_t n,
                                              const int32_t* __restrict__ perm,
                                              const double* __restrict__ Rin, int64_t ldr,
                                              float* __restrict__ Y, int64_t ldy,
                                              double* __restrict__ D, int64_t ldd, double scale,
                                              uint64_t* __restrict__ ybuf, uint32_t tag, int32_t* __restrict__ err) {
  __shared__ float ys[2][NR][64];
  __shared__ float part[4][NR][64];
  __shared__ int s_abort;
  const int nbk = (int)((n + 63) / 64);
  const int t = threadIdx.x, rr = t & 63, cq = t >> 6;
  const int bi = LOWER ? (int)blockIdx.x : (nbk - 1 - (int)blockIdx.x);
  const int64_t row = (int64_t)bi * 64 + rr;
  const bool rvalid = row < n;
  if (t == 0) s_abort = 0;
  float acc[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) acc[c] = 0.0f;
  __syncthreads();
  const int nsteps = LOWER ? bi : (nbk - 1 - bi);
  // diagonal block (wave 0: lane rr holds row rr) and the right-hand side are independent of the other
  // blocks: load them before waiting, so the dependency chain sees no HBM latency for them
  float Dg[64];
  float rhs0[NR];
  if (t < 64) {
#pragma unroll
    for (int q = 0; q < 64; q++) {
      const int64_t col = (int64_t)bi * 64 + q;
      Dg[q] = (rvalid && col < n) ? LU[col * lda + row] : (q == rr ? 1.0f : 0.0f);
    }
#pragma unroll
    for (int c = 0; c < NR; c++) {
      if (LOWER) rhs0[c] = rvalid ? (float)Rin[(int64_t)c * ldr + perm[row]] : 0.0f;
      else rhs0[c] = 0.0f;
    }
  }
  for (int s = 0; s < nsteps; s++) {
    const int bj = LOWER ? s : (nbk - 1 - s);
    // off-diagonal block values of my row, 16 columns of this wave's quarter (loaded before the wait)
    float Lv[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int64_t col = (int64_t)bj * 64 + cq * 16 + q;
      Lv[q] = (rvalid && col < n) ? LU[col * lda + row] : 0.0f;
    }
    float (*yb)[64] = ys[s & 1];
    for (int qq = t; qq < 64 * NR; qq += 256) {
      const int c = qq >> 6, col = qq & 63;
      const bool cv = (int64_t)bj * 64 + col < n;
      yb[c][col] = cv ? gs_get(ybuf + ((int64_t)bj * NR + c) * 64 + col, tag, &s_abort, s == nsteps - 1) : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q++)
#pragma unroll
      for (int c = 0; c < NR; c++) acc[c] = fmaf(Lv[q], yb[c][cq * 16 + q], acc[c]);
  }
#pragma unroll
  for (int c = 0; c < NR; c++) part[cq][c][rr] = acc[c];
  __syncthreads();
  if (t < 64) {
    float b[NR];
#pragma unroll
    for (int c = 0; c < NR; c++) {
      float rhs;
      if (LOWER) rhs = rhs0[c];
      else rhs = rvalid ? Y[(int64_t)c * ldy + row] : 0.0f;
      const float sum = ((part[0][c][rr] + part[1][c][rr]) + part[2][c][rr]) + part[3][c][rr];
      b[c] = rhs - sum;
    }
    if (LOWER) {
#pragma unroll
      for (int q = 0; q < 64; q++) {
#pragma unroll
        for (int c = 0; c < NR; c++) {
          const float yq = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b[c]), q));
          if (rr > q) b[c] = fmaf(-Dg[q], yq, b[c]);
        }
      }
    } else {
#pragma unroll
      for (int q = 63; q >= 0; q--) {
#pragma unroll
        for (int c = 0; c < NR; c++) {
          if (rr == q) b[c] = b[c] / Dg[q];
          const float xq = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b[c]), q));
          if (rr < q) b[c] = fmaf(-Dg[q], xq, b[c]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NR; c++) {
      gs_put(ybuf + ((int64_t)bi * NR + c) * 64 + rr, tag, b[c]);
      if (rvalid) {
        if (LOWER) Y[(int64_t)c * ldy + row] = b[c];
        else D[(int64_t)c * ldd + row] = scale * (double)b[c];
      }
    }
  }
  if (t == 0 && s_abort) err[0] = 1;
}

__global__ void k_zero1(int32_t* p) { p[0] = 0; }

template <int NR>
static void trsv_pair(int64_t n, const float* LU, int64_t lda, const int32_t* perm, const double* R, int64_t ldr,
                      double* D, int64_t ldd, double scale, Workspace& W, int32_t* err) {
  const int nbk = (int)cdiv(n, 64);
  const uint32_t tl = W.take_tags(1), tu = W.take_tags(1);
  hipEvent_t ev = W.prof_begin();
  hipLaunchKernelGGL((k_trsv<NR, true>), dim3(nbk), dim3(256), 0, W.stream, LU, lda, n, perm, R, ldr,
                     W.rhs32.p, n, (double*)nullptr, (int64_t)0, 1.0, W.ybuf.p, tl, err);
  LAUNCH_CK();
  hipLaunchKernelGGL((k_trsv<NR, false>), dim3(nbk), dim3(256), 0, W.stream, LU, lda, n, perm, (const double*)nullptr, (int64_t)0,
                     W.rhs32.p, n, D, ldd, scale, W.ybuf.p, tu, err);
  LAUNCH_CK();
  W.prof_end(ev, "getrs", 4.0 * (double)n * n + 24.0 * n * NR, 2.0 * (double)n * n * NR);
}

void getrs(int64_t n, int nrhs, const float* LU, int64_t lda, const int32_t* perm,
           const double* R, int64_t ldr, double* D, int64_t ldd, double scale, Workspace& W) {
  MP_REQUIRE(n > 0 && nrhs > 0, "getrs: n, nrhs > 0");
  MP_REQUIRE(lda >= n && ldr >= n && ldd >= n, "getrs: leading dimensions >= n");
  const int64_t nbk = cdiv(n, 64);
  if (W.ybuf.ensure(*W.tracker, (size_t)nbk * 8 * 64, "getrs_exchange"))
    HIP_CK(hipMemsetAsync(W.ybuf.p, 0, W.ybuf.n * 8, W.stream));
  W.rhs32.ensure(*W.tracker, (size_t)n * 8, "getrs_y32");
  W.dinfo.ensure(*W.tracker, 4, "getrf_info");
  int c = 0;
  while (c < nrhs) {
    const int left = nrhs - c;
    const double* Rc = R + (int64_t)c * ldr;
    double* Dc = D + (int64_t)c * ldd;
    int32_t* err = W.dinfo.p + 2;
    if (left >= 8) { trsv_pair<8>(n, LU, lda, perm, Rc, ldr, Dc, ldd, scale, W, err); c += 8; }
    else if (left >= 4) { trsv_pair<4>(n, LU, lda, perm, Rc, ldr, Dc, ldd, scale, W, err); c += 4; }
    else if (left >= 2) { trsv_pair<2>(n, LU, lda, perm, Rc, ldr, Dc, ldd, scale, W, err); c += 2; }
    else { trsv_pair<1>(n, LU, lda, perm, Rc, ldr, Dc, ldd, scale, W, err); c += 1; }
  }
}

}  // namespace mpk
