// k_getrf64.hip — FP64 blocked LU with the trailing update on FP64 MFMA (the plain-FP64 comparator of the
// mixed-precision solve, built on the GPU's own FP64 matrix cores).
//
// The reference's plain FP64 LU is the vendor getrf (native_lu64: factorize(cs64), src/mp_core.cuh:162-170).
// Here the blocked right-looking schedule and the trailing update (98.8% of the 2n^3/3 flops at n = 16384) are
// this library's: per block column, rocSOLVER factors the m x nb panel (a latency-bound call), rocSOLVER's laswp
// applies its interchanges left and right, rocBLAS solves U12 = L11^{-1} A12, and k_dgemm applies
// A22 -= L21 U12 on v_mfma_f64_16x16x4_f64.  With look-ahead the next panel's columns are updated first and the
// next panel factored on a second stream while the rest of the trailing matrix is updated.
#include "mp_kernels.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <mutex>

namespace mpk {

namespace {

typedef double f64x4_t __attribute__((ext_vector_type(4)));

// C (m2 x n2, column-major, ldc) -= L (m2 x kb, ldl) * U (kb x n2, ldu); kb % 4 == 0.
// Workgroup tile 128 x 128, four waves in 2 x 2, each wave 4 x 4 MFMA tiles of 16 x 16.  The kernel computes
// C^T tiles: the MFMA A operand is the negated U fragment (C columns), the B operand the L fragment (C rows),
// and register r of accumulator lane l holds D[4 r + l / 16][l % 16] (the lane layout of
// v_mfma_f64_16x16x4_f64), i.e. C[row = l % 16][col = 4 r + l / 16]: the 16 lanes of a row group touch 16
// consecutive rows of one column.  Operand loads of the next k step are issued before the
// current step's 16 products; edge rows and columns load from clamped addresses and are not stored.
__global__ __launch_bounds__(256) void k_dgemm(double* __restrict__ C, int64_t ldc, int m2, int n2, int kb,
                                               const double* __restrict__ L, int64_t ldl,
                                               const double* __restrict__ U, int64_t ldu) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int r_base = blockIdx.x * 128 + (w & 1) * 64;      // C rows of this wave
  const int c_base = blockIdx.y * 128 + (w >> 1) * 64;     // C columns of this wave
  const int li = lane & 15, lk = lane >> 4;
  int rows[4], cols[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    rows[i] = min(r_base + i * 16 + li, m2 - 1);
    cols[i] = min(c_base + i * 16 + li, n2 - 1);
  }
  // accumulators start from C^T
  f64x4_t acc[4][4];
#pragma unroll
  for (int ci = 0; ci < 4; ci++)
#pragma unroll
    for (int ri = 0; ri < 4; ri++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = min(c_base + ci * 16 + 4 * r + lk, n2 - 1);
        acc[ci][ri][r] = C[rows[ri] + (int64_t)col * ldc];
      }
  double ua[2][4], lb[2][4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    ua[0][i] = -U[lk + (int64_t)cols[i] * ldu];
    lb[0][i] = L[rows[i] + (int64_t)lk * ldl];
  }
  const int nk = kb / 4;
  for (int ks = 0; ks < nk; ks += 2) {
    // ping-pong operand sets: loads of step ks+1 in flight while step ks multiplies
    if (ks + 1 < nk) {
      const int k1 = (ks + 1) * 4 + lk;
#pragma unroll
      for (int i = 0; i < 4; i++) { ua[1][i] = -U[k1 + (int64_t)cols[i] * ldu]; lb[1][i] = L[rows[i] + (int64_t)k1 * ldl]; }
    }
#pragma unroll
    for (int ci = 0; ci < 4; ci++)
#pragma unroll
      for (int ri = 0; ri < 4; ri++)
        acc[ci][ri] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[0][ci], lb[0][ri], acc[ci][ri], 0, 0, 0);
    if (ks + 1 >= nk) break;
    if (ks + 2 < nk) {
      const int k2 = (ks + 2) * 4 + lk;
#pragma unroll
      for (int i = 0; i < 4; i++) { ua[0][i] = -U[k2 + (int64_t)cols[i] * ldu]; lb[0][i] = L[rows[i] + (int64_t)k2 * ldl]; }
    }
#pragma unroll
    for (int ci = 0; ci < 4; ci++)
#pragma unroll
      for (int ri = 0; ri < 4; ri++)
        acc[ci][ri] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua[1][ci], lb[1][ri], acc[ci][ri], 0, 0, 0);
  }
#pragma unroll
  for (int ci = 0; ci < 4; ci++)
#pragma unroll
    for (int ri = 0; ri < 4; ri++) {
      const int row = r_base + ri * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = c_base + ci * 16 + 4 * r + lk;
        if (row < m2 && col < n2) C[row + (int64_t)col * ldc] = acc[ci][ri][r];
      }
    }
}

// The same update with the operands staged through LDS: K is processed in chunks of 32, double-buffered; the
// workgroup loads chunk s+1 from global memory into registers (coalesced: L by rows, U by k) while its waves
// multiply chunk s from LDS, then stores it to the other buffer.  LDS: L chunk [32][128] and U chunk [32][129]
// (padded against bank conflicts of the k-major U stores), two buffers, 132 KB.  kb % 32 == 0.
constexpr int DG_KC = 32;
constexpr int DG_US = 129;
static size_t dgemm_lds_bytes() { return (size_t)2 * DG_KC * (128 + DG_US) * 8; }
__global__ __launch_bounds__(256) void k_dgemm_lds(double* __restrict__ C, int64_t ldc, int m2, int n2, int kb,
                                                   const double* __restrict__ L, int64_t ldl,
                                                   const double* __restrict__ U, int64_t ldu) {
  extern __shared__ double sm[];
  double* Ls = sm;                           // [2][KC][128]
  double* Us = sm + 2 * DG_KC * 128;         // [2][KC][129]
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int row0 = blockIdx.x * 128, col0 = blockIdx.y * 128;
  const int wr = (w & 1) * 64, wc = (w >> 1) * 64;       // wave tile inside the workgroup tile
  const int li = lane & 15, lk = lane >> 4;
  // C^T accumulators (same lane layout as k_dgemm)
  f64x4_t acc[4][4];
#pragma unroll
  for (int ci = 0; ci < 4; ci++)
#pragma unroll
    for (int ri = 0; ri < 4; ri++) {
      const int row = min(row0 + wr + ri * 16 + li, m2 - 1);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = min(col0 + wc + ci * 16 + 4 * r + lk, n2 - 1);
        acc[ci][ri][r] = C[row + (int64_t)col * ldc];
      }
    }
  // global -> register staging: L element e = t + 256 m: k = e / 128, row = e % 128 (rows contiguous);
  // U element e: k = e % 32, col = e / 32 (k contiguous within a column)
  double gl[16], gu[16];
  auto gload = [&](int k0) {
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int e = t + 256 * m;
      const int kl = e >> 7, rl = e & 127;
      gl[m] = L[min(row0 + rl, m2 - 1) + (int64_t)(k0 + kl) * ldl];
      const int ku = e & 31, cu = e >> 5;
      gu[m] = -U[(k0 + ku) + (int64_t)min(col0 + cu, n2 - 1) * ldu];
    }
  };
  auto lstore = [&](int buf) {
    double* lb = Ls + buf * DG_KC * 128;
    double* ub = Us + buf * DG_KC * DG_US;
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int e = t + 256 * m;
      lb[(e >> 7) * 128 + (e & 127)] = gl[m];
      ub[(e & 31) * DG_US + (e >> 5)] = gu[m];
    }
  };
  const int nch = kb / DG_KC;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ch++) {
    const int buf = ch & 1;
    if (ch + 1 < nch) gload((ch + 1) * DG_KC);
    const double* lb = Ls + buf * DG_KC * 128;
    const double* ub = Us + buf * DG_KC * DG_US;
#pragma unroll
    for (int ks = 0; ks < DG_KC / 4; ks++) {
      const int kk = ks * 4 + lk;
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        a[i] = ub[kk * DG_US + wc + i * 16 + li];
        b[i] = lb[kk * 128 + wr + i * 16 + li];
      }
#pragma unroll
      for (int ci = 0; ci < 4; ci++)
#pragma unroll
        for (int ri = 0; ri < 4; ri++)
          acc[ci][ri] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ci], b[ri], acc[ci][ri], 0, 0, 0);
    }
    if (ch + 1 < nch) lstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int ci = 0; ci < 4; ci++)
#pragma unroll
    for (int ri = 0; ri < 4; ri++) {
      const int row = row0 + wr + ri * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = col0 + wc + ci * 16 + 4 * r + lk;
        if (row < m2 && col < n2) C[row + (int64_t)col * ldc] = acc[ci][ri][r];
      }
    }
}

// panel-local 1-based pivots -> global 1-based pivots (rows of the whole matrix)
__global__ void k_ipiv_global(int32_t* ipiv, int kb, int off) {
  const int i = threadIdx.x + blockIdx.x * 256;
  if (i < kb) ipiv[i] += off;
}

#define RB_CK2(x) do { rocblas_status s_ = (x); if (s_ != rocblas_status_success) { std::ostringstream o_; \
  o_ << "rocBLAS/rocSOLVER status " << (int)s_ << " at " << __FILE__ << ":" << __LINE__ << " in " << #x; throw MPError(o_.str(), -2); } } while (0)

}  // namespace

void dgemm_update(int64_t m2, int64_t n2, int kb, const double* L, int64_t ldl, const double* U, int64_t ldu,
                  double* C, int64_t ldc, hipStream_t stream) {
  if (m2 <= 0 || n2 <= 0 || kb <= 0) return;
  MP_REQUIRE(kb % 4 == 0, "dgemm_update: kb must be a multiple of 4");
  MP_REQUIRE(m2 < (1ll << 31) && n2 < (1ll << 31), "dgemm_update: sizes too large");
  if (kb % DG_KC == 0) {
    static std::once_flag lds_once;      // contexts on several host threads may get here together
    hipError_t attr = hipSuccess;
    std::call_once(lds_once, [&] {
      attr = hipFuncSetAttribute((const void*)k_dgemm_lds, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dgemm_lds_bytes());
    });
    HIP_CK(attr);
    hipLaunchKernelGGL(k_dgemm_lds, dim3(cdiv(m2, 128), cdiv(n2, 128)), dim3(256), dgemm_lds_bytes(), stream, C, ldc,
                       (int)m2, (int)n2, kb, L, ldl, U, ldu);
  } else {
    hipLaunchKernelGGL(k_dgemm, dim3(cdiv(m2, 128), cdiv(n2, 128)), dim3(256), 0, stream, C, ldc, (int)m2, (int)n2, kb, L, ldl, U, ldu);
  }
  LAUNCH_CK();
}

// Blocked LU of the n x n FP64 matrix A (column-major, lda): P A = L U in place, ipiv 1-based global LAPACK
// pivots (device, n entries), *info = first zero pivot (1-based) or 0.  nb % 4 == 0.
void getrf64(int64_t n, double* A, int64_t lda, int32_t* ipiv, int nb, bool lookahead, void* rb_main, void* rb_panel,
             Workspace& W, int* info_out) {
  MP_REQUIRE(n > 0 && lda >= n && nb >= 4 && nb % 4 == 0, "getrf64: bad sizes");
  rocblas_handle rb = (rocblas_handle)rb_main;
  rocblas_handle rbp = (rocblas_handle)rb_panel;    // own workspace: the panel runs beside calls on rb
  W.getrf64_info.ensure(*W.tracker, (size_t)(n + nb - 1) / nb + 1, "getrf64 info");
  const int nblk = (int)((n + nb - 1) / nb);
  hipStream_t s0 = W.stream, s1 = W.pstream;
  const bool la = lookahead && s1 != nullptr && rbp != nullptr && nblk > 2;
  auto A_ = [&](int64_t r, int64_t c) { return A + r + c * lda; };
  auto panel = [&](int64_t k, hipStream_t s) {
    const int kb = (int)std::min<int64_t>(nb, n - k);
    rocblas_handle h = (s == s1) ? rbp : rb;
    RB_CK2(rocblas_set_stream(h, s));
    hipEvent_t e = W.prof_begin(s);
    RB_CK2(rocsolver_dgetrf(h, (rocblas_int)(n - k), kb, A_(k, k), (rocblas_int)lda, ipiv + k, W.getrf64_info.p + k / nb));
    hipLaunchKernelGGL(k_ipiv_global, dim3(cdiv(kb, 256)), dim3(256), 0, s, ipiv + k, kb, (int)k);
    LAUNCH_CK();
    W.prof_end(e, "lu64_panel", 16.0 * (n - k) * kb, (double)(n - k) * kb * kb, s);
  };
  // interchanges of block k on columns [c0, c1), then U12 and the update of those columns
  auto apply = [&](int64_t k, int64_t c0, int64_t c1, hipStream_t s) {
    const int kb = (int)std::min<int64_t>(nb, n - k);
    if (c1 <= c0) return;
    RB_CK2(rocblas_set_stream(rb, s));
    hipEvent_t e0 = W.prof_begin(s);
    RB_CK2(rocsolver_dlaswp(rb, (rocblas_int)(c1 - c0), A_(0, c0), (rocblas_int)lda, (rocblas_int)(k + 1), (rocblas_int)(k + kb), ipiv, 1));
    W.prof_end(e0, "lu64_laswp", 16.0 * kb * (c1 - c0), 0, s);
    const double one = 1.0;
    RB_CK2(rocblas_set_pointer_mode(rb, rocblas_pointer_mode_host));
    hipEvent_t e1 = W.prof_begin(s);
    RB_CK2(rocblas_dtrsm(rb, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_unit, kb,
                         (rocblas_int)(c1 - c0), &one, A_(k, k), (rocblas_int)lda, A_(k, c0), (rocblas_int)lda));
    W.prof_end(e1, "lu64_trsm", 16.0 * kb * (c1 - c0), (double)kb * kb * (c1 - c0), s);
    hipEvent_t e2 = W.prof_begin(s);
    dgemm_update(n - k - kb, c1 - c0, kb, A_(k + kb, k), lda, A_(k, c0), lda, A_(k + kb, c0), lda, s);
    W.prof_end(e2, "lu64_gemm", 16.0 * (n - k - kb) * (c1 - c0), 2.0 * (n - k - kb) * (c1 - c0) * kb, s);
  };
  auto swap_left = [&](int64_t k, hipStream_t s) {   // block k's interchanges on the factored columns [0, k)
    const int kb = (int)std::min<int64_t>(nb, n - k);
    if (k == 0) return;
    RB_CK2(rocblas_set_stream(rb, s));
    RB_CK2(rocsolver_dlaswp(rb, (rocblas_int)k, A_(0, 0), (rocblas_int)lda, (rocblas_int)(k + 1), (rocblas_int)(k + kb), ipiv, 1));
  };
  if (!la) {
    for (int b = 0; b < nblk; b++) {
      const int64_t k = (int64_t)b * nb;
      panel(k, s0);
      swap_left(k, s0);
      apply(k, k + nb, n, s0);
    }
  } else {
    panel(0, s0);
    for (int b = 0; b < nblk; b++) {
      const int64_t k = (int64_t)b * nb, kn = k + nb;
      swap_left(k, s0);
      if (kn < n) {
        const int64_t kn_end = std::min<int64_t>(kn + nb, n);
        apply(k, kn, kn_end, s0);                      // the next panel's columns first
        HIP_CK(hipEventRecord(W.ev_next, s0));
        HIP_CK(hipStreamWaitEvent(s1, W.ev_next, 0));
        panel(kn, s1);                                 // next panel on the second stream ...
        HIP_CK(hipEventRecord(W.ev_panel, s1));
        apply(k, kn_end, n, s0);                       // ... while the rest of the trailing matrix is updated
        HIP_CK(hipStreamWaitEvent(s0, W.ev_panel, 0));
      }
    }
    RB_CK2(rocblas_set_stream(rb, s0));
  }
  // first zero pivot over the blocks: rocSOLVER reports it per panel, relative to the panel
  std::vector<int32_t> bi(nblk);
  HIP_CK(hipMemcpyAsync(bi.data(), W.getrf64_info.p, (size_t)nblk * 4, hipMemcpyDeviceToHost, s0));
  HIP_CK(hipStreamSynchronize(s0));
  int info = 0;
  for (int b = 0; b < nblk && info == 0; b++) if (bi[b] > 0) info = b * nb + bi[b];
  if (info_out) *info_out = info;
}

}  // namespace mpk
