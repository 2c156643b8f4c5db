// k_resid.hip — fused FP64 and double-double residual kernels with the per-RHS metric epilogue.
//
// Computes, for a column-major A (FP64 or FP32 storage, promoted exactly to FP64):
//   R = B - A X   (or W = A X in matvec mode), T = |A||X| + |B|
// and per right-hand side: max|R|, max|X|, max_i |R_i|/T_i (0/0 -> 0, x/0 -> inf) with its row, sum R^2.
// This is the arithmetic of the reference's k_resid / k_colstats (src/mp_kernels.cuh:67-78, 92-96)
// and k_resid_dd (src/mp_kernels.cuh:83-87), laid out for MI355X instead of one thread per row:
//   * split-K over column chunks: grid = row-blocks x chunks, so a 16384^2 system launches ~2k
//     workgroups (the reference's one-thread-per-row launch fills 64 of 256 CUs);
//   * every wave streams 64 consecutive rows of a column (512 B coalesced loads), 8 columns unrolled
//     so each lane keeps 8 independent HBM loads in flight; X is read through the scalar cache;
//   * the chunk partials are combined in a fixed order by a second kernel, which also fuses the
//     metric epilogue, so results are deterministic and do not depend on scheduling.
// Summation order therefore differs from the reference's sequential-j order: the FP64 result agrees
// with it to within the usual (n+2)u|A||X| bound; the DD result agrees with the exact residual to
// ~n*2^-104 relative to |A||X|.
#include "mp_internal.hpp"
#include "mp_kernels.hpp"

namespace mpk {

__device__ __forceinline__ double dmax_nan(double a, double b) {  // NaN-propagating max
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}

// Partial sums over one column chunk.  P layout: P[((chunk*F + f)*NR + c)*n + i], F = DD ? 3 : 2,
// fields: 0 = sum a*x (FP64) or DD hi of -sum a*x, 1 = sum |a||x|, 2 = DD lo.
template <typename TA, int NR, bool DD>
__global__ __launch_bounds__(256) void k_resid_partial(const TA* __restrict__ A, int64_t n, int64_t lda,
                                                       const double* __restrict__ X, int64_t ldx, int64_t cpc,
                                                       double* __restrict__ P, const int* __restrict__ act) {
  if (act && *act == 0) return;
  constexpr int F = DD ? 3 : 2;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t chunk = blockIdx.y;
  const int64_t j0 = chunk * cpc;
  const int64_t jl = (j0 + cpc <= n) ? cpc : (n - j0);
  if (i >= n || jl <= 0) return;
  double s[NR], t[NR], lo[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) { s[c] = 0.0; t[c] = 0.0; lo[c] = 0.0; }
  const TA* ap = A + j0 * lda + i;
  const double* xp = X + j0;
  int64_t jj = 0;
  for (; jj + 8 <= jl; jj += 8) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; u++) a[u] = (double)__builtin_nontemporal_load(ap + (jj + u) * lda);
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const double aa = fabs(a[u]);
#pragma unroll
      for (int c = 0; c < NR; c++) {
        const double x = xp[c * ldx + jj + u];
        if (DD) {
          const double p = -a[u] * x;
          const double pe = fma(-a[u], x, -p);
          const double sn = s[c] + p;
          const double bb = sn - s[c];
          const double e = (s[c] - (sn - bb)) + (p - bb);
          s[c] = sn; lo[c] += e + pe;
        } else {
          s[c] = fma(a[u], x, s[c]);
        }
        t[c] = fma(aa, fabs(x), t[c]);
      }
    }
  }
  for (; jj < jl; jj++) {
    const double av = (double)ap[jj * lda];
    const double aa = fabs(av);
#pragma unroll
    for (int c = 0; c < NR; c++) {
      const double x = xp[c * ldx + jj];
      if (DD) {
        const double p = -av * x;
        const double pe = fma(-av, x, -p);
        const double sn = s[c] + p;
        const double bb = sn - s[c];
        const double e = (s[c] - (sn - bb)) + (p - bb);
        s[c] = sn; lo[c] += e + pe;
      } else {
        s[c] = fma(av, x, s[c]);
      }
      t[c] = fma(aa, fabs(x), t[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < NR; c++) {
    double* base = P + ((chunk * F) * NR + c) * n + i;
    base[0] = s[c];
    base[(int64_t)NR * n] = t[c];
    if (DD) base[(int64_t)2 * NR * n] = lo[c];
  }
}

// One pass over A for two single-column products: column 1 with k_resid_partial<TA, 1, false>'s FP64 sums into
// P1, column 2 with k_resid_partial<TA, 1, DD2>'s sums into P2 (each in its own layout).  Every accumulator sees
// the operations of its own kernel in the same order, so both sets of partials are bitwise theirs.
template <typename TA, bool DD2>
__global__ __launch_bounds__(256) void k_resid_partial_pair(const TA* __restrict__ A, int64_t n, int64_t lda,
                                                            const double* __restrict__ X1, const double* __restrict__ X2,
                                                            int64_t cpc, double* __restrict__ P1, double* __restrict__ P2) {
  constexpr int F2 = DD2 ? 3 : 2;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t chunk = blockIdx.y;
  const int64_t j0 = chunk * cpc;
  const int64_t jl = (j0 + cpc <= n) ? cpc : (n - j0);
  if (i >= n || jl <= 0) return;
  double s1 = 0.0, t1 = 0.0, s2 = 0.0, t2 = 0.0, lo2 = 0.0;
  const TA* ap = A + j0 * lda + i;
  const double* x1p = X1 + j0;
  const double* x2p = X2 + j0;
  auto col = [&](double av, double x1, double x2) {
    const double aa = fabs(av);
    s1 = fma(av, x1, s1);
    t1 = fma(aa, fabs(x1), t1);
    if (DD2) {
      const double p = -av * x2;
      const double pe = fma(-av, x2, -p);
      const double sn = s2 + p;
      const double bb = sn - s2;
      const double e = (s2 - (sn - bb)) + (p - bb);
      s2 = sn; lo2 += e + pe;
    } else {
      s2 = fma(av, x2, s2);
    }
    t2 = fma(aa, fabs(x2), t2);
  };
  int64_t jj = 0;
  for (; jj + 8 <= jl; jj += 8) {
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; u++) a[u] = (double)__builtin_nontemporal_load(ap + (jj + u) * lda);
#pragma unroll
    for (int u = 0; u < 8; u++) col(a[u], x1p[jj + u], x2p[jj + u]);
  }
  for (; jj < jl; jj++) col((double)ap[jj * lda], x1p[jj], x2p[jj]);
  double* b1 = P1 + (chunk * 2) * n + i;
  b1[0] = s1; b1[n] = t1;
  double* b2 = P2 + (chunk * F2) * n + i;
  b2[0] = s2; b2[n] = t2;
  if (DD2) b2[2 * n] = lo2;
}

// Combine the chunk partials in chunk order, write R/T (and the DD low part), and reduce the per-block
// metric partials.  grid = (ceil(n/256), NR); one workgroup handles 256 rows of one right-hand side.
// Sp layout: Sp[(c_global*nblk + blk)*6 + k], k: 0 max|R|, 1 max|X|, 2 max q, 3 sum R^2, 4 worst row.
template <int NR, bool DD>
__global__ __launch_bounds__(256) void k_resid_reduce(const double* __restrict__ P, int64_t n, int64_t nchunks,
                                                      const double* __restrict__ X, int64_t ldx,
                                                      const double* __restrict__ B, int64_t ldb,
                                                      double* __restrict__ R, double* __restrict__ Rlo, int64_t ldr,
                                                      double* __restrict__ T, int64_t ldt, int matvec,
                                                      double* __restrict__ Sp, int c_base, const int* __restrict__ act) {
  if (act && *act == 0) return;
  constexpr int F = DD ? 3 : 2;
  const int c = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0; int64_t wi = -1;
  if (i < n) {
    const double bv = (B && !matvec) ? B[(int64_t)c * ldb + i] : 0.0;
    double r, rlo = 0.0, tt = 0.0;
    // Chunks are combined in ascending order; eight chunks' loads are issued before their additions, so a
    // thread waits on one memory latency per eight chunks instead of one per chunk.
    const int64_t fs = (int64_t)F * NR * n;  // stride between consecutive chunks' fields
    const double* base = P + (int64_t)c * n + i;
    if (DD) {
      double H = bv, L = 0.0;
      int64_t k = 0;
      for (; k + 8 <= nchunks; k += 8) {
        double hv[8], lv[8], tv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const double* b = base + (k + u) * fs;
          hv[u] = b[0]; tv[u] = b[(int64_t)NR * n]; lv[u] = b[(int64_t)2 * NR * n];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          tt += tv[u];
          const double h = hv[u];
          const double sn = H + h; const double bb = sn - H; const double e = (H - (sn - bb)) + (h - bb);
          H = sn; L = L + (lv[u] + e);
        }
      }
      for (; k < nchunks; k++) {
        const double* b = base + k * fs;
        const double h = b[0], l = b[(int64_t)2 * NR * n];
        tt += b[(int64_t)NR * n];
        const double sn = H + h; const double bb = sn - H; const double e = (H - (sn - bb)) + (h - bb);
        H = sn; L = L + (l + e);
      }
      if (matvec) { H = -H; L = -L; }
      r = H + L; rlo = L - (r - H);
    } else {
      double s = 0.0;
      int64_t k = 0;
      for (; k + 8 <= nchunks; k += 8) {
        double sv[8], tv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) { const double* b = base + (k + u) * fs; sv[u] = b[0]; tv[u] = b[(int64_t)NR * n]; }
#pragma unroll
        for (int u = 0; u < 8; u++) { s += sv[u]; tt += tv[u]; }
      }
      for (; k < nchunks; k++) { const double* b = base + k * fs; s += b[0]; tt += b[(int64_t)NR * n]; }
      r = matvec ? s : (bv - s);
    }
    tt += fabs(bv);
    R[(int64_t)c * ldr + i] = r;
    if (DD && Rlo) Rlo[(int64_t)c * ldr + i] = rlo;
    if (T) T[(int64_t)c * ldt + i] = tt;
    const double ar = fabs(r), ax = fabs(X[(int64_t)c * ldx + i]);
    double q;
    if (ar == 0.0) q = 0.0; else if (tt == 0.0) q = INFINITY; else q = ar / tt;
    m0 = ar; m1 = ax; m2 = q; m3 = r * r; wi = i;
  }
  // block reduction: max with NaN propagation, worst row = smallest index among maximal q
  __shared__ double s0[256], s1[256], s2[256], s3[256];
  __shared__ int64_t si[256];
  const int t = threadIdx.x;
  s0[t] = m0; s1[t] = m1; s2[t] = m2; s3[t] = m3; si[t] = wi;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) {
      s0[t] = dmax_nan(s0[t], s0[t + st]);
      s1[t] = dmax_nan(s1[t], s1[t + st]);
      const double qa = s2[t], qb = s2[t + st];
      const int64_t ia = si[t], ib = si[t + st];
      bool take_b;
      if (ib < 0) take_b = false;
      else if (ia < 0) take_b = true;
      else if (qa != qa) take_b = false;
      else if (qb != qb) take_b = true;
      else take_b = (qb > qa) || (qb == qa && ib < ia);
      if (take_b) { s2[t] = qb; si[t] = ib; }
      s3[t] += s3[t + st];
    }
    __syncthreads();
  }
  if (t == 0 && Sp) {
    const int64_t nblk = gridDim.x;
    double* o = Sp + ((int64_t)(c_base + c) * nblk + blockIdx.x) * 6;
    o[0] = s0[0]; o[1] = s1[0]; o[2] = s2[0]; o[3] = s3[0]; o[4] = (double)si[0];
  }
}

// Final fixed-order reduction over blocks: stats[c*4 + {0..3}], worst[c].
__global__ __launch_bounds__(256) void k_resid_stats_final(const double* __restrict__ Sp, int64_t nblk, int nrhs,
                                                           double* __restrict__ stats, int32_t* __restrict__ worst) {
  const int c = blockIdx.x;
  if (c >= nrhs) return;
  __shared__ double s0[256], s1[256], s2[256], s3[256];
  __shared__ int64_t si[256];
  const int t = threadIdx.x;
  double m0 = 0, m1 = 0, m2 = 0, m3 = 0; int64_t wi = -1;
  for (int64_t b = t; b < nblk; b += 256) {
    const double* o = Sp + ((int64_t)c * nblk + b) * 6;
    m0 = dmax_nan(m0, o[0]); m1 = dmax_nan(m1, o[1]); m3 += o[3];
    const double q = o[2]; const int64_t qi = (int64_t)o[4];
    if (qi >= 0) {
      bool take;
      if (wi < 0) take = true;
      else if (m2 != m2) take = false;
      else if (q != q) take = true;
      else take = (q > m2) || (q == m2 && qi < wi);
      if (take) { m2 = q; wi = qi; }
    }
  }
  s0[t] = m0; s1[t] = m1; s2[t] = m2; s3[t] = m3; si[t] = wi;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) {
      s0[t] = dmax_nan(s0[t], s0[t + st]);
      s1[t] = dmax_nan(s1[t], s1[t + st]);
      const double qa = s2[t], qb = s2[t + st];
      const int64_t ia = si[t], ib = si[t + st];
      bool take_b;
      if (ib < 0) take_b = false;
      else if (ia < 0) take_b = true;
      else if (qa != qa) take_b = false;
      else if (qb != qb) take_b = true;
      else take_b = (qb > qa) || (qb == qa && ib < ia);
      if (take_b) { s2[t] = qb; si[t] = ib; }
      s3[t] += s3[t + st];
    }
    __syncthreads();
  }
  if (t == 0) {
    if (stats) { stats[c * 4 + 0] = s0[0]; stats[c * 4 + 1] = s1[0]; stats[c * 4 + 2] = s2[0]; stats[c * 4 + 3] = s3[0]; }
    if (worst) worst[c] = (int32_t)si[0];
  }
}

// Row sums of |A| (for ||A||_inf) and max|A| (for the global power-of-two scale), one pass.
template <typename TA>
__global__ __launch_bounds__(256) void k_absmax_partial(const TA* __restrict__ A, int64_t n, int64_t lda, int64_t cpc,
                                                        double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t j0 = blockIdx.y * cpc;
  const int64_t jl = (j0 + cpc <= n) ? cpc : (n - j0);
  double m = 0.0;
  if (i < n && jl > 0) {
    // eight columns' loads in flight per thread (the maximum does not depend on the order)
    const TA* col = A + j0 * lda + i;
    int64_t j = 0;
    for (; j + 8 <= jl; j += 8) {
      TA v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = col[(j + u) * lda];
#pragma unroll
      for (int u = 0; u < 8; u++) m = dmax_nan(m, fabs((double)v[u]));
    }
    for (; j < jl; j++) m = dmax_nan(m, fabs((double)col[j * lda]));
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) { if (threadIdx.x < st) sh[threadIdx.x] = dmax_nan(sh[threadIdx.x], sh[threadIdx.x + st]); __syncthreads(); }
  if (threadIdx.x == 0) out[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

__global__ void k_max_final(const double* __restrict__ in, int64_t cnt, double* __restrict__ out) {
  __shared__ double sh[256];
  double m = 0.0;
  for (int64_t k = threadIdx.x; k < cnt; k += 256) m = dmax_nan(m, in[k]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) { if (threadIdx.x < st) sh[threadIdx.x] = dmax_nan(sh[threadIdx.x], sh[threadIdx.x + st]); __syncthreads(); }
  if (threadIdx.x == 0) out[0] = sh[0];
}

}  // namespace mpk

// ------------------------------------------------------------------ host launchers
namespace mpk {

// Column chunks of the partial-sum pass: about 1024 workgroups in all (16 chunks of 1024 columns at n = 16384
// ran at 6.9 TB/s against 6.5 TB/s for 32 chunks of 512), at least 64 columns each.
static void pick_chunks(int64_t n, int64_t& cpc, int64_t& nchunks) {
  const int64_t rb = cdiv(n, 256);
  int64_t want = std::max<int64_t>(1, 1024 / rb);
  want = std::min<int64_t>(want, std::max<int64_t>(1, (n + 63) / 64));
  cpc = ((cdiv(n, want) + 7) / 8) * 8;
  if (cpc < 8) cpc = 8;
  nchunks = cdiv(n, cpc);
}

void product_chunks(int64_t n, int64_t& cpc, int64_t& nch) { pick_chunks(n, cpc, nch); }

void product_reduce(const double* P, int64_t n, int64_t nch, const double* X, double* w, Workspace& W, const int* act) {
  W.statp.ensure(*W.tracker, (size_t)cdiv(n, 256) * 6, "resid_stat_partials");
  hipLaunchKernelGGL((k_resid_reduce<1, false>), dim3(cdiv(n, 256), 1), dim3(256), 0, W.stream, P, n, nch, X, n,
                     (const double*)nullptr, (int64_t)0, w, (double*)nullptr, n, (double*)nullptr, (int64_t)0, 1,
                     W.statp.p, 0, act);
  LAUNCH_CK();
}

template <typename TA, int NR, bool DD>
static void launch_group_into(const ResidArgs& a, Workspace& W, int c0, double* part, double* statp, bool prof) {
  int64_t cpc, nch;
  pick_chunks(a.n, cpc, nch);
  const int64_t rb = cdiv(a.n, 256);
  const TA* A = (const TA*)a.A;
  hipEvent_t ev = prof ? W.prof_begin() : nullptr;
  hipLaunchKernelGGL((k_resid_partial<TA, NR, DD>), dim3(rb, nch), dim3(256), 0, W.stream,
                     A, a.n, a.lda, a.X + (int64_t)c0 * a.ldx, a.ldx, cpc, part, a.active);
  LAUNCH_CK();
  // algorithmic bytes: A once, X once, the chunk partials written (the reduce pass reads them back)
  if (prof)
    W.prof_end(ev, DD ? "residual_dd_partial" : (a.matvec ? "matvec_partial" : "residual_partial"),
               (double)sizeof(TA) * a.n * a.n + 8.0 * a.n * NR, (DD ? 20.0 : 4.0) * a.n * a.n * NR);
  hipLaunchKernelGGL((k_resid_reduce<NR, DD>), dim3(rb, NR), dim3(256), 0, W.stream,
                     part, a.n, nch, a.X + (int64_t)c0 * a.ldx, a.ldx,
                     a.B ? a.B + (int64_t)c0 * a.ldb : nullptr, a.ldb,
                     a.R + (int64_t)c0 * a.ldr, a.Rlo ? a.Rlo + (int64_t)c0 * a.ldr : nullptr, a.ldr,
                     a.T ? a.T + (int64_t)c0 * a.ldt : nullptr, a.ldt, a.matvec ? 1 : 0,
                     statp, c0, a.active);
  LAUNCH_CK();
}

template <typename TA, int NR, bool DD>
static void launch_group(const ResidArgs& a, Workspace& W, int c0) {
  int64_t cpc, nch;
  pick_chunks(a.n, cpc, nch);
  const int F = DD ? 3 : 2;
  W.part.ensure(*W.tracker, (size_t)nch * F * NR * a.n, "resid_partials");
  launch_group_into<TA, NR, DD>(a, W, c0, W.part.p, W.statp.p, true);
}

template <typename TA, bool DD>
static void launch_all(const ResidArgs& a, Workspace& W) {
  int c = 0;
  while (c < a.nrhs) {
    const int left = a.nrhs - c;
    if (left >= 8) { launch_group<TA, 8, DD>(a, W, c); c += 8; }
    else if (left >= 4) { launch_group<TA, 4, DD>(a, W, c); c += 4; }
    else if (left >= 2) { launch_group<TA, 2, DD>(a, W, c); c += 2; }
    else { launch_group<TA, 1, DD>(a, W, c); c += 1; }
  }
}

static void check_resid_args(const ResidArgs& a) {
  MP_REQUIRE(a.n > 0 && a.nrhs > 0, "n and nrhs must be positive");
  MP_REQUIRE(a.lda >= a.n && a.ldx >= a.n && a.ldr >= a.n, "leading dimensions must be >= n");
  MP_REQUIRE(a.A && a.X && a.R, "A, X and R are required");
  MP_REQUIRE(!a.T || a.ldt >= a.n, "ldt must be >= n");
  MP_REQUIRE(!a.B || a.ldb >= a.n, "ldb must be >= n");
}

void residual_pair(const ResidPairArgs& a, Workspace& W) {
  MP_REQUIRE(a.n > 0 && a.A && a.X1 && a.B1 && a.R1 && a.X2 && a.B2 && a.R2 && a.T2 && a.stats2 && a.worst2,
             "residual_pair: every operand is required");
  MP_REQUIRE(!a.dd2 || a.R2lo, "residual_pair: the double-double residual needs its low part");
  const int64_t n = a.n;
  int64_t cpc, nch;
  pick_chunks(n, cpc, nch);
  const int64_t rb = cdiv(n, 256);
  const int F2 = a.dd2 ? 3 : 2;
  W.part.ensure(*W.tracker, (size_t)nch * (2 + F2) * n, "resid_partials");
  W.statp.ensure(*W.tracker, (size_t)rb * 6, "resid_stat_partials");
  double* P1 = W.part.p;
  double* P2 = W.part.p + (size_t)nch * 2 * n;
  hipEvent_t ev = W.prof_begin();
  if (a.a32) {
    if (a.dd2) hipLaunchKernelGGL((k_resid_partial_pair<float, true>), dim3(rb, nch), dim3(256), 0, W.stream, (const float*)a.A, n, n, a.X1, a.X2, cpc, P1, P2);
    else hipLaunchKernelGGL((k_resid_partial_pair<float, false>), dim3(rb, nch), dim3(256), 0, W.stream, (const float*)a.A, n, n, a.X1, a.X2, cpc, P1, P2);
  } else {
    if (a.dd2) hipLaunchKernelGGL((k_resid_partial_pair<double, true>), dim3(rb, nch), dim3(256), 0, W.stream, (const double*)a.A, n, n, a.X1, a.X2, cpc, P1, P2);
    else hipLaunchKernelGGL((k_resid_partial_pair<double, false>), dim3(rb, nch), dim3(256), 0, W.stream, (const double*)a.A, n, n, a.X1, a.X2, cpc, P1, P2);
  }
  LAUNCH_CK();
  W.prof_end(ev, "residual_pair_partial", (a.a32 ? 4.0 : 8.0) * n * n + 16.0 * n, (a.dd2 ? 24.0 : 8.0) * n * n);
  // column 1 as residual() reduces a one-column FP64 residual without statistics, then column 2 with them
  hipLaunchKernelGGL((k_resid_reduce<1, false>), dim3(rb, 1), dim3(256), 0, W.stream, P1, n, nch, a.X1, n, a.B1, n, a.R1,
                     (double*)nullptr, n, (double*)nullptr, (int64_t)0, 0, W.statp.p, 0, (const int*)nullptr);
  LAUNCH_CK();
  if (a.dd2)
    hipLaunchKernelGGL((k_resid_reduce<1, true>), dim3(rb, 1), dim3(256), 0, W.stream, P2, n, nch, a.X2, n, a.B2, n, a.R2,
                       a.R2lo, n, a.T2, n, 0, W.statp.p, 0, (const int*)nullptr);
  else
    hipLaunchKernelGGL((k_resid_reduce<1, false>), dim3(rb, 1), dim3(256), 0, W.stream, P2, n, nch, a.X2, n, a.B2, n, a.R2,
                       (double*)nullptr, n, a.T2, n, 0, W.statp.p, 0, (const int*)nullptr);
  LAUNCH_CK();
  hipLaunchKernelGGL(k_resid_stats_final, dim3(1), dim3(256), 0, W.stream, W.statp.p, rb, 1, a.stats2, a.worst2);
  LAUNCH_CK();
}

void residual(const ResidArgs& a, Workspace& W) {
  check_resid_args(a);
  const int64_t nblk = cdiv(a.n, 256);
  W.statp.ensure(*W.tracker, (size_t)a.nrhs * nblk * 6, "resid_stat_partials");
  if (a.a32) { if (a.dd) launch_all<float, true>(a, W); else launch_all<float, false>(a, W); }
  else { if (a.dd) launch_all<double, true>(a, W); else launch_all<double, false>(a, W); }
  if (a.stats || a.worst) {
    hipLaunchKernelGGL(k_resid_stats_final, dim3(a.nrhs), dim3(256), 0, W.stream, W.statp.p, nblk, a.nrhs, a.stats, a.worst);
    LAUNCH_CK();
  }
}

// ------------------------------------------------------------------ residual without allocation
// A fixed device array holds the chunk partials and the per-block statistics of residual_noalloc, so that
// entry point allocates nothing.  Each GPU holds its own copy (at its own address) until the process ends.
// In host code `scratch` is only a stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, taken once per device when a context is created (resid_scratch_init).
static __device__ __align__(256) double scratch[RESID_SCRATCH_PART + RESID_SCRATCH_STAT];

double* resid_scratch_init() {
  double* ptr = nullptr;
  HIP_CK(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
  MP_REQUIRE(ptr != nullptr, "residual scratch: no device address");
  return ptr;
}

bool resid_noalloc_fits(int64_t n, int nrhs, bool dd) {
  if (n <= 0 || nrhs <= 0) return false;
  int64_t cpc, nch;
  pick_chunks(n, cpc, nch);
  const int64_t F = dd ? 3 : 2;
  return nch * F * n <= RESID_SCRATCH_PART && (int64_t)nrhs * cdiv(n, 256) * 6 <= RESID_SCRATCH_STAT;
}

// The residual of `residual` with the right-hand sides taken one at a time (each right-hand side's sums do not
// depend on how many are processed together, so the results are the same bits), partials in the scratch
// array at `scratch_dev` (its device address).  No allocation, no memset, no profiling events.
void residual_noalloc(const ResidArgs& a, Workspace& W, double* scratch_dev) {
  check_resid_args(a);
  MP_REQUIRE(scratch_dev != nullptr, "residual scratch was not set up for this device");
  if (!resid_noalloc_fits(a.n, a.nrhs, a.dd)) {
    std::ostringstream o;
    o << "invalid argument: n = " << a.n << ", nrhs = " << a.nrhs << " exceeds the scratch of the allocation-free residual ("
      << RESID_SCRATCH_PART << " partial words, " << RESID_SCRATCH_STAT << " statistics words)";
    throw MPError(o.str(), -1);
  }
  double* part = scratch_dev;
  double* statp = scratch_dev + RESID_SCRATCH_PART;
  for (int c = 0; c < a.nrhs; c++) {
    if (a.a32) { if (a.dd) launch_group_into<float, 1, true>(a, W, c, part, statp, false); else launch_group_into<float, 1, false>(a, W, c, part, statp, false); }
    else { if (a.dd) launch_group_into<double, 1, true>(a, W, c, part, statp, false); else launch_group_into<double, 1, false>(a, W, c, part, statp, false); }
  }
  if (a.stats || a.worst) {
    hipLaunchKernelGGL(k_resid_stats_final, dim3(a.nrhs), dim3(256), 0, W.stream, statp, cdiv(a.n, 256), a.nrhs, a.stats, a.worst);
    LAUNCH_CK();
  }
}

int64_t absmax_partials(int64_t n) {
  int64_t cpc, nch;
  pick_chunks(n, cpc, nch);
  return cdiv(n, 256) * nch;
}

void absmax_on(const void* A, bool a32, int64_t n, int64_t lda, double* out_dev, double* part, hipStream_t stream) {
  int64_t cpc, nch;
  pick_chunks(n, cpc, nch);
  const int64_t rb = cdiv(n, 256);
  if (a32) hipLaunchKernelGGL((k_absmax_partial<float>), dim3(rb, nch), dim3(256), 0, stream, (const float*)A, n, lda, cpc, part);
  else hipLaunchKernelGGL((k_absmax_partial<double>), dim3(rb, nch), dim3(256), 0, stream, (const double*)A, n, lda, cpc, part);
  LAUNCH_CK();
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, stream, part, (int64_t)rb * nch, out_dev);
  LAUNCH_CK();
}

void absmax(const void* A, bool a32, int64_t n, int64_t lda, double* out_dev, Workspace& W) {
  W.part.ensure(*W.tracker, (size_t)absmax_partials(n), "absmax_partials");
  absmax_on(A, a32, n, lda, out_dev, W.part.p, W.stream);
}

}  // namespace mpk
