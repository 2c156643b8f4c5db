// k_krylov.hip — FP64 vector kernels of the FGMRES inner solver: block dot products V^T w, w -/+= V h,
// 2-norms, scaling, and a device-resident Hessenberg / Givens / least-squares state.
//
// The reference runs one-pass modified Gram-Schmidt with one cublasDdot + one axpy launch per basis vector
// and copies every Hessenberg column to the host for the Givens rotations (src/mp_core.cuh:255-263).
// Here each Arnoldi step does classical Gram-Schmidt twice (CGS2): one block GEMV-T + one block GEMV per
// pass, with all reductions in a fixed order (chunk partials combined in chunk order) so results do not
// depend on scheduling.  The rotations and the recursive residual live on the device, and so does the decision
// to end a cycle: the Givens kernel clears an activity word when the step converged, broke down or met a
// non-finite value, and every kernel of an Arnoldi step given that word returns at once when it is clear.  The
// host can therefore enqueue step j+1 before it reads step j's three numbers (|g_{j+1}|/beta0, h_{j+1,j}, a
// finiteness flag); a step enqueued past the end of the cycle does nothing.
#include "mp_internal.hpp"
#include "mp_kernels.hpp"
#include "mp_krylov.hpp"

namespace mpk {

constexpr int KCH = 256;   // rows per chunk of the reductions

// part[chunk*ncols + i] = sum over the chunk's rows of V[r, i] * w[r], i < ncols
__global__ __launch_bounds__(256) void k_gemvt_partial(const double* __restrict__ V, int64_t ldv, int ncols,
                                                       const double* __restrict__ w, int64_t n, double* __restrict__ part,
                                                       const int* __restrict__ act) {
  if (act && *act == 0) return;
  __shared__ double red[4][64];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t r = (int64_t)blockIdx.x * KCH + t;
  const double wr = (r < n) ? w[r] : 0.0;
  // eight columns' loads are issued before their eight wave reductions
  for (int i0 = 0; i0 < ncols; i0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = (r < n && i0 + u < ncols) ? V[(int64_t)(i0 + u) * ldv + r] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + u;
      if (i >= ncols) break;
      double s = v[u] * wr;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == 0) red[wv][i & 63] = s;
      if ((i & 63) == 63 || i == ncols - 1) {
        __syncthreads();
        const int base = i & ~63;
        const int cnt = i - base + 1;
        if (t < cnt) part[(int64_t)blockIdx.x * ncols + base + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
        __syncthreads();
      }
    }
  }
}

// The product's reduction and the first CGS pass's dot products in one launch (256 rows per workgroup, the
// row block of both): w[r] = sum over the product's column chunks of their partials in chunk order (the
// residual reduction's product mode), then k_gemvt_partial's partials of V^T w for these rows.  Both are the
// same bits as the two kernels apart.
__global__ __launch_bounds__(256) void k_reduce_dots(const double* __restrict__ P, int64_t nch, double* __restrict__ w,
                                                     const double* __restrict__ V, int64_t ldv, int ncols, int64_t n,
                                                     double* __restrict__ part, const int* __restrict__ act) {
  if (act && *act == 0) return;
  __shared__ double red[4][64];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t r = (int64_t)blockIdx.x * KCH + t;
  double wr = 0.0;
  if (r < n) {
    const int64_t fs = 2 * n;            // field stride of the one-column FP64 partials
    const double* base = P + r;
    double s = 0.0;
    int64_t k = 0;
    for (; k + 8 <= nch; k += 8) {
      double sv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) sv[u] = base[(k + u) * fs];
#pragma unroll
      for (int u = 0; u < 8; u++) s += sv[u];
    }
    for (; k < nch; k++) s += base[k * fs];
    wr = s;
    w[r] = s;
  }
  for (int i0 = 0; i0 < ncols; i0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = (r < n && i0 + u < ncols) ? V[(int64_t)(i0 + u) * ldv + r] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + u;
      if (i >= ncols) break;
      double sm = v[u] * wr;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) sm += __shfl_xor(sm, off, 64);
      if (lane == 0) red[wv][i & 63] = sm;
      if ((i & 63) == 63 || i == ncols - 1) {
        __syncthreads();
        const int base = i & ~63;
        const int cnt = i - base + 1;
        if (t < cnt) part[(int64_t)blockIdx.x * ncols + base + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
        __syncthreads();
      }
    }
  }
}

// out[i] (+)= sum over chunks in order; if acc != 0 the result is added to out[i]
__global__ void k_gemvt_final(const double* __restrict__ part, int nchunks, int ncols, double* __restrict__ out, int acc,
                              const int* __restrict__ act) {
  if (act && *act == 0) return;
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= ncols) return;
  double s = 0.0;
  for (int c = 0; c < nchunks; c++) s += part[(int64_t)c * ncols + i];
  out[i] = acc ? out[i] + s : s;
}

// w[r] += sign * sum_i V[r, i] * h[i]
__global__ __launch_bounds__(256) void k_gemv_update(const double* __restrict__ V, int64_t ldv, int ncols,
                                                     const double* __restrict__ h, double* __restrict__ w, int64_t n, double sign,
                                                     const int* __restrict__ act) {
  if (act && *act == 0) return;
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int i = 0; i < ncols; i++) s = fma(V[(int64_t)i * ldv + r], h[i], s);
  w[r] = fma(sign, s, w[r]);
}

// One CGS pass's second half: h[i] = sum of the chunk partials of V^T w in chunk order (k_gemvt_final's sum,
// formed by every workgroup into LDS; workgroup 0 stores it), then w[r] += sign * sum_i V[r, i] h[i] (k_gemv_update's
// order).  With sq != nullptr the chunk partials of sum w^2 of the updated w follow (k_sumsq_partial's order; the
// workgroup of 256 rows is one chunk).  ncols <= 256.
__device__ __forceinline__ void cgs_update_body(const double* __restrict__ part, int nchunks, int ncols,
                                                const double* __restrict__ V, int64_t ldv, double* __restrict__ hout,
                                                double* __restrict__ w, int64_t n, double sign, double* __restrict__ sq) {
  __shared__ double hs[256];
  __shared__ double red[4];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  // Both sums keep their ascending order; each issues eight loads before the eight additions that use them.
  if (t < ncols) {
    double s = 0.0;
    int c = 0;
    for (; c + 8 <= nchunks; c += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = part[(int64_t)(c + u) * ncols + t];
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; c < nchunks; c++) s += part[(int64_t)c * ncols + t];
    hs[t] = s;
    if (blockIdx.x == 0) hout[t] = s;
  }
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * 256 + t;
  double wn = 0.0;
  if (r < n) {
    double s = 0.0;
    int i = 0;
    for (; i + 8 <= ncols; i += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = V[(int64_t)(i + u) * ldv + r];
#pragma unroll
      for (int u = 0; u < 8; u++) s = fma(v[u], hs[i + u], s);
    }
    for (; i < ncols; i++) s = fma(V[(int64_t)i * ldv + r], hs[i], s);
    wn = fma(sign, s, w[r]);
    w[r] = wn;
  }
  if (sq) {
    double q = (r < n) ? wn * wn : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off, 64);
    if (lane == 0) red[wv] = q;
    __syncthreads();
    if (t == 0) sq[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
  }
}

__global__ __launch_bounds__(256) void k_cgs_update(const double* __restrict__ part, int nchunks, int ncols,
                                                    const double* __restrict__ V, int64_t ldv, double* __restrict__ hout,
                                                    double* __restrict__ w, int64_t n, double sign, double* __restrict__ sq,
                                                    const int* __restrict__ act) {
  if (act && *act == 0) return;
  cgs_update_body(part, nchunks, ncols, V, ldv, hout, w, n, sign, sq);
}

// The first CGS pass's update and the second pass's dot products in one launch (256 rows per workgroup, the row
// block of both): k_cgs_update without the norm partials, then k_gemvt_partial's partials of V^T w for the
// updated rows, into part_out (not part_in: other workgroups may still be summing part_in).  The same bits as the
// two kernels apart.
__global__ __launch_bounds__(256) void k_cgs_update_dots(const double* __restrict__ part_in, int nchunks, int ncols,
                                                         const double* __restrict__ V, int64_t ldv, double* __restrict__ hout,
                                                         double* __restrict__ w, int64_t n, double sign,
                                                         double* __restrict__ part_out, const int* __restrict__ act) {
  if (act && *act == 0) return;
  __shared__ double hs[256];
  __shared__ double red[4][64];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (t < ncols) {
    double s = 0.0;
    int c = 0;
    for (; c + 8 <= nchunks; c += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = part_in[(int64_t)(c + u) * ncols + t];
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; c < nchunks; c++) s += part_in[(int64_t)c * ncols + t];
    hs[t] = s;
    if (blockIdx.x == 0) hout[t] = s;
  }
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * 256 + t;
  double wr = 0.0;
  if (r < n) {
    double s = 0.0;
    int i = 0;
    for (; i + 8 <= ncols; i += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = V[(int64_t)(i + u) * ldv + r];
#pragma unroll
      for (int u = 0; u < 8; u++) s = fma(v[u], hs[i + u], s);
    }
    for (; i < ncols; i++) s = fma(V[(int64_t)i * ldv + r], hs[i], s);
    wr = fma(sign, s, w[r]);
    w[r] = wr;
  }
  for (int i0 = 0; i0 < ncols; i0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = (r < n && i0 + u < ncols) ? V[(int64_t)(i0 + u) * ldv + r] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + u;
      if (i >= ncols) break;
      double sm = v[u] * wr;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) sm += __shfl_xor(sm, off, 64);
      if (lane == 0) red[wv][i & 63] = sm;
      if ((i & 63) == 63 || i == ncols - 1) {
        __syncthreads();
        const int base = i & ~63;
        const int cnt = i - base + 1;
        if (t < cnt) part_out[(int64_t)blockIdx.x * ncols + base + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
        __syncthreads();
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_sumsq_partial(const double* __restrict__ w, int64_t n, double* __restrict__ part,
                                                       const int* __restrict__ act) {
  if (act && *act == 0) return;
  __shared__ double red[4];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t r = (int64_t)blockIdx.x * KCH + t;
  double s = (r < n) ? w[r] * w[r] : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) red[wv] = s;
  __syncthreads();
  if (t == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// Start of a right-preconditioned solve in one pass: d = 0, rc = r, and k_sumsq_partial's chunk partials of r^2
// (the same order, so the norm that follows is that of rc bit for bit).
__global__ __launch_bounds__(256) void k_gmres_start(const double* __restrict__ r, double* __restrict__ rc,
                                                     double* __restrict__ d, int64_t n, double* __restrict__ part) {
  __shared__ double red[4];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t i = (int64_t)blockIdx.x * KCH + t;
  double v = 0.0;
  if (i < n) { v = r[i]; rc[i] = v; d[i] = 0.0; }
  double s = (i < n) ? v * v : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) red[wv] = s;
  __syncthreads();
  if (t == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// out[0] = sqrt(sum of partials)
__global__ void k_norm_final(const double* __restrict__ part, int nchunks, double* __restrict__ out, const int* __restrict__ act) {
  if (act && *act == 0) return;
  __shared__ double red[256];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int c = t; c < nchunks; c += 256) s += part[c];
  red[t] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) { if (t < st) red[t] += red[t + st]; __syncthreads(); }
  if (t == 0) out[0] = sqrt(red[0]);
}

// y[r] = x[r] * (*s) ; s == nullptr -> y = x * sv
__global__ void k_scale_copy(const double* __restrict__ x, double* __restrict__ y, int64_t n, const double* __restrict__ s, int inv,
                             const int* __restrict__ act) {
  if (act && *act == 0) return;
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const double v = s[0];
  y[r] = inv ? x[r] / v : x[r] * v;
}

__global__ void k_axpby(double* __restrict__ x, const double* __restrict__ d, int64_t n, double a) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < n) x[r] = fma(a, d[r], x[r]);
}
__global__ void k_axpby_to(double* __restrict__ y, const double* __restrict__ x, const double* __restrict__ d, int64_t n, double a) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < n) y[r] = fma(a, d[r], x[r]);   // k_axpby's value, into another vector
}
__global__ void k_copy_f64(const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < n) y[r] = x[r];
}
__global__ void k_fill_f64(double* __restrict__ x, int64_t n, double v) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < n) x[r] = v;
}

// Hessenberg column j: h[0..j] (projections) and hn = ||w||.  State: H[(m+1) x m] column-major, cs, sn, g.
// Applies the previous rotations, forms a new one (rho = hypot(a, b); rho == 0 -> (1, 0)), updates g and
// writes out[0] = |g[j+1]| / beta0, out[1] = hn, out[2] = 1 if every value involved is finite.  With an activity
// word, the step ends the cycle (clears it) when out[0] <= tol, hn == 0 or a value is not finite.
__global__ void k_givens_step(double* __restrict__ H, int m, int j, const double* __restrict__ h, const double* __restrict__ hn,
                              double* __restrict__ cs, double* __restrict__ sn, double* __restrict__ g, double beta0,
                              double* __restrict__ out, double tol, int* __restrict__ act) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (act && *act == 0) return;
  double* col = H + (int64_t)j * (m + 1);
  bool fin = true;
  for (int i = 0; i <= j; i++) { col[i] = h[i]; fin = fin && isfinite(h[i]); }
  const double hnv = hn[0];
  col[j + 1] = hnv;
  fin = fin && isfinite(hnv);
  for (int i = 0; i < j; i++) {
    const double a = col[i], b = col[i + 1];
    col[i] = cs[i] * a + sn[i] * b;
    col[i + 1] = -sn[i] * a + cs[i] * b;
  }
  const double a = col[j], b = col[j + 1];
  const double rho = hypot(a, b);
  if (rho == 0.0) { cs[j] = 1.0; sn[j] = 0.0; } else { cs[j] = a / rho; sn[j] = b / rho; }
  col[j] = rho; col[j + 1] = 0.0;
  const double g0 = g[j];
  g[j] = cs[j] * g0;
  g[j + 1] = -sn[j] * g0;
  out[0] = fabs(g[j + 1]) / beta0;
  out[1] = hnv;
  out[2] = fin ? 1.0 : 0.0;
  if (act && (!fin || hnv == 0.0 || out[0] <= tol)) *act = 0;
}

// The Arnoldi step's tail in one launch of 256 threads: h[i] += h2[i] (the two CGS passes), hn = sqrt of the sum
// of the w^2 chunk partials (k_norm_final's order), then k_givens_step's rotation update on thread 0; hn is also
// stored at hnout for the scaling of the next basis vector.  With hslot (coherent pinned host memory) the three
// step outputs also go there, followed by the step's tag with a system-scope release, so the host can wait for
// the step by reading the tag.
__device__ __forceinline__ void givens_tail(double* __restrict__ H, int m, int j, double* __restrict__ h,
                                            const double* __restrict__ h2, const double* __restrict__ sq, int nchunks,
                                            double* __restrict__ hnout, double* __restrict__ cs, double* __restrict__ sn,
                                            double* __restrict__ g, double beta0, double* __restrict__ out, double tol,
                                            int* __restrict__ act, double* __restrict__ hslot, unsigned long long tag,
                                            double* red) {
  const int t = threadIdx.x;
  if (t <= j) h[t] += h2[t];
  double s = 0.0;
  for (int c = t; c < nchunks; c += 256) s += sq[c];
  red[t] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) { if (t < st) red[t] += red[t + st]; __syncthreads(); }
  if (t != 0) return;
  const double hnv = sqrt(red[0]);
  hnout[0] = hnv;
  double* col = H + (int64_t)j * (m + 1);
  bool fin = true;
  for (int i = 0; i <= j; i++) { col[i] = h[i]; fin = fin && isfinite(h[i]); }
  col[j + 1] = hnv;
  fin = fin && isfinite(hnv);
  for (int i = 0; i < j; i++) {
    const double a = col[i], b = col[i + 1];
    col[i] = cs[i] * a + sn[i] * b;
    col[i + 1] = -sn[i] * a + cs[i] * b;
  }
  const double a = col[j], b = col[j + 1];
  const double rho = hypot(a, b);
  if (rho == 0.0) { cs[j] = 1.0; sn[j] = 0.0; } else { cs[j] = a / rho; sn[j] = b / rho; }
  col[j] = rho; col[j + 1] = 0.0;
  const double g0 = g[j];
  g[j] = cs[j] * g0;
  g[j + 1] = -sn[j] * g0;
  out[0] = fabs(g[j + 1]) / beta0;
  out[1] = hnv;
  out[2] = fin ? 1.0 : 0.0;
  if (act && (!fin || hnv == 0.0 || out[0] <= tol)) *act = 0;
  if (hslot) {
    hslot[0] = out[0]; hslot[1] = hnv; hslot[2] = out[2];
    __hip_atomic_store((unsigned long long*)(hslot + 3), tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__global__ __launch_bounds__(256) void k_givens_fused(double* __restrict__ H, int m, int j, double* __restrict__ h,
                                                      const double* __restrict__ h2, const double* __restrict__ sq, int nchunks,
                                                      double* __restrict__ hnout, double* __restrict__ cs, double* __restrict__ sn,
                                                      double* __restrict__ g, double beta0, double* __restrict__ out, double tol,
                                                      int* __restrict__ act, double* __restrict__ hslot, unsigned long long tag) {
  if (act && *act == 0) return;
  __shared__ double red[256];
  givens_tail(H, m, j, h, h2, sq, nchunks, hnout, cs, sn, g, beta0, out, tol, act, hslot, tag, red);
}

// The second CGS pass (k_cgs_update with the norm's partials) and the step's Givens update in one launch: every
// workgroup releases its writes and counts itself; the last one to finish acquires them, clears the counter for
// the next step and runs k_givens_fused's code.  The same bits as the two launches; nothing waits.
__global__ __launch_bounds__(256) void k_cgs_update_givens(const double* __restrict__ part, int nchunks, int ncols,
                                                           const double* __restrict__ V, int64_t ldv, double* __restrict__ hout,
                                                           double* __restrict__ w, int64_t n, double sign, double* __restrict__ sq,
                                                           unsigned int* __restrict__ counter, double* __restrict__ H, int m, int j,
                                                           double* __restrict__ h, double* __restrict__ hnout,
                                                           double* __restrict__ cs, double* __restrict__ sn, double* __restrict__ g,
                                                           double beta0, double* __restrict__ out, double tol, int* __restrict__ act,
                                                           double* __restrict__ hslot, unsigned long long tag) {
  if (act && *act == 0) return;
  cgs_update_body(part, nchunks, ncols, V, ldv, hout, w, n, sign, sq);
  __shared__ int s_last;
  __shared__ double red[256];
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  givens_tail(H, m, j, h, hout, sq, nchunks, hnout, cs, sn, g, beta0, out, tol, act, hslot, tag, red);
}

// Solve the k x k upper-triangular H y = g (back substitution, one thread).
__global__ void k_hess_solve(const double* __restrict__ H, int m, int k, const double* __restrict__ g, double* __restrict__ y) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = k - 1; i >= 0; i--) {
    double s = g[i];
    for (int l = i + 1; l < k; l++) s -= H[(int64_t)l * (m + 1) + i] * y[l];
    y[i] = s / H[(int64_t)i * (m + 1) + i];
  }
}

// Start of a cycle: g = beta e_1, H = 0, and the activity word (if any) set.
__global__ void k_gmres_init(double* __restrict__ g, double* __restrict__ H, int m, double beta, int* __restrict__ act) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int64_t hsz = (int64_t)(m + 1) * m;
  if (i == 0 && act) *act = 1;
  if (i <= m) g[i] = (i == 0) ? beta : 0.0;
  for (int64_t q = i; q < hsz; q += (int64_t)gridDim.x * 256) H[q] = 0.0;
}

// ------------------------------------------------------------------ host wrappers
Krylov::~Krylov() {
  for (hipEvent_t& e : ev) if (e) (void)hipEventDestroy(e);
  if (hstep) (void)hipHostFree(hstep);
}

void Krylov::ensure(int64_t n_, int m_) {
  // step slots in coherent pinned host memory (m <= 255 steps of 4 words); without them the steps' outputs
  // are copied back and waited for with events
  if (!hstep && hipHostMalloc((void**)&hstep, 256 * 4 * sizeof(double), hipHostMallocCoherent) != hipSuccess) {
    (void)hipGetLastError();
    hstep = nullptr;
  }
  if (n == n_ && m >= m_ && V.p) return;
  n = n_; m = m_;
  V.alloc(*W->tracker, (size_t)n * (m + 1), "gmres_V");
  Z.alloc(*W->tracker, (size_t)n * m, "gmres_Z");
  w.alloc(*W->tracker, (size_t)n, "gmres_w");
  sR.alloc(*W->tracker, (size_t)n, "gmres_sR");
  rcur.alloc(*W->tracker, (size_t)n, "gmres_rcur");
  H.alloc(*W->tracker, (size_t)(m + 1) * m + 8, "gmres_H");
  small.alloc(*W->tracker, (size_t)8 * (m + 2) + 64, "gmres_small");
  part.alloc(*W->tracker, (size_t)cdiv(n, KCH) * (m + 2) + 64, "gmres_part");
  part2.alloc(*W->tracker, (size_t)cdiv(n, KCH) * (m + 1) + 64, "gmres_part2");
  // the fused CGS / Givens launch's finish counter: zero here, and put back to zero by each step's last workgroup
  if (!gcount.p) {
    gcount.alloc(*W->tracker, 1, "gmres_finish_count");
    HIP_CK(hipMemsetAsync(gcount.p, 0, sizeof(unsigned), W->stream));
  }
}

void Krylov::dots(const double* Vb, int ncols, const double* x, double* out, bool accumulate, const int* act) {
  const unsigned nch = cdiv(n, KCH);
  hipLaunchKernelGGL(k_gemvt_partial, dim3(nch), dim3(256), 0, W->stream, Vb, n, ncols, x, n, part.p, act);
  LAUNCH_CK();
  hipLaunchKernelGGL(k_gemvt_final, dim3(cdiv(ncols, 64)), dim3(64), 0, W->stream, part.p, (int)nch, ncols, out, accumulate ? 1 : 0, act);
  LAUNCH_CK();
}
void Krylov::update(const double* Vb, int ncols, const double* h, double* x, double sign, const int* act) {
  hipLaunchKernelGGL(k_gemv_update, dim3(cdiv(n, 256)), dim3(256), 0, W->stream, Vb, n, ncols, h, x, n, sign, act);
  LAUNCH_CK();
}
void Krylov::norm(const double* x, double* out, const int* act) {
  const unsigned nch = cdiv(n, KCH);
  hipLaunchKernelGGL(k_sumsq_partial, dim3(nch), dim3(256), 0, W->stream, x, n, part.p, act);
  LAUNCH_CK();
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, W->stream, part.p, (int)nch, out, act);
  LAUNCH_CK();
}
void Krylov::start(const double* r, double* d, double* out) {
  const unsigned nch = cdiv(n, KCH);
  hipLaunchKernelGGL(k_gmres_start, dim3(nch), dim3(256), 0, W->stream, r, rcur.p, d, n, part.p);
  LAUNCH_CK();
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, W->stream, part.p, (int)nch, out, (const int*)nullptr);
  LAUNCH_CK();
}
void Krylov::scale_copy(const double* x, double* y, const double* s, bool inv, const int* act) {
  hipLaunchKernelGGL(k_scale_copy, dim3(cdiv(n, 256)), dim3(256), 0, W->stream, x, y, n, s, inv ? 1 : 0, act);
  LAUNCH_CK();
}
void Krylov::givens(int j, const double* h, const double* hn, double beta0, double* out, double tol, int* act) {
  double* cs = small.p; double* sn = cs + (m + 1); double* g = sn + (m + 1);
  hipLaunchKernelGGL(k_givens_step, dim3(1), dim3(64), 0, W->stream, H.p, m, j, h, hn, cs, sn, g, beta0, out, tol, act);
  LAUNCH_CK();
}
void Krylov::init_cycle(double beta, int* act) {
  double* g = small.p + 2 * (m + 1);
  hipLaunchKernelGGL(k_gmres_init, dim3(cdiv((size_t)(m + 1) * m + 1, 256)), dim3(256), 0, W->stream, g, H.p, m, beta, act);
  LAUNCH_CK();
}
// CGS2 of w against V[:, 0..j] and the step's Givens update: 3 launches (5 with MPSOLVE_CGS_FUSE=0) with the sums
// of the unfused kernels.
void Krylov::arnoldi_tail(int j, double beta0, double tol, int* act, double* hslot, unsigned long long tag,
                          const double* prod_part, int64_t prod_nch) {
  const unsigned nch = cdiv(n, KCH);
  const int ncols = j + 1;
  double* sq = part.p + (size_t)nch * (m + 1);
  double* cs = small.p; double* sn = cs + (m + 1); double* g = sn + (m + 1);
  if (prod_part)
    hipLaunchKernelGGL(k_reduce_dots, dim3(nch), dim3(256), 0, W->stream, prod_part, prod_nch, w.p, V.p, n, ncols, n, part.p,
                       (const int*)act);
  else
    hipLaunchKernelGGL(k_gemvt_partial, dim3(nch), dim3(256), 0, W->stream, V.p, n, ncols, w.p, n, part.p, (const int*)act);
  LAUNCH_CK();
  // MPSOLVE_CGS_FUSE=0: the first pass's update and the second pass's dot products as two launches (same sums)
  const bool cgs_fuse = !(std::getenv("MPSOLVE_CGS_FUSE") && std::getenv("MPSOLVE_CGS_FUSE")[0] == '0');
  if (cgs_fuse) {
    hipLaunchKernelGGL(k_cgs_update_dots, dim3(nch), dim3(256), 0, W->stream, part.p, (int)nch, ncols, V.p, n, h(), w.p, n, -1.0,
                       part2.p, (const int*)act);
    LAUNCH_CK();
  } else {
    hipLaunchKernelGGL(k_cgs_update, dim3(nch), dim3(256), 0, W->stream, part.p, (int)nch, ncols, V.p, n, h(), w.p, n, -1.0,
                       (double*)nullptr, (const int*)act);
    LAUNCH_CK();
    hipLaunchKernelGGL(k_gemvt_partial, dim3(nch), dim3(256), 0, W->stream, V.p, n, ncols, w.p, n, part2.p, (const int*)act);
    LAUNCH_CK();
  }
  if (cgs_fuse) {
    hipLaunchKernelGGL(k_cgs_update_givens, dim3(nch), dim3(256), 0, W->stream, part2.p, (int)nch, ncols, V.p, n, h2(), w.p, n,
                       -1.0, sq, gcount.p, H.p, m, j, h(), scal(), cs, sn, g, beta0, scal() + 2, tol, act, hslot, tag);
    LAUNCH_CK();
  } else {
    hipLaunchKernelGGL(k_cgs_update, dim3(nch), dim3(256), 0, W->stream, part2.p, (int)nch, ncols, V.p, n, h2(), w.p, n, -1.0,
                       sq, (const int*)act);
    LAUNCH_CK();
    hipLaunchKernelGGL(k_givens_fused, dim3(1), dim3(256), 0, W->stream, H.p, m, j, h(), h2(), sq, (int)nch, scal(), cs, sn, g,
                       beta0, scal() + 2, tol, act, hslot, tag);
    LAUNCH_CK();
  }
}
void Krylov::lsq(int k, double* y) {
  const double* g = small.p + 2 * (m + 1);
  hipLaunchKernelGGL(k_hess_solve, dim3(1), dim3(64), 0, W->stream, H.p, m, k, g, y);
  LAUNCH_CK();
}

void axpy(double* x, const double* d, int64_t n, double a, Workspace& W) {
  hipLaunchKernelGGL(k_axpby, dim3(cdiv(n, 256)), dim3(256), 0, W.stream, x, d, n, a);
  LAUNCH_CK();
}
void axpy_to(double* y, const double* x, const double* d, int64_t n, double a, Workspace& W) {
  hipLaunchKernelGGL(k_axpby_to, dim3(cdiv(n, 256)), dim3(256), 0, W.stream, y, x, d, n, a);
  LAUNCH_CK();
}
void copy_f64(const double* x, double* y, int64_t n, Workspace& W) {
  hipLaunchKernelGGL(k_copy_f64, dim3(cdiv(n, 256)), dim3(256), 0, W.stream, x, y, n);
  LAUNCH_CK();
}
void fill_f64(double* x, int64_t n, double v, Workspace& W) {
  hipLaunchKernelGGL(k_fill_f64, dim3(cdiv(n, 256)), dim3(256), 0, W.stream, x, n, v);
  LAUNCH_CK();
}

// ------------------------------------------------------------------ FP64 -> FP32 copy with a power-of-two scale
// out = fl32(A) * fl32(s) for an n-row x ncols block.  A workgroup covers 512 rows x 4 columns: each thread
// issues its eight loads (rows r and r + 256 of four columns) before it stores.
__global__ __launch_bounds__(256) void k_scale_convert(const double* __restrict__ A, int64_t n, int64_t ncols, int64_t lda,
                                                       float* __restrict__ out, int64_t ldo, double s) {
  const int64_t r0 = (int64_t)blockIdx.x * 512 + threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.y * 4;
  const float sf = (float)s;
  double v[8];
#pragma unroll
  for (int u = 0; u < 8; u++) {
    const int64_t r = r0 + 256 * (u & 1), c = c0 + (u >> 1);
    v[u] = (r < n && c < ncols) ? A[c * lda + r] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < 8; u++) {
    const int64_t r = r0 + 256 * (u & 1), c = c0 + (u >> 1);
    if (r < n && c < ncols) out[c * ldo + r] = (float)v[u] * sf;   // round to FP32, then the exact power-of-two scale
  }
}
void scale_convert_cols(const double* A, int64_t n, int64_t ncols, int64_t lda, float* out, int64_t ldo, double s,
                        hipStream_t stream) {
  if (n <= 0 || ncols <= 0) return;
  hipLaunchKernelGGL(k_scale_convert, dim3(cdiv(n, 512), (unsigned)cdiv(ncols, 4)), dim3(256), 0, stream, A, n, ncols, lda,
                     out, ldo, s);
  LAUNCH_CK();
}
void scale_convert_f64_to_f32(const double* A, int64_t n, int64_t lda, float* out, int64_t ldo, double s, Workspace& W) {
  scale_convert_cols(A, n, n, lda, out, ldo, s, W.stream);
}

}  // namespace mpk
