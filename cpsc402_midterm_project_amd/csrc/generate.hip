// generate.hip — deterministic test-system generation on the device (setup, outside every timed region).
//
// Families (the case fields of the reference's CaseSpec / generate_problem, src/mp_core.cuh:14-103):
//   prescribed diag(sigma) rotated by ceil(log2 n) + 4 stages of random disjoint Givens pairs on the columns,
//              transposed, then as many stages on the columns again (src/mp_core.cuh:36-58)
//   randsvd    A = U diag(sigma) V^T, U and V the Q factors (rocSOLVER geqrf/orgqr) of Gaussian matrices
//   dyadic     A = 2^-s C with integer C_ij in [-m, m]; with rhs=exact_sign, x* = +-1 and b = 2^-s C x* exactly
//   rbf        exp(-|p_i - p_j|^2 / (2 ell^2)) + nugget*delta_ij, points uniform in [0,1]^dim
//   grcar      -1 on the subdiagonal, 1 on the diagonal and k superdiagonals
//   wilkinson  1 on the diagonal and last column, -1 below the diagonal
// sigma from --spectrum (geometric kappa^(-i/(n-1)), clustered, one_small, identity).
// RHS modes: random (Gaussian), random_scaled (column c scaled by 2^(4c-2 nrhs)), xtrue_uniform
// (x_true ~ U(-1,1), b = A x_true), exact_sign (dyadic), unit_column (b = A e_j).
//
// Random streams.  Everything the reference draws on the host (rotation stages, rbf points, row/column
// scales, random right-hand sides, unit columns) is drawn here from the same xoshiro256** stream with the same
// seeds and in the same order, and the dyadic entries use the same counter hash, so a case seed names the same
// stored system in both builds (the transcendental functions come from the host C library).  The randsvd
// family, which the reference does not have, uses a counter-based stream of its own: value k of stream s for
// seed q is splitmix-style mixing of (q, s, k) -> 53-bit uniform; Gaussians by Box-Muller.
#include "mp_internal.hpp"
#include "mp_kernels.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace mpk {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline double unif(uint64_t seed, uint64_t stream, uint64_t k) {
  const uint64_t h = mix64(mix64(seed * 0x100000001B3ull + stream) ^ (k * 0xD6E8FEB86659FD93ull));
  return (double)(h >> 11) * 0x1.0p-53;
}
__host__ __device__ inline double gauss(uint64_t seed, uint64_t stream, uint64_t k) {
  double u1 = unif(seed, stream, 2 * k), u2 = unif(seed, stream, 2 * k + 1);
  if (u1 <= 0.0) u1 = 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// ---- the reference's streams (src/mp_common.cuh:34-43): splitmix64 steps and xoshiro256**
__host__ __device__ inline uint64_t splitmix_step(uint64_t& x) {
  x += 0x9E3779B97F4A7C15ull;
  uint64_t z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// counter hash of three words: the second splitmix64 output from a mixed start
__host__ __device__ inline uint64_t hash3(uint64_t a, uint64_t b, uint64_t c) {
  uint64_t x = (a * 0x9E3779B97F4A7C15ull) ^ ((b + 0x7F4A7C15ull) * 0xBF58476D1CE4E5B9ull) ^ ((c + 0x1234567ull) * 0x94D049BB133111EBull);
  splitmix_step(x);
  return splitmix_step(x);
}
class HostStream {
 public:
  explicit HostStream(uint64_t seed) { uint64_t x = seed; for (auto& w : s_) w = splitmix_step(x); }
  uint64_t next() {
    const uint64_t out = rotl(s_[1] * 5, 7) * 9, t = s_[1] << 17;
    s_[2] ^= s_[0]; s_[3] ^= s_[1]; s_[1] ^= s_[2]; s_[0] ^= s_[3]; s_[2] ^= t; s_[3] = rotl(s_[3], 45);
    return out;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }          // [0, 1)
  double normal() {                                                        // Box-Muller, first branch
    double u1;
    do { u1 = uniform(); } while (u1 <= 0.0);
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  int64_t below(int64_t n) { return (int64_t)(next() % (uint64_t)n); }
 private:
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t s_[4];
};

// ---- prescribed family: stages of disjoint column rotations.  Pair p of a stage rotates columns (c1, c2) by
// (c, s): a' = fma(c, a, s b), b' = fma(c, b, -(s a)), each product rounded once (src/mp_kernels.cuh:39-44).
__global__ __launch_bounds__(256) void k_rotate_pairs(double* __restrict__ A, int64_t n, const int32_t* __restrict__ pairs,
                                                      const double* __restrict__ cs) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const int p = blockIdx.y;
  const int64_t c1 = pairs[2 * p], c2 = pairs[2 * p + 1];
  const double c = cs[2 * p], sn = cs[2 * p + 1];
  const double a = A[c1 * n + r], b = A[c2 * n + r];
  A[c1 * n + r] = fma(c, a, __dmul_rn(sn, b));
  A[c2 * n + r] = fma(c, b, -__dmul_rn(sn, a));
}
// in-place transpose: workgroup (bi, bj), bj >= bi, swaps the 32 x 32 tiles (bi, bj) and (bj, bi) through LDS
__global__ __launch_bounds__(256) void k_transpose_sq(double* __restrict__ A, int64_t n) {
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bj < bi) return;
  __shared__ double tu[32][33], tl[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t r0 = (int64_t)bi * 32, c0 = (int64_t)bj * 32;
  for (int k = ty; k < 32; k += 8) {
    if (r0 + tx < n && c0 + k < n) tu[k][tx] = A[(c0 + k) * n + r0 + tx];      // tile (bi, bj): tu[k][t] = A(r0+t, c0+k)
    if (c0 + tx < n && r0 + k < n) tl[k][tx] = A[(r0 + k) * n + c0 + tx];      // tile (bj, bi): tl[k][t] = A(c0+t, r0+k)
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    if (r0 + tx < n && c0 + k < n) A[(c0 + k) * n + r0 + tx] = tl[tx][k];      // A(r0+tx, c0+k) <- A(c0+k, r0+tx)
    if (bi != bj && c0 + tx < n && r0 + k < n) A[(r0 + k) * n + c0 + tx] = tu[tx][k];
  }
}
static void rotation_stages(double* A, int64_t n, uint64_t seed, int side, int stages, Workspace& W, DevTracker& tr) {
  HostStream rng(seed * 1000003ull + 17ull * (uint64_t)side + 0xC0FFEEull);
  const int64_t np = n / 2;
  if (np == 0) return;
  std::vector<int32_t> perm(n), pairs((size_t)2 * np * stages);
  std::vector<double> cs((size_t)2 * np * stages);
  for (int st = 0; st < stages; st++) {
    for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
    for (int64_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[rng.below(i + 1)]);
    int32_t* pp = pairs.data() + (size_t)2 * np * st; double* cp = cs.data() + (size_t)2 * np * st;
    for (int64_t p = 0; p < np; p++) {
      pp[2 * p] = perm[2 * p]; pp[2 * p + 1] = perm[2 * p + 1];
      const double th = rng.uniform() * 6.283185307179586;
      cp[2 * p] = std::cos(th); cp[2 * p + 1] = std::sin(th);
    }
  }
  DBuf<int32_t> dp; DBuf<double> dc;
  dp.alloc(tr, pairs.size(), "gen_pairs"); dc.alloc(tr, cs.size(), "gen_cs");
  HIP_CK(hipMemcpyAsync(dp.p, pairs.data(), pairs.size() * 4, hipMemcpyHostToDevice, W.stream));
  HIP_CK(hipMemcpyAsync(dc.p, cs.data(), cs.size() * 8, hipMemcpyHostToDevice, W.stream));
  for (int st = 0; st < stages; st++) {
    hipLaunchKernelGGL(k_rotate_pairs, dim3(cdiv(n, 256), (unsigned)np), dim3(256), 0, W.stream, A, n,
                       dp.p + (size_t)2 * np * st, dc.p + (size_t)2 * np * st);
    LAUNCH_CK();
  }
  HIP_CK(hipStreamSynchronize(W.stream));
}

__global__ void k_gauss_fill(double* A, int64_t count, uint64_t seed, uint64_t stream) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < count) A[i] = gauss(seed, stream, (uint64_t)i);
}
__global__ void k_unif_fill(double* A, int64_t count, uint64_t seed, uint64_t stream, double lo, double hi) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < count) A[i] = lo + (hi - lo) * unif(seed, stream, (uint64_t)i);
}
__global__ void k_scale_cols(double* A, int64_t n, const double* s) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
  if (r < n) A[c * n + r] *= s[c];
}
// dyadic entries and signs from the counter hash (src/mp_kernels.cuh:48-49)
__device__ inline int dyadic_int(uint64_t seed, int64_t i, int64_t j, int m) {
  return (int)(hash3(seed, (uint64_t)i, (uint64_t)j) % (uint64_t)(2 * m + 1)) - m;
}
__device__ inline int dyadic_sign(uint64_t seed, int64_t j) { return (hash3(seed ^ 0xA5A5A5A5ull, (uint64_t)j, 7ull) & 1ull) ? 1 : -1; }
__global__ void k_dyadic_A(double* A, int64_t n, uint64_t seed, int m, double scale) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
  if (r < n) A[c * n + r] = (double)dyadic_int(seed, r, c, m) * scale;
}
// b_i = 2^-s sum_j C_ij x*_j with exact 64-bit integer accumulation
__global__ void k_dyadic_b(double* b, double* x, int64_t n, uint64_t seed, int m, double scale, int nrhs, uint64_t rseed) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  for (int c = 0; c < nrhs; c++) {
    long long s = 0;
    for (int64_t j = 0; j < n; j++) s += (long long)dyadic_int(seed, r, j, m) * dyadic_sign(rseed + c, j);
    b[c * n + r] = (double)s * scale;
    x[c * n + r] = (double)dyadic_sign(rseed + c, r);
  }
}
__global__ void k_rbf(double* A, int64_t n, const double* pts, int dim, double ell, double nugget) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
  if (r >= n) return;
  double d2 = 0;
  for (int k = 0; k < dim; k++) { const double t = pts[r + k * n] - pts[c + k * n]; d2 = fma(t, t, d2); }
  double v = exp(-d2 / (2.0 * ell * ell));
  if (r == c) v += nugget;
  A[c * n + r] = v;
}
__global__ void k_band(double* A, int64_t n, int family, int k) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
  if (r >= n) return;
  double v = 0;
  if (family == 0) { if (r == c + 1) v = -1; else if (c >= r && c <= r + k) v = 1; }        // grcar
  else { if (r == c) v = 1; else if (r > c) v = -1; if (c == n - 1) v = 1; }                  // wilkinson
  A[c * n + r] = v;
}

__global__ void k_scale_inplace(double* x, int64_t n, double s) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] *= s;
}
void axpy_scale_inplace(double* x, int64_t n, double s, Workspace& W) {
  hipLaunchKernelGGL(k_scale_inplace, dim3(cdiv(n, 256)), dim3(256), 0, W.stream, x, n, s);
  LAUNCH_CK();
}

double spectrum_value(const std::string& sp, int64_t i, int64_t n, double kappa) {
  if (sp == "geometric") return (n == 1) ? 1.0 : std::pow(kappa, -(double)i / (double)(n - 1));
  if (sp == "clustered") { const int64_t k = std::max<int64_t>(2, n / 64); return (i >= n - k) ? 1.0 / kappa : 1.0; }
  if (sp == "one_small") return (i == n - 1) ? 1.0 / kappa : 1.0;
  if (sp == "identity") return 1.0;
  throw MPError("unknown spectrum " + sp, -1);
}

// A[i + j n] *= r[i] * c[j] (powers of two: exact)
__global__ void k_rowcol_scale(double* __restrict__ A, int64_t n, const double* __restrict__ r, const double* __restrict__ c) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  if (i < n) A[i + j * n] *= r[i] * c[j];
}

// Generates A (FP64, n x n), B (n x nrhs) and optionally X_exact into device buffers.  Returns has_exact.
bool generate(const GenSpec& g, void* rb_handle, Workspace& W, DevTracker& tr, double* A, double* B, double* Xe) {
  rocblas_handle rb = (rocblas_handle)rb_handle;
  const int64_t n = g.n; const int nrhs = g.nrhs; const size_t nn = (size_t)n * n;
  bool has_exact = false;
  HIP_CK(hipStreamSynchronize(W.stream));
  if (rocblas_set_stream(rb, W.stream) != rocblas_status_success) throw MPError("rocblas_set_stream failed", -2);
  auto rbck = [](rocblas_status s, const char* w) { if (s != rocblas_status_success) throw MPError(std::string("rocBLAS/rocSOLVER failure in ") + w, -2); };
  if (g.family == "prescribed") {
    std::vector<double> hs(n);
    for (int64_t i = 0; i < n; i++) hs[i] = spectrum_value(g.spectrum, i, n, g.kappa);
    DBuf<double> sig; sig.alloc(tr, n, "gen_sigma");
    HIP_CK(hipMemsetAsync(A, 0, nn * 8, W.stream));
    HIP_CK(hipMemcpyAsync(sig.p, hs.data(), n * 8, hipMemcpyHostToDevice, W.stream));
    HIP_CK(hipMemcpy2DAsync(A, (size_t)(n + 1) * 8, sig.p, 8, 8, (size_t)n, hipMemcpyDeviceToDevice, W.stream));
    const int stages = (int)std::ceil(std::log2((double)n)) + 4;
    rotation_stages(A, n, g.seed, 1, stages, W, tr);
    const unsigned nt = (unsigned)cdiv(n, 32);
    hipLaunchKernelGGL(k_transpose_sq, dim3(nt, nt), dim3(256), 0, W.stream, A, n);
    LAUNCH_CK();
    rotation_stages(A, n, g.seed, 2, stages, W, tr);
  } else if (g.family == "randsvd") {
    // U and V depend on (n, seed) only: with keep_q they stay in the context, so systems that differ only in
    // kappa (or the right-hand side) reuse them; the stored A is the same bits either way
    DBuf<double> Ut, Vt, T, tau, sig;
    const bool hit = g.keep_q && W.gen_valid && W.gen_n == n && W.gen_seed == g.seed;
    double *U, *V;
    if (g.keep_q) {
      if (!hit) { W.gen_valid = false; W.genU.alloc(tr, nn, "gen_U_kept"); W.genV.alloc(tr, nn, "gen_V_kept"); }
      U = W.genU.p; V = W.genV.p;
    } else {
      Ut.alloc(tr, nn, "gen_U"); Vt.alloc(tr, nn, "gen_V"); U = Ut.p; V = Vt.p;
    }
    if (!hit) {
      tau.alloc(tr, n, "gen_tau");
      hipLaunchKernelGGL(k_gauss_fill, dim3(cdiv(nn, 256)), dim3(256), 0, W.stream, U, (int64_t)nn, g.seed, (uint64_t)1);
      hipLaunchKernelGGL(k_gauss_fill, dim3(cdiv(nn, 256)), dim3(256), 0, W.stream, V, (int64_t)nn, g.seed, (uint64_t)2);
      rbck(rocsolver_dgeqrf(rb, (rocblas_int)n, (rocblas_int)n, U, (rocblas_int)n, tau.p), "dgeqrf");
      rbck(rocsolver_dorgqr(rb, (rocblas_int)n, (rocblas_int)n, (rocblas_int)n, U, (rocblas_int)n, tau.p), "dorgqr");
      rbck(rocsolver_dgeqrf(rb, (rocblas_int)n, (rocblas_int)n, V, (rocblas_int)n, tau.p), "dgeqrf");
      rbck(rocsolver_dorgqr(rb, (rocblas_int)n, (rocblas_int)n, (rocblas_int)n, V, (rocblas_int)n, tau.p), "dorgqr");
      if (g.keep_q) { W.gen_n = n; W.gen_seed = g.seed; W.gen_valid = true; }
    }
    sig.alloc(tr, n, "gen_sigma");
    std::vector<double> hs(n);
    for (int64_t i = 0; i < n; i++) hs[i] = spectrum_value(g.spectrum, i, n, g.kappa);
    HIP_CK(hipMemcpyAsync(sig.p, hs.data(), n * 8, hipMemcpyHostToDevice, W.stream));
    double* Us = U;
    if (g.keep_q) {   // the kept U stays unscaled: U diag(sigma) goes to a temporary
      T.alloc(tr, nn, "gen_U_sigma");
      HIP_CK(hipMemcpyAsync(T.p, U, nn * 8, hipMemcpyDeviceToDevice, W.stream));
      Us = T.p;
    }
    hipLaunchKernelGGL(k_scale_cols, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, Us, n, sig.p);
    const double one = 1.0, zero = 0.0;
    rbck(rocblas_dgemm(rb, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)n, (rocblas_int)n, (rocblas_int)n,
                       &one, Us, (rocblas_int)n, V, (rocblas_int)n, &zero, A, (rocblas_int)n), "dgemm");
    HIP_CK(hipStreamSynchronize(W.stream));
  } else if (g.family == "dyadic") {
    hipLaunchKernelGGL(k_dyadic_A, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, A, n, g.seed, g.intmax, std::ldexp(1.0, -g.dyadic_exp));
  } else if (g.family == "rbf") {
    DBuf<double> pts; pts.alloc(tr, (size_t)n * g.dim, "gen_pts");
    HostStream rng(g.seed * 7919ull + 0xB1ull);
    std::vector<double> hp((size_t)n * g.dim);
    for (auto& v : hp) v = rng.uniform();
    HIP_CK(hipMemcpyAsync(pts.p, hp.data(), hp.size() * 8, hipMemcpyHostToDevice, W.stream));
    hipLaunchKernelGGL(k_rbf, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, A, n, pts.p, g.dim, g.ell, g.nugget);
    HIP_CK(hipStreamSynchronize(W.stream));
  } else if (g.family == "grcar" || g.family == "wilkinson") {
    hipLaunchKernelGGL(k_band, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, A, n, g.family == "grcar" ? 0 : 1, g.grcar_k);
  } else {
    throw MPError("unknown or unsupported family " + g.family + " (supported: prescribed, randsvd, dyadic, rbf, grcar, wilkinson)", -4);
  }
  LAUNCH_CK();
  if (g.rowcol_span > 0) {
    // the reference's --rowcol-span (src/mp_core.cuh:71-72): A <- D_r A D_c with row and column factors
    // 2^(u - span/2), u uniform in [0, span]; applied before the right-hand side is formed.
    if (g.rhs == "exact_sign") throw MPError("exact_sign with scaling not supported", -1);
    HostStream rng(g.seed * 31ull + 5ull);
    std::vector<double> r(n), c(n);
    for (int64_t i = 0; i < n; i++) r[i] = std::ldexp(1.0, (int)rng.below(g.rowcol_span + 1) - g.rowcol_span / 2);
    for (int64_t i = 0; i < n; i++) c[i] = std::ldexp(1.0, (int)rng.below(g.rowcol_span + 1) - g.rowcol_span / 2);
    DBuf<double> dr, dc; dr.alloc(tr, n, "gen_dr"); dc.alloc(tr, n, "gen_dc");
    HIP_CK(hipMemcpyAsync(dr.p, r.data(), n * 8, hipMemcpyHostToDevice, W.stream));
    HIP_CK(hipMemcpyAsync(dc.p, c.data(), n * 8, hipMemcpyHostToDevice, W.stream));
    hipLaunchKernelGGL(k_rowcol_scale, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, A, n, dr.p, dc.p);
    LAUNCH_CK();
    HIP_CK(hipStreamSynchronize(W.stream));
  }
  const size_t nb = (size_t)n * nrhs;
  if (g.rhs == "random" || g.rhs == "random_scaled") {
    // Gaussian B in column-major order from one stream; random_scaled scales column c by 2^(4c - 2 nrhs)
    HostStream rng(g.seed * 104729ull + 0xBEEFull + (uint64_t)nrhs);
    std::vector<double> hb(nb);
    for (int c = 0; c < nrhs; c++)
      for (int64_t i = 0; i < n; i++) {
        const double v = rng.normal();
        hb[i + (size_t)c * n] = (g.rhs == "random_scaled") ? std::ldexp(v, 4 * c - 2 * nrhs) : v;
      }
    HIP_CK(hipMemcpyAsync(B, hb.data(), nb * 8, hipMemcpyHostToDevice, W.stream));
    HIP_CK(hipStreamSynchronize(W.stream));
  } else if (g.rhs == "xtrue_uniform") {
    hipLaunchKernelGGL(k_unif_fill, dim3(cdiv(nb, 256)), dim3(256), 0, W.stream, Xe, (int64_t)nb, g.seed, (uint64_t)4, -1.0, 1.0);
    const double one = 1.0, zero = 0.0;
    rbck(rocblas_dgemm(rb, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)nrhs, (rocblas_int)n,
                       &one, A, (rocblas_int)n, Xe, (rocblas_int)n, &zero, B, (rocblas_int)n), "dgemm");
    has_exact = true;
  } else if (g.rhs == "exact_sign") {
    if (g.family != "dyadic") throw MPError("exact_sign rhs requires the dyadic family", -1);
    if ((double)n * g.intmax > 9007199254740992.0) throw MPError("dyadic row-sum bound violates exact representability", -1);
    hipLaunchKernelGGL(k_dyadic_b, dim3(cdiv(n, 256)), dim3(256), 0, W.stream, B, Xe, n, g.seed, g.intmax, std::ldexp(1.0, -g.dyadic_exp), nrhs, g.seed * 3 + 11);
    has_exact = true;
  } else if (g.rhs == "unit_column") {
    HIP_CK(hipMemsetAsync(Xe, 0, nb * 8, W.stream));
    HostStream rng(g.seed * 17ull + 3ull);
    for (int c = 0; c < nrhs; c++) {
      const int64_t j = rng.below(n);
      HIP_CK(hipMemcpyAsync(B + (size_t)c * n, A + (size_t)j * n, n * 8, hipMemcpyDeviceToDevice, W.stream));
      const double one = 1.0;
      HIP_CK(hipMemcpyAsync(Xe + (size_t)c * n + j, &one, 8, hipMemcpyHostToDevice, W.stream));
      HIP_CK(hipStreamSynchronize(W.stream));
    }
    has_exact = true;
  } else {
    throw MPError("unknown rhs mode " + g.rhs, -1);
  }
  LAUNCH_CK();
  if (g.global_exp != 0) {
    // exact power-of-two range transformation of the stored system (the reference's --global-exp)
    const double s = std::ldexp(1.0, g.global_exp);
    axpy_scale_inplace(A, (int64_t)nn, s, W);
    axpy_scale_inplace(B, (int64_t)nb, s, W);
  }
  HIP_CK(hipStreamSynchronize(W.stream));
  return has_exact;
}

}  // namespace mpk
