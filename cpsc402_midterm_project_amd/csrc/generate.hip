// generate.hip — deterministic test-system generation on the device (setup, outside every timed region).
//
// Families (the case fields of the reference's CaseSpec / generate_problem, src/mp_core.cuh:14-103):
//   randsvd    A = U diag(sigma) V^T, U and V the Q factors (rocSOLVER geqrf/orgqr) of Gaussian matrices,
//              sigma from --spectrum (geometric kappa^(-i/(n-1)), clustered, one_small, identity)
//   dyadic     A = 2^-s C with integer C_ij in [-m, m]; with rhs=exact_sign, x* = +-1 and b = 2^-s C x* exactly
//   rbf        exp(-|p_i - p_j|^2 / (2 ell^2)) + nugget*delta_ij, points uniform in [0,1]^dim
//   grcar      -1 on the subdiagonal, 1 on the diagonal and k superdiagonals
//   wilkinson  1 on the diagonal and last column, -1 below the diagonal
// RHS modes: random (Gaussian), random_scaled (column c scaled by 2^(4c-2 nrhs)), xtrue_uniform
// (x_true ~ U(-1,1), b = A x_true), exact_sign (dyadic), unit_column (b = A e_j).
// Random numbers are a counter-based stream: value k of stream s for seed q is splitmix-style mixing of
// (q, s, k) -> 53-bit uniform; Gaussians by Box-Muller.  Streams do not depend on the launch shape.
#include "mp_internal.hpp"
#include "mp_kernels.hpp"

#include <random>
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
__device__ inline int dyadic_int(uint64_t seed, int64_t i, int64_t j, int m) {
  return (int)(mix64(mix64(seed ^ 0xA24BAED4963EE407ull) + (uint64_t)i * 0x9FB21C651E98DF25ull + (uint64_t)j) % (uint64_t)(2 * m + 1)) - m;
}
__device__ inline int dyadic_sign(uint64_t seed, int64_t j) { return (mix64(seed * 31 + (uint64_t)j) & 1ull) ? 1 : -1; }
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
  if (g.family == "randsvd") {
    DBuf<double> U, V, tau, sig;
    U.alloc(tr, nn, "gen_U"); V.alloc(tr, nn, "gen_V"); tau.alloc(tr, n, "gen_tau"); sig.alloc(tr, n, "gen_sigma");
    hipLaunchKernelGGL(k_gauss_fill, dim3(cdiv(nn, 256)), dim3(256), 0, W.stream, U.p, (int64_t)nn, g.seed, (uint64_t)1);
    hipLaunchKernelGGL(k_gauss_fill, dim3(cdiv(nn, 256)), dim3(256), 0, W.stream, V.p, (int64_t)nn, g.seed, (uint64_t)2);
    rbck(rocsolver_dgeqrf(rb, (rocblas_int)n, (rocblas_int)n, U.p, (rocblas_int)n, tau.p), "dgeqrf");
    rbck(rocsolver_dorgqr(rb, (rocblas_int)n, (rocblas_int)n, (rocblas_int)n, U.p, (rocblas_int)n, tau.p), "dorgqr");
    rbck(rocsolver_dgeqrf(rb, (rocblas_int)n, (rocblas_int)n, V.p, (rocblas_int)n, tau.p), "dgeqrf");
    rbck(rocsolver_dorgqr(rb, (rocblas_int)n, (rocblas_int)n, (rocblas_int)n, V.p, (rocblas_int)n, tau.p), "dorgqr");
    std::vector<double> hs(n);
    for (int64_t i = 0; i < n; i++) hs[i] = spectrum_value(g.spectrum, i, n, g.kappa);
    HIP_CK(hipMemcpyAsync(sig.p, hs.data(), n * 8, hipMemcpyHostToDevice, W.stream));
    hipLaunchKernelGGL(k_scale_cols, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, U.p, n, sig.p);
    const double one = 1.0, zero = 0.0;
    rbck(rocblas_dgemm(rb, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)n, (rocblas_int)n, (rocblas_int)n,
                       &one, U.p, (rocblas_int)n, V.p, (rocblas_int)n, &zero, A, (rocblas_int)n), "dgemm");
    HIP_CK(hipStreamSynchronize(W.stream));
  } else if (g.family == "dyadic") {
    hipLaunchKernelGGL(k_dyadic_A, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, A, n, g.seed, g.intmax, std::ldexp(1.0, -g.dyadic_exp));
  } else if (g.family == "rbf") {
    DBuf<double> pts; pts.alloc(tr, (size_t)n * g.dim, "gen_pts");
    hipLaunchKernelGGL(k_unif_fill, dim3(cdiv((size_t)n * g.dim, 256)), dim3(256), 0, W.stream, pts.p, (int64_t)n * g.dim, g.seed, (uint64_t)7, 0.0, 1.0);
    hipLaunchKernelGGL(k_rbf, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, A, n, pts.p, g.dim, g.ell, g.nugget);
    HIP_CK(hipStreamSynchronize(W.stream));
  } else if (g.family == "grcar" || g.family == "wilkinson") {
    hipLaunchKernelGGL(k_band, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, A, n, g.family == "grcar" ? 0 : 1, g.grcar_k);
  } else {
    throw MPError("unknown or unsupported family " + g.family + " (supported: randsvd, dyadic, rbf, grcar, wilkinson)", -4);
  }
  LAUNCH_CK();
  if (g.rowcol_span > 0) {
    // the reference's --rowcol-span (src/mp_core.cuh:71-72): A <- D_r A D_c with row and column factors
    // 2^(u - span/2), u uniform in [0, span]; applied before the right-hand side is formed.  The stream is
    // this library's own (seeded from the case seed).
    if (g.rhs == "exact_sign") throw MPError("exact_sign with scaling not supported", -1);
    std::mt19937_64 rng(g.seed * 31ull + 5ull);
    std::uniform_int_distribution<int> u(0, g.rowcol_span);
    std::vector<double> r(n), c(n);
    for (int64_t i = 0; i < n; i++) r[i] = std::ldexp(1.0, u(rng) - g.rowcol_span / 2);
    for (int64_t i = 0; i < n; i++) c[i] = std::ldexp(1.0, u(rng) - g.rowcol_span / 2);
    DBuf<double> dr, dc; dr.alloc(tr, n, "gen_dr"); dc.alloc(tr, n, "gen_dc");
    HIP_CK(hipMemcpyAsync(dr.p, r.data(), n * 8, hipMemcpyHostToDevice, W.stream));
    HIP_CK(hipMemcpyAsync(dc.p, c.data(), n * 8, hipMemcpyHostToDevice, W.stream));
    hipLaunchKernelGGL(k_rowcol_scale, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, A, n, dr.p, dc.p);
    LAUNCH_CK();
    HIP_CK(hipStreamSynchronize(W.stream));
  }
  const size_t nb = (size_t)n * nrhs;
  if (g.rhs == "random" || g.rhs == "random_scaled") {
    hipLaunchKernelGGL(k_gauss_fill, dim3(cdiv(nb, 256)), dim3(256), 0, W.stream, B, (int64_t)nb, g.seed, (uint64_t)3 + (uint64_t)nrhs);
    if (g.rhs == "random_scaled") {
      std::vector<double> s(nrhs);
      for (int c = 0; c < nrhs; c++) s[c] = std::ldexp(1.0, 4 * c - 2 * nrhs);
      DBuf<double> ds; ds.alloc(tr, nrhs, "gen_cs");
      HIP_CK(hipMemcpyAsync(ds.p, s.data(), nrhs * 8, hipMemcpyHostToDevice, W.stream));
      hipLaunchKernelGGL(k_scale_cols, dim3(cdiv(n, 256), (unsigned)nrhs), dim3(256), 0, W.stream, B, n, ds.p);
      HIP_CK(hipStreamSynchronize(W.stream));
    }
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
    for (int c = 0; c < nrhs; c++) {
      const int64_t j = (int64_t)(unif(g.seed, 5, (uint64_t)c) * (double)n) % n;
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
