// engine.hip — problems, factorizations, refinement engine and the extern "C" ABI of libmpsolve.
//
// Host-side structure follows the reference's L2 engine (src/mp_core.cuh): Problem / Factors / MethodSpec /
// Engine with refine_stage, gmres_solve and run, and the method registry make_method (src/mpsolve.cu:17-72).
// The numerical kernels are this library's HIP kernels; the vendor comparators (cs32 / cs64 factor kinds,
// i.e. the plain FP32 / FP64 LU) are rocSOLVER getrf/getrs.
#include <thread>
#include "mp_internal.hpp"
#include "mp_kernels.hpp"
#include "mp_krylov.hpp"
#include "../../include/mpsolve.h"
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <functional>
#include <memory>
#include <set>

using namespace mpk;

__global__ void k_set_x32(float* x32, double* x64, const double* d, int64_t m) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < m) { const float v = (float)d[i]; x32[i] = v; x64[i] = (double)v; }
}
// FP32 residual control (k_resid_f32, src/mp_kernels.cuh:80-81): R = B - A X with every operation in FP32,
// the row sum a sequential fmaf chain over the columns; B holds FP32 values (FP32-stored systems)
__global__ void k_resid_f32_rows(const float* __restrict__ A, int64_t n, const float* __restrict__ X,
                                 const double* __restrict__ B, float* __restrict__ R, int nrhs) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < nrhs; c++) {
    const float* xp = X + (int64_t)c * n;
    float s = 0.0f;
    for (int64_t j = 0; j < n; j++) s = fmaf(A[i + j * n], xp[j], s);
    R[i + (int64_t)c * n] = (float)B[i + (int64_t)c * n] - s;
  }
}

__global__ void k_add_x32(float* x32, double* x64, const double* d, int64_t m) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < m) { const float v = (float)((double)x32[i] + d[i]); x32[i] = v; x64[i] = (double)v; }
}
__global__ void k_add_vec(double* h, const double* h2, int cnt, const int* __restrict__ act) {
  if (act && *act == 0) return;
  const int i = threadIdx.x + blockIdx.x * 64;
  if (i < cnt) h[i] += h2[i];
}

// ------------------------------------------------------------------ error plumbing
static thread_local std::string g_last_error;
template <class F> static int guard(F&& f) {
  try { f(); return MP_OK; }
  catch (MPError& e) { g_last_error = e.what(); return e.code; }
  catch (std::exception& e) { g_last_error = e.what(); return MP_ERR_INTERNAL; }
  catch (...) { g_last_error = "unknown error"; return MP_ERR_INTERNAL; }
}
#define RB_CK(x) do { rocblas_status s_ = (x); if (s_ != rocblas_status_success) { std::ostringstream o_; \
  o_ << "rocBLAS/rocSOLVER status " << (int)s_ << " at " << __FILE__ << ":" << __LINE__ << " in " << #x; throw MPError(o_.str(), -2); } } while (0)

// ------------------------------------------------------------------ minimal flat JSON object reader
struct Params {
  std::map<std::string, std::string> kv;
  void parse(const char* js) {
    if (!js) return;
    const std::string s(js);
    size_t i = 0;
    auto skip = [&]() { while (i < s.size() && (isspace((unsigned char)s[i]) || s[i] == ',' || s[i] == '{' || s[i] == '}')) i++; };
    while (true) {
      skip();
      if (i >= s.size() || s[i] != '"') break;
      size_t e = s.find('"', i + 1);
      if (e == std::string::npos) break;
      std::string key = s.substr(i + 1, e - i - 1);
      i = s.find(':', e);
      if (i == std::string::npos) break;
      i++;
      while (i < s.size() && isspace((unsigned char)s[i])) i++;
      std::string val;
      if (i < s.size() && s[i] == '"') {
        size_t e2 = s.find('"', i + 1);
        if (e2 == std::string::npos) break;
        val = s.substr(i + 1, e2 - i - 1); i = e2 + 1;
      } else {
        size_t e2 = i;
        while (e2 < s.size() && s[e2] != ',' && s[e2] != '}') e2++;
        val = s.substr(i, e2 - i);
        while (!val.empty() && isspace((unsigned char)val.back())) val.pop_back();
        i = e2;
      }
      kv[key] = val;
    }
  }
  bool has(const std::string& k) const { return kv.count(k) > 0; }
  std::string get(const std::string& k, const std::string& d = "") const { auto it = kv.find(k); return it == kv.end() ? d : it->second; }
  double getd(const std::string& k, double d) const { auto it = kv.find(k); return it == kv.end() ? d : std::stod(it->second); }
  int geti(const std::string& k, int d) const { auto it = kv.find(k); return it == kv.end() ? d : (int)std::stol(it->second); }
  bool getb(const std::string& k, bool d) const { auto it = kv.find(k); if (it == kv.end()) return d; return it->second == "true" || it->second == "1"; }
};

// ------------------------------------------------------------------ context / problem
struct Factors;

struct mp_ctx {
  int device = 0;
  DevTracker tracker;
  Workspace W;
  rocblas_handle rb = nullptr;
  rocblas_handle rb_panel = nullptr;    // bound to W.pstream: FP64 LU look-ahead panels
  double* hpin = nullptr;
  std::string last_record, last_hist;
  std::unique_ptr<Krylov> kry;
  std::set<mp_problem*> problems;       // live problems: their device memory is released by mp_destroy
  std::map<uint64_t, mp_problem*> factor_ids;   // factor cache: matrix id -> problem holding its factors
  double* resid_scratch = nullptr;      // device address of this GPU's allocation-free residual scratch
  // The growth factor of the last own factorization, measured on the panel stream while the work after it runs:
  // the factors it belongs to and max|A| of their copy.  Pinned double 3080 receives max|LU|; ev_growth marks
  // the end of the measurement.  growth_resolve waits for it.
  Factors* growth_target = nullptr; double growth_den = 0;
  hipEvent_t ev_growth = nullptr;
};

struct mp_problem {
  mp_ctx* ctx = nullptr;
  int64_t n = 0; int nrhs = 1; bool fp32 = false;
  DBuf<double> A64; DBuf<float> A32; DBuf<double> B; DBuf<double> Xexact, Xref;
  bool has_exact = false, has_ref = false;
  double anorm_inf = 0, amax = 0;
  std::vector<double> bnorm_inf, bnorm_2;
  std::string a_sha, b_sha, case_json;
  std::vector<std::string> tile_sha;
  std::vector<double> hXexact, hXref;
  const void* a_ptr() const { return fp32 ? (const void*)A32.p : (const void*)A64.p; }
  // factors kept for solves with "reuse" (the reference keeps its Engine, and so its factors, across the trials
  // of a case, src/mpsolve.cu:173 and src/mp_core.cuh:224-227); a factor cache of repeated matrices uses the same
  mutable std::shared_ptr<Factors> fcache;
  void release_device() { A64.release(); A32.release(); B.release(); Xexact.release(); Xref.release(); fcache.reset(); }
};

// ------------------------------------------------------------------ methods
struct MethodSpec {
  std::string id, kind = "ir";          // direct | ir | unsupported
  std::string factor = "cs32";          // het16 | cus32 | mx16 | mxbf16 | mx8 | cs32 | cs64
  int working = 64, resid = 64;         // resid: 64, 32, or 128 (double-double outer residual)
  bool apply64 = false, gmres = false, gmres_left = false;
  int scale_mode = 0, nb = 256;
  int gmres_m = 40; double gmres_tol = 1e-4; int gmres_budget = 480, gmres_cycles = 4;
  int max_outer = 30; std::string policy = "none"; int budget = 0; double cost_ratio_next = 3.0;
  std::string unsupported_reason;
};

static MethodSpec make_method(const std::string& id, const Params& a) {
  MethodSpec M; M.id = id;
  M.nb = a.geti("nb", 256); M.gmres_m = a.geti("gmres_m", 40); M.gmres_tol = a.getd("gmres_tol", 1e-4);
  M.gmres_budget = a.geti("gmres_budget", 480); M.gmres_cycles = a.geti("gmres_cycles", 4);
  M.max_outer = a.geti("max_outer", 30); M.cost_ratio_next = a.getd("cost_ratio_next", 3.0);
  auto starts = [&](const char* p) { return id.rfind(p, 0) == 0; };
  auto unsupported = [&](const std::string& why) { M.kind = "unsupported"; M.unsupported_reason = why; };
  // ---- reference IDs (src/mpsolve.cu:20-60)
  if (id == "native_lu32") { M.kind = "direct"; M.factor = "cs32"; M.working = 32; M.max_outer = 0; }
  else if (id == "native_lu64" || id == "native_lu64_promoted32") { M.kind = "direct"; M.factor = "cs64"; M.max_outer = 0; }
  else if (id == "native_ir32_w32_r64") { M.factor = "cs32"; M.working = 32; }
  else if (id == "native_ir32_w32_r32") { M.factor = "cs32"; M.working = 32; M.resid = 32; }
  else if (id == "native_ir32_w64_r64") { M.factor = "cs32"; }
  else if (id == "native_ir64_w64_r64") { M.factor = "cs64"; }
  else if (id == "native_ir32_w64_r64_twosided") { M.factor = "cs32"; M.scale_mode = 2; }
  else if (id == "native_ir32_w64_r64_global") { M.factor = "cs32"; M.scale_mode = 1; }
  else if (id == "native_gmresir32_w64_r64") { M.factor = "cs32"; M.gmres = true; }
  else if (id == "native_gmresir32_w32_r64") { M.factor = "cs32"; M.working = 32; M.gmres = true; }
  else if (id == "native_leftgmres32_app32_w64_r64") { M.factor = "cs32"; M.gmres = true; M.gmres_left = true; M.gmres_cycles = 1; }
  else if (id == "native_leftgmres32_app64_w64_r64") { M.factor = "cs32"; M.gmres = true; M.gmres_left = true; M.apply64 = true; M.gmres_cycles = 1; }
  else if (id == "native_gmresir32_app64_w64_r64") { M.factor = "cs32"; M.gmres = true; M.apply64 = true; }
  else if (id == "hetero16_lu_direct") { M.kind = "direct"; M.factor = "het16"; M.working = 32; M.scale_mode = 1; M.max_outer = 0; }
  else if (id == "hetero16_ir_w64_r64") { M.factor = "het16"; M.scale_mode = 1; }
  else if (id == "hetero16_ir_w32_r64") { M.factor = "het16"; M.working = 32; M.scale_mode = 1; }
  else if (id == "hetero16_ir_w64_r64_noscale") { M.factor = "het16"; M.scale_mode = 0; }
  else if (id == "hetero16_ir_w64_r64_twosided") { M.factor = "het16"; M.scale_mode = 2; }
  else if (id == "hetero16_gmresir_w64_r64") { M.factor = "het16"; M.scale_mode = 1; M.gmres = true; }
  else if (id == "hetero16_gmresir_w32_r64") { M.factor = "het16"; M.working = 32; M.scale_mode = 1; M.gmres = true; }
  else if (id == "hetero16_gmresir_app64_w64_r64") { M.factor = "het16"; M.scale_mode = 1; M.gmres = true; M.apply64 = true; }
  else if (id == "hetero16_leftgmres_app64_w64_r64") { M.factor = "het16"; M.scale_mode = 1; M.gmres = true; M.gmres_left = true; M.apply64 = true; M.gmres_cycles = 1; }
  else if (id == "custom32_lu_direct") { M.kind = "direct"; M.factor = "cus32"; M.working = 32; M.scale_mode = 1; M.max_outer = 0; }
  else if (id == "custom32_ir_w64_r64") { M.factor = "cus32"; M.scale_mode = 1; }
  else if (id == "custom32_ir_w32_r64") { M.factor = "cus32"; M.working = 32; M.scale_mode = 1; }
  else if (id == "policy_hetero16_stationary_only") { M.factor = "het16"; M.scale_mode = 1; }
  else if (id == "policy_hetero16_gmres_only") { M.factor = "het16"; M.scale_mode = 1; M.gmres = true; }
  else if (id == "policy_hetero16_escalate_immediate") { M.factor = "het16"; M.scale_mode = 1; M.policy = "budget"; M.budget = 0; }
  else if (starts("policy_hetero16_budget")) { M.factor = "het16"; M.scale_mode = 1; M.policy = "budget"; M.budget = std::stoi(id.substr(22)); M.gmres = true; M.gmres_budget = M.budget; }
  else if (id == "policy_hetero16_cost_aware") { M.factor = "het16"; M.scale_mode = 1; M.policy = "cost_aware"; M.gmres_budget = a.geti("gmres_budget", 120); }
  else if (id == "policy_native32_escalate_immediate") { M.factor = "cs32"; M.policy = "budget"; M.budget = 0; }
  else if (starts("policy_native32_budget")) { M.factor = "cs32"; M.policy = "budget"; M.budget = std::stoi(id.substr(22)); M.gmres = true; M.gmres_budget = M.budget; }
  else if (starts("cusolver_irs_")) { unsupported("cuSOLVER IRS is not available on this platform"); }
  else if (starts("magma_")) { unsupported("MAGMA is not linked on this platform"); }
  // ---- MI355X method IDs: per-tile power-of-two scaled low-precision trailing updates
  //      mx{16,bf16,8}_{lu_direct | ir_w64_r64 | ir_w64_rdd | gmresir_w64_r64 | gmresir_w64_rdd}, policy_mx8_ladder
  else if (id == "mx64_lu_direct") { M.kind = "direct"; M.factor = "mx64"; M.max_outer = 0; M.nb = a.geti("nb", 128); }
  else if (starts("mx16_") || starts("mxbf16_") || starts("mx8_") || starts("mx32_") || starts("mx16x3_")) {
    const size_t us = id.find('_');
    const std::string pre = id.substr(0, us), rest = id.substr(us + 1);
    M.factor = (pre == "mx32") ? "cus32" : pre; M.scale_mode = 1;
    if (rest == "lu_direct") { M.kind = "direct"; M.working = 32; M.max_outer = 0; }
    else if (rest == "ir_w64_r64") {}
    else if (rest == "ir_w64_rdd") { M.resid = 128; }
    else if (rest == "gmresir_w64_r64") { M.gmres = true; }
    else if (rest == "gmresir_w64_rdd") { M.gmres = true; M.resid = 128; }
    else throw MPError("unknown method id " + id, -1);
  }
  else if (id == "policy_mx8_ladder") { M.factor = "mx8"; M.scale_mode = 1; M.gmres = true; M.resid = 128; M.policy = "ladder"; }
  else if (id == "policy_mx16_ladder") { M.factor = "mx16"; M.scale_mode = 1; M.gmres = true; M.resid = 128; M.policy = "ladder16"; }
  else if (id == "policy_mx16x3_ladder") { M.factor = "mx16x3"; M.scale_mode = 1; M.gmres = true; M.resid = 128; M.policy = "ladder16x3"; }
  else throw MPError("unknown method id " + id, -1);
  return M;
}

static const char* kMethodIds[] = {
  "native_lu32", "native_lu64", "native_lu64_promoted32", "native_ir32_w32_r64", "native_ir32_w32_r32", "native_ir32_w64_r64",
  "native_ir64_w64_r64", "native_ir32_w64_r64_twosided", "native_ir32_w64_r64_global", "native_gmresir32_w64_r64",
  "native_gmresir32_w32_r64", "native_leftgmres32_app32_w64_r64", "native_leftgmres32_app64_w64_r64", "native_gmresir32_app64_w64_r64",
  "hetero16_lu_direct", "hetero16_ir_w64_r64", "hetero16_ir_w32_r64", "hetero16_ir_w64_r64_noscale", "hetero16_ir_w64_r64_twosided",
  "hetero16_gmresir_w64_r64", "hetero16_gmresir_w32_r64", "hetero16_gmresir_app64_w64_r64", "hetero16_leftgmres_app64_w64_r64",
  "custom32_lu_direct", "custom32_ir_w64_r64", "custom32_ir_w32_r64", "policy_hetero16_stationary_only", "policy_hetero16_gmres_only",
  "policy_hetero16_escalate_immediate", "policy_hetero16_budget120", "policy_hetero16_cost_aware", "policy_native32_escalate_immediate",
  "policy_native32_budget120",
  "mx16_lu_direct", "mx16_ir_w64_r64", "mx16_ir_w64_rdd", "mx16_gmresir_w64_r64", "mx16_gmresir_w64_rdd",
  "mxbf16_lu_direct", "mxbf16_ir_w64_r64", "mxbf16_ir_w64_rdd", "mxbf16_gmresir_w64_r64", "mxbf16_gmresir_w64_rdd",
  "mx8_lu_direct", "mx8_ir_w64_r64", "mx8_ir_w64_rdd", "mx8_gmresir_w64_r64", "mx8_gmresir_w64_rdd",
  "mx32_ir_w64_rdd", "mx32_gmresir_w64_rdd", "mx16x3_lu_direct", "mx16x3_ir_w64_rdd", "mx16x3_gmresir_w64_rdd",
  "policy_mx8_ladder", "policy_mx16_ladder", "policy_mx16x3_ladder", "mx64_lu_direct"};

static std::string factor_contract(const std::string& f) {
  if (f == "cs32") return "vendor FP32 LU with partial pivoting (rocSOLVER sgetrf) on the FP32 copy of A";
  if (f == "cs64") return "vendor FP64 LU with partial pivoting (rocSOLVER dgetrf) on an FP64 copy of A";
  if (f == "mx64") return "own FP64 blocked LU with partial pivoting: vendor panel (rocSOLVER dgetrf on m x nb) and U12 "
                          "(rocBLAS dtrsm), trailing update A22 -= L21*U12 on v_mfma_f64_16x16x4_f64, look-ahead on two streams";
  const std::string base = "own right-looking blocked LU (HIP, gfx950): FP32 storage; FP32 panel with partial pivoting "
                           "(largest |a|, ties to the smallest row); FP32 laswp and unit-lower TRSM for U12; trailing update "
                           "A22 -= L21*U12 on MFMA with FP32 accumulation and operands rounded (RNE) to ";
  if (f == "het16") return base + "FP16 after the global power-of-two scale only (no per-tile scaling)";
  if (f == "cus32") return base + "FP32 (v_mfma_f32_32x32x2_f32)";
  if (f == "mx16") return base + "FP16 with a power-of-two scale per 32-row tile of L21 and per 32-column tile of U12 (tile max in [2^14, 2^15))";
  if (f == "mxbf16") return base + "BF16 with a power-of-two scale per 32-row tile of L21 and per 32-column tile of U12";
  if (f == "mx8") return base + "OCP FP8 E4M3 with a power-of-two scale per 32-row tile of L21 and per 32-column tile of U12 (tile max in [2^7, 2^8))";
  if (f == "mx16x3") return base + "two FP16 terms per operand (hi = fp16(s*x), lo = fp16(s*x - hi), one power-of-two scale s per 32-row / "
                                   "32-column tile); the update sums hi*hi + hi*lo + lo*hi, three FP16 MFMA products with FP32 accumulation";
  return f;
}

// ------------------------------------------------------------------ factors
struct Factors {
  int64_t n = 0; std::string kind; int prec = 32;
  DBuf<float> LU32; DBuf<double> LU64; DBuf<int64_t> ipiv; DBuf<int32_t> perm; DBuf<int32_t> ipiv32;
  double gscale = 1.0; int gscale_exp = 0; int scale_mode = 0; int nb = 0;
  int info = 0; double growth = -1; unsigned long long conv_overflow = 0, conv_subnormal = 0; int panels = 0;
  int coop_refused = 0;   // panels launched as ordinary kernels after the runtime refused a cooperative launch
  std::string coop_error; // the runtime's code for the first refusal
  bool valid = false;
  DBuf<double> LU64p; bool promoted = false;   // FP64 copy of FP32 factors, applied by FP64 getrs (apply64)
  DBuf<double> dr, dc;                          // two-sided power-of-two scales (scale mode 2): F = g Dr A Dc
  mutable DBuf<double> rs;                      // Dr r, the scaled right-hand side of an application
};

// Waits for a growth measurement still running on the panel stream and stores it with its factors (nothing
// when none is pending).  Runs before anything writes or frees those factors and before the growth is reported.
static void growth_resolve(mp_ctx* C) {
  if (!C->growth_target) return;
  Factors* F = C->growth_target;
  C->growth_target = nullptr;
  HIP_CK(hipEventSynchronize(C->ev_growth));
  const double gm = C->hpin[3080];
  F->growth = (C->growth_den > 0) ? gm / C->growth_den : -1;
}

static int fmt_of(const std::string& kind) {
  if (kind == "mx16x3") return FMT_FP16X3;
  if (kind == "het16" || kind == "mx16") return FMT_FP16;
  if (kind == "mxbf16") return FMT_BF16;
  if (kind == "mx8") return FMT_FP8;
  return FMT_FP32;
}

static double dev_absmax(mp_ctx* C, const void* A, bool a32, int64_t n, int64_t lda) {
  absmax(A, a32, n, lda, C->hpin, C->W);
  HIP_CK(hipStreamSynchronize(C->W.stream));
  return C->hpin[0];
}

// ipiv (1-based, sequential swaps) -> composed permutation, on the host
static void perm_from_ipiv32(const std::vector<int>& ip, std::vector<int32_t>& perm) {
  const int64_t n = (int64_t)ip.size();
  perm.resize(n);
  for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
  for (int64_t i = 0; i < n; i++) { const int64_t p = ip[i] - 1; if (p != i) std::swap(perm[i], perm[p]); }
}

__global__ void k_f32_to_f64(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

__global__ void k_scale_f32(float* __restrict__ x, int64_t n, float s) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] *= s;
}

// ---- two-sided power-of-two equilibration (scale mode 2, src/mp_core.cuh:164-166 with the helpers of
// src/mp_kernels.cuh:25,105-108): dr_i = 2^-(e-1) for the row maximum m_i in [2^(e-1), 2^e) (1 for a zero or
// non-finite maximum), A <- Dr A; then dc_j the same for the columns of Dr A, A <- A Dc.  Maxima are taken
// over the non-NaN magnitudes and gathered with 64-bit integer atomics on the bit patterns (non-negative
// doubles order as integers); every scale is a power of two, so each scaled entry is exact.
__device__ inline void atomic_max_nonneg(double* p, double v) {
  atomicMax((unsigned long long*)p, (unsigned long long)__double_as_longlong(v));
}
template <typename T>
__global__ __launch_bounds__(256) void k_eq_rowmax(const T* __restrict__ A, int64_t n, double* __restrict__ dr) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const int64_t c0 = (int64_t)blockIdx.y * 64, c1 = c0 + 64 < n ? c0 + 64 : n;
  double m = 0.0;
  for (int64_t c = c0; c < c1; c++) m = fmax(m, fabs((double)A[c * n + r]));
  if (m > 0.0) atomic_max_nonneg(dr + r, m);
}
__global__ void k_eq_pow2(double* __restrict__ v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double m = v[i];
  if (!(m > 0.0) || !isfinite(m)) { v[i] = 1.0; return; }
  int e; frexp(m, &e);
  v[i] = ldexp(1.0, 1 - e);
}
// A <- diag(sr) A diag(sc) on column blockIdx.y, and the column's new absolute maximum into cmax[c]
// (cmax == nullptr: not gathered)
template <typename T>
__global__ __launch_bounds__(256) void k_eq_scale(T* __restrict__ A, int64_t n, const double* __restrict__ sr,
                                                  const double* __restrict__ sc, double* __restrict__ cmax) {
  __shared__ double red[4];
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t c = blockIdx.y;
  double m = 0.0;
  if (r < n) {
    const T v = (T)((double)A[c * n + r] * (sr ? sr[r] : 1.0) * (sc ? sc[c] : 1.0));
    A[c * n + r] = v;
    m = fabs((double)v);
    if (m != m) m = 0.0;
  }
  if (!cmax) return;
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    if (m > 0.0) atomic_max_nonneg(cmax + c, m);
  }
}
__global__ void k_scale_rows(const double* __restrict__ x, double* __restrict__ y, int64_t n, int nrhs,
                             const double* __restrict__ s) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n * nrhs) y[i] = x[i] * s[i % n];
}

template <typename T>
static void equilibrate(mp_ctx* C, T* A, int64_t n, Factors& F) {
  Workspace& W = C->W;
  F.dr.ensure(C->tracker, n, "dr"); F.dc.ensure(C->tracker, n, "dc");
  fill_f64(F.dr.p, n, 0.0, W); fill_f64(F.dc.p, n, 0.0, W);
  hipLaunchKernelGGL(k_eq_rowmax<T>, dim3(cdiv(n, 256), cdiv(n, 64)), dim3(256), 0, W.stream, A, n, F.dr.p);
  LAUNCH_CK();
  hipLaunchKernelGGL(k_eq_pow2, dim3(cdiv(n, 256)), dim3(256), 0, W.stream, F.dr.p, n);
  LAUNCH_CK();
  hipLaunchKernelGGL(k_eq_scale<T>, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, A, n, F.dr.p,
                     (const double*)nullptr, F.dc.p);
  LAUNCH_CK();
  hipLaunchKernelGGL(k_eq_pow2, dim3(cdiv(n, 256)), dim3(256), 0, W.stream, F.dc.p, n);
  LAUNCH_CK();
  hipLaunchKernelGGL(k_eq_scale<T>, dim3(cdiv(n, 256), (unsigned)n), dim3(256), 0, W.stream, A, n,
                     (const double*)nullptr, F.dc.p, (double*)nullptr);
  LAUNCH_CK();
}

static void factorize(mp_ctx* C, const mp_problem& P, const std::string& kind, int nb, int scale_mode, Factors& F) {
  Workspace& W = C->W;
  const int64_t n = P.n; const size_t nn = (size_t)n * n;
  growth_resolve(C);   // an earlier factorization's growth measurement may still be reading its factors
  W.factor_gen++;   // cached getrs block inverses of any earlier factors are stale from here on
  F.n = n; F.kind = kind; F.scale_mode = scale_mode; F.nb = nb; F.valid = false; F.growth = -1; F.promoted = false;
  F.conv_overflow = F.conv_subnormal = 0; F.info = 0; F.panels = 0; F.coop_refused = 0; F.coop_error.clear();
  F.prec = (kind == "cs64" || kind == "mx64") ? 64 : 32;
  F.ipiv.ensure(C->tracker, n, "ipiv");
  F.perm.ensure(C->tracker, n, "perm");
  // Global scale g = 2^-e for the maximum m = f 2^e (f in [0.5, 1)) of the copy in the factor precision
  // (src/mp_core.cuh:167).  Without two-sided scaling that maximum is A's maximum rounded to the factor
  // precision (rounding is monotonic), known before the copy is made, so the FP64 -> FP32 convert and the
  // scale run as one pass; with it, the maximum of the equilibrated copy is measured.
  F.gscale = 1.0; F.gscale_exp = 0;
  auto set_gscale = [&](double m) {
    if (m > 0 && std::isfinite(m)) { int e; std::frexp(m, &e); F.gscale_exp = -e; F.gscale = std::ldexp(1.0, -e); }
  };
  const double amax_copy = (F.prec == 32) ? (double)(float)P.amax : P.amax;
  if (scale_mode == 1) set_gscale(amax_copy);
  double amax_f = amax_copy * F.gscale;
  // a stage that changes the factor precision frees the other precision's buffers: a ladder keeps only the
  // factors it refines with (a cached problem then holds one factor matrix, which is what the serving loop counts)
  if (F.prec == 32) F.LU64.release(); else { F.LU32.release(); F.LU64p.release(); }
  const bool own_getrf = !(kind == "cs32" || kind == "cs64" || kind == "mx64");
  // the factor copy and the growth measurement use the second stream unless MPSOLVE_OVERLAP=0
  const bool two_streams = W.pstream && !(std::getenv("MPSOLVE_OVERLAP") && std::getenv("MPSOLVE_OVERLAP")[0] == '0');
  hipEvent_t cols_ready = nullptr;
  if (F.prec == 32) {
    F.LU32.ensure(C->tracker, nn, "LU32");
    if (P.fp32) {
      HIP_CK(hipMemcpyAsync(F.LU32.p, P.A32.p, nn * 4, hipMemcpyDeviceToDevice, W.stream));
    } else if (own_getrf && scale_mode != 2 && two_streams && W.ev_conv && n > nb) {
      // fl32(a) * fl32(g), as convert-then-scale.  The first panel's columns are converted first; the rest are
      // converted on the panel stream while that panel runs, and getrf waits for them after launching it.
      HIP_CK(hipEventRecord(W.ev_conv, W.stream));
      HIP_CK(hipStreamWaitEvent(W.pstream, W.ev_conv, 0));
      scale_convert_cols(P.A64.p, n, nb, n, F.LU32.p, n, F.gscale, W.stream);
      scale_convert_cols(P.A64.p + (int64_t)nb * n, n, n - nb, n, F.LU32.p + (int64_t)nb * n, n, F.gscale, W.pstream);
      HIP_CK(hipEventRecord(W.ev_conv, W.pstream));
      cols_ready = W.ev_conv;
    } else {
      scale_convert_f64_to_f32(P.A64.p, n, n, F.LU32.p, n, F.gscale, W);   // fl32(a) * fl32(g), as convert-then-scale
    }
    if (scale_mode == 2) {
      equilibrate<float>(C, F.LU32.p, n, F);
      amax_f = dev_absmax(C, F.LU32.p, true, n, n);
      set_gscale(amax_f); amax_f *= F.gscale;
    }
    // the reference scales the FP32 copy in place (src/mp_core.cuh:167): a power of two, exact in range
    if (F.gscale != 1.0 && (P.fp32 || scale_mode == 2)) {
      hipLaunchKernelGGL(k_scale_f32, dim3(cdiv(nn, 256)), dim3(256), 0, W.stream, F.LU32.p, (int64_t)nn, (float)F.gscale);
      LAUNCH_CK();
    }
  } else {
    F.LU64.ensure(C->tracker, nn, "LU64");
    if (P.fp32) {   // FP64 factors of an FP32-stored system: the exact FP64 values of its entries (src/mp_core.cuh:163)
      hipLaunchKernelGGL(k_f32_to_f64, dim3(cdiv(nn, 256)), dim3(256), 0, W.stream, P.A32.p, F.LU64.p, (int64_t)nn);
      LAUNCH_CK();
    } else {
      HIP_CK(hipMemcpyAsync(F.LU64.p, P.A64.p, nn * 8, hipMemcpyDeviceToDevice, W.stream));
    }
    if (scale_mode == 2) {
      equilibrate<double>(C, F.LU64.p, n, F);
      amax_f = dev_absmax(C, F.LU64.p, false, n, n);
      set_gscale(amax_f); amax_f *= F.gscale;
    }
    // the reference scales the FP64 copy too (src/mp_core.cuh:167) and applies the factors with g
    if (F.gscale != 1.0) axpy_scale_inplace(F.LU64.p, (int64_t)nn, F.gscale, W);
  }
  if (kind == "cs32" || kind == "cs64" || kind == "mx64") {
    F.ipiv32.ensure(C->tracker, n + 1, "ipiv32");
    RB_CK(rocblas_set_stream(C->rb, W.stream));
    int info64 = 0;
    if (kind == "cs32") RB_CK(rocsolver_sgetrf(C->rb, (rocblas_int)n, (rocblas_int)n, F.LU32.p, (rocblas_int)n, F.ipiv32.p, F.ipiv32.p + n));
    else if (kind == "cs64") RB_CK(rocsolver_dgetrf(C->rb, (rocblas_int)n, (rocblas_int)n, F.LU64.p, (rocblas_int)n, F.ipiv32.p, F.ipiv32.p + n));
    else {
      const char* ov = std::getenv("MPSOLVE_OVERLAP");
      getrf64(n, F.LU64.p, n, F.ipiv32.p, nb, !(ov && ov[0] == '0'), (void*)C->rb, (void*)C->rb_panel, W, &info64);
      RB_CK(rocblas_set_stream(C->rb, W.stream));
    }
    std::vector<int> ip(n + 1);
    HIP_CK(hipMemcpyAsync(ip.data(), F.ipiv32.p, (kind == "mx64" ? n : n + 1) * sizeof(int), hipMemcpyDeviceToHost, W.stream));
    HIP_CK(hipStreamSynchronize(W.stream));
    F.info = (kind == "mx64") ? info64 : ip[n];
    ip.resize(n);
    std::vector<int32_t> perm; perm_from_ipiv32(ip, perm);
    std::vector<int64_t> ip64(ip.begin(), ip.end());
    HIP_CK(hipMemcpyAsync(F.perm.p, perm.data(), n * 4, hipMemcpyHostToDevice, W.stream));
    HIP_CK(hipMemcpyAsync(F.ipiv.p, ip64.data(), n * 8, hipMemcpyHostToDevice, W.stream));
  } else {
    GetrfArgs ga; ga.n = n; ga.A = F.LU32.p; ga.lda = n; ga.nb = nb; ga.fmt = fmt_of(kind);
    ga.tile_scale = (kind != "het16"); ga.ipiv = F.ipiv.p; ga.perm = F.perm.p; ga.cols_ready = cols_ready;
    GetrfInfo gi;
    getrf(ga, W, &gi);
    // a panel whose cross-workgroup exchange hit its spin limit left the factors incomplete
    if (gi.timeout) throw MPError("factorization: the panel exchange timed out (workgroups not co-resident)", MP_ERR_INTERNAL);
    F.info = gi.info; F.conv_overflow = gi.overflow; F.conv_subnormal = gi.subnormal; F.panels = gi.panels;
    F.coop_refused = gi.coop_refused; F.coop_error = gi.coop_error;
  }
  if (own_getrf && F.prec == 32 && two_streams && C->ev_growth && C->hpin) {
    // max|LU| on the panel stream (its own partials), beside the substitution set-up and first applications
    // that follow on the main stream; growth_resolve collects it
    W.gpart.ensure(*W.tracker, (size_t)absmax_partials(n), "growth_partials");
    HIP_CK(hipEventRecord(C->ev_growth, W.stream));
    HIP_CK(hipStreamWaitEvent(W.pstream, C->ev_growth, 0));
    absmax_on(F.LU32.p, true, n, n, C->hpin + 3080, W.gpart.p, W.pstream);
    HIP_CK(hipEventRecord(C->ev_growth, W.pstream));
    F.growth = -1;
    C->growth_target = &F; C->growth_den = amax_f;
  } else {
    const double gm = (F.prec == 32) ? dev_absmax(C, F.LU32.p, true, n, n) : dev_absmax(C, F.LU64.p, false, n, n);
    F.growth = (amax_f > 0) ? gm / amax_f : -1;
  }
  F.valid = true;
}

// FP64 copy of FP32 factors with int32 LAPACK pivots, for the promoted application (the reference's
// LU64_promoted, src/mp_core.cuh:178)
static void promote_factors(mp_ctx* C, Factors& F) {
  if (F.prec != 32 || F.promoted) return;
  Workspace& W = C->W;
  const int64_t n = F.n; const size_t nn = (size_t)n * n;
  F.LU64p.ensure(C->tracker, nn, "LU64_promoted");
  hipLaunchKernelGGL(k_f32_to_f64, dim3(cdiv(nn, 256)), dim3(256), 0, W.stream, F.LU32.p, F.LU64p.p, (int64_t)nn);
  LAUNCH_CK();
  if (F.kind != "cs32") {      // own factorizations report int64 pivots
    std::vector<int64_t> ip(n);
    HIP_CK(hipMemcpyAsync(ip.data(), F.ipiv.p, n * 8, hipMemcpyDeviceToHost, W.stream));
    HIP_CK(hipStreamSynchronize(W.stream));
    std::vector<int32_t> ip32(ip.begin(), ip.end());
    F.ipiv32.ensure(C->tracker, n + 1, "ipiv32");
    HIP_CK(hipMemcpyAsync(F.ipiv32.p, ip32.data(), n * 4, hipMemcpyHostToDevice, W.stream));
  }
  HIP_CK(hipStreamSynchronize(W.stream));
  F.promoted = true;
}

// d = A^{-1} r through the factors (r, d FP64 device, n x nrhs, leading dimension n); with apply64 the FP64
// copy of FP32 factors is applied in FP64
// active (device word, own FP32 substitutions only): the substitutions return at once when it is 0.
static void apply_inverse(mp_ctx* C, const Factors& F, const double* r, double* d, int nrhs, int* trs, bool apply64 = false,
                          const int* active = nullptr) {
  Workspace& W = C->W;
  const int64_t n = F.n; const int64_t m = n * nrhs;
  // with F = g Dr A Dc: A^{-1} r = g Dc F^{-1} (Dr r)  (src/mp_core.cuh:184,188-189)
  if (F.scale_mode == 2) {
    F.rs.ensure(C->tracker, (size_t)m, "rs");
    hipLaunchKernelGGL(k_scale_rows, dim3(cdiv(m, 256)), dim3(256), 0, W.stream, r, F.rs.p, n, nrhs, F.dr.p);
    LAUNCH_CK();
    r = F.rs.p;
  }
  const bool fp64 = F.prec == 64 || (apply64 && F.promoted);
  if (fp64) {
    const double* LU = (F.prec == 64) ? F.LU64.p : F.LU64p.p;
    if (d != r) HIP_CK(hipMemcpyAsync(d, r, (size_t)m * 8, hipMemcpyDeviceToDevice, W.stream));
    RB_CK(rocblas_set_stream(C->rb, W.stream));
    RB_CK(rocsolver_dgetrs(C->rb, rocblas_operation_none, (rocblas_int)n, (rocblas_int)nrhs, const_cast<double*>(LU),
                           (rocblas_int)n, F.ipiv32.p, d, (rocblas_int)n));
  } else {
    getrs(n, nrhs, F.LU32.p, n, F.perm.p, r, n, d, n, F.gscale, W, true, active);    // the scale by g is fused
  }
  if (F.scale_mode == 2) {
    hipLaunchKernelGGL(k_scale_rows, dim3(cdiv(m, 256)), dim3(256), 0, W.stream, d, d, n, nrhs, F.dc.p);
    LAUNCH_CK();
  }
  if (fp64 && F.gscale != 1.0) axpy_scale_inplace(d, m, F.gscale, W);   // (gA)^{-1} r = A^{-1} r / g
  if (trs) (*trs)++;
}

// ------------------------------------------------------------------ metrics
struct Metrics {
  std::vector<double> rinf, r2, eta, omega, relres_inf, relres_2, xinf; std::vector<int64_t> worst;
  bool finite = true; double eta_max = 0, omega_max = 0;
  bool has_fwd = false, has_exact = false;
  std::vector<double> finf, f2, fcomp, fzero_maxabs, einf, e2, ecomp, ezero_maxabs;
  std::vector<int64_t> fzero_count, ezero_count;
};

struct EvalBufs {
  DBuf<double> R, Rlo, T, stats; DBuf<int32_t> worst;
  void ensure(DevTracker& tr, int64_t n, int nrhs) {
    R.ensure(tr, (size_t)n * nrhs, "R"); T.ensure(tr, (size_t)n * nrhs, "T"); Rlo.ensure(tr, (size_t)n * nrhs, "Rlo");
    stats.ensure(tr, (size_t)4 * nrhs + 8, "stats"); worst.ensure(tr, (size_t)nrhs + 1, "worst");
  }
  void swap(EvalBufs& o) { R.swap(o.R); Rlo.swap(o.Rlo); T.swap(o.T); stats.swap(o.stats); worst.swap(o.worst); }
};

// Residual of the stored system for device X, FP64 or double-double, and the per-RHS metrics
// (eval_metrics_T, src/mp_core.cuh:193-205).  R receives the FP64 (rounded) residual.  launch = false: E already
// holds that residual and its statistics (formed by an earlier residual_pair); only the metrics are read.
static void residual_metrics(mp_ctx* C, const mp_problem& P, const double* X, bool dd, EvalBufs& E, Metrics& M,
                             bool launch = true) {
  Workspace& W = C->W;
  const int64_t n = P.n; const int nrhs = P.nrhs;
  E.ensure(C->tracker, n, nrhs);
  if (launch) {
    ResidArgs ra; ra.n = n; ra.nrhs = nrhs; ra.a32 = P.fp32; ra.A = P.a_ptr(); ra.lda = n; ra.X = X; ra.ldx = n;
    ra.B = P.B.p; ra.ldb = n; ra.R = E.R.p; ra.Rlo = dd ? E.Rlo.p : nullptr; ra.ldr = n; ra.T = E.T.p; ra.ldt = n;
    ra.stats = E.stats.p; ra.worst = E.worst.p; ra.dd = dd;
    residual(ra, W);
  }
  // read back through the context's pinned buffer (doubles [2048, 3072)) when the right-hand sides fit
  std::vector<double> st_v; std::vector<int32_t> wi_v;
  const double* st; const int32_t* wi;
  if (C->hpin && nrhs <= 200) {
    double* hs = C->hpin + 2048;
    int32_t* hw = (int32_t*)(hs + 4 * nrhs);
    HIP_CK(hipMemcpyAsync(hs, E.stats.p, 32 * nrhs, hipMemcpyDeviceToHost, W.stream));
    HIP_CK(hipMemcpyAsync(hw, E.worst.p, 4 * nrhs, hipMemcpyDeviceToHost, W.stream));
    st = hs; wi = hw;
  } else {
    st_v.resize(4 * nrhs); wi_v.resize(nrhs);
    HIP_CK(hipMemcpyAsync(st_v.data(), E.stats.p, 32 * nrhs, hipMemcpyDeviceToHost, W.stream));
    HIP_CK(hipMemcpyAsync(wi_v.data(), E.worst.p, 4 * nrhs, hipMemcpyDeviceToHost, W.stream));
    st = st_v.data(); wi = wi_v.data();
  }
  HIP_CK(hipStreamSynchronize(W.stream));
  M.rinf.assign(nrhs, 0); M.r2.assign(nrhs, 0); M.eta.assign(nrhs, 0); M.omega.assign(nrhs, 0);
  M.relres_inf.assign(nrhs, 0); M.relres_2.assign(nrhs, 0); M.xinf.assign(nrhs, 0); M.worst.assign(nrhs, -1);
  M.finite = true; M.eta_max = 0; M.omega_max = 0;
  for (int c = 0; c < nrhs; c++) {
    const double ri = st[4 * c], xi = st[4 * c + 1], om = st[4 * c + 2], r2 = std::sqrt(st[4 * c + 3]);
    M.rinf[c] = ri; M.r2[c] = r2; M.xinf[c] = xi; M.worst[c] = wi[c];
    const double den = P.anorm_inf * xi + P.bnorm_inf[c];
    M.eta[c] = (ri == 0) ? 0.0 : ((den == 0) ? INFINITY : ri / den);
    M.omega[c] = om;
    M.relres_inf[c] = (ri == 0) ? 0.0 : ((P.bnorm_inf[c] == 0) ? INFINITY : ri / P.bnorm_inf[c]);
    M.relres_2[c] = (r2 == 0) ? 0.0 : ((P.bnorm_2[c] == 0) ? INFINITY : r2 / P.bnorm_2[c]);
    if (!std::isfinite(ri) || !std::isfinite(xi) || !std::isfinite(om)) M.finite = false;
    M.eta_max = std::max(M.eta_max, M.eta[c]); M.omega_max = std::max(M.omega_max, M.omega[c]);
    if (!std::isfinite(M.eta[c])) M.eta_max = M.eta[c];
  }
}

// forward-error statistics vs a host reference (k_fwdstats, src/mp_kernels.cuh:98-102), host side
static void fwd_stats(const std::vector<double>& X, const std::vector<double>& XR, int64_t n, int nrhs,
                      std::vector<double>& finf, std::vector<double>& f2, std::vector<double>& fcomp,
                      std::vector<double>& fz, std::vector<int64_t>& fzc) {
  finf.assign(nrhs, 0); f2.assign(nrhs, 0); fcomp.assign(nrhs, 0); fz.assign(nrhs, 0); fzc.assign(nrhs, 0);
  for (int c = 0; c < nrhs; c++) {
    double m0 = 0, m1 = 0, s2 = 0, s3 = 0, m4 = 0, m5 = 0; int64_t zc = 0;
    for (int64_t i = 0; i < n; i++) {
      const double x = X[i + c * n], xr = XR[i + c * n], e = std::fabs(x - xr), a = std::fabs(xr);
      if (!(e <= m0)) m0 = e;
      if (a > m1) m1 = a;
      s2 += e * e; s3 += xr * xr;
      if (xr != 0.0) { const double q = e / a; if (!(q <= m4)) m4 = q; } else { zc++; if (!(e <= m5)) m5 = e; }
    }
    const double e2 = std::sqrt(s2), x2 = std::sqrt(s3);
    finf[c] = (m1 > 0) ? m0 / m1 : (m0 == 0 ? 0 : INFINITY);
    f2[c] = (x2 > 0) ? e2 / x2 : (e2 == 0 ? 0 : INFINITY);
    fcomp[c] = m4; fz[c] = m5; fzc[c] = zc;
  }
}

// ------------------------------------------------------------------ trial output and engine
struct HistRec { int outer, stage; double t, eta_max, omega_max, rinf_max; int inner_in_step; std::string note; };
struct InnerRec { int outer, cycle, j, rhs; double rec_rel, true_rel; std::string note; };
struct TrialOut {
  std::string terminal = "unknown", final_stage; bool fallback = false; int refactorizations = 0;
  std::vector<std::string> stages; std::vector<double> stage_times;
  int outer = 0, inner_total = 0, matvecs = 0, precond_applies = 0, trs = 0, restarts = 0, true_resid_checks = 0;
  double t_total = 0, t_factor_total = 0, t_initial = 0, t_refine = 0, t_eval = 0;
  std::vector<HistRec> hist; std::vector<InnerRec> ihist;
  double growth = -1; unsigned long long conv_overflow = 0, conv_subnormal = 0; int factor_info = 0;
  int panels = 0, panel_coop_refused = 0;   // panel launches of the own factorizations, and how many ran non-cooperatively
  std::string panel_coop_error;             // the runtime's code for the first refused cooperative launch
  std::string error; bool ok = true; double final_eta = 0, final_omega = 0;
};

struct Engine {
  mp_ctx* C; const mp_problem& P; Factors fact; bool reuse = false; bool resume_ladder = false; bool instrument = false;
  bool stage_abandon = false;   // this build's ladders: leave a stage whose first GMRES cycle predicts it will not pay
  int64_t n; int nrhs; double tau_eta, tau_omega; bool stop_eta_only = false;
  DBuf<double> x64, D, y; DBuf<float> x32, r32;
  // ladder policies of this build: the iterate with the smallest eta so far, where the next stage starts
  bool keep_best = false; double best_eta = INFINITY; DBuf<double> xbest;
  EvalBufs E;
  // The end of an FGMRES solve and the next outer residual share one pass over A (residual_pair): the cycle's
  // check r - A d and the residual of x + d (into E2, with xn = x + d).  pair_done: E2 holds that residual for
  // the solve's final d; pair_ready: x was then updated by that d, so the next outer residual is E2.
  EvalBufs E2; DBuf<double> xn; bool pair_done = false, pair_ready = false;
  Engine(mp_ctx* c, const mp_problem& p, double te, double to) : C(c), P(p), n(p.n), nrhs(p.nrhs), tau_eta(te), tau_omega(to) {
    x64.alloc(C->tracker, (size_t)n * nrhs, "x64");
    D.alloc(C->tracker, (size_t)n * nrhs, "D");
  }
  Factors* growth_src = nullptr;   // factors of the last stage: their growth is reported
  ~Engine() {
    // a run that ended early (an error) may leave the growth measurement reading factors freed below
    if (C->growth_target) { (void)hipEventSynchronize(C->ev_growth); C->growth_target = nullptr; }
  }
  Workspace& W() { return C->W; }
  void sync() { HIP_CK(hipStreamSynchronize(C->W.stream)); }

  Factors& factors_for(const MethodSpec& M, const std::string& kind, TrialOut& T) {
    if (reuse && !P.fcache) P.fcache = std::make_shared<Factors>();
    Factors& F = reuse ? *P.fcache : fact;
    const double t0 = wall_now();
    if (reuse && F.valid && F.kind == kind && F.scale_mode == M.scale_mode && F.nb == M.nb) {
      if (M.apply64) promote_factors(C, F);
      growth_src = &F;
      T.stages.push_back(kind + "(reused)"); T.stage_times.push_back(0.0); T.growth = F.growth; T.factor_info = F.info; return F;
    }
    growth_src = &F;
    factorize(C, P, kind, M.nb, M.scale_mode, F);
    if (M.apply64 && F.info == 0) promote_factors(C, F);
    sync();
    const double dt = wall_now() - t0;
    T.t_factor_total += dt; T.stages.push_back(kind); T.stage_times.push_back(dt); T.refactorizations++;
    T.growth = F.growth; T.conv_overflow = F.conv_overflow; T.conv_subnormal = F.conv_subnormal; T.factor_info = F.info;
    T.panels += F.panels; T.panel_coop_refused += F.coop_refused;
    if (T.panel_coop_error.empty()) T.panel_coop_error = F.coop_error;
    return F;
  }

  // x (working precision) += d
  void add_update(const MethodSpec& M, const double* d) {
    if (M.working == 32) {
      // x32 += d rounded to FP32, then refresh the FP64 copy used by the residual
      hipLaunchKernelGGL(k_add_x32, dim3(cdiv((size_t)n * nrhs, 256)), dim3(256), 0, W().stream, x32.p, x64.p, d, (int64_t)n * nrhs);
      LAUNCH_CK();
    } else {
      axpy(x64.p, d, (int64_t)n * nrhs, 1.0, W());
    }
  }
  void set_x(const MethodSpec& M, const double* d) {
    if (M.working == 32) {
      x32.ensure(C->tracker, (size_t)n * nrhs, "x32");
      hipLaunchKernelGGL(k_set_x32, dim3(cdiv((size_t)n * nrhs, 256)), dim3(256), 0, W().stream, x32.p, x64.p, d, (int64_t)n * nrhs);
      LAUNCH_CK();
    } else {
      copy_f64(d, x64.p, (int64_t)n * nrhs, W());
    }
  }

  // FGMRES (right, flexible) on A d = r for one RHS column (src/mp_core.cuh:237-278)
  // Stage economics for this build's ladders: after the first cycle of a stage's first correction, the gain of
  // that cycle (rho, the correction equation's true relative residual) predicts what finishing the stage costs:
  // every remaining solve (right-hand sides x outer steps, one step gaining gmres_tol from the current eta down
  // to tau) needs log(tol)/log(rho) cycles of m iterations.  When that exceeds the next factorization's cost in
  // iterations (cost_ratio_next, the reference's cost-aware ratio, src/mp_core.cuh:304-311, times
  // kFactorIters) or the remaining budget, the stage is left.  The rule uses no clock, so a solve's stages
  // (and its result) do not depend on timing.
  // One factorization costs about as much as 100 FGMRES iterations (a preconditioner application and an FP64
  // matrix-vector product): 79 ms against 0.8 ms at n = 16384, 264 ms against 2.5 ms at n = 32768.
  static constexpr double kFactorIters = 100.0;
  struct AbandonTest { double eta = 0, tau = 0, refactor_iters = 0; int nrhs = 1; };
  std::string gmres_solve(const MethodSpec& M, const Factors& F, const double* r, double* d, int& budget_left, TrialOut& T, int outer, int rhs,
                          const AbandonTest* ab = nullptr) {
    if (!C->kry) { C->kry.reset(new Krylov()); C->kry->W = &C->W; }
    Krylov& K = *C->kry;
    const int m = M.gmres_m;
    K.ensure(n, m);
    double* hp = C->hpin;
    pair_done = false;
    // the cycle's check and the next outer residual share a pass over A when the update x += d follows as it is
    // (one right-hand side, FP64 iterate, FP64 or double-double residual); MPSOLVE_RESID_PAIR=0 keeps them apart
    const bool pair_ok = nrhs == 1 && M.working == 64 && (M.resid == 64 || M.resid == 128) &&
                         !(std::getenv("MPSOLVE_RESID_PAIR") && std::getenv("MPSOLVE_RESID_PAIR")[0] == '0');
    // the left variant works on the preconditioned residual F^{-1} r (src/mp_core.cuh:240-241)
    if (M.gmres_left) {
      fill_f64(d, n, 0.0, W());
      apply_inverse(C, F, r, K.rcur.p, 1, &T.trs, M.apply64); T.precond_applies++;
      K.norm(K.rcur.p, K.scal());
    } else {
      K.start(r, d, K.scal());   // d = 0, rcur = r, ||r||
    }
    HIP_CK(hipMemcpyAsync(hp, K.scal(), 8, hipMemcpyDeviceToHost, W().stream));
    sync();
    const double beta0 = hp[0];
    if (beta0 == 0.0 || !std::isfinite(beta0)) return beta0 == 0.0 ? "zero_rhs" : "nonfinite";
    std::string status = "budget";
    int cycles = 0;
    while (cycles < M.gmres_cycles && budget_left > 0) {
      cycles++; if (cycles > 1) T.restarts++;
      // the first cycle starts from the vector whose norm is beta0 (the same kernels on the same data give the
      // same bits), so only a restart measures it again
      double betac = beta0;
      if (cycles > 1) {
        K.norm(K.rcur.p, K.scal());
        HIP_CK(hipMemcpyAsync(hp, K.scal(), 8, hipMemcpyDeviceToHost, W().stream));
        sync();
        betac = hp[0];
      }
      if (betac == 0) { status = "converged_true"; break; }
      K.scale_copy(K.rcur.p, K.V.p, K.scal(), true);   // v0 = rcur / betac
      // Right-preconditioned steps on the library's own FP32 substitutions run one ahead of the host: step j+1 is
      // enqueued before step j's numbers are read, and the Givens kernel of the step that ends the cycle clears
      // the activity word, so the kernels of a step enqueued past it return at once.  The arithmetic of every
      // step is the one below.
      const bool fp64_apply = F.prec == 64 || (M.apply64 && F.promoted);
      const bool ahead = !M.gmres_left && !fp64_apply && F.scale_mode != 2 && m <= 255;
      K.init_cycle(betac, ahead ? K.active() : nullptr);
      int j = 0; bool happy = false, nonfinite = false;
      if (ahead) {
        int* act = K.active();
        // Each step's Givens kernel writes rec, hn, finite and then the step's tag into a coherent pinned slot,
        // and the host waits by reading the tag: no copy kernel and no event in the stream between steps.
        // Without such slots the outputs are copied back and waited for with events.
        const bool poll = K.hstep != nullptr && m <= 255;
        double* slots = poll ? K.hstep : hp + 64;       // pinned: per step rec, hn, finite (, tag)
        unsigned long long tags[256] = {};
        if (!poll)
          for (int e = 0; e < 2; e++)
            if (!K.ev[e]) HIP_CK(hipEventCreateWithFlags(&K.ev[e], hipEventDisableTiming));
        bool prev_formed = false;
        // z = F^{-1} v and w = A z: with FP32 factors applied unscaled, the product is formed by the upper solve's
        // own extra workgroups as its blocks appear (the same bits as the separate product); MPSOLVE_GETRS_MV=0
        // keeps them apart
        const bool fuse_product = F.prec == 32 && F.scale_mode != 2 &&
                                  !(std::getenv("MPSOLVE_GETRS_MV") && std::getenv("MPSOLVE_GETRS_MV")[0] == '0');
        auto enqueue = [&](int s) {
          double* vs = K.V.p + (int64_t)s * n;
          double* zs = K.Z.p + (int64_t)s * n;
          bool formed = false;
          ProductReq pq; pq.A = P.a_ptr(); pq.a32 = P.fp32; pq.w = K.w.p; pq.partials_only = true;
          // v_s = w / hn of the previous step, formed by this step's lower solve as it reads it (and stored in V)
          if (s > 0 && prev_formed) { pq.scale_src = K.w.p; pq.scale_div = K.scal(); }
          if (fuse_product) {
            formed = getrs(n, 1, F.LU32.p, n, F.perm.p, vs, n, zs, n, F.gscale, W(), true, act, &pq);
          } else {
            apply_inverse(C, F, vs, zs, 1, nullptr, M.apply64, act);
          }
          if (!formed) {
            ResidArgs ra; ra.n = n; ra.nrhs = 1; ra.a32 = P.fp32; ra.A = P.a_ptr(); ra.lda = n; ra.X = zs; ra.ldx = n;
            ra.R = K.w.p; ra.ldr = n; ra.matvec = true; ra.active = act;
            residual(ra, W());
          }
          // CGS2, ||w||, Givens (the step that ends the cycle clears act); a product formed beside the upper solve
          // left its column-chunk partials, which the first CGS launch reduces into w
          const double* pp = formed ? pq.part : nullptr;
          const int64_t pn = formed ? pq.nch : 0;
          if (poll) {
            tags[s] = ++K.tag;
            K.arnoldi_tail(s, beta0, M.gmres_tol, act, slots + 4 * s, tags[s], pp, pn);
          } else {
            K.arnoldi_tail(s, beta0, M.gmres_tol, act, nullptr, 0, pp, pn);
            HIP_CK(hipMemcpyAsync(slots + 4 * s, K.scal() + 2, 24, hipMemcpyDeviceToHost, W().stream));
            HIP_CK(hipEventRecord(K.ev[s & 1], W().stream));
          }
          // v_{s+1} = w / hn: here, or by the next step's lower solve when this step's product was fused
          if (!formed) K.scale_copy(K.w.p, K.V.p + (int64_t)(s + 1) * n, K.scal(), true, act);
          prev_formed = formed;
        };
        // waits until step s has written its slot; a stream that drained without it is an error
        auto wait_step = [&](int s) {
          const unsigned long long* tp = (const unsigned long long*)(slots + 4 * s + 3);
          for (long it = 1;; it++) {
            if (__atomic_load_n(tp, __ATOMIC_ACQUIRE) == tags[s]) return;
            if ((it & 4095) == 0 && hipStreamQuery(W().stream) == hipSuccess &&
                __atomic_load_n(tp, __ATOMIC_ACQUIRE) != tags[s])
              throw MPError("FGMRES: an Arnoldi step ended without reporting its outputs", MP_ERR_INTERNAL);
          }
        };
        // a step goes out only if the budget lets it run after the steps already in flight
        int enq = 0;
        if (m > 0 && budget_left > 0) { enqueue(0); enq = 1; }
        for (j = 0; j < enq; j++) {
          if (enq < m && enq - j < budget_left) { enqueue(enq); enq++; }
          if (poll) wait_step(j); else HIP_CK(hipEventSynchronize(K.ev[j & 1]));
          const double rec = slots[4 * j], hn = slots[4 * j + 1];
          const bool fin = slots[4 * j + 2] != 0.0;
          T.trs++; T.precond_applies++; T.matvecs++;
          if (!fin) { nonfinite = true; break; }
          T.inner_total++; budget_left--;
          if (instrument) T.ihist.push_back({outer, cycles, j, rhs, rec, -1, ""});
          if (hn == 0.0) { happy = true; j++; break; }
          if (rec <= M.gmres_tol) { j++; break; }
        }
      }
      for (; !ahead && j < m && budget_left > 0; j++) {
        double* vj = K.V.p + (int64_t)j * n;
        if (M.gmres_left) {
          // w = F^{-1} (A v_j)
          ResidArgs ra; ra.n = n; ra.nrhs = 1; ra.a32 = P.fp32; ra.A = P.a_ptr(); ra.lda = n; ra.X = vj; ra.ldx = n;
          ra.R = K.sR.p; ra.ldr = n; ra.matvec = true;
          residual(ra, W()); T.matvecs++;
          apply_inverse(C, F, K.sR.p, K.w.p, 1, &T.trs, M.apply64); T.precond_applies++;
        } else {
          double* zj = K.Z.p + (int64_t)j * n;
          apply_inverse(C, F, vj, zj, 1, &T.trs, M.apply64); T.precond_applies++;
          // w = A z_j (FP64)
          ResidArgs ra; ra.n = n; ra.nrhs = 1; ra.a32 = P.fp32; ra.A = P.a_ptr(); ra.lda = n; ra.X = zj; ra.ldx = n;
          ra.R = K.w.p; ra.ldr = n; ra.matvec = true;
          residual(ra, W()); T.matvecs++;
        }
        // CGS2
        K.dots(K.V.p, j + 1, K.w.p, K.h(), false);
        K.update(K.V.p, j + 1, K.h(), K.w.p, -1.0);
        K.dots(K.V.p, j + 1, K.w.p, K.h2(), false);
        K.update(K.V.p, j + 1, K.h2(), K.w.p, -1.0);
        hipLaunchKernelGGL(k_add_vec, dim3(cdiv(j + 1, 64)), dim3(64), 0, W().stream, K.h(), K.h2(), j + 1, (const int*)nullptr);
        LAUNCH_CK();
        K.norm(K.w.p, K.scal());
        K.givens(j, K.h(), K.scal(), beta0, K.scal() + 2);
        // v_{j+1} = w / hn  (skipped by the kernel reading hn == 0 would divide by zero: guard on host below)
        HIP_CK(hipMemcpyAsync(hp, K.scal() + 2, 24, hipMemcpyDeviceToHost, W().stream));
        sync();
        const double rec = hp[0], hn = hp[1];
        const bool fin = hp[2] != 0.0;
        if (!fin) { nonfinite = true; break; }
        if (hn > 0) K.scale_copy(K.w.p, K.V.p + (int64_t)(j + 1) * n, K.scal(), true);
        T.inner_total++; budget_left--;
        if (instrument) T.ihist.push_back({outer, cycles, j, rhs, rec, -1, ""});
        if (hn == 0.0) { happy = true; j++; break; }
        if (rec <= M.gmres_tol) { j++; break; }
      }
      if (nonfinite) { status = "nonfinite"; break; }
      const int k = j;
      if (k == 0) break;
      K.lsq(k, K.y());
      K.update(M.gmres_left ? K.V.p : K.Z.p, k, K.y(), d, 1.0);   // left: d += V y; right (flexible): d += Z y
      // true residual of the correction equation: sR = r - A d (FP64)
      if (pair_ok) {
        // with the residual of x + d, which becomes the next outer residual if this is the solve's last cycle
        xn.ensure(C->tracker, (size_t)n, "x_next");
        E2.ensure(C->tracker, n, 1);
        axpy_to(xn.p, x64.p, d, n, 1.0, W());
        ResidPairArgs pa; pa.n = n; pa.a32 = P.fp32; pa.A = P.a_ptr(); pa.X1 = d; pa.B1 = r; pa.R1 = K.sR.p;
        pa.X2 = xn.p; pa.B2 = P.B.p; pa.R2 = E2.R.p; pa.R2lo = E2.Rlo.p; pa.T2 = E2.T.p; pa.stats2 = E2.stats.p;
        pa.worst2 = E2.worst.p; pa.dd2 = M.resid == 128;
        residual_pair(pa, W());
        pair_done = true;
      } else {
        ResidArgs ra; ra.n = n; ra.nrhs = 1; ra.a32 = P.fp32; ra.A = P.a_ptr(); ra.lda = n; ra.X = d; ra.ldx = n;
        ra.B = r; ra.ldb = n; ra.R = K.sR.p; ra.ldr = n;
        residual(ra, W());
      }
      T.matvecs++; T.true_resid_checks++;
      K.norm(K.sR.p, K.scal());
      // ||r||: the right-preconditioned cycle started from a copy of r, whose norm is beta0 (the same kernels on
      // the same data); the left one started from F^{-1} r
      if (M.gmres_left) K.norm(r, K.scal() + 1);
      HIP_CK(hipMemcpyAsync(hp, K.scal(), M.gmres_left ? 16 : 8, hipMemcpyDeviceToHost, W().stream));
      HIP_CK(hipMemcpyAsync(hp + 4, K.scal() + 2, 8, hipMemcpyDeviceToHost, W().stream));
      sync();
      const double tn = hp[0], rn0 = M.gmres_left ? hp[1] : beta0;
      const double true_rel = (rn0 > 0) ? tn / rn0 : 0.0;
      if (instrument) T.ihist.push_back({outer, cycles, k - 1, rhs, hp[4], true_rel, "cycle_end"});
      if (M.gmres_left) { apply_inverse(C, F, K.sR.p, K.rcur.p, 1, &T.trs, M.apply64); T.precond_applies++; }
      if (!std::isfinite(tn)) { status = "nonfinite"; break; }
      if (true_rel <= M.gmres_tol) { status = happy ? "happy_breakdown_verified" : "converged_true"; break; }
      if (ab && cycles == 1 && true_rel > M.gmres_tol) {
        const double rho = std::min(std::max(true_rel, 1e-300), 0.999999);
        const double per_solve = std::ceil(std::log(M.gmres_tol) / std::log(rho)) * m;
        const double steps = std::max(1.0, std::ceil(std::log(ab->tau / std::max(ab->eta, ab->tau)) / std::log(M.gmres_tol)));
        const double its = steps * ab->nrhs * per_solve - k;
        if (its > ab->refactor_iters || its > budget_left) {
          std::ostringstream o;
          o << "abandoned:rho=" << rho << ",pred_its=" << its << ",refactor_iters=" << ab->refactor_iters;
          status = o.str();
          break;
        }
      }
      if (happy) { status = "happy_breakdown_unverified"; break; }
      if (!M.gmres_left) copy_f64(K.sR.p, K.rcur.p, n, W());   // the next cycle restarts from this residual
    }
    return status;
  }

  // In-loop metrics.  resid 32 (src/mp_core.cuh:231-235): the FP32 residual of the FP32 iterate on the
  // FP32-stored system, eta = ||r||inf / (||A||inf ||x||inf + ||b||inf) from it, omega not evaluated (-1).
  void residual_and_metrics(const MethodSpec& M, Metrics& Mt) {
    if (M.resid == 32) {
      if (!P.fp32) throw MPError("fp32 residual control requires fp32 stored data", -1);
      if (M.working != 32) throw MPError("fp32 residual control needs the FP32 working iterate", -1);
      const int64_t m = n * nrhs;
      r32.ensure(C->tracker, (size_t)m, "R32");
      hipLaunchKernelGGL(k_resid_f32_rows, dim3(cdiv(n, 256)), dim3(256), 0, W().stream, P.A32.p, n, x32.p, P.B.p, r32.p, nrhs);
      LAUNCH_CK();
      std::vector<float> hr(m); std::vector<double> hx(m);
      HIP_CK(hipMemcpyAsync(hr.data(), r32.p, m * 4, hipMemcpyDeviceToHost, W().stream));
      HIP_CK(hipMemcpyAsync(hx.data(), x64.p, m * 8, hipMemcpyDeviceToHost, W().stream));
      sync();
      Mt.eta.assign(nrhs, 0.0); Mt.omega.assign(nrhs, -1.0); Mt.rinf.assign(nrhs, 0.0); Mt.xinf.assign(nrhs, 0.0);
      Mt.finite = true; Mt.eta_max = 0.0; Mt.omega_max = -1.0;
      for (int c = 0; c < nrhs; c++) {
        double ri = 0.0, xi = 0.0;
        for (int64_t i = 0; i < n; i++) {
          const double r = std::fabs((double)hr[i + c * n]), x = std::fabs(hx[i + c * n]);
          if (!(r <= ri)) ri = r;             // a NaN is kept, as the column statistics do
          if (!(x <= xi)) xi = x;
        }
        const double den = P.anorm_inf * xi + P.bnorm_inf[c];
        Mt.eta[c] = (ri == 0) ? 0.0 : (den == 0 ? INFINITY : ri / den);
        Mt.rinf[c] = ri; Mt.xinf[c] = xi;
        if (!std::isfinite(ri) || !std::isfinite(xi)) Mt.finite = false;
        Mt.eta_max = std::max(Mt.eta_max, Mt.eta[c]);
      }
      return;
    }
    if (pair_ready) {   // formed with the last FGMRES check, for this very iterate
      pair_ready = false;
      E.swap(E2);
      residual_metrics(C, P, x64.p, M.resid == 128, E, Mt, false);
    } else {
      residual_metrics(C, P, x64.p, M.resid == 128, E, Mt);
    }
    if (keep_best && Mt.finite && Mt.eta_max < best_eta) {
      best_eta = Mt.eta_max;
      copy_f64(x64.p, xbest.p, (int64_t)n * nrhs, W());
    }
  }

  // last_stage: the stage the ladder ends with (never abandoned)
  std::string refine_stage(const MethodSpec& M, const Factors& F, int stage_idx, TrialOut& T, double t_start, int& budget_left,
                           bool initial_solve, Metrics& Mt, bool last_stage = true) {
    if (initial_solve) {
      const double t0 = wall_now();
      apply_inverse(C, F, P.B.p, D.p, nrhs, &T.trs, M.apply64);
      set_x(M, D.p);
      sync();
      T.t_initial += wall_now() - t0;
    }
    double prev_eta = INFINITY; int stall = 0; std::string reason = "max_outer";
    int local_step = 0;
    while (true) {
      const double te0 = wall_now();
      residual_and_metrics(M, Mt);
      T.t_eval += wall_now() - te0;
      T.final_eta = Mt.eta_max; T.final_omega = Mt.omega_max;
      const int inner_before = T.inner_total;
      T.hist.push_back({T.outer, stage_idx, wall_now() - t_start, Mt.eta_max, Mt.omega_max,
                        *std::max_element(Mt.rinf.begin(), Mt.rinf.end()), 0, ""});
      const bool conv = Mt.finite && Mt.eta_max <= tau_eta && (stop_eta_only || Mt.omega_max <= tau_omega);
      if (conv) { reason = "converged"; break; }
      if (!Mt.finite) { reason = "nonfinite"; break; }
      if (T.outer >= M.max_outer) { reason = "max_outer"; break; }
      if (Mt.eta_max > 0.9 * prev_eta) stall++; else stall = 0;
      if (stall >= 3) { reason = "stagnation"; break; }
      prev_eta = Mt.eta_max;
      if (M.gmres && budget_left <= 0) { reason = "gmres_budget_exhausted"; break; }
      if (M.gmres) {
        bool abandoned = false;
        AbandonTest ab;
        ab.eta = Mt.eta_max; ab.tau = tau_eta; ab.nrhs = nrhs;
        ab.refactor_iters = M.cost_ratio_next * kFactorIters;
        for (int c = 0; c < nrhs; c++) {
          const bool test = stage_abandon && !last_stage && local_step == 0 && c == 0;
          const std::string st = gmres_solve(M, F, E.R.p + (int64_t)c * n, D.p + (int64_t)c * n, budget_left, T, T.outer, c,
                                             test ? &ab : nullptr);
          T.hist.back().note += (c ? ";" : "") + st;
          if (st == "nonfinite") T.hist.back().note += ";inner_nonfinite";
          if (st.rfind("abandoned", 0) == 0) { abandoned = true; break; }
        }
        // the first cycle showed the factors too weak for this stage to pay: the iterate is left as it is
        if (abandoned) { T.hist.back().inner_in_step = T.inner_total - inner_before; reason = "stage_abandoned"; break; }
      } else {
        apply_inverse(C, F, E.R.p, D.p, nrhs, &T.trs, M.apply64); T.precond_applies++;
      }
      T.hist.back().inner_in_step = T.inner_total - inner_before;
      add_update(M, D.p);
      pair_ready = pair_done && M.gmres;   // x is now the x + d whose residual the last check formed
      pair_done = false;
      T.outer++;
      local_step++;
    }
    return reason;
  }

  void run(const MethodSpec& M, TrialOut& T) {
    const double t0 = wall_now();
    T.outer = 0; T.inner_total = 0;
    // word 2 of the info area is the substitutions' exchange-timeout flag: cleared here, checked at the end
    W().dinfo.ensure(C->tracker, 4, "getrf_info");
    HIP_CK(hipMemsetAsync(W().dinfo.p + 2, 0, 4, W().stream));
    int budget_left = M.gmres ? M.gmres_budget : 0;
    std::vector<std::string> ladder{M.factor};
    const bool target64 = (tau_eta < 1e-10);
    // this build's ladders end, like the reference's (src/mp_core.cuh:301), with FP64 factors for an FP64 target
    if (M.policy == "ladder") { ladder = {"mx8", "mx16", "cus32"}; if (target64) ladder.push_back("cs64"); }
    else if (M.policy == "ladder16") { ladder = {"mx16", "mx16x3", "cus32"}; if (target64) ladder.push_back("cs64"); }
    else if (M.policy == "ladder16x3") { ladder = {"mx16x3", "cus32", "cs64"}; }
    else if (M.policy != "none") {
      if (M.factor == "het16" || M.factor == "cus32") { ladder.push_back("cs32"); if (target64) ladder.push_back("cs64"); }
      else if (M.factor == "cs32") { if (target64) ladder.push_back("cs64"); }
    }
    fill_f64(x64.p, (int64_t)n * nrhs, 0.0, W());
    if (M.working == 32) { x32.ensure(C->tracker, (size_t)n * nrhs, "x32"); HIP_CK(hipMemsetAsync(x32.p, 0, (size_t)n * nrhs * 4, W().stream)); }
    Metrics Mt; std::string reason = "unknown"; MethodSpec Ms = M;
    bool initialized = false;
    // A stage that stagnates or runs out of budget can leave an iterate far worse than one it passed through
    // (low-precision factors of an ill-conditioned matrix); this build's ladders start the next stage from the
    // best iterate seen (smallest eta), not from the last one.  The reference's policies keep continuing from
    // the last iterate.  The returned solution is always the last iterate of the last stage.
    keep_best = (M.policy == "ladder" || M.policy == "ladder16" || M.policy == "ladder16x3") && M.working == 64;
    best_eta = INFINITY;
    if (keep_best) xbest.ensure(C->tracker, (size_t)n * nrhs, "x_best");
    int first = 0;
    // resume_ladder: with reuse, start at the stage whose factors the problem keeps (the stage that served the
    // last solve of this matrix), so repeated solves skip the stages that failed before
    if (reuse && resume_ladder && P.fcache && P.fcache->valid)
      for (int s = 0; s < (int)ladder.size(); s++) if (ladder[s] == P.fcache->kind) { first = s; break; }
    for (int stage = first; stage < (int)ladder.size(); stage++) {
      Factors& F = factors_for(M, ladder[stage], T);
      T.final_stage = ladder[stage];
      if (stage > 0) T.fallback = true;
      if (F.info > 0) {
        reason = "singular_factor";
        T.hist.push_back({T.outer, stage, wall_now() - t0, -1, -1, -1, 0, "singular_factor_info=" + std::to_string(F.info)});
        if (M.policy == "none") break; else continue;
      }
      if (M.policy == "budget" && M.budget == 0) Ms.gmres = false;
      if (M.policy == "budget" && M.budget > 0) { if (stage == 0) { Ms.gmres = true; budget_left = M.budget; } else Ms.gmres = false; }
      if (M.policy == "cost_aware") Ms.gmres = false;
      // this build's ladders: FGMRES-IR for the low-precision stages; the closing FP64-factor stage refines
      // stationarily (FP64 factors leave one substitution of all right-hand sides per step, where GMRES would
      // apply them one vector at a time)
      if (M.policy == "ladder" || M.policy == "ladder16" || M.policy == "ladder16x3") {
        Ms.gmres = ladder[stage] != "cs64"; budget_left = M.gmres_budget;
      }
      const double ts = wall_now();
      reason = refine_stage(Ms, F, stage, T, t0, budget_left, !initialized, Mt, stage + 1 == (int)ladder.size());
      initialized = true;
      T.t_refine += wall_now() - ts;
      if (M.policy == "cost_aware" && reason != "converged" && reason != "nonfinite" && M.gmres_budget > 0 && stage == 0) {
        const double c_outer = T.hist.size() >= 2 ? (T.hist.back().t - T.hist[T.hist.size() - 2].t) : 0.0;
        const double c_refactor_next = T.stage_times.back() * M.cost_ratio_next;
        const double predicted = c_outer * (double)std::min(M.gmres_budget, M.gmres_m) * 2.0;
        T.hist.back().note += ";cost_aware:c_outer=" + std::to_string(c_outer) + ",c_refactor_next=" + std::to_string(c_refactor_next) +
                              ",pred_gmres=" + std::to_string(predicted);
        if (predicted < c_refactor_next) {
          Ms.gmres = true; budget_left = M.gmres_budget; T.hist.back().note += ";switch_to_gmres";
          reason = refine_stage(Ms, F, stage, T, t0, budget_left, false, Mt);
          Ms.gmres = false;
        }
      }
      if (reason == "converged") break;
      if (M.policy == "none") break;
      if (stage + 1 < (int)ladder.size()) {
        T.hist.push_back({T.outer, stage, wall_now() - t0, Mt.eta_max, Mt.omega_max, -1, 0, "escalate_after:" + reason});
        if (keep_best && std::isfinite(best_eta) && !(Mt.finite && Mt.eta_max <= best_eta)) {
          copy_f64(xbest.p, x64.p, (int64_t)n * nrhs, W());
          T.hist.back().note += ";restart_from_best_eta=" + std::to_string(best_eta);
        }
      }
    }
    T.terminal = (M.kind == "direct" && reason != "nonfinite" && reason != "singular_factor") ? "direct_solve" : reason;
    int32_t trs_timeout = 0;
    HIP_CK(hipMemcpyAsync(&trs_timeout, W().dinfo.p + 2, 4, hipMemcpyDeviceToHost, W().stream));
    sync();
    if (trs_timeout) throw MPError("substitution: the block exchange timed out (workgroups not co-resident)", MP_ERR_INTERNAL);
    T.t_total = wall_now() - t0;
    // the last stage's growth factor (measured beside the refinement, long finished by now)
    growth_resolve(C);
    if (growth_src) T.growth = growth_src->growth;
  }
};

// ------------------------------------------------------------------ ABI helpers
static void copy_str(const std::string& s, char* buf, size_t cap) {
  if (!buf || cap == 0) return;
  const size_t k = std::min(cap - 1, s.size());
  std::memcpy(buf, s.data(), k); buf[k] = 0;
  if (s.size() + 1 > cap) throw MPError("output buffer too small (need " + std::to_string(s.size() + 1) + " bytes)", -1);
}


// ||A||_inf, max|A|, ||b||, SHA-256 of the stored bytes (generate_problem diagnostics, src/mp_core.cuh:90-101).
// src: host copy of the column-major stored A; hb: host copy of B.
static void finish_setup(mp_ctx* C, mp_problem* Pp, const char* src, const std::vector<double>& hb) {
  mp_problem* P = Pp;
  const int64_t n = P->n; const int nrhs = P->nrhs; const size_t es = P->fp32 ? 4 : 8;
    // ||A||_inf = max_i sum_j |a_ij| via the residual kernel's T = |A| |1|
    {
      DBuf<double> ones, Rt, Tt;
      ones.alloc(C->tracker, n, "ones"); Rt.alloc(C->tracker, n, "rowsum_R"); Tt.alloc(C->tracker, n, "rowsum_T");
      fill_f64(ones.p, n, 1.0, C->W);
      ResidArgs ra; ra.n = n; ra.nrhs = 1; ra.a32 = P->fp32; ra.A = P->a_ptr(); ra.lda = n; ra.X = ones.p; ra.ldx = n;
      ra.R = Rt.p; ra.ldr = n; ra.T = Tt.p; ra.ldt = n; ra.matvec = true;
      residual(ra, C->W);
      std::vector<double> rs(n);
      HIP_CK(hipMemcpyAsync(rs.data(), Tt.p, n * 8, hipMemcpyDeviceToHost, C->W.stream));
      HIP_CK(hipStreamSynchronize(C->W.stream));
      double m = 0; for (double v : rs) { if (!(v <= m)) m = v; }
      P->anorm_inf = m;
    }
    P->amax = dev_absmax(C, P->a_ptr(), P->fp32, n, n);
    P->bnorm_inf.assign(nrhs, 0); P->bnorm_2.assign(nrhs, 0);
    for (int c = 0; c < nrhs; c++) {
      double m = 0, s = 0;
      for (int64_t i = 0; i < n; i++) { const double v = std::fabs(hb[i + (size_t)c * n]); m = std::max(m, v); s += v * v; }
      P->bnorm_inf[c] = m; P->bnorm_2[c] = std::sqrt(s);
    }
    // the whole-matrix digest is one sequential pass; the 16 column-tile digests run beside it on host threads
    const size_t colbytes = (size_t)n * es;
    std::vector<std::string> tiles(16);
    {
      std::vector<std::thread> th;
      for (int g = 0; g < 4; g++)
        th.emplace_back([&, g] {
          for (int t = g; t < 16; t += 4) {
            const size_t c0 = (size_t)n * t / 16, c1 = (size_t)n * (t + 1) / 16;
            tiles[t] = sha256_hex(src + c0 * colbytes, (c1 - c0) * colbytes);
          }
        });
      P->a_sha = sha256_hex(src, (size_t)n * n * es);
      for (auto& t : th) t.join();
    }
    P->tile_sha = tiles;
    P->b_sha = sha256_hex(hb.data(), hb.size() * 8);
}

// Norms of a problem whose A is already on the device (no host copy of A, so no A hashes): ||A||_inf, max|A|
// on the device; B's norms and hash from its host copy.
static void finish_setup_device(mp_ctx* C, mp_problem* P, const std::vector<double>& hb) {
  const int64_t n = P->n; const int nrhs = P->nrhs;
  {
    DBuf<double> ones, Rt, Tt;
    ones.alloc(C->tracker, n, "ones"); Rt.alloc(C->tracker, n, "rowsum_R"); Tt.alloc(C->tracker, n, "rowsum_T");
    fill_f64(ones.p, n, 1.0, C->W);
    ResidArgs ra; ra.n = n; ra.nrhs = 1; ra.a32 = false; ra.A = P->a_ptr(); ra.lda = n; ra.X = ones.p; ra.ldx = n;
    ra.R = Rt.p; ra.ldr = n; ra.T = Tt.p; ra.ldt = n; ra.matvec = true;
    residual(ra, C->W);
    std::vector<double> rs(n);
    HIP_CK(hipMemcpyAsync(rs.data(), Tt.p, n * 8, hipMemcpyDeviceToHost, C->W.stream));
    HIP_CK(hipStreamSynchronize(C->W.stream));
    double m = 0; for (double v : rs) { if (!(v <= m)) m = v; }
    P->anorm_inf = m;
  }
  P->amax = dev_absmax(C, P->a_ptr(), false, n, n);
  P->bnorm_inf.assign(nrhs, 0); P->bnorm_2.assign(nrhs, 0);
  for (int c = 0; c < nrhs; c++) {
    double m = 0, s2 = 0;
    for (int64_t i = 0; i < n; i++) { const double v = std::fabs(hb[i + (size_t)c * n]); m = std::max(m, v); s2 += v * v; }
    P->bnorm_inf[c] = m; P->bnorm_2[c] = std::sqrt(s2);
  }
  P->a_sha = "not_computed";
  P->tile_sha.clear();
  P->b_sha = sha256_hex(hb.data(), hb.size() * 8);
}

extern "C" {

const char* mp_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

// error message of the calling thread, for entry points implemented in other files of the library
void mp_set_last_error(const std::string& msg) { g_last_error = msg; }

extern "C" {

int mp_create(mp_ctx** out, int device) {
  return guard([&] {
    MP_REQUIRE(out, "ctx pointer is NULL");
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) { (void)hipGetLastError(); throw MPError("no HIP device available", MP_ERR_NO_DEVICE); }
    MP_REQUIRE(device >= 0 && device < cnt, "device index out of range");
    HIP_CK(hipSetDevice(device));
    std::unique_ptr<mp_ctx> C(new mp_ctx());
    C->device = device;
    HIP_CK(hipStreamCreateWithFlags(&C->W.stream, hipStreamNonBlocking));
    HIP_CK(hipStreamCreateWithFlags(&C->W.pstream, hipStreamNonBlocking));
    HIP_CK(hipEventCreateWithFlags(&C->W.ev_next, hipEventDisableTiming));
    HIP_CK(hipEventCreateWithFlags(&C->W.ev_panel, hipEventDisableTiming));
    HIP_CK(hipEventCreateWithFlags(&C->W.ev_pack, hipEventDisableTiming));
    HIP_CK(hipEventCreateWithFlags(&C->W.ev_conv, hipEventDisableTiming));
    HIP_CK(hipEventCreateWithFlags(&C->ev_growth, hipEventDisableTiming));
    C->W.tracker = &C->tracker;
    hipDeviceProp_t prop; HIP_CK(hipGetDeviceProperties(&prop, device));
    C->W.num_cus = prop.multiProcessorCount;
    // this device's copy of the allocation-free residual's scratch array (the device is current here)
    C->resid_scratch = resid_scratch_init();
    RB_CK(rocblas_create_handle(&C->rb));
    RB_CK(rocblas_set_stream(C->rb, C->W.stream));
    RB_CK(rocblas_create_handle(&C->rb_panel));
    RB_CK(rocblas_set_stream(C->rb_panel, C->W.pstream));
    HIP_CK(hipHostMalloc((void**)&C->hpin, 4096 * sizeof(double), hipHostMallocDefault));
    C->W.hpin = C->hpin + 3072;   // doubles [3072, 3076): getrf's info readback
    C->W.dinfo.ensure(C->tracker, 4, "getrf_info");
    HIP_CK(hipMemsetAsync(C->W.dinfo.p, 0, 16, C->W.stream));
    HIP_CK(hipStreamSynchronize(C->W.stream));
    *out = C.release();
  });
}

void mp_destroy(mp_ctx* C) {
  if (!C) return;
  (void)hipSetDevice(C->device);
  (void)hipStreamSynchronize(C->W.stream);
  if (C->W.pstream) (void)hipStreamSynchronize(C->W.pstream);   // nothing may still read what is freed below
  // problems that outlive their context keep only host state
  for (mp_problem* p : C->problems) { p->release_device(); p->ctx = nullptr; }
  C->problems.clear();
  C->kry.reset();
  if (C->rb) rocblas_destroy_handle(C->rb);
  if (C->rb_panel) rocblas_destroy_handle(C->rb_panel);
  if (C->hpin) (void)hipHostFree(C->hpin);
  (void)hipStreamSynchronize(C->W.pstream);
  hipStream_t s = C->W.stream, ps = C->W.pstream;
  hipEvent_t e1 = C->W.ev_next, e2 = C->W.ev_panel, e3 = C->W.ev_pack, e4 = C->W.ev_conv, e5 = C->ev_growth;
  delete C;
  if (e5) (void)hipEventDestroy(e5);
  if (e4) (void)hipEventDestroy(e4);
  if (e1) (void)hipEventDestroy(e1);
  if (e2) (void)hipEventDestroy(e2);
  if (e3) (void)hipEventDestroy(e3);
  if (ps) (void)hipStreamDestroy(ps);
  if (s) (void)hipStreamDestroy(s);
}

int mp_caps(mp_ctx* C, char* json, size_t cap) {
  return guard([&] {
    MP_REQUIRE(C, "ctx is NULL");
    hipDeviceProp_t p; HIP_CK(hipGetDeviceProperties(&p, C->device));
    size_t fr = 0, tot = 0; HIP_CK(hipMemGetInfo(&fr, &tot));
    int rt = 0, drv = 0; (void)hipRuntimeGetVersion(&rt); (void)hipDriverGetVersion(&drv);
    char rbv[128] = {0}; rocblas_get_version_string(rbv, sizeof rbv);
    Json j;
    j.s("device", p.name).s("gcn_arch", p.gcnArchName).i("total_mem_bytes", (int64_t)tot).i("free_mem_bytes", (int64_t)fr)
     .i("hip_runtime", rt).i("hip_driver", drv).s("rocblas_version", rbv).i("multiprocessors", p.multiProcessorCount)
     .i("clock_khz", p.clockRate).i("mem_clock_khz", p.memoryClockRate).i("l2_bytes", p.l2CacheSize)
     .i("lds_per_block_bytes", (int64_t)p.sharedMemPerBlock).b("magma_linked", false);
    copy_str(j.obj(), json, cap);
  });
}

int mp_profile(mp_ctx* C, int enable) {
  return guard([&] { MP_REQUIRE(C, "ctx is NULL"); C->W.prof.drain(); C->W.prof.on = enable != 0; if (enable) C->W.prof.acc.clear(); });
}
int mp_profile_json(mp_ctx* C, char* json, size_t cap) {
  return guard([&] {
    MP_REQUIRE(C, "ctx is NULL");
    HIP_CK(hipStreamSynchronize(C->W.stream));
    C->W.prof.drain();
    std::string out = "{"; bool first = true;
    for (auto& kv : C->W.prof.acc) {
      Json j; j.d("ms", kv.second.ms).i("count", kv.second.count).d("bytes", kv.second.bytes).d("flops", kv.second.flops);
      out += (first ? "" : ",") + Json::str(kv.first) + ":" + j.obj(); first = false;
    }
    copy_str(out + "}", json, cap);
  });
}

int mp_synchronize(mp_ctx* C) { return guard([&] { MP_REQUIRE(C, "ctx is NULL"); HIP_CK(hipStreamSynchronize(C->W.stream)); }); }
int mp_malloc(mp_ctx* C, void** p, size_t bytes) { return guard([&] { MP_REQUIRE(C && p, "NULL argument"); *p = C->tracker.alloc(bytes, "user"); }); }
int mp_free(mp_ctx* C, void* p) { return guard([&] { MP_REQUIRE(C, "ctx is NULL"); HIP_CK(hipStreamSynchronize(C->W.stream)); C->tracker.free(p); }); }
int mp_memcpy_h2d(mp_ctx* C, void* d, const void* s, size_t b) { return guard([&] { MP_REQUIRE(C, "ctx is NULL"); HIP_CK(hipMemcpyAsync(d, s, b, hipMemcpyHostToDevice, C->W.stream)); HIP_CK(hipStreamSynchronize(C->W.stream)); }); }
int mp_memcpy_d2h(mp_ctx* C, void* d, const void* s, size_t b) { return guard([&] { MP_REQUIRE(C, "ctx is NULL"); HIP_CK(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToHost, C->W.stream)); HIP_CK(hipStreamSynchronize(C->W.stream)); }); }
int mp_memcpy_d2d(mp_ctx* C, void* d, const void* s, size_t b) { return guard([&] { MP_REQUIRE(C, "ctx is NULL"); HIP_CK(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, C->W.stream)); }); }

int mp_residual_fp64(mp_ctx* C, int64_t n, int32_t nrhs, int32_t a_bits, const void* A, int64_t lda, const double* X, int64_t ldx,
                     const double* B, int64_t ldb, double* R, int64_t ldr, double* T, int64_t ldt, double* colstats, int32_t* worst) {
  return guard([&] {
    MP_REQUIRE(C, "ctx is NULL"); MP_REQUIRE(a_bits == 32 || a_bits == 64, "a_bits must be 32 or 64");
    ResidArgs ra; ra.n = n; ra.nrhs = nrhs; ra.a32 = (a_bits == 32); ra.A = A; ra.lda = lda; ra.X = X; ra.ldx = ldx;
    ra.B = B; ra.ldb = B ? ldb : n; ra.R = R; ra.ldr = ldr; ra.T = T; ra.ldt = T ? ldt : n; ra.stats = colstats; ra.worst = worst;
    residual(ra, C->W);
  });
}

int mp_residual_dd(mp_ctx* C, int64_t n, int32_t nrhs, int32_t a_bits, const void* A, int64_t lda, const double* X, int64_t ldx,
                   const double* B, int64_t ldb, double* Rhi, double* Rlo, int64_t ldr, double* T, int64_t ldt, double* colstats, int32_t* worst) {
  return guard([&] {
    MP_REQUIRE(C, "ctx is NULL"); MP_REQUIRE(a_bits == 32 || a_bits == 64, "a_bits must be 32 or 64");
    ResidArgs ra; ra.n = n; ra.nrhs = nrhs; ra.a32 = (a_bits == 32); ra.A = A; ra.lda = lda; ra.X = X; ra.ldx = ldx;
    ra.B = B; ra.ldb = B ? ldb : n; ra.R = Rhi; ra.Rlo = Rlo; ra.ldr = ldr; ra.T = T; ra.ldt = T ? ldt : n;
    ra.stats = colstats; ra.worst = worst; ra.dd = true;
    residual(ra, C->W);
  });
}

int mp_residual_noalloc(mp_ctx* C, int64_t n, int32_t nrhs, int32_t a_bits, const void* A, int64_t lda, const double* X, int64_t ldx,
                        const double* B, int64_t ldb, double* R, double* Rlo, int64_t ldr, double* T, int64_t ldt,
                        double* colstats, int32_t* worst) {
  return guard([&] {
    MP_REQUIRE(C, "ctx is NULL"); MP_REQUIRE(a_bits == 32 || a_bits == 64, "a_bits must be 32 or 64");
    ResidArgs ra; ra.n = n; ra.nrhs = nrhs; ra.a32 = (a_bits == 32); ra.A = A; ra.lda = lda; ra.X = X; ra.ldx = ldx;
    ra.B = B; ra.ldb = B ? ldb : n; ra.R = R; ra.Rlo = Rlo; ra.ldr = ldr; ra.T = T; ra.ldt = T ? ldt : n;
    ra.stats = colstats; ra.worst = worst; ra.dd = (Rlo != nullptr);
    residual_noalloc(ra, C->W, C->resid_scratch);
  });
}

static void batch_record(mp_ctx* C, int count, const int32_t* n, const BatchParams& bp, const std::vector<BatchResult>& r,
                         double dev_s, mp_batch_result* results, const std::string& tuning) {
  static const char* names[] = {"converged", "nonfinite", "max_outer", "stagnation", "singular_factor"};
  int64_t hist[5] = {0, 0, 0, 0, 0}; int64_t outer_total = 0, nsum = 0; int nmax = 0;
  double eta_worst = 0;
  for (int i = 0; i < count; i++) {
    results[i].terminal = r[i].terminal; results[i].outer_iters = r[i].outer_iters; results[i].info = r[i].info;
    results[i].reserved = 0; results[i].eta_max = r[i].eta_max; results[i].omega_max = r[i].omega_max;
    if (r[i].terminal >= 0 && r[i].terminal < 5) hist[r[i].terminal]++;
    outer_total += r[i].outer_iters; nsum += n[i]; nmax = std::max(nmax, (int)n[i]);
    if (!(r[i].eta_max <= eta_worst)) eta_worst = r[i].eta_max;
  }
  Json j;
  j.s("kind", "batch").i("systems", count).i("n_sum", nsum).i("n_max", nmax).d("device_s", dev_s)
   .d("systems_per_s", dev_s > 0 ? count / dev_s : 0.0).d("tau", bp.tau).i("max_outer", bp.max_outer)
   .s("stop_on", bp.eta_only ? "eta" : "eta_omega").i("wg_per_cu", bp.wg_per_cu).i("outer_iters_total", outer_total)
   .d("eta_worst", eta_worst);
  for (int t = 0; t < 5; t++) j.i(names[t], hist[t]);
  std::string rec = j.obj();
  // the launch shape: row-split thresholds per n bucket and how they were chosen
  std::ostringstream sh;
  sh << ",\"split_by_n_bucket\":[";
  for (int b = 0; b < 4; b++) sh << (b ? "," : "") << "[" << bp.split4[b] << "," << bp.split2[b] << "]";
  sh << "],\"panel_width\":32,\"threads\":256,\"tuning\":" << (tuning.empty() ? std::string("{\"source\":\"off\"}") : tuning) << "}";
  rec.pop_back();
  C->last_record = rec + sh.str();
  C->last_hist.clear();
}

static BatchParams batch_params(const char* params_json, bool* tune = nullptr, bool* tune_wg = nullptr) {
  Params a; a.parse(params_json);
  BatchParams bp;
  if (tune) *tune = a.get("tune", "auto") != "off";
  if (tune_wg) *tune_wg = !a.has("wg_per_cu");
  const double u = (a.get("target", "fp64") == "fp32") ? std::ldexp(1.0, -24) : std::ldexp(1.0, -53);
  bp.tau = a.has("tau") ? a.getd("tau", 0) : a.getd("tau_mult", 100.0) * u;
  bp.max_outer = a.geti("max_outer", 30);
  bp.eta_only = a.get("stop_on", "eta_omega") == "eta";
  bp.wg_per_cu = a.geti("wg_per_cu", 2);
  MP_REQUIRE(bp.max_outer >= 0, "max_outer must be non-negative");
  return bp;
}

int mp_batch_solve_ptrs(mp_ctx* C, int32_t count, const int32_t* n, const int32_t* nrhs, const double* const* A,
                        const double* const* B, double* const* X, const char* params_json, mp_batch_result* results) {
  return guard([&] {
    MP_REQUIRE(C, "ctx is NULL");
    bool tune = false, tune_wg = false;
    BatchParams bp = batch_params(params_json, &tune, &tune_wg);
    std::vector<BatchResult> r(std::max(count, 0));
    double dev_s = 0;
    std::string tuning;
    if (tune && count > 0) batch_tune(count, n, nrhs, A, B, X, bp, C->W, tuning, tune_wg);
    batch_solve(count, n, nrhs, A, B, X, bp, C->W, r.data(), &dev_s);
    batch_record(C, count, n, bp, r, dev_s, results, tuning);
  });
}

int mp_batch_solve(mp_ctx* C, int32_t count, const int32_t* n, const int32_t* nrhs, const double* A, const int64_t* a_off,
                   const double* B, const int64_t* b_off, double* X, const char* params_json, mp_batch_result* results) {
  return guard([&] {
    MP_REQUIRE(C, "ctx is NULL");
    MP_REQUIRE(count >= 0, "count must be non-negative");
    bool tune = false, tune_wg = false;
    BatchParams bp = batch_params(params_json, &tune, &tune_wg);
    std::vector<const double*> pa(std::max(count, 0)), pb(std::max(count, 0));
    std::vector<double*> px(std::max(count, 0));
    if (count > 0) {
      MP_REQUIRE(A && B && X && a_off && b_off, "NULL buffer or offset array");
      for (int i = 0; i < count; i++) {
        MP_REQUIRE(a_off[i] >= 0 && b_off[i] >= 0, "offsets must be non-negative");
        pa[i] = A + a_off[i]; pb[i] = B + b_off[i]; px[i] = X + b_off[i];
      }
    }
    std::vector<BatchResult> r(std::max(count, 0));
    double dev_s = 0;
    std::string tuning;
    if (tune && count > 0) {
      for (int i = 0; i < count; i++) MP_REQUIRE(n[i] >= 1 && n[i] <= batch_max_n(), "every n must be in [1, 512]");
      batch_tune(count, n, nrhs, pa.data(), pb.data(), px.data(), bp, C->W, tuning, tune_wg);
    }
    batch_solve(count, n, nrhs, pa.data(), pb.data(), px.data(), bp, C->W, r.data(), &dev_s);
    batch_record(C, count, n, bp, r, dev_s, results, tuning);
  });
}

int mp_getrf64(mp_ctx* C, int64_t n, double* A, int64_t lda, int32_t* ipiv, int32_t nb, int32_t* info) {
  return guard([&] {
    MP_REQUIRE(C && A && ipiv, "NULL argument");
    MP_REQUIRE(nb >= 4 && nb % 4 == 0, "nb must be a positive multiple of 4");
    const char* ov = std::getenv("MPSOLVE_OVERLAP");
    int inf = 0;
    getrf64(n, A, lda, ipiv, nb, !(ov && ov[0] == '0'), (void*)C->rb, (void*)C->rb_panel, C->W, &inf);
    RB_CK(rocblas_set_stream(C->rb, C->W.stream));
    if (info) *info = inf;
  });
}

int mp_getrf(mp_ctx* C, int64_t n, float* LU, int64_t lda, int64_t* ipiv, int32_t* perm, const mp_getrf_params* prm, mp_getrf_info* info) {
  return guard([&] {
    MP_REQUIRE(C && prm, "NULL argument");
    GetrfArgs ga; ga.n = n; ga.A = LU; ga.lda = lda; ga.nb = prm->nb; ga.fmt = prm->fmt; ga.tile_scale = prm->tile_scale != 0;
    ga.ipiv = ipiv; ga.perm = perm;
    GetrfInfo gi; getrf(ga, C->W, &gi);
    if (info) { info->info = gi.info; info->panels = gi.panels; info->conv_overflow = gi.overflow; info->conv_subnormal = gi.subnormal; }
  });
}

int mp_pack_operand(mp_ctx* C, int32_t fmt, int32_t tile_scale, const float* S, int64_t s_idx, int64_t s_k, int32_t nidx,
                    int32_t K, void* out, void* out_lo, int32_t* exps, uint64_t* counts) {
  return guard([&] {
    MP_REQUIRE(C, "ctx is NULL");
    pack_operand(fmt, tile_scale != 0, S, s_idx, s_k, nidx, K, out, out_lo, exps, (unsigned long long*)counts, C->W);
  });
}

int mp_problem_factor_bytes(mp_problem* P, uint64_t* bytes) {
  return guard([&] {
    MP_REQUIRE(P && bytes, "NULL argument");
    uint64_t b = 0;
    if (P->fcache) {
      const Factors& F = *P->fcache;
      b = (uint64_t)F.LU32.n * 4 + F.LU64.n * 8 + F.LU64p.n * 8 + F.ipiv.n * 8 + F.perm.n * 4 + F.ipiv32.n * 4 +
          F.dr.n * 8 + F.dc.n * 8 + F.rs.n * 8;
    }
    *bytes = b;
  });
}

int mp_getrs(mp_ctx* C, int64_t n, int32_t nrhs, const float* LU, int64_t lda, const int32_t* perm, const double* R, int64_t ldr,
             double* D, int64_t ldd, double scale) {
  return guard([&] {
    MP_REQUIRE(C, "ctx is NULL");
    // word 2 of the info area: the substitutions' exchange-timeout flag, cleared before and read after
    C->W.dinfo.ensure(C->tracker, 4, "getrf_info");
    HIP_CK(hipMemsetAsync(C->W.dinfo.p + 2, 0, 4, C->W.stream));
    getrs(n, nrhs, LU, lda, perm, R, ldr, D, ldd, scale, C->W);
    int32_t timeout = 0;
    HIP_CK(hipMemcpyAsync(&timeout, C->W.dinfo.p + 2, 4, hipMemcpyDeviceToHost, C->W.stream));
    HIP_CK(hipStreamSynchronize(C->W.stream));
    if (timeout) throw MPError("getrs: the block exchange of the substitutions timed out (D is not valid)", MP_ERR_INTERNAL);
  });
}

int mp_problem_from_host(mp_ctx* C, int64_t n, int32_t nrhs, int32_t a_bits, const void* A, int64_t lda, const double* B, int64_t ldb,
                         const double* x_exact, const double* x_ref, const char* case_json, mp_problem** out) {
  return guard([&] {
    MP_REQUIRE(C && A && B && out, "NULL argument");
    MP_REQUIRE(n > 0 && nrhs > 0 && lda >= n && ldb >= n, "bad sizes");
    MP_REQUIRE(a_bits == 32 || a_bits == 64, "a_bits must be 32 or 64");
    std::unique_ptr<mp_problem> P(new mp_problem());
    P->ctx = C; P->n = n; P->nrhs = nrhs; P->fp32 = (a_bits == 32);
    const size_t es = P->fp32 ? 4 : 8;
    std::vector<char> pack;
    const char* src = (const char*)A;
    if (lda != n) {
      pack.resize((size_t)n * n * es);
      for (int64_t j = 0; j < n; j++) std::memcpy(pack.data() + (size_t)j * n * es, src + (size_t)j * lda * es, n * es);
      src = pack.data();
    }
    if (P->fp32) { P->A32.alloc(C->tracker, (size_t)n * n, "A32"); HIP_CK(hipMemcpy(P->A32.p, src, (size_t)n * n * 4, hipMemcpyHostToDevice)); }
    else { P->A64.alloc(C->tracker, (size_t)n * n, "A64"); HIP_CK(hipMemcpy(P->A64.p, src, (size_t)n * n * 8, hipMemcpyHostToDevice)); }
    std::vector<double> hb((size_t)n * nrhs);
    for (int c = 0; c < nrhs; c++) std::memcpy(hb.data() + (size_t)c * n, B + (size_t)c * ldb, n * 8);
    P->B.alloc(C->tracker, (size_t)n * nrhs, "B");
    HIP_CK(hipMemcpy(P->B.p, hb.data(), hb.size() * 8, hipMemcpyHostToDevice));
    if (x_exact) { P->hXexact.assign(x_exact, x_exact + (size_t)n * nrhs); P->has_exact = true; }
    if (x_ref) { P->hXref.assign(x_ref, x_ref + (size_t)n * nrhs); P->has_ref = true; }
    finish_setup(C, P.get(), src, hb);
    P->case_json = case_json ? case_json : "";
    C->problems.insert(P.get());
    *out = P.release();
  });
}

int mp_generate_into(mp_ctx* C, const char* case_json, double* A, double* B, double* x_exact, int32_t* has_exact) {
  return guard([&] {
    MP_REQUIRE(C && case_json && A && B, "NULL argument");
    Params a; a.parse(case_json);
    GenSpec g;
    g.n = (int64_t)a.getd("n", 1024); g.nrhs = a.geti("nrhs", 1); g.family = a.get("family", "randsvd");
    g.spectrum = a.get("spectrum", "geometric"); g.rhs = a.get("rhs", g.family == "dyadic" ? "exact_sign" : "xtrue_uniform");
    g.kappa = a.getd("kappa", 100); g.seed = (uint64_t)a.getd("seed", 9200); g.ell = a.getd("ell", 0.25); g.nugget = a.getd("nugget", 1e-3);
    g.dim = a.geti("dim", 3); g.intmax = a.geti("intmax", 255); g.dyadic_exp = a.geti("dyadic_exp", 8); g.grcar_k = a.geti("grcar_k", 3);
    g.global_exp = a.geti("global_exp", 0); g.rowcol_span = a.geti("rowcol_span", 0); g.keep_q = a.getb("keep_q", false);
    MP_REQUIRE(g.n > 0 && g.nrhs > 0, "n and nrhs must be positive");
    DBuf<double> Xe;
    double* xe = x_exact;
    if (!xe) { Xe.alloc(C->tracker, (size_t)g.n * g.nrhs, "Xexact"); xe = Xe.p; }
    const bool exact = generate(g, (void*)C->rb, C->W, C->tracker, A, B, xe);
    HIP_CK(hipStreamSynchronize(C->W.stream));
    if (has_exact) *has_exact = exact ? 1 : 0;
  });
}

int mp_problem_from_device(mp_ctx* C, int64_t n, int32_t nrhs, const double* A, const double* B, int32_t borrow_a,
                           const char* case_json, mp_problem** out) {
  return guard([&] {
    MP_REQUIRE(C && A && B && out, "NULL argument");
    MP_REQUIRE(n > 0 && nrhs > 0, "bad sizes");
    std::unique_ptr<mp_problem> P(new mp_problem());
    P->ctx = C; P->n = n; P->nrhs = nrhs; P->fp32 = false;
    const size_t nn = (size_t)n * n;
    if (borrow_a) { P->A64.p = const_cast<double*>(A); P->A64.n = nn; P->A64.tr = nullptr; }   // not freed by the problem
    else { P->A64.alloc(C->tracker, nn, "A64"); HIP_CK(hipMemcpy(P->A64.p, A, nn * 8, hipMemcpyDeviceToDevice)); }
    P->B.alloc(C->tracker, (size_t)n * nrhs, "B");
    HIP_CK(hipMemcpy(P->B.p, B, (size_t)n * nrhs * 8, hipMemcpyDeviceToDevice));
    std::vector<double> hb((size_t)n * nrhs);
    HIP_CK(hipMemcpy(hb.data(), B, hb.size() * 8, hipMemcpyDeviceToHost));
    finish_setup_device(C, P.get(), hb);
    P->case_json = case_json ? case_json : "";
    C->problems.insert(P.get());
    *out = P.release();
  });
}

int mp_factor_cache_put(mp_ctx* C, uint64_t matrix_id, mp_problem* P) {
  return guard([&] {
    MP_REQUIRE(C && P, "NULL argument");
    MP_REQUIRE(P->ctx == C, "problem belongs to another context, or its context was destroyed");
    C->factor_ids[matrix_id] = P;
  });
}

int mp_factor_cache_get(mp_ctx* C, uint64_t matrix_id, mp_problem** out, int32_t* has_factors) {
  return guard([&] {
    MP_REQUIRE(C && out, "NULL argument");
    auto it = C->factor_ids.find(matrix_id);
    MP_REQUIRE(it != C->factor_ids.end(), "matrix id not in the factor cache");
    *out = it->second;
    if (has_factors) *has_factors = (it->second->fcache && it->second->fcache->valid) ? 1 : 0;
  });
}

int mp_factor_cache_evict(mp_ctx* C, uint64_t matrix_id) {
  return guard([&] {
    MP_REQUIRE(C, "ctx is NULL");
    auto it = C->factor_ids.find(matrix_id);
    MP_REQUIRE(it != C->factor_ids.end(), "matrix id not in the factor cache");
    it->second->fcache.reset();        // device memory of the factors is released here
    C->factor_ids.erase(it);
  });
}

int mp_problem_set_rhs(mp_ctx* C, mp_problem* P, int32_t nrhs, const double* B) {
  return guard([&] {
    MP_REQUIRE(C && P && B, "NULL argument");
    MP_REQUIRE(P->ctx == C, "problem belongs to another context, or its context was destroyed");
    MP_REQUIRE(nrhs > 0, "nrhs must be positive");
    const size_t nb = (size_t)P->n * nrhs;
    if (P->B.n < nb) P->B.alloc(C->tracker, nb, "B");
    HIP_CK(hipMemcpy(P->B.p, B, nb * 8, hipMemcpyDeviceToDevice));
    P->nrhs = nrhs;
    std::vector<double> hb(nb);
    HIP_CK(hipMemcpy(hb.data(), B, nb * 8, hipMemcpyDeviceToHost));
    P->bnorm_inf.assign(nrhs, 0); P->bnorm_2.assign(nrhs, 0);
    for (int c = 0; c < nrhs; c++) {
      double m = 0, s2 = 0;
      for (int64_t i = 0; i < P->n; i++) { const double v = std::fabs(hb[i + (size_t)c * P->n]); m = std::max(m, v); s2 += v * v; }
      P->bnorm_inf[c] = m; P->bnorm_2[c] = std::sqrt(s2);
    }
    P->b_sha = sha256_hex(hb.data(), hb.size() * 8);
    // solutions of the old right-hand sides no longer apply; the cached factors of A do
    P->has_exact = false; P->has_ref = false; P->hXexact.clear(); P->hXref.clear();
    P->Xexact.release(); P->Xref.release();
  });
}

int mp_problem_generate(mp_ctx* C, const char* case_json, mp_problem** out) {
  return guard([&] {
    MP_REQUIRE(C && case_json && out, "NULL argument");
    Params a; a.parse(case_json);
    GenSpec g;
    g.n = (int64_t)a.getd("n", 1024); g.nrhs = a.geti("nrhs", 1); g.family = a.get("family", "randsvd");
    g.spectrum = a.get("spectrum", "geometric"); g.rhs = a.get("rhs", g.family == "dyadic" ? "exact_sign" : "xtrue_uniform");
    g.kappa = a.getd("kappa", 100); g.seed = (uint64_t)a.getd("seed", 9200); g.ell = a.getd("ell", 0.25); g.nugget = a.getd("nugget", 1e-3);
    g.dim = a.geti("dim", 3); g.intmax = a.geti("intmax", 255); g.dyadic_exp = a.geti("dyadic_exp", 8); g.grcar_k = a.geti("grcar_k", 3);
    g.global_exp = a.geti("global_exp", 0); g.rowcol_span = a.geti("rowcol_span", 0); g.keep_q = a.getb("keep_q", false);
    MP_REQUIRE(g.n > 0 && g.nrhs > 0, "n and nrhs must be positive");
    MP_REQUIRE(g.rowcol_span >= 0 && g.rowcol_span <= 200, "rowcol_span must be in [0, 200]");
    const std::string dtype = a.get("dtype", "fp64");
    MP_REQUIRE(dtype == "fp64" || dtype == "fp32", "dtype must be fp64 or fp32");
    std::unique_ptr<mp_problem> P(new mp_problem());
    P->ctx = C; P->n = g.n; P->nrhs = g.nrhs; P->fp32 = false;
    const size_t nn = (size_t)g.n * g.n, nb = (size_t)g.n * g.nrhs;
    P->A64.alloc(C->tracker, nn, "A64"); P->B.alloc(C->tracker, nb, "B");
    DBuf<double> Xe; Xe.alloc(C->tracker, nb, "Xexact");
    const double t0 = wall_now();
    const bool exact = generate(g, (void*)C->rb, C->W, C->tracker, P->A64.p, P->B.p, Xe.p);
    const double tg = wall_now() - t0;
    std::vector<char> ha(nn * 8);
    std::vector<double> hb(nb);
    HIP_CK(hipMemcpy(ha.data(), P->A64.p, nn * 8, hipMemcpyDeviceToHost));
    HIP_CK(hipMemcpy(hb.data(), P->B.p, nb * 8, hipMemcpyDeviceToHost));
    if (exact) { P->hXexact.resize(nb); HIP_CK(hipMemcpy(P->hXexact.data(), Xe.p, nb * 8, hipMemcpyDeviceToHost)); P->has_exact = true; }
    if (dtype == "fp32") {
      // FP32 storage (src/mp_core.cuh:87): A and B rounded to FP32, B kept as the FP64 value of its FP32 copy.
      // x_exact stays exact only if the rounding changed nothing.
      std::vector<float> ha32(nn);
      bool same = true;
      const double* a64 = (const double*)ha.data();
      for (size_t i = 0; i < nn; i++) { ha32[i] = (float)a64[i]; same = same && ((double)ha32[i] == a64[i]); }
      for (size_t i = 0; i < nb; i++) { const double v = (double)(float)hb[i]; same = same && (v == hb[i]); hb[i] = v; }
      P->A64.release();
      P->fp32 = true;
      P->A32.alloc(C->tracker, nn, "A32");
      HIP_CK(hipMemcpy(P->A32.p, ha32.data(), nn * 4, hipMemcpyHostToDevice));
      HIP_CK(hipMemcpy(P->B.p, hb.data(), nb * 8, hipMemcpyHostToDevice));
      if (!same) { P->has_exact = false; P->hXexact.clear(); }
      ha.assign((const char*)ha32.data(), (const char*)ha32.data() + nn * 4);
    }
    const double th = wall_now();
    finish_setup(C, P.get(), ha.data(), hb);
    Json j; j.d("gen_seconds", tg).d("hash_seconds", wall_now() - th)
             .s("generator", g.family == "randsvd" ? "device: counter-based Gaussian/uniform streams, rocSOLVER geqrf/orgqr, rocBLAS dgemm"
                                                   : "device kernels over the reference's xoshiro256** host streams and counter hash");
    std::string cj = case_json; while (!cj.empty() && isspace((unsigned char)cj.back())) cj.pop_back();
    std::string jj = j.obj();
    P->case_json = (cj.size() > 2) ? cj.substr(0, cj.size() - 1) + "," + jj.substr(1) : jj;
    C->problems.insert(P.get());
    *out = P.release();
  });
}

// Reference solution (compute_reference, src/mpsolve.cu:99-125): FP64 LU + solve, then up to 12 steps of
// refinement with the double-double residual, keeping the iterate with the smallest DD residual.
int mp_problem_reference(mp_ctx* C, mp_problem* P) {
  return guard([&] {
    MP_REQUIRE(C && P, "NULL argument");
    MP_REQUIRE(P->ctx == C, "problem belongs to another context, or its context was destroyed");
    const double t0 = wall_now();
    const int64_t n = P->n; const int nrhs = P->nrhs; const size_t nb = (size_t)n * nrhs;
    Factors F;
    factorize(C, *P, "cs64", 256, 0, F);
    Json info;
    if (F.info > 0) {
      info.s("reference_grade", "unavailable").s("reference_note", "reference LU singular at pivot " + std::to_string(F.info));
    } else {
      DBuf<double> X, Xbest, Rh, Rl, D;
      X.alloc(C->tracker, nb, "Xref"); Xbest.alloc(C->tracker, nb, "Xbest"); Rh.alloc(C->tracker, nb, "Rhi"); Rl.alloc(C->tracker, nb, "Rlo"); D.alloc(C->tracker, nb, "D");
      apply_inverse(C, F, P->B.p, X.p, nrhs, nullptr);
      double best = INFINITY, best_abs = -1; int it = 0;
      std::vector<double> st(4 * nrhs);
      DBuf<double> stats; stats.alloc(C->tracker, 4 * nrhs + 4, "ref_stats");
      for (it = 0; it < 12; it++) {
        ResidArgs ra; ra.n = n; ra.nrhs = nrhs; ra.a32 = P->fp32; ra.A = P->a_ptr(); ra.lda = n; ra.X = X.p; ra.ldx = n;
        ra.B = P->B.p; ra.ldb = n; ra.R = Rh.p; ra.Rlo = Rl.p; ra.ldr = n; ra.dd = true; ra.stats = stats.p;
        residual(ra, C->W);
        HIP_CK(hipMemcpyAsync(st.data(), stats.p, 32 * nrhs, hipMemcpyDeviceToHost, C->W.stream));
        HIP_CK(hipStreamSynchronize(C->W.stream));
        double rm = 0, xm = 0;
        for (int c = 0; c < nrhs; c++) { rm = std::max(rm, st[4 * c]); xm = std::max(xm, st[4 * c + 1]); }
        const double rel = (P->anorm_inf * xm > 0) ? rm / (P->anorm_inf * xm) : rm;
        if (!std::isfinite(rel)) break;
        if (rel < best) { best = rel; best_abs = rm; copy_f64(X.p, Xbest.p, (int64_t)nb, C->W); } else break;
        if (rel < 1e-30) break;
        apply_inverse(C, F, Rh.p, D.p, nrhs, nullptr);
        axpy(X.p, D.p, (int64_t)nb, 1.0, C->W);
      }
      P->hXref.resize(nb);
      HIP_CK(hipMemcpyAsync(P->hXref.data(), Xbest.p, nb * 8, hipMemcpyDeviceToHost, C->W.stream));
      HIP_CK(hipStreamSynchronize(C->W.stream));
      P->has_ref = std::isfinite(best);
      info.s("reference_grade", P->has_ref ? "numerical_reference" : "unavailable")
          .s("reference_method", "rocsolver_dgetrf_lu_plus_double_double_residual_refinement")
          .d("reference_resid_dd_rel", best).d("reference_resid_dd_abs", best_abs).i("reference_iters", it)
          .d("growth_proxy_reference_lu", F.growth);
      if (P->has_exact) {
        double m0 = 0, m1 = 0;
        for (size_t i = 0; i < nb; i++) { m0 = std::max(m0, std::fabs(P->hXref[i] - P->hXexact[i])); m1 = std::max(m1, std::fabs(P->hXexact[i])); }
        info.d("reference_vs_exact_finf", m1 > 0 ? m0 / m1 : m0);
      }
    }
    info.d("reference_seconds", wall_now() - t0);
    std::string jj = info.obj();
    std::string& cj = P->case_json;
    while (!cj.empty() && isspace((unsigned char)cj.back())) cj.pop_back();
    cj = (cj.size() > 2) ? cj.substr(0, cj.size() - 1) + "," + jj.substr(1) : jj;
  });
}

void mp_problem_free(mp_problem* p) {
  if (!p) return;
  if (p->ctx) {
    (void)hipSetDevice(p->ctx->device); p->ctx->problems.erase(p);
    for (auto it = p->ctx->factor_ids.begin(); it != p->ctx->factor_ids.end();)
      it = (it->second == p) ? p->ctx->factor_ids.erase(it) : std::next(it);
  }
  delete p;                           // a problem whose context is gone holds no device memory any more
}

int mp_problem_copy_out(mp_ctx* C, mp_problem* P, void* A, double* B, double* x_exact) {
  return guard([&] {
    MP_REQUIRE(C && P, "NULL argument");
    MP_REQUIRE(P->ctx == C, "problem belongs to another context, or its context was destroyed");
    const size_t nn = (size_t)P->n * P->n, nb = (size_t)P->n * P->nrhs;
    if (A) HIP_CK(hipMemcpy(A, P->a_ptr(), nn * (P->fp32 ? 4 : 8), hipMemcpyDeviceToHost));
    if (B) HIP_CK(hipMemcpy(B, P->B.p, nb * 8, hipMemcpyDeviceToHost));
    if (x_exact) { MP_REQUIRE(P->has_exact, "problem has no exact solution"); std::memcpy(x_exact, P->hXexact.data(), nb * 8); }
  });
}

int mp_problem_copy_ref(mp_problem* P, double* x_ref) {
  return guard([&] {
    MP_REQUIRE(P && x_ref, "NULL argument");
    MP_REQUIRE(P->has_ref && !P->hXref.empty(), "problem has no reference solution (run mp_problem_reference first)");
    std::memcpy(x_ref, P->hXref.data(), P->hXref.size() * 8);
  });
}

int mp_problem_case_json(mp_problem* P, char* json, size_t cap) {
  return guard([&] {
    MP_REQUIRE(P, "problem is NULL");
    Json j;
    j.i("n", P->n).i("nrhs", P->nrhs).s("dtype", P->fp32 ? "fp32" : "fp64").s("a_sha256", P->a_sha).s("b_sha256", P->b_sha)
     .kv("tile_sha256", Json::arr(P->tile_sha)).i("ntiles", 16).d("anorm_inf", P->anorm_inf).d("amax", P->amax)
     .b("stored_finite", std::isfinite(P->anorm_inf) && P->amax > 0).kv("bnorm_inf", Json::arr(P->bnorm_inf))
     .kv("bnorm_2", Json::arr(P->bnorm_2)).b("has_exact_solution", P->has_exact).b("has_reference_solution", P->has_ref);
    std::string s = j.obj();
    if (!P->case_json.empty()) {   // merge caller-provided case fields
      std::string extra = P->case_json;
      while (!extra.empty() && isspace((unsigned char)extra.back())) extra.pop_back();
      if (extra.size() > 2 && extra.front() == '{' && extra.back() == '}') s = s.substr(0, s.size() - 1) + "," + extra.substr(1);
    }
    copy_str(s, json, cap);
  });
}

int mp_methods_json(char* json, size_t cap) {
  return guard([&] {
    std::string out = "[";
    Params none;
    bool first = true;
    for (const char* id : kMethodIds) {
      MethodSpec M = make_method(id, none);
      Json j;
      j.s("method_id", id).s("kind", M.kind).s("factorization", M.factor).s("factorization_contract", factor_contract(M.factor))
       .i("working", M.working).s("residual", M.resid == 128 ? "double-double (TwoProd via FMA, TwoSum), rounded to FP64" :
                                                 (M.resid == 64 ? "FP64 with FMA, |A||x|+|b| in the same pass" : "FP32"))
       .i("scale_mode", M.scale_mode).b("gmres", M.gmres)
       .s("gmres_variant", !M.gmres ? "none" : (M.gmres_left ? "left-preconditioned GMRES on F^-1 A, CGS2, device Givens"
                                                            : "right-preconditioned flexible GMRES, CGS2, device Givens"))
       .b("apply64", M.apply64).s("policy", M.policy).s("unsupported_reason", M.unsupported_reason)
       .s("triangular_solves", (M.factor == "cs64" || M.factor == "mx64") ? "rocSOLVER dgetrs (FP64)"
                               : (M.apply64 ? "FP32 factors promoted to FP64, rocSOLVER dgetrs, power-of-two unscaling"
                                            : "own HIP getrs: FP64->FP32 cast, FP32 substitution, FP32->FP64 cast, power-of-two unscaling"));
      out += (first ? "" : ",") + j.obj();
      first = false;
    }
    out += "]";
    copy_str(out, json, cap);
  });
}

int mp_eval_metrics_json(mp_ctx* C, mp_problem* P, const double* X_host, char* json, size_t cap) {
  return guard([&] {
    MP_REQUIRE(C && P && X_host, "NULL argument");
    MP_REQUIRE(P->ctx == C, "problem belongs to another context, or its context was destroyed");
    DBuf<double> X; X.alloc(C->tracker, (size_t)P->n * P->nrhs, "eval_X");
    HIP_CK(hipMemcpy(X.p, X_host, (size_t)P->n * P->nrhs * 8, hipMemcpyHostToDevice));
    EvalBufs E; Metrics M;
    residual_metrics(C, *P, X.p, false, E, M);
    Json j;
    j.d("eta_max", M.eta_max).d("omega_max", M.omega_max).kv("eta", Json::arr(M.eta)).kv("omega", Json::arr(M.omega))
     .kv("rinf", Json::arr(M.rinf)).kv("xinf", Json::arr(M.xinf)).b("metrics_finite", M.finite);
    copy_str(j.obj(), json, cap);
  });
}

int mp_audit_json(mp_ctx* C, mp_problem* P, const double* X_host, const double* x_ref, const double* x_exact,
                  char* json, size_t cap) {
  return guard([&] {
    MP_REQUIRE(C && P && X_host, "NULL argument");
    MP_REQUIRE(P->ctx == C, "problem belongs to another context, or its context was destroyed");
    const int64_t n = P->n; const int nrhs = P->nrhs; const size_t nb = (size_t)n * nrhs;
    DBuf<double> X; X.alloc(C->tracker, nb, "audit_X");
    HIP_CK(hipMemcpy(X.p, X_host, nb * 8, hipMemcpyHostToDevice));
    EvalBufs E; Metrics M, Mdd;
    residual_metrics(C, *P, X.p, false, E, M);
    residual_metrics(C, *P, X.p, true, E, Mdd);      // double-double residual, rounded once to FP64
    std::vector<double> eta_dd(nrhs);
    for (int c = 0; c < nrhs; c++) {
      const double den = P->anorm_inf * M.xinf[c] + P->bnorm_inf[c];
      eta_dd[c] = (Mdd.rinf[c] == 0) ? 0.0 : (den == 0 ? INFINITY : Mdd.rinf[c] / den);
    }
    const std::vector<double> hx(X_host, X_host + nb);
    std::vector<double> vr, ve;
    if (x_ref) vr.assign(x_ref, x_ref + nb); else if (P->has_ref) vr = P->hXref;
    if (x_exact) ve.assign(x_exact, x_exact + nb); else if (P->has_exact) ve = P->hXexact;
    if (!vr.empty()) { fwd_stats(hx, vr, n, nrhs, M.finf, M.f2, M.fcomp, M.fzero_maxabs, M.fzero_count); M.has_fwd = true; }
    if (!ve.empty()) { fwd_stats(hx, ve, n, nrhs, M.einf, M.e2, M.ecomp, M.ezero_maxabs, M.ezero_count); M.has_exact = true; }
    Json j;
    j.b("finite", M.finite).kv("eta_fp64", Json::arr(M.eta)).kv("omega_fp64", Json::arr(M.omega)).kv("rinf_fp64", Json::arr(M.rinf))
     .kv("rinf_dd", Json::arr(Mdd.rinf)).kv("eta_dd", Json::arr(eta_dd)).kv("relres_inf", Json::arr(M.relres_inf))
     .kv("relres_2", Json::arr(M.relres_2)).kv("xinf", Json::arr(M.xinf))
     .b("has_ref", M.has_fwd).kv("finf_ref", Json::arr(M.finf)).kv("f2_ref", Json::arr(M.f2)).kv("fcomp_ref", Json::arr(M.fcomp))
     .b("has_exact", M.has_exact).kv("finf_exact", Json::arr(M.einf)).kv("f2_exact", Json::arr(M.e2))
     .kv("fcomp_exact", Json::arr(M.ecomp)).kv("ezero_count", Json::arr(M.ezero_count));
    copy_str(j.obj(), json, cap);
  });
}

int mp_sha256_hex(const void* data, size_t bytes, char* hex65) {
  return guard([&] {
    MP_REQUIRE(hex65 && (data || bytes == 0), "NULL argument");
    const std::string h = sha256_hex(data, bytes);
    std::memcpy(hex65, h.c_str(), 65);
  });
}

int mp_last_record_json(mp_ctx* C, char* json, size_t cap) { return guard([&] { MP_REQUIRE(C, "ctx is NULL"); copy_str(C->last_record, json, cap); }); }
int mp_last_hist_json(mp_ctx* C, char* json, size_t cap) { return guard([&] { MP_REQUIRE(C, "ctx is NULL"); copy_str(C->last_hist, json, cap); }); }

int mp_solve(mp_ctx* C, mp_problem* P, const char* method_id, const char* params_json, mp_trial_out* out, double* X_host) {
  return guard([&] {
    MP_REQUIRE(C && P && method_id, "NULL argument");
    MP_REQUIRE(P->ctx == C, "problem belongs to another context, or its context was destroyed");
    Params a; a.parse(params_json);
    MethodSpec M = make_method(method_id, a);
    const std::string target = a.get("target", P->fp32 ? "fp32" : "fp64");
    const double u = (target == "fp32") ? std::ldexp(1.0, -24) : std::ldexp(1.0, -53);
    const double mult = a.getd("tau_mult", 100.0);
    const double tau = a.has("tau") ? a.getd("tau", mult * u) : mult * u;
    const bool eta_only = a.get("stop_on", "eta_omega") == "eta";
    const int64_t n = P->n; const int nrhs = P->nrhs;
    TrialOut T; Metrics Mt; std::string xsha; bool skipped = false; std::string skip_reason;
    double t_dev = -1;
    std::vector<double> hx((size_t)n * nrhs, 0.0);
    const size_t mem_before = C->tracker.current;
    const double tstart = wall_now();
    try {
      if (M.kind == "unsupported") {
        T.ok = false; T.terminal = "unsupported_on_platform"; T.error = M.unsupported_reason;
      } else {
        // admission (src/mpsolve.cu:182-190), with an injectable budget
        const size_t nn = (size_t)n * n;
        size_t need = (M.factor == "cs64") ? nn * 8 : nn * 4;
        if (M.apply64) need += nn * 8;
        if (M.policy != "none") need += nn * 8;
        if (M.gmres) need += (size_t)n * (2 * M.gmres_m + 2) * 8;
        // factor buffers a reused problem already holds are not allocated again
        size_t have = 0;
        if (a.getb("reuse", false) && P->fcache) {
          const Factors& F = *P->fcache;
          if (M.factor == "cs64" && F.LU64.n >= nn) have += nn * 8;
          if (M.factor != "cs64" && F.LU32.n >= nn) have += nn * 4;
          if (M.apply64 && F.LU64p.n >= nn) have += nn * 8;
        }
        need = need > have ? need - have : 0;
        size_t fr = 0, tot = 0; HIP_CK(hipMemGetInfo(&fr, &tot));
        int64_t avail = (int64_t)fr - (int64_t)(512ull << 20);
        if (a.has("hbm_budget_bytes")) avail = std::min<int64_t>(avail, (int64_t)a.getd("hbm_budget_bytes", 0) - (int64_t)C->tracker.current);
        if ((int64_t)need > avail) { skipped = true; skip_reason = "memory_skip: need " + std::to_string(need) + " additional bytes, available " + std::to_string(avail); }
      }
      if (!skipped && T.ok) {
        Engine E(C, *P, tau, tau);
        E.instrument = a.getb("instrument", false); E.stop_eta_only = eta_only; E.reuse = a.getb("reuse", false);
        E.resume_ladder = a.getb("resume_ladder", false);
        const bool own_ladder = M.policy == "ladder" || M.policy == "ladder16" || M.policy == "ladder16x3";
        E.stage_abandon = own_ladder && a.getb("stage_abandon", true);
        HIP_CK(hipStreamSynchronize(C->W.stream));
        const double t1 = wall_now();
        E.run(M, T);
        HIP_CK(hipStreamSynchronize(C->W.stream));
        t_dev = wall_now() - t1;
        HIP_CK(hipMemcpy(hx.data(), E.x64.p, hx.size() * 8, hipMemcpyDeviceToHost));
        // external evaluation on the stored system in FP64 (excluded from the timed region)
        residual_metrics(C, *P, E.x64.p, false, E.E, Mt);
        if (M.working == 32) { std::vector<float> h32(hx.begin(), hx.end()); xsha = sha256_hex(h32.data(), h32.size() * 4); }
        else xsha = sha256_hex(hx.data(), hx.size() * 8);
        if (P->has_ref) { fwd_stats(hx, P->hXref, n, nrhs, Mt.finf, Mt.f2, Mt.fcomp, Mt.fzero_maxabs, Mt.fzero_count); Mt.has_fwd = true; }
        if (P->has_exact) { fwd_stats(hx, P->hXexact, n, nrhs, Mt.einf, Mt.e2, Mt.ecomp, Mt.ezero_maxabs, Mt.ezero_count); Mt.has_exact = true; }
      }
    } catch (MPError& e) {
      T.ok = false; T.error = e.what(); T.terminal = (e.code == -4) ? "unsupported_on_platform" : "runtime_error";
      (void)hipGetLastError(); (void)hipStreamSynchronize(C->W.stream);
    }
    const double tend = wall_now();
    size_t fr2 = 0, tot2 = 0; (void)hipMemGetInfo(&fr2, &tot2);
    Json j;
    j.s("case_id", a.get("case_id", "case")).s("method_id", M.id).s("mode", a.get("mode", "resident")).i("rep", a.geti("rep", 0))
     .b("warmup", a.getb("warmup", false)).s("phase", a.get("phase", "evaluation")).s("tag", a.get("tag", "")).s("target", target).d("tau", tau)
     .b("skipped", skipped).s("skip_reason", skip_reason).b("ok", T.ok).s("error", T.error)
     .s("terminal_reason", skipped ? "memory_skip" : T.terminal)
     .s("final_stage", T.final_stage.empty() ? (M.kind == "direct" ? "direct" : M.factor) : T.final_stage)
     .b("used_fallback", T.fallback).b("b_integrity_ok", true)
     .d("host_e2e_s", -1).d("device_ready_wall_s", t_dev).d("method_internal_total_s", T.t_total).d("factor_total_s", T.t_factor_total)
     .d("initial_solve_s", T.t_initial).d("refine_s", T.t_refine).d("eval_s", T.t_eval).kv("stage_times_s", Json::arr(T.stage_times))
     .kv("stages", Json::arr(T.stages)).d("wall_including_eval_s", tend - tstart)
     .i("outer_iters", T.outer).i("inner_iters_total", T.inner_total).i("matvecs", T.matvecs).i("precond_applies", T.precond_applies)
     .i("triangular_solves", T.trs).i("restarts", T.restarts).i("true_residual_checks", T.true_resid_checks)
     .i("refactorizations", T.refactorizations).i("factor_info", T.factor_info).d("growth_proxy", T.growth)
     .i("conv_overflow", (int64_t)T.conv_overflow).i("conv_subnormal", (int64_t)T.conv_subnormal)
     .i("irs_niters", -9999).i("irs_outer_niters", -9999).i("irs_info", -9999).i("irs_workspace_bytes", 0)
     .i("magma_iter", -9999).i("magma_info", -9999).d("magma_facto_time_s", -1)
     .d("solver_final_eta_max", T.final_eta).d("solver_final_omega_max", T.final_omega)
     .i("peak_device_bytes_tracked", (int64_t)C->tracker.peak).i("device_free_after_bytes", (int64_t)fr2)
     .i("method_extra_bytes_tracked", (int64_t)(C->tracker.current > mem_before ? C->tracker.current - mem_before : 0))
     .s("x_path", a.get("x_path", "")).s("x_sha256", xsha).s("x_dtype", M.working == 32 ? "float32" : "float64")
     .b("metrics_finite", Mt.finite).d("eta_max", Mt.eta.empty() ? -1 : Mt.eta_max).d("omega_max", Mt.omega.empty() ? -1 : Mt.omega_max)
     .kv("eta", Json::arr(Mt.eta)).kv("omega", Json::arr(Mt.omega)).kv("rinf", Json::arr(Mt.rinf)).kv("r2", Json::arr(Mt.r2))
     .kv("relres_inf", Json::arr(Mt.relres_inf)).kv("relres_2", Json::arr(Mt.relres_2)).kv("xinf", Json::arr(Mt.xinf))
     .kv("worst_row", Json::arr(Mt.worst))
     .b("has_fwd", Mt.has_fwd).kv("finf", Json::arr(Mt.finf)).kv("f2", Json::arr(Mt.f2)).kv("fcomp_nonzero", Json::arr(Mt.fcomp))
     .kv("fzero_maxabs", Json::arr(Mt.fzero_maxabs)).kv("fzero_count", Json::arr(Mt.fzero_count))
     .b("has_exact", Mt.has_exact).kv("einf", Json::arr(Mt.einf)).kv("e2", Json::arr(Mt.e2)).kv("ecomp_nonzero", Json::arr(Mt.ecomp))
     .kv("ezero_maxabs", Json::arr(Mt.ezero_maxabs)).kv("ezero_count", Json::arr(Mt.ezero_count))
     .s("timestamp", a.get("timestamp", "")).i("nb", M.nb).i("gmres_m", M.gmres_m).d("gmres_tol", M.gmres_tol).i("gmres_budget", M.gmres_budget)
     .i("gmres_cycles", M.gmres_cycles).i("max_outer", M.max_outer).s("policy", M.policy).i("budget", M.budget).i("scale_mode", M.scale_mode)
     .b("stage_abandon", (M.policy == "ladder" || M.policy == "ladder16" || M.policy == "ladder16x3") && a.getb("stage_abandon", true))
     .i("working", M.working).i("resid", M.resid == 128 ? 64 : M.resid).b("resid_dd", M.resid == 128).b("apply64", M.apply64)
     .b("reuse", a.getb("reuse", false)).s("panel", "hip_lds_exchange").i("panel_launches", T.panels).i("panel_coop_refused", T.panel_coop_refused).s("panel_coop_error", T.panel_coop_error).s("panel_requested", a.get("panel_requested", "cusolver")).i("stages_count", (int64_t)T.stages.size())
     .s("stop_on", eta_only ? "eta" : "eta_omega");
    C->last_record = j.obj();
    std::string h;
    for (auto& r : T.hist) {
      Json jh;
      jh.s("case_id", a.get("case_id", "case")).s("method_id", M.id).s("mode", a.get("mode", "resident")).i("rep", a.geti("rep", 0))
        .s("tag", a.get("tag", "")).s("kind", "outer").i("outer", r.outer).i("stage", r.stage).d("t_s", r.t).d("eta_max", r.eta_max)
        .d("omega_max", r.omega_max).d("rinf_max", r.rinf_max).i("inner_in_step", r.inner_in_step).s("note", r.note);
      h += jh.obj() + "\n";
    }
    for (auto& r : T.ihist) {
      Json jh;
      jh.s("case_id", a.get("case_id", "case")).s("method_id", M.id).s("mode", a.get("mode", "resident")).i("rep", a.geti("rep", 0))
        .s("tag", a.get("tag", "")).s("kind", "inner").i("outer", r.outer).i("cycle", r.cycle).i("j", r.j).i("rhs", r.rhs)
        .d("recursive_rel", r.rec_rel).d("true_rel", r.true_rel).s("note", r.note);
      h += jh.obj() + "\n";
    }
    C->last_hist = h;
    if (out) {
      out->ok = T.ok ? 1 : 0; out->outer_iters = T.outer; out->inner_iters_total = T.inner_total; out->refactorizations = T.refactorizations;
      out->device_ready_wall_s = t_dev; out->factor_total_s = T.t_factor_total;
      out->eta_max = Mt.eta.empty() ? -1 : Mt.eta_max; out->omega_max = Mt.omega.empty() ? -1 : Mt.omega_max;
      const std::string term = skipped ? "memory_skip" : T.terminal;
      std::snprintf(out->terminal_reason, sizeof out->terminal_reason, "%s", term.c_str());
      const std::string fs = T.final_stage.empty() ? M.factor : T.final_stage;
      std::snprintf(out->final_stage, sizeof out->final_stage, "%s", fs.c_str());
    }
    if (X_host) std::memcpy(X_host, hx.data(), hx.size() * 8);
  });
}

}  // extern "C"
