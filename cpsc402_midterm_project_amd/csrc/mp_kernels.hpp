// mp_kernels.hpp — host-side interfaces of the HIP kernels in libmpsolve.
#pragma once
#include "mp_internal.hpp"

namespace mpk {

// Scratch owned by a context: one HIP stream, grow-only buffers, the tag counter of the exchange protocol.
// Optional per-kernel timing: when enabled, selected launches are bracketed by HIP events recorded on the
// context stream; durations are accumulated per kernel name when the events are drained.
struct KernelProfiler {
  bool on = false;
  struct Pending { std::string name; hipEvent_t a, b; double bytes, flops; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  struct Acc { double ms = 0, bytes = 0, flops = 0; long count = 0; };
  std::map<std::string, Acc> acc;
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e; HIP_CK(hipEventCreate(&e)); return e;
  }
  void drain() {
    for (auto& p : pending) {
      HIP_CK(hipEventSynchronize(p.b));
      float ms = 0; HIP_CK(hipEventElapsedTime(&ms, p.a, p.b));
      Acc& a = acc[p.name]; a.ms += ms; a.bytes += p.bytes; a.flops += p.flops; a.count++;
      pool.push_back(p.a); pool.push_back(p.b);
    }
    pending.clear();
  }
};

// Many independent systems (n <= 512, nrhs <= 16) in one persistent launch (batch_solve below).
// split4 / split2 per n bucket (n - 1) / 128: trailing sizes at or below which a system's update splits its rows
// over 4 / 2 thread groups
struct BatchParams {
  double tau = 100 * 1.1102230246251565e-16; int max_outer = 30; bool eta_only = false; int wg_per_cu = 2;
  int split4[4] = {64, 64, 64, 64}, split2[4] = {128, 128, 128, 128};
};

struct Workspace {
  hipStream_t stream = nullptr;
  hipStream_t pstream = nullptr;          // second stream: panel factorizations overlapped with trailing updates
  hipEvent_t ev_next = nullptr, ev_panel = nullptr, ev_pack = nullptr;   // cross-stream ordering (timing disabled)
  hipEvent_t ev_conv = nullptr;   // the factor copy's columns past the first panel converted (second stream)
  KernelProfiler prof;
  DevTracker* tracker = nullptr;
  void* hpin = nullptr;   // 32 pinned host bytes for the factorization's info readback (nullptr: pageable copies)
  int num_cus = 256;
  int coop_refused = 0;   // cooperative panel launches refused by the runtime in the current factorization
  std::string coop_error;  // the runtime's code for the first refusal
  DBuf<double> part;     // residual / absmax partial sums
  DBuf<double> statp;    // residual metric partials
  DBuf<double> gpart;    // partials of the factors' maximum, measured on the second stream
  // getrf
  DBuf<uint64_t> xbuf;   // tagged-word exchange area of the panel kernel (zeroed at allocation)
  DBuf<uint8_t> packL, packU;   // trailing-update operands in MFMA fragment order
  DBuf<uint8_t> packLlo, packUlo; // low parts of the two-term FP16 operands
  DBuf<int32_t> scaleL, scaleU; // per-32-row / per-32-column power-of-two exponents
  // second set of the packed L operand: steps alternate between the two, so the next step's L21 can be packed
  // on the panel stream while this step's remaining trailing update still reads its own
  DBuf<uint8_t> packL2, packLlo2; DBuf<int32_t> scaleL2;
  DBuf<unsigned long long> cnt; // [0] overflow, [1] subnormal counts of the operand conversion
  DBuf<int32_t> dinfo;   // [0] info, [1] panel exchange timeout flag, [2] substitution exchange timeout flag
  DBuf<int32_t> affect;  // per-panel swap plan (src,dst pairs)
  DBuf<int32_t> psnap;   // per step: inverse of the row permutation after that step (deferred L interchanges)
  DBuf<int32_t> parrive; // per panel: workgroups started
  DBuf<float> linv;      // inverses of the 32x32 diagonal blocks of L11
  // getrs
  DBuf<uint64_t> ybuf;   // tagged partial-solution words of the triangular solves
  DBuf<float> rhs32;
  DBuf<float> tinv;      // inverses of the 128 x 128 diagonal blocks of L and U
  DBuf<float> tchain;    // chain matrices Linv_i L_{i,i-1} and Uinv_i U_{i,i+1} of the blocked substitution
  DBuf<float> tinv3, tchain3;   // the same for 256-row blocks
  bool gs_force = false;     // forced exchange timeout of the substitutions (tests only)
  std::string gs_fail_env; long gs_fail_calls = 0;
  const float* tinv_lu = nullptr; int64_t tinv_n = 0, tinv_lda = 0; uint64_t tinv_gen = 0; bool tinv_v3 = false;
  // batched small systems
  DBuf<uint8_t> bsys;    // per-system descriptors in processing order
  DBuf<float> bwork;     // per-system factors (row-major origin order + column-major LAPACK order)
  DBuf<double> bresid;   // per-system residuals
  DBuf<uint8_t> bres;    // per-system outcomes
  DBuf<int32_t> bqueue;  // work-queue head of the persistent launch
  std::map<std::string, BatchParams> batch_tuned;   // autotuned launch shapes by batch signature
  // randsvd Q factors kept for later generations of the same (n, seed) (case option keep_q)
  DBuf<double> genU, genV; int64_t gen_n = 0; uint64_t gen_seed = 0; bool gen_valid = false;
  // FP64 LU
  DBuf<int32_t> getrf64_info;   // per-panel info of the vendor panel factorizations
  uint64_t factor_gen = 1;   // incremented by every factorization written through getrf
  uint32_t tag = 1;      // next free exchange tag (never 0)
  // bracket a launch with profiling events (no-op when profiling is off)
  hipEvent_t prof_begin(hipStream_t s = nullptr) {
    if (!prof.on) return nullptr; hipEvent_t e = prof.get(); HIP_CK(hipEventRecord(e, s ? s : stream)); return e;
  }
  void prof_end(hipEvent_t a, const char* name, double bytes, double flops, hipStream_t s = nullptr) {
    if (!prof.on || !a) return;
    hipEvent_t b = prof.get(); HIP_CK(hipEventRecord(b, s ? s : stream));
    prof.pending.push_back({name, a, b, bytes, flops});
    if (prof.pending.size() > 4096) prof.drain();
  }
  uint32_t take_tags(uint32_t count) {
    if ((uint64_t)tag + count + 1 >= 0xFFFFFFF0ull) tag = 1;
    uint32_t t = tag; tag += count + 1; return t;
  }
};

struct ResidArgs {
  int64_t n = 0; int nrhs = 1; bool a32 = false; const void* A = nullptr; int64_t lda = 0;
  const double* X = nullptr; int64_t ldx = 0; const double* B = nullptr; int64_t ldb = 0;
  double* R = nullptr; double* Rlo = nullptr; int64_t ldr = 0; double* T = nullptr; int64_t ldt = 0;
  double* stats = nullptr; int32_t* worst = nullptr; bool dd = false; bool matvec = false;
  const int* active = nullptr;   // device word: the kernels return at once when it is 0 (nullptr: always run)
};
// Two one-column residuals of the same n x n A (leading dimension n) in one pass over A: R1 = B1 - A X1 in FP64
// (no statistics) and R2 = B2 - A X2 in FP64 or double-double with T2 and the statistics, each the same bits as
// residual() gives for it alone.
struct ResidPairArgs {
  int64_t n = 0; bool a32 = false; const void* A = nullptr;
  const double* X1 = nullptr; const double* B1 = nullptr; double* R1 = nullptr;
  const double* X2 = nullptr; const double* B2 = nullptr; double* R2 = nullptr; double* R2lo = nullptr; double* T2 = nullptr;
  double* stats2 = nullptr; int32_t* worst2 = nullptr; bool dd2 = false;
};
void residual_pair(const ResidPairArgs& a, Workspace& W);
void residual(const ResidArgs& a, Workspace& W);
// Allocation-free residual: partials in a fixed device array of RESID_SCRATCH_PART + RESID_SCRATCH_STAT doubles
// (16 MiB + 2 MiB), shared by every call on one GPU.  resid_scratch_init (with the device current) returns that
// array's device address; resid_noalloc_fits says whether a size fits it.
constexpr int64_t RESID_SCRATCH_PART = 2LL << 20;
constexpr int64_t RESID_SCRATCH_STAT = 256LL << 10;
double* resid_scratch_init();
bool resid_noalloc_fits(int64_t n, int nrhs, bool dd);
void residual_noalloc(const ResidArgs& a, Workspace& W, double* scratch_dev);
void absmax(const void* A, bool a32, int64_t n, int64_t lda, double* out_dev, Workspace& W);
// the same on another stream, with its own partials buffer (at least absmax_partials(n) doubles)
void absmax_on(const void* A, bool a32, int64_t n, int64_t lda, double* out_dev, double* part, hipStream_t stream);
int64_t absmax_partials(int64_t n);

// FMT_FP16X3: operands split into FP16 hi + lo parts (same tile scale), update = hi*hi + hi*lo + lo*hi on
// FP16 MFMA with FP32 accumulation.
// LDS flag words that kernels poll: the explicit address space keeps volatile accesses on the LDS path
// (ds_read / ds_write).  Through a generic pointer they become FLAT operations, which count against the
// vector-memory counter and make every flag read wait for the wave's outstanding global loads and stores.
typedef __attribute__((address_space(3))) int lds_i32;
#define LDS_FLAG(p) ((volatile lds_i32*)(p))

enum TrailFmt { FMT_FP16 = 0, FMT_BF16 = 1, FMT_FP8 = 2, FMT_FP32 = 3, FMT_FP16X3 = 4 };

struct GetrfArgs {
  int64_t n = 0; float* A = nullptr; int64_t lda = 0; int nb = 256; int fmt = FMT_FP16; bool tile_scale = true;
  int64_t* ipiv = nullptr;       // 1-based global LAPACK pivots (device)
  int32_t* perm = nullptr;       // composed row permutation: row i of P*A is row perm[i] of A (device)
  // when set, columns [nb, n) of A are still being written on another stream and this event marks their end:
  // the first panel (columns [0, nb)) starts at once and everything else waits for the event
  hipEvent_t cols_ready = nullptr;
};
// coop_refused: panels whose cooperative launch the runtime refused (another process on the device), which then
// ran as ordinary launches with bounded exchange waits
struct GetrfInfo { int info = 0; int timeout = 0; unsigned long long overflow = 0, subnormal = 0; int panels = 0; int coop_refused = 0;
                   std::string coop_error; };
void getrf(const GetrfArgs& a, Workspace& W, GetrfInfo* out);
void pack_operand(int fmt, bool tile_scale, const float* S, int64_t sI, int64_t sK, int nidx, int K, void* out,
                  void* out_lo, int32_t* exps, unsigned long long* counts, Workspace& W);

// d = (LU)^{-1} P r  with FP32 factors and FP32 arithmetic, FP64 input/output; d is scaled by `scale`.
// reuse_inverse: the diagonal-block inverses computed for the same LU buffer since the last getrf are reused.
// active: device word; the substitutions return at once when it is 0 (nullptr: always run).
// product: when given (one right-hand side, 256-row blocks), also w = A d in FP64 (the residual kernels' product
// mode, the same bits), formed by extra workgroups of the upper solve as its blocks appear; returns whether it was.
// partials_only: the product's column-chunk partials are left in part (nch chunks) for the caller to reduce.
// scale_src: the right-hand side is first formed as scale_src / scale_div[0] and stored into R (which must then be
// writable), by the lower solve's threads as they read it.
struct ProductReq { const void* A = nullptr; bool a32 = false; double* w = nullptr; bool partials_only = false;
                    const double* part = nullptr; int64_t nch = 0;
                    const double* scale_src = nullptr; const double* scale_div = nullptr; };
bool getrs(int64_t n, int nrhs, const float* LU, int64_t lda, const int32_t* perm,
           const double* R, int64_t ldr, double* D, int64_t ldd, double scale, Workspace& W, bool reuse_inverse = false,
           const int* active = nullptr, ProductReq* product = nullptr);
// the residual kernels' column split of an n x n product (chunks of cpc columns) and the reduction of its
// partials into w = A x (product mode)
void product_chunks(int64_t n, int64_t& cpc, int64_t& nch);
void product_reduce(const double* P, int64_t n, int64_t nch, const double* X, double* w, Workspace& W, const int* act);

// Many independent systems (n <= 512, nrhs <= 16) in one persistent launch: FP32 LU (nb = 32, FP32 trailing
// update) and stationary refinement with FP64 residuals until eta <= tau (and omega <= tau unless eta_only).
struct BatchResult { int32_t terminal = 0, outer_iters = 0, info = 0; double eta_max = 0, omega_max = 0; };
int batch_max_n();
int batch_max_rhs();
// System i: A[i] (n x n, lda n), B[i] and X[i] (n x nrhs, ld n), device pointers in host arrays.
void batch_solve(int count, const int32_t* n, const int32_t* nrhs, const double* const* A, const double* const* B,
                 double* const* X, const BatchParams& p, Workspace& W, BatchResult* out, double* device_s);
// Picks the launch shape of a batch (workgroups per CU if tune_wg, row-split thresholds per n bucket) by timing
// candidates on a sample of it; cached per batch signature in the workspace.  detail: JSON of the choice.
void batch_tune(int count, const int32_t* n, const int32_t* nrhs, const double* const* A, const double* const* B,
                double* const* X, BatchParams& p, Workspace& W, std::string& detail, bool tune_wg);

// FP64 blocked LU (P A = L U in place, 1-based global ipiv on the device) with the trailing update on FP64 MFMA;
// rb_main / rb_panel: rocBLAS handles of the context (the second one serves the look-ahead panel stream).
void getrf64(int64_t n, double* A, int64_t lda, int32_t* ipiv, int nb, bool lookahead, void* rb_main, void* rb_panel,
             Workspace& W, int* info);
// C (m2 x n2) -= L (m2 x kb) * U (kb x n2), FP64, column-major, on FP64 MFMA; kb % 4 == 0.
void dgemm_update(int64_t m2, int64_t n2, int kb, const double* L, int64_t ldl, const double* U, int64_t ldu,
                  double* C, int64_t ldc, hipStream_t stream);

struct GenSpec {
  int64_t n = 1024; int nrhs = 1; std::string family = "randsvd", spectrum = "geometric", rhs = "xtrue_uniform";
  double kappa = 100; uint64_t seed = 9200; double ell = 0.25, nugget = 1e-3; int dim = 3, intmax = 255, dyadic_exp = 8, grcar_k = 3;
  int global_exp = 0, rowcol_span = 0;
  bool keep_q = false;   // randsvd: keep U and V in the context for later generations with the same n and seed
};
// device generation of A (n x n), B and X_exact (n x nrhs); returns whether X_exact holds the exact solution
bool generate(const GenSpec& g, void* rocblas_handle, Workspace& W, DevTracker& tr, double* A, double* B, double* Xe);
void axpy_scale_inplace(double* x, int64_t n, double s, Workspace& W);
double spectrum_value(const std::string& sp, int64_t i, int64_t n, double kappa);

void scale_convert_f64_to_f32(const double* A, int64_t n, int64_t lda, float* out, int64_t ldo, double s, Workspace& W);
// the same for columns [0, ncols) of an n-row block, on the given stream
void scale_convert_cols(const double* A, int64_t n, int64_t ncols, int64_t lda, float* out, int64_t ldo, double s,
                        hipStream_t stream);

}  // namespace mpk
