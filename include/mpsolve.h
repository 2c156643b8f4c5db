/* mpsolve.h — C ABI of libmpsolve, the MI355X (gfx950) mixed-precision dense linear solver.
 *
 * The reference (aloupeep/CPSC402_Midterm_Project) exposes its hot path only as header-only C++ inside the
 * mpsolve executable; these entry points are what an FFI for that path binds.  Each declaration names the
 * reference interface it replaces.  Conventions:
 *   - every function returns MP_OK (0) or a negative mp_status; mp_last_error() gives the message of the
 *     last failure on the calling thread (replaces MPError exceptions, src/mp_common.cuh:25-28);
 *   - numerical outcomes (singular factor, non-finite values, stagnation) are solve OUTCOMES reported in
 *     mp_trial_out / the JSON record, not errors (src/mpsolve.cu:217, src/mp_core.cuh:285-314);
 *   - matrices are column-major; "device" pointers are HIP device memory of the context's GPU; kernels are
 *     enqueued on the context's stream and mp_synchronize() waits for them;
 *   - one context per host thread; the library owns the device buffers behind mp_ctx / mp_problem handles.
 */
#ifndef MPSOLVE_H
#define MPSOLVE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct mp_ctx mp_ctx;
typedef struct mp_problem mp_problem;

enum {
  MP_OK = 0,
  MP_ERR_INVALID = -1,      /* bad argument */
  MP_ERR_HIP = -2,          /* HIP runtime failure */
  MP_ERR_ALLOC = -3,        /* device allocation failed */
  MP_ERR_UNSUPPORTED = -4,  /* method / option not available on this platform */
  MP_ERR_INTERNAL = -5,     /* kernel-level failure (e.g. exchange timeout) */
  MP_ERR_NO_DEVICE = -6
};

/* trailing-update operand formats of the blocked LU */
enum { MP_FMT_FP16 = 0, MP_FMT_BF16 = 1, MP_FMT_FP8E4M3 = 2, MP_FMT_FP32 = 3,
       MP_FMT_FP16X3 = 4 /* two-term FP16 operands: hi*hi + hi*lo + lo*hi on FP16 MFMA */ };

/* ---------------------------------------------------------------- context
 * Replaces Ctx::init (src/mp_core.cuh:9-11) and the `caps` command (src/mpsolve.cu:137-143). */
int mp_create(mp_ctx** ctx, int device);
void mp_destroy(mp_ctx* ctx);
const char* mp_last_error(void);
int mp_caps(mp_ctx* ctx, char* json, size_t cap);
int mp_synchronize(mp_ctx* ctx);
/* Kernel timing with HIP events recorded on the context stream around the trailing GEMM, panel, laswp, TRSM,
 * pack, residual and getrs launches.  mp_profile(ctx, 1) clears and enables; mp_profile_json returns
 * {"kernel": {"ms", "count", "bytes", "flops"}} with the algorithmic bytes / flops of the timed launches.
 * No reference counterpart (the reference has only host stage timers, src/mp_core.cuh:219-222). */
int mp_profile(mp_ctx* ctx, int enable);
int mp_profile_json(mp_ctx* ctx, char* json, size_t cap);
/* device memory helpers for callers without their own allocator (DBuf, src/mp_common.cuh:74) */
int mp_malloc(mp_ctx* ctx, void** dptr, size_t bytes);
int mp_free(mp_ctx* ctx, void* dptr);
int mp_memcpy_h2d(mp_ctx* ctx, void* dst, const void* src, size_t bytes);
int mp_memcpy_d2h(mp_ctx* ctx, void* dst, const void* src, size_t bytes);
int mp_memcpy_d2d(mp_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---------------------------------------------------------------- standalone kernels (device pointers)
 * R = B - A X (B may be NULL: R = -A X), T = |A||X| + |B|, FP64 with FMA; A is FP64 (a_bits 64) or FP32
 * (a_bits 32).  colstats[4*c + {0,1,2,3}] = max|R|, max|X|, max |R_i|/T_i (0/0 -> 0, x/0 -> inf), sum R^2;
 * worst[c] = row of the maximal ratio (smallest index on ties).  T, colstats, worst may be NULL.
 * Replaces k_resid + dev_resid + k_colstats (src/mp_kernels.cuh:67-78, 92-96, 113-114). */
int mp_residual_fp64(mp_ctx* ctx, int64_t n, int32_t nrhs, int32_t a_bits, const void* A, int64_t lda,
                     const double* X, int64_t ldx, const double* B, int64_t ldb,
                     double* R, int64_t ldr, double* T, int64_t ldt, double* colstats, int32_t* worst);
/* Double-double residual: (Rhi + Rlo) = B - A X to ~2^-104 relative to |A||X|; Rhi is the FP64 rounding of
 * the sum and |Rlo| <= ulp(Rhi)/2.  Replaces k_resid_dd + resid_dd_norms (src/mp_kernels.cuh:83-87,
 * src/mp_core.cuh:207-208). */
int mp_residual_dd(mp_ctx* ctx, int64_t n, int32_t nrhs, int32_t a_bits, const void* A, int64_t lda,
                   const double* X, int64_t ldx, const double* B, int64_t ldb,
                   double* Rhi, double* Rlo, int64_t ldr, double* T, int64_t ldt, double* colstats, int32_t* worst);
/* The residual of mp_residual_fp64 (Rlo == NULL) or mp_residual_dd (Rlo != NULL) with the same arguments and
 * the same result bits, allocating nothing: no device or host allocation, no memset, no profiling events.
 * No counterpart in the reference; it is for callers that cannot allocate (code under HIP graph capture,
 * serving loops that must not call hipMalloc).  Its partial sums live in one fixed device array per GPU
 * (16 MiB + 2 MiB, held until the process ends; its address is taken when a context is created).  Limit: the
 * chunk partials of one right-hand side must fit 2^21 doubles (every n <= 1048576 for FP64, n <= 699050 for
 * double-double) and nrhs * ceil(n / 256) * 6 <= 2^18 (nrhs <= 682 at n = 16384); larger sizes return
 * MP_ERR_INVALID before any launch.
 * Right-hand sides are processed one at a time (A is read once per right-hand side).  Every call on one GPU
 * uses the same array: calls that overlap on one GPU (other streams, other contexts or threads) race. */
int mp_residual_noalloc(mp_ctx* ctx, int64_t n, int32_t nrhs, int32_t a_bits, const void* A, int64_t lda,
                        const double* X, int64_t ldx, const double* B, int64_t ldb, double* R, double* Rlo,
                        int64_t ldr, double* T, int64_t ldt, double* colstats, int32_t* worst);

typedef struct {
  int32_t nb;          /* panel width (1..256) */
  int32_t fmt;         /* MP_FMT_* of the trailing-update operands */
  int32_t tile_scale;  /* 1: per-32-row/column power-of-two operand scaling; 0: none */
} mp_getrf_params;
typedef struct {
  int32_t info;        /* 0, or 1-based column of the first exactly-zero pivot */
  int32_t panels;
  uint64_t conv_overflow, conv_subnormal;   /* operand conversion events */
} mp_getrf_info;
/* In-place blocked LU with partial pivoting of an n x n FP32 matrix: P A = L U, unit lower L.
 * ipiv: 1-based LAPACK pivots (int64, device); perm: row i of P A is row perm[i] of A (int32, device).
 * Replaces custom_blocked_lu (src/mp_core.cuh:129-150). */
int mp_getrf(mp_ctx* ctx, int64_t n, float* LU, int64_t lda, int64_t* ipiv, int32_t* perm,
             const mp_getrf_params* params, mp_getrf_info* info);
/* The operand conversion of the trailing update on its own (enqueued on the context stream): replaces
 * k_f32_to_f16_2d (src/mp_kernels.cuh:14-18), the round-to-nearest FP32 -> FP16 pack of L21 and U12 with its
 * overflow (|x| > 65504) and subnormal (0 < |x| < 2^-14) counters.  Source element (idx, kk) is
 * S[idx * s_idx + kk * s_k] (device FP32; one stride 1, the other at least the extent), idx < nidx, kk < K <= 256.
 * Each tile of 32 indices is scaled by 2^s (s = exps[tile], chosen from the tile's largest |x| so it lands in
 * the top binade below the format's maximum; s = 0 when tile_scale is 0 or for fmt 3) and rounded to nearest
 * even into fmt (0 FP16, 1 BF16, 2 OCP FP8-e4m3 saturating, 3 FP32, 4 two-term FP16: out = fp16(x), out_lo =
 * fp16(x - out)).  Output element e of lane l of k-block b of tile t sits at ((t * nkb + b) * 64 + l) * 8 + e
 * (element size of fmt, nkb = ceil(K / 16)) and holds idx = 32 t + l % 32 and kk = 16 b + 8 (l / 32) + e
 * (FP32: kk = 16 b + 2 e + l / 32); positions past nidx or K hold 0.  counts[0] / counts[1] (device, caller
 * zeroes) are incremented by the scaled values above the format's maximum / in its subnormal range. */
int mp_pack_operand(mp_ctx* ctx, int32_t fmt, int32_t tile_scale, const float* S, int64_t s_idx, int64_t s_k, int32_t nidx,
                    int32_t K, void* out, void* out_lo, int32_t* exps, uint64_t* counts);
/* Plain FP64 LU (P A = L U in place, A n x n FP64 device, column-major lda): the comparator the mixed-precision
 * solve is measured against (native_lu64, src/mpsolve.cu:22 and src/mp_core.cuh:162-170, cusolverDnXgetrf).
 * Blocked by nb (a multiple of 4): the vendor panel (rocSOLVER dgetrf on m x nb), rocBLAS dtrsm for U12, and
 * the trailing update on this library's FP64 MFMA kernel, with look-ahead.  ipiv: n 1-based global LAPACK pivots
 * (device); *info: first zero pivot (1-based) or 0. */
int mp_getrf64(mp_ctx* ctx, int64_t n, double* A, int64_t lda, int32_t* ipiv, int32_t nb, int32_t* info);
/* D = scale * U^{-1} L^{-1} P R with FP32 factors and FP32 arithmetic, FP64 R and D (device).
 * Replaces apply_inverse for FP32 factors (src/mp_core.cuh:182-190).  Returns MP_ERR_INTERNAL when a wait of
 * the substitutions' block exchange reached its bound (workgroups that never became resident); D is then not
 * valid.  Synchronizes the context stream. */
int mp_getrs(mp_ctx* ctx, int64_t n, int32_t nrhs, const float* LU, int64_t lda, const int32_t* perm,
             const double* R, int64_t ldr, double* D, int64_t ldd, double scale);

/* ---------------------------------------------------------------- problems
 * A (FP64 or FP32 storage, a_bits) and B (FP64) are copied to the device once; ||A||_inf, max|A|, ||b||
 * and SHA-256 of the column-major bytes are computed at setup (generate_problem, src/mp_core.cuh:51-103).
 * x_exact / x_ref (host, may be NULL) enable the forward-error metrics. */
int mp_problem_from_host(mp_ctx* ctx, int64_t n, int32_t nrhs, int32_t a_bits, const void* A, int64_t lda,
                         const double* B, int64_t ldb, const double* x_exact, const double* x_ref,
                         const char* case_json, mp_problem** out);
/* Device generation of a seeded test system from a case object (JSON): "family" (randsvd | dyadic | rbf |
 * grcar | wilkinson), "n", "nrhs", "kappa", "spectrum", "seed", "rhs" (xtrue_uniform | random | random_scaled |
 * exact_sign | unit_column), "intmax", "dyadic_exp", "ell", "nugget", "dim", "grcar_k", "global_exp",
 * "rowcol_span" (power-of-two row and column scaling), "dtype" (fp64 | fp32 storage), "keep_q" (randsvd: the
 * context keeps the two n x n orthogonal factors, 16 n^2 bytes, until a keep_q generation of another (n, seed)
 * or mp_destroy, and reuses them for systems of the same n and seed, e.g. another kappa; the matrix is the same).
 * Replaces generate_problem (src/mp_core.cuh:51-103); the random streams are this library's own. */
int mp_problem_generate(mp_ctx* ctx, const char* case_json, mp_problem** out);
/* The same seeded generation into caller-provided device buffers of the context's GPU: A (n x n), B and, if not
 * NULL, x_exact (n x nrhs); *has_exact (may be NULL) tells whether x_exact holds the exact solution.  No
 * problem object and no host copies (for building batches and request pools in HBM). */
int mp_generate_into(mp_ctx* ctx, const char* case_json, double* A, double* B, double* x_exact, int32_t* has_exact);
/* A problem over an FP64 system already in device memory (n x n A, n x nrhs B, leading dimension n).  B is
 * copied; A is copied unless borrow_a != 0, in which case the caller keeps A alive and unchanged for the
 * problem's lifetime.  Norms are computed on the device; the record's a_sha256 is "not_computed". */
int mp_problem_from_device(mp_ctx* ctx, int64_t n, int32_t nrhs, const double* A, const double* B, int32_t borrow_a,
                           const char* case_json, mp_problem** out);
/* Replaces the right-hand sides (device, n x nrhs) and drops the exact / reference solutions; factors cached by
 * solves with "reuse" stay valid (they depend on A only). */
int mp_problem_set_rhs(mp_ctx* ctx, mp_problem* p, int32_t nrhs, const double* B);
/* Factor cache by matrix id (no counterpart in the reference; a serving extension of its --reuse).  put
 * registers a problem under an id (the caller keeps ownership; freeing the problem drops its ids); solves of
 * that problem with "reuse" keep their factors in it.  get returns the problem and whether it holds factors;
 * evict releases the factors' device memory and forgets the id. */
int mp_factor_cache_put(mp_ctx* ctx, uint64_t matrix_id, mp_problem* p);
int mp_factor_cache_get(mp_ctx* ctx, uint64_t matrix_id, mp_problem** p, int32_t* has_factors);
int mp_factor_cache_evict(mp_ctx* ctx, uint64_t matrix_id);
/* FP64 LU solution refined with double-double residuals (best of <= 12 steps), used for the forward-error
 * metrics; replaces compute_reference (src/mpsolve.cu:99-125) without the kappa_2 estimate. */
int mp_problem_reference(mp_ctx* ctx, mp_problem* p);
void mp_problem_free(mp_problem* p);
/* copy the stored system back to the host (column-major; A in the stored dtype); any pointer may be NULL */
int mp_problem_copy_out(mp_ctx* ctx, mp_problem* p, void* A, double* B, double* x_exact);
/* the reference solution of mp_problem_reference (n x nrhs FP64); MP_ERR_INVALID when there is none.  The
 * reference writes the same vector to Xref_<case>.f64 (src/mpsolve.cu:168-170). */
int mp_problem_copy_ref(mp_problem* p, double* x_ref);
/* case record (JSON object: sizes, norms, hashes, case parameters passed at creation) */
int mp_problem_case_json(mp_problem* p, char* json, size_t cap);

/* ---------------------------------------------------------------- solves
 * Runs one trial of a method ID (make_method, src/mpsolve.cu:17-72, plus the MI355X IDs listed by
 * mp_methods_json) with the reference's stopping rules and terminal reasons (Engine::run,
 * src/mp_core.cuh:299-314).  params_json overrides parameters ("nb", "gmres_m", "gmres_tol",
 * "gmres_budget", "gmres_cycles", "max_outer", "tau_mult", "target", "cost_ratio_next"); NULL = defaults.
 * "reuse": true keeps the problem's factors and reuses them in later solves of the same factor kind, nb and
 * scale mode (the reference's --reuse, src/mpsolve.cu:173; its Engine lives across the trials of a case).
 * The timed region (device_ready_wall_s) starts with inputs resident and ends after the last kernel.
 * X_host (n x nrhs FP64, may be NULL) receives the solution.  The full trial record (the reference's
 * field names, src/mpsolve.cu:221-229) is returned by mp_last_record_json, the histories by
 * mp_last_hist_json. */
typedef struct {
  int32_t ok;                 /* 0 on runtime_error / unsupported */
  int32_t outer_iters, inner_iters_total, refactorizations;
  double device_ready_wall_s, factor_total_s, eta_max, omega_max;
  char terminal_reason[48];
  char final_stage[48];
} mp_trial_out;
int mp_solve(mp_ctx* ctx, mp_problem* p, const char* method_id, const char* params_json,
             mp_trial_out* out, double* X_host);
/* Device bytes held by the factors a problem keeps for "reuse" solves (0 when it keeps none).  The reference
 * tracks every device buffer in one global tracker (g_tracker, src/mp_common.cuh:71) and subtracts reusable
 * factor buffers in its admission estimate (src/mpsolve.cu:182-187); a factor cache of many problems needs
 * the per-problem figure. */
int mp_problem_factor_bytes(mp_problem* p, uint64_t* bytes);
int mp_last_record_json(mp_ctx* ctx, char* json, size_t cap);
int mp_last_hist_json(mp_ctx* ctx, char* jsonl, size_t cap);
/* precision contracts of every method ID (methods_catalog.py:50-100), as a JSON array */
int mp_methods_json(char* json, size_t cap);

/* ---------------------------------------------------------------- batches of small systems
 * Solves `count` independent systems A_i X_i = B_i (n_i <= 512, nrhs_i <= 16) in one persistent launch: each
 * system is factorized in FP32 with partial pivoting (the reference's custom_blocked_lu with nb = 32 and the FP32
 * trailing update of its cublasSgemm path, src/mp_core.cuh:129-150) and refined by the stationary loop of
 * Engine::refine_stage (src/mp_core.cuh:281-296: FP64 residual, FP32 correction, the same stopping tests and
 * terminal reasons).  The reference has no batched entry point; this replaces running its per-system path once
 * per system.  A, B, X are device buffers of the context's GPU: system i reads A + a_off[i] (n_i x n_i,
 * column-major, lda n_i) and B + b_off[i], and writes X + b_off[i] (n_i x nrhs_i, ld n_i).  n, nrhs, a_off,
 * b_off and results are host arrays of `count` entries.  params_json: "tau" or "tau_mult" (x u of "target",
 * default 100), "max_outer" (30), "stop_on" ("eta_omega" | "eta"), "wg_per_cu" (resident workgroups per CU, 2).
 * mp_last_record_json then returns the batch record (device seconds, systems per second, terminal counts). */
typedef struct {
  int32_t terminal;       /* 0 converged, 1 nonfinite, 2 max_outer, 3 stagnation, 4 singular_factor */
  int32_t outer_iters;    /* refinement steps taken */
  int32_t info;           /* first zero pivot of the FP32 factorization (1-based column), 0 if none */
  int32_t reserved;
  double eta_max, omega_max;   /* backward errors of the returned X over its right-hand sides (-1 if singular) */
} mp_batch_result;
int mp_batch_solve(mp_ctx* ctx, int32_t count, const int32_t* n, const int32_t* nrhs, const double* A, const int64_t* a_off,
                   const double* B, const int64_t* b_off, double* X, const char* params_json, mp_batch_result* results);
/* The same with one device pointer per system (host arrays A[i], B[i], X[i]), for systems that live in
 * separate allocations. */
int mp_batch_solve_ptrs(mp_ctx* ctx, int32_t count, const int32_t* n, const int32_t* nrhs, const double* const* A,
                        const double* const* B, double* const* X, const char* params_json, mp_batch_result* results);

/* ---------------------------------------------------------------- serving loop
 * Continuous-batching request loop on one GPU (no counterpart in the reference, which solves one system per
 * process, src/mpsolve.cu:130-240).  Requests name a matrix id and FP64 A (n x n) and B (n x nrhs) already in
 * device memory (leading dimension n, alive and unchanged until the request completes; A for as long as its id
 * may be repeated), and a host buffer that receives X (n x nrhs).  Systems with n <= 512 and nrhs <= 16 are
 * grouped into persistent batched launches (mp_batch_solve) while larger ones are solved one at a time by
 * "method" on a second context, so both kinds overlap on the device; a repeated matrix id reuses its cached
 * factors.  Every request is refined to eta <= "tau" (default 1e-15).  Admission keeps the bytes the server
 * holds within "hbm_budget_bytes" (default 200 GiB), releasing the least recently used cached factors first.
 * Large requests run smallest n first ("schedule": "shortest_first", or "fifo"); one that has waited longer than
 * "max_wait_s" (default 0.6) goes next.
 * params_json: "method", "tau", "hbm_budget_bytes", "max_batch", "solve" {...} and "batch" {...} (extra
 * parameters of mp_solve / mp_batch_solve).  Latencies use the clock of mp_server_clock (seconds). */
typedef struct mp_server mp_server;
typedef struct {
  uint64_t request_id;
  int32_t terminal;       /* 0 converged, 1 nonfinite, 2 max_outer, 3 stagnation, 4 singular_factor, 5 error, 6 other */
  int32_t path;           /* 0 batched small system, 1 large system */
  int32_t cache_hit;      /* 1: factors reused from an earlier request with the same matrix id */
  int32_t outer_iters;
  double eta_max, omega_max;
  double queue_s, service_s, latency_s;   /* arrival -> start, start -> done, arrival -> done */
} mp_server_completion;
double mp_server_clock(void);
int mp_server_create(int device, const char* params_json, mp_server** out);
/* arrival_s: the request's arrival time on mp_server_clock (negative: now) */
int mp_server_submit(mp_server* s, uint64_t matrix_id, int64_t n, int32_t nrhs, const double* A, const double* B,
                     double* X_host, double arrival_s, uint64_t* request_id);
/* up to max completions; waits up to wait_s seconds when none is ready */
int mp_server_poll(mp_server* s, mp_server_completion* out, int32_t max, double wait_s, int32_t* count);
int mp_server_stats_json(mp_server* s, char* json, size_t cap);
/* finishes the queued requests, then releases the cached factors and the server's contexts */
int mp_server_destroy(mp_server* s);

/* FP64 metrics of a host solution on the stored system (eval_metrics, src/mp_core.cuh:193-205):
 * eta/omega per RHS and their maxima, written as a JSON object. */
int mp_eval_metrics_json(mp_ctx* ctx, mp_problem* p, const double* X_host, char* json, size_t cap);

/* Independent evaluation of a saved solution (the per-solution record of mpaudit, src/mpaudit.cu:28-34):
 * FP64 eta/omega/rinf, double-double rinf and eta_dd, relres, and forward errors against x_ref / x_exact
 * (host n x nrhs; NULL: the problem's own reference / exact solution, if it has one). */
int mp_audit_json(mp_ctx* ctx, mp_problem* p, const double* X_host, const double* x_ref, const double* x_exact,
                  char* json, size_t cap);

/* SHA-256 of host bytes as 64 hex digits plus NUL (the x_sha256 / a_sha256 of the records,
 * src/mp_common.cuh:45-52). */
int mp_sha256_hex(const void* data, size_t bytes, char* hex65);

#ifdef __cplusplus
}
#endif
#endif /* MPSOLVE_H */
