// k_getrf.hip — blocked right-looking LU with partial pivoting on FP32 storage, low-precision trailing update.
//
// Follows the algorithm of the reference's custom_blocked_lu (src/mp_core.cuh:129-150): per panel of kb
// columns: panel LU with partial pivoting (FP32), row interchanges applied to the columns outside the
// panel, U12 <- L11^{-1} U12 (FP32), L21/U12 rounded to the trailing-update format, A22 -= L21*U12 with
// FP32 accumulation.  What is MI355X-specific:
//   * k_panel: one persistent launch per panel.  Each workgroup keeps its slice of rows of the m x kb panel
//     resident in LDS; the per-column argmax is agreed on by exchanging tagged 64-bit words (32-bit payload +
//     32-bit step tag) so a reader needs no separate flag and no fence: the tag says the payload is current.
//     Pivot rule: largest |a|, ties to the smallest row index (LAPACK i?amax); a NaN counts as +inf.
//   * k_laswp: the panel emits the swap plan (destination row <- source row) for the rows it moved; the
//     interchange of the other columns is a gather/scatter over that plan, one column per workgroup pass.
//   * k_trsm: U12 column blocks live in LDS; the unit-lower solve is blocked by 32 with the update loop
//     reading L11 through the scalar cache.  Its summation order equals plain forward substitution.
//   * k_pack: L21 and U12 are rounded (RNE) into FP16 / BF16 / OCP FP8-e4m3 (or kept FP32) after scaling
//     each 32-row tile of L21 and each 32-column tile of U12 by its own power of two, so the tile maximum
//     lands just under the top of the format's range; the values are stored in MFMA fragment order so
//     every lane's operand load is one contiguous 8/16/32-byte vector.
//   * k_gemm: v_mfma_f32_32x32x16_{f16,bf16,fp8_fp8} (or 32x32x2_f32) computes C^T tiles, so each FP32 C
//     load/store of the read-modify-write is 128 contiguous bytes per 32 lanes.  The exact power-of-two
//     tile scales are undone in the epilogue.
#include "mp_internal.hpp"
#include "mp_kernels.hpp"

namespace mpk {

constexpr int MAX_KB = 512;

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef float f32x8_t __attribute__((ext_vector_type(8)));

// Phase markers of the panel kernel: empty in the library build; a probe build may define them to record
// per-step timestamps.
#ifndef PANEL_MARK
#define PANEL_MARK(phase)
#endif
#ifndef PANEL2_MARK
#define PANEL2_MARK(col, phase)
#endif
#ifndef PANEL2_COUNT
#define PANEL2_COUNT(col, value)
#endif
#ifndef PANEL2_WGMARK
#define PANEL2_WGMARK(slot)
#endif
#ifndef PANEL2_MARKB
#define PANEL2_MARKB(col, phase)
#endif
#ifndef NP_MARK
#define NP_MARK(phase)
#endif
constexpr long SPIN_MAX = 1L << 22;

// ------------------------------------------------------------------ tagged-word exchange
__device__ __forceinline__ void put_word(uint64_t* p, uint32_t tag, uint32_t payload) {
  __hip_atomic_store(p, ((uint64_t)tag << 32) | (uint64_t)payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t get_word(const uint64_t* p, uint32_t tag, volatile lds_i32* s_abort) {
  for (long it = 0; it < SPIN_MAX; it++) {
    const uint64_t w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(w >> 32) == tag) return (uint32_t)w;
    if (*s_abort) return 0u;
  }
  *s_abort = 1;
  return 0u;
}

__device__ __forceinline__ void amax_combine(float& v, int& i, float v2, int i2) {
  const bool take = (i2 >= 0) && (i < 0 || v2 > v || (v2 == v && i2 < i));
  if (take) { v = v2; i = i2; }
}
__device__ __forceinline__ void amax_wave(float& v, int& i) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float v2 = __shfl_xor(v, off, 64);
    const int i2 = __shfl_xor(i, off, 64);
    amax_combine(v, i, v2, i2);
  }
}

struct PanelArgs {
  float* A; int64_t lda; int64_t k; int kb; int m; int G; int R;
  uint64_t* xbuf; int64_t* ipiv; int32_t* dinfo;
  int32_t* plan;   // [0] = count, then (dst, src) panel-local row pairs
};

// Head word of a workgroup's pivot candidate: [63:40] 24-bit tag, [39:32] local row, [31:0] |value| bits
// (negative = no candidate).  Row / identity words: [63:32] 32-bit tag, [31:0] payload.
// Tags are the global column index + 1: unique within one factorization; the exchange area is cleared at
// the start of every factorization, so a stale word can never carry the tag being waited for.
__device__ __forceinline__ uint64_t get_head(const uint64_t* p, uint32_t tag24, volatile lds_i32* s_abort) {
  for (long it = 0; it < SPIN_MAX; it++) {
    const uint64_t w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(w >> 40) == tag24) return w;
    if (*s_abort) return 0ull;
    __builtin_amdgcn_s_sleep(1);
  }
  *s_abort = 1;
  return 0ull;
}

// Wave-wide maximum with DPP (row_shr 1/2/4/8 inside rows of 16 lanes, then row_bcast 15/31 across rows):
// a handful of VALU ops instead of a chain of LDS-routed shuffles.  Inputs are never NaN here.
// Inputs are non-negative floats or the marker -1: on those, the order of the bit patterns as signed integers
// is the order of the values, so an integer max (one DPP-fused instruction per stage, no canonicalization).
__device__ __forceinline__ float wave_max_dpp(float v) {
  int x = __float_as_int(v);
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x111, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x112, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x114, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x118, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x142, 0xa, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x143, 0xc, 0xf, false));
  return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}
// argmax over the wave: largest value, ties to the lowest lane (callers lay indices out increasing with the
// lane, so that is the smallest index).  Invalid lanes carry v = -1.  Returns (value, index) in every lane.
__device__ __forceinline__ void wave_argmax(float& v, int& i) {
  const float m = wave_max_dpp(v);
  const unsigned long long mask = __ballot(v == m && i >= 0);
  if (mask) { const int lane = __ffsll((long long)mask) - 1; i = __builtin_amdgcn_readlane(i, lane); v = m; }
  else { v = -1.0f; i = -1; }
}

// Poll up to four tagged words at once (p[i] == nullptr: not needed); each lane keeps all of its loads in
// flight instead of waiting for them one after the other.
__device__ __forceinline__ void get_words4(const uint64_t* const p[4], uint32_t tag, uint32_t v[4], volatile lds_i32* s_abort) {
  bool done[4];
#pragma unroll
  for (int i = 0; i < 4; i++) done[i] = (p[i] == nullptr);
  for (long it = 0; it < SPIN_MAX; it++) {
    uint64_t w[4];
#pragma unroll
    for (int i = 0; i < 4; i++) w[i] = done[i] ? 0ull : __hip_atomic_load(p[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool all = true;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (!done[i] && (uint32_t)(w[i] >> 32) == tag) { v[i] = (uint32_t)w[i]; done[i] = true; }
      all = all && done[i];
    }
    if (all || *s_abort) return;
    __builtin_amdgcn_s_sleep(1);
  }
  *s_abort = 1;
}

// grid = G workgroups x 256 threads (all co-resident: cooperative launch), dynamic LDS = R*kb floats + R ints.
// G <= 128 and R <= 256.  Workgroup g owns panel rows [g*R, g*R + R).  Per column j: (1) local argmax; waves 2-3 publish the head
// word, the candidate row (kb words + identity) and, in the owner of row j, the old row j; (2) waves 0-1
// read the G head words and every workgroup picks the same winner (largest |a|, ties to the smallest row);
// (3) waves 0-1 read the winner's row (the owner of the pivot row also the old row j) and the rows are
// interchanged; (4) column j is scaled and the rank-1 update applied to the rows below.  Two dependent round
// trips per column.  Stores and loads share one counter per wave on this architecture, so the publishing
// waves never poll and the polling waves never store to global memory inside the loop.
__global__ __launch_bounds__(256) void k_panel(PanelArgs a) {
  extern __shared__ float lds[];
  __shared__ float prow[MAX_KB];
  __shared__ int s_piv[MAX_KB];
  __shared__ float wv[2][4];
  __shared__ int wi[2][4];
  __shared__ int s_abort;
  __shared__ int s_prid, s_jrid, s_info;
  const int t = threadIdx.x, g = blockIdx.x;
  const int R = a.R, kb = a.kb, m = a.m, G = a.G;
  const int r0 = g * R;
  const int rows = max(0, min(R, m - r0));
  int* rid = (int*)(lds + (size_t)R * kb);   // panel-local source row of each row's current content
  float* Ap = a.A + a.k + a.k * a.lda;
  if (t == 0) { s_abort = 0; s_info = 0; }
  for (int idx = t; idx < R * kb; idx += 256) {
    const int c = idx / R, r = idx - c * R;
    lds[idx] = (r < rows) ? Ap[(int64_t)(r0 + r) + (int64_t)c * a.lda] : 0.0f;
  }
  for (int r = t; r < R; r += 256) rid[r] = r0 + r;
  const int SW = kb + 1;
  const int64_t PSTRIDE = (int64_t)G + (int64_t)G * SW + SW;
  const int lane = t & 63, wave = t >> 6;
  const bool poller = t < 128;              // waves 0-1
  const int pt = t, st = t - 128;            // poller / publisher thread index
  const int ur = t % R, cg = t / R, ncg = 256 / R;
  float l_prev = 0.0f; int j_prev = -1; bool act_prev = false;   // deferred store of the scaled column
  for (int j = 0; j < kb; j++) {
    const uint32_t tag = (uint32_t)(a.k + j + 1);
    const uint32_t tag24 = tag & 0xFFFFFFu;
    uint64_t* heads = a.xbuf + (int64_t)(j & 1) * PSTRIDE;
    uint64_t* rowsl = heads + G;
    uint64_t* diag = rowsl + (int64_t)G * SW;
    PANEL_MARK(0);
    __syncthreads();                       // step j-1 finished reading column j-1 and updating column j
    if (act_prev && cg == 0) lds[j_prev * R + ur] = l_prev;   // multipliers of column j-1
    // (1) local argmax of column j over rows with panel-local index >= j
    float v = -1.0f; int vi = -1;
    if (t < rows) {
      const int pr = r0 + t;
      if (pr >= j) { float x = fabsf(lds[j * R + t]); if (x != x) x = INFINITY; v = x; vi = pr; }
    }
    wave_argmax(v, vi);
    if (lane == 0) { wv[0][wave] = v; wi[0][wave] = vi; }
    __syncthreads();
    PANEL_MARK(1);
    float bv = wv[0][0]; int bi = wi[0][0];
    for (int w = 1; w < 4; w++) amax_combine(bv, bi, wv[0][w], wi[0][w]);
    const bool own_j = (j >= r0 && j < r0 + rows);
    if (!poller) {
      if (st == 0) {
        const uint64_t hv = (bi >= 0) ? (uint64_t)__float_as_uint(bv) : (uint64_t)__float_as_uint(-1.0f);
        const uint64_t lr = (bi >= 0) ? (uint64_t)(bi - r0) : 0ull;
        __hip_atomic_store(heads + g, ((uint64_t)tag24 << 40) | (lr << 32) | hv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (bi >= 0) {
        for (int c = st; c < kb; c += 128) put_word(rowsl + (int64_t)g * SW + c, tag, __float_as_uint(lds[c * R + (bi - r0)]));
        if (st == 127) put_word(rowsl + (int64_t)g * SW + kb, tag, (uint32_t)rid[bi - r0]);
      }
      if (own_j) {
        for (int c = st; c < kb; c += 128) put_word(diag + c, tag, __float_as_uint(lds[c * R + (j - r0)]));
        if (st == 126) put_word(diag + kb, tag, (uint32_t)rid[j - r0]);
      }
    }
    PANEL_MARK(2);
    // (2) agree on the pivot from the G head words (waves 0-1, one head per lane; G <= 128)
    if (poller) {
      float gv = -1.0f; int gi = -1;
      if (pt < G) {
        const uint64_t w = get_head(heads + pt, tag24, LDS_FLAG(&s_abort));
        const float hv = __uint_as_float((uint32_t)w);
        if (hv >= 0.0f) {
          gi = pt * R + (int)((w >> 32) & 0xFFull);
          gv = hv;
          if (gi >= m) { gi = -1; gv = -1.0f; }
        }
      }
      wave_argmax(gv, gi);
      if (lane == 0) { wv[1][wave] = gv; wi[1][wave] = gi; }
    }
    PANEL_MARK(3);
    __syncthreads();
    float pvv = wv[1][0]; int p = wi[1][0];
    amax_combine(pvv, p, wv[1][1], wi[1][1]);
    if (p < 0) p = j;
    const int gp = p / R;
    PANEL_MARK(4);
    // (3) pivot row and the interchange of rows j and p (waves 0-1 read; the owner of row p copies)
    if (poller) {
      const bool need_diag = (p != j && gp == g && !own_j);
      const uint64_t* pw[4] = {nullptr, nullptr, nullptr, nullptr};
      uint32_t vw[4] = {0u, 0u, 0u, 0u};
      const int c0 = pt, c1 = pt + 128;    // kb <= 256
      if (c0 < kb) { pw[0] = rowsl + (int64_t)gp * SW + c0; if (need_diag) pw[2] = diag + c0; }
      if (c1 < kb) { pw[1] = rowsl + (int64_t)gp * SW + c1; if (need_diag) pw[3] = diag + c1; }
      get_words4(pw, tag, vw, LDS_FLAG(&s_abort));
      if (c0 < kb) {
        prow[c0] = __uint_as_float(vw[0]);
        if (p != j && gp == g) lds[c0 * R + (p - r0)] = own_j ? lds[c0 * R + (j - r0)] : __uint_as_float(vw[2]);
      }
      if (c1 < kb) {
        prow[c1] = __uint_as_float(vw[1]);
        if (p != j && gp == g) lds[c1 * R + (p - r0)] = own_j ? lds[c1 * R + (j - r0)] : __uint_as_float(vw[3]);
      }
      if (pt == 127 && p != j) {
        const uint64_t* pi[4] = {nullptr, nullptr, nullptr, nullptr};
        uint32_t vi2[4] = {0u, 0u, 0u, 0u};
        if (own_j && gp != g) pi[0] = rowsl + (int64_t)gp * SW + kb;
        if (gp == g && !own_j) pi[1] = diag + kb;
        get_words4(pi, tag, vi2, LDS_FLAG(&s_abort));
        if (own_j) s_prid = (gp == g) ? rid[p - r0] : (int)vi2[0];
        if (gp == g) s_jrid = own_j ? rid[j - r0] : (int)vi2[1];
      }
    }
    __syncthreads();
    PANEL_MARK(5);
    if (p != j) {
      if (own_j) { for (int c = t; c < kb; c += 256) lds[c * R + (j - r0)] = prow[c]; }
      if (t == 127) { if (own_j) rid[j - r0] = s_prid; if (gp == g) rid[p - r0] = s_jrid; }
    }
    if (t == 0) {
      s_piv[j] = p;
      if (prow[j] == 0.0f && s_info == 0) s_info = (int)(a.k + j + 1);
    }
    // (4) l = a_rj / pivot (stored at the start of the next step), rank-1 update of the rows below
    const float piv = prow[j];
    const int pr = r0 + ur;
    const bool active = (ur < rows) && (pr > j) && (piv != 0.0f);
    float l = 0.0f;
    if (active) {
      const float x = lds[j * R + ur];
      l = (fabsf(piv) >= 1.17549435e-38f) ? x * (1.0f / piv) : x / piv;
      // 8 columns per iteration: all LDS loads are issued before the stores (the stores may alias prow as
      // far as the compiler knows, which would otherwise serialise every iteration on LDS latency)
      int c = j + 1 + cg;
      for (; c + 7 * ncg < kb; c += 8 * ncg) {
        float av[8], uv[8];
#pragma unroll
        for (int q = 0; q < 8; q++) { av[q] = lds[(c + q * ncg) * R + ur]; uv[q] = prow[c + q * ncg]; }
#pragma unroll
        for (int q = 0; q < 8; q++) lds[(c + q * ncg) * R + ur] = fmaf(-l, uv[q], av[q]);
      }
      for (; c < kb; c += ncg) lds[c * R + ur] = fmaf(-l, prow[c], lds[c * R + ur]);
    }
    l_prev = l; j_prev = j; act_prev = active;
    PANEL_MARK(6);
  }
  __syncthreads();
  if (act_prev && cg == 0) lds[j_prev * R + ur] = l_prev;
  __syncthreads();
  for (int idx = t; idx < R * kb; idx += 256) {
    const int c = idx / R, r = idx - c * R;
    if (r < rows) Ap[(int64_t)(r0 + r) + (int64_t)c * a.lda] = lds[idx];
  }
  if (g == 0) {
    for (int jj = t; jj < kb; jj += 256) a.ipiv[a.k + jj] = a.k + s_piv[jj] + 1;
    if (t == 0 && s_info != 0 && a.dinfo[0] == 0) a.dinfo[0] = s_info;
  }
  // swap plan: every row whose content moved is a (dst, src) pair
  for (int r = t; r < rows; r += 256) {
    const int src = rid[r];
    if (src != r0 + r) {
      const int slot = atomicAdd(a.plan, 1);
      a.plan[1 + 2 * slot] = r0 + r;
      a.plan[2 + 2 * slot] = src;
    }
  }
  if (t == 0 && s_abort) a.dinfo[1] = 1;
}

// ------------------------------------------------------------------ register-resident panel
// k_panel2: the same pivoting and the same floating-point operations per element as k_panel, arranged so a
// column step costs one exchange of head words plus one read of the winner's row, with no LDS traffic
// proportional to the panel width.
//   * Workgroup g owns panel rows [128 g, 128 g + 128).  Wave 0 holds rows lane and lane + 64 of the
//     current 32-column sub-block in registers, rotated one place per step so column j is always x[0]
//     (every register index is a compile-time constant).  The whole panel row also lives in LDS (P).
//   * Rows never move: each carries its current position `pos` (LAPACK's sequential-interchange order).
//     At step j the winner's position becomes j and the row that held position j takes the winner's old
//     position, so "ties to the smallest index" compares positions.
//   * Exchange per step, all in wave 0: it finds its workgroup's candidate (DPP argmax) and pushes a
//     16-byte head (tag, position, the candidate's value and its next-column value) into every workgroup's
//     mailbox, then writes the candidate's 32 values to its row slot (four replicas).  After reading its own
//     mailbox every workgroup picks the same winner; the winner's value and next-column value are enough to
//     form the multipliers and column j+1 and to find the next candidate, while the winner's row (needed
//     for the rest of the rank-1 update and for the next head's next-column value) is being fetched.
//     Mailboxes sit 4 KB apart and replicas 64 KB apart, so the polling of 100+ workgroups spreads over
//     memory channels instead of queueing on a few cache lines.
//   * After 32 steps the panel columns right of the sub-block get the deferred update
//     a -= sum_d l_d u_d (d in step order, so each element sees the same fmaf chain as 32 rank-1 updates):
//     the owners publish their pivot rows (four replicas, double-buffered by sub-block), every workgroup
//     forms the 32 U rows by forward substitution.  The next 32 columns are updated at once by all four
//     waves; the columns beyond them are updated by waves 2-3 while waves 0-1 factor the next sub-block.
constexpr int P2_R = 128;
constexpr int P2_PS = 129;          // LDS column stride of the panel slice (odd: column-per-lane access is conflict-free)
constexpr int P2_MAXG = 256;        // workgroups = exchange participants
constexpr int P2_MBOX = 512;        // words from one mailbox to the next (4 KB)
constexpr int P2_NREP = 4;          // replicas of published rows
constexpr int P2_US = 36;           // padded row length of the U staging tile (16-byte aligned rows)
// exchange area (64-bit words): mailboxes [parity][reader][P2_MBOX], candidate rows [parity][replica]
// [participant][32], sub-block rows [sub-block parity][replica][32 rows][256 columns]
constexpr int64_t P2_OFF_ROWS = 2LL * P2_MAXG * P2_MBOX;
constexpr int64_t P2_OFF_SUB = P2_OFF_ROWS + 2LL * P2_NREP * P2_MAXG * 32;
constexpr int64_t P2_SUBREP = 32 * 256;
constexpr int64_t P2_WORDS = P2_OFF_SUB + 2 * P2_NREP * P2_SUBREP;

struct Panel2Args {
  float* A; int64_t lda; int64_t k; int kb; int m; int G;
  uint64_t* xbuf;    // P2_WORDS words, zeroed before the factorization (tags are unique within it)
  int64_t* ipiv; int32_t* dinfo; int32_t* plan;
  int32_t* arrive;   // incremented once by every workgroup as it starts (nullptr: not counted)
};

// LDS layout (bytes): the fixed-size arrays first, at compile-time offsets (16-byte aligned), so their
// addresses are instruction immediates; the kb-dependent panel slice P last.
constexpr int P2L_MISC = 0;                              // 16 ints: flags and step values
constexpr int P2L_PROW = P2L_MISC + 16 * 4;              // 32 floats
constexpr int P2L_CROW = P2L_PROW + 32 * 4;              // 32 floats
constexpr int P2L_SPIV = P2L_CROW + 32 * 4;              // 256 ints
constexpr int P2L_PIVL = P2L_SPIV + 256 * 4;             // 32 ints
constexpr int P2L_PSTEP = P2L_PIVL + 32 * 4;             // P2_R ints
constexpr int P2L_PSTEPBG = P2L_PSTEP + P2_R * 4;        // P2_R ints
constexpr int P2L_SPOS = P2L_PSTEPBG + P2_R * 4;         // P2_R ints
constexpr int P2L_US = P2L_SPOS + P2_R * 4;              // [32][P2_US] floats
constexpr int P2L_LST = P2L_US + 32 * P2_US * 4;         // [32][32] floats
constexpr int P2L_USF = P2L_LST + 32 * 32 * 4;           // [2][8][P2_US] floats
constexpr int P2L_P = P2L_USF + 2 * 8 * P2_US * 4;       // [kb][P2_PS] floats
static_assert(P2L_US % 16 == 0 && P2L_LST % 16 == 0 && P2L_USF % 16 == 0 && P2L_P % 16 == 0, "LDS alignment");
__host__ __device__ __forceinline__ int p2_pfloats(int kb) { return (kb * P2_PS + 3) & ~3; }
static size_t panel2_lds_bytes(int kb) { return (size_t)P2L_P + (size_t)p2_pfloats(kb) * 4; }

__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
__device__ __forceinline__ void lds_wait() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

__device__ __forceinline__ int wave_min_dpp(int x) {
  x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x111, 0xf, 0xf, false));
  x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x112, 0xf, 0xf, false));
  x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x114, 0xf, 0xf, false));
  x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x118, 0xf, 0xf, false));
  x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x142, 0xa, 0xf, false));
  x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(x, 63);
}
// Pivot-search key of a value: its |value| bits, NaN counted as infinity.  On non-negative floats the order
// of the bit patterns as integers is the order of the values; -1 marks "no candidate".
__device__ __forceinline__ int p2_key(float v) {
  const uint32_t b = __float_as_uint(v) & 0x7FFFFFFFu;
  return (int)(b > 0x7F800000u ? 0x7F800000u : b);
}
// Wave-wide maximum of a signed integer: one DPP-modified v_max per stage (row_shr 1/2/4/8 inside rows of 16
// lanes, then row_bcast 15/31 across rows; a lane whose DPP source is invalid keeps its value), two wait
// states before each stage reads the previous stage's result.  All 64 lanes must be active.
__device__ __forceinline__ int wave_max_i32(int x) {
  int r;
  asm volatile(
      "v_mov_b32 %0, %1\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_max_i32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
      "s_nop 1"
      : "=&v"(r) : "v"(x));
  return __builtin_amdgcn_readlane(r, 63);
}
// Wave argmax of (key, pos): largest key, ties to the smallest pos; key < 0 marks "no candidate".  Returns the
// winning lane (wave-uniform) or -1, and the winner's pos.  All 64 lanes must be active.
__device__ __forceinline__ int wave_argmax_key(int key, int pos, int& bpos) {
  const int mx = wave_max_i32(key);
  if (mx < 0) { bpos = 0x7FFFFFFF; return -1; }
  const unsigned long long tie = __ballot(key == mx);
  int lane;
  if (__popcll(tie) == 1) {
    lane = __ffsll((long long)tie) - 1;
  } else {
    const int mn = wave_min_dpp(key == mx ? pos : 0x7FFFFFFF);
    lane = __ffsll((long long)__ballot(key == mx && pos == mn)) - 1;
  }
  lane = __builtin_amdgcn_readfirstlane(lane);
  bpos = __builtin_amdgcn_readlane(pos, lane);
  return lane;
}
// 16-byte head slots of the panel exchange, written and read with one vector access at agent scope (sc1):
// both halves carry the tag, so a slot read while it is being written is simply read again.  The loads wait
// for their data inside the asm block (their destination registers are written by the memory system later).
typedef unsigned long long p2u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void p2_head_load2(const uint64_t* pa, const uint64_t* pb, p2u64x2& ra, p2u64x2& rb) {
  asm volatile("global_load_dwordx4 %0, %2, off sc1\n\t"
               "global_load_dwordx4 %1, %3, off sc1\n\t"
               "s_waitcnt vmcnt(0)"
               : "=&v"(ra), "=&v"(rb) : "v"(pa), "v"(pb) : "memory");
}
__device__ __forceinline__ void p2_head_store(uint64_t* p, uint64_t w0, uint64_t w1) {
  const p2u64x2 v = {w0, w1};
  asm volatile("global_store_dwordx4 %0, %1, off sc1\n\ts_nop 0" :: "v"(p), "v"(v) : "memory");
}
// A tagged word whose first load (`first`) was issued earlier: polled until it carries the tag.
__device__ __forceinline__ uint32_t p2_wait_word(const uint64_t* p, uint64_t first, uint32_t tag, volatile lds_i32* s_abort) {
  if ((uint32_t)(first >> 32) == tag) return (uint32_t)first;
  for (long it = 0; it < SPIN_MAX; it++) {
    if (*s_abort) return 0u;
    __builtin_amdgcn_s_sleep(1);
    const uint64_t w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(w >> 32) == tag) return (uint32_t)w;
  }
  *s_abort = 1;
  return 0u;
}

// Poll four tagged U-row words (q, q + 1, q + 2, q + 3 of one column: stride 256 words) into u; a word that
// never arrives sets the abort flag and reads as zero.
__device__ __forceinline__ void p2_poll4(const uint64_t* xs, uint32_t tag, volatile lds_i32* s_abort, float (&u)[4]) {
  uint32_t need = 0xFu;
  for (long it = 0; it < SPIN_MAX && need; it++) {
    uint64_t w[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
      w[i] = (need >> i & 1u) ? __hip_atomic_load(xs + i * 256, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
#pragma unroll
    for (int i = 0; i < 4; i++)
      if ((need >> i & 1u) && (uint32_t)(w[i] >> 32) == tag) { u[i] = __uint_as_float((uint32_t)w[i]); need &= ~(1u << i); }
    if (!need || *s_abort) break;
    __builtin_amdgcn_s_sleep(1);
  }
  if (need) {
    *s_abort = 1;
#pragma unroll
    for (int i = 0; i < 4; i++) if (need >> i & 1u) u[i] = 0.0f;
  }
}
// Forward substitution of U columns spread over lanes: within each half-wave, lane = 8 * column + qb holds
// the U rows 4 qb .. 4 qb + 3 of its column.  At step d the lane holding row d (final since step d - 1)
// broadcasts u_d to its column's lanes (ds_swizzle), and every row q > d takes u_q -= l_{q,d} u_d: per row
// the operations of sequential forward substitution, in increasing d.  Steps whose pivot was zero are
// skipped.  LsT[d][q] = l_{q,d}.  All 64 lanes must be active.
template <int D>
__device__ __forceinline__ void p2_subst_step(const float* LsT, int zmask, int qb, float (&u)[4]) {
  const float ud = __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(u[D & 3]), 0x18 | ((D >> 2) << 5)));
  if (!((zmask >> D) & 1)) {
    const float4 lq = *(const float4*)(LsT + D * 32 + 4 * qb);
    const float le[4] = {lq.x, lq.y, lq.z, lq.w};
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const float nv = fmaf(-le[e], ud, u[e]);
      u[e] = (4 * qb + e > D) ? nv : u[e];
    }
  }
  if constexpr (D + 1 < 31) p2_subst_step<D + 1>(LsT, zmask, qb, u);
}
__device__ __forceinline__ void p2_subst_lanes(const float* LsT, int zmask, int qb, float (&u)[4]) {
  p2_subst_step<0>(LsT, zmask, qb, u);
}
// Two rows' updates of one column: y -= sum_d l_d u_d in increasing d (steps in zmask skipped), u = Uc[0..31]
// read as broadcasts shared by both rows.
__device__ __forceinline__ void p2_chain2(float& ya, float& yb, const float (&la)[32], const float (&lb)[32],
                                          const float* Uc, int zmask) {
  const float4* u4 = (const float4*)Uc;
  if (zmask == 0) {
#pragma unroll
    for (int q4 = 0; q4 < 8; q4++) {
      const float4 v = u4[q4];
      const float ue[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int e = 0; e < 4; e++) { ya = fmaf(-la[4 * q4 + e], ue[e], ya); yb = fmaf(-lb[4 * q4 + e], ue[e], yb); }
    }
  } else {
#pragma unroll
    for (int q4 = 0; q4 < 8; q4++) {
      const float4 v = u4[q4];
      const float ue[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const bool skip = (zmask >> (4 * q4 + e) & 1) != 0;
        const float na = fmaf(-la[4 * q4 + e], ue[e], ya), nb = fmaf(-lb[4 * q4 + e], ue[e], yb);
        ya = skip ? ya : na; yb = skip ? yb : nb;
      }
    }
  }
}
__device__ __forceinline__ uint32_t p2_emask(int dlim, int zmask) {
  if (dlim <= 0) return 0u;
  return (dlim >= 32 ? 0xFFFFFFFFu : ((1u << dlim) - 1u)) & ~(uint32_t)zmask;
}

__global__ __launch_bounds__(256) void k_panel2(Panel2Args a) {
  extern __shared__ float lds[];
  const int kb = a.kb, m = a.m, G = a.G;
  char* lb = (char*)lds;
  float* P = (float*)(lb + P2L_P);                        // [kb][P2_PS]
  float* Us = (float*)(lb + P2L_US);                      // [32][P2_US] U rows of the next 32 columns
  float* LsT = (float*)(lb + P2L_LST);                    // [32][32]: LsT[d][q] = l_{q,d}
  float* prow = (float*)(lb + P2L_PROW);                  // [32] previous step's pivot row
  float* crow = (float*)(lb + P2L_CROW);                  // [32] this workgroup's pivot row staging
  int* s_piv = (int*)(lb + P2L_SPIV);                     // [256]
  int* piv_local = (int*)(lb + P2L_PIVL);                 // [32]
  int* pstep = (int*)(lb + P2L_PSTEP);                    // [128] step at which a row became pivot (or -1 / 32)
  int* pstep_bg = (int*)(lb + P2L_PSTEPBG);               // [128] copy for the background update
  int* spos = (int*)(lb + P2L_SPOS);                      // [128]
  int* misc = (int*)(lb + P2L_MISC);                      // [16]
  float* UsF = (float*)(lb + P2L_USF);                    // [2 waves][8][P2_US] background U columns
  volatile lds_i32* s_abort = LDS_FLAG(misc + 0);
  int* s_info = misc + 3; int* s_zmask = misc + 4;

  const int t = threadIdx.x, g = blockIdx.x, lane = t & 63, wave = t >> 6;
  const int r0 = g * P2_R;
  const int rows = max(0, min(P2_R, m - r0));
  float* Ap = a.A + a.k + a.k * a.lda;
  uint64_t* mbox = a.xbuf;
  uint64_t* rowsl = a.xbuf + P2_OFF_ROWS;
  const int myrep = g % P2_NREP;
  PANEL2_WGMARK(0);
  if (t == 0 && a.arrive != nullptr) __hip_atomic_fetch_add(a.arrive, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (t == 0) { *s_abort = 0; *s_info = 0; *s_zmask = 0; }
  {
    // the slice into LDS: thread (row r, column parity ch) keeps 32 loads in flight before its first LDS write
    // (one load per iteration would wait a full memory round trip 128 times).  Rows past the panel read the
    // last row and store zero; columns past kb read column kb - 1 and are not stored.
    const int r = t & (P2_R - 1), ch = t >> 7;
    const float* src = Ap + (int64_t)(r0 + min(r, rows - 1));
    for (int cb = 0; cb < kb; cb += 64) {
      float v[32];
#pragma unroll
      for (int i = 0; i < 32; i++) v[i] = src[(int64_t)min(cb + 2 * i + ch, kb - 1) * a.lda];
#pragma unroll
      for (int i = 0; i < 32; i++) {
        const int c = cb + 2 * i + ch;
        if (c < kb) P[c * P2_PS + r] = (r < rows) ? v[i] : 0.0f;
      }
    }
  }
  // rows of wave 0: set 0 = row lane, set 1 = row lane + 64
  const bool v0 = (wave == 0) && lane < rows, v1 = (wave == 0) && lane + 64 < rows;
  int pos0 = r0 + lane, pos1 = r0 + lane + 64;
  int bg_c0 = -1, bg_zmask = 0;                           // background update pending for waves 2-3
  lds_barrier();
  PANEL2_WGMARK(1);

  for (int c0 = 0; c0 < kb; c0 += 32) {
    const int w = min(32, kb - c0);
    int ps0 = -1, ps1 = -1;                               // -1: pivot before this sub-block (or no row)
    if (wave >= 2) {
      // far columns of the previous sub-block, 8 per pass per wave: the wave polls and substitutes the pass's 8
      // U columns (p2_subst_lanes layout), stages them in LDS and updates rows lane and lane + 64 of those
      // columns, the rows' multipliers held in registers.  A row that was a pivot step of that sub-block takes
      // its substituted U value (the same operations on the same operands), earlier pivots stay as they are.
      if (bg_c0 >= 0) {
        const uint32_t btag = (uint32_t)(a.k + bg_c0 + 1);
        const uint64_t* xsb = a.xbuf + P2_OFF_SUB + (int64_t)(((bg_c0 >> 5) & 1) * P2_NREP + myrep) * P2_SUBREP;
        const int wb = wave - 2, cl = lane >> 3, qb = lane & 7;
        float* Uw = UsF + wb * 8 * P2_US;
        const int psa = pstep_bg[lane], psb = pstep_bg[lane + 64];
        float la[32], lb[32];
#pragma unroll
        for (int d = 0; d < 32; d++) { la[d] = P[(bg_c0 + d) * P2_PS + lane]; lb[d] = P[(bg_c0 + d) * P2_PS + lane + 64]; }
        for (int fb = bg_c0 + 64 + 8 * wb; fb < kb; fb += 16) {
          const int col = fb + cl;
          float u[4] = {0.0f, 0.0f, 0.0f, 0.0f};
          if (col < kb) p2_poll4(xsb + (4 * qb) * 256 + col, btag, s_abort, u);
          p2_subst_lanes(LsT, bg_zmask, qb, u);
          *(float4*)(Uw + cl * P2_US + 4 * qb) = make_float4(u[0], u[1], u[2], u[3]);
          lds_wait();
          const int nc = min(8, kb - fb);
          for (int c = 0; c < nc; c++) {
            const int fc = fb + c;
            const float* Uc = Uw + c * P2_US;
            float ya = P[fc * P2_PS + lane], yb = P[fc * P2_PS + lane + 64];
            p2_chain2(ya, yb, la, lb, Uc, bg_zmask);
            if (psa >= 0) P[fc * P2_PS + lane] = (psa == 32) ? ya : Uc[min(psa, 31)];
            if (psb >= 0) P[fc * P2_PS + lane + 64] = (psb == 32) ? yb : Uc[min(psb, 31)];
          }
        }
      }
    } else if (wave == 0) {
      // Wave 0 runs every column step end to end: it polls its mailbox, picks the winner, reads the winner's
      // row, updates its two rows per lane and pushes the next head, so no hand-off between waves lies on the
      // chain of steps.  The only stores it has in flight while polling are the previous head and candidate
      // row, which complete within the exchange round trip.
      float x0[32], x1[32];
#pragma unroll
      for (int q = 0; q < 32; q++) {
        x0[q] = (c0 + q < kb) ? P[(c0 + q) * P2_PS + lane] : 0.0f;
        x1[q] = (c0 + q < kb) ? P[(c0 + q) * P2_PS + lane + 64] : 0.0f;
      }
      ps0 = (v0 && pos0 >= c0) ? 32 : -1;
      ps1 = (v1 && pos1 >= c0) ? 32 : -1;
      if (lane == 0) *s_zmask = 0;
      if (lane < 32) piv_local[lane] = -1;
      // Step order: the candidate of column j+1 is found from the rows' column-j+1 values alone (updated
      // first) and its head pushed out; the rest of the rank-1 update, the multipliers and the candidate's
      // row (which readers fetch only after the heads round trip) follow while the exchange is in flight.
      int blane = -1;                                     // candidate lane of the current column (-1: none)
      bool bsel = false;                                  // candidate is the lane's second row
      float cur_c = 0.0f;                                 // the candidate's value in the column it competes for
      int pos_c = 0xFFFF;                                 // its position (0xFFFF: no candidate)
      // The candidate among the rows in play (cur*: their values in the column), by a wave argmax of |value|
      // (NaN counted as infinity; ties to the smaller position, as isamax).
      auto find_candidate = [&](float cur0, float cur1, bool in0, bool in1) {
        const int k0 = in0 ? p2_key(cur0) : -1, k1 = in1 ? p2_key(cur1) : -1;
        const bool take1 = (k1 > k0) || (k1 == k0 && k1 >= 0 && pos1 < pos0);
        int bpos;
        PANEL2_MARKB(c0 + 31, 2);
        blane = wave_argmax_key(take1 ? k1 : k0, take1 ? pos1 : pos0, bpos);
        PANEL2_MARKB(c0 + 31, 3);
        bsel = (blane >= 0) && (__builtin_amdgcn_readlane((int)take1, blane) != 0);
        pos_c = blane >= 0 ? bpos : 0xFFFF;
        cur_c = blane >= 0 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(take1 ? cur1 : cur0), blane)) : 0.0f;
      };
      // The head of column j: [tag | position | value] and [tag | next-column value], one 16-byte slot in every
      // reader's mailbox.  The next-column value is the candidate's column j+1 value with every update through
      // step j-1, so the step that picks the winner can update its rows' column j+1 before the winner's row
      // arrives.  Both halves carry the tag: a reader accepts a slot only when both do.
      auto head_push = [&](int j, float nxt_c) {
        const uint64_t tg = (uint64_t)((uint32_t)(a.k + j + 1) & 0xFFFFu) << 48;
        const uint64_t w0 = tg | ((uint64_t)(uint32_t)(pos_c & 0xFFFF) << 32) | (uint64_t)__float_as_uint(cur_c);
        const uint64_t w1 = tg | (uint64_t)__float_as_uint(nxt_c);
        uint64_t* mrow = mbox + (size_t)(j & 1) * P2_MAXG * P2_MBOX + 2 * g;
        for (int rd = lane; rd < G; rd += 64) p2_head_store(mrow + (size_t)rd * P2_MBOX, w0, w1);
      };
      // the candidate's 32 values (x[0] = column j), staged in LDS, four replicas
      auto publish_row = [&](int j) {
        if (blane < 0) return;
        const uint32_t tag = (uint32_t)(a.k + j + 1);
        const int par = j & 1;
        if (lane == blane) {
          float4* c4 = (float4*)crow;
          if (bsel) {
#pragma unroll
            for (int q = 0; q < 8; q++) c4[q] = make_float4(x1[4 * q], x1[4 * q + 1], x1[4 * q + 2], x1[4 * q + 3]);
          } else {
#pragma unroll
            for (int q = 0; q < 8; q++) c4[q] = make_float4(x0[4 * q], x0[4 * q + 1], x0[4 * q + 2], x0[4 * q + 3]);
          }
        }
        lds_wait();
        const uint32_t mine = __float_as_uint(crow[lane & 31]);
        const int rep = lane >> 5;
        put_word(rowsl + (((size_t)(par * P2_NREP + rep) * P2_MAXG + g) * 32 + (lane & 31)), tag, mine);
        put_word(rowsl + (((size_t)(par * P2_NREP + rep + 2) * P2_MAXG + g) * 32 + (lane & 31)), tag, mine);
      };
      find_candidate(x0[0], x1[0], v0 && pos0 >= c0, v1 && pos1 >= c0);
      head_push(c0, blane >= 0 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bsel ? x1[1] : x0[1]), blane)) : 0.0f);
      publish_row(c0);
      for (int jj = 0; jj < w; jj++) {
        const int j = c0 + jj;
        const uint32_t tag = (uint32_t)(a.k + j + 1);
        const uint32_t tag16 = tag & 0xFFFFu;
        const int par = j & 1;
        const int ql = lane & 31;
        PANEL2_MARK(j, 4);
        // own mailbox: the G heads of column j (slot s at words 2s, 2s + 1)
        const uint64_t* mb = mbox + ((size_t)par * P2_MAXG + g) * P2_MBOX;
        uint64_t hw0[4] = {0ull, 0ull, 0ull, 0ull}, hw1[4] = {0ull, 0ull, 0ull, 0ull};
        uint32_t need = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) if (lane + 64 * i < G) need |= 1u << i;
        const int npair = G > 128 ? 2 : 1;
        long it_heads = 0;
        for (long it = 0; it < SPIN_MAX; it++, it_heads++) {
          for (int h = 0; h < npair; h++) {
            p2u64x2 sa, sb;
            p2_head_load2(mb + 2 * (lane + 128 * h), mb + 2 * (lane + 128 * h + 64), sa, sb);
            const int ia = 2 * h, ib = 2 * h + 1;
            if ((need >> ia & 1u) && (uint32_t)(sa.x >> 48) == tag16 && (uint32_t)(sa.y >> 48) == tag16) {
              hw0[ia] = sa.x; hw1[ia] = sa.y; need &= ~(1u << ia);
            }
            if ((need >> ib & 1u) && (uint32_t)(sb.x >> 48) == tag16 && (uint32_t)(sb.y >> 48) == tag16) {
              hw0[ib] = sb.x; hw1[ib] = sb.y; need &= ~(1u << ib);
            }
          }
          if (__ballot(need != 0u) == 0ull) break;
          if (*s_abort) break;
          __builtin_amdgcn_s_sleep(1);
        }
        if (__ballot(need != 0u) != 0ull) *s_abort = 1;
        PANEL2_COUNT(j, it_heads);
        (void)it_heads;
        // the winner: the lane's best head, then the wave's
        int kk = -1, kp = 0x7FFFFFFF, ks = -1;
        uint32_t kv = 0u, kn = 0u;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int pp = lane + 64 * i;
          const int hpos = (int)((hw0[i] >> 32) & 0xFFFFull);
          const uint32_t vb = (uint32_t)hw0[i];
          const int key = (pp < G && !(need >> i & 1u) && hpos != 0xFFFF) ? p2_key(__uint_as_float(vb)) : -1;
          const bool better = key > kk || (key == kk && key >= 0 && hpos < kp);
          kk = better ? key : kk; kp = better ? hpos : kp; ks = better ? pp : ks;
          kv = better ? vb : kv; kn = better ? (uint32_t)hw1[i] : kn;
        }
        PANEL2_MARK(j, 5);
        int p;
        PANEL2_MARKB(j, 5);
        const int wl = wave_argmax_key(kk, kp, p);
        PANEL2_MARKB(j, 6);
        int pw = (wl >= 0) ? __builtin_amdgcn_readlane(ks, wl) : -1;
        if (wl < 0 || p < j || p >= m) { *s_abort = 1; p = j; pw = -1; }
        // the pivot and the pivot row's column j+1 value come with the head
        const float piv = (wl >= 0) ? __uint_as_float(__builtin_amdgcn_readlane(kv, wl)) : 0.0f;
        const float u1 = (wl >= 0) ? __uint_as_float(__builtin_amdgcn_readlane(kn, wl)) : 0.0f;
        PANEL2_MARK(j, 0);
        // the winner's 32 values (lane q < 32: column j + q; lanes 32-63 read the same words): the load is in
        // flight while the multipliers, column j+1 and the next candidate are formed
        const uint64_t* rp = rowsl + (((size_t)(par * P2_NREP + myrep) * P2_MAXG + max(pw, 0)) * 32 + ql);
        const uint64_t rw_first = __hip_atomic_load(rp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const float rinv = 1.0f / piv;
        PANEL2_MARKB(j, 0);
        // positions after the interchange: the winner takes position j, the row that held j takes the
        // winner's old position p; rows in play for column j+1 are those past position j
        const bool win0 = (pw == g) && lane == blane && !bsel, win1 = (pw == g) && lane == blane && bsel;
        if (win0) { pos0 = j; ps0 = jj; }
        if (win1) { pos1 = j; ps1 = jj; }
        if (!win0 && v0 && pos0 == j) pos0 = p;
        if (!win1 && v1 && pos1 == j) pos1 = p;
        const bool in0 = v0 && pos0 > j, in1 = v1 && pos1 > j;
        // rows in play take the multiplier; a zero pivot leaves every row unchanged
        const bool a0 = piv != 0.0f && in0, a1 = piv != 0.0f && in1;
        const bool normal = fabsf(piv) >= 1.17549435e-38f;
        const float l0 = normal ? x0[0] * rinv : x0[0] / piv;
        const float l1 = normal ? x1[0] * rinv : x1[0] / piv;
        PANEL2_MARKB(j, 1);
        const int pblane = blane;                          // this workgroup's candidate of column j
        const bool pbsel = bsel;
        const bool more = jj + 1 < w;
        // column j+1 first (same operands as the full update below), then the candidate of column j+1
        if (more) find_candidate(a0 ? fmaf(-l0, u1, x0[1]) : x0[1], a1 ? fmaf(-l1, u1, x1[1]) : x1[1], in0, in1);
        float pv = 0.0f;
        if (pw >= 0) pv = __uint_as_float(p2_wait_word(rp, rw_first, tag, s_abort));
        PANEL2_MARK(j, 2);
        if (lane < 32) prow[lane] = pv;
        const float lc = (more && blane >= 0) ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bsel ? l1 : l0), blane)) : 0.0f;
        if (more) {
          // the candidate's column j+2 value after this step: the operands of its row published below
          float nxt = 0.0f;
          if (blane >= 0) {
            const float x2c = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bsel ? x1[2] : x0[2]), blane));
            const float pr2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pv), 2));
            nxt = (piv != 0.0f) ? fmaf(-lc, pr2, x2c) : x2c;
          }
          head_push(j + 1, nxt);
        }
        PANEL2_MARK(j, 6);
        if (lane == 0) {
          s_piv[j] = p;
          if (piv == 0.0f) { *s_zmask |= 1 << jj; if (*s_info == 0) *s_info = (int)(a.k + j + 1); }
        }
        if (pw == g) {
          // this workgroup's candidate won: its remaining sub-block values (staged in crow) are final
          const int rw = pblane + (pbsel ? 64 : 0);
          const int qmax = min(32 - jj, kb - j);          // columns j .. j+qmax-1 exist in this sub-block
          if (lane < qmax) P[(j + lane) * P2_PS + rw] = crow[lane];
          if (lane == 0) piv_local[jj] = rw;
        }
        if (more && blane >= 0) {
          // the new candidate's row after this step's update, one value per lane from its current values
          // (same operands as the full update below), published right behind the head
          if (lane == blane) {
            float4* c4 = (float4*)crow;
            if (bsel) {
#pragma unroll
              for (int q = 0; q < 8; q++) c4[q] = make_float4(x1[4 * q], x1[4 * q + 1], x1[4 * q + 2], x1[4 * q + 3]);
            } else {
#pragma unroll
              for (int q = 0; q < 8; q++) c4[q] = make_float4(x0[4 * q], x0[4 * q + 1], x0[4 * q + 2], x0[4 * q + 3]);
            }
          }
          lds_wait();
          const float xo = crow[min(ql + 1, 31)];
          const float prq = prow[min(ql + 1, 31)];        // pivot-row value this lane contributes to the next row
          float nv = (piv != 0.0f) ? fmaf(-lc, prq, xo) : xo;
          if (ql == 31) nv = 0.0f;
          lds_wait();                                     // every lane's read is done before crow is overwritten
          if (lane < 32) crow[ql] = nv;                   // the candidate's row as the pivot row of step j + 1
          const uint32_t tag1 = (uint32_t)(a.k + j + 2);
          const int par1 = (j + 1) & 1, rep = lane >> 5;
          put_word(rowsl + (((size_t)(par1 * P2_NREP + rep) * P2_MAXG + g) * 32 + ql), tag1, __float_as_uint(nv));
          put_word(rowsl + (((size_t)(par1 * P2_NREP + rep + 2) * P2_MAXG + g) * 32 + ql), tag1, __float_as_uint(nv));
        }
        PANEL2_MARK(j, 1);
        float pr[32];
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const float4 v4 = ((const float4*)prow)[q];
          pr[4 * q] = v4.x; pr[4 * q + 1] = v4.y; pr[4 * q + 2] = v4.z; pr[4 * q + 3] = v4.w;
        }
        if (piv != 0.0f) {
          if (a0) P[j * P2_PS + lane] = l0;
          if (a1) P[j * P2_PS + lane + 64] = l1;
          const float m0 = a0 ? l0 : 0.0f, m1 = a1 ? l1 : 0.0f;
#pragma unroll
          for (int q = 1; q < 32; q++) { x0[q - 1] = fmaf(-m0, pr[q], x0[q]); x1[q - 1] = fmaf(-m1, pr[q], x1[q]); }
        } else {
          if (v0 && pos0 > j) P[j * P2_PS + lane] = x0[0];
          if (v1 && pos1 > j) P[j * P2_PS + lane + 64] = x1[0];
#pragma unroll
          for (int q = 1; q < 32; q++) { x0[q - 1] = x0[q]; x1[q - 1] = x1[q]; }
        }
        x0[31] = 0.0f; x1[31] = 0.0f;
        lds_wait();
        PANEL2_MARK(j, 3);
      }
    }
    if (wave == 0) { pstep[lane] = ps0; pstep[lane + 64] = ps1; }
    const int rem = kb - c0 - 32;
    PANEL2_MARK(c0 + 1, 7);
    lds_barrier();                                        // P, pstep, piv_local, zmask complete; background done
    bg_c0 = -1;
    if (rem <= 0) continue;                               // only the last sub-block can be partial

    // deferred update of the columns [c0 + 32, kb): the next 32 now, the rest in the background
    const uint32_t stag = (uint32_t)(a.k + c0 + 1);
    const int zmask = *s_zmask;
    uint64_t* xsp = a.xbuf + P2_OFF_SUB + (int64_t)((c0 >> 5) & 1) * P2_NREP * P2_SUBREP;
    if (t < P2_R) pstep_bg[t] = pstep[t];
    const int nw = min(32, rem);                          // columns updated now
    if (wave >= 2) {
      // the owners publish their pivot rows (columns [c0, kb)) to every replica
      const int tq = t - 128;
      for (int q = 0; q < 32; q++) {
        const int r = piv_local[q];
        if (r >= 0)
          for (int c = c0 + tq; c < kb; c += 128) {
            const uint32_t bits = __float_as_uint(P[c * P2_PS + r]);
#pragma unroll
            for (int rep = 0; rep < P2_NREP; rep++) put_word(xsp + rep * P2_SUBREP + q * 256 + c, stag, bits);
          }
      }
    } else {
      // waves 0-1, one polling round: LsT[d][q] = l_{q,d} (pivot row q's values in the sub-block's columns) and
      // the raw U words of the next nw columns (Us[col][q], substituted below)
      const uint64_t* xs = xsp + myrep * P2_SUBREP;
      uint32_t vals[16];
      uint32_t need = 0xFFFFu;
#pragma unroll
      for (int i = 8; i < 16; i++) if (((t + P2_R * (i - 8)) & 31) >= nw) { need &= ~(1u << i); vals[i] = 0u; }
      for (long it = 0; it < SPIN_MAX && need; it++) {
        uint64_t wv[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
          const int idx = t + P2_R * (i & 7), q = idx >> 5, d = idx & 31;
          const int col = (i < 8) ? c0 + d : c0 + 32 + d;
          wv[i] = (need >> i & 1u) ? __hip_atomic_load(xs + q * 256 + col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
        }
#pragma unroll
        for (int i = 0; i < 16; i++)
          if ((need >> i & 1u) && (uint32_t)(wv[i] >> 32) == stag) { vals[i] = (uint32_t)wv[i]; need &= ~(1u << i); }
        if (!need || *s_abort) break;
        __builtin_amdgcn_s_sleep(1);
      }
      if (need) { *s_abort = 1; for (int i = 0; i < 16; i++) if (need >> i & 1u) vals[i] = 0u; }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int idx = t + P2_R * i, q = idx >> 5, d = idx & 31;
        LsT[d * 32 + q] = __uint_as_float(vals[i]);
        Us[d * P2_US + q] = __uint_as_float(vals[i + 8]);
      }
    }
    lds_barrier();
    PANEL2_MARK(c0 + 2, 7);
    // next nw columns: forward substitution of the U rows, all four waves (8 columns each), in place
    {
      const int cl = lane >> 3, qb = lane & 7, col = 8 * wave + cl;
      float4* up = (float4*)(Us + col * P2_US + 4 * qb);
      const float4 v = *up;
      float u[4] = {v.x, v.y, v.z, v.w};
      p2_subst_lanes(LsT, zmask, qb, u);
      if (col < nw) *up = make_float4(u[0], u[1], u[2], u[3]);
    }
    lds_barrier();
    PANEL2_MARK(c0 + 3, 7);
    {
      // thread per (row, half of the columns): the row's 32 multipliers stay in registers, the U columns are
      // read as broadcasts; each element gets a -= l_d u_d in increasing d (steps with a zero pivot skipped).
      // A row that was pivot step ps of this sub-block takes U row ps (the same operations on the same
      // operands); rows pivoted earlier keep their values.
      const int r = t & (P2_R - 1), ch = t >> 7;
      const int ps = pstep[r];
      if (ps >= 0) {
        float l[32];
#pragma unroll
        for (int d = 0; d < 32; d++) l[d] = P[(c0 + d) * P2_PS + r];
        // four columns at a time: independent accumulation chains (columns past nw are clamped reads whose
        // results are not stored)
        for (int c4 = 0; c4 < 16; c4 += 4) {
          if (ch * 16 + c4 >= nw) break;
          float acc[4];
          const float4* u4[4];
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int col = min(ch * 16 + c4 + i, nw - 1);
            acc[i] = P[(c0 + 32 + col) * P2_PS + r];
            u4[i] = (const float4*)(Us + col * P2_US);
          }
          if (zmask == 0) {
#pragma unroll
            for (int q4 = 0; q4 < 8; q4++) {
#pragma unroll
              for (int i = 0; i < 4; i++) {
                const float4 v = u4[i][q4];
                acc[i] = fmaf(-l[4 * q4], v.x, acc[i]); acc[i] = fmaf(-l[4 * q4 + 1], v.y, acc[i]);
                acc[i] = fmaf(-l[4 * q4 + 2], v.z, acc[i]); acc[i] = fmaf(-l[4 * q4 + 3], v.w, acc[i]);
              }
            }
          } else {
#pragma unroll
            for (int q4 = 0; q4 < 8; q4++) {
#pragma unroll
              for (int i = 0; i < 4; i++) {
                const float4 v = u4[i][q4];
                const float ue[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 4; e++) {
                  const int d = 4 * q4 + e;
                  const float nv = fmaf(-l[d], ue[e], acc[i]);
                  acc[i] = (zmask >> d & 1) ? acc[i] : nv;
                }
              }
            }
          }
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int col = ch * 16 + c4 + i;
            if (col < nw) P[(c0 + 32 + col) * P2_PS + r] = (ps == 32) ? acc[i] : Us[col * P2_US + min(ps, 31)];
          }
        }
      }
    }
    lds_barrier();
    PANEL2_MARK(c0 + 4, 7);
    if (rem > 32) { bg_c0 = c0; bg_zmask = zmask; }
    PANEL2_MARK(c0, 7);
  }

  // results: rows to their final positions, pivots, info, swap plan
  PANEL2_WGMARK(2);
  if (wave == 0) { spos[lane] = pos0; spos[lane + 64] = pos1; }
  lds_barrier();
  const bool ok = (*s_abort == 0);
  if (ok) {
    // thread (row r, column parity ch): 32 LDS reads in flight, then 32 stores to the row's final position
    const int r = t & (P2_R - 1), ch = t >> 7;
    if (r < rows) {
      float* dst = Ap + (int64_t)spos[r];
      for (int cb = 0; cb < kb; cb += 64) {
        float v[32];
#pragma unroll
        for (int i = 0; i < 32; i++) v[i] = P[min(cb + 2 * i + ch, kb - 1) * P2_PS + r];
#pragma unroll
        for (int i = 0; i < 32; i++) {
          const int c = cb + 2 * i + ch;
          if (c < kb) dst[(int64_t)c * a.lda] = v[i];
        }
      }
    }
    if (g == 0) {
      for (int jj = t; jj < kb; jj += 256) a.ipiv[a.k + jj] = a.k + s_piv[jj] + 1;
      if (t == 0 && *s_info != 0 && a.dinfo[0] == 0) a.dinfo[0] = *s_info;
    }
    if (wave == 0) {
      if (v0 && pos0 != r0 + lane) {
        const int slot = atomicAdd(a.plan, 1);
        a.plan[1 + 2 * slot] = pos0; a.plan[2 + 2 * slot] = r0 + lane;
      }
      if (v1 && pos1 != r0 + lane + 64) {
        const int slot = atomicAdd(a.plan, 1);
        a.plan[1 + 2 * slot] = pos1; a.plan[2 + 2 * slot] = r0 + lane + 64;
      }
    }
  }
  if (t == 0 && *s_abort) a.dinfo[1] = 1;
  PANEL2_WGMARK(3);
}

// Inverse of diagonal block b of L11 into Linv (column-major 32 x 32); threads 0..31 compute, all threads
// of the block (nthr of them, a multiple of 32) help load and take part in the barrier.
__device__ __forceinline__ void trsm_inv_block(const float* __restrict__ A, int64_t lda, int64_t k, int kb,
                                               float* __restrict__ Linv, int b, int c, int nthr, float (*T)[33]) {
  const int rb = 32 * b, rib = min(32, kb - rb);
  for (int idx = c; idx < 32 * 32; idx += nthr) {
    const int q = idx / 32, r = idx - q * 32;
    T[r][q] = (r < rib && q < rib) ? A[(k + rb + r) + (k + rb + q) * lda] : 0.0f;
  }
  __syncthreads();
  if (c < 32) {
    float x[32];
#pragma unroll
    for (int r = 0; r < 32; r++) x[r] = (r == c) ? 1.0f : 0.0f;
#pragma unroll
    for (int r = 1; r < 32; r++) {
      float acc = x[r];
#pragma unroll
      for (int q = 0; q < r; q++) if (q >= c) acc = fmaf(-T[r][q], x[q], acc);
      if (r > c) x[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < 32; r++) Linv[(size_t)b * 1024 + c * 32 + r] = (r < rib && c < rib) ? x[r] : (r == c ? 1.0f : 0.0f);
  }
}

// ------------------------------------------------------------------ row interchanges outside the panel
// plan[0] = number of moved rows, then (dst, src) pairs of panel-local rows: new[dst] = old[src].
// Columns [ra0, ra1) and [rb0, rb1) of A (rows k .. n-1), plus, with do_perm, the composed permutation vector
// (one extra "column").  Workgroups stride over the columns.  The first ninv workgroups instead invert the
// diagonal blocks of L11 for the TRSM (L11's columns are not among the interchanged ones), which saves the
// TRSM path a dependent launch.  zero_plan (optional): the swap-plan counter of the next panel, cleared here
// (its previous readers precede this launch on the stream).  inv_snap (optional, with do_perm): after the
// permutation vector is updated, the workgroup that updated it records its inverse for the rows at positions
// k .. n-1 (inv_snap[perm[p]] = p), the map k_lperm uses to put the L columns of this step in final order.
__global__ __launch_bounds__(256) void k_laswp(float* __restrict__ A, int64_t lda, int64_t k,
                                               int64_t ra0, int64_t ra1, int64_t rb0, int64_t rb1, int do_perm,
                                               const int32_t* __restrict__ plan, int32_t* __restrict__ perm,
                                               int kb, int ninv, float* __restrict__ Linv, int32_t* __restrict__ zero_plan,
                                               int32_t* __restrict__ inv_snap, int64_t n) {
  if (zero_plan != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *zero_plan = 0;
  if ((int)blockIdx.x < ninv) {
    __shared__ float T[32][33];
    trsm_inv_block(A, lda, k, kb, Linv, blockIdx.x, threadIdx.x, 256, T);
    return;
  }
  const int cnt = plan[0];
  const int t = threadIdx.x;
  const int bx = (int)blockIdx.x - ninv, gx = (int)gridDim.x - ninv;
  const int64_t na = ra1 - ra0, nb2 = rb1 - rb0, ncols = na + nb2;
  int dst[4], src[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int e = t + 256 * q;
    dst[q] = (e < cnt) ? plan[1 + 2 * e] : -1;
    src[q] = (e < cnt) ? plan[2 + 2 * e] : -1;
  }
  // LB columns per pass: every load of the pass (unconditional, from row src or, for an unused slot, row 0 of
  // the column) is in flight before the barrier, then the stores.  Passes touch disjoint columns, so only the
  // load -> store order inside a pass needs the barrier.  The permutation vector is "column" ncols.
  constexpr int LB = 8;
  int sl[4];
#pragma unroll
  for (int q = 0; q < 4; q++) sl[q] = dst[q] >= 0 ? src[q] : 0;
  for (int64_t c0 = (int64_t)bx * LB; c0 < ncols + do_perm; c0 += (int64_t)gx * LB) {
    float v[LB][4];
    float* cp[LB];
#pragma unroll
    for (int j = 0; j < LB; j++) {
      const int64_t cc = min(c0 + j, ncols - 1 >= 0 ? ncols - 1 : 0);
      const int64_t col = (cc < na) ? ra0 + cc : rb0 + (cc - na);
      cp[j] = A + col * lda + k;
    }
    if (ncols > 0) {
#pragma unroll
      for (int j = 0; j < LB; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) v[j][q] = cp[j][sl[q]];
    }
    const bool with_perm = do_perm && c0 <= ncols && ncols < c0 + LB;
    int32_t pv[4];
    int32_t* pp = perm + k;
    if (with_perm) {
#pragma unroll
      for (int q = 0; q < 4; q++) pv[q] = pp[sl[q]];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < LB; j++) {
      if (c0 + j < ncols) {
#pragma unroll
        for (int q = 0; q < 4; q++) if (dst[q] >= 0) cp[j][dst[q]] = v[j][q];
      }
    }
    if (with_perm) {
#pragma unroll
      for (int q = 0; q < 4; q++) if (dst[q] >= 0) pp[dst[q]] = pv[q];
      if (inv_snap != nullptr) {
        __syncthreads();   // this workgroup's permutation stores before its reads below
        for (int64_t i = t; i < n - k; i += 256) inv_snap[pp[i]] = (int32_t)(k + i);
      }
    }
  }
}

// ------------------------------------------------------------------ deferred interchanges of the L columns
// Step b's panel writes its L columns in the row order after step b (position p holds original row perm_b[p]);
// the later steps' interchanges, which permute only positions >= (b + 1) * nb, are not applied to them on the
// way.  Here each column block b < nblk - 1 is put in final order at once: position q >= r0 = (b + 1) * nb
// takes the value of position inv_b[perm[q]] (inv_b: the inverse of perm_b, recorded by step b's interchange
// launch), which is the same data movement as the sequence of interchanges.  A workgroup owns `cpw` columns of
// one block: the gather map in LDS, then per column the segment rows r0 .. n-1 staged in LDS by coalesced loads
// and written back in the new order by coalesced stores.  LDS: 8 * (n - r0) bytes.
__global__ __launch_bounds__(1024) void k_lperm(float* __restrict__ A, int64_t lda, int64_t n, int nb, int cpw,
                                                const int32_t* __restrict__ perm, const int32_t* __restrict__ inv_all) {
  extern __shared__ int32_t lp_sm[];
  const int wpb = nb / cpw;
  const int b = (int)blockIdx.x / wpb, cw = (int)blockIdx.x - b * wpb;
  const int64_t r0 = (int64_t)(b + 1) * nb;
  const int len = (int)(n - r0);
  if (len <= 0) return;
  int32_t* map = lp_sm;
  float* col = (float*)(lp_sm + len);
  const int32_t* inv = inv_all + (int64_t)b * n;
  const int t = threadIdx.x;
  for (int i = t; i < len; i += 1024) {
    const int32_t s = inv[perm[r0 + i]] - (int32_t)r0;
    map[i] = min(max(s, 0), len - 1);
  }
  constexpr int U = 8;
  for (int c = 0; c < cpw; c++) {
    float* Ac = A + ((int64_t)b * nb + (int64_t)cw * cpw + c) * lda + r0;
    __syncthreads();   // the map is complete, and the previous column's gathers are done
    for (int i0 = 0; i0 < len; i0 += 1024 * U) {
      float v[U];
#pragma unroll
      for (int u = 0; u < U; u++) { const int i = i0 + u * 1024 + t; v[u] = i < len ? Ac[i] : 0.0f; }
#pragma unroll
      for (int u = 0; u < U; u++) { const int i = i0 + u * 1024 + t; if (i < len) col[i] = v[u]; }
    }
    __syncthreads();
    for (int i = t; i < len; i += 1024) Ac[i] = col[map[i]];
  }
}

// Holds the stream until `target` workgroups of a panel have started (or about a millisecond has passed), so
// that kernels launched behind it, whose LDS request does not fit beside a panel workgroup, take the other CUs
// instead of the CUs the panel is waiting for.
__global__ void k_gate(const int32_t* cnt, int target) {
  if (threadIdx.x == 0)
    for (int it = 0; it < 1024 && __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; it++)
      __builtin_amdgcn_s_sleep(8);
}

__global__ void k_zero_i32(int32_t* p, int cnt) { const int i = blockIdx.x * 256 + threadIdx.x; if (i < cnt) p[i] = 0; }
// The state a factorization starts from, in one launch: the panel exchange area and both swap plans cleared,
// the panel arrival counters, info words 0-1 and the conversion counters cleared, the permutation the identity.
__global__ __launch_bounds__(256) void k_getrf_init(unsigned long long* __restrict__ xbuf, int64_t nx, int32_t* __restrict__ plans,
                                                    int64_t nplan, int32_t* __restrict__ parrive, int64_t narrive,
                                                    int32_t* __restrict__ perm, int64_t n, int32_t* __restrict__ dinfo,
                                                    unsigned long long* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x, st = (int64_t)gridDim.x * 256;
  for (int64_t i = t; i < nx; i += st) xbuf[i] = 0ull;
  for (int64_t i = t; i < nplan; i += st) plans[i] = 0;
  for (int64_t i = t; i < narrive; i += st) parrive[i] = 0;
  for (int64_t i = t; i < n; i += st) perm[i] = (int32_t)i;
  if (t < 2) { dinfo[t] = 0; cnt[t] = 0ull; }
}

// ------------------------------------------------------------------ U12 <- L11^{-1} U12 (unit lower, FP32)
// Two kernels.  k_trsm_inv: one wave per 32x32 diagonal block T of L11 computes T^{-1} column by column
// (lane c: x_c = 1, x_r = sum_{q=c}^{r-1} fmaf(-T_rq, x_q) in increasing q, r = c+1..31), into Linv.
// k_trsm: a workgroup owns 32 columns of U12 (LDS X[kb][33]); per 32-row block b:
//   X_b <- T_b^{-1} X_b         (row r: acc = X_r; acc = fmaf(Tinv_rq, X_q, acc) for q = 0..r-1)
//   X_r -= L_{r,b} X_b, r below  (acc = fmaf(-L_rq, X_q, acc) for q = 0..31)
__global__ __launch_bounds__(64) void k_trsm_inv(const float* __restrict__ A, int64_t lda, int64_t k, int kb, float* __restrict__ Linv) {
  __shared__ float T[32][33];
  trsm_inv_block(A, lda, k, kb, Linv, blockIdx.x, threadIdx.x, 64, T);
}

__global__ __launch_bounds__(256) void k_trsm(float* __restrict__ A, int64_t lda, int64_t n, int64_t k, int kb,
                                              const float* __restrict__ Linv) {
  extern __shared__ float sm[];
  float* X = sm;                           // [kb][33]
  float* Lb = sm + (size_t)kb * 33;        // [kb][33]
  __shared__ float Ti[32][33];
  const int64_t c0 = k + kb + (int64_t)blockIdx.x * 32;
  const int ncols = (int)min((int64_t)32, n - c0);
  const int t = threadIdx.x, cc = t & 31, rg = t >> 5;   // 8 row groups
  for (int idx = t; idx < kb * 32; idx += 256) {
    const int c = idx / kb, r = idx - c * kb;
    X[r * 33 + c] = (c < ncols) ? A[(k + r) + (c0 + c) * lda] : 0.0f;
  }
  const int nblk = (kb + 31) / 32;
  for (int b = 0; b < nblk; b++) {
    const int rb = 32 * b;
    const int rib = min(32, kb - rb);
    const int below = kb - rb;
    __syncthreads();
    for (int idx = t; idx < below * 32; idx += 256) {
      const int c = idx / below, r = idx - c * below;
      Lb[r * 33 + c] = (c < rib) ? A[(k + rb + r) + (k + rb + c) * lda] : 0.0f;
    }
    for (int idx = t; idx < 1024; idx += 256) { const int c = idx >> 5, r = idx & 31; Ti[r][c] = Linv[(size_t)b * 1024 + idx]; }
    __syncthreads();
    float nv[4];
#pragma unroll
    for (int s2 = 0; s2 < 4; s2++) {
      const int r = rg + 8 * s2;
      float acc = X[(rb + r) * 33 + cc];
      for (int q = 0; q < r; q++) acc = fmaf(Ti[r][q], X[(rb + q) * 33 + cc], acc);
      nv[s2] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int s2 = 0; s2 < 4; s2++) { const int r = rg + 8 * s2; if (r < rib) X[(rb + r) * 33 + cc] = nv[s2]; }
    __syncthreads();
    for (int r = rb + 32 + rg; r < kb; r += 8) {
      float acc = X[r * 33 + cc];
#pragma unroll 8
      for (int q = 0; q < 32; q++) acc = fmaf(-Lb[(r - rb) * 33 + q], X[(rb + q) * 33 + cc], acc);
      X[r * 33 + cc] = acc;
    }
  }
  __syncthreads();
  for (int idx = t; idx < kb * 32; idx += 256) {
    const int c = idx / kb, r = idx - c * kb;
    if (c < ncols) A[(k + r) + (c0 + c) * lda] = X[r * 33 + c];
  }
}

// k_trsm_mfma: the same blocked algorithm as k_trsm (X_b <- T_b^{-1} X_b, then X_r -= L_rb X_b for the row
// blocks below) with the 32 x 32 products on v_mfma_f32_32x32x2_f32.  A workgroup of four waves covers 128
// columns of U12, one wave per 32-column strip; the strip's kb x 32 values stay in accumulator layout
// (register a of lane l holds row 8(a/4) + 4(l/32) + a%4 of column l%32 of each 32-row block).  That
// register a is exactly the B operand of k-step a when the k index of step a for lane half h is taken as
// R(h, a) = 8(a/4) + 4h + a%4, so X_b feeds the products straight from registers; the A operand is read in
// the same k order from LDS, where the workgroup first stages T_b^{-1} and -L_rb for every block pair
// (36 blocks of 32 x 33 floats).  Loads and stores of the strips go through LDS so they stay coalesced.
// The matrix unit accumulates the products, so the rounding differs from the sequential FMA order of
// k_trsm (same algorithm, FP32).
constexpr int TM_BLK = 32 * 33;
static size_t trsm_mfma_lds_bytes() { return (size_t)36 * TM_BLK * 4; }
// Columns [cb, ce) of U12 (cb >= k + kb); grid = cdiv(ce - cb, 128).
__global__ __launch_bounds__(256) void k_trsm_mfma(float* __restrict__ A, int64_t lda, int64_t cb, int64_t ce, int64_t k, int kb,
                                                   const float* __restrict__ Linv) {
  extern __shared__ float sm[];
  const int t = threadIdx.x, w = t >> 6, l = t & 63, h = l >> 5, li = l & 31;
  const int64_t n = ce;
  const int64_t c0 = cb + (int64_t)blockIdx.x * 128 + w * 32;        // first column of this wave's strip
  const int64_t ncol = max((int64_t)0, min((int64_t)32, n - c0));
  const int nrb = (kb + 31) / 32;
  float* tile = sm + w * TM_BLK;                                      // staging tile of this wave
  f32x16_t X[8];
  // strip load: lane l reads row l%32 of columns l/32 + 2q (coalesced over rows), transposed through LDS
#pragma unroll
  for (int rb = 0; rb < 8; rb++) {
    if (rb < nrb) {
      // edge tiles (partial rows or columns) take a masked path; the common path has no condition on the
      // loaded values (a masked load gets sunk into a branch and waited for on its own)
      if ((rb + 1) * 32 <= kb && ncol == 32) {
        float v[16];
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = A[(k + rb * 32 + li) + (c0 + h + 2 * q) * lda];
#pragma unroll
        for (int q = 0; q < 16; q++) tile[(h + 2 * q) * 33 + li] = v[q];
      } else {
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int col = h + 2 * q, row = rb * 32 + li;
          const int64_t cc = min(c0 + col, n - 1);
          const float v = A[(k + min(row, kb - 1)) + cc * lda];
          tile[col * 33 + li] = (row < kb && col < ncol) ? v : 0.0f;
        }
      }
      __syncthreads();
#pragma unroll
      for (int a = 0; a < 16; a++) X[rb][a] = tile[li * 33 + 8 * (a >> 2) + 4 * h + (a & 3)];
      __syncthreads();
    } else {
#pragma unroll
      for (int a = 0; a < 16; a++) X[rb][a] = 0.0f;
    }
  }
  // operand blocks: index rb (rb + 1) / 2 + b for b <= rb; [row][col] with stride 33.  One block row at a
  // time: all of its loads are issued before the LDS stores that consume them (a load consumed on the spot
  // is waited for on its own).
#pragma unroll
  for (int rb = 0; rb < 8; rb++) {
    if (rb >= nrb) break;
    float v[8][4];
#pragma unroll
    for (int bb = 0; bb < 8; bb++) {
      if (bb > rb) break;
#pragma unroll
      for (int e4 = 0; e4 < 4; e4++) {
        const int e = t + 256 * e4, col = e >> 5, row = e & 31;
        if (bb == rb) {
          v[bb][e4] = Linv[(size_t)bb * 1024 + col * 32 + row];
        } else {
          const int gr = rb * 32 + row, gc = bb * 32 + col;
          v[bb][e4] = -A[(k + min(gr, kb - 1)) + (k + min(gc, kb - 1)) * lda];
        }
      }
    }
#pragma unroll
    for (int bb = 0; bb < 8; bb++) {
      if (bb > rb) break;
#pragma unroll
      for (int e4 = 0; e4 < 4; e4++) {
        const int e = t + 256 * e4, col = e >> 5, row = e & 31;
        sm[(rb * (rb + 1) / 2 + bb) * TM_BLK + row * 33 + col] = v[bb][e4];
      }
    }
    // rows / columns of L beyond kb (partial last block) are zero; only the last block row can have them
    if ((kb & 31) != 0 && rb == nrb - 1) {
      __syncthreads();
      for (int e = t; e < (rb + 1) * 1024; e += 256) {
        const int bb = e >> 10, col = (e >> 5) & 31, row = e & 31;
        if (bb < rb && (rb * 32 + row >= kb || bb * 32 + col >= kb)) sm[(rb * (rb + 1) / 2 + bb) * TM_BLK + row * 33 + col] = 0.0f;
      }
    }
  }
  __syncthreads();
  auto kidx = [&](int s) { return 8 * (s >> 2) + 4 * h + (s & 3); };
#pragma unroll
  for (int b = 0; b < 8; b++) {
    if (b >= nrb) break;
    {
      const float* M = sm + (b * (b + 1) / 2 + b) * TM_BLK + li * 33;
      f32x16_t y;
#pragma unroll
      for (int a = 0; a < 16; a++) y[a] = 0.0f;
#pragma unroll
      for (int s2 = 0; s2 < 16; s2++) y = __builtin_amdgcn_mfma_f32_32x32x2f32(M[kidx(s2)], X[b][s2], y, 0, 0, 0);
      X[b] = y;
    }
#pragma unroll
    for (int rb = b + 1; rb < 8; rb++) {
      if (rb >= nrb) break;
      const float* M = sm + (rb * (rb + 1) / 2 + b) * TM_BLK + li * 33;
#pragma unroll
      for (int s2 = 0; s2 < 16; s2++) X[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(M[kidx(s2)], X[b][s2], X[rb], 0, 0, 0);
    }
  }
  __syncthreads();
  // strip store through the staging tile
#pragma unroll
  for (int rb = 0; rb < 8; rb++) {
    if (rb >= nrb) break;
#pragma unroll
    for (int a = 0; a < 16; a++) tile[li * 33 + 8 * (a >> 2) + 4 * h + (a & 3)] = X[rb][a];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int col = h + 2 * q, row = rb * 32 + li;
      if (row < kb && col < ncol) A[(k + row) + (c0 + col) * lda] = tile[col * 33 + li];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ operand rounding
__device__ __forceinline__ uint16_t to_bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
// OCP FP8 E4M3 (bias 7, max 448, no infinities, S.1111.111 = NaN), round to nearest even, saturating.
__device__ __forceinline__ uint8_t to_e4m3_rne(float x) {
  const uint32_t sgn = (__float_as_uint(x) >> 24) & 0x80u;
  float a = fabsf(x);
  if (a != a) return (uint8_t)(sgn | 0x7fu);
  if (a > 448.0f) a = 448.0f;
  if (a < 0.015625f) {                       // below 2^-6: subnormal grid m * 2^-9
    const int mq = (int)rintf(a * 512.0f);   // 0..8 (8 becomes the smallest normal)
    return (uint8_t)(sgn | (uint32_t)mq);
  }
  int e; const float f = frexpf(a, &e);      // a = f * 2^e, f in [0.5, 1)
  int ex = e - 1;                             // a = (2f) * 2^ex, 2f in [1, 2)
  int mq = (int)rintf((2.0f * f - 1.0f) * 8.0f);
  if (mq == 8) { mq = 0; ex += 1; }
  return (uint8_t)(sgn | (uint32_t)((ex + 7) << 3) | (uint32_t)mq);
}

template <int FMT> struct FmtInfo;
template <> struct FmtInfo<FMT_FP16> { static constexpr int FB = 16, ES = 2, TOPEXP = 14; static constexpr float MAXV = 65504.0f, MINN = 6.103515625e-05f; };
template <> struct FmtInfo<FMT_BF16> { static constexpr int FB = 16, ES = 2, TOPEXP = 14; static constexpr float MAXV = 3.3895313892515355e+38f, MINN = 1.17549435e-38f; };
template <> struct FmtInfo<FMT_FP8>  { static constexpr int FB = 8,  ES = 1, TOPEXP = 7;  static constexpr float MAXV = 448.0f, MINN = 0.015625f; };
template <> struct FmtInfo<FMT_FP32> { static constexpr int FB = 32, ES = 4, TOPEXP = 0;  static constexpr float MAXV = 3.40282347e+38f, MINN = 1.17549435e-38f; };
template <> struct FmtInfo<FMT_FP16X3> { static constexpr int FB = 16, ES = 2, TOPEXP = 14; static constexpr float MAXV = 65504.0f, MINN = 6.103515625e-05f; };

// Pack one 32-index tile.  Source element (idx, kk) = S[idx*sI + kk*sK], idx < nidx, kk < K.
// Output: for k-block kb16 and lane l, element e lives at ((tile*nkb + kb16)*64 + l)*8 + e and holds
//   idx = tile*32 + l%32,  kk = kb16*16 + (l/32)*8 + e     (16- and 8-bit formats, MFMA 32x32x16)
//   idx = tile*32 + l%32,  kk = kb16*16 + 2*e + l/32       (FP32, eight MFMA 32x32x2 steps)
// exps[tile] = power-of-two exponent s applied before rounding (value = packed * 2^-s).
template <int FMT>
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ S, int64_t sI, int64_t sK, int nidx, int K,
                                              int nkb, uint8_t* __restrict__ out, uint8_t* __restrict__ out_lo,
                                              int32_t* __restrict__ exps, unsigned long long* __restrict__ cnt, int tile_scale) {
  using FI = FmtInfo<FMT>;
  extern __shared__ float tile[];   // [32][K+1]
  __shared__ float red[256];
  const int t = threadIdx.x;
  const int tl = blockIdx.x;
  const int KP = nkb * 16;
  float m = 0.0f;
  // 32 x KP source elements, at most 32 per thread (KP <= 256): every load is issued before the first one is
  // used (indices clamped in range, out-of-range elements zeroed afterwards), so a thread waits for memory
  // once instead of once per element
  float vals[32];
#pragma unroll
  for (int j = 0; j < 32; j++) {
    const int q = t + 256 * j;
    int ii, kk;
    if (sI == 1) { kk = q / 32; ii = q - kk * 32; } else { ii = q / KP; kk = q - ii * KP; }
    const int idx = min(tl * 32 + ii, nidx - 1);
    vals[j] = (q < 32 * KP) ? S[(int64_t)idx * sI + (int64_t)min(kk, K - 1) * sK] : 0.0f;
  }
#pragma unroll
  for (int j = 0; j < 32; j++) {
    const int q = t + 256 * j;
    if (q < 32 * KP) {
      int ii, kk;
      if (sI == 1) { kk = q / 32; ii = q - kk * 32; } else { ii = q / KP; kk = q - ii * KP; }
      const float v = (tl * 32 + ii < nidx && kk < K) ? vals[j] : 0.0f;
      tile[ii * (KP + 1) + kk] = v;
      const float av = fabsf(v);
      m = (av != av || av > m) ? av : m;
    }
  }
  red[t] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) { const float a2 = red[t], b2 = red[t + st]; red[t] = (b2 != b2 || b2 > a2) ? b2 : a2; }
    __syncthreads();
  }
  const float tmax = red[0];
  int s = 0;
  if (tile_scale && FMT != FMT_FP32 && tmax > 0.0f && tmax <= 3.40282347e+38f) {
    int e; frexpf(tmax, &e);               // tmax in [2^(e-1), 2^e)
    s = FI::TOPEXP - (e - 1);
    s = max(-120, min(120, s));
  }
  if (t == 0) exps[tl] = s;
  const float sc = ldexpf(1.0f, s);
  unsigned long long n_over = 0, n_sub = 0;
  // one (k-block, lane) fragment of 8 elements per thread and pass: 8 LDS reads, one vector store per output
  for (int u = t; u < nkb * 64; u += 256) {
    const int kblk = u >> 6, l = u & 63;
    const float* trow = tile + (l & 31) * (KP + 1);
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int kk = (FMT == FMT_FP32) ? (kblk * 16 + 2 * e + (l >> 5)) : (kblk * 16 + (l >> 5) * 8 + e);   // FP16X3: as FP16
      v[e] = trow[kk] * sc;
      const float av = fabsf(v[e]);
      if (av > FI::MAXV) n_over++; else if (av > 0.0f && av < FI::MINN) n_sub++;
    }
    const size_t o = ((size_t)(tl * nkb + kblk) * 64 + l) * 8;
    if (FMT == FMT_FP16 || FMT == FMT_FP16X3) {
      union { uint16_t h[8]; uint4 w; } hi, lo;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const _Float16 hv = (_Float16)v[e];
        hi.h[e] = __builtin_bit_cast(uint16_t, hv);
        if (FMT == FMT_FP16X3) lo.h[e] = __builtin_bit_cast(uint16_t, (_Float16)(v[e] - (float)hv));   // v - hi is exact in FP32
      }
      *(uint4*)((uint16_t*)out + o) = hi.w;
      if (FMT == FMT_FP16X3) *(uint4*)((uint16_t*)out_lo + o) = lo.w;
    } else if (FMT == FMT_BF16) {
      union { uint16_t h[8]; uint4 w; } hi;
#pragma unroll
      for (int e = 0; e < 8; e++) hi.h[e] = to_bf16_rne(v[e]);
      *(uint4*)((uint16_t*)out + o) = hi.w;
    } else if (FMT == FMT_FP8) {
      union { uint8_t h[8]; uint2 w; } q8;
#pragma unroll
      for (int e = 0; e < 8; e++) q8.h[e] = to_e4m3_rne(v[e]);
      *(uint2*)(out + o) = q8.w;
    } else {
      float4* d = (float4*)((float*)out + o);
      d[0] = make_float4(v[0], v[1], v[2], v[3]);
      d[1] = make_float4(v[4], v[5], v[6], v[7]);
    }
  }
  if (n_over) atomicAdd(cnt, n_over);
  if (n_sub) atomicAdd(cnt + 1, n_sub);
}

// ------------------------------------------------------------------ the next panel's columns in one launch
// k_next_prep: for 32 columns per workgroup of the next panel's range [c0, c0 + ncols), everything the
// look-ahead does before their trailing update, in one launch with the strip held in registers:
//   1. when given the swap plan (new[dst] = old[src], panel-local rows), the row interchanges of panel k (the
//      caller may instead have applied them with a wide k_laswp launch, each column's scattered rows being one
//      cache line per element, which 8 workgroups cannot fetch fast enough): the top kb rows
//      are loaded through the map the plan defines, the moved rows below are gathered, and written back once
//      every load of the workgroup has completed;
//   2. U12 <- L11^{-1} U12 with the arithmetic of k_trsm_mfma: wave w owns 32-row block w of the strip in
//      accumulator layout and, for b < w in increasing b, adds (-L_wb) X_b from a tile wave b published in LDS,
//      then forms X_w = T_w^{-1} X_w.  The products and their order per block are k_trsm_mfma's, so the result
//      is the same bits; the chain between blocks is one block's 16 products instead of a whole strip's;
//   3. the packed FP16 / BF16 / FP8 / FP32 (or two-term FP16) U operand of the strip's 32-column tile, with its
//      power-of-two exponent and the conversion counters, as k_pack writes it.
// The diagonal-block inverses are formed per workgroup (one per wave) and the first workgroup stores them for
// the TRSM of the remaining columns.  kb and ncols are multiples of 32, kb <= 256.
constexpr int NP_THREADS = 512;
// block row of block index blk in the order (0,0), (1,0), (1,1), (2,0), ... (rb * (rb + 1) / 2 + bb)
__host__ __device__ constexpr int tri_row(int blk) { int rb = 0; while ((rb + 1) * (rb + 2) / 2 <= blk) rb++; return rb; }
static size_t next_prep_lds_bytes() { return (size_t)36 * TM_BLK * 4 + 64 * 16 * 4 + 256 * 4; }

template <int FMT>
__global__ __launch_bounds__(NP_THREADS) void k_next_prep(float* __restrict__ A, int64_t lda, int64_t k, int kb, int64_t c0,
                                                          const int32_t* __restrict__ plan, int32_t* __restrict__ zero_plan,
                                                          float* __restrict__ Linv_out, uint8_t* __restrict__ outU,
                                                          uint8_t* __restrict__ outUlo, int32_t* __restrict__ expsU,
                                                          unsigned long long* __restrict__ cnt, int tile_scale) {
  using FI = FmtInfo<FMT>;
  extern __shared__ float sm[];
  float* Xex = sm + 36 * TM_BLK;                     // [64 lanes][16]: the block a wave publishes
  int* rmap = (int*)(Xex + 64 * 16);                 // [256]: source row of each of the top kb rows
  __shared__ float red[NP_THREADS / 64];
  __shared__ int s_misc[2];
  const int t = threadIdx.x, w = t >> 6, l = t & 63, h = l >> 5, li = l & 31;
  const int nrb = kb / 32;
  const int64_t cs = c0 + (int64_t)blockIdx.x * 32;  // first column of this strip
  NP_MARK(0);
  if (blockIdx.x == 0 && t == 0 && zero_plan != nullptr) *zero_plan = 0;
  const int npairs = plan != nullptr ? plan[0] : 0;   // no plan: the rows are already interchanged
  // row map of the top kb rows: identity, then the plan's destinations above kb
  for (int r = t; r < kb; r += NP_THREADS) rmap[r] = r;
  __syncthreads();
  for (int e = t; e < npairs; e += NP_THREADS) {
    const int d = plan[1 + 2 * e];
    if (d < kb) rmap[d] = plan[2 + 2 * e];
  }
  __syncthreads();
  NP_MARK(1);
  float* Ac = A + k;                                 // panel-local row r of column c: Ac[r + c * lda]
  // (1) every load first: the moved rows below kb (two rounds of pairs per thread) and the strip rows
  int dlo[2] = {-1, -1};
  float vlo[2][32];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = t + NP_THREADS * q;
    const int d = (e < npairs) ? plan[1 + 2 * e] : -1;
    if (d >= kb) {
      dlo[q] = d;
      const int sr = plan[2 + 2 * e];
#pragma unroll
      for (int c = 0; c < 32; c++) vlo[q][c] = Ac[sr + (cs + c) * lda];
    }
  }
  f32x16_t X;
  if (w < nrb) {
#pragma unroll
    for (int a = 0; a < 16; a++) {
      const int row = 32 * w + 8 * (a >> 2) + 4 * h + (a & 3);
      X[a] = Ac[rmap[row] + (cs + li) * lda];
    }
  }
  // L11 operand blocks into LDS: block (rb, bb), bb < rb, holds -L_rb,bb; the diagonal blocks get T^{-1} below.
  // Element e = t + 512 j of the block sequence lies in block j / 2 (a compile-time index), at offset
  // t + 512 (j % 2); the loads of 24 elements are issued before any of them is stored (a load consumed at once
  // is waited for on its own).  Blocks past kb read block (0, 0) instead, which always lies in the panel, and
  // store nothing.
  {
    const int nblk = nrb * (nrb + 1) / 2;
#pragma unroll
    for (int j0 = 0; j0 < 72; j0 += 24) {
      float v[24];
#pragma unroll
      for (int jj = 0; jj < 24; jj++) {
        const int j = j0 + jj, blk = j >> 1, rb = tri_row(blk), bb = blk - rb * (rb + 1) / 2;
        const int off = t + NP_THREADS * (j & 1), col = off >> 5, row = off & 31;
        const int rbc = blk < nblk ? rb : 0, bbc = blk < nblk ? bb : 0;
        if (bb < rb) v[jj] = A[(k + rbc * 32 + row) + (k + bbc * 32 + col) * lda];
      }
#pragma unroll
      for (int jj = 0; jj < 24; jj++) {
        const int j = j0 + jj, blk = j >> 1, rb = tri_row(blk), bb = blk - rb * (rb + 1) / 2;
        const int off = t + NP_THREADS * (j & 1), col = off >> 5, row = off & 31;
        if (bb < rb && blk < nblk) sm[blk * TM_BLK + row * 33 + col] = -v[jj];
      }
    }
  }
  NP_MARK(7);
  // diagonal-block inverses: wave w inverts T_w (lane c < 32: column c by forward substitution, the order of
  // trsm_inv_block), staged in its own diagonal slot
  if (w < nrb) {
    float* T = sm + (w * (w + 1) / 2 + w) * TM_BLK;
    {
      float tv[16];
#pragma unroll
      for (int i = 0; i < 16; i++) { const int e = l + 64 * i, q = e >> 5, r = e & 31; tv[i] = A[(k + 32 * w + r) + (k + 32 * w + q) * lda]; }
#pragma unroll
      for (int i = 0; i < 16; i++) { const int e = l + 64 * i, q = e >> 5, r = e & 31; T[r * 33 + q] = tv[i]; }
    }
    lds_wait();
    // right-looking: at step q, x[q] has all its terms and every later row takes -T[r][q] x[q].  Per row the
    // terms arrive in increasing q from the same starting value, so each x[r] is the fmaf chain of the
    // row-by-row substitution above, while a step's 31 - q updates are independent (31 dependent steps
    // instead of 496).
    float x[32];
    if (l < 32) {
#pragma unroll
      for (int r = 0; r < 32; r++) x[r] = (r == l) ? 1.0f : 0.0f;
#pragma unroll
      for (int q = 0; q < 31; q++) {
        const float xq = x[q];
        const bool on = q >= l;
#pragma unroll
        for (int r = q + 1; r < 32; r++) {
          const float nv = fmaf(-T[r * 33 + q], xq, x[r]);
          x[r] = on ? nv : x[r];
        }
        __builtin_amdgcn_sched_barrier(0);             // a step's reads stay with it (no hoisting of all 496)
      }
    }
    lds_wait();                                        // every lane's reads of T precede the overwrite
    if (l < 32) {
#pragma unroll
      for (int r = 0; r < 32; r++) T[r * 33 + l] = x[r];
      if (blockIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < 32; r++) Linv_out[(size_t)w * 1024 + l * 32 + r] = x[r];
      }
    }
  }
  __syncthreads();                                     // loads of this workgroup complete; LDS operands ready
  NP_MARK(2);
  // the moved rows below kb
#pragma unroll
  for (int q = 0; q < 2; q++) {
    if (dlo[q] >= 0) {
#pragma unroll
      for (int c = 0; c < 32; c++) Ac[dlo[q] + (cs + c) * lda] = vlo[q][c];
    }
  }
  NP_MARK(3);
  // (2) the triangular solve, block by block
  auto kidx = [&](int s2) { return 8 * (s2 >> 2) + 4 * h + (s2 & 3); };
  for (int b = 0; b < nrb; b++) {
    if (w == b) {
      const float* M = sm + (b * (b + 1) / 2 + b) * TM_BLK + li * 33;
      f32x16_t y;
#pragma unroll
      for (int a = 0; a < 16; a++) y[a] = 0.0f;
#pragma unroll
      for (int s2 = 0; s2 < 16; s2++) y = __builtin_amdgcn_mfma_f32_32x32x2f32(M[kidx(s2)], X[s2], y, 0, 0, 0);
      X = y;
#pragma unroll
      for (int a = 0; a < 16; a++) Xex[l * 16 + a] = X[a];
    }
    __syncthreads();
    if (w > b && w < nrb) {
      const float* M = sm + (w * (w + 1) / 2 + b) * TM_BLK + li * 33;
      float xb[16];
#pragma unroll
      for (int a = 0; a < 16; a++) xb[a] = Xex[l * 16 + a];
#pragma unroll
      for (int s2 = 0; s2 < 16; s2++) X = __builtin_amdgcn_mfma_f32_32x32x2f32(M[kidx(s2)], xb[s2], X, 0, 0, 0);
    }
    __syncthreads();                                   // Xex is rewritten by the next block's owner
  }
  NP_MARK(4);
  // U12 back into A, and the strip staged row-major in LDS (the operand blocks are no longer needed)
  float* St = sm;                                      // [256 rows][33]
  if (w < nrb) {
#pragma unroll
    for (int a = 0; a < 16; a++) {
      const int row = 32 * w + 8 * (a >> 2) + 4 * h + (a & 3);
      Ac[row + (cs + li) * lda] = X[a];
      St[row * 33 + li] = X[a];
    }
  }
  // (3) tile maximum (NaN counts as the largest) and the power-of-two exponent of k_pack
  float m = 0.0f;
  if (w < nrb) {
#pragma unroll
    for (int a = 0; a < 16; a++) { const float av = fabsf(X[a]); m = (av != av || av > m) ? av : m; }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const float o = __shfl_xor(m, off, 64);
    m = (o != o || o > m) ? o : m;
  }
  if (l == 0) red[w] = m;
  __syncthreads();
  if (t == 0) {
    float tm = red[0];
    for (int i = 1; i < NP_THREADS / 64; i++) { const float o = red[i]; tm = (o != o || o > tm) ? o : tm; }
    int sc = 0;
    if (tile_scale && FMT != FMT_FP32 && tm > 0.0f && tm <= 3.40282347e+38f) {
      int e2; frexpf(tm, &e2);
      sc = FI::TOPEXP - (e2 - 1);
      sc = max(-120, min(120, sc));
    }
    s_misc[0] = sc;
    expsU[blockIdx.x] = sc;
  }
  __syncthreads();
  NP_MARK(5);
  const float scl = ldexpf(1.0f, s_misc[0]);
  const int nkb = kb / 16;
  unsigned long long n_over = 0, n_sub = 0;
  for (int u = t; u < nkb * 64; u += NP_THREADS) {
    const int kblk = u >> 6, ln = u & 63;
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int kk = (FMT == FMT_FP32) ? (kblk * 16 + 2 * e + (ln >> 5)) : (kblk * 16 + (ln >> 5) * 8 + e);
      v[e] = St[kk * 33 + (ln & 31)] * scl;
      const float av = fabsf(v[e]);
      if (av > FI::MAXV) n_over++; else if (av > 0.0f && av < FI::MINN) n_sub++;
    }
    const size_t o = ((size_t)((size_t)blockIdx.x * nkb + kblk) * 64 + ln) * 8;
    if (FMT == FMT_FP16 || FMT == FMT_FP16X3) {
      union { uint16_t h[8]; uint4 w4; } hi, lo;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const _Float16 hv = (_Float16)v[e];
        hi.h[e] = __builtin_bit_cast(uint16_t, hv);
        if (FMT == FMT_FP16X3) lo.h[e] = __builtin_bit_cast(uint16_t, (_Float16)(v[e] - (float)hv));
      }
      *(uint4*)((uint16_t*)outU + o) = hi.w4;
      if (FMT == FMT_FP16X3) *(uint4*)((uint16_t*)outUlo + o) = lo.w4;
    } else if (FMT == FMT_BF16) {
      union { uint16_t h[8]; uint4 w4; } hi;
#pragma unroll
      for (int e = 0; e < 8; e++) hi.h[e] = to_bf16_rne(v[e]);
      *(uint4*)((uint16_t*)outU + o) = hi.w4;
    } else if (FMT == FMT_FP8) {
      union { uint8_t h[8]; uint2 w2; } q8;
#pragma unroll
      for (int e = 0; e < 8; e++) q8.h[e] = to_e4m3_rne(v[e]);
      *(uint2*)(outU + o) = q8.w2;
    } else {
      float4* d = (float4*)((float*)outU + o);
      d[0] = make_float4(v[0], v[1], v[2], v[3]);
      d[1] = make_float4(v[4], v[5], v[6], v[7]);
    }
  }
  NP_MARK(6);
  if (n_over) atomicAdd(cnt, n_over);
  if (n_sub) atomicAdd(cnt + 1, n_sub);
}

// ------------------------------------------------------------------ trailing update on MFMA

template <int FMT> struct Frag;
template <> struct Frag<FMT_FP16> { half8_t v;
  __device__ __forceinline__ void load(const uint8_t* p) { v = *(const half8_t*)p; } };
template <> struct Frag<FMT_BF16> { bf16x8_t v;
  __device__ __forceinline__ void load(const uint8_t* p) { v = *(const bf16x8_t*)p; } };
template <> struct Frag<FMT_FP16X3> { half8_t v;
  __device__ __forceinline__ void load(const uint8_t* p) { v = *(const half8_t*)p; } };
template <> struct Frag<FMT_FP8> { long v;
  __device__ __forceinline__ void load(const uint8_t* p) { v = *(const long*)p; } };
template <> struct Frag<FMT_FP32> { f32x8_t v;
  __device__ __forceinline__ void load(const uint8_t* p) { v = *(const f32x8_t*)p; } };

__device__ __forceinline__ void mma(f32x16_t& acc, const Frag<FMT_FP16>& a, const Frag<FMT_FP16>& b) {
  acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.v, b.v, acc, 0, 0, 0); }
__device__ __forceinline__ void mma(f32x16_t& acc, const Frag<FMT_FP16X3>& a, const Frag<FMT_FP16X3>& b) {
  acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.v, b.v, acc, 0, 0, 0); }
__device__ __forceinline__ void mma(f32x16_t& acc, const Frag<FMT_BF16>& a, const Frag<FMT_BF16>& b) {
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.v, b.v, acc, 0, 0, 0); }
__device__ __forceinline__ void mma(f32x16_t& acc, const Frag<FMT_FP8>& a, const Frag<FMT_FP8>& b) {
  acc = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a.v, b.v, acc, 0, 0, 0); }
__device__ __forceinline__ void mma(f32x16_t& acc, const Frag<FMT_FP32>& a, const Frag<FMT_FP32>& b) {
#pragma unroll
  for (int s = 0; s < 8; s++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.v[s], b.v[s], acc, 0, 0, 0);
}

// C[m2 x n2] (FP32, column-major ldc) -= (L21 * U12) using packed operands PL (row tiles) / PU (column tiles).
// Workgroup tile 128 x 128, four waves in 2 x 2, each wave 2 x 2 MFMA tiles of 32 x 32.
// The MFMA A operand is the U12 fragment (its 32 indices are C columns) and the B operand the L21 fragment
// (C rows), so accumulator register a of lane l holds C[row = l%32][col = 8(a/4) + 4(l/32) + a%4].
//
// Tile order (xcdmap): workgroups are dealt to the 8 XCDs in turn (launch index mod 8), each XCD with its own
// L2.  Tile ids are renumbered so that XCD x takes the x-th eighth of an order that walks bands of
// ceil(gx / 8) row tiles column by column: an XCD then keeps its band's packed L21 (about 2 MB at m2 = 16K) in
// its L2 and reads each U12 column tile once per band, instead of cycling through every row tile.  A
// bijection of the tiles: every tile is computed once, with the same arithmetic.
template <int FMT>
__global__ __launch_bounds__(256) void k_gemm(float* __restrict__ C, int64_t ldc, int m2, int n2, int nkb,
                                              const uint8_t* __restrict__ PL, const uint8_t* __restrict__ PU,
                                              const uint8_t* __restrict__ PLlo, const uint8_t* __restrict__ PUlo,
                                              const int32_t* __restrict__ eL, const int32_t* __restrict__ eU, int xcdmap) {
  constexpr int FB = FmtInfo<FMT>::FB;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int bx = blockIdx.x, by = blockIdx.y;
  if (xcdmap) {
    const int gx = gridDim.x, gy = gridDim.y, T = gx * gy;
    const int pid = blockIdx.x + blockIdx.y * gx;
    const int q = T / 8, r = T % 8, xcd = pid % 8;
    const int np = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + pid / 8;
    const int H = (gx + 7) / 8;
    const int band = np / (H * gy), off = np - band * H * gy;
    const int hb = min(H, gx - band * H);
    by = off / hb;
    bx = band * H + (off - by * hb);
  }
  const int rt0 = (bx * 128 + (w & 1) * 64) / 32;
  const int ct0 = (by * 128 + (w >> 1) * 64) / 32;
  f32x16_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int a = 0; a < 16; a++) acc[i][j][a] = 0.0f;
  const size_t kstride = (size_t)64 * FB;
  const uint8_t* pu0 = PU + (size_t)(ct0 + 0) * nkb * kstride + (size_t)lane * FB;
  const uint8_t* pu1 = PU + (size_t)(ct0 + 1) * nkb * kstride + (size_t)lane * FB;
  const uint8_t* pl0 = PL + (size_t)(rt0 + 0) * nkb * kstride + (size_t)lane * FB;
  const uint8_t* pl1 = PL + (size_t)(rt0 + 1) * nkb * kstride + (size_t)lane * FB;
  // C tile values for the epilogue are loaded before the K loop so their HBM latency overlaps the MFMAs.
  // The loads are unconditional (edge tiles read clamped addresses whose values are never stored): a
  // guarded load becomes a branch with a full wait behind it, which would serialise the 64 loads.
  float cv[2][2][16];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int row = min((rt0 + j) * 32 + (lane & 31), m2 - 1);
#pragma unroll
      for (int a = 0; a < 16; a++) {
        const int col = min((ct0 + i) * 32 + 8 * (a >> 2) + 4 * (lane >> 5) + (a & 3), n2 - 1);
        cv[i][j][a] = __builtin_nontemporal_load(C + (int64_t)col * ldc + row);
      }
    }
  constexpr bool X3 = (FMT == FMT_FP16X3);
  const uint8_t* qu0 = X3 ? PUlo + (pu0 - PU) : nullptr;
  const uint8_t* qu1 = X3 ? PUlo + (pu1 - PU) : nullptr;
  const uint8_t* ql0 = X3 ? PLlo + (pl0 - PL) : nullptr;
  const uint8_t* ql1 = X3 ? PLlo + (pl1 - PL) : nullptr;
  Frag<FMT> u0, u1, l0, l1, v0, v1, m0, m1;   // v / m: low parts of the two-term operands
  u0.load(pu0); u1.load(pu1); l0.load(pl0); l1.load(pl1);
  if (X3) { v0.load(qu0); v1.load(qu1); m0.load(ql0); m1.load(ql1); }
  for (int kb16 = 0; kb16 < nkb; kb16++) {
    // next k-block (the last step reloads its own block: the packed operands are padded, always in bounds)
    const size_t o = (size_t)min(kb16 + 1, nkb - 1) * kstride;
    Frag<FMT> nu0, nu1, nl0, nl1, nv0, nv1, nm0, nm1;
    nu0.load(pu0 + o); nu1.load(pu1 + o); nl0.load(pl0 + o); nl1.load(pl1 + o);
    if (X3) { nv0.load(qu0 + o); nv1.load(qu1 + o); nm0.load(ql0 + o); nm1.load(ql1 + o); }
    mma(acc[0][0], u0, l0); mma(acc[0][1], u0, l1);
    mma(acc[1][0], u1, l0); mma(acc[1][1], u1, l1);
    if (X3) {
      // cross terms hi*lo + lo*hi of the two-term operands (lo*lo ~ u16^2 is dropped)
      mma(acc[0][0], u0, m0); mma(acc[0][1], u0, m1);
      mma(acc[1][0], u1, m0); mma(acc[1][1], u1, m1);
      mma(acc[0][0], v0, l0); mma(acc[0][1], v0, l1);
      mma(acc[1][0], v1, l0); mma(acc[1][1], v1, l1);
      v0 = nv0; v1 = nv1; m0 = nm0; m1 = nm1;
    }
    u0 = nu0; u1 = nu1; l0 = nl0; l1 = nl1;
  }
#pragma unroll
  for (int i = 0; i < 2; i++) {
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int ct = ct0 + i, rt = rt0 + j;
      const float sc = ldexpf(1.0f, -(eL[rt] + eU[ct]));
      const int row = rt * 32 + (lane & 31);
      if (row >= m2) continue;
#pragma unroll
      for (int a = 0; a < 16; a++) {
        const int col = ct * 32 + 8 * (a >> 2) + 4 * (lane >> 5) + (a & 3);
        if (col < n2) __builtin_nontemporal_store(fmaf(-acc[i][j][a], sc, cv[i][j][a]), C + (int64_t)col * ldc + row);
      }
    }
  }
}

// ------------------------------------------------------------------ host driver
static int lds_limit_bytes() { return 160 * 1024; }

// XCD-banded tile order of the trailing update (MPSOLVE_GEMM_XCD=0: launch order)
static bool gemm_xcd_order() {
  static const bool on = !(getenv("MPSOLVE_GEMM_XCD") && atoi(getenv("MPSOLVE_GEMM_XCD")) == 0);
  return on;
}


// Packed operand buffers for a trailing update of m2 rows x m2 columns: tiles padded to whole 128-wide
// workgroup tiles.
template <int FMT>
static void pack_buffers(int kb, int64_t m2, Workspace& W) {
  constexpr int ES = FmtInfo<FMT>::ES;
  const int nkb = (kb + 15) / 16;
  const int nT = (int)((cdiv(m2, 128)) * 4);
  const size_t bytes = (size_t)nT * nkb * 64 * 8 * ES;
  W.packL.ensure(*W.tracker, bytes, "packL");
  W.packL2.ensure(*W.tracker, bytes, "packL_b");
  W.packU.ensure(*W.tracker, bytes, "packU");
  if (FMT == FMT_FP16X3) {
    W.packLlo.ensure(*W.tracker, bytes, "packL_lo"); W.packLlo2.ensure(*W.tracker, bytes, "packL_lo_b");
    W.packUlo.ensure(*W.tracker, bytes, "packU_lo");
  }
  W.scaleL.ensure(*W.tracker, nT, "scaleL");
  W.scaleL2.ensure(*W.tracker, nT, "scaleL_b");
  W.scaleU.ensure(*W.tracker, nT, "scaleU");
}

// the packed-L set of step k (k / nb even: the first set, odd: the second)
struct LPack { uint8_t* L; uint8_t* Llo; int32_t* s; };
static LPack lpack(Workspace& W, int64_t k, int nb) {
  if (((k / nb) & 1) == 0) return {W.packL.p, W.packLlo.p, W.scaleL.p};
  return {W.packL2.p, W.packLlo2.p, W.scaleL2.p};
}

// L21 (rows k+kb .. n-1 of the panel columns) into row tiles 0 ..
template <int FMT>
static void pack_L(float* A, int64_t lda, int64_t k, int kb, int64_t m2, bool tile_scale, Workspace& W, const char* pname,
                   LPack lp, hipStream_t stream) {
  constexpr int ES = FmtInfo<FMT>::ES;
  const int nkb = (kb + 15) / 16;
  const int nT = (int)((cdiv(m2, 128)) * 4);
  const size_t lds = (size_t)32 * (nkb * 16 + 1) * sizeof(float);
  hipEvent_t evp = W.prof_begin(stream);
  hipLaunchKernelGGL((k_pack<FMT>), dim3(nT), dim3(256), lds, stream,
                     A + (k + kb) + k * lda, (int64_t)1, lda, (int)m2, kb, nkb, lp.L,
                     FMT == FMT_FP16X3 ? lp.Llo : nullptr, lp.s, W.cnt.p, tile_scale ? 1 : 0);
  LAUNCH_CK();
  W.prof_end(evp, pname, (4.0 + ES) * (double)m2 * kb, 0.0, stream);
}

// U12 columns [cb, ce) (cb - (k + kb) a multiple of 32) into column tiles starting at (cb - k - kb) / 32;
// the tiles up to the next multiple of four past ce are written too (zero padding of the last workgroup tile)
template <int FMT>
static void pack_U(float* A, int64_t lda, int64_t k, int kb, int64_t cb, int64_t ce, bool last, bool tile_scale, Workspace& W,
                   const char* pname) {
  constexpr int ES = FmtInfo<FMT>::ES;
  const int nkb = (kb + 15) / 16;
  const int64_t t0 = (cb - k - kb) / 32;
  const int64_t nidx = ce - cb;
  const int nT = (int)(last ? cdiv(nidx, 128) * 4 : cdiv(nidx, 32));
  const size_t tb = (size_t)nkb * 64 * 8 * ES;
  const size_t lds = (size_t)32 * (nkb * 16 + 1) * sizeof(float);
  hipEvent_t evp = W.prof_begin();
  hipLaunchKernelGGL((k_pack<FMT>), dim3(nT), dim3(256), lds, W.stream,
                     A + k + cb * lda, lda, (int64_t)1, (int)nidx, kb, nkb, W.packU.p + t0 * tb,
                     FMT == FMT_FP16X3 ? W.packUlo.p + t0 * tb : nullptr, W.scaleU.p + t0, W.cnt.p, tile_scale ? 1 : 0);
  LAUNCH_CK();
  W.prof_end(evp, pname, (4.0 + ES) * (double)nidx * kb, 0.0);
}

// One operand pack on its own (the conversion step of the trailing update, exposed for parity checks):
// nidx x K source elements S[idx * sI + kk * sK] into tiles of 32 indices in MFMA fragment order.
void pack_operand(int fmt, bool tile_scale, const float* S, int64_t sI, int64_t sK, int nidx, int K, void* out,
                  void* out_lo, int32_t* exps, unsigned long long* counts, Workspace& W) {
  MP_REQUIRE(S && out && exps && counts, "pack: NULL argument");
  MP_REQUIRE(nidx > 0 && K > 0 && K <= 256, "pack: need nidx > 0 and 0 < K <= 256");
  MP_REQUIRE((sI == 1 && sK >= nidx) || (sK == 1 && sI >= K), "pack: one of the strides must be 1, the other >= the extent");
  MP_REQUIRE(fmt >= FMT_FP16 && fmt <= FMT_FP16X3, "pack: unknown format");
  MP_REQUIRE(fmt != FMT_FP16X3 || out_lo, "pack: the two-term format needs out_lo");
  const int nkb = (K + 15) / 16;
  const int nT = (int)cdiv(nidx, 32);
  const size_t lds = (size_t)32 * (nkb * 16 + 1) * sizeof(float);
  const int ts = tile_scale ? 1 : 0;
  uint8_t* o = (uint8_t*)out; uint8_t* ol = (uint8_t*)out_lo;
  switch (fmt) {
    case FMT_FP16: hipLaunchKernelGGL((k_pack<FMT_FP16>), dim3(nT), dim3(256), lds, W.stream, S, sI, sK, nidx, K, nkb, o, nullptr, exps, counts, ts); break;
    case FMT_BF16: hipLaunchKernelGGL((k_pack<FMT_BF16>), dim3(nT), dim3(256), lds, W.stream, S, sI, sK, nidx, K, nkb, o, nullptr, exps, counts, ts); break;
    case FMT_FP8: hipLaunchKernelGGL((k_pack<FMT_FP8>), dim3(nT), dim3(256), lds, W.stream, S, sI, sK, nidx, K, nkb, o, nullptr, exps, counts, ts); break;
    case FMT_FP32: hipLaunchKernelGGL((k_pack<FMT_FP32>), dim3(nT), dim3(256), lds, W.stream, S, sI, sK, nidx, K, nkb, o, nullptr, exps, counts, ts); break;
    default: hipLaunchKernelGGL((k_pack<FMT_FP16X3>), dim3(nT), dim3(256), lds, W.stream, S, sI, sK, nidx, K, nkb, o, ol, exps, counts, ts); break;
  }
  LAUNCH_CK();
}

// A22[:, cb .. ce) -= L21 * U12[:, cb .. ce)
template <int FMT>
static void gemm_cols(float* A, int64_t lda, int64_t k, int kb, int64_t m2, int64_t cb, int64_t ce, Workspace& W, const char* gname,
                      LPack lp, size_t apart_lds = 0) {
  constexpr int ES = FmtInfo<FMT>::ES, FB = FmtInfo<FMT>::FB;
  const int nkb = (kb + 15) / 16;
  const int64_t t0 = (cb - k - kb) / 32, n2 = ce - cb;
  const size_t tb = (size_t)nkb * 64 * FB;
  // algorithmic traffic of the update: C read + written (8 B/element) plus both packed operands once
  const double gb = 8.0 * (double)m2 * n2 + (FMT == FMT_FP16X3 ? 2.0 : 1.0) * ES * (double)(m2 + n2) * kb;
  const double gf = (FMT == FMT_FP16X3 ? 6.0 : 2.0) * (double)m2 * n2 * kb;
  static bool attr = false;
  if (!attr) {
    HIP_CK(hipFuncSetAttribute((const void*)k_gemm<FMT>, hipFuncAttributeMaxDynamicSharedMemorySize, lds_limit_bytes() - 8192));
    attr = true;
  }
  hipEvent_t ev = W.prof_begin();
  hipLaunchKernelGGL((k_gemm<FMT>), dim3((unsigned)cdiv(m2, 128), (unsigned)cdiv(n2, 128)), dim3(256), apart_lds, W.stream,
                     A + (k + kb) + cb * lda, lda, (int)m2, (int)n2, nkb, lp.L, W.packU.p + t0 * tb,
                     FMT == FMT_FP16X3 ? lp.Llo : nullptr, FMT == FMT_FP16X3 ? W.packUlo.p + t0 * tb : nullptr,
                     lp.s, W.scaleU.p + t0, gemm_xcd_order() ? 1 : 0);
  LAUNCH_CK();
  W.prof_end(ev, gname, gb, gf);
}

// with_L: pack L21 of step k here (on the main stream) before the update; otherwise it was packed ahead
template <int FMT>
static void update_cols(float* A, int64_t lda, int64_t k, int kb, int nb, int64_t m2, int64_t cb, int64_t ce, bool last,
                        bool with_L, bool nxt, bool tile_scale, Workspace& W, bool u_packed = false, size_t apart_lds = 0) {
  // the next panel's columns (look-ahead) are profiled under their own names.  The packed-L pointers are
  // taken after pack_buffers, which allocates the sets on first use.
  if (with_L) pack_buffers<FMT>(kb, m2, W);
  const LPack lp = lpack(W, k, nb);
  if (with_L) pack_L<FMT>(A, lda, k, kb, m2, tile_scale, W, nxt ? "pack_next" : "pack", lp, W.stream);
  if (!u_packed) pack_U<FMT>(A, lda, k, kb, cb, ce, last, tile_scale, W, nxt ? "pack_next" : "pack");
  gemm_cols<FMT>(A, lda, k, kb, m2, cb, ce, W, nxt ? "gemm_next" : "trailing_gemm", lp, apart_lds);
}

// The next panel's columns [cb, cb + ncols) brought up to date except for their trailing update (k_next_prep):
// TRSM and the packed U operand of step k in one launch (and the interchanges too when `plan` is given).
template <int FMT>
static void next_prep(float* A, int64_t lda, int64_t k, int kb, int nb, int64_t m2, int64_t cb, int ncols, const int32_t* plan,
                      int32_t* zero_plan, bool tile_scale, Workspace& W) {
  static bool attr = false;
  if (!attr) {
    HIP_CK(hipFuncSetAttribute((const void*)k_next_prep<FMT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)next_prep_lds_bytes()));
    attr = true;
  }
  pack_buffers<FMT>(kb, m2, W);
  constexpr int ES = FmtInfo<FMT>::ES;
  const int nkb = kb / 16;
  const size_t tb = (size_t)nkb * 64 * 8 * ES;
  const int64_t t0 = (cb - k - kb) / 32;
  hipEvent_t ev = W.prof_begin();
  hipLaunchKernelGGL((k_next_prep<FMT>), dim3((unsigned)(ncols / 32)), dim3(NP_THREADS), next_prep_lds_bytes(), W.stream,
                     A, lda, k, kb, cb, plan, zero_plan, W.linv.p, W.packU.p + t0 * tb,
                     FMT == FMT_FP16X3 ? W.packUlo.p + t0 * tb : nullptr, W.scaleU.p + t0, W.cnt.p, tile_scale ? 1 : 0);
  LAUNCH_CK();
  W.prof_end(ev, "prep_next", (16.0 + 4.0 + ES) * (double)kb * ncols, (double)kb * kb * ncols);
}

static void update_cols_fmt(int fmt, float* A, int64_t lda, int64_t k, int kb, int nb, int64_t m2, int64_t cb, int64_t ce,
                            bool last, bool with_L, bool nxt, bool tile_scale, Workspace& W, bool u_packed = false,
                            size_t apart_lds = 0) {
  switch (fmt) {
    case FMT_FP16: update_cols<FMT_FP16>(A, lda, k, kb, nb, m2, cb, ce, last, with_L, nxt, tile_scale, W, u_packed, apart_lds); break;
    case FMT_BF16: update_cols<FMT_BF16>(A, lda, k, kb, nb, m2, cb, ce, last, with_L, nxt, tile_scale, W, u_packed, apart_lds); break;
    case FMT_FP8:  update_cols<FMT_FP8>(A, lda, k, kb, nb, m2, cb, ce, last, with_L, nxt, tile_scale, W, u_packed, apart_lds); break;
    case FMT_FP16X3: update_cols<FMT_FP16X3>(A, lda, k, kb, nb, m2, cb, ce, last, with_L, nxt, tile_scale, W, u_packed, apart_lds); break;
    default:       update_cols<FMT_FP32>(A, lda, k, kb, nb, m2, cb, ce, last, with_L, nxt, tile_scale, W, u_packed, apart_lds); break;
  }
}
static void next_prep_fmt(int fmt, float* A, int64_t lda, int64_t k, int kb, int nb, int64_t m2, int64_t cb, int ncols,
                          const int32_t* plan, int32_t* zero_plan, bool tile_scale, Workspace& W) {
  switch (fmt) {
    case FMT_FP16: next_prep<FMT_FP16>(A, lda, k, kb, nb, m2, cb, ncols, plan, zero_plan, tile_scale, W); break;
    case FMT_BF16: next_prep<FMT_BF16>(A, lda, k, kb, nb, m2, cb, ncols, plan, zero_plan, tile_scale, W); break;
    case FMT_FP8:  next_prep<FMT_FP8>(A, lda, k, kb, nb, m2, cb, ncols, plan, zero_plan, tile_scale, W); break;
    case FMT_FP16X3: next_prep<FMT_FP16X3>(A, lda, k, kb, nb, m2, cb, ncols, plan, zero_plan, tile_scale, W); break;
    default:       next_prep<FMT_FP32>(A, lda, k, kb, nb, m2, cb, ncols, plan, zero_plan, tile_scale, W); break;
  }
}

// L21 of step k packed ahead on `stream` (the panel stream, right after step k's panel), into step k's set
template <int FMT>
static void pack_L_ahead(float* A, int64_t lda, int64_t k, int kb, int nb, int64_t m2, bool tile_scale, Workspace& W,
                         hipStream_t stream) {
  pack_buffers<FMT>(kb, m2, W);
  pack_L<FMT>(A, lda, k, kb, m2, tile_scale, W, "pack_ahead", lpack(W, k, nb), stream);
}
static void pack_L_ahead_fmt(int fmt, float* A, int64_t lda, int64_t k, int kb, int nb, int64_t m2, bool tile_scale,
                             Workspace& W, hipStream_t stream) {
  switch (fmt) {
    case FMT_FP16: pack_L_ahead<FMT_FP16>(A, lda, k, kb, nb, m2, tile_scale, W, stream); break;
    case FMT_BF16: pack_L_ahead<FMT_BF16>(A, lda, k, kb, nb, m2, tile_scale, W, stream); break;
    case FMT_FP8:  pack_L_ahead<FMT_FP8>(A, lda, k, kb, nb, m2, tile_scale, W, stream); break;
    case FMT_FP16X3: pack_L_ahead<FMT_FP16X3>(A, lda, k, kb, nb, m2, tile_scale, W, stream); break;
    default:       pack_L_ahead<FMT_FP32>(A, lda, k, kb, nb, m2, tile_scale, W, stream); break;
  }
}

constexpr int PLAN_WORDS = 1 + 4 * 256 + 8;

// A refused cooperative launch of a valid panel grid (its size was checked against the CUs) means the runtime
// cannot promise co-residency now (another process holds the device, or its resources are taken): the panel then
// takes an ordinary launch with bounded waits.  Any other failure is an error of the call and is raised.  The
// first refusal's code is kept for the trial record.
static bool coop_refusal(hipError_t e, Workspace& W) {
  (void)hipGetLastError();
  const bool refusal = e == hipErrorCooperativeLaunchTooLarge || e == hipErrorLaunchOutOfResources ||
                       e == hipErrorNotSupported || e == hipErrorOutOfMemory || e == hipErrorNotReady;
  if (!refusal) throw MPError(std::string("getrf: cooperative panel launch failed: ") + hipGetErrorName(e), -5);
  if (W.coop_refused++ == 0) W.coop_error = hipGetErrorName(e);
  return true;
}

static bool panel2_usable(const GetrfArgs& a, const Workspace& W, int64_t k, int kb, bool allow_p2) {
  const int64_t G2 = cdiv(a.n - k, P2_R);
  return allow_p2 && a.n < 65536 && G2 <= std::min<int64_t>(P2_MAXG, W.num_cus) &&
         panel2_lds_bytes(kb) <= (size_t)lds_limit_bytes();
}

// Panel factorization of columns [k, k + kb) on `stream`.  Without `cooperative` (panel2 only) the launch
// may share the device with other kernels.
static void run_panel(const GetrfArgs& a, Workspace& W, int64_t k, int kb, int32_t* plan, hipStream_t stream,
                      bool cooperative, bool allow_p2, int rmin) {
  const int64_t n = a.n;
  const int m = (int)(n - k);
  const int64_t G2 = cdiv(m, P2_R);
  if (panel2_usable(a, W, k, kb, allow_p2)) {
    Panel2Args pa;
    pa.A = a.A; pa.lda = a.lda; pa.k = k; pa.kb = kb; pa.m = m; pa.G = (int)G2;
    pa.xbuf = W.xbuf.p;
    pa.ipiv = a.ipiv; pa.dinfo = W.dinfo.p; pa.plan = plan;
    pa.arrive = W.parrive.p ? W.parrive.p + k / a.nb : nullptr;
    void* args[] = {&pa};
    hipEvent_t e0 = W.prof_begin(stream);
    // The runtime refuses a cooperative launch while another process shares the device; the ordinary launch
    // is then used.  Every wait of the exchange is bounded, so a workgroup that never became resident ends
    // the panel with the timeout flag (reported as an error), never with a hang.
    bool launched = false;
    if (cooperative) {
      const hipError_t e = hipLaunchCooperativeKernel((const void*)k_panel2, dim3((unsigned)G2), dim3(256), args,
                                                      (unsigned)panel2_lds_bytes(kb), stream);
      launched = e == hipSuccess;
      if (!launched) coop_refusal(e, W);
    }
    if (!launched) {
      // overlapped with trailing-update kernels that never wait on anything: workgroups that are not yet
      // resident start as those kernels drain, and the resident ones spin meanwhile
      hipLaunchKernelGGL(k_panel2, dim3((unsigned)G2), dim3(256), panel2_lds_bytes(kb), stream, pa);
      LAUNCH_CK();
    }
    W.prof_end(e0, "panel", 8.0 * (double)m * kb, (double)m * kb * kb, stream);
    return;
  }
  int R = 0;
  for (int cand : {64, 128, 256}) {
    if (cand < rmin) continue;
    const int64_t G = cdiv(m, cand);
    const size_t lds = (size_t)cand * kb * 4 + (size_t)cand * 4;
    if (G <= std::min(128, W.num_cus) && lds <= (size_t)lds_limit_bytes() - 8192) { R = cand; break; }
  }
  MP_REQUIRE(R > 0, "getrf: panel does not fit on the device (n too large for this nb)");
  MP_REQUIRE(cooperative, "getrf: the fallback panel needs a cooperative launch");
  const int G = (int)cdiv(m, R);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(256), 0, stream, plan, 1);
  PanelArgs pa;
  pa.A = a.A; pa.lda = a.lda; pa.k = k; pa.kb = kb; pa.m = m; pa.G = G; pa.R = R;
  pa.xbuf = W.xbuf.p; pa.ipiv = a.ipiv; pa.dinfo = W.dinfo.p; pa.plan = plan;
  void* args[] = {&pa};
  const size_t lds = (size_t)R * kb * 4 + (size_t)R * 4;
  hipEvent_t e0 = W.prof_begin(stream);
  const hipError_t ce = hipLaunchCooperativeKernel((const void*)k_panel, dim3(G), dim3(256), args, (unsigned)lds, stream);
  if (ce != hipSuccess) {
    coop_refusal(ce, W);          // refused: an ordinary launch with bounded waits, as above
    HIP_CK(hipLaunchKernel((const void*)k_panel, dim3(G), dim3(256), args, lds, stream));
  }
  W.prof_end(e0, "panel", 8.0 * (double)m * kb, (double)m * kb * kb, stream);
}

static void laswp_ranges(const GetrfArgs& a, Workspace& W, int64_t k, int kb, int64_t ra0, int64_t ra1,
                         int64_t rb0, int64_t rb1, bool perm, const int32_t* plan, bool with_inv,
                         int32_t* zero_plan, const char* pname = "laswp", int32_t* inv_snap = nullptr,
                         size_t apart_lds = 0) {
  const int64_t ncols = (ra1 - ra0) + (rb1 - rb0) + (perm ? 1 : 0);
  const int ninv = with_inv ? (int)cdiv(kb, 32) : 0;
  if (ncols <= 0 && ninv == 0 && zero_plan == nullptr) return;
  // at most two workgroups per CU, 8 columns per pass each: the panel launched beside this one on the other
  // stream finds room on every CU instead of queueing behind thousands of short workgroups
  const int64_t nwg = std::max<int64_t>(1, std::min<int64_t>(cdiv(ncols, 8), 2 * (int64_t)W.num_cus));
  hipEvent_t e1 = W.prof_begin();
  hipLaunchKernelGGL(k_laswp, dim3((unsigned)(ninv + nwg)), dim3(256), apart_lds, W.stream,
                     a.A, a.lda, k, ra0, ra1, rb0, rb1, perm ? 1 : 0, plan, a.perm, kb, ninv, W.linv.p, zero_plan,
                     perm ? inv_snap : nullptr, a.n);
  LAUNCH_CK();
  W.prof_end(e1, pname, 16.0 * (double)(ncols - (perm ? 1 : 0)) * kb, 0.0);
}

static void trsm_cols(const GetrfArgs& a, Workspace& W, int64_t k, int kb, int64_t cb, int64_t ce, bool fma,
                      const char* pname = "trsm") {
  if (ce <= cb) return;
  hipEvent_t e2 = W.prof_begin();
  if (fma) {
    // the FMA kernel handles the whole trailing width only
    const size_t tl = (size_t)kb * 33 * 4 * 2;
    hipLaunchKernelGGL(k_trsm, dim3(cdiv(ce - (k + kb), 32)), dim3(256), tl, W.stream, a.A, a.lda, ce, k, kb, W.linv.p);
  } else {
    hipLaunchKernelGGL(k_trsm_mfma, dim3((unsigned)cdiv(ce - cb, 128)), dim3(256), trsm_mfma_lds_bytes(), W.stream,
                       a.A, a.lda, cb, ce, k, kb, W.linv.p);
  }
  LAUNCH_CK();
  W.prof_end(e2, pname, 8.0 * (double)kb * (ce - cb), (double)kb * kb * (ce - cb));
}

// Right-looking blocked LU.  With look-ahead (default), step k first brings the next panel's columns up to
// date (interchanges, TRSM, operand packing, update), releases the next panel factorization on the second
// stream, and then updates the remaining columns on the main stream while that panel runs.
void getrf(const GetrfArgs& a, Workspace& W, GetrfInfo* out) {
  const int64_t n = a.n;
  MP_REQUIRE(n > 0 && a.A && a.ipiv && a.perm, "getrf: n > 0 and A, ipiv, perm required");
  MP_REQUIRE(a.lda >= n, "getrf: lda >= n");
  MP_REQUIRE(a.nb >= 1 && a.nb <= 256, "getrf: nb must be in [1, 256]");
  MP_REQUIRE(a.fmt >= FMT_FP16 && a.fmt <= FMT_FP16X3, "getrf: unknown trailing format");
  const int nb = a.nb;
  W.factor_gen++;
  W.coop_refused = 0;
  W.coop_error.clear();
  W.dinfo.ensure(*W.tracker, 4, "getrf_info");
  W.cnt.ensure(*W.tracker, 2, "getrf_counts");
  W.affect.ensure(*W.tracker, 2 * PLAN_WORDS, "swap_plan");
  W.linv.ensure(*W.tracker, (size_t)cdiv(nb, 32) * 1024, "trsm_block_inverses");
  W.parrive.ensure(*W.tracker, (size_t)cdiv(n, nb), "panel_arrivals");
  static bool attr_set = false;
  if (!attr_set) {
    // dynamic limit leaves room for each kernel's static shared memory (at most 8 KB here)
    HIP_CK(hipFuncSetAttribute((const void*)k_panel, hipFuncAttributeMaxDynamicSharedMemorySize, lds_limit_bytes() - 8192));
    HIP_CK(hipFuncSetAttribute((const void*)k_trsm, hipFuncAttributeMaxDynamicSharedMemorySize, lds_limit_bytes() - 8192));
    HIP_CK(hipFuncSetAttribute((const void*)k_panel2, hipFuncAttributeMaxDynamicSharedMemorySize, lds_limit_bytes()));
    HIP_CK(hipFuncSetAttribute((const void*)k_trsm_mfma, hipFuncAttributeMaxDynamicSharedMemorySize, (int)trsm_mfma_lds_bytes()));
    HIP_CK(hipFuncSetAttribute((const void*)k_lperm, hipFuncAttributeMaxDynamicSharedMemorySize, lds_limit_bytes()));
    HIP_CK(hipFuncSetAttribute((const void*)k_laswp, hipFuncAttributeMaxDynamicSharedMemorySize, lds_limit_bytes() - 8192));
    attr_set = true;
  }
  // exchange area for the largest panel, cleared once per factorization (tags restart at 1)
  {
    const int64_t Gmax = std::min<int64_t>(128, cdiv(n, 64));
    const int64_t need = std::max<int64_t>(2 * (Gmax + Gmax * (nb + 1) + (nb + 1)), P2_WORDS);
    W.xbuf.ensure(*W.tracker, (size_t)need, "panel_exchange");
    // Words 0-1 of the info area (info, panel exchange timeout) belong to this factorization; word 2, the
    // substitutions' timeout flag, is owned by whoever runs them (a refinement may refactorize between
    // substitutions) and stays set.  Both swap plans start empty (later each is cleared by the interchange launch
    // of the step before its panel).
    const int64_t most = std::max<int64_t>(need, std::max<int64_t>(n, 2 * PLAN_WORDS));
    hipLaunchKernelGGL(k_getrf_init, dim3((unsigned)std::min<int64_t>(2048, cdiv(most, 256))), dim3(256), 0, W.stream,
                       (unsigned long long*)W.xbuf.p, need, W.affect.p, (int64_t)(2 * PLAN_WORDS), W.parrive.p, cdiv(n, nb),
                       a.perm, n, W.dinfo.p, (unsigned long long*)W.cnt.p);
    LAUNCH_CK();
  }
  int rmin = 64;
  if (const char* e = getenv("MPSOLVE_PANEL_RMIN")) rmin = std::max(64, std::min(256, atoi(e)));
  const bool allow_p2 = !(getenv("MPSOLVE_PANEL") && atoi(getenv("MPSOLVE_PANEL")) == 1);
  const bool use_fma_trsm = getenv("MPSOLVE_TRSM") && atoi(getenv("MPSOLVE_TRSM")) == 1;
  const bool overlap = !(getenv("MPSOLVE_OVERLAP") && atoi(getenv("MPSOLVE_OVERLAP")) == 0) && W.pstream && !use_fma_trsm;
  const bool fused_next = !(getenv("MPSOLVE_FUSED_NEXT") && atoi(getenv("MPSOLVE_FUSED_NEXT")) == 0);
  // The factored L columns are put in final row order once, after the last panel (k_lperm), when a column
  // segment and its gather map fit in LDS; otherwise (or with MPSOLVE_LPERM=0) every step's interchanges are
  // applied to them as they come.  Each step's interchange launch then records the inverse permutation.
  const bool defer_left = n > nb && (int64_t)8 * (n - nb) <= (int64_t)lds_limit_bytes() &&
                          !(getenv("MPSOLVE_LPERM") && atoi(getenv("MPSOLVE_LPERM")) == 0);
  if (defer_left) W.psnap.ensure(*W.tracker, (size_t)cdiv(n, nb) * (size_t)n, "perm_inverse_steps");
  auto snap_of = [&](int64_t k) { return defer_left ? W.psnap.p + (k / nb) * n : nullptr; };
  auto left_of = [&](int64_t k) { return defer_left ? (int64_t)0 : k; };   // end of the left column range
  const bool laswp_apart = !(getenv("MPSOLVE_LASWP_APART") && atoi(getenv("MPSOLVE_LASWP_APART")) == 0);
  // the remaining columns' update is kept off the panel's CUs too once its rows are few enough that it still
  // finishes within the panel on the CUs left over (its time grows as m2^2, a panel's about as m2)
  const int64_t gemm_apart_m = getenv("MPSOLVE_GEMM_APART_M") ? atoll(getenv("MPSOLVE_GEMM_APART_M")) : 12288;
  auto plan_of = [&](int64_t k) { return W.affect.p + ((k / nb) & 1) * PLAN_WORDS; };
  int panels = 0;
  run_panel(a, W, 0, (int)std::min<int64_t>(nb, n), plan_of(0), W.stream, true, allow_p2, rmin);
  panels++;
  // the columns right of the first panel may still be in preparation on another stream
  if (a.cols_ready) HIP_CK(hipStreamWaitEvent(W.stream, a.cols_ready, 0));
  // L21 of the current step was packed ahead on the panel stream (after its panel, while the previous step's
  // remaining update ran on the main stream); steps alternate between two packed-L sets for that
  bool l_ahead = false;
  for (int64_t k = 0; k < n; k += nb) {
    const int kb = (int)std::min<int64_t>(nb, n - k);
    const int64_t m2 = n - k - kb;
    int32_t* plan = plan_of(k);
    if (m2 == 0) { laswp_ranges(a, W, k, kb, 0, left_of(k), 0, 0, true, plan, false, nullptr, "laswp", snap_of(k)); break; }
    const int kbn = (int)std::min<int64_t>(nb, m2);            // next panel width
    const int64_t cN = k + kb + kbn;                            // end of the next panel's columns
    const bool split = overlap && (kbn % 128) == 0 && m2 > kbn && panel2_usable(a, W, k + kb, kbn, allow_p2);
    // the diagonal-block inverses for the TRSM are formed by extra workgroups of the first interchange launch
    // and the next panel's swap-plan counter is cleared by it (its last reader was the previous step's)
    int32_t* next_plan = plan_of(k + kb);
    // the next panel's interchanges, TRSM and U pack in one launch (k_next_prep) when the widths allow
    const bool fused = split && fused_next && (kb % 32) == 0 && (kbn % 32) == 0;
    if (fused) {
      // the interchanges touch one cache line per element: they go wide (a workgroup per column) ahead of the
      // fused TRSM + pack, which then reads contiguous rows
      laswp_ranges(a, W, k, kb, k + kb, cN, 0, 0, false, plan, false, next_plan, "laswp_next");
      next_prep_fmt(a.fmt, a.A, a.lda, k, kb, nb, m2, k + kb, kbn, nullptr, nullptr, a.tile_scale, W);
    } else if (split) {
      laswp_ranges(a, W, k, kb, k + kb, cN, 0, 0, false, plan, !use_fma_trsm, next_plan, "laswp_next");
    } else {
      laswp_ranges(a, W, k, kb, 0, left_of(k), k + kb, n, true, plan, !use_fma_trsm, next_plan, "laswp", snap_of(k));
    }
    if (use_fma_trsm) {
      hipEvent_t e2 = W.prof_begin();
      hipLaunchKernelGGL(k_trsm_inv, dim3(cdiv(kb, 32)), dim3(64), 0, W.stream, a.A, a.lda, k, kb, W.linv.p);
      LAUNCH_CK();
      W.prof_end(e2, "trsm_inv", 0.0, 0.0);
    }
    // the update needs this step's packed L21: when it was packed ahead, wait for that pack only here, so the
    // interchanges and the TRSM above run while it finishes
    if (split) {
      if (!fused) trsm_cols(a, W, k, kb, k + kb, cN, false, "trsm_next");
      if (l_ahead) HIP_CK(hipStreamWaitEvent(W.stream, W.ev_pack, 0));
      update_cols_fmt(a.fmt, a.A, a.lda, k, kb, nb, m2, k + kb, cN, false, !l_ahead, true, a.tile_scale, W, fused);
      HIP_CK(hipEventRecord(W.ev_next, W.stream));
      HIP_CK(hipStreamWaitEvent(W.pstream, W.ev_next, 0));
      run_panel(a, W, k + kb, kbn, plan_of(k + kb), W.pstream, false, allow_p2, rmin);
      panels++;
      HIP_CK(hipEventRecord(W.ev_panel, W.pstream));
      // the next step's L21 is final once its panel is written back (its own interchanges touch other columns)
      const int64_t m2n = m2 - kbn;
      l_ahead = m2n > 0;
      if (l_ahead) {
        pack_L_ahead_fmt(a.fmt, a.A, a.lda, k + kb, kbn, nb, m2n, a.tile_scale, W, W.pstream);
        HIP_CK(hipEventRecord(W.ev_pack, W.pstream));
      }
      // interchanges of the remaining columns: the trailing ones now (the TRSM and the update read them); the
      // factored L columns left of the panel (nothing reads them before the end) either at the end, all steps
      // at once (k_lperm), or after the update, where their scattered one-element-per-line traffic does not
      // meet the next panel's start
      // Kept off the panel's CUs: an LDS request that does not fit beside a panel workgroup, issued once all of
      // the panel's workgroups have started (their polling loads would otherwise queue behind these scattered
      // one-element accesses in the same CU's memory pipeline)
      size_t apart = 0;
      if (laswp_apart) {
        apart = (size_t)std::min<int64_t>(lds_limit_bytes() - 8192,
                                          std::max<int64_t>(0, (int64_t)lds_limit_bytes() - (int64_t)panel2_lds_bytes(kbn) + 256));
        hipLaunchKernelGGL(k_gate, dim3(1), dim3(64), 0, W.stream, W.parrive.p + (k + kb) / nb, (int)cdiv(n - (k + kb), P2_R));
      }
      laswp_ranges(a, W, k, kb, 0, 0, cN, n, defer_left, plan, false, nullptr, "laswp", snap_of(k), apart);
      trsm_cols(a, W, k, kb, cN, n, false);
      update_cols_fmt(a.fmt, a.A, a.lda, k, kb, nb, m2, cN, n, true, false, false, a.tile_scale, W, false,
                      m2 <= gemm_apart_m ? apart : 0);
      if (!defer_left) laswp_ranges(a, W, k, kb, 0, k, 0, 0, true, plan, false, nullptr, "laswp_left");
      HIP_CK(hipStreamWaitEvent(W.stream, W.ev_panel, 0));
    } else {
      trsm_cols(a, W, k, kb, k + kb, n, use_fma_trsm);
      if (l_ahead) HIP_CK(hipStreamWaitEvent(W.stream, W.ev_pack, 0));
      update_cols_fmt(a.fmt, a.A, a.lda, k, kb, nb, m2, k + kb, n, true, !l_ahead, false, a.tile_scale, W);
      run_panel(a, W, k + kb, kbn, plan_of(k + kb), W.stream, true, allow_p2, rmin);
      panels++;
      l_ahead = false;
    }
  }
  if (defer_left) {
    const int64_t nblk = cdiv(n, nb);
    int cpw = 16;
    while (nb % cpw) cpw >>= 1;
    // algorithmic bytes: every L element below its diagonal block read and written once
    double moved = 0.0;
    for (int64_t b = 0; b + 1 < nblk; b++) moved += (double)nb * (double)(n - (b + 1) * nb);
    hipEvent_t e = W.prof_begin();
    hipLaunchKernelGGL(k_lperm, dim3((unsigned)((nblk - 1) * (nb / cpw))), dim3(1024), (size_t)8 * (n - nb), W.stream,
                       a.A, a.lda, n, nb, cpw, a.perm, W.psnap.p);
    LAUNCH_CK();
    W.prof_end(e, "lperm", 8.0 * moved, 0.0);
  }
  // info words and conversion counters, read back through pinned memory where the context provides it (a copy
  // into pageable memory is staged and synchronous)
  int32_t hinfo_s[2] = {0, 0};
  unsigned long long hc_s[2] = {0, 0};
  unsigned long long* hc = W.hpin ? (unsigned long long*)W.hpin : hc_s;
  int32_t* hinfo = W.hpin ? (int32_t*)((char*)W.hpin + 16) : hinfo_s;
  HIP_CK(hipMemcpyAsync(hinfo, W.dinfo.p, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, W.stream));
  HIP_CK(hipMemcpyAsync(hc, W.cnt.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, W.stream));
  HIP_CK(hipStreamSynchronize(W.stream));
  if (out) {
    out->info = hinfo[0]; out->timeout = hinfo[1]; out->overflow = hc[0]; out->subnormal = hc[1]; out->panels = panels;
    out->coop_refused = W.coop_refused;
    out->coop_error = W.coop_error;
  }
  if (hinfo[1]) throw MPError("getrf: panel exchange timed out (workgroups not co-resident?)", -5);
}

}  // namespace mpk
