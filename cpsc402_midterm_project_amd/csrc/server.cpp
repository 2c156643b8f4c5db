// server.cpp — continuous-batching request loop on one GPU, built on the library's C ABI.
//
// The reference runs one system per process invocation (src/mpsolve.cu:130-240); a serving loop is this build's
// extension of that path.  Requests (matrix id, FP64 A and B resident in device memory, a host buffer for X)
// are queued by size and served by two workers, each with its own library context (own HIP streams and
// workspace), so their work overlaps on the device:
//   * systems with n <= 512 and nrhs <= 16 are collected while the previous batch runs and solved together by
//     one persistent batched launch (mp_batch_solve_ptrs): requests that arrive while others are on the GPU
//     join the next launch as soon as it starts;
//   * larger systems are solved one at a time by the configured method on the whole GPU (default: two-term FP16
//     LU + FGMRES-IR with double-double residuals, escalating to FP32 and then FP64 factors when a stage stops
//     short of tau); a matrix id seen before reuses its problem and the factors cached with it
//     (mp_problem_set_rhs + "reuse", starting at the stage that served it last), so a repeat costs only the
//     refinement.
// Admission: the bytes the server holds (cached factors, the running large solve, batch buffers) stay within
// the HBM budget; the least recently used cached factors are released first, and a large request waits at
// the head of its queue until its bytes fit.
#include "mpsolve.h"
#include "mp_internal.hpp"

#include <hip/hip_runtime_api.h>

#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

void mp_set_last_error(const std::string& msg);

namespace {

double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

struct Req {
  uint64_t id, matrix;
  int64_t n; int32_t nrhs;
  const double* A; const double* B; double* X;
  double arrival, start = 0;
};

int terminal_code(const char* s) {
  const std::string t(s);
  if (t == "converged") return 0;
  if (t == "nonfinite") return 1;
  if (t == "max_outer") return 2;
  if (t == "stagnation") return 3;
  if (t == "singular_factor") return 4;
  return 6;
}

std::string json_escape(const std::string& s) {
  std::string r;
  for (char c : s) { if (c == '"' || c == '\\') r += '\\'; r += c; }
  return r;
}

}  // namespace

struct mp_server {
  int device = 0;
  mp_ctx* big = nullptr;
  mp_ctx* small = nullptr;
  std::string method = "policy_mx16x3_ladder";
  std::string solve_params, batch_params;
  size_t budget = (size_t)200 << 30;
  int max_batch = 4096;
  bool sched_fifo = false;      // large requests in arrival order instead of smallest first
  double max_wait_s = 0.6;      // a large request waiting longer than this goes next

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req> qbig, qsmall;
  std::deque<mp_server_completion> done;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread tbig, tsmall;

  // factor cache of the large path: matrix id -> problem over the caller's A, factors kept by "reuse"
  struct Entry { mp_problem* P; size_t bytes; uint64_t last; };
  std::unordered_map<uint64_t, Entry> cache;
  size_t cache_bytes = 0, big_bytes = 0, small_bytes = 0;
  uint64_t tick = 0;
  // statistics
  int64_t n_small = 0, n_big = 0, n_hits = 0, n_evict = 0, n_batches = 0, n_errors = 0;
  std::string last_error;

  size_t held() const { return cache_bytes + big_bytes + small_bytes; }

  void finish(const Req& r, int path, int hit, int term, int outer, double eta, double omega, double t_done) {
    mp_server_completion c{};
    c.request_id = r.id; c.terminal = term; c.path = path; c.cache_hit = hit; c.outer_iters = outer;
    c.eta_max = eta; c.omega_max = omega;
    c.queue_s = r.start - r.arrival; c.service_s = t_done - r.start; c.latency_s = t_done - r.arrival;
    done.push_back(c);
  }

  // bytes a large request holds while it runs: factors (FP32, or FP64 at the last stage of a ladder), the
  // FGMRES basis (m = 30) and the solve vectors
  static size_t big_need(int64_t n, int nrhs) {
    return (size_t)n * n * 8 + (size_t)n * (2 * 30 + 2) * 8 + (size_t)n * nrhs * 8 * 8;
  }

  void evict_lru_locked() {
    auto victim = cache.end();
    for (auto it = cache.begin(); it != cache.end(); ++it)
      if (victim == cache.end() || it->second.last < victim->second.last) victim = it;
    if (victim == cache.end()) return;
    mp_problem_free(victim->second.P);
    cache_bytes -= victim->second.bytes;
    cache.erase(victim);
    n_evict++;
  }

  void run_big() {
    (void)hipSetDevice(device);
    for (;;) {
      Req r;
      mp_problem* P = nullptr;
      bool hit = false;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (stop && qbig.empty()) return;
          if (!qbig.empty()) {
            // next large request: the smallest system first (its solve time grows like n^3), unless the oldest
            // one has waited longer than max_wait_s, which then goes first
            size_t pick = 0;
            if (!(sched_fifo || now_s() - qbig.front().arrival > max_wait_s))
              for (size_t q = 1; q < qbig.size(); q++) if (qbig[q].n < qbig[pick].n) pick = q;
            const Req& h = qbig[pick];
            auto it = cache.find(h.matrix);
            const size_t full = big_need(h.n, h.nrhs);
            const size_t need = full - (it != cache.end() ? std::min(full, it->second.bytes) : 0);   // beyond what it holds
            while (held() + need > budget && !cache.empty()) {
              // release other matrices' factors, least recently used first (never the one about to be reused)
              auto keep = cache.find(h.matrix);
              if (keep != cache.end() && cache.size() == 1) break;
              if (keep != cache.end()) keep->second.last = ~0ull;
              evict_lru_locked();
              if (keep != cache.end()) keep->second.last = tick;
            }
            if (held() + need <= budget || (big_bytes == 0 && small_bytes == 0)) {
              r = h; qbig.erase(qbig.begin() + (std::ptrdiff_t)pick);   // the others keep their arrival order
              big_bytes = full;
              auto it2 = cache.find(r.matrix);
              if (it2 != cache.end()) {
                P = it2->second.P; hit = true;
                cache_bytes -= it2->second.bytes;     // counted in big_bytes while it runs
                cache.erase(it2);
              }
              break;
            }
          }
          cv_work.wait(lk);
        }
      }
      r.start = now_s();
      int st = MP_OK;
      if (hit) st = mp_problem_set_rhs(big, P, r.nrhs, r.B);
      else st = mp_problem_from_device(big, r.n, r.nrhs, r.A, r.B, 1, "{\"source\":\"server\"}", &P);
      mp_trial_out out{};
      if (st == MP_OK) st = mp_solve(big, P, method.c_str(), solve_params.c_str(), &out, r.X);
      const double t_done = now_s();
      std::lock_guard<std::mutex> lk(mu);
      big_bytes = 0;
      n_big++;
      if (hit) n_hits++;
      if (st != MP_OK) {
        n_errors++; last_error = mp_last_error();
        if (P) mp_problem_free(P);
        finish(r, 1, hit ? 1 : 0, 5, 0, -1, -1, t_done);
      } else {
        uint64_t fb = (uint64_t)r.n * r.n * (std::strcmp(out.final_stage, "cs64") == 0 ? 8 : 4);   // factors kept
        (void)mp_problem_factor_bytes(P, &fb);
        cache[r.matrix] = Entry{P, fb, ++tick};
        cache_bytes += fb;
        finish(r, 1, hit ? 1 : 0, out.ok ? terminal_code(out.terminal_reason) : 5, out.outer_iters, out.eta_max,
               out.omega_max, t_done);
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }

  void run_small() {
    (void)hipSetDevice(device);
    std::vector<Req> batch;
    for (;;) {
      size_t need = 0;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !qsmall.empty(); });
        if (stop && qsmall.empty()) return;
        batch.clear();
        while (!qsmall.empty() && (int)batch.size() < max_batch) {
          const Req& h = qsmall.front();
          const size_t b = (size_t)h.n * (h.n + 32) * 8 + (size_t)h.n * h.nrhs * 24;
          if (!batch.empty() && held() + need + b > budget) break;
          need += b;
          batch.push_back(h); qsmall.pop_front();
        }
        small_bytes = need;
      }
      const double t0 = now_s();
      const int cnt = (int)batch.size();
      std::vector<int32_t> n(cnt), nr(cnt);
      std::vector<const double*> A(cnt), B(cnt);
      std::vector<double*> X(cnt);
      size_t tot = 0;
      for (int i = 0; i < cnt; i++) { batch[i].start = t0; n[i] = (int32_t)batch[i].n; nr[i] = batch[i].nrhs; tot += (size_t)n[i] * nr[i]; }
      void* dx = nullptr;
      std::vector<mp_batch_result> res(cnt);
      int st = mp_malloc(small, &dx, tot * 8);
      std::vector<double> hx;
      if (st == MP_OK) {
        size_t off = 0;
        for (int i = 0; i < cnt; i++) { A[i] = batch[i].A; B[i] = batch[i].B; X[i] = (double*)dx + off; off += (size_t)n[i] * nr[i]; }
        st = mp_batch_solve_ptrs(small, cnt, n.data(), nr.data(), A.data(), B.data(), X.data(), batch_params.c_str(), res.data());
        if (st == MP_OK) { hx.resize(tot); st = mp_memcpy_d2h(small, hx.data(), dx, tot * 8); }
        (void)mp_free(small, dx);
      }
      if (st == MP_OK) {
        size_t off = 0;
        for (int i = 0; i < cnt; i++) {
          std::memcpy(batch[i].X, hx.data() + off, (size_t)n[i] * nr[i] * 8);
          off += (size_t)n[i] * nr[i];
        }
      }
      const double t_done = now_s();
      std::lock_guard<std::mutex> lk(mu);
      small_bytes = 0;
      n_small += cnt; n_batches++;
      if (st != MP_OK) { n_errors += cnt; last_error = mp_last_error(); }
      for (int i = 0; i < cnt; i++) {
        if (st != MP_OK) finish(batch[i], 0, 0, 5, 0, -1, -1, t_done);
        else finish(batch[i], 0, 0, res[i].terminal, res[i].outer_iters, res[i].eta_max, res[i].omega_max, t_done);
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

namespace {
template <class F> int sguard(F&& f) {
  try { f(); return MP_OK; }
  catch (MPError& e) { mp_set_last_error(e.what()); return e.code; }
  catch (std::exception& e) { mp_set_last_error(e.what()); return MP_ERR_INTERNAL; }
}

// flat JSON values of the server parameters (numbers and strings)
std::string jget(const std::string& js, const std::string& key, const std::string& def) {
  const std::string pat = "\"" + key + "\"";
  size_t p = js.find(pat);
  if (p == std::string::npos) return def;
  p = js.find(':', p + pat.size());
  if (p == std::string::npos) return def;
  p++;
  while (p < js.size() && isspace((unsigned char)js[p])) p++;
  if (p < js.size() && js[p] == '"') { const size_t e = js.find('"', p + 1); return js.substr(p + 1, e - p - 1); }
  if (p < js.size() && js[p] == '{') {                 // nested object: returned verbatim
    int depth = 0; size_t e = p;
    for (; e < js.size(); e++) { if (js[e] == '{') depth++; else if (js[e] == '}' && --depth == 0) break; }
    return js.substr(p, e - p + 1);
  }
  size_t e = p;
  while (e < js.size() && js[e] != ',' && js[e] != '}') e++;
  return js.substr(p, e - p);
}
}  // namespace

extern "C" {

double mp_server_clock(void) { return now_s(); }

int mp_server_create(int device, const char* params_json, mp_server** out) {
  return sguard([&] {
    MP_REQUIRE(out, "server pointer is NULL");
    const std::string js = params_json ? params_json : "{}";
    std::unique_ptr<mp_server> S(new mp_server());
    S->device = device;
    S->method = jget(js, "method", S->method);
    S->budget = (size_t)std::stod(jget(js, "hbm_budget_bytes", std::to_string((double)S->budget)));
    S->max_batch = std::stoi(jget(js, "max_batch", "4096"));
    S->sched_fifo = jget(js, "schedule", "shortest_first") == "fifo";
    S->max_wait_s = std::stod(jget(js, "max_wait_s", "0.6"));
    MP_REQUIRE(S->max_batch >= 1, "max_batch must be positive");
    std::string sp = jget(js, "solve", "{}");
    std::string bp = jget(js, "batch", "{}");
    const std::string tau = jget(js, "tau", "1e-15");
    // every request is refined to the same backward-error target by both paths
    auto with = [&](std::string o, const std::string& extra) {
      while (!o.empty() && isspace((unsigned char)o.back())) o.pop_back();
      if (o.size() <= 2) return "{" + extra + "}";
      return o.substr(0, o.size() - 1) + "," + extra + "}";
    };
    S->solve_params = with(sp, "\"tau\":" + tau + ",\"stop_on\":\"eta\",\"reuse\":true,\"resume_ladder\":true");
    // no launch-shape tuning inside the request loop unless asked for (a new batch signature would add its tuning
    // time to the latency of the requests in that batch)
    const std::string tune = bp.find("\"tune\"") == std::string::npos ? ",\"tune\":\"off\"" : "";
    S->batch_params = with(bp, "\"tau\":" + tau + ",\"stop_on\":\"eta\"" + tune);
    if (mp_create(&S->big, device) != MP_OK) throw MPError(std::string("mp_create: ") + mp_last_error(), MP_ERR_NO_DEVICE);
    if (mp_create(&S->small, device) != MP_OK) {
      const std::string m = mp_last_error();
      mp_destroy(S->big);
      throw MPError("mp_create: " + m, MP_ERR_NO_DEVICE);
    }
    mp_server* p = S.get();
    p->tbig = std::thread([p] { p->run_big(); });
    p->tsmall = std::thread([p] { p->run_small(); });
    *out = S.release();
  });
}

int mp_server_submit(mp_server* S, uint64_t matrix_id, int64_t n, int32_t nrhs, const double* A, const double* B,
                     double* X_host, double arrival_s, uint64_t* request_id) {
  return sguard([&] {
    MP_REQUIRE(S && A && B && X_host, "NULL argument");
    MP_REQUIRE(n > 0 && nrhs > 0, "n and nrhs must be positive");
    Req r{};
    r.matrix = matrix_id; r.n = n; r.nrhs = nrhs; r.A = A; r.B = B; r.X = X_host;
    r.arrival = arrival_s >= 0 ? arrival_s : now_s();
    std::lock_guard<std::mutex> lk(S->mu);
    MP_REQUIRE(!S->stop, "server is shutting down");
    r.id = S->next_id++;
    if (n <= 512 && nrhs <= 16) S->qsmall.push_back(r); else S->qbig.push_back(r);
    if (request_id) *request_id = r.id;
    S->cv_work.notify_all();
  });
}

int mp_server_poll(mp_server* S, mp_server_completion* out, int32_t max, double wait_s, int32_t* count) {
  return sguard([&] {
    MP_REQUIRE(S && out && count && max > 0, "bad argument");
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->done.empty() && wait_s > 0)
      S->cv_done.wait_for(lk, std::chrono::duration<double>(wait_s), [&] { return !S->done.empty(); });
    int k = 0;
    while (k < max && !S->done.empty()) { out[k++] = S->done.front(); S->done.pop_front(); }
    *count = k;
  });
}

int mp_server_stats_json(mp_server* S, char* json, size_t cap) {
  return sguard([&] {
    MP_REQUIRE(S && json && cap > 0, "bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    char b[1024];
    std::snprintf(b, sizeof b,
                  "{\"small_requests\":%lld,\"large_requests\":%lld,\"batches\":%lld,\"cache_hits\":%lld,\"evictions\":%lld,"
                  "\"errors\":%lld,\"cached_matrices\":%zu,\"cache_bytes\":%zu,\"hbm_budget_bytes\":%zu,\"queued_small\":%zu,"
                  "\"queued_large\":%zu,\"method\":\"%s\",\"last_error\":\"%s\"}",
                  (long long)S->n_small, (long long)S->n_big, (long long)S->n_batches, (long long)S->n_hits,
                  (long long)S->n_evict, (long long)S->n_errors, S->cache.size(), S->cache_bytes, S->budget,
                  S->qsmall.size(), S->qbig.size(), json_escape(S->method).c_str(), json_escape(S->last_error).c_str());
    MP_REQUIRE(std::strlen(b) + 1 <= cap, "output buffer too small");
    std::memcpy(json, b, std::strlen(b) + 1);
  });
}

int mp_server_destroy(mp_server* S) {
  return sguard([&] {
    if (!S) return;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->tbig.joinable()) S->tbig.join();
    if (S->tsmall.joinable()) S->tsmall.join();
    for (auto& kv : S->cache) mp_problem_free(kv.second.P);
    S->cache.clear();
    if (S->big) mp_destroy(S->big);
    if (S->small) mp_destroy(S->small);
    delete S;
  });
}

}  // extern "C"
