// cli.cpp — the `mpsolve` command-line driver, built on libmpsolve's C ABI only.
//
// The reference's hot path is reachable only through its process boundary: `mpsolve <cmd> --key value ...`
// writing case/trial/history records and binary vectors under --out, which campaign.py drives
// (src/mpsolve.cu:130-240).  This driver keeps that boundary: the same commands (run, regen, caps, selftest),
// the same flags with the same defaults, the same output file names and formats, and trial records with the
// reference's field names (they come from mp_solve).  Differences, by design of this build:
//   * --modes host times the upload of A and B plus the solve plus the download of X, with the solve itself
//     reading the resident copy (the same bytes cross the bus as in the reference's host mode).
#include "mpsolve.h"

#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// ------------------------------------------------------------------ arguments (src/mpsolve.cu:8-13 semantics)
// `--key value`; a bare `--flag` (next token absent or starting with "--") means "1"; values may be negative.
struct Args {
  std::string cmd;
  std::map<std::string, std::string> kv;
  std::string get(const std::string& k, const std::string& d = "") const { auto it = kv.find(k); return it == kv.end() ? d : it->second; }
  double getd(const std::string& k, double d) const { auto it = kv.find(k); return it == kv.end() ? d : std::stod(it->second); }
  long long geti(const std::string& k, long long d) const { auto it = kv.find(k); return it == kv.end() ? d : std::stoll(it->second); }
  bool has(const std::string& k) const { return kv.count(k) > 0; }
};

// mpsolve: argv[1] is the command.  mpaudit (src/mpaudit.cu:9) takes options only: first = 1, cmd "audit".
Args parse_args(int argc, char** argv, int first = 2) {
  Args a;
  if (first == 2 && argc > 1) a.cmd = argv[1];
  if (first == 1) a.cmd = "audit";
  for (int i = first; i < argc; i++) {
    const std::string s = argv[i];
    if (s.rfind("--", 0) != 0) continue;
    const std::string k = s.substr(2);
    if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) a.kv[k] = argv[++i];
    else a.kv[k] = "1";
  }
  return a;
}

std::vector<std::string> split(const std::string& s, char d) {
  std::vector<std::string> r; std::stringstream ss(s); std::string t;
  while (std::getline(ss, t, d)) if (!t.empty()) r.push_back(t);
  return r;
}

// ------------------------------------------------------------------ small JSON writer (non-finite -> null)
struct Obj {
  std::ostringstream o; bool first = true;
  static std::string esc(const std::string& s) {
    std::string r;
    for (char c : s) { if (c == '"') r += "\\\""; else if (c == '\\') r += "\\\\"; else if (c == '\n') r += "\\n"; else r += c; }
    return r;
  }
  static std::string num(double v) {
    if (!std::isfinite(v)) return "null";
    char b[40]; std::snprintf(b, sizeof b, "%.17g", v); std::string s = b;
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    return s;
  }
  Obj& raw(const std::string& k, const std::string& v) { if (!first) o << ","; first = false; o << "\"" << esc(k) << "\":" << v; return *this; }
  Obj& s(const std::string& k, const std::string& v) { return raw(k, "\"" + esc(v) + "\""); }
  Obj& d(const std::string& k, double v) { return raw(k, num(v)); }
  Obj& i(const std::string& k, long long v) { return raw(k, std::to_string(v)); }
  Obj& b(const std::string& k, bool v) { return raw(k, v ? "true" : "false"); }
  std::string str() const { return "{" + o.str() + "}"; }
};

// {a...} + {b...} -> {a..., b...} (b's keys appended; both are flat objects produced here or by the library)
std::string merge(const std::string& a, const std::string& b) {
  if (a.size() <= 2) return b;
  if (b.size() <= 2) return a;
  return a.substr(0, a.size() - 1) + "," + b.substr(1);
}

void check(int st, const char* what) {
  if (st != MP_OK) throw std::runtime_error(std::string(what) + ": " + mp_last_error());
}

// JSON text from the library, growing the buffer while the library reports it too small (histories of
// instrumented GMRES runs can reach megabytes).
std::string call_json(int (*fn)(mp_ctx*, char*, size_t), mp_ctx* C, size_t cap = 1 << 20) {
  for (;; cap *= 4) {
    std::vector<char> buf(cap);
    const int st = fn(C, buf.data(), buf.size());
    if (st == MP_OK) return std::string(buf.data());
    if (cap >= ((size_t)1 << 32) || std::string(mp_last_error()).find("buffer too small") == std::string::npos) check(st, "json");
  }
}

double wall_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

std::string iso_now() {
  char b[40]; std::time_t t = std::time(nullptr); std::tm tm; gmtime_r(&t, &tm);
  std::strftime(b, sizeof b, "%Y-%m-%dT%H:%M:%SZ", &tm);
  return b;
}

void write_bin(const std::string& path, const void* data, size_t bytes) {
  std::ofstream f(path, std::ios::binary);
  f.write((const char*)data, (std::streamsize)bytes);
  if (!f) throw std::runtime_error("cannot write " + path);
}

// Context and problem handles released in the right order on every exit path.
struct Session {
  mp_ctx* C = nullptr;
  mp_problem* P = nullptr;
  ~Session() { if (P) mp_problem_free(P); if (C) mp_destroy(C); }
};

// The case object passed to mp_problem_generate: the reference's flag names with dashes as underscores.
std::string case_spec(const Args& a, std::string& family_generated) {
  const std::string family = a.get("family", "prescribed");
  family_generated = family;
  Obj c;
  c.s("case_id", a.get("case-id", "case")).i("n", a.geti("n", 1024)).i("nrhs", a.geti("nrhs", 1))
   .s("family", family_generated).s("family_requested", family).s("spectrum", a.get("spectrum", "geometric"))
   .s("dtype", a.get("dtype", "fp64")).s("rhs", a.get("rhs", "random")).d("kappa", a.getd("kappa", 100)).d("kappa_nominal", a.getd("kappa", 100))
   .i("seed", a.geti("seed", 9200)).d("ell", a.getd("ell", 0.25)).d("nugget", a.getd("nugget", 1e-3)).i("dim", a.geti("dim", 3))
   .i("intmax", a.geti("intmax", 255)).i("dyadic_exp", a.geti("dyadic-exp", 8)).i("grcar_k", a.geti("grcar-k", 3))
   .i("global_exp", a.geti("global-exp", 0)).i("rowcol_span", a.geti("rowcol-span", 0));
  return c.str();
}

int cmd_caps(Session& S) {
  std::printf("%s\n", call_json(mp_caps, S.C).c_str());
  return 0;
}

std::string problem_json(mp_problem* P) {
  std::vector<char> buf(1 << 20);
  check(mp_problem_case_json(P, buf.data(), buf.size()), "mp_problem_case_json");
  return std::string(buf.data());
}

// Raw text of a top-level field of a flat JSON object: a string with its quotes, a number, a literal or a
// flat array; "null" when absent.
std::string json_raw(const std::string& j, const std::string& key) {
  const std::string pat = "\"" + key + "\":";
  size_t p = j.find(pat);
  if (p == std::string::npos) return "null";
  p += pat.size();
  size_t e = p;
  if (j[p] == '"') { e = p + 1; while (e < j.size() && (j[e] != '"' || j[e - 1] == '\\')) e++; e++; }
  else if (j[p] == '[') { e = j.find(']', p); e = (e == std::string::npos) ? j.size() : e + 1; }
  else { while (e < j.size() && j[e] != ',' && j[e] != '}') e++; }
  return j.substr(p, e - p);
}

bool read_bin(const std::string& path, void* data, size_t bytes) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.read((char*)data, (std::streamsize)bytes);
  return (bool)f;
}

// mpaudit (src/mpaudit.cu:8-39): the stored problem regenerated in this fresh process with its hashes, an
// optional dump of A, then one record per saved solution of --x-list ("<path> <float32|float64>" lines) with
// FP64 and double-double metrics and forward errors against --xref / --xexact.  Never runs a solver.
int cmd_audit(Session& S, const Args& a, long long n, long long nrhs) {
  const std::string cj = problem_json(S.P);
  Obj head;
  head.s("case_id", a.get("case-id", "case")).raw("regenerated_a_sha256", json_raw(cj, "a_sha256"))
      .raw("regenerated_b_sha256", json_raw(cj, "b_sha256")).raw("regenerated_tile_sha256", json_raw(cj, "tile_sha256"))
      .raw("gen_seconds", json_raw(cj, "gen_seconds")).raw("hash_seconds", json_raw(cj, "hash_seconds"))
      .raw("anorm_inf", json_raw(cj, "anorm_inf")).s("timestamp", iso_now()).s("audit_process_pid", std::to_string((long)getpid()));
  std::printf("%s\n", head.str().c_str());
  std::fflush(stdout);
  const bool f32 = json_raw(cj, "dtype") == "\"fp32\"";
  const size_t nb = (size_t)n * nrhs;
  if (a.has("dump-a")) {
    std::vector<double> ha((size_t)n * n);
    check(mp_problem_copy_out(S.C, S.P, ha.data(), nullptr, nullptr), "mp_problem_copy_out");
    write_bin(a.get("dump-a"), ha.data(), ha.size() * (f32 ? 4 : 8));
  }
  std::vector<double> xr, xe;
  if (a.has("xref")) { xr.resize(nb); if (!read_bin(a.get("xref"), xr.data(), nb * 8)) xr.clear(); }
  if (a.has("xexact")) { xe.resize(nb); if (!read_bin(a.get("xexact"), xe.data(), nb * 8)) xe.clear(); }
  std::ifstream lst(a.get("x-list"));
  std::string path, dt;
  std::vector<char> buf(1 << 16);
  while (lst >> path >> dt) {
    const bool x32 = dt == "float32";
    std::vector<char> raw(nb * (x32 ? 4 : 8));
    std::ifstream f(path, std::ios::binary);
    if (!f) { std::printf("%s\n", Obj().s("x_path", path).s("error", "cannot open").str().c_str()); continue; }
    f.read(raw.data(), (std::streamsize)raw.size());
    if (!f) { std::printf("%s\n", Obj().s("x_path", path).s("error", "short read").str().c_str()); continue; }
    char hex[65];
    check(mp_sha256_hex(raw.data(), raw.size(), hex), "mp_sha256_hex");
    std::vector<double> X(nb);
    if (x32) { const float* p = (const float*)raw.data(); for (size_t i = 0; i < nb; i++) X[i] = p[i]; }
    else std::memcpy(X.data(), raw.data(), nb * 8);
    for (;;) {
      const int st = mp_audit_json(S.C, S.P, X.data(), xr.empty() ? nullptr : xr.data(), xe.empty() ? nullptr : xe.data(),
                                   buf.data(), buf.size());
      if (st == MP_OK) break;
      if (std::string(mp_last_error()).find("buffer too small") == std::string::npos || buf.size() > ((size_t)1 << 30)) check(st, "mp_audit_json");
      buf.resize(buf.size() * 8);
    }
    std::printf("%s\n", merge(Obj().s("x_path", path).s("x_sha256", hex).s("x_dtype", dt).str(), std::string(buf.data())).c_str());
    std::fflush(stdout);
  }
  return 0;
}

int cmd_regen(Session& S) {
  std::printf("%s\n", problem_json(S.P).c_str());
  return 0;
}

// Orthogonality / density check of the generator (src/mpsolve.cu:148): run with an identity spectrum, A is
// orthogonal and max |A^T A - I| measures the generator's orthogonality.
int cmd_selftest(Session& S, long long n, bool f32) {
  if (f32) throw std::runtime_error("selftest checks the FP64 generator: run it with --dtype fp64");
  std::vector<double> h((size_t)n * n);
  check(mp_problem_copy_out(S.C, S.P, h.data(), nullptr, nullptr), "mp_problem_copy_out");
  double maxdev = 0, minabs = INFINITY, anorm = 0; size_t zeros = 0;
  for (long long i = 0; i < n; i++)
    for (long long j = i; j < n; j++) {
      double s = 0;
      for (long long k = 0; k < n; k++) s += h[k + (size_t)i * n] * h[k + (size_t)j * n];
      maxdev = std::max(maxdev, std::fabs(s - (i == j ? 1.0 : 0.0)));
    }
  for (double v : h) { if (v == 0) zeros++; minabs = std::min(minabs, std::fabs(v)); }
  for (long long i = 0; i < n; i++) { double r = 0; for (long long j = 0; j < n; j++) r += std::fabs(h[i + (size_t)j * n]); anorm = std::max(anorm, r); }
  std::printf("{\"selftest\":\"generator\",\"n\":%lld,\"max_orthogonality_deviation\":%.3e,\"zero_entries\":%zu,\"min_abs\":%.3e,\"anorm_inf\":%.6f}\n",
              n, maxdev, zeros, minabs, anorm);
  return 0;
}

// Replaces "host_e2e_s":<value> in a record produced by mp_solve.
std::string set_field(const std::string& rec, const std::string& key, const std::string& val) {
  const std::string pat = "\"" + key + "\":";
  const size_t p = rec.find(pat);
  if (p == std::string::npos) return rec;
  size_t e = p + pat.size();
  while (e < rec.size() && rec[e] != ',' && rec[e] != '}') e++;
  return rec.substr(0, p + pat.size()) + val + rec.substr(e);
}

int cmd_run(Session& S, const Args& a, const std::string& case_id, long long n, long long nrhs) {
  const std::string out = a.get("out", ".");
  mkdir(out.c_str(), 0755);
  const std::string target = a.get("target", a.get("dtype", "fp64"));
  const double u = (target == "fp32") ? std::ldexp(1.0, -24) : std::ldexp(1.0, -53);
  const double mult = a.getd("tau-mult", 100);
  const double tau = mult * u;
  const long long reps = a.geti("reps", 1), warm = a.geti("warmup", 1);
  const std::vector<std::string> modes = split(a.get("modes", "resident"), ',');
  const std::vector<std::string> mids = split(a.get("methods", "native_lu64"), ',');
  long long mem_budget = (long long)a.getd("mem-budget-bytes", 0);
  if (mem_budget <= 0) {   // 80% of free device memory, as the reference records it (src/mpsolve.cu:154)
    const std::string caps = call_json(mp_caps, S.C);
    const size_t p = caps.find("\"free_mem_bytes\":");
    if (p != std::string::npos) mem_budget = (long long)(0.8 * std::stod(caps.substr(p + 17)));
  }
  if (a.get("reference", "auto") != "none") check(mp_problem_reference(S.C, S.P), "mp_problem_reference");

  // case record
  {
    Obj extra;
    extra.s("target", target).d("tau", tau).d("tau_mult", mult).i("mem_budget_bytes", mem_budget).s("timestamp", iso_now());
    std::ofstream f(out + "/case_" + case_id + ".json");
    f << merge(problem_json(S.P), extra.str()) << "\n";
  }
  // B, X_exact, X_ref, optionally A
  const size_t nb = (size_t)n * nrhs;
  const bool f32 = a.get("dtype", "fp64") == "fp32";       // files of A and B in the stored dtype
  {
    std::vector<double> hb(nb), hx(nb);
    std::string cj = problem_json(S.P);
    const bool has_exact = cj.find("\"has_exact_solution\":true") != std::string::npos;
    check(mp_problem_copy_out(S.C, S.P, nullptr, hb.data(), has_exact ? hx.data() : nullptr), "mp_problem_copy_out");
    if (f32) { std::vector<float> b32(hb.begin(), hb.end()); write_bin(out + "/B_" + case_id + ".f32", b32.data(), nb * 4); }
    else write_bin(out + "/B_" + case_id + ".f64", hb.data(), nb * 8);
    if (has_exact) write_bin(out + "/Xexact_" + case_id + ".f64", hx.data(), nb * 8);
    if (mp_problem_copy_ref(S.P, hx.data()) == MP_OK) write_bin(out + "/Xref_" + case_id + ".f64", hx.data(), nb * 8);
    if (a.has("embed-a")) {
      std::vector<double> ha((size_t)n * n);       // large enough for either dtype
      check(mp_problem_copy_out(S.C, S.P, ha.data(), nullptr, nullptr), "mp_problem_copy_out");
      write_bin(out + "/A_" + case_id + (f32 ? ".f32" : ".f64"), ha.data(), ha.size() * (f32 ? 4 : 8));
    }
  }
  std::ofstream trials(out + "/trials_" + case_id + ".jsonl", std::ios::app), hist(out + "/hist_" + case_id + ".jsonl", std::ios::app);
  const long long order_seed = a.geti("order-seed", 1);
  const std::string phase = a.get("phase", "evaluation"), tag = a.get("tag", "");
  const int nm = (int)mids.size();
  std::vector<double> X(nb);
  std::vector<double> hA, hB, Xdl;             // host copies for the host mode (uploaded each trial)
  for (const std::string& mode : modes) {
    const bool host = (mode == "host");
    if (host && hA.empty()) {
      hA.resize((size_t)n * n); hB.resize(nb); Xdl.resize(nb);
      check(mp_problem_copy_out(S.C, S.P, hA.data(), hB.data(), nullptr), "mp_problem_copy_out");
    }
    for (long long rep = -warm; rep < reps; rep++) {
      const bool warmup = rep < 0;
      const long long rot = (order_seed + rep + warm) % std::max(1, nm);
      for (int mi = 0; mi < nm; mi++) {
        const std::string& id = mids[(size_t)((mi + rot) % nm)];
        const std::string xpath = warmup ? "" : "X_" + case_id + "_" + id + "_" + mode + "_r" + std::to_string(rep) + (tag.empty() ? "" : "_" + tag);
        Obj p;
        p.s("case_id", case_id).s("mode", mode).i("rep", rep).b("warmup", warmup).s("phase", phase).s("tag", tag)
         .s("target", target).d("tau_mult", mult).i("nb", a.geti("nb", 256)).i("gmres_m", a.geti("gmres-m", 40))
         .d("gmres_tol", a.getd("gmres-tol", 1e-4)).i("gmres_budget", a.geti("gmres-budget", 480))
         .i("gmres_cycles", a.geti("gmres-cycles", 4)).i("max_outer", a.geti("max-outer", 30))
         .d("cost_ratio_next", a.getd("cost-ratio-next", 3.0)).b("instrument", a.has("instrument")).b("reuse", a.has("reuse"))
         .s("panel_requested", a.get("panel", "cusolver")).s("timestamp", iso_now());
        mp_trial_out T;
        double t_up = 0, t_down = 0;
        void* dA = nullptr; void* dB = nullptr;
        if (host) {
          // the host mode's bus traffic: A and B up, X down (src/mpsolve.cu:197, 212-213)
          const size_t abytes = (size_t)n * n * (f32 ? 4 : 8), bbytes = nb * (f32 ? 4 : 8);   // stored dtype
          check(mp_malloc(S.C, &dA, abytes), "mp_malloc"); check(mp_malloc(S.C, &dB, nb * 8), "mp_malloc");
          const double t0 = wall_now();
          check(mp_memcpy_h2d(S.C, dA, hA.data(), abytes), "upload A"); check(mp_memcpy_h2d(S.C, dB, hB.data(), bbytes), "upload B");
          t_up = wall_now() - t0;
        }
        // the record's x_path names the file written below (after the solve reports the working dtype)
        const int st = mp_solve(S.C, S.P, id.c_str(), merge(p.str(), Obj().s("x_path", xpath.empty() ? "" : xpath + ".f64").str()).c_str(), &T, X.data());
        std::string rec = st == MP_OK ? call_json(mp_last_record_json, S.C) : std::string();
        const bool x32 = rec.find("\"x_dtype\":\"float32\"") != std::string::npos;
        if (host) {
          // X comes back in the working dtype: time a download of that many bytes (the solution itself has
          // already been returned by mp_solve, outside its timed region)
          const double t0 = wall_now();
          check(mp_memcpy_d2h(S.C, Xdl.data(), dB, nb * (x32 ? 4 : 8)), "download X");
          t_down = wall_now() - t0;
          check(mp_free(S.C, dA), "mp_free"); check(mp_free(S.C, dB), "mp_free");
        }
        check(st, "mp_solve");
        if (host && T.device_ready_wall_s >= 0) rec = set_field(rec, "host_e2e_s", Obj::num(t_up + T.device_ready_wall_s + t_down));
        std::string xfile = xpath.empty() ? "" : xpath + (x32 ? ".f32" : ".f64");
        if (x32 && !xfile.empty()) rec = set_field(rec, "x_path", "\"" + xfile + "\"");
        if (!warmup && T.ok && !xfile.empty()) {
          if (x32) { std::vector<float> xf(nb); for (size_t i = 0; i < nb; i++) xf[i] = (float)X[i]; write_bin(out + "/" + xfile, xf.data(), nb * 4); }
          else write_bin(out + "/" + xfile, X.data(), nb * 8);
        }
        trials << rec << "\n"; trials.flush();
        if (!warmup) { hist << call_json(mp_last_hist_json, S.C); hist.flush(); }
        std::fprintf(stderr, "[%s] %s %s rep=%lld %s dev=%.4fs e2e=%.4fs outer=%d inner=%d eta=%.2e omega=%.2e\n", case_id.c_str(), id.c_str(),
                     mode.c_str(), rep, T.terminal_reason, T.device_ready_wall_s, host ? t_up + T.device_ready_wall_s + t_down : -1.0,
                     T.outer_iters, T.inner_iters_total, T.eta_max, T.omega_max);
      }
    }
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
#ifdef MPSOLVE_AUDIT_MAIN
  const Args a = parse_args(argc, argv, 1);
#else
  const Args a = parse_args(argc, argv);
#endif
  if (a.cmd.empty()) { std::fprintf(stderr, "usage: mpsolve run|regen|caps|selftest|audit --key value ...\n"); return 2; }
  try {
    Session S;
    check(mp_create(&S.C, (int)a.geti("device", 0)), "mp_create");
    if (a.cmd == "caps") return cmd_caps(S);
    std::string fam;
    const std::string spec = case_spec(a, fam);
    check(mp_problem_generate(S.C, spec.c_str(), &S.P), "mp_problem_generate");
    const long long n = a.geti("n", 1024), nrhs = a.geti("nrhs", 1);
    if (a.cmd == "regen") return cmd_regen(S);
    if (a.cmd == "selftest") return cmd_selftest(S, n, a.get("dtype", "fp64") == "fp32");
    if (a.cmd == "audit") return cmd_audit(S, a, n, nrhs);
    return cmd_run(S, a, a.get("case-id", "case"), n, nrhs);   // any other command runs trials, as the reference does
  } catch (std::exception& e) {
    std::fprintf(stderr, "FATAL: %s\n", e.what());
    std::printf("%s\n", Obj().s("fatal", e.what()).str().c_str());
    return 1;
  }
}
