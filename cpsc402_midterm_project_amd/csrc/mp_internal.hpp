// mp_internal.hpp — shared host/device infrastructure of libmpsolve (MI355X / gfx950).
//
// Error handling replaces the reference's C++ exceptions at the ABI edge (src/mp_common.cuh:25-28):
// every internal failure throws MPError, and each extern "C" entry point converts it into a negative
// status plus a thread-local message.  Device memory is tracked per context (the reference keeps one
// global tracker, src/mp_common.cuh:70-73).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <map>
#include <stdexcept>
#include <sstream>
#include <chrono>
#include <algorithm>

struct MPError : std::runtime_error {
  int code;
  MPError(const std::string& m, int c = -5) : std::runtime_error(m), code(c) {}
};

#define HIP_CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { std::ostringstream o_; \
  o_ << "HIP error " << hipGetErrorString(e_) << " (" << (int)e_ << ") at " << __FILE__ << ":" << __LINE__ << " in " << #x; \
  throw MPError(o_.str(), -2); } } while (0)
#define MP_REQUIRE(cond, msg) do { if (!(cond)) throw MPError(std::string("invalid argument: ") + (msg), -1); } while (0)

static inline double wall_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------------ device allocation tracking (per context)
struct DevTracker {
  size_t current = 0, peak = 0;
  std::map<void*, size_t> sizes;
  void* alloc(size_t bytes, const char* what) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      std::ostringstream o; o << "hipMalloc failed for " << what << " (" << bytes << " bytes): " << hipGetErrorString(e);
      throw MPError(o.str(), -3);
    }
    sizes[p] = bytes; current += bytes; peak = std::max(peak, current);
    return p;
  }
  void free(void* p) {
    if (!p) return;
    auto it = sizes.find(p);
    if (it != sizes.end()) { current -= it->second; sizes.erase(it); }
    (void)hipFree(p);
  }
};

template <typename T> struct DBuf {
  T* p = nullptr; size_t n = 0; DevTracker* tr = nullptr;
  DBuf() {}
  DBuf(const DBuf&) = delete; DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void alloc(DevTracker& t, size_t n_, const char* what) { release(); tr = &t; n = n_; p = (T*)t.alloc(n_ * sizeof(T), what); }
  // grow-only: keeps the buffer when it is already large enough
  bool ensure(DevTracker& t, size_t n_, const char* what) { if (p && n >= n_) return false; alloc(t, n_, what); return true; }
  void release() { if (p && tr) tr->free(p); p = nullptr; n = 0; }
  void swap(DBuf& o) { std::swap(p, o.p); std::swap(n, o.n); std::swap(tr, o.tr); }
};

// ------------------------------------------------------------------ JSON emitter (finite numbers with %.17g, non-finite -> null)
struct Json {
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
  static std::string num(int64_t v) { return std::to_string(v); }
  static std::string str(const std::string& s) { return "\"" + esc(s) + "\""; }
  static std::string arr(const std::vector<double>& v) { std::string r = "["; for (size_t i = 0; i < v.size(); i++) { if (i) r += ","; r += num(v[i]); } return r + "]"; }
  static std::string arr(const std::vector<int64_t>& v) { std::string r = "["; for (size_t i = 0; i < v.size(); i++) { if (i) r += ","; r += num(v[i]); } return r + "]"; }
  static std::string arr(const std::vector<std::string>& v) { std::string r = "["; for (size_t i = 0; i < v.size(); i++) { if (i) r += ","; r += str(v[i]); } return r + "]"; }
  Json& kv(const std::string& k, const std::string& raw) { if (!first) o << ","; first = false; o << "\"" << esc(k) << "\":" << raw; return *this; }
  Json& s(const std::string& k, const std::string& v) { return kv(k, str(v)); }
  Json& d(const std::string& k, double v) { return kv(k, num(v)); }
  Json& i(const std::string& k, int64_t v) { return kv(k, num(v)); }
  Json& b(const std::string& k, bool v) { return kv(k, v ? "true" : "false"); }
  std::string obj() const { return "{" + o.str() + "}"; }
};

// ------------------------------------------------------------------ SHA-256 (FIPS 180-4), for the input/output hashes in records
std::string sha256_hex(const void* data, size_t len);

// ------------------------------------------------------------------ launch helpers
static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }
#define LAUNCH_CK() HIP_CK(hipGetLastError())
