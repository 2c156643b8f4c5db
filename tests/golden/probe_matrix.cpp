// probe_matrix: writes the system of the reference's capability probe (src/probe/probe_cusolver.cu:31-35) so
// the tests can compare the probe's published error magnitudes (logs/probe_cusolver_2048.jsonl) with this
// library's solves of the same matrix.  Test infrastructure: a std::mt19937_64(42) stream through
// std::normal_distribution<double>(0, 1), A filled in column-major order, 2 sqrt(n) added to the diagonal,
// then b.  libstdc++ provides both distributions, so the values are the probe's.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

int main(int argc, char** argv) {
  if (argc != 4) { std::fprintf(stderr, "usage: probe_matrix n A.f64 b.f64\n"); return 2; }
  const long n = std::atol(argv[1]);
  std::mt19937_64 rng(42);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<double> A((size_t)n * n), b(n);
  for (auto& v : A) v = nd(rng);
  for (long i = 0; i < n; i++) A[i + (size_t)i * n] += 2.0 * std::sqrt((double)n);
  for (auto& v : b) v = nd(rng);
  FILE* fa = std::fopen(argv[2], "wb");
  FILE* fb = std::fopen(argv[3], "wb");
  if (!fa || !fb) return 1;
  const bool ok = std::fwrite(A.data(), 8, A.size(), fa) == A.size() && std::fwrite(b.data(), 8, b.size(), fb) == b.size();
  std::fclose(fa); std::fclose(fb);
  return ok ? 0 : 1;
}
