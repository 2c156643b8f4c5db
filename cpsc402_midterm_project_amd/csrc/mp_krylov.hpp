// mp_krylov.hpp — device state of the FGMRES inner solver (basis V, Z, Hessenberg, rotations).
#pragma once
#include "mp_kernels.hpp"

namespace mpk {

struct Krylov {
  Workspace* W = nullptr;
  int64_t n = 0; int m = 0;
  DBuf<double> V, Z, w, sR, rcur, H, small, part;
  DBuf<double> part2;   // the second CGS pass's dot-product partials (written while part is still being read)
  DBuf<unsigned> gcount;  // workgroups of the fused second CGS pass that have finished (the last runs the Givens step)
  hipEvent_t ev[2] = {nullptr, nullptr};             // step-output events of the steps run ahead of the host
  // coherent pinned host slots [step][rec, hn, finite, tag] written by the step's Givens kernel, and the last tag
  // handed out (tags only grow, so a slot never shows a stale step as the awaited one)
  double* hstep = nullptr; unsigned long long tag = 0;
  Krylov() = default;
  Krylov(const Krylov&) = delete;
  Krylov& operator=(const Krylov&) = delete;
  ~Krylov();
  // small layout: cs[m+1] | sn[m+1] | g[m+1] | h[m+1] (projections) | y[m+1] | scalars...
  double* cs() { return small.p; }
  double* g() { return small.p + 2 * (m + 1); }
  double* h() { return small.p + 3 * (m + 1); }
  double* h2() { return small.p + 4 * (m + 1); }
  double* y() { return small.p + 5 * (m + 1); }
  double* scal() { return small.p + 6 * (m + 1); }   // 8 scalars: [0] hn, [1] norm tmp, [2..4] givens out
  // activity word of the Arnoldi steps (set at a cycle's start, cleared by the step that ends the cycle)
  int* active() { return (int*)(small.p + 6 * (m + 1) + 8); }
  void ensure(int64_t n_, int m_);
  // act: kernels return at once when *act == 0 (nullptr: always run)
  void dots(const double* Vb, int ncols, const double* x, double* out, bool accumulate, const int* act = nullptr);
  void update(const double* Vb, int ncols, const double* h, double* x, double sign, const int* act = nullptr);
  void norm(const double* x, double* out, const int* act = nullptr);
  // d = 0, rcur = r and out[0] = ||r|| (norm's sums), in two launches
  void start(const double* r, double* d, double* out);
  void scale_copy(const double* x, double* y, const double* s, bool inv, const int* act = nullptr);
  void givens(int j, const double* h, const double* hn, double beta0, double* out, double tol = -1.0, int* act = nullptr);
  void init_cycle(double beta, int* act = nullptr);
  // One Arnoldi step after w = A z_j: CGS2 against V[:, 0..j] (h, h2, w), hn = ||w|| in scal()[0] and the Givens
  // update with its outputs in scal()[2..4], in 3 launches; the same sums as dots / update / norm / givens.  j < 256.
  // hslot (coherent pinned host memory, 4 words): the step's outputs and then its tag are also written there.
  // prod_part: w is first formed from the product's column-chunk partials (prod_nch chunks, one FP64 column),
  // in the same launch as the first dot products.
  void arnoldi_tail(int j, double beta0, double tol, int* act, double* hslot = nullptr, unsigned long long tag = 0,
                    const double* prod_part = nullptr, int64_t prod_nch = 0);
  void lsq(int k, double* y);
};

void axpy(double* x, const double* d, int64_t n, double a, Workspace& W);
void copy_f64(const double* x, double* y, int64_t n, Workspace& W);
// y = x + a d with axpy's rounding (y may not alias x or d)
void axpy_to(double* y, const double* x, const double* d, int64_t n, double a, Workspace& W);
void fill_f64(double* x, int64_t n, double v, Workspace& W);

}  // namespace mpk
