// mp_krylov.hpp — device state of the FGMRES inner solver (basis V, Z, Hessenberg, rotations).
#pragma once
#include "mp_kernels.hpp"

namespace mpk {

struct Krylov {
  Workspace* W = nullptr;
  int64_t n = 0; int m = 0;
  DBuf<double> V, Z, w, sR, rcur, H, small, part;
  // small layout: cs[m+1] | sn[m+1] | g[m+1] | h[m+1] (projections) | y[m+1] | scalars...
  double* cs() { return small.p; }
  double* g() { return small.p + 2 * (m + 1); }
  double* h() { return small.p + 3 * (m + 1); }
  double* h2() { return small.p + 4 * (m + 1); }
  double* y() { return small.p + 5 * (m + 1); }
  double* scal() { return small.p + 6 * (m + 1); }   // 8 scalars: [0] hn, [1] norm tmp, [2..4] givens out
  void ensure(int64_t n_, int m_);
  void dots(const double* Vb, int ncols, const double* x, double* out, bool accumulate);
  void update(const double* Vb, int ncols, const double* h, double* x, double sign);
  void norm(const double* x, double* out);
  void scale_copy(const double* x, double* y, const double* s, bool inv);
  void givens(int j, const double* h, const double* hn, double beta0, double* out);
  void init_cycle(double beta);
  void lsq(int k, double* y);
};

void axpy(double* x, const double* d, int64_t n, double a, Workspace& W);
void copy_f64(const double* x, double* y, int64_t n, Workspace& W);
void fill_f64(double* x, int64_t n, double v, Workspace& W);

}  // namespace mpk
