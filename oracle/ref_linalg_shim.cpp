// ref_linalg_shim.cpp -- extern "C" entry points over the reference's own
// linalg.cpp (trunk/src/linalg.cpp:95-126 InvertMatrix, :249-307
// EigenRealGeneral, :1304-1315 CalcCijk), compiled together with that file by
// oracle/Makefile into oracle/_ref/.  Test infrastructure only: it lets the
// tests pin the oracle's eigen -> P(t) step against the reference's EISPACK path.
#include <cmath>
#include <vector>

#include "defs.h"
#include "linalg.h"

namespace {
struct Rows {
  std::vector<double> buf;
  std::vector<double *> ptr;
  Rows(int n, const double *src = nullptr) : buf(size_t(n) * n), ptr(n) {
    for (int i = 0; i < n; i++) ptr[i] = &buf[size_t(i) * n];
    if (src)
      for (size_t k = 0; k < buf.size(); k++) buf[k] = src[k];
  }
};
}  // namespace

// Q (n x n, row-major) -> eigenvalues (real parts), V, Vinv as the reference's
// CalcEigenStuff (model.cpp:827-856) computes them.  Returns the EISPACK rc
// (0 ok, RC_COMPLEX_EVAL for complex eigenvalues).
extern "C" int ref_eigen(int n, const double *Q, double *lam, double *V, double *Vinv) {
  Rows a(n, Q), u(n), tv(n), inv(n);
  std::vector<double> vi(n), work(n), col(n);
  std::vector<int> iwork(n), indx(n);
  int rc = EigenRealGeneral(n, a.ptr.data(), lam, vi.data(), u.ptr.data(), iwork.data(), work.data());
  if (rc != 0) return rc;
  for (size_t k = 0; k < u.buf.size(); k++) tv.buf[k] = u.buf[k];
  InvertMatrix(tv.ptr.data(), n, col.data(), indx.data(), inv.ptr.data());
  for (size_t k = 0; k < u.buf.size(); k++) {
    V[k] = u.buf[k];
    Vinv[k] = inv.buf[k];
  }
  return 0;
}

// P(t) from the reference's c_ijk path (CalcCijk + the AltCalcPmat sum,
// model.cpp:1469-1597): P_ij = max(0, sum_k c_ijk exp(lam_k * t)).
extern "C" void ref_pmat(int n, const double *V, const double *Vinv, const double *lam, double t, double *P) {
  Rows u(n, V), v(n, Vinv);
  std::vector<double> c(size_t(n) * n * n), ex(n);
  CalcCijk(c.data(), n, (const double **)u.ptr.data(), (const double **)v.ptr.data());
  for (int k = 0; k < n; k++) ex[k] = std::exp(lam[k] * t);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double s = 0.0;
      for (int k = 0; k < n; k++) s += c[(size_t(i) * n + j) * n + k] * ex[k];
      P[i * n + j] = s > 0.0 ? s : 0.0;
    }
}
