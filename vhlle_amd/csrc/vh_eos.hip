// Tabulated EoS on the host: the TABLE branch of EoSs (src/eos.cpp:27-74).
// The reference reads columns e, p, T [, mu], clamps p < 0 to 0, and builds
// GSL natural cubic splines (gsl_interp_cspline) of p(e), T(e), mu(e).  The
// spline coefficients below follow the published natural-spline construction
// of that routine: c_i = y''(x_i)/2 solves, for i = 1 .. n-2,
//   h_i-1 c_i-1 + 2 (h_i-1 + h_i) c_i + h_i c_i+1
//     = 3 ((y_i+1 - y_i)/h_i - (y_i - y_i-1)/h_i-1),   c_0 = c_n-1 = 0,
// by the symmetric tridiagonal LDL^T recurrence, and on each interval
//   b_i = (y_i+1 - y_i)/h_i - h_i (c_i+1 + 2 c_i)/3,  d_i = (c_i+1 - c_i)/(3 h_i).
#include "vh_eos.h"

#include <cmath>
#include <cstring>

namespace vh {
namespace {

void natural_spline(const double *x, const double *y, int n, double *b, double *c, double *d) {
  const int m = n - 2;  // unknowns c_1 .. c_n-2
  std::vector<double> diag(m > 0 ? m : 1), off(m > 0 ? m : 1), g(m > 0 ? m : 1);
  for (int i = 0; i < m; i++) {
    const double h_i = x[i + 1] - x[i], h_ip1 = x[i + 2] - x[i + 1];
    const double ydiff_i = y[i + 1] - y[i], ydiff_ip1 = y[i + 2] - y[i + 1];
    const double g_i = h_i != 0. ? 1. / h_i : 0., g_ip1 = h_ip1 != 0. ? 1. / h_ip1 : 0.;
    off[i] = h_ip1;
    diag[i] = 2. * (h_ip1 + h_i);
    g[i] = 3. * (ydiff_ip1 * g_ip1 - ydiff_i * g_i);
  }
  c[0] = c[n - 1] = 0.;
  if (m == 1) {
    c[1] = g[0] / diag[0];
  } else if (m > 1) {
    // LDL^T of the symmetric tridiagonal matrix (diag, off)
    std::vector<double> gamma(m), alpha(m), z(m);
    alpha[0] = diag[0];
    gamma[0] = off[0] / alpha[0];
    for (int i = 1; i < m - 1; i++) {
      alpha[i] = diag[i] - off[i - 1] * gamma[i - 1];
      gamma[i] = off[i] / alpha[i];
    }
    alpha[m - 1] = diag[m - 1] - off[m - 2] * gamma[m - 2];
    z[0] = g[0];
    for (int i = 1; i < m; i++) z[i] = g[i] - gamma[i - 1] * z[i - 1];
    for (int i = 0; i < m; i++) z[i] = z[i] / alpha[i];
    c[m] = z[m - 1];
    for (int i = m - 2; i >= 0; i--) c[i + 1] = z[i] - gamma[i] * c[i + 2];
  }
  for (int i = 0; i < n - 1; i++) {
    const double dx = x[i + 1] - x[i], dy = y[i + 1] - y[i];
    b[i] = (dy / dx) - dx * (c[i + 1] + 2. * c[i]) / 3.;
    d[i] = (c[i + 1] - c[i]) / (3. * dx);
  }
  b[n - 1] = d[n - 1] = 0.;
}

}  // namespace

int eos_table_host(const vh_eos_t &eos, std::vector<double> &buf, int &n) {
  n = eos.n;
  if (n < 3 || !eos.e || !eos.p || !eos.T) return VH_ERR_ARG;
  for (int i = 1; i < n; i++)
    if (!(eos.e[i] > eos.e[i - 1])) return VH_ERR_ARG;  // GSL rejects such a table
  buf.assign((size_t)13 * n, 0.);
  double *x = buf.data();
  std::memcpy(x, eos.e, n * sizeof(double));
  for (int f = 0; f < 3; f++) {
    double *y = x + (size_t)(1 + 4 * f) * n;
    for (int i = 0; i < n; i++) {
      if (f == 0) y[i] = eos.p[i] < 0. ? 0. : eos.p[i];  // src/eos.cpp:49
      else if (f == 1) y[i] = eos.T[i];
      else y[i] = eos.mu ? eos.mu[i] : 0.;
    }
    natural_spline(x, y, n, y + n, y + 2 * n, y + 3 * n);
  }
  return VH_OK;
}

int make_dev_eos(const vh_eos_t *eos, int *st, DevEos &de, double **dbuf) {
  *dbuf = nullptr;
  if (!eos) return VH_ERR_ARG;
  de = DevEos();
  de.kind = eos->kind;
  const double pi_ = 3.14159265358979312;  // C_PI, src/inc.h
  de.cnst = (16 + 0.5 * 21.0 * 2.5) * (pi_ * pi_) / 30.0 /
            (0.197326968 * 0.197326968 * 0.197326968);
  de.tab.x = nullptr;
  de.tab.n = 0;
  de.tab.hint = nullptr;
  de.tab.K = 0;
  de.tab.l0 = de.tab.inv = 0.f;
  de.st = st;
  if (eos->kind == VH_EOS_IDEAL) return VH_OK;
  if (eos->kind != VH_EOS_TABLE) return VH_ERR_ARG;
  std::vector<double> buf;
  int n = 0;
  const int rc = eos_table_host(*eos, buf, n);
  if (rc != VH_OK) return rc;
  // search hints: K bins uniform in log2(e) from the first positive node
  const double *x = buf.data();
  const int K = 4096;
  const double xlo = x[0] > 0. ? x[0] : x[1];
  const double l0 = std::log2(xlo), l1 = std::log2(x[n - 1]);
  std::vector<int> hint(K);
  for (int k = 0; k < K; k++) {
    const double v = std::exp2(l0 + (l1 - l0) * k / K);
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
      const int mid = (hi + lo) / 2;
      if (x[mid] > v)
        hi = mid;
      else
        lo = mid;
    }
    hint[k] = lo;
  }
  const size_t nb = buf.size() * sizeof(double), bytes = nb + K * sizeof(int);
  if (hipMalloc(dbuf, bytes) != hipSuccess) return VH_ERR_HIP;
  if (hipMemcpy(*dbuf, buf.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(reinterpret_cast<char *>(*dbuf) + nb, hint.data(), K * sizeof(int),
                hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(*dbuf);
    *dbuf = nullptr;
    return VH_ERR_HIP;
  }
  de.tab.x = *dbuf;
  de.tab.n = n;
  de.tab.hint = reinterpret_cast<const int *>(reinterpret_cast<const char *>(*dbuf) + nb);
  de.tab.K = K;
  de.tab.l0 = (float)l0;
  de.tab.inv = (float)(K / (l1 - l0));
  return VH_OK;
}

}  // namespace vh

extern "C" int vh_eos_table(const vh_eos_t *eos, double *buf) {
  // the spline table of a VH_EOS_TABLE EoS on the host (no GPU involved)
  if (!eos || !buf || eos->kind != VH_EOS_TABLE) return VH_ERR_ARG;
  std::vector<double> t;
  int n = 0;
  const int rc = vh::eos_table_host(*eos, t, n);
  if (rc != VH_OK) return rc;
  std::memcpy(buf, t.data(), t.size() * sizeof(double));
  return VH_OK;
}
