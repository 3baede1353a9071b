// Host side of the tabulated EoSs: validation, the natural cubic spline
// coefficients of EoSs (VH_EOS_TABLE), the device copies of the EoSChiral and
// EoSHadron tables (VH_EOS_CHIRAL, VH_EOS_HADRON) used by DevEos, and
// vh_eos_eval.
//
// The TABLE branch of EoSs (src/eos.cpp:27-74) reads columns e, p, T [, mu],
// clamps p < 0 to 0, and builds GSL natural cubic splines (gsl_interp_cspline)
// of p(e), T(e), mu(e).  The spline coefficients below follow the published
// natural-spline construction of that routine: c_i = y''(x_i)/2 solves, for
// i = 1 .. n-2,
//   h_i-1 c_i-1 + 2 (h_i-1 + h_i) c_i + h_i c_i+1
//     = 3 ((y_i+1 - y_i)/h_i - (y_i - y_i-1)/h_i-1),   c_0 = c_n-1 = 0,
// by the symmetric tridiagonal LDL^T recurrence, and on each interval
//   b_i = (y_i+1 - y_i)/h_i - h_i (c_i+1 + 2 c_i)/3,  d_i = (c_i+1 - c_i)/(3 h_i).
#include "vh_eos.h"

#include <cfloat>
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

// the interval SplineTab::interval returns for v (x_i <= v < x_i+1, the end
// intervals outside the table)
int interval_of(const double *x, int n, double v) {
  if (!(v >= x[1])) return 0;
  if (v >= x[n - 2]) return n - 2;
  int lo = 1, hi = n - 2;  // x_lo <= v < x_hi
  while (hi > lo + 1) {
    const int mid = (hi + lo) / 2;
    if (x[mid] > v)
      hi = mid;
    else
      lo = mid;
  }
  return lo;
}

// The cubic of interval i of function f re-centred on xf = (float)x_i: the
// Taylor coefficients of the same polynomial at xf, in double, then rounded.
// False if one is not a finite float.
bool recentred(const double *x, int n, int f, int i, float out[5]) {
  const double *y = x + (size_t)(1 + 4 * f) * n, *b = y + n, *c = y + 2 * n, *d = y + 3 * n;
  const float xf = (float)x[i];
  const double t = (double)xf - x[i];
  const double v[5] = {(double)xf, y[i] + t * (b[i] + t * (c[i] + t * d[i])),
                       b[i] + t * (2. * c[i] + 3. * d[i] * t), c[i] + 3. * d[i] * t, d[i]};
  for (int k = 0; k < 5; k++) {
    out[k] = (float)v[k];
    if (!std::isfinite(out[k])) return false;
  }
  return true;
}

// The direct lookup of SplineTab: the bins, tried with fK = 4096, 8192, ...
// 2^18; false if no bin count leaves every bin with at most one node.  A
// bin's entry serves every v the device may map to it: its range widened by a
// relative 1e-4 on both sides, far above the error of the FP32 bin index (the
// rounding of v to FP32 and of the log2).  c64: the interleaved FP64
// coefficients per interval; c32: the FP32 records per bin (empty when a
// coefficient does not fit a float).
bool direct_lookup(const double *x, int n, double xlo, SplineTab &t, std::vector<SplineBin> &bins,
                   std::vector<double> &c64, std::vector<float> &c32) {
  const float fxlo = (float)xlo;
  const double l1 = std::log2(x[n - 1]);
  const float fl0 = (float)std::log2((double)fxlo);
  int fK = 4096;
  for (; fK <= (1 << 18); fK *= 2) {
    const float finv = (float)(fK / (l1 - (double)fl0));
    if (!(finv > 0.f) || !std::isfinite(finv)) return false;
    bins.assign((size_t)fK + 1, SplineBin{0., 0, 0.f});
    bool ok = true;
    for (int k = -1; k < fK; k++) {
      double a, b;  // the v that map to entry k + 1
      if (k < 0) {
        a = -INFINITY;
        b = (double)fxlo * (1. + 1e-4);
      } else {
        a = std::exp2((double)fl0 + k / (double)finv) * (1. - 1e-4);
        b = k == fK - 1 ? INFINITY : std::exp2((double)fl0 + (k + 1) / (double)finv) * (1. + 1e-4);
        if (k == 0) a = std::fmin(a, (double)fxlo * (1. - 1e-4));
      }
      const int ia = interval_of(x, n, a), ib = interval_of(x, n, b);
      if (ib > ia + 1) {
        ok = false;
        break;
      }
      bins[k + 1] = SplineBin{ib > ia ? x[ib] : NAN, ia, ib > ia ? (float)x[ib] : NAN};
    }
    if (!ok) continue;
    t.finv = finv;
    break;
  }
  if (fK > (1 << 18)) return false;
  t.fK = fK;
  t.fl0 = fl0;
  t.fxlo = fxlo;
  c64.assign((size_t)3 * n * 6, 0.);
  for (int f = 0; f < 3; f++) {
    const double *y = x + (size_t)(1 + 4 * f) * n, *b = y + n, *c = y + 2 * n, *d = y + 3 * n;
    for (int i = 0; i < n - 1; i++) {
      double *r = c64.data() + ((size_t)f * n + i) * 6;
      r[0] = x[i];
      r[1] = y[i];
      r[2] = b[i];
      r[3] = c[i];
      r[4] = d[i];
    }
  }
  c32.assign((size_t)2 * (fK + 1) * 12, 0.f);
  for (int f = 0; f < 2; f++)
    for (int k = 0; k <= fK; k++) {
      const SplineBin &e = bins[k];
      const bool two = !std::isnan(e.xs);
      float A[5], B[5];
      if (!recentred(x, n, f, e.ia, A) || !recentred(x, n, f, e.ia + (two ? 1 : 0), B)) {
        c32.clear();
        return true;
      }
      float *r = c32.data() + ((size_t)f * (fK + 1) + k) * 12;
      r[0] = two ? e.xsf : FLT_MAX;  // v >= xs takes cubic B
      for (int j = 0; j < 5; j++) {
        r[1 + j] = A[j];
        r[6 + j] = B[j];
      }
    }
  return true;
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

namespace {

bool chiral_ok(const vh_eos_chiral_tab_t *t) {
  return t && t->ne >= 2 && t->nn >= 2 && t->emax > t->emin && t->nmax > t->nmin && t->p &&
         t->T && t->mub && t->mus;
}

// the scalars of one EoSaux table, its fields pointing at the descriptor's
void chiral_scalars(const vh_eos_chiral_tab_t &t, ChiralTab &c) {
  c.ne = t.ne;
  c.nn = t.nn;
  c.emin = t.emin;
  c.nmin = t.nmin;
  c.de = (t.emax - t.emin) / (t.ne - 1);  // src/eoChiral.cpp:104-105
  c.dn = (t.nmax - t.nmin) / (t.nn - 1);
  c.p = t.p;
  c.T = t.T;
  c.mub = t.mub;
  c.mus = t.mus;
}

void chiral_append(const vh_eos_chiral_tab_t &t, std::vector<double> &v) {
  const size_t M = (size_t)t.ne * t.nn;
  for (const double *f : {t.p, t.T, t.mub, t.mus}) v.insert(v.end(), f, f + M);
}

}  // namespace

// The scalars of a VH_EOS_CHIRAL or VH_EOS_HADRON EoS into de, whose table
// pointers are left at the descriptor's host arrays; with pack, also the
// concatenated field arrays for the device copy: chiral small then big (p, T,
// mu_b, mu_s each), hadron p, T, mu_b, mu_q, mu_s; stat: the hadron status
// flags (zeros when the descriptor has none).
int charged_tables_host(const vh_eos_t &eos, DevEos &de, std::vector<double> *v,
                        std::vector<int> *stat) {
  if (v) v->clear();
  if (stat) stat->clear();
  if (eos.kind == VH_EOS_CHIRAL) {
    if (!chiral_ok(eos.small) || !chiral_ok(eos.big)) return VH_ERR_ARG;
    chiral_scalars(*eos.small, de.small);
    chiral_scalars(*eos.big, de.big);
    if (v) {
      chiral_append(*eos.small, *v);
      chiral_append(*eos.big, *v);
    }
    return VH_OK;
  }
  if (eos.kind != VH_EOS_HADRON) return VH_ERR_ARG;
  const vh_eos_hadron_tab_t *t = eos.hadron;
  if (!t || t->ne < 2 || t->nnb < 3 || t->nnq < 3 || !(t->logemax > t->logemin) ||
      !(t->lognmax > 0.) || !(t->e0 > 0.) || !(t->n0 > 0.) || !t->p || !t->T || !t->mub ||
      !t->muq || !t->mus)
    return VH_ERR_ARG;
  HadronTab &h = de.had;
  h.ne = t->ne;
  h.nnb = t->nnb;
  h.nnq = t->nnq;
  h.e0 = t->e0;
  h.n0 = t->n0;
  h.logemin = t->logemin;
  h.lognmax = t->lognmax;
  // the constructor's derived constants (src/eoHadron.cpp:50-54) and the
  // per-call steps (src/eoHadron.cpp:93-96)
  h.dxe = (t->logemax - t->logemin) / (t->ne - 1);
  h.dxnb = (2.0 * t->lognmax) / (t->nnb - 1);
  h.dxnq = (2.0 * t->lognmax) / (t->nnq - 1);
  h.nb_abs_min = t->n0 * std::exp(2.0 * h.dxnb - t->lognmax);
  h.nq_abs_min = t->n0 * std::exp(2.0 * h.dxnq - t->lognmax);
  h.e_min = t->e0 * std::exp(t->logemin);
  h.p = t->p;
  h.T = t->T;
  h.mub = t->mub;
  h.muq = t->muq;
  h.mus = t->mus;
  h.status = t->status;
  const size_t M = (size_t)t->ne * t->nnb * t->nnq;
  if (v)
    for (const double *f : {t->p, t->T, t->mub, t->muq, t->mus}) v->insert(v->end(), f, f + M);
  if (stat) {
    if (t->status)
      stat->assign(t->status, t->status + M);
    else
      stat->assign(M, 0);
  }
  return VH_OK;
}

// points the tables of de into the concatenated arrays at v (and st)
void point_tables(DevEos &de, const double *v, const int *st) {
  if (de.kind == VH_EOS_CHIRAL) {
    for (ChiralTab *t : {&de.small, &de.big}) {
      const size_t M = (size_t)t->ne * t->nn;
      t->p = v;
      t->T = v + M;
      t->mub = v + 2 * M;
      t->mus = v + 3 * M;
      v += 4 * M;
    }
  } else if (de.kind == VH_EOS_HADRON) {
    const size_t M = (size_t)de.had.ne * de.had.nnb * de.had.nnq;
    de.had.p = v;
    de.had.T = v + M;
    de.had.mub = v + 2 * M;
    de.had.muq = v + 3 * M;
    de.had.mus = v + 4 * M;
    de.had.status = st;
  }
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
  de.tab.bins = nullptr;
  de.tab.c64 = nullptr;
  de.tab.c32 = nullptr;
  de.tab.fK = 0;
  de.st = st;
  if (eos->kind == VH_EOS_IDEAL) return VH_OK;
  if (eos->kind == VH_EOS_CHIRAL || eos->kind == VH_EOS_HADRON) {
    std::vector<double> v;
    std::vector<int> stat;
    const int rc = charged_tables_host(*eos, de, &v, &stat);
    if (rc != VH_OK) return rc;
    const size_t nb = v.size() * sizeof(double), bytes = nb + stat.size() * sizeof(int);
    if (hipMalloc(dbuf, bytes) != hipSuccess) return VH_ERR_HIP;
    if (hipMemcpy(*dbuf, v.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
        (!stat.empty() && hipMemcpy(reinterpret_cast<char *>(*dbuf) + nb, stat.data(),
                                    stat.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)) {
      (void)hipFree(*dbuf);
      *dbuf = nullptr;
      return VH_ERR_HIP;
    }
    point_tables(de, *dbuf, reinterpret_cast<const int *>(reinterpret_cast<const char *>(*dbuf) + nb));
    return VH_OK;
  }
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
  // v >= fxmin and v <= fxmax hold for a float v exactly when v >= x_0 and
  // v <= x_n-1 hold in double (the range test of SplineTab::eval)
  float lo = (float)x[0], hi = (float)x[n - 1];
  if ((double)lo < x[0]) lo = std::nextafter(lo, INFINITY);
  if ((double)hi > x[n - 1]) hi = std::nextafter(hi, -INFINITY);
  de.tab.fxmin = lo;
  de.tab.fxmax = hi;
  std::vector<SplineBin> bins;
  std::vector<double> c64;
  std::vector<float> c32;
  if (!direct_lookup(x, n, xlo, de.tab, bins, c64, c32)) {
    bins.clear();
    c64.clear();
    c32.clear();
  }
  const size_t nb = buf.size() * sizeof(double), nh = (nb + K * sizeof(int) + 255) / 256 * 256;
  const size_t nbin = bins.size() * sizeof(SplineBin), n64 = c64.size() * sizeof(double);
  const size_t bytes = nh + nbin + n64 + c32.size() * sizeof(float);
  if (hipMalloc(dbuf, bytes) != hipSuccess) return VH_ERR_HIP;
  char *base = reinterpret_cast<char *>(*dbuf);
  if (hipMemcpy(*dbuf, buf.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(base + nb, hint.data(), K * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      (!bins.empty() &&
       (hipMemcpy(base + nh, bins.data(), nbin, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(base + nh + nbin, c64.data(), n64, hipMemcpyHostToDevice) != hipSuccess)) ||
      (!c32.empty() && hipMemcpy(base + nh + nbin + n64, c32.data(), c32.size() * sizeof(float),
                                 hipMemcpyHostToDevice) != hipSuccess)) {
    (void)hipFree(*dbuf);
    *dbuf = nullptr;
    return VH_ERR_HIP;
  }
  de.tab.bins = bins.empty() ? nullptr : reinterpret_cast<const SplineBin *>(base + nh);
  de.tab.c64 = bins.empty() ? nullptr : reinterpret_cast<const double2 *>(base + nh + nbin);
  de.tab.c32 = c32.empty() ? nullptr : reinterpret_cast<const float4 *>(base + nh + nbin + n64);
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

namespace {

// EoS::eos and EoS::s (src/eos.cpp:18-25) at n points on the GPU
template <bool CH>
__global__ void __launch_bounds__(256)
k_eos_eval(vh::DevEos eos, const double *__restrict__ in, long n, double *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double e = in[4 * i], nb = in[4 * i + 1], nq = in[4 * i + 2], ns = in[4 * i + 3];
  double T, mub, muq, mus, p;
  eos.eos<CH>(e, nb, nq, ns, T, mub, muq, mus, p);
  double *o = out + 6 * i;
  o[0] = p;
  o[1] = T;
  o[2] = mub;
  o[3] = muq;
  o[4] = mus;
  o[5] = T > 0.0 ? (e + p - mub * nb - muq * nq - mus * ns) / T : 0.;
}

// the same on the host: the lookups of kinds 2 and 3 are the device code
// compiled for the CPU, reading the descriptor's arrays in place (no copy, so
// a call per point is cheap); kinds 0 and 1 evaluate p = e/3 and the splines
int eos_eval_host(const vh_eos_t &d, const double *in, long n, double *out) {
  vh::DevEos de = vh::DevEos();
  de.kind = d.kind;
  std::vector<double> tab;
  int nt = 0;
  if (d.kind == VH_EOS_CHIRAL || d.kind == VH_EOS_HADRON) {
    // the lookups read the descriptor's arrays in place
    const int rc = vh::charged_tables_host(d, de, nullptr, nullptr);
    if (rc != VH_OK) return rc;
  } else if (d.kind == VH_EOS_TABLE) {
    const int rc = vh::eos_table_host(d, tab, nt);
    if (rc != VH_OK) return rc;
  } else if (d.kind != VH_EOS_IDEAL) {
    return VH_ERR_ARG;
  }
  const double pi_ = 3.14159265358979312;  // C_PI, src/inc.h
  const double cnst = (16 + 0.5 * 21.0 * 2.5) * (pi_ * pi_) / 30.0 /
                      (0.197326968 * 0.197326968 * 0.197326968);
  auto spline = [&](int f, double x) {  // gsl_interp_bsearch + the cubic
    const double *xs = tab.data();
    int lo = 0, hi = nt - 1;
    while (hi > lo + 1) {
      const int mid = (hi + lo) / 2;
      if (xs[mid] > x)
        hi = mid;
      else
        lo = mid;
    }
    const double *y = xs + (size_t)(1 + 4 * f) * nt, *b = y + nt, *c = y + 2 * nt, *dd = y + 3 * nt;
    const double delx = x - xs[lo];
    return y[lo] + delx * (b[lo] + delx * (c[lo] + delx * dd[lo]));
  };
  for (long i = 0; i < n; i++) {
    const double e = in[4 * i], nb = in[4 * i + 1], nq = in[4 * i + 2], ns = in[4 * i + 3];
    double T = 0., mub = 0., muq = 0., mus = 0., p = 0.;
    if (d.kind == VH_EOS_CHIRAL) {
      vh::chiral_eos<true>(de.small, de.big, e, nb, p, T, mub, muq, mus);
    } else if (d.kind == VH_EOS_HADRON) {
      de.had.get<true>(e, nb, nq, p, T, mub, muq, mus);
    } else if (d.kind == VH_EOS_TABLE) {
      p = spline(0, e);
      T = spline(1, e);
    } else {
      p = e / 3.;
      T = e > 0. ? 1.0 * std::pow(e / cnst, 0.25) : 0.;
    }
    double *o = out + 6 * i;
    o[0] = p;
    o[1] = T;
    o[2] = mub;
    o[3] = muq;
    o[4] = mus;
    o[5] = T > 0.0 ? (e + p - mub * nb - muq * nq - mus * ns) / T : 0.;
  }
  return VH_OK;
}

}  // namespace

extern "C" int vh_eos_eval(const vh_eos_t *eos, const double *in, long n, double *out,
                           int where) {
  if (!eos || n < 0 || (n > 0 && (!in || !out)) || (where != 0 && where != 1)) return VH_ERR_ARG;
  if (where == 0) return eos_eval_host(*eos, in, n, out);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return VH_ERR_NODEVICE;
  vh::DevEos de;
  double *tab = nullptr, *buf = nullptr;
  int rc = vh::make_dev_eos(eos, nullptr, de, &tab);
  if (rc != VH_OK) return rc;
  if (n > 0) {
    const size_t bi = (size_t)4 * n * sizeof(double), bo = (size_t)6 * n * sizeof(double);
    if (hipMalloc(&buf, bi + bo) != hipSuccess) {
      rc = VH_ERR_HIP;
    } else if (hipMemcpy(buf, in, bi, hipMemcpyHostToDevice) != hipSuccess) {
      rc = VH_ERR_HIP;
    } else {
      if (de.charged())
        hipLaunchKernelGGL(k_eos_eval<true>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, de,
                           buf, n, buf + 4 * n);
      else
        hipLaunchKernelGGL(k_eos_eval<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                           de, buf, n, buf + 4 * n);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpy(out, buf + 4 * n, bo, hipMemcpyDeviceToHost) != hipSuccess)
        rc = VH_ERR_HIP;
    }
  }
  if (buf) (void)hipFree(buf);
  if (tab) (void)hipFree(tab);
  return rc;
}
