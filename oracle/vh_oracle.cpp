// CPU oracle for the vHLLE hydro step -- TEST INFRASTRUCTURE ONLY (see
// vh_oracle.h).  Serial restatement of yukarpenko/vhlle's Hydro::performStep
// in the reference's loop order and floating-point operation order.  Built
// with -O3 -ffp-contract=off and no -march flags (x86-64 SSE2, like the
// reference Makefile:3), so that each expression rounds exactly as the
// reference's does.  Every function cites the reference lines it follows.
#include "vh_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// ----------------------------------------------------------------- constants
// Q component slots (src/inc.h:4-10)
enum { kT = 0, kX = 1, kY = 2, kZ = 3, kNB = 4, kNQ = 5, kNS = 6 };
enum { kPredict = 0, kCorrect = 1 };  // src/inc.h:12-13

// packed symmetric 4x4 index (src/cll.cpp:22-31), without the bounds exit
inline int sym(int i, int j) {
  return j < i ? (i * (i + 1)) / 2 + j : (j * (j + 1)) / 2 + i;
}

// slope limiter (src/cll.cpp:12-19)
inline double minmod(double a, double b) {
  if (a * b <= 0.) return 0.;
  return std::fabs(a) > std::fabs(b) ? b : a;
}

// sign() of src/hdo.cpp:34-41
inline double sgn(double x) { return x > 0 ? 1. : (x < 0. ? -1. : 0.); }

// ------------------------------------------------------------------- EoS
// EoSs (src/eos.h:44-71, src/eos.cpp:27-120).  kind 0: the SIMPLE branch,
// ideal massless gas p = e/3, T = (e/cnst)^(1/4).  kind 1: the TABLE branch,
// GSL natural cubic splines (gsl_interp_cspline; GSL is not part of the
// reference tree and its version is unpinned) of p(e), T(e), mu(e), restated
// from the published natural-spline construction: c_i = y''(x_i)/2 from
//   h_i-1 c_i-1 + 2 (h_i-1 + h_i) c_i + h_i c_i+1
//     = 3 ((y_i+1 - y_i)/h_i - (y_i - y_i-1)/h_i-1),   c_0 = c_n-1 = 0
// (symmetric tridiagonal LDL^T), b_i = dy/h - h (c_i+1 + 2 c_i)/3,
// d_i = (c_i+1 - c_i)/(3 h), y = y_i + dx (b_i + dx (c_i + dx d_i)); interval
// search as gsl_interp_bsearch.  cs2(e) = dpe(e) = (p(1.1 e) - p(e))/(0.1 e).
// Outside the table GSL's domain error aborts the reference; here the end
// interval is used and out_of_range counts it.
struct Spline {
  std::vector<double> x, y, b, c, d;
  void init(const double *xs, const double *ys, int n) {
    x.assign(xs, xs + n);
    y.assign(ys, ys + n);
    b.assign(n, 0.);
    c.assign(n, 0.);
    d.assign(n, 0.);
    const int m = n - 2;
    std::vector<double> diag(m > 0 ? m : 1), off(m > 0 ? m : 1), g(m > 0 ? m : 1);
    for (int i = 0; i < m; i++) {
      const double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
      const double y0 = y[i + 1] - y[i], y1 = y[i + 2] - y[i + 1];
      const double g0 = h0 != 0. ? 1. / h0 : 0., g1 = h1 != 0. ? 1. / h1 : 0.;
      off[i] = h1;
      diag[i] = 2. * (h1 + h0);
      g[i] = 3. * (y1 * g1 - y0 * g0);
    }
    if (m == 1) {
      c[1] = g[0] / diag[0];
    } else if (m > 1) {
      std::vector<double> gam(m), alp(m), z(m);
      alp[0] = diag[0];
      gam[0] = off[0] / alp[0];
      for (int i = 1; i < m - 1; i++) {
        alp[i] = diag[i] - off[i - 1] * gam[i - 1];
        gam[i] = off[i] / alp[i];
      }
      alp[m - 1] = diag[m - 1] - off[m - 2] * gam[m - 2];
      z[0] = g[0];
      for (int i = 1; i < m; i++) z[i] = g[i] - gam[i - 1] * z[i - 1];
      for (int i = 0; i < m; i++) z[i] = z[i] / alp[i];
      c[m] = z[m - 1];
      for (int i = m - 2; i >= 0; i--) c[i + 1] = z[i] - gam[i] * c[i + 2];
    }
    for (int i = 0; i < n - 1; i++) {
      const double dx = x[i + 1] - x[i], dy = y[i + 1] - y[i];
      b[i] = (dy / dx) - dx * (c[i + 1] + 2. * c[i]) / 3.;
      d[i] = (c[i + 1] - c[i]) / (3. * dx);
    }
  }
  double eval(double v, long &out_of_range) const {
    const int n = (int)x.size();
    if (!(v >= x[0] && v <= x[n - 1])) out_of_range++;
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
      const int mid = (hi + lo) / 2;
      if (x[mid] > v)
        hi = mid;
      else
        lo = mid;
    }
    const double delx = v - x[lo];
    return y[lo] + delx * (b[lo] + delx * (c[lo] + delx * d[lo]));
  }
};

// EoSaux (src/eoChiral.cpp:38-154): one uniform (e, nb) table of p, T, mu_b,
// mu_s with bilinear interpolation; indices clamped to the table (so outside
// it the end cell's bilinear form extrapolates); p < 0 clamped to 0; e < 0
// gives zeros.
struct ChiralAux {
  int ne = 0, nn = 0;
  double emin = 0., emax = 0., nmin = 0., nmax = 0.;
  std::vector<double> ptab, Ttab, mubtab, mustab;  // [ie * nn + in]
  void init(const or_eos_chiral_tab_t &t) {
    ne = t.ne;
    nn = t.nn;
    emin = t.emin;
    emax = t.emax;
    nmin = t.nmin;
    nmax = t.nmax;
    const size_t M = (size_t)ne * nn;
    ptab.assign(t.p, t.p + M);
    Ttab.assign(t.T, t.T + M);
    mubtab.assign(t.mub, t.mub + M);
    mustab.assign(t.mus, t.mus + M);
  }
  // EoSaux::get (src/eoChiral.cpp:99-129); with only_p, EoSaux::p (:131-154)
  void get(double e, double nb, double &p, double &T, double &mub, double &mus,
           bool only_p) const {
    if (e < 0.) {
      T = mub = mus = p = 0.;
      return;
    }
    const double de = (emax - emin) / (ne - 1);
    const double dn = (nmax - nmin) / (nn - 1);
    int ie = (int)((e - emin) / de);
    int in = (int)((nb - nmin) / dn);
    if (ie < 0) ie = 0;
    if (in < 0) in = 0;
    if (ie > ne - 2) ie = ne - 2;
    if (in > nn - 2) in = nn - 2;
    const double em = e - emin - ie * de;
    const double nm = nb - nmin - in * dn;
    const double we[2] = {1. - em / de, em / de};
    const double wn[2] = {1. - nm / dn, nm / dn};
    T = mub = mus = p = 0.0;
    for (int je = 0; je < 2; je++)
      for (int jn = 0; jn < 2; jn++) {
        const size_t k = (size_t)(ie + je) * nn + (in + jn);
        p += we[je] * wn[jn] * ptab[k];
        if (only_p) continue;
        T += we[je] * wn[jn] * Ttab[k];
        mub += we[je] * wn[jn] * mubtab[k];
        mus += we[je] * wn[jn] * mustab[k];
      }
    if (p < 0.0) p = 0.0;
  }
};

// EoSHadron (src/eoHadron.cpp:13-208): a table of p, T, mu_b, mu_q, mu_s on
// nodes uniform in log(e/e0) and in log-mapped nb, nq (linear near n = 0).
struct HadronTable {
  double e0 = 1., n0 = 1., logemin = 0., logemax = 0., lognmax = 0.;
  int ne = 0, nnb = 0, nnq = 0;
  double nb_abs_min = 0., nq_abs_min = 0., e_min = 0.;
  std::vector<double> ptab, Ttab, mubtab, muqtab, mustab;
  std::vector<int> statustab;
  int index3(int ie, int inb, int inq) const { return inq + nnq * inb + nnq * nnb * ie; }
  void init(const or_eos_hadron_tab_t &t) {
    ne = t.ne;
    nnb = t.nnb;
    nnq = t.nnq;
    e0 = t.e0;
    n0 = t.n0;
    logemin = t.logemin;
    logemax = t.logemax;
    lognmax = t.lognmax;
    const size_t M = (size_t)ne * nnb * nnq;
    ptab.assign(t.p, t.p + M);
    Ttab.assign(t.T, t.T + M);
    mubtab.assign(t.mub, t.mub + M);
    muqtab.assign(t.muq, t.muq + M);
    mustab.assign(t.mus, t.mus + M);
    statustab.assign(M, 0);
    if (t.status) statustab.assign(t.status, t.status + M);
    // the constructor's derived constants (src/eoHadron.cpp:50-54)
    const double dxnb = (2.0 * lognmax) / (nnb - 1);
    const double dxnq = (2.0 * lognmax) / (nnq - 1);
    nb_abs_min = n0 * std::exp(2.0 * dxnb - lognmax);
    nq_abs_min = n0 * std::exp(2.0 * dxnq - lognmax);
    e_min = e0 * std::exp(logemin);
  }
  // position of a charge density on its mapped axis (src/eoHadron.cpp:98-111)
  static double xpos(double n, double n0, double lognmax, double dxn, double n_abs_min) {
    if (n == 0.0) return 0.;
    const double zn = std::log(std::fabs(n) / n0);
    if (zn >= -lognmax + 2.0 * dxn) return n / std::fabs(n) * 0.5 * (zn + lognmax);
    return n / n_abs_min * dxn;
  }
  // EoSHadron::eos (src/eoHadron.cpp:78-152); with only_p, EoSHadron::p (:154-208)
  void get(double e, double nb, double nq, double &T, double &mub, double &muq, double &mus,
           double &p, bool only_p) const {
    if (e <= 0.) {
      T = mub = muq = mus = p = 0.;
      return;
    }
    if (e < e_min) {
      p = e / e_min * ptab[index3(0, nnb / 2, nnq / 2)];
      T = e / e_min * Ttab[index3(0, nnb / 2, nnq / 2)];
      mub = muq = mus = 0.;
      return;
    }
    const double xe = std::log(e / e0);
    const double dxe = (logemax - logemin) / (ne - 1);
    const double dxnb = (2.0 * lognmax) / (nnb - 1);
    const double dxnq = (2.0 * lognmax) / (nnq - 1);
    const double xnb = xpos(nb, n0, lognmax, dxnb, nb_abs_min);
    const double xnq = xpos(nq, n0, lognmax, dxnq, nq_abs_min);
    int ie = (int)((xe - logemin) / dxe);
    int inb = (int)((xnb + lognmax) / dxnb);
    int inq = (int)((xnq + lognmax) / dxnq);
    if (ie < 0) ie = 0;
    if (inb < 0) inb = 0;
    if (inq < 0) inq = 0;
    if (ie > ne - 2) ie = ne - 2;
    if (inb > nnb - 2) inb = nnb - 2;
    if (inq > nnq - 2) inq = nnq - 2;
    const double em = xe - logemin - ie * dxe;
    const double nbm = xnb + lognmax - inb * dxnb;
    const double nqm = xnq + lognmax - inq * dxnq;
    if (statustab[index3(ie, inb, inq)] == 1) {
      T = mub = muq = mus = p = 0.;
      return;
    }
    const double we[2] = {1. - em / dxe, em / dxe};
    const double wnb[2] = {1. - nbm / dxnb, nbm / dxnb};
    const double wnq[2] = {1. - nqm / dxnq, nqm / dxnq};
    T = mub = muq = mus = p = 0.0;
    for (int je = 0; je < 2; je++)
      for (int jnb = 0; jnb < 2; jnb++)
        for (int jnq = 0; jnq < 2; jnq++) {
          const int k = index3(ie + je, inb + jnb, inq + jnq);
          const double w = we[je] * wnb[jnb] * wnq[jnq];
          p += w * ptab[k];
          if (only_p) continue;
          T += w * Ttab[k];
          mub += w * mubtab[k];
          muq += w * muqtab[k];
          mus += w * mustab[k];
        }
    if (p < 0.0) p = 0.0;
  }
};

struct Eos {
  int kind = 0;
  // cnst of EoSs::t (src/eos.cpp:98-99); GCC folds pow(C_PI,2) and
  // pow(0.197326968,3) at compile time with correct rounding.
  double cnst;
  Spline sp, sT, smu;
  ChiralAux small, big;
  HadronTable had;
  mutable long out_of_range = 0;
  explicit Eos(const or_eos_t *d = nullptr) {
    const double pi_ = 3.14159265358979312;  // C_PI, src/inc.h:18
    cnst = (16 + 0.5 * 21.0 * 2.5) * (pi_ * pi_) / 30.0 /
           (0.197326968 * 0.197326968 * 0.197326968);
    if (d && d->kind == 1) {
      kind = 1;
      const int n = d->n;
      std::vector<double> p(n), mu(n, 0.);
      for (int i = 0; i < n; i++) p[i] = d->p[i] < 0. ? 0. : d->p[i];  // src/eos.cpp:49
      if (d->mu) mu.assign(d->mu, d->mu + n);
      sp.init(d->e, p.data(), n);
      sT.init(d->e, d->T, n);
      smu.init(d->e, mu.data(), n);
    } else if (d && d->kind == 2) {
      kind = 2;
      small.init(*d->small);
      big.init(*d->big);
    } else if (d && d->kind == 3) {
      kind = 3;
      had.init(*d->hadron);
    }
  }
  // EoSs::p(e) / EoSs::t(e): the kinds whose p does not depend on the charges
  double p(double e) const { return kind == 1 ? sp.eval(e, out_of_range) : e / 3.; }
  double T(double e) const {
    if (kind == 1) return sT.eval(e, out_of_range);
    return e > 0. ? 1.0 * std::pow(e / cnst, 0.25) : 0.;
  }
  // EoSChiral::eos (src/eoChiral.cpp:131-145) and ::p (:147-154)
  void chiral(double e, double nb, double &T_, double &mub, double &muq, double &mus,
              double &p_, bool only_p) const {
    if (e < 1.46 && nb < 0.3)
      small.get(e, nb, p_, T_, mub, mus, only_p);
    else if (e < 146. && nb < 6.)
      big.get(e, nb, p_, T_, mub, mus, only_p);
    else {
      p_ = 0.2964 * e;
      T_ = only_p ? 0. : 0.15120476935 * std::pow(e, 0.25);
      mub = mus = 0.0;
    }
    muq = 0.0;
  }
  // EoS::p(e, nb, nq, ns)
  double p(double e, double nb, double nq, double ns) const {
    double T_, mub, muq, mus, p_;
    if (kind == 2) {
      chiral(e, nb, T_, mub, muq, mus, p_, true);
      return p_;
    }
    if (kind == 3) {
      had.get(e, nb, nq, T_, mub, muq, mus, p_, true);
      return p_;
    }
    return p(e);
  }
  // EoS::eos(e, nb, nq, ns, T, mub, muq, mus, p)
  void eos(double e, double nb, double nq, double ns, double &T_, double &mub, double &muq,
           double &mus, double &p_) const {
    if (kind == 2) {
      chiral(e, nb, T_, mub, muq, mus, p_, false);
    } else if (kind == 3) {
      had.get(e, nb, nq, T_, mub, muq, mus, p_, false);
    } else {
      p_ = p(e);
      T_ = T(e);
      mub = muq = mus = 0.;
    }
  }
  // EoS::s (src/eos.cpp:18-25)
  double s(double e, double nb, double nq, double ns) const {
    double T_, mub, muq, mus, p_;
    eos(e, nb, nq, ns, T_, mub, muq, mus, p_);
    if (T_ > 0.0) return (e + p_ - mub * nb - muq * nq - mus * ns) / T_;
    return 0.;
  }
  // EoSs::cs2(e) = dpe(e); the chiral and hadron EoSs keep EoS::cs2(e) = 1/3
  // (src/eos.h:37)
  double cs2e(double e) const {
    if (kind == 1) return (p(e * 1.1) - p(e)) / (0.1 * e);
    return 1. / 3.;
  }
};

// (C2P_COUNT) iteration counter hook for or_c2p
thread_local int g_c2p_iters = 0;

// ---------------------------------------------------------------- con2prim
// transformPV (src/rmn.cpp:28-139).  Q is modified in place when |Q_vec| > Q_t
// (src/rmn.cpp:65-70), like the reference.
void con2prim(const Eos &eos, double Q[7], double &e, double &p,
              double &nb, double &nq, double &ns, double &vx, double &vy,
              double &vz) {
  const int maxit = 100;
  const double dpe = 1. / 3.;
  const double corrf = 0.9999;
  double vl = 0., vh = 1.;
  double M = std::sqrt(Q[kX] * Q[kX] + Q[kY] * Q[kY] + Q[kZ] * Q[kZ]);
  g_c2p_iters = 0;
  if (Q[kT] <= 0.) {
    e = p = 0.;
    vx = vy = vz = 0.;
    nb = nq = ns = 0.;
    return;
  }
  if (M == 0.) {
    e = Q[kT];
    vx = vy = vz = 0.;
    nb = Q[kNB];
    nq = Q[kNQ];
    ns = Q[kNS];
    p = eos.p(e, nb, nq, ns);
    return;
  }
  if (M > Q[kT]) {
    Q[kX] *= corrf * Q[kT] / M;
    Q[kY] *= corrf * Q[kT] / M;
    Q[kZ] *= corrf * Q[kT] / M;
    M = Q[kT] * corrf;
  }
  double v = 0.5 * (vl + vh);
  e = Q[kT] - M * v;
  if (e < 0.) e = 0.;
  nb = Q[kNB] * std::sqrt(1 - v * v);
  nq = Q[kNQ] * std::sqrt(1 - v * v);
  ns = Q[kNS] * std::sqrt(1 - v * v);
  p = eos.p(e, nb, nq, ns);
  double f = (Q[kT] + p) * v - M;
  double df = (Q[kT] + p) - M * v * dpe;
  double dvold = vh - vl;
  double dv = dvold;
  for (int it = 0; it < maxit; it++) {
    g_c2p_iters = it + 1;
    const bool bisect = (f + df * (vh - v)) * (f + df * (vl - v)) >= 0. ||
                        std::fabs(2. * f) > std::fabs(dvold * df);
    dvold = dv;
    if (bisect) {
      dv = 0.5 * (vh - vl);
      v = vl + dv;
    } else {
      dv = f / df;
      v -= dv;
    }
    if (std::fabs(dv) < 0.00001) break;
    e = Q[kT] - M * v;
    if (e < 0.) e = 0.;
    nb = Q[kNB] * std::sqrt(1 - v * v);
    nq = Q[kNQ] * std::sqrt(1 - v * v);
    ns = Q[kNS] * std::sqrt(1 - v * v);
    p = eos.p(e, nb, nq, ns);
    f = (Q[kT] + p) * v - M;
    df = (Q[kT] + p) - M * v * dpe;
    if (f > 0.)
      vh = v;
    else
      vl = v;
  }
  vx = v * Q[kX] / M;
  vy = v * Q[kY] / M;
  vz = v * Q[kZ] / M;
  e = Q[kT] - M * v;
  p = eos.p(e, nb, nq, ns);  // the previous iterate's nb (src/rmn.cpp:125-126)
  nb = Q[kNB] * std::sqrt(1 - vx * vx - vy * vy - vz * vz);
  nq = Q[kNQ] * std::sqrt(1 - vx * vx - vy * vy - vz * vz);
  ns = Q[kNS] * std::sqrt(1 - vx * vx - vy * vy - vz * vz);
}

// transformCV (src/rmn.cpp:253-263)
void prim2con(double e, double p, double nb, double nq, double ns, double vx,
              double vy, double vz, double Q[7]) {
  const double g2 = 1. / (1 - vx * vx - vy * vy - vz * vz);
  Q[kT] = (e + p * (vx * vx + vy * vy + vz * vz)) * g2;
  Q[kX] = (e + p) * vx * g2;
  Q[kY] = (e + p) * vy * g2;
  Q[kZ] = (e + p) * vz * g2;
  Q[kNB] = nb * std::sqrt(g2);
  Q[kNQ] = nq * std::sqrt(g2);
  Q[kNS] = ns * std::sqrt(g2);
}

// ------------------------------------------------------ transport coeffs
// TransportCoeff (src/trancoeff.h:6-45, src/trancoeff.cpp:19-100)
struct Trc {
  or_trc_t c;
  const Eos *eos;
  double zetaS(double e, double T, double s, double P) const {
    if (c.zetaSparam == 0)
      return c.zetaS0 * (1. / 3. - eos->cs2e(e)) / (std::exp((0.16 - T) / 0.001) + 1.);
    if (c.zetaSparam == 1) {
      const double Tp = 0.180;
      if (T < 0.180)
        return 0.03 + (0.08 * std::exp(((T / Tp) - 1.) / (0.0025))) +
               (0.22 * std::exp(((T / Tp) - 1) / (0.022)));
      else if (T < 0.200)
        return 27.55 * (T / Tp) - 13.45 - (13.77 * (T / Tp) * (T / Tp));
      return 0.001 + (0.9 * std::exp(-((T / Tp) - 1.) / (0.0025))) +
             (0.25 * std::exp(-((T / Tp) - 1.) / (0.13)));
    }
    if (c.zetaSparam == 2) {
      const double Tpk = 0.165, Tw = 0.010, Bn = 0.24, Bw = 1.5;
      if (T > Tpk)
        return Bn * ((Bw * Bw) / ((((T / Tpk) - 1.) * ((T / Tpk) - 1.)) + (Bw * Bw)));
      return Bn * (std::exp(-((T - Tpk) / Tw) * ((T - Tpk) / Tw)));
    }
    if (c.zetaSparam == 3) {
      const double Tpk2 = 0.160, Bn2 = 0.13, B1 = 0.01, B2 = 0.12;
      if (T < Tpk2) return Bn2 * std::exp(-((T - Tpk2) * (T - Tpk2) / (B1 * B1)));
      return Bn2 * std::exp(-((T - Tpk2) * (T - Tpk2) / (B2 * B2)));
    }
    if (c.zetaSparam == 4) {
      // operator precedence kept as in src/trancoeff.cpp:62,64
      if (std::fabs(s) < 1e-10 || T < 1e-1) return 0;
      const double sig = e < c.zetaSPeakEpsilon ? c.zetaSSigmaMinus : c.zetaSSigmaPlus;
      const double d = c.zetaSScaleBeta *
                       (std::pow(e, 0.25) - std::pow(c.zetaSPeakEpsilon, 0.25));
      return ((e + P) / (s * T)) * c.zetaS0 * std::exp((d * d) / 2.0 * (sig * sig));
    }
    if (c.zetaSparam == 5) {
      // T-peaked asymmetric Gaussian (new in this build):
      // zetaS0 * exp(-(T-Tpk)^2 / (2 w^2)), w = width below/above the peak
      if (T <= 0.) return 0.;
      const double w = T < c.zetaSPeakT ? c.zetaSWidthMinus : c.zetaSWidthPlus;
      const double d = (T - c.zetaSPeakT) / w;
      return c.zetaS0 * std::exp(-0.5 * d * d);
    }
    return 0.;
  }
  double etaS(double e, double rho, double T, double muB, double s, double P) const {
    if (c.etaSparam == 0) return c.etaS0;
    if (c.etaSparam == 1) return c.etaSMin + ((T > c.T0) ? c.ah * (T - c.T0) : c.al * (T - c.T0));
    if (c.etaSparam == 2)
      return std::max(0.0, c.etaSMin + ((e > c.etaSEpsilonMin)
                                            ? ((c.ah * (e - c.etaSEpsilonMin) + c.aRho * rho))
                                            : c.al * (e - c.etaSEpsilonMin) + c.aRho * rho));
    if (c.etaSparam == 3) {
      if (std::fabs(s) < 1e-10 || T < 1e-1) return 0;
      const double Tc = c.T0 + c.etaSShiftMuB * (muB / 0.938);
      return ((e + P) / (s * T)) *
             std::max(0.0, (c.etaSMin + ((T > Tc) ? c.ah * (T - Tc) : c.al * (T - Tc))) *
                               (1 + c.etaSScaleMuB * (muB / 0.938)));
    }
    return 0.;
  }
  void getEta(double e, double rho, double T, double muB, double s, double P,
              double &es, double &zs) const {
    es = etaS(e, rho, T, muB, s, P);
    zs = zetaS(e, T, s, P);
  }
  // getTau (src/trancoeff.cpp:92-100); abs() of a double resolves to |x| in
  // the reference build (GSL's header pulls in <stdlib.h>).
  void getTau(double e, double rho, double T, double muB, double s, double P,
              double &taupi, double &tauPi) const {
    if (T > 1e-1 && std::fabs(e + P) > 1.e-10) {
      taupi = (5. / 5.068) * s * etaS(e, rho, T, muB, s, P) / (e + P);
      tauPi = (6. / 5.068) * zetaS(e, T, s, P) * s / (e + P);
    } else {
      taupi = tauPi = 0.;
    }
  }
  bool isViscous() const {
    return c.etaS0 > 0. || c.zetaS0 > 0. || c.etaSparam > 0 || c.zetaSparam > 0;
  }
};

// ------------------------------------------------------------------- cells
// The per-cell record of src/cll.h:42-61 (AoS, as in the reference).
struct OCell {
  double Q[7], Qh[7], Qprev[7];
  double pi[10], piH[10], Pi, PiH;
  double pi0[10], piH0[10], Pi0, PiH0;
  double flux[7];
  double m[3], dm[3];
  double vcc;
  double db[16];  // Cell::dbeta (src/cll.h:50), [i][j] at 4 i + j, as setDbeta left it
  OCell() {
    std::memset(this, 0, sizeof(*this));
    vcc = 1.;
  }
  double maxM() const { return std::max(m[0], std::max(m[1], m[2])); }
  void add_m(int d, double inc) {  // Cell::addM (src/cll.h:146-150)
    m[d] += inc;
    if (m[d] > 0.9) m[0] = m[1] = m[2] = 1.;
  }
  void add_flux(const double f[7]) {
    for (int i = 0; i < 7; i++) flux[i] += f[i];
  }
  void clear_flux() {
    for (int i = 0; i < 7; i++) flux[i] = 0.;
  }
  // Cell::setQ / setQh (src/cll.h:75-86)
  static void set_clamped(double dst[7], const double src[7]) {
    for (int i = 0; i < 7; i++) dst[i] = src[i];
    if (dst[kT] < 0.)
      for (int i = 0; i < 7; i++) dst[i] = 0.;
  }
};

struct Grid {
  int nx, ny, nz;
  double minx, maxx, miny, maxy, minz, maxz, dx, dy, dz;
  std::vector<OCell> c;
  // Fluid::getCell with index clamping (src/fld.h:70-78)
  OCell &at(int ix, int iy, int iz) {
    ix = ix > 0 ? ix : 0;
    ix = ix < nx ? ix : nx - 1;
    iy = iy > 0 ? iy : 0;
    iy = iy < ny ? iy : ny - 1;
    iz = iz > 0 ? iz : 0;
    iz = iz < nz ? iz : nz - 1;
    return c[ix + nx * iy + nx * ny * iz];
  }
  // neighbour along axis d (1..3) at offset s, clamped (src/fld.cpp:89-99)
  OCell &nb(int ix, int iy, int iz, int d, int s) {
    return at(ix + (d == 1 ? s : 0), iy + (d == 2 ? s : 0), iz + (d == 3 ? s : 0));
  }
};

// ------------------------------------------------------------------- hydro
struct Solver {
  Grid g;
  Eos eos;
  Trc trc;
  double dt = 0, tau = 0;
  int status = 0;  // 0 ok, 1 NaN exit(1), 2 e>1e10 exit(0)
  bool vort = false;  // Hydro::enableVorticity (src/hdo.cpp:60-65)
  // Branch counters (test instrumentation, no effect on the results):
  // [0], [1] negative-pressure re-evaluations of hlle_flux in the predictor
  // and the corrector (src/hdo.cpp:134-140), [2] cells whose pi was rescaled
  // by the ISformal limiter (src/hdo.cpp:925-932), [3] cells whose Pi was
  // clamped to +-p (:933-937).
  long counts[4] = {0, 0, 0, 0};
  // HLLE signal speeds max(-bl, br) of the interfaces that formed a flux
  // (src/hdo.cpp:181-188, 210-227, 255-264): [0..5] the running step's
  // predictor/corrector maxima per axis, [6..11] the last step's, [12..14]
  // the per-axis maxima since the last reset.
  double sig[15] = {0};

  // Cell::getPrimVar{Left,Right,HLeft,HRight} (src/cll.cpp:106-204):
  // MUSCL half-slope reconstruction of Q (or Qh), face state (Q -+ dQ)/tau.
  void face_state(int ix, int iy, int iz, int d, bool useQh, int side,
                  double tau_, double prim[8]) {
    OCell &c = g.at(ix, iy, iz);
    OCell &nx_ = g.nb(ix, iy, iz, d, +1);
    OCell &pv = g.nb(ix, iy, iz, d, -1);
    const double *q = useQh ? c.Qh : c.Q;
    const double *qr = useQh ? nx_.Qh : nx_.Q;
    const double *ql = useQh ? pv.Qh : pv.Q;
    double dQ[7], s[7];
    for (int i = 0; i < 7; i++) dQ[i] = minmod((qr[i] - q[i]) / 2., (q[i] - ql[i]) / 2.);
    if (side < 0)
      for (int i = 0; i < 7; i++) s[i] = (q[i] - dQ[i]) / tau_;
    else
      for (int i = 0; i < 7; i++) s[i] = (q[i] + dQ[i]) / tau_;
    con2prim(eos, s, prim[0], prim[1], prim[2], prim[3], prim[4], prim[5], prim[6], prim[7]);
  }

  // Hydro::hlle_flux (src/hdo.cpp:77-305); the left cell is (ix,iy,iz), the
  // right cell its +1 neighbour along axis d.
  void hlle(int ix, int iy, int iz, int d, int mode) {
    OCell &L = g.at(ix, iy, iz);
    OCell &R = g.nb(ix, iy, iz, d, +1);
    const int rx = ix + (d == 1), ry = iy + (d == 2), rz = iz + (d == 3);
    const double dta = mode == 0 ? dt / 2. : dt;
    double l[8], r[8];
    const bool h = mode != kPredict;
    face_state(ix, iy, iz, d, h, +1, tau, l);
    face_state(rx, ry, rz, d, h, -1, tau, r);
    double &el = l[0], &pl = l[1], &vxl = l[5], &vyl = l[6], &vzl = l[7];
    double &er = r[0], &pr = r[1], &vxr = r[5], &vyr = r[6], &vzr = r[7];
    const double El = (el + pl) / (1 - vxl * vxl - vyl * vyl - vzl * vzl);
    const double Er = (er + pr) / (1 - vxr * vxr - vyr * vyr - vzr * vzr);
    const double tauFactor = mode == kPredict ? tau + 0.25 * dt : tau + 0.5 * dt;
    if (el < 0.) {
      el = 0.;
      pl = 0.;
    }
    if (er < 0.) {
      er = 0.;
      pr = 0.;
    }
    if (el > 1e10) {
      status = 2;
      return;
    }
    if (el == 0. && er == 0.) return;
    if (pr < 0.) {  // the reference re-evaluates both states from Q (not Qh)
                    // in either mode (src/hdo.cpp:134-140)
      counts[mode == kPredict ? 0 : 1]++;
      face_state(ix, iy, iz, d, false, +1, tau, l);
      face_state(rx, ry, rz, d, false, -1, tau, r);
    }
    if (L.m[d - 1] < 1. && R.m[d - 1] < 1.) return;

    const double gl = 1. / std::sqrt(1 - vxl * vxl - vyl * vyl - vzl * vzl);
    const double gr = 1. / std::sqrt(1 - vxr * vxr - vyr * vyr - vzr * vzr);
    // conserved densities U (order: t, x, y, z, b, q, s)
    double Ul[7], Ur[7], Fl[7], Fr[7];
    Ul[kX] = gl * gl * (el + pl) * vxl;
    Ul[kY] = gl * gl * (el + pl) * vyl;
    Ul[kZ] = gl * gl * (el + pl) * vzl;
    Ul[kT] = gl * gl * (el + pl) - pl;
    Ul[kNB] = gl * l[2];
    Ul[kNQ] = gl * l[3];
    Ul[kNS] = gl * l[4];
    Ur[kX] = gr * gr * (er + pr) * vxr;
    Ur[kY] = gr * gr * (er + pr) * vyr;
    Ur[kZ] = gr * gr * (er + pr) * vzr;
    Ur[kT] = gr * gr * (er + pr) - pr;
    Ur[kNB] = gr * r[2];
    Ur[kNQ] = gr * r[3];
    Ur[kNS] = gr * r[4];
    const double cs2 = 1. / 3.;            // EoS::cs2() (src/eos.h:34)
    const double cs = std::sqrt(1. / 3.);  // EoS::cs()
    double bl, br, dx;
    const double vnl = d == 1 ? vxl : (d == 2 ? vyl : vzl);
    const double vnr = d == 1 ? vxr : (d == 2 ? vyr : vzr);
    if (d != 3) {
      for (int i = 0; i < 7; i++) {
        Fl[i] = (i == kT) ? Ul[kT] * vnl + pl * vnl : Ul[i] * vnl;
        Fr[i] = (i == kT) ? Ur[kT] * vnr + pr * vnr : Ur[i] * vnr;
      }
      Fl[d] = Ul[d] * vnl + pl;
      Fr[d] = Ur[d] * vnr + pr;
    } else {
      const double t1 = tauFactor;
      for (int i = 0; i < 7; i++) {
        Fl[i] = (i == kT) ? Ul[kT] * vnl / t1 + pl * vnl / t1 : Ul[i] * vnl / t1;
        Fr[i] = (i == kT) ? Ur[kT] * vnr / t1 + pr * vnr / t1 : Ur[i] * vnr / t1;
      }
      Fl[kZ] = Ul[kZ] * vnl / t1 + pl / t1;
      Fr[kZ] = Ur[kZ] * vnr / t1 + pr / t1;
    }
    const double sEl = std::sqrt(El), sEr = std::sqrt(Er);
    const double dv = vnl - vnr;
    const double sumE = std::sqrt(El) + std::sqrt(Er);
    const double csb = std::sqrt(cs2 + 0.5 * std::sqrt(El * Er) / (sumE * sumE) * (dv * dv));
    const double vb = (sEl * vnl + sEr * vnr) / (std::sqrt(El) + std::sqrt(Er));
    bl = std::min(0., std::min((vb - csb) / (1 - vb * csb), (vnl - cs) / (1 - vnl * cs)));
    br = std::max(0., std::max((vb + csb) / (1 + vb * csb), (vnr + cs) / (1 + vnr * cs)));
    if (d == 3) {
      bl = 1. / tau * bl;
      br = 1. / tau * br;
      if (el == 0.) bl = -1. / tau;
      if (er == 0.) br = 1. / tau;
      dx = g.dz;
    } else {
      if (el == 0.) bl = -1.;
      if (er == 0.) br = 1.;
      dx = d == 1 ? g.dx : g.dy;
    }
    if (bl == 0. && br == 0.) return;
    double F[7];
    for (int i = 0; i < 7; i++)
      F[i] = tauFactor * dta / dx * (-bl * br * (Ul[i] - Ur[i]) + br * Fl[i] - bl * Fr[i]) /
             (-bl + br);
    if (F[kNB] != F[kNB]) {
      status = 1;
      return;
    }
    double mF[7];
    for (int i = 0; i < 7; i++) mF[i] = -F[i];
    L.add_flux(mF);
    R.add_flux(F);
    const double sp = std::max(-bl, br);
    double &slot = sig[3 * mode + d - 1];
    if (sp > slot) slot = sp;
  }

  // Hydro::source + source_step (src/hdo.cpp:307-353)
  void source_step(int ix, int iy, int iz, int mode) {
    OCell &c = g.at(ix, iy, iz);
    const double sdt = mode == kPredict ? dt / 2. : dt;
    const double tau1 = mode == kPredict ? tau : tau + 0.5 * dt;
    const double *Q = mode == kPredict ? c.Q : c.Qh;
    double q[7], S[7], pr[8];
    for (int i = 0; i < 7; i++) q[i] = Q[i] / tau1;
    con2prim(eos, q, pr[0], pr[1], pr[2], pr[3], pr[4], pr[5], pr[6], pr[7]);
    const double p = pr[1], vz = pr[7];
    S[kT] = -q[kT] * vz * vz - p * (1. + vz * vz);
    S[kX] = 0.;
    S[kY] = 0.;
    S[kZ] = -q[kZ];
    S[kNB] = S[kNQ] = S[kNS] = 0.;
    for (int i = 0; i < 7; i++) S[i] *= sdt;
    if (S[kNB] != S[kNB]) {
      status = 1;
      return;
    }
    c.add_flux(S);
  }

  // con2prim of a cell's Q/Qh/Qprev divided by tau (Cell::getPrimVar*,
  // src/cll.cpp:90-104,206-220)
  void prim_of(const double *Qsrc, double tau_, double pr[8]) {
    double q[7];
    for (int i = 0; i < 7; i++) q[i] = Qsrc[i] / tau_;
    con2prim(eos, q, pr[0], pr[1], pr[2], pr[3], pr[4], pr[5], pr[6], pr[7]);
  }

  // Hydro::NSquant (src/hdo.cpp:381-652).  db: the matrix dbeta[i][j] =
  // d_i (u^j / T) shared by all cells of one ISformal pass (NULL: vorticity
  // off).  Like the reference, the function leaves db untouched when it
  // returns at the vacuum check, and leaves rows 1-3 untouched when a
  // neighbour pair has e <= 0 (the Christoffel terms of row 3 are still added),
  // so such entries keep the values of the cell before in the loop order.
  void nsquant(int ix, int iy, int iz, double piNS[4][4], double &PiNS,
               double dmu[4][4], double &du, double (*db)[4] = nullptr) {
    const double VMIN = 1e-2, UDIFF = 3.0;
    const double gmn[4][4] = {{1, 0, 0, 0}, {0, -1, 0, 0}, {0, 0, -1, 0}, {0, 0, 0, -1}};
    OCell &c = g.at(ix, iy, iz);
    if (g.at(ix + 1, iy, iz).maxM() <= 0.9 || g.at(ix, iy + 1, iz).maxM() <= 0.9 ||
        g.at(ix - 1, iy, iz).maxM() <= 0.9 || g.at(ix, iy - 1, iz).maxM() <= 0.9 ||
        g.at(ix + 1, iy + 1, iz).maxM() <= 0.9 || g.at(ix + 1, iy - 1, iz).maxM() <= 0.9 ||
        g.at(ix - 1, iy + 1, iz).maxM() <= 0.9 || g.at(ix - 1, iy - 1, iz).maxM() <= 0.9) {
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) piNS[i][j] = dmu[i][j] = 0.;
      PiNS = du = 0.;
      return;
    }
    double p0[8], p1[8], ph[8];
    prim_of(c.Qprev, tau - dt, p0);
    prim_of(c.Q, tau, p1);
    prim_of(c.Qh, tau - 0.5 * dt, ph);
    double e0 = p0[0], e1 = ph[0];  // e1 is overwritten by the H-centre state
    const double s = eos.s(ph[0], ph[2], ph[3], ph[4]);
    double T, mub, muq, mus, p;
    eos.eos(ph[0], ph[2], ph[3], ph[4], T, mub, muq, mus, p);
    double etaS, zetaS;
    trc.getEta(ph[0], ph[2], T, mub, s, p, etaS, zetaS);
    double ut0 = 1.0 / std::sqrt(1.0 - p0[5] * p0[5] - p0[6] * p0[6] - p0[7] * p0[7]);
    double ux0 = ut0 * p0[5], uy0 = ut0 * p0[6], uz0 = ut0 * p0[7];
    double ut1 = 1.0 / std::sqrt(1.0 - p1[5] * p1[5] - p1[6] * p1[6] - p1[7] * p1[7]);
    double ux1 = ut1 * p1[5], uy1 = ut1 * p1[6], uz1 = ut1 * p1[7];
    double Tv0 = 0., Tv1 = 0., mu_[3], pv;
    if (db) {  // temperatures of the previous and current states (src/hdo.cpp:420-425)
      eos.eos(p0[0], p0[2], p0[3], p0[4], Tv0, mu_[0], mu_[1], mu_[2], pv);
      eos.eos(p1[0], p1[2], p1[3], p1[4], Tv1, mu_[0], mu_[1], mu_[2], pv);
      // d_tau (u^nu / T) (src/hdo.cpp:451-458)
      db[0][0] = (ut1 / Tv1 - ut0 / Tv0) / dt;
      db[0][1] = (ux1 / Tv1 - ux0 / Tv0) / dt;
      db[0][2] = (uy1 / Tv1 - uy0 / Tv0) / dt;
      db[0][3] = (uz1 / Tv1 - uz0 / Tv0) / dt;
      if (e1 <= 0. || e0 <= 0. || Tv1 <= 0. || Tv0 <= 0.)
        db[0][0] = db[0][1] = db[0][2] = db[0][3] = 0.;
    }
    double u[4];
    u[0] = 1.0 / std::sqrt(1.0 - ph[5] * ph[5] - ph[6] * ph[6] - ph[7] * ph[7]);
    u[1] = u[0] * ph[5];
    u[2] = u[0] * ph[6];
    u[3] = u[0] * ph[7];
    // time derivatives
    dmu[0][0] = (ut1 * ut1 - ut0 * ut0) / 2. / u[0] / dt;
    dmu[0][1] = (ux1 * ux1 - ux0 * ux0) / 2. / u[1] / dt;
    dmu[0][2] = (uy1 * uy1 - uy0 * uy0) / 2. / u[2] / dt;
    dmu[0][3] = (uz1 * uz1 - uz0 * uz0) / 2. / u[3] / dt;
    if (std::fabs(0.5 * (ut1 + ut0) / ut1) > UDIFF) dmu[0][0] = (ut1 - ut0) / dt;
    if (std::fabs(u[1]) < VMIN || std::fabs(0.5 * (ux1 + ux0) / ux1) > UDIFF)
      dmu[0][1] = (ux1 - ux0) / dt;
    if (std::fabs(u[2]) < VMIN || std::fabs(0.5 * (uy1 + uy0) / uy1) > UDIFF)
      dmu[0][2] = (uy1 - uy0) / dt;
    if (std::fabs(u[3]) < VMIN || std::fabs(0.5 * (uz1 + uz0) / uz1) > UDIFF)
      dmu[0][3] = (uz1 - uz0) / dt;
    if (e1 <= 0. || e0 <= 0.) dmu[0][0] = dmu[0][1] = dmu[0][2] = dmu[0][3] = 0.;
    // space derivatives, neighbours' Qh / tau_{n+1}
    for (int d = 1; d <= 3; d++) {
      double a[8], b[8];
      prim_of(g.nb(ix, iy, iz, d, +1).Qh, tau, a);
      prim_of(g.nb(ix, iy, iz, d, -1).Qh, tau, b);
      e1 = a[0];
      e0 = b[0];
      if (db) {  // neighbour temperatures (src/hdo.cpp:474-481)
        eos.eos(e1, a[2], a[3], a[4], Tv1, mu_[0], mu_[1], mu_[2], pv);
        eos.eos(e0, b[2], b[3], b[4], Tv0, mu_[0], mu_[1], mu_[2], pv);
      }
      if (e1 > 0. && e0 > 0.) {
        ut0 = 1.0 / std::sqrt(1.0 - b[5] * b[5] - b[6] * b[6] - b[7] * b[7]);
        ux0 = ut0 * b[5];
        uy0 = ut0 * b[6];
        uz0 = ut0 * b[7];
        ut1 = 1.0 / std::sqrt(1.0 - a[5] * a[5] - a[6] * a[6] - a[7] * a[7]);
        ux1 = ut1 * a[5];
        uy1 = ut1 * a[6];
        uz1 = ut1 * a[7];
        if (db) {  // d_x, d_y, d_eta (u^nu / T) (src/hdo.cpp:495-503,541-549,587-595)
          const double den = d == 1 ? g.dx : d == 2 ? g.dy : g.dz * (tau + 0.5 * dt);
          db[d][0] = 0.5 * (ut1 / Tv1 - ut0 / Tv0) / den;
          db[d][1] = 0.5 * (ux1 / Tv1 - ux0 / Tv0) / den;
          db[d][2] = 0.5 * (uy1 / Tv1 - uy0 / Tv0) / den;
          db[d][3] = 0.5 * (uz1 / Tv1 - uz0 / Tv0) / den;
          if (e1 <= 0. || e0 <= 0. || Tv1 <= 0. || Tv0 <= 0.)
            db[d][0] = db[d][1] = db[d][2] = db[d][3] = 0.;
        }
        if (d < 3) {
          const double h = d == 1 ? g.dx : g.dy;
          dmu[d][0] = 0.25 * (ut1 * ut1 - ut0 * ut0) / u[0] / h;
          dmu[d][1] = 0.25 * (ux1 * ux1 - ux0 * ux0) / u[1] / h;
          dmu[d][2] = 0.25 * (uy1 * uy1 - uy0 * uy0) / u[2] / h;
          dmu[d][3] = 0.25 * (uz1 * uz1 - uz0 * uz0) / u[3] / h;
          if (std::fabs(0.5 * (ut1 + ut0) / u[0]) > UDIFF) dmu[d][0] = 0.5 * (ut1 - ut0) / h;
          if (std::fabs(u[1]) < VMIN || std::fabs(0.5 * (ux1 + ux0) / u[1]) > UDIFF)
            dmu[d][1] = 0.5 * (ux1 - ux0) / h;
          if (std::fabs(u[2]) < VMIN || std::fabs(0.5 * (uy1 + uy0) / u[2]) > UDIFF)
            dmu[d][2] = 0.5 * (uy1 - uy0) / h;
          if (std::fabs(u[3]) < VMIN || std::fabs(0.5 * (uz1 + uz0) / u[3]) > UDIFF)
            dmu[d][3] = 0.5 * (uz1 - uz0) / h;
        } else {
          const double dz = g.dz, tz = tau + 0.5 * dt;
          dmu[3][0] = 0.25 * (ut1 * ut1 - ut0 * ut0) / u[0] / dz / tz;
          dmu[3][1] = 0.25 * (ux1 * ux1 - ux0 * ux0) / u[1] / dz / tz;
          dmu[3][2] = 0.25 * (uy1 * uy1 - uy0 * uy0) / u[2] / dz / tz;
          dmu[3][3] = 0.25 * (uz1 * uz1 - uz0 * uz0) / u[3] / dz / tz;
          if (std::fabs(0.5 * (ut1 + ut0) / u[0]) > UDIFF) dmu[3][0] = 0.5 * (ut1 - ut0) / dz / tz;
          if (std::fabs(u[1]) < VMIN || std::fabs(0.5 * (ux1 + ux0) / u[1]) > UDIFF)
            dmu[3][1] = 0.5 * (ux1 - ux0) / dz / tz;
          if (std::fabs(u[2]) < VMIN || std::fabs(0.5 * (uy1 + uy0) / u[2]) > UDIFF)
            dmu[3][2] = 0.5 * (uy1 - uy0) / dz / tz;
          if (std::fabs(u[3]) < VMIN || std::fabs(0.5 * (uz1 + uz0) / u[3]) > UDIFF)
            dmu[3][3] = 0.5 * (uz1 - uz0) / dz / tz;
        }
      } else {
        dmu[d][0] = dmu[d][1] = dmu[d][2] = dmu[d][3] = 0.;
      }
    }
    // Christoffel terms (src/hdo.cpp:606-607)
    dmu[3][0] += u[3] / (tau - 0.5 * dt);
    dmu[3][3] += u[0] / (tau - 0.5 * dt);
    if (db && T > 0.) {  // src/hdo.cpp:608-611
      db[3][0] += u[3] / (T * (tau - 0.5 * dt));
      db[3][3] += u[0] / (T * (tau - 0.5 * dt));
    }
    // projector Z (src/hdo.cpp:613-628) and sigma contraction (:630-637)
    static thread_local double Z[4][4][4][4];
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++)
        for (int k = 0; k < 4; k++)
          for (int l = 0; l < 4; l++) Z[i][j][k][l] = 0.0;
    for (int mu = 0; mu < 4; mu++)
      for (int nu = 0; nu < 4; nu++)
        for (int lam = 0; lam < 4; lam++)
          for (int rho = 0; rho < 4; rho++) {
            if (nu == rho) Z[mu][nu][lam][rho] += 0.5 * (gmn[mu][lam] - u[mu] * u[lam]);
            if (mu == rho) Z[mu][nu][lam][rho] += 0.5 * (gmn[nu][lam] - u[nu] * u[lam]);
            if (lam == rho) Z[mu][nu][lam][rho] -= (gmn[mu][nu] - u[mu] * u[nu]) / 3.0;
          }
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        piNS[i][j] = 0.0;
        for (int k = 0; k < 4; k++)
          for (int l = 0; l < 4; l++)
            piNS[i][j] += Z[i][j][k][l] * dmu[k][l] * 2.0 * etaS * s / 5.068;
      }
    PiNS = -zetaS * s * (dmu[0][0] + dmu[1][1] + dmu[2][2] + dmu[3][3]) / 5.068;
    du = dmu[0][0] + dmu[1][1] + dmu[2][2] + dmu[3][3];
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++)
        if (std::isinf(piNS[i][j]) || std::isnan(piNS[i][j])) status = 1;
  }

  // Hydro::ISformal (src/hdo.cpp:682-963), vorticity off, no FORMAL_SOLUTION
  void isformal() {
    const double gm[4] = {1., -1., -1., -1.};
    double dbs[4][4] = {};  // the pass's shared dbeta matrix (src/hdo.cpp:685-687)
    // loop #1: relaxation + sources
    for (int ix = 0; ix < g.nx; ix++)
      for (int iy = 0; iy < g.ny; iy++)
        for (int iz = 0; iz < g.nz; iz++) {
          OCell &c = g.at(ix, iy, iz);
          double pr[8];
          prim_of(c.Qh, tau - dt / 2., pr);
          double e = pr[0], nb = pr[2], nq = pr[3], ns = pr[4];
          const double vx = pr[5], vy = pr[6], vz = pr[7];
          if (e <= 0.) {
            for (int i = 0; i < 10; i++) c.piH0[i] = c.pi0[i] = 0.0;
            c.PiH0 = c.Pi0 = 0.0;
            if (vort)  // Cell::resetDbeta (src/cll.h:186-196)
              for (int k = 0; k < 16; k++) c.db[k] = 0.0;
            continue;
          }
          const double gamma = 1.0 / std::sqrt(1.0 - vx * vx - vy * vy - vz * vz);
          double u[4] = {gamma, gamma * vx, gamma * vy, gamma * vz};
          double fl[7] = {0, 0, 0, 0, 0., 0., 0.};
          for (int i = 0; i < 4; i++) fl[i] = (tau - dt) * (c.pi[sym(0, i)] + c.Pi * u[0] * u[i]);
          fl[0] += -(tau - dt) * c.Pi;
          c.add_flux(fl);
          double piNS[4][4], PiNS, dmu[4][4], du;
          nsquant(ix, iy, iz, piNS, PiNS, dmu, du, vort ? dbs : nullptr);
          if (status) return;
          if (vort)  // Cell::setDbeta: copy clamped to [-100, 100] (src/cll.h:173-183)
            for (int i = 0; i < 4; i++)
              for (int j = 0; j < 4; j++)
                c.db[4 * i + j] = std::max(std::min(dbs[i][j], 100.), -100.);
          double T, mub, muq, mus, p;
          eos.eos(e, nb, nq, ns, T, mub, muq, mus, p);
          double etaS, zetaS;
          const double s = eos.s(e, nb, nq, ns);
          trc.getEta(e, nb, T, mub, s, p, etaS, zetaS);
          const double eta = etaS * s;
          double sig[4][4];
          for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
              sig[i][j] = 0.5 * piNS[i][j] / eta * 5.068;
              if (eta <= 0.0) sig[i][j] = 0.0;
            }
          double taupi, tauPi;
          trc.getTau(e, nb, T, mub, s, p, taupi, tauPi);
          // getOther / getOtherBulk (src/trancoeff.h:28-37)
          double dpp = 4. / 3., tpp = 10. / 7., lpP = 6. / 5.;
          double phi7 = 9. / 70. / eos.p(e, nb, nq, ns);
          if (std::isinf(phi7)) phi7 = 0.0;
          phi7 = phi7 / taupi;
          if (taupi < 0.5 * dt) dpp = tpp = lpP = phi7 = 0.0;
          double dPP = 2. / 3., lPp = 8. / 5. * (1. / 3. - eos.cs2e(e));
          if (tauPi < 0.5 * dt) dPP = lPp = 0.0;
          double Delta[10];
          const double *pi = c.pi;
          double *pH = c.piH0;
          // half step
          for (int i = 0; i < 4; i++)
            for (int j = 0; j <= i; j++) {
              Delta[sym(i, j)] = -u[i] * u[j];
              if (i == j) Delta[sym(i, j)] += gm[i];
              if (taupi > 0.5 * dt)
                pH[sym(i, j)] = pi[sym(i, j)] - (pi[sym(i, j)] - piNS[i][j]) * dt / 2.0 / gamma / taupi;
              else
                pH[sym(i, j)] = piNS[i][j];
            }
          if (tauPi > 0.5 * dt)
            c.PiH0 = c.Pi - (c.Pi - PiNS) * dt / 2.0 / gamma / tauPi;
          else
            c.PiH0 = PiNS;
          double tau1 = tau - dt * 0.75;
          pH[sym(0, 0)] += -2. * vz * pi[sym(0, 3)] / tau1 * dt / 2.;
          pH[sym(3, 3)] += -(2. * vz / tau1 * pi[sym(0, 3)]) * dt / 2.;
          pH[sym(3, 0)] += -(vz / tau1 * pi[sym(0, 0)] + vz / tau1 * pi[sym(3, 3)]) * dt / 2.;
          pH[sym(1, 0)] += -vz / tau1 * pi[sym(1, 3)] * dt / 2.;
          pH[sym(2, 0)] += -vz / tau1 * pi[sym(2, 3)] * dt / 2.;
          pH[sym(3, 1)] += -(vz / tau1 * pi[sym(0, 1)]) * dt / 2.;
          pH[sym(3, 2)] += -(vz / tau1 * pi[sym(0, 2)]) * dt / 2.;
          for (int i = 0; i < 4; i++)
            for (int j = 0; j <= i; j++) {
              double &t = pH[sym(i, j)];
              t += (-dpp * pi[sym(i, j)] * du + lpP * c.Pi * sig[i][j]) / gamma * 0.5 * dt;
              for (int k = 0; k < 4; k++) {
                t += (phi7 * pi[sym(i, k)] * pi[sym(j, k)] * gm[k] -
                      tpp * 0.5 * (pi[sym(i, k)] * sig[j][k] * gm[k] + pi[sym(j, k)] * sig[i][k] * gm[k])) /
                     gamma * 0.5 * dt;
                for (int l = 0; l < 4; l++)
                  t += (-pi[sym(i, k)] * u[j] - pi[sym(j, k)] * u[i]) * u[l] * dmu[l][k] * gm[k] / gamma * 0.5 * dt -
                       1. / 3. * Delta[sym(i, j)] * pi[sym(k, l)] * (phi7 * pi[sym(k, l)] - tpp * sig[k][l]) *
                           gm[k] * gm[l] / gamma * 0.5 * dt;
              }
            }
          for (int k = 0; k < 4; k++)
            for (int l = 0; l < 4; l++)
              c.PiH0 += lPp * pi[sym(k, l)] * sig[k][l] * gm[k] * gm[l] / gamma * 0.5 * dt;
          c.PiH0 += -dPP * c.Pi * du / gamma * 0.5 * dt;
          // full step
          double *p0 = c.pi0;
          for (int i = 0; i < 4; i++)
            for (int j = 0; j <= i; j++) {
              if (taupi > 0.5 * dt)
                p0[sym(i, j)] = pi[sym(i, j)] - (pH[sym(i, j)] - piNS[i][j]) * dt / gamma / taupi;
              else
                p0[sym(i, j)] = piNS[i][j];
            }
          if (tauPi > 0.5 * dt)
            c.Pi0 = c.Pi - (c.PiH0 - PiNS) * dt / gamma / tauPi;
          else
            c.Pi0 = PiNS;
          tau1 = tau - dt * 0.5;
          p0[sym(0, 0)] += -2. * vz / tau1 * pH[sym(0, 3)] * dt;
          p0[sym(3, 3)] += -(2. * vz / tau1 * pH[sym(0, 3)]) * dt;
          p0[sym(3, 0)] += -(vz / tau1 * pH[sym(0, 0)] + vz / tau1 * pH[sym(3, 3)]) * dt;
          p0[sym(1, 0)] += -vz / tau1 * pH[sym(1, 3)] * dt;
          p0[sym(2, 0)] += -vz / tau1 * pH[sym(2, 3)] * dt;
          p0[sym(3, 1)] += -(vz / tau1 * pH[sym(0, 1)]) * dt;
          p0[sym(3, 2)] += -(vz / tau1 * pH[sym(0, 2)]) * dt;
          for (int i = 0; i < 4; i++)
            for (int j = 0; j <= i; j++) {
              double &t = p0[sym(i, j)];
              t += (-dpp * pH[sym(i, j)] * du + lpP * c.PiH0 * sig[i][j]) / gamma * dt;
              for (int k = 0; k < 4; k++) {
                t += (phi7 * pi[sym(i, k)] * pH[sym(j, k)] * gm[k] -
                      tpp * 0.5 * (pH[sym(i, k)] * sig[j][k] * gm[k] + pH[sym(j, k)] * sig[i][k] * gm[k])) /
                     gamma * dt;
                for (int l = 0; l < 4; l++)
                  t += ((-pH[sym(i, k)] * u[j] - pH[sym(j, k)] * u[i]) * u[l] * dmu[l][k] * gm[k] -
                        1. / 3. * Delta[sym(i, j)] * pH[sym(k, l)] * (phi7 * pH[sym(k, l)] - tpp * sig[k][l]) *
                            gm[k] * gm[l]) /
                       gamma * dt;
              }
            }
          for (int k = 0; k < 4; k++)
            for (int l = 0; l < 4; l++)
              c.Pi0 += lPp * pH[sym(k, l)] * sig[k][l] * gm[k] * gm[l] / gamma * dt;
          c.Pi0 += -dPP * c.PiH0 * du / gamma * dt;
        }
    advect();
  }

  // ISformal loop #2 (src/hdo.cpp:866-962): upwind advection of pi0, piH0,
  // Pi0, PiH0 from the 8 cells toward -sign(v), the limiter, the bulk clamp
  // and the -tau dQ_{n+1} flux term
  void advect() {
    for (int ix = 0; ix < g.nx; ix++)
      for (int iy = 0; iy < g.ny; iy++)
        for (int iz = 0; iz < g.nz; iz++) {
          OCell &c = g.at(ix, iy, iz);
          double pr[8];
          prim_of(c.Qh, tau - 0.5 * dt, pr);
          const double e = pr[0], p = pr[1], vx = pr[5], vy = pr[6], vz = pr[7];
          if (e <= 0.) continue;
          const double xm = -vx * dt / g.dx;
          const double ym = -vy * dt / g.dy;
          const double zm = -vz * dt / g.dz / (tau - 0.5 * dt);
          const double xmH = -vx * dt / g.dx / 2.0;
          const double ymH = -vy * dt / g.dy / 2.0;
          const double zmH = -vz * dt / g.dz / (tau - 0.5 * dt) / 2.0;
          const double wx[2] = {(1. - std::fabs(xm)), std::fabs(xm)};
          const double wy[2] = {(1. - std::fabs(ym)), std::fabs(ym)};
          const double wz[2] = {(1. - std::fabs(zm)), std::fabs(zm)};
          const double wxH[2] = {(1. - std::fabs(xmH)), std::fabs(xmH)};
          const double wyH[2] = {(1. - std::fabs(ymH)), std::fabs(ymH)};
          const double wzH[2] = {(1. - std::fabs(zmH)), std::fabs(zmH)};
          double pi[4][4], piH[4][4], Pi = 0.0, PiH = 0.0;
          for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) pi[i][j] = piH[i][j] = 0.0;
          for (int jx = 0; jx < 2; jx++)
            for (int jy = 0; jy < 2; jy++)
              for (int jz = 0; jz < 2; jz++) {
                OCell &c1 = g.at(ix + jx * sgn(xm), iy + jy * sgn(ym), iz + jz * sgn(zm));
                for (int i = 0; i < 4; i++)
                  for (int j = 0; j < 4; j++) {
                    pi[i][j] += wx[jx] * wy[jy] * wz[jz] * c1.pi0[sym(i, j)];
                    piH[i][j] += wxH[jx] * wyH[jy] * wzH[jz] * c1.piH0[sym(i, j)];
                  }
                Pi += wx[jx] * wy[jy] * wz[jz] * c1.Pi0;
                PiH += wxH[jx] * wyH[jy] * wzH[jz] * c1.PiH0;
              }
          const double maxT0 = std::max((e + p) / (1. - vx * vx - vy * vy - vz * vz) - p,
                                        (e + p) * (vx * vx + vy * vy + vz * vz) /
                                                (1. - vx * vx - vy * vy - vz * vz) +
                                            p);
          double maxpi = 0.;
          for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
              if (std::fabs(pi[i][j]) > maxpi) maxpi = std::fabs(pi[i][j]);
          bool rescaled = false;
          if (maxT0 / maxpi < 1.0) {
            for (int i = 0; i < 4; i++)
              for (int j = 0; j < 4; j++) {
                pi[i][j] = 0.1 * pi[i][j] * maxT0 / maxpi;
                piH[i][j] = 0.1 * piH[i][j] * maxT0 / maxpi;
              }
            rescaled = true;
            counts[2]++;
          }
          if (std::fabs(Pi) > p) {
            counts[3]++;
            if (Pi != 0.) Pi = 1 * Pi / std::fabs(Pi) * p;
            if (PiH != 0.) PiH = 1 * PiH / std::fabs(PiH) * p;
            rescaled = true;
          }
          c.vcc = rescaled ? maxT0 / maxpi : 1.;
          for (int i = 0; i < 4; i++)
            for (int j = 0; j <= i; j++) {
              c.pi[sym(i, j)] = pi[i][j];
              c.piH[sym(i, j)] = piH[i][j];
            }
          c.Pi = Pi;
          c.PiH = PiH;
          const double gamma = 1.0 / std::sqrt(1.0 - vx * vx - vy * vy - vz * vz);
          const double u[4] = {gamma, gamma * vx, gamma * vy, gamma * vz};
          double fl[7] = {0, 0, 0, 0, 0., 0., 0.};
          for (int i = 0; i < 4; i++) fl[i] = -tau * (c.pi[sym(0, i)] + c.Pi * u[0] * u[i]);
          fl[0] += tau * c.Pi;
          c.add_flux(fl);
        }
  }

  // Hydro::visc_flux (src/hdo.cpp:966-1014)
  void visc_flux(int ix, int iy, int iz, int d) {
    OCell &L = g.at(ix, iy, iz);
    OCell &R = g.nb(ix, iy, iz, d, +1);
    if (L.m[d - 1] < 1. && R.m[d - 1] < 1.) return;
    double dxa = d == 1 ? g.dx : (d == 2 ? g.dy : g.dz * (tau + 0.5 * dt));
    double a[8], b[8];
    prim_of(L.Qh, tau - dt / 2., a);
    prim_of(R.Qh, tau - dt / 2., b);
    double vx = 0.5 * (a[5] + b[5]);
    double vy = 0.5 * (a[6] + b[6]);
    double vz = 0.5 * (a[7] + b[7]);
    const double v = std::sqrt(vx * vx + vy * vy + vz * vz);
    if (v > 1.) {
      vx = 0.99 * vx / v;
      vy = 0.99 * vy / v;
      vz = 0.99 * vz / v;
    }
    const double gamma = 1. / std::sqrt(1. - v * v);
    const double u[4] = {gamma, gamma * vx, gamma * vy, gamma * vz};
    const double gm[4] = {1., -1., -1., -1.};
    double F[7] = {0, 0, 0, 0, 0., 0., 0.};
    for (int m = 0; m < 4; m++) {
      F[m] = 0.5 * (L.piH[sym(m, d)] + R.piH[sym(m, d)]);
      if (m == d) F[m] += -0.5 * (L.PiH + R.PiH) * gm[m];
      F[m] += 0.5 * (L.PiH + R.PiH) * u[m] * u[d];
    }
    for (int m = 0; m < 4; m++) F[m] = F[m] * (tau - 0.5 * dt) * dt / dxa;
    double mF[7] = {-F[0], -F[1], -F[2], -F[3], 0., 0., 0.};
    double pF[7] = {F[0], F[1], F[2], F[3], 0., 0., 0.};
    L.add_flux(mF);
    R.add_flux(pF);
  }

  // Hydro::visc_source_step (src/hdo.cpp:355-375)
  void visc_source_step(int ix, int iy, int iz) {
    OCell &c = g.at(ix, iy, iz);
    double pr[8];
    prim_of(c.Qh, tau - dt / 2., pr);
    if (pr[0] <= 0.) return;
    const double vx = pr[5], vy = pr[6], vz = pr[7];
    double u[4];
    u[0] = 1. / std::sqrt(1. - vx * vx - vy * vy - vz * vz);
    u[1] = u[0] * vx;
    u[2] = u[0] * vy;
    u[3] = u[0] * vz;
    double k[7] = {0, 0, 0, 0, 0., 0., 0.};
    k[kT] = -c.piH[sym(3, 3)] + c.PiH * (-1.0 - u[3] * u[3]);
    k[kX] = 0.;
    k[kY] = 0.;
    k[kZ] = -(c.piH[sym(0, 3)] + c.PiH * u[0] * u[3]);
    for (int i = 0; i < 4; i++) k[i] *= dt;
    c.add_flux(k);
  }

  // Fluid::correctImagCells (src/fld.cpp:211-273): ghost layers of Q and Qh
  // copied from the 3rd cell inwards, axes in order z, y, x.
  void ghost_ideal() {
    auto copy2 = [&](OCell &src, OCell &d1, OCell &d0) {
      double q[7], qh[7];
      std::memcpy(q, src.Q, sizeof q);
      OCell::set_clamped(d1.Q, q);
      OCell::set_clamped(d0.Q, q);
      std::memcpy(qh, src.Qh, sizeof qh);
      OCell::set_clamped(d1.Qh, qh);
      OCell::set_clamped(d0.Qh, qh);
    };
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) {
        copy2(g.at(ix, iy, 2), g.at(ix, iy, 1), g.at(ix, iy, 0));
        copy2(g.at(ix, iy, nz - 3), g.at(ix, iy, nz - 2), g.at(ix, iy, nz - 1));
      }
    for (int ix = 0; ix < nx; ix++)
      for (int iz = 0; iz < nz; iz++) {
        copy2(g.at(ix, 2, iz), g.at(ix, 1, iz), g.at(ix, 0, iz));
        copy2(g.at(ix, ny - 3, iz), g.at(ix, ny - 2, iz), g.at(ix, ny - 1, iz));
      }
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++) {
        copy2(g.at(2, iy, iz), g.at(1, iy, iz), g.at(0, iy, iz));
        copy2(g.at(nx - 3, iy, iz), g.at(nx - 2, iy, iz), g.at(nx - 1, iy, iz));
      }
  }

  // Fluid::correctImagCellsFull (src/fld.cpp:275-382): Q, pi, Pi
  void ghost_full() {
    auto copyf = [&](OCell &src, OCell &d1, OCell &d0) {
      double q[7], p[10], P;
      std::memcpy(q, src.Q, sizeof q);
      OCell::set_clamped(d1.Q, q);
      OCell::set_clamped(d0.Q, q);
      std::memcpy(p, src.pi, sizeof p);
      P = src.Pi;
      // the reference writes ghost 0 first, then ghost 1 (src/fld.cpp:288-294)
      std::memcpy(d0.pi, p, sizeof p);
      std::memcpy(d1.pi, p, sizeof p);
      d0.Pi = P;
      d1.Pi = P;
    };
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) {
        copyf(g.at(ix, iy, 2), g.at(ix, iy, 1), g.at(ix, iy, 0));
        copyf(g.at(ix, iy, nz - 3), g.at(ix, iy, nz - 2), g.at(ix, iy, nz - 1));
      }
    for (int ix = 0; ix < nx; ix++)
      for (int iz = 0; iz < nz; iz++) {
        copyf(g.at(ix, 2, iz), g.at(ix, 1, iz), g.at(ix, 0, iz));
        copyf(g.at(ix, ny - 3, iz), g.at(ix, ny - 2, iz), g.at(ix, ny - 1, iz));
      }
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++) {
        copyf(g.at(2, iy, iz), g.at(1, iy, iz), g.at(0, iy, iz));
        copyf(g.at(nx - 3, iy, iz), g.at(nx - 2, iy, iz), g.at(nx - 1, iy, iz));
      }
  }

  // Fluid::updateM (src/fld.cpp:384-423)
  void update_m() {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++)
        for (int iz = 0; iz < nz; iz++) {
          OCell &c = g.at(ix, iy, iz);
          c.dm[0] = c.dm[1] = c.dm[2] = 0.;
          if (c.maxM() < 1.) {
            if (g.at(ix + 1, iy, iz).m[0] >= 1. || g.at(ix - 1, iy, iz).m[0] >= 1.) c.dm[0] = dt / g.dx;
            if (g.at(ix, iy + 1, iz).m[1] >= 1. || g.at(ix, iy - 1, iz).m[1] >= 1.) c.dm[1] = dt / g.dy;
            if (g.at(ix, iy, iz + 1).m[2] >= 1. || g.at(ix, iy, iz - 1).m[2] >= 1.)
              c.dm[2] = dt / g.dz / tau;
            if (c.dm[0] == 0. && c.dm[1] == 0.) {
              if (g.at(ix + 1, iy + 1, iz).maxM() >= 1. || g.at(ix + 1, iy - 1, iz).maxM() >= 1. ||
                  g.at(ix - 1, iy + 1, iz).maxM() >= 1. || g.at(ix - 1, iy - 1, iz).maxM() >= 1.) {
                c.dm[0] = 0.707 * dt / g.dx;
                c.dm[1] = 0.707 * dt / g.dy;
              }
            }
          }
        }
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++)
        for (int iz = 0; iz < nz; iz++) {
          OCell &c = g.at(ix, iy, iz);
          c.add_m(0, c.dm[0]);
          c.add_m(1, c.dm[1]);
          c.add_m(2, c.dm[2]);
        }
  }

  // Hydro::performStep (src/hdo.cpp:1016-1134)
  void step() {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    update_m();
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
          OCell &c = g.at(ix, iy, iz);
          std::memcpy(c.Qprev, c.Q, sizeof c.Q);
          c.clear_flux();
        }
    for (int mode = kPredict; mode <= kCorrect; mode++) {
      for (int iy = 0; iy < ny; iy++)
        for (int iz = 0; iz < nz; iz++)
          for (int ix = 0; ix < nx - 1; ix++) {
            hlle(ix, iy, iz, 1, mode);
            if (status) return;
          }
      for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++)
          for (int iy = 0; iy < ny - 1; iy++) {
            hlle(ix, iy, iz, 2, mode);
            if (status) return;
          }
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
          for (int iz = 0; iz < nz - 1; iz++) {
            hlle(ix, iy, iz, 3, mode);
            if (status) return;
          }
      for (int iy = 0; iy < ny; iy++)
        for (int iz = 0; iz < nz; iz++)
          for (int ix = 0; ix < nx; ix++) {
            OCell &c = g.at(ix, iy, iz);
            source_step(ix, iy, iz, mode);
            if (status) return;
            if (mode == kPredict) {
              for (int i = 0; i < 7; i++) c.Qh[i] = c.Q[i] + c.flux[i];  // updateQtoQhByFlux
            } else if (!(c.Q[0] + c.flux[0] < 0.)) {                    // updateByFlux
              for (int i = 0; i < 7; i++) c.Q[i] += c.flux[i];
            }
            c.clear_flux();
          }
    }
    tau += dt;
    for (int k = 0; k < 6; k++) {
      sig[6 + k] = sig[k];
      if (sig[k] > sig[12 + k % 3]) sig[12 + k % 3] = sig[k];
      sig[k] = 0.;
    }
    ghost_ideal();
    if (trc.isViscous()) {
      isformal();
      if (status) return;
      for (int iy = 0; iy < ny; iy++)
        for (int iz = 0; iz < nz; iz++)
          for (int ix = 0; ix < nx - 1; ix++) visc_flux(ix, iy, iz, 1);
      for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++)
          for (int iy = 0; iy < ny - 1; iy++) visc_flux(ix, iy, iz, 2);
      for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
          for (int iz = 0; iz < nz - 1; iz++) visc_flux(ix, iy, iz, 3);
      for (int iy = 0; iy < ny; iy++)
        for (int iz = 0; iz < nz; iz++)
          for (int ix = 0; ix < nx; ix++) {
            visc_source_step(ix, iy, iz);
            OCell &c = g.at(ix, iy, iz);
            // Cell::updateByViscFlux (src/cll.cpp:76-84)
            if (std::fabs(c.flux[0]) <= 0.5 * c.Q[0]) {
              for (int i = 0; i < 7; i++) c.Q[i] += c.flux[i];
            } else if (c.flux[0] != 0.) {
              const double fac = std::fabs(0.5 * c.Q[0] / c.flux[0]);
              for (int i = 0; i < 7; i++) c.Q[i] += fac * c.flux[i];
            }
            c.clear_flux();
          }
    }
    ghost_full();
  }
};

}  // namespace

// ------------------------------------------------------------------- C API
extern "C" {

void *or_create(int nx, int ny, int nz, double minx, double maxx, double miny,
                double maxy, double minz, double maxz, const or_eos_t *eos,
                const or_trc_t *trc) {
  if ((eos && (eos->kind < 0 || eos->kind > 3)) || nx < 1 || ny < 1 || nz < 1) return nullptr;
  Solver *s = new Solver;
  s->eos = Eos(eos);
  Grid &g = s->g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.minx = minx;
  g.maxx = maxx;
  g.miny = miny;
  g.maxy = maxy;
  g.minz = minz;
  g.maxz = maxz;
  // Fluid ctor (src/fld.cpp:77-79); nz=1 gives dz=inf like the reference
  g.dx = (maxx - minx) / (nx - 1);
  g.dy = (maxy - miny) / (ny - 1);
  g.dz = (maxz - minz) / (nz - 1);
  g.c.assign((size_t)nx * ny * nz, OCell());
  s->trc.c = *trc;
  s->trc.eos = &s->eos;
  return s;
}

void or_destroy(void *h) { delete static_cast<Solver *>(h); }

long or_ncell(void *h) { return (long)static_cast<Solver *>(h)->g.c.size(); }

void or_set_state(void *h, const or_state_t *st) {
  Solver *s = static_cast<Solver *>(h);
  for (double &v : s->sig) v = 0.;  // a new state: no signal speeds yet
  const long N = (long)s->g.c.size();
  for (long n = 0; n < N; n++) {
    OCell &c = s->g.c[n];
    for (int k = 0; k < 7; k++) {
      if (st->Q) c.Q[k] = st->Q[k * N + n];
      if (st->Qh) c.Qh[k] = st->Qh[k * N + n];
      if (st->Qprev) c.Qprev[k] = st->Qprev[k * N + n];
    }
    for (int k = 0; k < 10; k++) {
      if (st->pi) c.pi[k] = st->pi[k * N + n];
      if (st->piH) c.piH[k] = st->piH[k * N + n];
      if (st->pi0) c.pi0[k] = st->pi0[k * N + n];
      if (st->piH0) c.piH0[k] = st->piH0[k * N + n];
    }
    if (st->Pi) c.Pi = st->Pi[n];
    if (st->PiH) c.PiH = st->PiH[n];
    if (st->Pi0) c.Pi0 = st->Pi0[n];
    if (st->PiH0) c.PiH0 = st->PiH0[n];
    for (int k = 0; k < 3; k++)
      if (st->m) c.m[k] = st->m[k * N + n];
    if (st->viscCorrCut) c.vcc = st->viscCorrCut[n];
  }
}

void or_get_state(void *h, or_state_t *st) {
  Solver *s = static_cast<Solver *>(h);
  const long N = (long)s->g.c.size();
  for (long n = 0; n < N; n++) {
    const OCell &c = s->g.c[n];
    for (int k = 0; k < 7; k++) {
      if (st->Q) st->Q[k * N + n] = c.Q[k];
      if (st->Qh) st->Qh[k * N + n] = c.Qh[k];
      if (st->Qprev) st->Qprev[k * N + n] = c.Qprev[k];
    }
    for (int k = 0; k < 10; k++) {
      if (st->pi) st->pi[k * N + n] = c.pi[k];
      if (st->piH) st->piH[k * N + n] = c.piH[k];
      if (st->pi0) st->pi0[k * N + n] = c.pi0[k];
      if (st->piH0) st->piH0[k * N + n] = c.piH0[k];
    }
    if (st->Pi) st->Pi[n] = c.Pi;
    if (st->PiH) st->PiH[n] = c.PiH;
    if (st->Pi0) st->Pi0[n] = c.Pi0;
    if (st->PiH0) st->PiH0[n] = c.PiH0;
    for (int k = 0; k < 3; k++)
      if (st->m) st->m[k * N + n] = c.m[k];
    if (st->viscCorrCut) st->viscCorrCut[n] = c.vcc;
  }
}

void or_set_vorticity(void *h, int on) {
  Solver *s = static_cast<Solver *>(h);
  s->vort = on != 0;
  for (OCell &c : s->g.c)
    for (int k = 0; k < 16; k++) c.db[k] = 0.;  // Cell::enableVorticity (src/cll.h:167-170)
}

void or_get_dbeta(void *h, double *out) {
  Solver *s = static_cast<Solver *>(h);
  const long N = (long)s->g.c.size();
  for (long n = 0; n < N; n++)
    for (int k = 0; k < 16; k++) out[k * N + n] = s->g.c[n].db[k];
}

void or_set_prim(void *h, double tau, const double *e, const double *nb,
                 const double *nq, const double *ns, const double *vx,
                 const double *vy, const double *vz) {
  Solver *s = static_cast<Solver *>(h);
  for (double &v : s->sig) v = 0.;
  const long N = (long)s->g.c.size();
  for (long n = 0; n < N; n++) {
    OCell &c = s->g.c[n];
    // Cell::setPrimVar (src/cll.cpp:222-241)
    const double g2 = 1. / (1 - vx[n] * vx[n] - vy[n] * vy[n] - vz[n] * vz[n]);
    const double p = s->eos.p(e[n], nb[n], nq[n], ns[n]);
    c.Q[kT] = tau * (e[n] + p * (vx[n] * vx[n] + vy[n] * vy[n] + vz[n] * vz[n])) * g2;
    c.Q[kX] = tau * (e[n] + p) * vx[n] * g2;
    c.Q[kY] = tau * (e[n] + p) * vy[n] * g2;
    c.Q[kZ] = tau * (e[n] + p) * vz[n] * g2;
    c.Q[kNB] = tau * nb[n] * std::sqrt(g2);
    c.Q[kNQ] = tau * nq[n] * std::sqrt(g2);
    c.Q[kNS] = tau * ns[n] * std::sqrt(g2);
    std::memcpy(c.Qprev, c.Q, sizeof c.Q);  // saveQprev
    if (e[n] > 0.) c.m[0] = c.m[1] = c.m[2] = 1.;
  }
}

int or_run(void *h, double *tau, double dt, int nsteps) {
  Solver *s = static_cast<Solver *>(h);
  s->tau = *tau;
  s->dt = dt;
  s->status = 0;
  for (int i = 0; i < nsteps && !s->status; i++) s->step();
  *tau = s->tau;
  return s->status;
}

// ---- single functions of the step, for unit tests of the restatement
void or_ghost(void *h, int full) {
  Solver *sv = static_cast<Solver *>(h);
  if (full)
    sv->ghost_full();
  else
    sv->ghost_ideal();
}

void or_update_m(void *h, double tau, double dt) {
  Solver *sv = static_cast<Solver *>(h);
  sv->tau = tau;
  sv->dt = dt;
  sv->update_m();
}

void or_advect(void *h, double tau, double dt) {
  Solver *sv = static_cast<Solver *>(h);
  sv->tau = tau;
  sv->dt = dt;
  for (auto &c : sv->g.c) c.clear_flux();
  sv->advect();
}

int or_hlle(void *h, int ix, int iy, int iz, int d, int mode, double tau, double dt,
            double *flux) {
  Solver *sv = static_cast<Solver *>(h);
  sv->tau = tau;
  sv->dt = dt;
  sv->status = 0;
  for (auto &c : sv->g.c) c.clear_flux();
  sv->hlle(ix, iy, iz, d, mode);
  const OCell &r = sv->g.at(ix + (d == 1), iy + (d == 2), iz + (d == 3));
  int any = 0;
  for (int k = 0; k < 7; k++) {
    flux[k] = r.flux[k];
    any |= r.flux[k] != 0.;
  }
  return sv->status ? -sv->status : any;
}

void or_branch_counts(void *h, long *out) {
  const Solver *sv = static_cast<Solver *>(h);
  for (int k = 0; k < 4; k++) out[k] = sv->counts[k];
}

void or_signal_speed(void *h, int reset, double *out) {
  Solver *sv = static_cast<Solver *>(h);
  for (int k = 0; k < 9; k++) out[k] = sv->sig[6 + k];
  if (reset)
    for (int k = 12; k < 15; k++) sv->sig[k] = 0.;
}

int or_check_dtau(void *h, double dt) {
  Solver *s = static_cast<Solver *>(h);
  return (dt > s->g.dx / 2. || dt > s->g.dy / 2.) ? 1 : 0;
}

void or_c2p(const or_eos_t *d, double *Q, long n, double *out, int *iters) {
  const Eos eos(d);
  for (long i = 0; i < n; i++) {
    double *o = out + 8 * i;
    con2prim(eos, Q + 7 * i, o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7]);
    if (iters) iters[i] = g_c2p_iters;
  }
}

void or_p2c(const double *prim, long n, double *Q) {
  for (long i = 0; i < n; i++) {
    const double *p = prim + 8 * i;
    prim2con(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], Q + 7 * i);
  }
}

void or_eos_spline(const or_eos_t *d, int f, const double *v, long n, double *out) {
  const Eos eos(d);
  const Spline &sp = f == 0 ? eos.sp : (f == 1 ? eos.sT : eos.smu);
  long oor = 0;
  for (long i = 0; i < n; i++) out[i] = sp.eval(v[i], oor);
}

void or_eos(const or_eos_t *d, const double *e, long n, double *out) {
  const Eos eos(d);
  for (long i = 0; i < n; i++) {
    double T, mub, muq, mus, p;
    eos.eos(e[i], 0., 0., 0., T, mub, muq, mus, p);
    out[4 * i + 0] = p;
    out[4 * i + 1] = T;
    out[4 * i + 2] = eos.s(e[i], 0., 0., 0.);
    out[4 * i + 3] = eos.cs2e(e[i]);
  }
}

void or_eos_eval(const or_eos_t *d, const double *in, long n, double *out) {
  const Eos eos(d);
  for (long i = 0; i < n; i++) {
    const double *a = in + 4 * i;
    double *o = out + 6 * i;
    eos.eos(a[0], a[1], a[2], a[3], o[1], o[2], o[3], o[4], o[0]);
    o[5] = eos.s(a[0], a[1], a[2], a[3]);
  }
}

void or_trcoeff(const or_trc_t *c, const or_eos_t *d, const double *in, long n, double *out) {
  const Eos eos(d);
  Trc t;
  t.c = *c;
  t.eos = &eos;
  for (long i = 0; i < n; i++) {
    const double *a = in + 6 * i;
    double es, zs, tp, tP;
    t.getEta(a[0], a[1], a[2], a[3], a[4], a[5], es, zs);
    t.getTau(a[0], a[1], a[2], a[3], a[4], a[5], tp, tP);
    out[4 * i + 0] = es;
    out[4 * i + 1] = zs;
    out[4 * i + 2] = tp;
    out[4 * i + 3] = tP;
  }
}

}  // extern "C"
