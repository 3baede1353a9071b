// Host side of the vHLLE class API on the GPU (see vhlle.h).
#include "vhlle.h"

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <sys/stat.h>

// ------------------------------------------------------------- helpers
static void vh_or_die(int rc, const char *what);

// Writes n records of ncol doubles as "%24.15g" columns, one record a line
// (the reference's freezeout.dat / beta.dat format).  Each number goes
// through std::to_chars (general format, precision 15: by definition the
// digits of printf's %.15g in the C locale), right-aligned to 24 characters.
// Up to 16 host threads format separate record ranges into their own buffers,
// written in record order, so the file is the same as one thread's.
// Formatting is the cost (0.67 s for 53 000 records with fprintf), not the
// write.
static void write_records(std::FILE *f, const double *rows, long n, int ncol) {
  if (!f || n <= 0) return;
  const long per = 4096;
  const int nt = (int)std::max(1L, std::min<long>({16L, (long)std::thread::hardware_concurrency(),
                                                    (n + per - 1) / per}));
  std::vector<std::string> out(nt);
  auto work = [&](int t) {
    const long a = n * t / nt, b = n * (t + 1) / nt;
    std::string &o = out[t];
    o.reserve((size_t)(b - a) * (ncol * 24 + 1));
    char buf[40];
    for (long i = a; i < b; i++) {
      for (int k = 0; k < ncol; k++) {
        const auto r = std::to_chars(buf, buf + sizeof buf, rows[i * ncol + k],
                                     std::chars_format::general, 15);
        const size_t len = (size_t)(r.ptr - buf);
        if (len < 24) o.append(24 - len, ' ');
        o.append(buf, len);
      }
      o.push_back('\n');
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
  for (auto &o : out) std::fwrite(o.data(), 1, o.size(), f);
  std::fflush(f);
}

int index44(const int &i, const int &j) {
  if (i > 3 || j > 3 || i < 0 || j < 0) {
    std::cout << "index44: i j " << i << " " << j << std::endl;
    std::exit(1);
  }
  return j < i ? (i * (i + 1)) / 2 + j : (j * (j + 1)) / 2 + i;
}

// ------------------------------------------------------------------ EoS
double EoS::s(double e, double nb, double nq, double ns) {
  double T, mub, muq, mus, p_;
  eos(e, nb, nq, ns, T, mub, muq, mus, p_);
  if (T > 0.0) return (e + p_ - mub * nb - muq * nq - mus * ns) / T;
  return 0.;
}

double EoSIdeal::t(double e) {
  const double cnst = (16 + 0.5 * 21.0 * 2.5) * (C_PI * C_PI) / 30.0 /
                      (0.197326968 * 0.197326968 * 0.197326968);
  return e > 0. ? 1.0 * std::pow(e / cnst, 0.25) : 0.;
}

void EoSIdeal::eos(double e, double, double, double, double &T, double &mub, double &muq,
                   double &mus, double &p_) {
  p_ = e / 3.;
  T = t(e);
  mub = muq = mus = 0.;
}

vh_eos_t EoSIdeal::device() const {
  vh_eos_t d{};
  d.kind = VH_EOS_IDEAL;
  return d;
}

EoSs::EoSs(std::string fname, int ncols) {
  std::ifstream finput(fname.c_str(), std::ios::in);
  if (!finput) {
    std::cerr << "can't open input file \"" << fname.c_str() << "\"" << std::endl;
    std::exit(1);
  }
  double e, p, T, mu;
  while (true) {
    mu = 0.;
    if (!(finput >> e >> p >> T)) break;
    if (ncols != 3 && !(finput >> mu)) break;
    if (p < 0.) p = 0.;
    if (!te_.empty() && e <= te_.back())
      std::cout << "WARNING non-monotonic EoS table " << te_.size() << "  " << te_.back() << "  "
                << e << std::endl;
    te_.push_back(e);
    tp_.push_back(p);
    tT_.push_back(T);
    tmu_.push_back(mu);
  }
  build();
}

EoSs::EoSs(const vh_eos_t &d) {
  if (d.kind == VH_EOS_TABLE && d.n > 0 && d.e && d.p && d.T) {
    te_.assign(d.e, d.e + d.n);
    tp_.assign(d.p, d.p + d.n);
    tT_.assign(d.T, d.T + d.n);
    tmu_.assign(d.n, 0.);
    if (d.mu) tmu_.assign(d.mu, d.mu + d.n);
  }
  build();
}

void EoSs::build() {
  n_ = (int)te_.size();
  vh_eos_t d = device();
  tab_.assign((size_t)13 * (n_ > 0 ? n_ : 1), 0.);
  if (vh_eos_table(&d, tab_.data()) != VH_OK) {
    std::cerr << "EoSs: the table needs at least 3 rows with increasing e" << std::endl;
    std::exit(1);
  }
}

double EoSs::eval(int f, double e) const {
  const double *x = tab_.data();
  if (!(e >= x[0] && e <= x[n_ - 1])) {  // gsl_spline_eval's domain error
    std::cerr << "gsl: interpolation error: e = " << e << " outside the EoS table [" << x[0]
              << ", " << x[n_ - 1] << "]" << std::endl;
    std::exit(1);
  }
  int lo = 0, hi = n_ - 1;
  while (hi > lo + 1) {
    const int mid = (hi + lo) / 2;
    if (x[mid] > e)
      hi = mid;
    else
      lo = mid;
  }
  const double *y = x + (size_t)(1 + 4 * f) * n_, *b = y + n_, *c = y + 2 * n_, *dd = y + 3 * n_;
  const double delx = e - x[lo];
  return y[lo] + delx * (b[lo] + delx * (c[lo] + delx * dd[lo]));
}

double EoSs::p(double e) { return eval(0, e); }
double EoSs::dpe(double e) { return (eval(0, e * 1.1) - eval(0, e)) / (0.1 * e); }
double EoSs::t(double e) { return eval(1, e); }
double EoSs::mu(double e) { return eval(2, e); }

void EoSs::eos(double e, double, double, double, double &T, double &mub, double &muq,
               double &mus, double &pr) {
  pr = p(e);
  T = t(e);
  mub = muq = mus = 0.;
}

vh_eos_t EoSs::device() const {
  vh_eos_t d{};
  d.kind = VH_EOS_TABLE;
  d.n = (int)te_.size();
  d.e = te_.data();
  d.p = tp_.data();
  d.T = tT_.data();
  d.mu = tmu_.data();
  return d;
}

// -------------------------------------------------- EoSChiral, EoSHadron
namespace {
// EoS::eos of a charge-dependent EoS at one point, on the host
void eval_host(const vh_eos_t &d, double e, double nb, double nq, double ns, double &T,
               double &mub, double &muq, double &mus, double &p) {
  const double in[4] = {e, nb, nq, ns};
  double out[6];
  vh_or_die(vh_eos_eval(&d, in, 1, out, 0), "vh_eos_eval");
  p = out[0];
  T = out[1];
  mub = out[2];
  muq = out[3];
  mus = out[4];
}
}  // namespace

// EoSaux::EoSaux (src/eoChiral.cpp:38-84): nn blocks of ne rows
void EoSChiral::Tab::read(const char *fname, int ne, int nn) {
  std::ifstream fin(fname);
  if (!fin.good()) {
    std::cout << "I/O error with " << fname << std::endl;
    std::exit(1);
  }
  const size_t M = (size_t)ne * nn;
  p.assign(M, 0.);
  T.assign(M, 0.);
  mub.assign(M, 0.);
  mus.assign(M, 0.);
  std::vector<double> e(ne), n(nn);
  double s, a;
  for (int in = 0; in < nn; in++)
    for (int ie = 0; ie < ne; ie++) {
      const size_t k = (size_t)ie * nn + in;
      fin >> T[k] >> mub[k] >> e[ie] >> p[k] >> n[in] >> s >> mus[k] >> a;
      T[k] /= 1000.0;
      mub[k] *= 3.0 / 1000.0;
      mus[k] /= 1000.0;
      p[k] *= 0.146;
    }
  d.ne = ne;
  d.nn = nn;
  d.emin = e[0] * 0.146;
  d.emax = e[ne - 1] * 0.146;
  d.nmin = n[0] * 0.15;
  d.nmax = n[nn - 1] * 0.15;
  d.p = p.data();
  d.T = T.data();
  d.mub = mub.data();
  d.mus = mus.data();
  std::cout << "EoSaux: table " << fname << " read, [emin,emax,nmin,nmax] = " << d.emin << "  "
            << d.emax << "  " << d.nmin << "  " << d.nmax << std::endl;
}

void EoSChiral::Tab::copy(const vh_eos_chiral_tab_t &t) {
  const size_t M = (size_t)t.ne * t.nn;
  p.assign(t.p, t.p + M);
  T.assign(t.T, t.T + M);
  mub.assign(t.mub, t.mub + M);
  mus.assign(t.mus, t.mus + M);
  d = t;
  d.p = p.data();
  d.T = T.data();
  d.mub = mub.data();
  d.mus = mus.data();
}

EoSChiral::EoSChiral() {  // src/eoChiral.cpp:156-159
  big_.read("eos/chiraleos.dat", 2001, 401);
  small_.read("eos/chiralsmall.dat", 201, 201);
}

EoSChiral::EoSChiral(const vh_eos_chiral_tab_t &small, const vh_eos_chiral_tab_t &big) {
  small_.copy(small);
  big_.copy(big);
}

vh_eos_t EoSChiral::device() const {
  vh_eos_t d{};
  d.kind = VH_EOS_CHIRAL;
  d.small = &small_.d;
  d.big = &big_.d;
  return d;
}

void EoSChiral::eos(double e, double nb, double nq, double ns, double &T, double &mub,
                    double &muq, double &mus, double &p_) {
  eval_host(device(), e, nb, nq, ns, T, mub, muq, mus, p_);
}

double EoSChiral::p(double e, double nb, double nq, double ns) {
  double T, mub, muq, mus, p_;
  eval_host(device(), e, nb, nq, ns, T, mub, muq, mus, p_);
  return p_;
}

// EoSHadron::EoSHadron (src/eoHadron.cpp:13-65)
EoSHadron::EoSHadron(const char *filename) {
  std::ifstream fin(filename);
  if (!fin.good()) {
    std::cout << "I/O error with " << filename << std::endl;
    std::exit(1);
  }
  fin >> d_.ne >> d_.nnb >> d_.nnq >> d_.e0 >> d_.n0 >> d_.logemin >> d_.logemax >> d_.lognmax;
  const size_t M = (size_t)d_.ne * d_.nnb * d_.nnq;
  for (auto *v : {&hp_, &hT_, &hmub_, &hmuq_, &hmus_}) v->assign(M, 0.);
  hstatus_.assign(M, 0);
  std::vector<double> e(d_.ne), nb(d_.nnb), nq(d_.nnq);
  for (int ie = 0; ie < d_.ne; ie++)
    for (int inb = 0; inb < d_.nnb; inb++)
      for (int inq = 0; inq < d_.nnq; inq++) {
        const size_t k = inq + (size_t)d_.nnq * (inb + (size_t)d_.nnb * ie);
        fin >> e[ie] >> nb[inb] >> nq[inq] >> hp_[k] >> hT_[k] >> hmub_[k] >> hmuq_[k] >> hmus_[k] >>
            hstatus_[k];
      }
  const double emin = e[0], emax = e[d_.ne - 1], nbmax = nb[d_.nnb - 1];
  if (std::fabs((d_.e0 * std::exp(d_.logemin) - emin) / emin) > 1e-5 ||
      std::fabs((d_.e0 * std::exp(d_.logemax) - emax) / emax) > 1e-5 ||
      std::fabs(d_.n0 * std::exp(d_.lognmax) - nbmax) > 1e-4) {
    std::cout << "EoSHadron: wrong eps or nb range: " << emin << " " << emax << " " << nb[0]
              << " " << nbmax << std::endl;
    std::exit(1);
  }
  std::cout << "EoHadron: table " << filename << " read, [emin,emax,nmin,nmax] = " << emin
            << "  " << emax << "  " << nb[0] << "  " << nbmax << "  " << nq[0] << "  "
            << nq[d_.nnq - 1] << std::endl;
  point();
}

EoSHadron::EoSHadron(const vh_eos_hadron_tab_t &t) {
  d_ = t;
  const size_t M = (size_t)t.ne * t.nnb * t.nnq;
  hp_.assign(t.p, t.p + M);
  hT_.assign(t.T, t.T + M);
  hmub_.assign(t.mub, t.mub + M);
  hmuq_.assign(t.muq, t.muq + M);
  hmus_.assign(t.mus, t.mus + M);
  hstatus_.assign(M, 0);
  if (t.status) hstatus_.assign(t.status, t.status + M);
  point();
}

void EoSHadron::point() {
  d_.p = hp_.data();
  d_.T = hT_.data();
  d_.mub = hmub_.data();
  d_.muq = hmuq_.data();
  d_.mus = hmus_.data();
  d_.status = hstatus_.data();
}

vh_eos_t EoSHadron::device() const {
  vh_eos_t d{};
  d.kind = VH_EOS_HADRON;
  d.hadron = &d_;
  return d;
}

void EoSHadron::eos(double e, double nb, double nq, double ns, double &T, double &mub,
                    double &muq, double &mus, double &p) {
  eval_host(device(), e, nb, nq, ns, T, mub, muq, mus, p);
}

double EoSHadron::p(double e, double nb, double nq, double ns) {
  double T, mub, muq, mus, pr;
  eval_host(device(), e, nb, nq, ns, T, mub, muq, mus, pr);
  return pr;
}

// ----------------------------------------------------- con2prim (host)
// Bisection/Newton root of v = M / (Q_t + p(Q_t - M v, n)), from v = 0.5,
// stopping at |dv| < 1e-5 (src/rmn.cpp:28-139).
void transformPV(EoS *eos, double Q[7], double &e, double &p, double &nb, double &nq,
                 double &ns, double &vx, double &vy, double &vz) {
  const double dpe = 1. / 3., corrf = 0.9999;
  double M = std::sqrt(Q[X_] * Q[X_] + Q[Y_] * Q[Y_] + Q[Z_] * Q[Z_]);
  if (Q[T_] <= 0.) {
    e = p = 0.;
    vx = vy = vz = 0.;
    nb = nq = ns = 0.;
    return;
  }
  if (M == 0.) {
    e = Q[T_];
    vx = vy = vz = 0.;
    nb = Q[NB_];
    nq = Q[NQ_];
    ns = Q[NS_];
    p = eos->p(e, nb, nq, ns);
    return;
  }
  if (M > Q[T_]) {
    Q[X_] *= corrf * Q[T_] / M;
    Q[Y_] *= corrf * Q[T_] / M;
    Q[Z_] *= corrf * Q[T_] / M;
    M = Q[T_] * corrf;
  }
  double lo = 0., hi = 1., v = 0.5 * (lo + hi);
  auto refresh = [&](double &f, double &df) {
    e = Q[T_] - M * v;
    if (e < 0.) e = 0.;
    const double g = std::sqrt(1 - v * v);
    nb = Q[NB_] * g;
    nq = Q[NQ_] * g;
    ns = Q[NS_] * g;
    p = eos->p(e, nb, nq, ns);
    f = (Q[T_] + p) * v - M;
    df = (Q[T_] + p) - M * v * dpe;
  };
  double f, df;
  refresh(f, df);
  double step_prev = hi - lo, step = step_prev;
  for (int it = 0; it < 100; it++) {
    const bool bisect = (f + df * (hi - v)) * (f + df * (lo - v)) >= 0. ||
                        std::fabs(2. * f) > std::fabs(step_prev * df);
    step_prev = step;
    if (bisect) {
      step = 0.5 * (hi - lo);
      v = lo + step;
    } else {
      step = f / df;
      v -= step;
    }
    if (std::fabs(step) < 0.00001) break;
    refresh(f, df);
    if (f > 0.)
      hi = v;
    else
      lo = v;
  }
  vx = v * Q[X_] / M;
  vy = v * Q[Y_] / M;
  vz = v * Q[Z_] / M;
  e = Q[T_] - M * v;
  p = eos->p(e, nb, nq, ns);
  const double g = std::sqrt(1 - vx * vx - vy * vy - vz * vz);
  nb = Q[NB_] * g;
  nq = Q[NQ_] * g;
  ns = Q[NS_] * g;
  if (e < 0. || std::sqrt(vx * vx + vy * vy + vz * vz) > 1.) {
    std::cout << Q[T_] << "  " << Q[X_] << "  " << Q[Y_] << "  " << Q[Z_] << "  " << Q[NB_]
              << std::endl;
    std::cout << "transformRF::Error\n";
  }
}

void transformCV(double e, double p, double nb, double nq, double ns, double vx, double vy,
                 double vz, double Q[]) {
  const double g2 = 1. / (1 - vx * vx - vy * vy - vz * vz);
  Q[T_] = (e + p * (vx * vx + vy * vy + vz * vz)) * g2;
  Q[X_] = (e + p) * vx * g2;
  Q[Y_] = (e + p) * vy * g2;
  Q[Z_] = (e + p) * vz * g2;
  Q[NB_] = nb * std::sqrt(g2);
  Q[NQ_] = nq * std::sqrt(g2);
  Q[NS_] = ns * std::sqrt(g2);
}

// ------------------------------------------------------- TransportCoeff
double TransportCoeff::zetaS(double e, double T, double s, double P) {
  switch (zetaSparam) {
    case 0:
      return zetaS0 * (1. / 3. - eos->cs2(e)) / (std::exp((0.16 - T) / 0.001) + 1.);
    case 1: {
      const double Tp = 0.180;
      if (T < 0.180)
        return 0.03 + (0.08 * std::exp(((T / Tp) - 1.) / (0.0025))) +
               (0.22 * std::exp(((T / Tp) - 1) / (0.022)));
      if (T < 0.200) return 27.55 * (T / Tp) - 13.45 - (13.77 * (T / Tp) * (T / Tp));
      return 0.001 + (0.9 * std::exp(-((T / Tp) - 1.) / (0.0025))) +
             (0.25 * std::exp(-((T / Tp) - 1.) / (0.13)));
    }
    case 2: {
      const double Tpk = 0.165, Tw = 0.010, Bn = 0.24, Bw = 1.5;
      if (T > Tpk) return Bn * ((Bw * Bw) / ((((T / Tpk) - 1.) * ((T / Tpk) - 1.)) + (Bw * Bw)));
      return Bn * (std::exp(-((T - Tpk) / Tw) * ((T - Tpk) / Tw)));
    }
    case 3: {
      const double Tpk2 = 0.160, Bn2 = 0.13, B1 = 0.01, B2 = 0.12;
      if (T < Tpk2) return Bn2 * std::exp(-((T - Tpk2) * (T - Tpk2) / (B1 * B1)));
      return Bn2 * std::exp(-((T - Tpk2) * (T - Tpk2) / (B2 * B2)));
    }
    case 4: {
      if (std::fabs(s) < 1e-10 || T < 1e-1) return 0;
      const double sig = e < zetaSPeakEpsilon ? zetaSSigmaMinus : zetaSSigmaPlus;
      const double d = zetaSScaleBeta * (std::pow(e, 0.25) - std::pow(zetaSPeakEpsilon, 0.25));
      return ((e + P) / (s * T)) * zetaS0 * std::exp((d * d) / 2.0 * (sig * sig));
    }
    case 5: {
      if (T <= 0.) return 0.;
      const double w = T < zetaSPeakT ? zetaSWidthMinus : zetaSWidthPlus;
      const double d = (T - zetaSPeakT) / w;
      return zetaS0 * std::exp(-0.5 * d * d);
    }
    default:
      return 0.;
  }
}

double TransportCoeff::etaS(double e, double rho, double T, double muB, double s, double P) {
  switch (etaSparam) {
    case 0:
      return etaS0;
    case 1:
      return etaSMin + ((T > T0) ? ah * (T - T0) : al * (T - T0));
    case 2:
      return std::max(0.0, etaSMin + ((e > etaSEpsilonMin)
                                          ? ((ah * (e - etaSEpsilonMin) + aRho * rho))
                                          : al * (e - etaSEpsilonMin) + aRho * rho));
    case 3: {
      if (std::fabs(s) < 1e-10 || T < 1e-1) return 0;
      const double Tc = T0 + etaSShiftMuB * (muB / 0.938);
      return ((e + P) / (s * T)) *
             std::max(0.0, (etaSMin + ((T > Tc) ? ah * (T - Tc) : al * (T - Tc))) *
                               (1 + etaSScaleMuB * (muB / 0.938)));
    }
    default:
      return 0.;
  }
}

void TransportCoeff::getEta(double e, double rho, double T, double muB, double s, double P,
                            double &_etaS, double &_zetaS) {
  _etaS = etaS(e, rho, T, muB, s, P);
  _zetaS = zetaS(e, T, s, P);
}

void TransportCoeff::getTau(double e, double rho, double T, double muB, double s, double P,
                            double &_taupi, double &_tauPi) {
  if (T > 1e-1 && std::fabs(e + P) > 1.e-10) {
    _taupi = (5. / 5.068) * s * etaS(e, rho, T, muB, s, P) / (e + P);
    _tauPi = (6. / 5.068) * zetaS(e, T, s, P) * s / (e + P);
  } else {
    _taupi = _tauPi = 0.;
  }
}

void TransportCoeff::getOther(double e, double nb, double nq, double ns, double &deltapipi,
                              double &taupipi, double &lambdapiPi, double &phi7) {
  deltapipi = 4. / 3.;
  taupipi = 10. / 7.;
  lambdapiPi = 6. / 5.;
  phi7 = 9. / 70. / eos->p(e, nb, nq, ns);
  if (std::isinf(phi7)) phi7 = 0.0;
}

void TransportCoeff::getOtherBulk(double e, double, double, double, double &delPiPi,
                                  double &lamPipi) {
  delPiPi = 2. / 3.;
  lamPipi = 8. / 5. * (1. / 3. - eos->cs2(e));
}

vh_trcoeff_t TransportCoeff::device() const {
  vh_trcoeff_t d;
  d.etaS0 = etaS0;
  d.zetaS0 = zetaS0;
  d.ah = ah;
  d.al = al;
  d.aRho = aRho;
  d.T0 = T0;
  d.etaSMin = etaSMin;
  d.etaSEpsilonMin = etaSEpsilonMin;
  d.etaSShiftMuB = etaSShiftMuB;
  d.etaSScaleMuB = etaSScaleMuB;
  d.zetaSPeakEpsilon = zetaSPeakEpsilon;
  d.zetaSScaleBeta = zetaSScaleBeta;
  d.zetaSSigmaMinus = zetaSSigmaMinus;
  d.zetaSSigmaPlus = zetaSSigmaPlus;
  d.etaSparam = etaSparam;
  d.zetaSparam = zetaSparam;
  d.zetaSPeakT = zetaSPeakT;
  d.zetaSWidthMinus = zetaSWidthMinus;
  d.zetaSWidthPlus = zetaSWidthPlus;
  return d;
}

// ----------------------------------------------------------------- Cell
#define CELL_RD(vec, k) (f->hostRead(), f->vec[(size_t)(k) * f->N + n])

void Cell::setQ(double *_Q) {
  f->hostWrite();
  const bool z = _Q[T_] < 0.;
  for (int i = 0; i < 7; i++) f->hQ[(size_t)i * f->N + n] = z ? 0. : _Q[i];
}
void Cell::setQh(double *_Qh) {
  f->hostWrite();
  const bool z = _Qh[T_] < 0.;
  for (int i = 0; i < 7; i++) f->hQh[(size_t)i * f->N + n] = z ? 0. : _Qh[i];
}
void Cell::getQ(double *_Q) {
  for (int i = 0; i < 7; i++) _Q[i] = CELL_RD(hQ, i);
}
void Cell::getQh(double *_Qh) {
  for (int i = 0; i < 7; i++) _Qh[i] = CELL_RD(hQh, i);
}
void Cell::getQprev(double *_Qp) {
  for (int i = 0; i < 7; i++) _Qp[i] = CELL_RD(hQprev, i);
}
void Cell::saveQprev(void) {
  f->hostWrite();
  for (int i = 0; i < 7; i++) f->hQprev[(size_t)i * f->N + n] = f->hQ[(size_t)i * f->N + n];
}
double Cell::getpi(const int &i, const int &j) { return CELL_RD(hpi, index44(i, j)); }
double Cell::getpiH(const int &i, const int &j) { return CELL_RD(hpiH, index44(i, j)); }
double Cell::getPi(void) { return CELL_RD(hPi, 0); }
double Cell::getPiH(void) { return CELL_RD(hPiH, 0); }
void Cell::setpi(const int &i, const int &j, const double &val) {
  f->hostWrite();
  f->hpi[(size_t)index44(i, j) * f->N + n] = val;
}
void Cell::setpiH(const int &i, const int &j, const double &val) {
  f->hostWrite();
  f->hpiH[(size_t)index44(i, j) * f->N + n] = val;
}
void Cell::setPi(const double &val) {
  f->hostWrite();
  f->hPi[n] = val;
}
void Cell::setPiH(const double &val) {
  f->hostWrite();
  f->hPiH[n] = val;
}
void Cell::setAllM(double value) {
  f->hostWrite();
  for (int d = 0; d < 3; d++) f->hm[(size_t)d * f->N + n] = value;
}
double Cell::getM(int dir) { return CELL_RD(hm, dir - 1); }
double Cell::getMaxM(void) {
  return std::max(getM(1), std::max(getM(2), getM(3)));
}
void Cell::importVars(Cell *c) {  // src/cll.cpp:54-68
  double q[7];
  c->getQ(q);
  for (int i = 0; i < 7; i++) {
    f->hostWrite();
    f->hQ[(size_t)i * f->N + n] = q[i];
  }
  c->getQh(q);
  for (int i = 0; i < 7; i++) f->hQh[(size_t)i * f->N + n] = q[i];
  c->getQprev(q);
  for (int i = 0; i < 7; i++) f->hQprev[(size_t)i * f->N + n] = q[i];
  setPi(c->getPi());
  setPiH(c->getPiH());
  for (int i = 0; i < 4; i++)
    for (int j = 0; j <= i; j++) {
      setpi(i, j, c->getpi(i, j));
      setpiH(i, j, c->getpiH(i, j));
    }
  for (int d = 1; d <= 3; d++) f->hm[(size_t)(d - 1) * f->N + n] = c->getM(d);
  setViscCorrCutFlag(c->getViscCorrCutFlag());
}
void Cell::getPrimVar(EoS *eos, double tau, double &_e, double &_p, double &_nb, double &_nq,
                      double &_ns, double &_vx, double &_vy, double &_vz) {
  double q[7];
  getQ(q);
  for (int i = 0; i < 7; i++) q[i] = q[i] / tau;
  transformPV(eos, q, _e, _p, _nb, _nq, _ns, _vx, _vy, _vz);
}
void Cell::getPrimVarHCenter(EoS *eos, double tau, double &_e, double &_p, double &_nb,
                             double &_nq, double &_ns, double &_vx, double &_vy, double &_vz) {
  double q[7];
  getQh(q);
  for (int i = 0; i < 7; i++) q[i] = q[i] / tau;
  transformPV(eos, q, _e, _p, _nb, _nq, _ns, _vx, _vy, _vz);
}
void Cell::getPrimVarPrev(EoS *eos, double tau, double &_e, double &_p, double &_nb,
                          double &_nq, double &_ns, double &_vx, double &_vy, double &_vz) {
  double q[7];
  getQprev(q);
  for (int i = 0; i < 7; i++) q[i] = q[i] / tau;
  transformPV(eos, q, _e, _p, _nb, _nq, _ns, _vx, _vy, _vz);
}
void Cell::setPrimVar(EoS *eos, double tau, double _e, double _nb, double _nq, double _ns,
                      double _vx, double _vy, double _vz) {  // src/cll.cpp:222-241
  const double g2 = 1. / (1 - _vx * _vx - _vy * _vy - _vz * _vz);
  const double p = eos->p(_e, _nb, _nq, _ns);
  double q[7];
  q[T_] = tau * (_e + p * (_vx * _vx + _vy * _vy + _vz * _vz)) * g2;
  q[X_] = tau * (_e + p) * _vx * g2;
  q[Y_] = tau * (_e + p) * _vy * g2;
  q[Z_] = tau * (_e + p) * _vz * g2;
  q[NB_] = tau * _nb * std::sqrt(g2);
  q[NQ_] = tau * _nq * std::sqrt(g2);
  q[NS_] = tau * _ns * std::sqrt(g2);
  if (std::isinf(q[NB_]) || std::isnan(q[NB_])) {
    std::cout << "init error!\n";
    std::cout << "e = " << _e << " p = " << p << " vx = " << _vx << " vy = " << _vy
              << " vz = " << _vz << " gamma2 = " << g2 << std::endl;
    return;
  }
  f->hostWrite();
  for (int i = 0; i < 7; i++) f->hQ[(size_t)i * f->N + n] = q[i];
}
void Cell::setViscCorrCutFlag(double value) {
  f->hostWrite();
  f->hvcc[n] = value;
}
double Cell::getViscCorrCutFlag(void) { return CELL_RD(hvcc, 0); }
void Cell::enableVorticity() { f->enableVorticity(); }
double Cell::getDbeta(int i, int j) {
  if (!f->vorticityOn) throw std::runtime_error("Error: dbeta is nullptr. Cannot get values.");
  if (f->db_stale) {
    f->hdb.resize(16 * f->N);
    vh_or_die(vh_get_dbeta(f->ctx, f->hdb.data()), "vh_get_dbeta");
    f->db_stale = false;
  }
  return f->hdb[(4 * i + j) * f->N + n];
}
void Cell::Dump(double tau) {
  double q[7], qh[7];
  getQ(q);
  getQh(qh);
  std::cout << "---------cell values dump-------\n";
  std::cout << std::setw(5) << ix << std::setw(5) << iy << std::setw(5) << iz << std::endl;
  for (int i = 0; i < 7; i++) std::cout << std::setw(14) << q[i] / tau;
  std::cout << std::endl;
  for (int i = 0; i < 7; i++) std::cout << std::setw(14) << qh[i] / tau;
  std::cout << "\n--------------------------------\n";
}
#undef CELL_RD

// ---------------------------------------------------------------- Fluid
static void vh_or_die(int rc, const char *what) {
  if (rc != VH_OK) {
    std::cerr << "vhlle_gpu: " << what << " failed (code " << rc << ")" << std::endl;
    std::exit(1);
  }
}

Fluid::Fluid(EoS *_eos, EoS *_eosH, TransportCoeff *_trcoeff, int _nx, int _ny, int _nz,
             double _minx, double _maxx, double _miny, double _maxy, double _minz,
             double _maxz, double _dt, double eCrit)
    : eos(_eos), eosH(_eosH), trcoeff(_trcoeff), nx(_nx), ny(_ny), nz(_nz),
      minx(_minx), maxx(_maxx), miny(_miny), maxy(_maxy), minz(_minz), maxz(_maxz),
      dt(_dt), ecrit(eCrit) {
  dx = (maxx - minx) / (nx - 1);
  dy = (maxy - miny) / (ny - 1);
  dz = (maxz - minz) / (nz - 1);
  N = (long)nx * ny * nz;
  vh_grid_t g{nx, ny, nz, minx, maxx, miny, maxy, minz, maxz};
  vh_eos_t e = eos->device();
  vh_trcoeff_t t = trcoeff->device();
  vh_or_die(vh_create(&g, &e, &t, VH_PREC_FP64, &ctx), "vh_create");
  setSurfaceEos();
  // Cell ctor defaults (src/cll.cpp:33-52)
  hQ.assign(7 * N, 0.);
  hQh.assign(7 * N, 0.);
  hQprev.assign(7 * N, 0.);
  hpi.assign(10 * N, 0.);
  hpiH.assign(10 * N, 0.);
  hPi.assign(N, 0.);
  hPiH.assign(N, 0.);
  hm.assign(3 * N, 0.);
  hvcc.assign(N, 1.);
  host_dirty = true;
}

// eosH, the EoS of outputSurface / outputCorona (SWAP_EOS, src/fld.cpp:39)
void Fluid::setSurfaceEos() {
  if (!eosH || eosH == eos) return;
  const vh_eos_t d = eosH->device();
  vh_or_die(vh_set_surface_eos(ctx, &d), "vh_set_surface_eos");
}

Fluid::~Fluid() {
  closeOutput();
  if (ctx) vh_destroy(ctx);
}

Cell *Fluid::getCell(int ix, int iy, int iz) {
  ix = ix > 0 ? ix : 0;
  ix = ix < nx ? ix : nx - 1;
  iy = iy > 0 ? iy : 0;
  iy = iy < ny ? iy : ny - 1;
  iz = iz > 0 ? iz : 0;
  iz = iz < nz ? iz : nz - 1;
  if (cells.empty()) {
    cells.resize(N);
    for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          Cell &c = cells[x + (long)nx * (y + (long)ny * z)];
          c.f = this;
          c.n = x + (long)nx * (y + (long)ny * z);
          c.ix = x;
          c.iy = y;
          c.iz = z;
        }
  }
  return &cells[ix + (long)nx * iy + (long)nx * ny * iz];
}

void Fluid::hostRead() {
  if (dev_newer) syncFromDevice();
}

void Fluid::hostWrite() {
  if (dev_newer) syncFromDevice();
  host_dirty = true;
}

namespace {
// Fluid::checkOutputDirectory (src/fld.cpp:198-209): an old file is removed
void remove_old(const std::string &path) {
  std::FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return;
  std::fclose(f);
  const std::string name = path.substr(path.find_last_of('/') + 1);
  std::cout << "\033[33mWarning! A '" << name << "' is present in your output directory.\n"
            << "         It will be deleted automatically.\n\033[0m";
  std::remove(path.c_str());
}
}  // namespace

void Fluid::initOutput(const char *dir, double tau0, bool hsOnly) {
  closeOutput();
  const std::string d(dir);
  std::cout << "mkdir returns: " << (mkdir(dir, 0755) == 0) << std::endl;
  remove_old(d + "/freezeout.dat");
  remove_old(d + "/freezeout.dat.unfinished");
  const std::string fn = d + "/freezeout.dat.unfinished";
  ffreeze = std::fopen(fn.c_str(), "w");
  if (!ffreeze) std::cout << "cannot open " << fn << std::endl;
  if (vorticityOn) {  // src/fld.cpp:131-136
    const std::string fb = d + "/beta.dat";
    fbeta = std::fopen(fb.c_str(), "w");
    if (!fbeta) std::cout << "cannot open " << fb << std::endl;
  }
  if (hsOnly) return;
  static const char *names[9] = {"outx.dat",   "outy.dat",      "outz.dat",
                                 "outdiag.dat", "out2D.dat",     "out.aniz.dat",
                                 "outx.visc.dat", "outy.visc.dat", "diag.visc.dat"};
  for (int k = 0; k < 9; k++) fgp[k] = std::fopen((d + "/" + names[k]).c_str(), "w");
  outputGnuplot(tau0);
  if (fgp[5]) std::fprintf(fgp[5], "#  tau  <<v_T>>  e_p  e'_p  (to compare with SongHeinz)\n");
}

void Fluid::renameOutput(const char *dir) {
  const std::string d(dir);
  if (ffreeze) std::fflush(ffreeze);
  if (std::rename((d + "/freezeout.dat.unfinished").c_str(), (d + "/freezeout.dat").c_str()) != 0)
    std::perror("Error renaming freezeout.dat.unfinished!");
}

void Fluid::outputGnuplot(double tau) {
  if (!fgp[0]) return;
  syncToDevice();
  // the cells of the four lines: x, y, diagonal (getCell clamps iy), eta
  std::vector<long> idx;
  auto at = [&](int ix, int iy, int iz) {
    ix = std::min(std::max(ix, 0), nx - 1);
    iy = std::min(std::max(iy, 0), ny - 1);
    iz = std::min(std::max(iz, 0), nz - 1);
    return ix + (long)nx * (iy + (long)ny * iz);
  };
  for (int ix = 0; ix < nx; ix++) idx.push_back(at(ix, ny / 2, nz / 2));
  for (int iy = 0; iy < ny; iy++) idx.push_back(at(nx / 2, iy, nz / 2));
  for (int ix = 0; ix < nx; ix++) idx.push_back(at(ix, ix, nz / 2));
  for (int iz = 0; iz < nz; iz++) idx.push_back(at(nx / 2, ny / 2, iz));
  std::vector<double> v(idx.size() * VH_CELL_COLS);
  vh_or_die(vh_cells(ctx, tau, idx.data(), (long)idx.size(), v.data()), "vh_cells");
  // getCMFvariables (src/fld.cpp:50-60) and EoS::eos of one listed cell
  struct Row {
    double e, nb, t, mub, vx, vy, Y;
    const double *pi;
    double Pi, vcc;
  };
  auto row = [&](size_t k) {
    const double *c = &v[k * VH_CELL_COLS];
    const long cell = idx[k];
    const double eta = getZ((int)(cell / ((long)nx * ny)));
    Row r;
    r.e = c[0];
    r.nb = c[2];
    double vx = c[5], vy = c[6], vz = c[7];
    r.Y = eta + 1. / 2. * std::log((1. + vz) / (1. - vz));
    r.vx = vx * std::cosh(r.Y - eta) / std::cosh(r.Y);
    r.vy = vy * std::cosh(r.Y - eta) / std::cosh(r.Y);
    double muq, mus, p;
    eos->eos(r.e, r.nb, c[3], c[4], r.t, r.mub, muq, mus, p);
    r.pi = c + 8;
    r.Pi = c[18];
    r.vcc = c[19];
    return r;
  };
  auto tail = [&](std::FILE *f, const Row &r) {  // pi^{00..33}, Pi, viscCorrCut
    static const int order[10][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 1},
                                     {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}};
    for (auto &ij : order) std::fprintf(f, "%14g", r.pi[index44(ij[0], ij[1])]);
    std::fprintf(f, "%14g%14g\n", r.Pi, r.vcc);
  };
  size_t k = 0;
  for (int ix = 0; ix < nx; ix++, k++) {
    const Row r = row(k);
    std::fprintf(fgp[0], "%14g%14g%14g%14g%14g%14g%14g%14g", tau, getX(ix), r.vx, r.vy, r.e, r.nb,
                 r.t, r.mub);
    tail(fgp[0], r);
  }
  std::fprintf(fgp[0], "\n");
  for (int iy = 0; iy < ny; iy++, k++) {
    const Row r = row(k);
    std::fprintf(fgp[1], "%14g%14g%14g%14g%14g%14g%14g%14g", tau, getY(iy), r.vy, r.vx, r.e, r.nb,
                 r.t, r.mub);
    tail(fgp[1], r);
  }
  std::fprintf(fgp[1], "\n");
  for (int ix = 0; ix < nx; ix++, k++) {
    const Row r = row(k);
    std::fprintf(fgp[3], "%14g%14g%14g%14g%14g%14g%14g%14g\n", tau, std::sqrt(2.) * getY(ix), r.vx,
                 r.vy, r.e, r.nb, r.t, r.mub);
    tail(fgp[3], r);
  }
  std::fprintf(fgp[3], "\n");
  for (int iz = 0; iz < nz; iz++, k++) {
    const Row r = row(k);
    std::fprintf(fgp[2], "%14g%14g%14g%14g%14g%14g%14g%14g", tau, getZ(iz), r.Y, r.vx, r.e, r.nb,
                 r.t, r.mub);
    tail(fgp[2], r);
  }
  std::fprintf(fgp[2], "\n");
  for (int q = 0; q < 4; q++) std::fflush(fgp[q]);
}

void Fluid::gridTotals(double tau, bool central, double out[16]) {
  vh_or_die(vh_fluid_stats(ctx, tau, ecrit, central ? 1 : 0, out), "vh_fluid_stats");
  if (std::isinf(out[0])) {  // src/fld.cpp:562-566
    std::cout << "EEinf" << std::endl;
    std::exit(1);
  }
  const double dV = dx * dy * dz;
  out[0] *= dV;  // E
  out[1] *= dV;  // Efull
  out[2] *= dV;  // S
  out[4] *= dV;  // Nb1
  out[5] *= dV;  // Nb2
  if (fgp[5])
    std::fprintf(fgp[5], "%12g%14g%14g%14g\n", tau, out[6] / out[7], out[8] / out[9],
                 out[10] / out[11]);
  if (fgp[5]) std::fflush(fgp[5]);
}

void Fluid::adoptOutput(Fluid *from) {
  closeOutput();
  ffreeze = from->ffreeze;
  fbeta = from->fbeta;
  from->ffreeze = from->fbeta = nullptr;
  for (int k = 0; k < 9; k++) {
    fgp[k] = from->fgp[k];
    from->fgp[k] = nullptr;
  }
  num_corona_cells = from->num_corona_cells;
  EtotSurf = from->EtotSurf;
  nbSurf = from->nbSurf;
  vEff = from->vEff;
}

void Fluid::closeOutput() {
  if (ffreeze) std::fclose(ffreeze);
  if (fbeta) std::fclose(fbeta);
  ffreeze = fbeta = nullptr;
  for (auto &f : fgp) {
    if (f) std::fclose(f);
    f = nullptr;
  }
}

void Fluid::enableVorticity() {
  if (!vorticityOn) vh_or_die(vh_set_vorticity(ctx, 1), "vh_set_vorticity");
  vorticityOn = true;
  db_stale = true;
}

void Fluid::printDbetaHeader() {
  if (num_corona_cells == -1)
    throw std::runtime_error(
        "Error: num_corona_cells not set, cannot print header of beta.dat");
  if (!fbeta) return;
  std::fprintf(fbeta,
               "# The derivatives of \u03b2 are given in Cartesian coordinates and include "
               "a factor of 1/2, such that \u2202\u2098\u03b2\u2099=1/2 * "
               "\u2202\u2098(u\u2099/T)\n");
  std::fprintf(fbeta, "# Number of corona cells: %d\n", num_corona_cells);
  std::fprintf(fbeta,
               "#  \u03c4  x  y  \u03b7  d\u03a3[0]  d\u03a3[1]  d\u03a3[2]  "
               "d\u03a3[3]  u[0]  u[1]  u[2]  u[3]  T  \u03bcB  \u03bcQ  \u03bcS  ");
  static const char *sub[4] = {"\u2080", "\u2081", "\u2082", "\u2083"};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) std::fprintf(fbeta, "\u2202%s\u03b2%s  ", sub[i], sub[j]);
  std::fprintf(fbeta, "\u03f5\n");
  std::fflush(fbeta);
}

int Fluid::outputCorona(double tau, bool extendFO) {
  syncToDevice();
  if (fgp[4]) std::fprintf(fgp[4], "\n");  // out2D.dat (src/fld.cpp:888)
  long n = 0;
  vh_fo_stats_t st;
  const int ncol = extendFO ? 29 : 27;
  vh_or_die(vh_corona(ctx, tau, ecrit, extendFO, nullptr, 0, &n, &st), "vh_corona");
  if (ffreeze && n > 0) {
    std::vector<double> rows((size_t)n * ncol);
    vh_or_die(vh_corona(ctx, tau, ecrit, extendFO, rows.data(), n, &n, &st), "vh_corona");
    write_records(ffreeze, rows.data(), n, ncol);
  }
  // the reference adds the corona to EtotSurf and vEff; its n_b sum goes to
  // a function-local static (src/fld.cpp:877), not to the Fluid's nbSurf
  EtotSurf += st.e_surf;
  vEff += st.v_eff;
  num_corona_cells = (int)n;
  double T[16];
  gridTotals(tau, false, T);
  std::fflush(stdout);
  std::cout << std::setw(10) << "tau" << std::setw(13) << "E" << std::setw(13) << "Efull"
            << std::setw(13) << "Nb" << std::setw(13) << "Sfull" << std::setw(13) << "EtotSurf"
            << std::setw(13) << "elements" << std::setw(10) << "susp." << std::setw(13) << "%cut"
            << std::endl;
  std::cout << std::setw(10) << tau << std::setw(13) << T[0] << std::setw(13) << T[1]
            << std::setw(13) << st.nb_surf << std::setw(13) << T[2] << std::setw(13) << EtotSurf
            << std::endl;
  std::cout << "corona elements : " << n << std::endl;
  return (int)n;
}

int Fluid::outputSurface(double tau, bool extendFO) {
  syncToDevice();
  if (fgp[4]) std::fprintf(fgp[4], "\n");  // out2D.dat (src/fld.cpp:543)
  long n = 0;
  vh_fo_stats_t st;
  const int ncol = extendFO ? 29 : 27;
  const bool wb = vorticityOn && fbeta, want = ffreeze || wb;
  for (;;) {
    if (want) {
      fo_rows.resize((size_t)fo_cap * ncol);
      if (wb) fo_beta.resize((size_t)fo_cap * VH_BETA_COLS);
    }
    const bool has = want && fo_cap > 0;
    vh_or_die(vh_freezeout_beta(ctx, tau, dt, ecrit, extendFO, VH_FO_QPREV,
                                has ? fo_rows.data() : nullptr,
                                has && wb ? fo_beta.data() : nullptr, has ? fo_cap : 0, &n, &st),
              "vh_freezeout");
    if (!want || n <= fo_cap) break;
    fo_cap = n + n / 4 + 1024;  // room for the surface to grow, then fetch again
  }
  const std::vector<double> &rows = fo_rows, &beta = fo_beta;
  if (want && n > 0) {
    write_records(ffreeze, rows.data(), n, ncol);
    if (wb) write_records(fbeta, beta.data(), n, VH_BETA_COLS);  // src/fld.cpp:663-834
  }
  if (st.err) std::cout << "#### Error (surface): hypercube construction failed ####\n";
  EtotSurf += st.e_surf;
  nbSurf += st.nb_surf;
  vEff += st.v_eff;
  double T[16];
  gridTotals(tau, true, T);
  // src/fld.cpp:852-857
  std::printf("%10g%13g%13g%13g%13g%13g%10ld%10ld%13g\n", tau, T[0], T[1], nbSurf, T[2],
              EtotSurf, n, st.nsusp, (double)((float)T[15] / (float)T[14]));
  std::fflush(stdout);
  return (int)n;
}

void Fluid::configure(int precision, int c2pMode, const float *model) {
  vh_grid_t g{nx, ny, nz, minx, maxx, miny, maxy, minz, maxz};
  vh_eos_t e = eos->device();
  vh_trcoeff_t t = trcoeff->device();
  if (ctx) vh_destroy(ctx);
  ctx = nullptr;
  vh_or_die(vh_create(&g, &e, &t, precision, &ctx), "vh_create");
  setSurfaceEos();
  if (model) vh_or_die(vh_set_c2p_model(ctx, model), "vh_set_c2p_model");
  if (vorticityOn) vh_or_die(vh_set_vorticity(ctx, 1), "vh_set_vorticity");
  if (c2pMode != VH_C2P_REFERENCE) vh_or_die(vh_set_c2p_mode(ctx, c2pMode), "vh_set_c2p_mode");
  host_dirty = true;  // the host mirror is the state to upload
  dev_newer = false;
}

void Fluid::syncToDevice() {
  if (!host_dirty) return;
  vh_state_t s{hQ.data(), hQh.data(), hQprev.data(), hpi.data(), hpiH.data(),
               hPi.data(), hPiH.data(), hm.data(), hvcc.data()};
  vh_or_die(vh_upload(ctx, &s), "vh_upload");
  host_dirty = false;
}

void Fluid::syncFromDevice() {
  if (!dev_newer) return;
  vh_state_t s{hQ.data(), hQh.data(), hQprev.data(), hpi.data(), hpiH.data(),
               hPi.data(), hPiH.data(), hm.data(), hvcc.data()};
  vh_or_die(vh_download(ctx, &s), "vh_download");
  dev_newer = false;
}

// The ghost fills and updateM are stages of performStep on the device; as
// stand-alone calls they act on the host mirror with the reference's
// sequential semantics (src/fld.cpp:211-423).
void Fluid::correctImagCells(void) {
  double Q[7], Qh[7];
  auto edge = [&](Cell *s, Cell *a, Cell *b) {
    s->getQ(Q);
    a->setQ(Q);
    b->setQ(Q);
    s->getQh(Qh);
    a->setQh(Qh);
    b->setQh(Qh);
  };
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) {
      edge(getCell(ix, iy, 2), getCell(ix, iy, 1), getCell(ix, iy, 0));
      edge(getCell(ix, iy, nz - 3), getCell(ix, iy, nz - 2), getCell(ix, iy, nz - 1));
    }
  for (int ix = 0; ix < nx; ix++)
    for (int iz = 0; iz < nz; iz++) {
      edge(getCell(ix, 2, iz), getCell(ix, 1, iz), getCell(ix, 0, iz));
      edge(getCell(ix, ny - 3, iz), getCell(ix, ny - 2, iz), getCell(ix, ny - 1, iz));
    }
  for (int iy = 0; iy < ny; iy++)
    for (int iz = 0; iz < nz; iz++) {
      edge(getCell(2, iy, iz), getCell(1, iy, iz), getCell(0, iy, iz));
      edge(getCell(nx - 3, iy, iz), getCell(nx - 2, iy, iz), getCell(nx - 1, iy, iz));
    }
}

void Fluid::correctImagCellsFull(void) {
  double Q[7], pi[4][4], Pi;
  auto edge = [&](Cell *s, Cell *a, Cell *b) {  // a = ghost 1, b = ghost 0
    s->getQ(Q);
    a->setQ(Q);
    b->setQ(Q);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) pi[i][j] = s->getpi(i, j);
    Pi = s->getPi();
    for (int i = 0; i < 4; i++)
      for (int j = 0; j <= i; j++) {
        b->setpi(i, j, pi[i][j]);
        a->setpi(i, j, pi[i][j]);
      }
    b->setPi(Pi);
    a->setPi(Pi);
  };
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) {
      edge(getCell(ix, iy, 2), getCell(ix, iy, 1), getCell(ix, iy, 0));
      edge(getCell(ix, iy, nz - 3), getCell(ix, iy, nz - 2), getCell(ix, iy, nz - 1));
    }
  for (int ix = 0; ix < nx; ix++)
    for (int iz = 0; iz < nz; iz++) {
      edge(getCell(ix, 2, iz), getCell(ix, 1, iz), getCell(ix, 0, iz));
      edge(getCell(ix, ny - 3, iz), getCell(ix, ny - 2, iz), getCell(ix, ny - 1, iz));
    }
  for (int iy = 0; iy < ny; iy++)
    for (int iz = 0; iz < nz; iz++) {
      edge(getCell(2, iy, iz), getCell(1, iy, iz), getCell(0, iy, iz));
      edge(getCell(nx - 3, iy, iz), getCell(nx - 2, iy, iz), getCell(nx - 1, iy, iz));
    }
}

void Fluid::updateM(double tau, double _dt) {
  std::vector<double> dm(3 * N, 0.);
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++) {
        Cell *c = getCell(ix, iy, iz);
        double *d = &dm[3 * c->n];
        if (c->getMaxM() < 1.) {
          if (getCell(ix + 1, iy, iz)->getM(X_) >= 1. || getCell(ix - 1, iy, iz)->getM(X_) >= 1.)
            d[0] = _dt / dx;
          if (getCell(ix, iy + 1, iz)->getM(Y_) >= 1. || getCell(ix, iy - 1, iz)->getM(Y_) >= 1.)
            d[1] = _dt / dy;
          if (getCell(ix, iy, iz + 1)->getM(Z_) >= 1. || getCell(ix, iy, iz - 1)->getM(Z_) >= 1.)
            d[2] = _dt / dz / tau;
          if (d[0] == 0. && d[1] == 0.) {
            if (getCell(ix + 1, iy + 1, iz)->getMaxM() >= 1. ||
                getCell(ix + 1, iy - 1, iz)->getMaxM() >= 1. ||
                getCell(ix - 1, iy + 1, iz)->getMaxM() >= 1. ||
                getCell(ix - 1, iy - 1, iz)->getMaxM() >= 1.) {
              d[0] = 0.707 * _dt / dx;
              d[1] = 0.707 * _dt / dy;
            }
          }
        }
      }
  hostWrite();
  for (long c = 0; c < N; c++)
    for (int k = 0; k < 3; k++) {
      double &mk = hm[(size_t)k * N + c];
      mk += dm[3 * c + k];
      if (mk > 0.9)
        for (int j = 0; j < 3; j++) hm[(size_t)j * N + c] = 1.;
    }
}

// ---------------------------------------------------------------- Hydro
Hydro::Hydro(Fluid *_f, EoS *_eos, TransportCoeff *_trcoeff, double _t0, double _dt)
    : f(_f), eos(_eos), trcoeff(_trcoeff), dt(_dt), tau(_t0) {}

Hydro::~Hydro() {}

void Hydro::setDtau(double deltaTau) {  // src/hdo.cpp:67-75
  dt = deltaTau;
  if (dt > f->getDx() / 2. || dt > f->getDy() / 2.) {
    std::cout << "too big delta_tau " << dt << "  " << f->getDx() << "  " << f->getDy()
              << "  " << tau * f->getDz() << std::endl;
    std::exit(1);
  }
}

// maps device status bits to the reference's diagnostics and exits
void Hydro::check(int rc, int status) {
  vh_or_die(rc, "vh_step");
  if (status & VH_ST_BIG_E) {
    std::cout << "e>1e10; debug info below:\n";
    std::exit(0);
  }
  if (status & VH_ST_NAN_FLUX) {
    std::cout << "---- error in hlle_flux: f_nb undefined!\n";
    std::exit(1);
  }
  if (status & VH_ST_NAN_SOURCE) {
    std::cout << "---- error in source_step: k_nb undefined!\n";
    std::exit(1);
  }
  if (status & VH_ST_NAN_PI) {
    std::cout << "hydro:NSquant: inf/nan\n";
    std::exit(1);
  }
}

void Hydro::performStep(void) {
  f->syncToDevice();
  int status = 0;
  const int rc = vh_step(f->context(), &tau, dt, &status);
  f->markDeviceNewer();
  check(rc, status);
}

void Hydro::performSteps(int nsteps) {
  f->syncToDevice();
  int status = 0;
  const int rc = vh_run(f->context(), &tau, dt, nsteps, &status);
  f->markDeviceNewer();
  check(rc, status);
}

// ------------------------------------------------------------- ICGubser
// Gubser flow at proper time tau with q = 1 and t = 1 as in the reference
// (src/icGubser.cpp:17-62), energy density times scale.
void ICGubser::setIC(Fluid *f, EoS *eos, double tau) {
  double avv_num = 0., avv_den = 0., Etotal = 0.0;
  for (int ix = 0; ix < f->getNX(); ix++)
    for (int iy = 0; iy < f->getNY(); iy++)
      for (int iz = 0; iz < f->getNZ(); iz++) {
        Cell *c = f->getCell(ix, iy, iz);
        const double x = f->getX(ix), y = f->getY(iy), eta = f->getZ(iz);
        const double _t = 1.0, q = 1.0, r = std::sqrt(x * x + y * y);
        const double k = 2. * q * q * _t * r / (1. + q * q * _t * _t + q * q * r * r);
        const double qq = q * q * (_t * _t - r * r);
        double e = 4. * q * q / (1. + 2. * q * q * (_t * _t + r * r) + qq * qq) / _t;
        if (e < 0.) e = 0.;
        e = scale_ * std::pow(e, 4. / 3.);
        const double vx = x / (r + 1e-50) * k, vy = y / (r + 1e-50) * k, vz = 0.0;
        avv_num += std::sqrt(vx * vx + vy * vy) * e;
        avv_den += e;
        c->setPrimVar(eos, tau, e, 0., 0., 0., vx, vy, vz);
        const double p = eos->p(e, 0., 0., 0.);
        const double g2 = 1.0 / (1.0 - vx * vx - vy * vy - vz * vz);
        Etotal += ((e + p) * g2 * (std::cosh(eta) + vz * std::sinh(eta)) - p * std::cosh(eta));
        c->saveQprev();
        if (e > 0.) c->setAllM(1.);
      }
  std::cout << "average initial flow = " << avv_num / avv_den << std::endl;
  std::cout << "total energy = " << Etotal * f->getDx() * f->getDy() * f->getDz() * tau
            << std::endl;
}

// ------------------------------------------------------------ VtkOutput
namespace {
// kind of each quantity VtkOutput writes (src/vtk.h:33-46)
const std::map<std::string, char> &vtk_quantities() {
  static const std::map<std::string, char> q = {
      {"eps", 's'}, {"mub", 's'}, {"muq", 's'}, {"mus", 's'}, {"nb", 's'}, {"nq", 's'},
      {"ns", 's'},  {"p", 's'},   {"Pi", 's'},  {"pi", 't'},  {"T", 's'},  {"v", 'v'}};
  return q;
}
}  // namespace

bool VtkOutput::is_quantity_implemented(const std::string &quantity) {
  return vtk_quantities().count(quantity) != 0;
}

std::string VtkOutput::make_filename(const std::string &descr, int counter) {
  char suffix[24];
  std::snprintf(suffix, sizeof(suffix), "_taustep%05i.vtk", counter);
  return path_ + "/" + descr + suffix;
}

void VtkOutput::write(const Hydro &h, const std::string &quantities) {
  Fluid *f = h.getFluid();
  const int nx = f->getNX(), ny = f->getNY(), nz = f->getNZ();
  const long N = f->ncell();
  std::vector<double> prim;  // e, p, nb, nq, ns, vx, vy, vz (8 x N), fetched once
  auto fields = [&]() -> const std::vector<double> & {
    if (prim.empty()) {
      f->syncToDevice();
      prim.resize(8 * N);
      vh_or_die(vh_prim_fields(f->context(), h.getTau(), prim.data()), "vh_prim_fields");
    }
    return prim;
  };
  std::stringstream ss(quantities);
  std::string q;
  while (std::getline(ss, q, ',')) {
    if (!is_quantity_implemented(q)) {
      std::cout << "Given quantity '" << q << "' is not an "
                << "implemented VTK quantity. This entry will be skipped." << std::endl;
      continue;
    }
    const std::string fn = make_filename(q, vtk_output_counter_);
    std::FILE *out = std::fopen(fn.c_str(), "w");
    if (!out) {
      std::cout << "cannot open " << fn << std::endl;
      continue;
    }
    std::fprintf(out,
                 "# vtk DataFile Version 2.0\n%s\nASCII\nDATASET STRUCTURED_POINTS\n"
                 "DIMENSIONS %d %d %d\nSPACING %g %g %g\nORIGIN %g %g %g\nPOINT_DATA %d\n",
                 q.c_str(), nx, ny, nz, f->getDx(), f->getDy(), f->getDz(), xmin_, ymin_,
                 etamin_, nx * ny * nz);
    const char kind = vtk_quantities().at(q);
    if (kind == 's') {
      std::fprintf(out, "SCALARS %s double 1\nLOOKUP_TABLE default\n", q.c_str());
      const std::vector<double> &pr = fields();
      for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++) {
          for (int ix = 0; ix < nx; ix++) {
            const long c = ix + (long)nx * (iy + (long)ny * iz);
            const double e = pr[c], p = pr[N + c], nb = pr[2 * N + c], nq = pr[3 * N + c],
                         ns = pr[4 * N + c];
            double v = 0.;
            if (q == "eps") v = e;
            else if (q == "nb") v = nb;
            else if (q == "nq") v = nq;
            else if (q == "ns") v = ns;
            else if (q == "p") v = p;
            else if (q == "Pi") v = f->getCell(ix, iy, iz)->getPi();
            else {  // EoS::eos of the cell's state (src/vtk.cpp:64-76)
              double T, mub, muq, mus, pp;
              eos_->eos(e, nb, nq, ns, T, mub, muq, mus, pp);
              v = q == "mub" ? mub : q == "muq" ? muq : q == "mus" ? mus : T;
            }
            std::fprintf(out, "%.3f ", v);
          }
          std::fprintf(out, "\n");
        }
    } else if (kind == 'v') {
      std::fprintf(out, "VECTORS %s double\n", q.c_str());
      const std::vector<double> &pr = fields();
      for (long c = 0; c < N; c++)
        std::fprintf(out, "%.3f %.3f %.3f\n", pr[5 * N + c], pr[6 * N + c], pr[7 * N + c]);
    } else {
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
          std::fprintf(out, "SCALARS %s%d%d double 1\nLOOKUP_TABLE default\n", q.c_str(), i, j);
          for (int iz = 0; iz < nz; iz++)
            for (int iy = 0; iy < ny; iy++) {
              for (int ix = 0; ix < nx; ix++)
                std::fprintf(out, "%.3f ", f->getCell(ix, iy, iz)->getpi(i, j));
              std::fprintf(out, "\n");
            }
        }
    }
    std::fclose(out);
  }
  vtk_output_counter_++;
}
