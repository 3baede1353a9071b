// Device-side physics of the vHLLE hydro step for gfx950 (FP64 faithful
// path).  Every function reproduces the reference's floating-point operation
// order so that, compiled with -ffp-contract=off, results agree with the
// reference's x86-64 build to rounding (sqrt and division are correctly
// rounded on both; pow/exp come from the ROCm device library).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/vhlle_gpu.h"

namespace vh {

// Q component slots (src/inc.h:4-10)
enum { T_ = 0, X_ = 1, Y_ = 2, Z_ = 3, NB_ = 4, NQ_ = 5, NS_ = 6 };

// packed symmetric 4x4 index (src/cll.cpp:22-31), compile-time foldable
__host__ __device__ constexpr inline int sym(int i, int j) {
  return j < i ? (i * (i + 1)) / 2 + j : (j * (j + 1)) / 2 + i;
}

// slope limiter (src/cll.cpp:12-19)
__device__ inline double minmod(double a, double b) {
  if (a * b <= 0.) return 0.;
  return fabs(a) > fabs(b) ? b : a;
}

// sign (src/hdo.cpp:34-41), as an int offset
__device__ inline int isgn(double x) { return x > 0 ? 1 : (x < 0. ? -1 : 0); }

// std::min / std::max semantics (argument order matters for NaN and -0)
__device__ inline double smin(double a, double b) { return (b < a) ? b : a; }
__device__ inline double smax(double a, double b) { return (a < b) ? b : a; }

// --------------------------------------------------------------------- EoS
// Natural cubic spline of a tabulated function (gsl_interp_cspline, the TABLE
// branch of EoSs, src/eos.cpp:59-69): on [x_i, x_i+1],
//   y = y_i + dx (b_i + dx (c_i + dx d_i)),  dx = v - x_i,
// with c_i = y''(x_i)/2 from the natural tridiagonal system and b_i, d_i from
// c_i, c_i+1 (computed once on the host).  Interval search as gsl_interp_bsearch
// (x_i <= v < x_i+1; v = x_n-1 falls in the last interval).
// The search starts from a hint: hint[k] is the interval of the k-th of K
// bins uniform in log2(e) over [2^l0, x_n-1]; a walk of a step or two then
// lands on the interval the bisection would return (x_i <= v < x_i+1, or the
// end interval outside the table), so the result does not depend on the hint.
struct __align__(16) SplineBin {
  double xs;  // the node inside the bin's range (NaN: none)
  int ia;     // the interval at the start of the range
  float xsf;  // xs rounded to FP32 (NaN: none)
};
struct SplineTab {
  const double *x;  // n nodes, then per function: y, b, c, d (n each)
  int n;
  const int *hint;  // K bins (null: bisection)
  int K;
  float l0, inv;
  __device__ inline int interval(double v) const {
    if (hint) {
      const float kf = (__log2f((float)v) - l0) * inv;
      const int k = kf > 0.f ? (kf < (float)(K - 1) ? (int)kf : K - 1) : 0;
      int i = hint[k];
      while (i < n - 2 && x[i + 1] <= v) i++;
      while (i > 0 && x[i] > v) i--;
      return i;
    }
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
      const int mid = (hi + lo) / 2;
      if (x[mid] > v)
        hi = mid;
      else
        lo = mid;
    }
    return lo;
  }
  // Direct lookup (bins non-null when the host could build it): fK bins
  // uniform in log2(e) over [fxlo, x_n-1] (fl0 = log2 fxlo, finv = fK / log2
  // range), each holding at most one node, with a relative margin of 1e-4 on
  // both sides that covers the error of the FP32 bin index.  bins[0] serves
  // v < fxlo, bins[k + 1] bin k: the first interval ia of the bin's range and
  // the node xs that ends it (NaN when none does), so that the interval of v
  // is ia + (v >= xs), with no search.  The coefficients of interval i follow
  // interleaved, one cache line apart at most:
  //   c64[f n + i]: three double2 (x_i, y_i) (b_i, c_i) (d_i, 0), functions
  //       0-2, the spline's own (eval and eval_d return exactly the values of
  //       the interval search);
  // The FP32 lookup of the mixed-precision kernels reads one record per bin
  // instead (a lookup inside a Newton iteration waits on one load, not two):
  //   c32[f (fK + 1) + k]: three float4 (xs, xa, ya, ba) (ca, da, xb, yb)
  //       (bb, cb, db, 0), functions 0-1, the cubics of the intervals ia and
  //       ia + 1, each re-centred on its node rounded to FP32 and rounded to
  //       FP32 (null when a coefficient does not fit a float: evalf is then
  //       not used).
  const SplineBin *bins;
  const double2 *c64;
  const float4 *c32;
  int fK;
  float fl0, finv, fxlo;
  // the table's range as FP32 bounds, set for every table: a float v has
  // v >= fxmin and v <= fxmax exactly when x_0 <= v <= x_n-1
  float fxmin, fxmax;
  __device__ inline int bin(float vf) const {
    int k = 0;
    if (vf >= fxlo) {
      const float kf = (__log2f(vf) - fl0) * finv;
      k = 1 + (kf > 0.f ? (kf < (float)(fK - 1) ? (int)kf : fK - 1) : 0);
    }
    return k;
  }
  // the node and coefficients of the interval of v, from the direct lookup
  __device__ inline void cubic(int f, double v, double &x0, double &y0, double &b0, double &c0,
                               double &d0) const {
    const SplineBin B = bins[bin((float)v)];
    const int i = B.ia + (v >= B.xs ? 1 : 0);
    const double2 *C = c64 + ((long)f * n + i) * 3;
    const double2 r0 = C[0], r1 = C[1], r2 = C[2];
    x0 = r0.x;
    y0 = r0.y;
    b0 = r1.x;
    c0 = r1.y;
    d0 = r2.x;
  }
  // function f (0: p, 1: T, 2: mu) at v; *out_of_range set outside [x_0, x_n-1]
  __device__ inline double eval(int f, double v, bool &out_of_range) const {
    out_of_range = !(v >= x[0] && v <= x[n - 1]);
    if (bins) {
      double x0, y0, b0, c0, d0;
      cubic(f, v, x0, y0, b0, c0, d0);
      const double delx = v - x0;
      return y0 + delx * (b0 + delx * (c0 + delx * d0));
    }
    const int i = interval(v);
    const double *y = x + (long)(1 + 4 * f) * n;
    const double *b = y + n, *c = y + 2 * n, *d = y + 3 * n;
    const double delx = v - x[i];
    return y[i] + delx * (b[i] + delx * (c[i] + delx * d[i]));
  }
  // f and its derivative in FP32 (c32 non-null)
  __device__ inline float evalf(int f, float v, float &dy, bool &out_of_range) const {
    out_of_range = !(v >= fxmin && v <= fxmax);
    const float4 *R = c32 + ((long)f * (fK + 1) + bin(v)) * 3;
    const float4 a = R[0], b = R[1], c = R[2];
    const bool hi = v >= a.x;
    const float x0 = hi ? b.z : a.y, y0 = hi ? b.w : a.z, b0 = hi ? c.x : a.w,
                c0 = hi ? c.y : b.x, d0 = hi ? c.z : b.y;
    const float dx = v - x0;
    dy = b0 + dx * (2.f * c0 + 3.f * d0 * dx);
    return y0 + dx * (b0 + dx * (c0 + dx * d0));
  }
  // f and its derivative at v (one interval search)
  __device__ inline double eval_d(int f, double v, double &dy, bool &out_of_range) const {
    out_of_range = !(v >= x[0] && v <= x[n - 1]);
    if (bins) {
      double x0, y0, b0, c0, d0;
      cubic(f, v, x0, y0, b0, c0, d0);
      const double delx = v - x0;
      dy = b0 + delx * (2. * c0 + 3. * d0 * delx);
      return y0 + delx * (b0 + delx * (c0 + delx * d0));
    }
    const int i = interval(v);
    const double *y = x + (long)(1 + 4 * f) * n;
    const double *b = y + n, *c = y + 2 * n, *d = y + 3 * n;
    const double delx = v - x[i];
    dy = b[i] + delx * (2. * c[i] + 3. * d[i] * delx);
    return y[i] + delx * (b[i] + delx * (c[i] + delx * d[i]));
  }
};

// e / 3, correctly rounded, without a division.  y = RN(1/3) = (1 - 2^-54)/3,
// so q = RN(e y) lies within one ulp of e/3; then the remainder e - 3q is
// exact (one FMA) and RN(q + (e - 3q) y) is the correctly rounded quotient
// (Markstein's correction theorem; no quotient is a rounding midpoint, since
// 3 times an odd 54-bit integer does not fit 53 bits).  Outside
// [2^-1000, 2^1000] in magnitude (subnormal quotients, 0, inf, NaN) it is the
// division itself.  Bit-identical to e / 3. on every input.
__device__ inline double div3(double e) {
  const double y = 0.33333333333333331483;  // RN(1/3)
  const double a = fabs(e);
  if (!(a >= 0x1p-1000 && a <= 0x1p1000)) return e / 3.;
  const double q = e * y;
  const double r = fma(-q, 3., e);
  return fma(r, y, q);
}

// One table of EoSChiral (the auxiliary class EoSaux, src/eoChiral.cpp:38-154):
// ne x nn nodes uniform in e and n_b; the fields p, T, mu_b, mu_s are each
// ne*nn long, at [ie*nn + in].  de and dn are (emax-emin)/(ne-1) and
// (nmax-nmin)/(nn-1), formed on the host as the reference forms them per call.
// Written for host and device, so the C++ facade evaluates the same lookups.
struct ChiralTab {
  const double *p, *T, *mub, *mus;
  int ne, nn;
  double emin, nmin, de, dn;
  // EoSaux::get (ALL) and EoSaux::p (src/eoChiral.cpp:99-154), in the
  // reference's operation order: the bilinear sums run over (je, jn) from 0.0.
  template <bool ALL>
  __host__ __device__ __forceinline__ void get(double e, double nb, double &p_, double &T_, double &mub_,
                                      double &mus_) const {
    p_ = T_ = mub_ = mus_ = 0.;
    if (e < 0.) return;
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
    double sp = 0.0, sT = 0.0, sb = 0.0, ss = 0.0;
#pragma unroll
    for (int je = 0; je < 2; je++)
#pragma unroll
      for (int jn = 0; jn < 2; jn++) {
        const long k = (long)(ie + je) * nn + (in + jn);
        const double w = we[je] * wn[jn];
        sp += w * p[k];
        if (ALL) {
          sT += w * T[k];
          sb += w * mub[k];
          ss += w * mus[k];
        }
      }
    if (sp < 0.0) sp = 0.0;
    p_ = sp;
    T_ = sT;
    mub_ = sb;
    mus_ = ss;
  }
};

// The table of EoSHadron (src/eoHadron.cpp:13-208): ne x nnb x nnq nodes
// uniform in x_e = log(e/e0) and in the log-mapped x_nb, x_nq; the fields p,
// T, mu_b, mu_q, mu_s are each ne*nnb*nnq long, in index3 order inq +
// nnq*(inb + nnb*ie); status holds the node flags (null: none set).  dxe, dxnb, dxnq, nb_abs_min, nq_abs_min
// and e_min are formed on the host as the reference's constructor does.
struct HadronTab {
  const double *p, *T, *mub, *muq, *mus;
  const int *status;
  int ne, nnb, nnq;
  double e0, n0, logemin, lognmax, dxe, dxnb, dxnq, nb_abs_min, nq_abs_min, e_min;
  __host__ __device__ inline long index3(int ie, int inb, int inq) const {
    return inq + (long)nnq * inb + (long)nnq * nnb * ie;
  }
  // inverse of the log mapping of a charge density (src/eoHadron.cpp:98-111)
  __host__ __device__ static inline double xmap(double n, double n0, double lognmax, double dx,
                                                double n_abs_min) {
    if (n != 0.0) {
      const double z = log(fabs(n) / n0);
      if (z >= -lognmax + 2.0 * dx) return n / fabs(n) * 0.5 * (z + lognmax);
      return n / n_abs_min * dx;
    }
    return 0.;
  }
  // EoSHadron::eos (ALL) and EoSHadron::p (src/eoHadron.cpp:78-208)
  template <bool ALL>
  __host__ __device__ __forceinline__ void get(double e, double nb, double nq, double &p_, double &T_,
                                      double &mub_, double &muq_, double &mus_) const {
    p_ = T_ = mub_ = muq_ = mus_ = 0.;
    if (e <= 0.) return;
    if (e < e_min) {
      const long k0 = index3(0, nnb / 2, nnq / 2);
      p_ = e / e_min * p[k0];
      if (ALL) T_ = e / e_min * T[k0];
      return;
    }
    const double xe = log(e / e0);
    const double xnb = xmap(nb, n0, lognmax, dxnb, nb_abs_min);
    const double xnq = xmap(nq, n0, lognmax, dxnq, nq_abs_min);
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
    if (status && status[index3(ie, inb, inq)] == 1) return;
    const double we[2] = {1. - em / dxe, em / dxe};
    const double wnb[2] = {1. - nbm / dxnb, nbm / dxnb};
    const double wnq[2] = {1. - nqm / dxnq, nqm / dxnq};
    double sp = 0.0, sT = 0.0, sb = 0.0, sq = 0.0, ss = 0.0;
#pragma unroll
    for (int je = 0; je < 2; je++)
#pragma unroll
      for (int jnb = 0; jnb < 2; jnb++)
#pragma unroll
        for (int jnq = 0; jnq < 2; jnq++) {
          const long k = index3(ie + je, inb + jnb, inq + jnq);
          const double w = we[je] * wnb[jnb] * wnq[jnq];
          sp += w * p[k];
          if (ALL) {
            sT += w * T[k];
            sb += w * mub[k];
            sq += w * muq[k];
            ss += w * mus[k];
          }
        }
    if (sp < 0.0) sp = 0.0;
    p_ = sp;
    T_ = sT;
    mub_ = sb;
    muq_ = sq;
    mus_ = ss;
  }
};

// EoSChiral::eos / EoSChiral::p (src/eoChiral.cpp:131-154): the table by
// region, else the power law; mu_q = 0.
template <bool ALL>
__host__ __device__ __forceinline__ void chiral_eos(const ChiralTab &small, const ChiralTab &big,
                                                   double e, double nb, double &p, double &T,
                                                   double &mub, double &muq, double &mus) {
  if (e < 1.46 && nb < 0.3) {
    small.get<ALL>(e, nb, p, T, mub, mus);
  } else if (e < 146. && nb < 6.) {
    big.get<ALL>(e, nb, p, T, mub, mus);
  } else {
    p = 0.2964 * e;
    T = ALL ? 0.15120476935 * pow(e, 0.25) : 0.;
    mub = mus = 0.0;
  }
  muq = 0.0;
}

// x / b for many x and one b, at a third of the cost of an IEEE division
// and with the same result: y = RN(1/b) once; then q = RN(x y) is within an
// ulp of x/b, the remainder r = x - b q is exact (one FMA), and RN(q + r y)
// is the correctly rounded quotient (Markstein's theorem: y within half an
// ulp of 1/b, q faithful; P. W. Markstein, IBM J. Res. Dev. 34 (1990) 111).
// A zero, infinite or NaN remainder returns q, which is then x/b itself
// (signed zero, infinity, NaN).  For b in [2^-100, 2^100] the result is the
// IEEE quotient except where x y or r underflows (|x| below about 2^-969),
// where it may differ by an ulp of the subnormal range.
struct Divider {
  double b, y;
  __device__ explicit Divider(double b_) : b(b_), y(1.0 / b_) {}
  __device__ double operator()(double x) const {
    const double q = x * y;
    const double r = fma(-q, b, x);
    // classes: signaling and quiet NaN, -inf, -0, +0, +inf
    return __builtin_amdgcn_class(r, 1 | 2 | 4 | 32 | 64 | 512) ? q : fma(r, y, q);
  }
};

// Device EoS (kinds VH_EOS_IDEAL, VH_EOS_TABLE, VH_EOS_CHIRAL, VH_EOS_HADRON,
// include/vhlle_gpu.h).
// Ideal massless gas: the SIMPLE branch of EoSs (src/eos.cpp:87-111), cnst
// computed on the host exactly as the reference's compile-time constant.
// Table: the TABLE branch (src/eos.cpp:27-120), p, T, mu splines of e; an
// evaluation outside the table raises VH_ST_EOS_RANGE in *st (the reference
// aborts there) and uses the end interval's cubic.
// Chiral / hadron: the bilinear and log-trilinear lookups above; these
// depend on the charge densities.
// p(e), T(e) and the FP32 forms are the charge-free evaluations of kinds 0
// and 1 (the mixed-precision kernels, which reject kinds 2 and 3); the FP64
// path calls p(e, nb, nq, ns), eos(...) and s(...), which hold for every kind.
struct DevEos {
  int kind;
  double cnst;
  SplineTab tab;
  ChiralTab small, big;
  HadronTab had;
  int *st;  // status word for VH_ST_EOS_RANGE (may be null)
  __host__ __device__ inline bool charged() const { return kind >= VH_EOS_CHIRAL; }
  __device__ inline double spline(int f, double e) const {
    bool out;
    const double v = tab.eval(f, e, out);
    if (out && st) atomicOr(st, VH_ST_EOS_RANGE);
    return v;
  }
  __device__ double p(double e) const { return kind == VH_EOS_TABLE ? spline(0, e) : div3(e); }
  __device__ double T(double e) const {
    if (kind == VH_EOS_TABLE) return spline(1, e);
    return e > 0. ? 1.0 * pow(e / cnst, 0.25) : 0.;
  }
  // The charge-aware forms, compiled per class of EoS: CH = false for kinds 0
  // and 1 (p, T of e alone, mu = 0), CH = true for kinds 2 and 3.  A kernel
  // that evaluates them is instantiated for both classes and launched with
  // CH = charged(), so that the lookups of kinds 2 and 3 cost the other
  // kinds' kernels no registers.
  // EoS::p(e, nb, nq, ns)
  template <bool CH>
  __device__ double p(double e, double nb, double nq, double ns) const {
    if (!CH) return p(e);
    if (kind == VH_EOS_CHIRAL) {
      double p_, T_, mub, muq, mus;
      chiral_eos<false>(small, big, e, nb, p_, T_, mub, muq, mus);
      return p_;
    }
    double p_, T_, mub, muq, mus;
    had.get<false>(e, nb, nq, p_, T_, mub, muq, mus);
    return p_;
  }
  // EoS::eos(e, nb, nq, ns, T, mub, muq, mus, p)
  template <bool CH>
  __device__ void eos(double e, double nb, double nq, double ns, double &T_, double &mub,
                      double &muq, double &mus, double &p_) const {
    if (!CH) {
      p_ = p(e);
      T_ = T(e);
      mub = muq = mus = 0.;
    } else if (kind == VH_EOS_CHIRAL) {
      chiral_eos<true>(small, big, e, nb, p_, T_, mub, muq, mus);
    } else {
      had.get<true>(e, nb, nq, p_, T_, mub, muq, mus);
    }
  }
  template <bool CH>
  __device__ double T(double e, double nb, double nq, double ns) const {
    if (!CH) return T(e);
    double T_, mub, muq, mus, p_;
    eos<CH>(e, nb, nq, ns, T_, mub, muq, mus, p_);
    return T_;
  }
  // EoS::s (src/eos.cpp:18-25)
  template <bool CH>
  __device__ double s(double e, double nb, double nq, double ns) const {
    double T_, mub, muq, mus, p_;
    eos<CH>(e, nb, nq, ns, T_, mub, muq, mus, p_);
    if (T_ > 0.0) return (e + p_ - mub * nb - muq * nq - mus * ns) / T_;
    return 0.;
  }
  // p and the exact dp/de (the spline's derivative; 1/3 for the ideal gas),
  // for the Newton updates of the con2prim polish.  (The reference's own
  // iteration keeps the estimate 1/3 for every EoS, src/rmn.cpp:35.)
  __device__ double p_dpde(double e, double &dp) const {
    if (kind == VH_EOS_TABLE) {
      bool out;
      const double v = tab.eval_d(0, e, dp, out);
      if (out && st) atomicOr(st, VH_ST_EOS_RANGE);
      return v;
    }
    dp = 1. / 3.;
    return div3(e);
  }
  // EoSs::cs2(e) = dpe(e) (src/eos.cpp:89-96); 1/3 for the ideal gas and for
  // the chiral and hadron EoSs, which keep the base class's cs2(e)
  // (src/eos.h:37)
  __device__ double cs2e(double e) const {
    if (kind == VH_EOS_TABLE) return (spline(0, e * 1.1) - spline(0, e)) / (0.1 * e);
    return 1. / 3.;
  }
  // FP32 forms (mixed precision: interface states of the FP32 flux, the
  // thermodynamics of the FP32 relaxation step)
  // table function f in FP32: the FP32 records when the table has them
  __device__ inline float splinef(int f, float e, float &dy) const {
    if (!tab.c32) {
      double d;
      const double v = (double)e;
      bool out;
      const double r = tab.eval_d(f, v, d, out);
      if (out && st) atomicOr(st, VH_ST_EOS_RANGE);
      dy = (float)d;
      return (float)r;
    }
    bool out;
    const float r = tab.evalf(f, e, dy, out);
    if (out && st) atomicOr(st, VH_ST_EOS_RANGE);
    return r;
  }
  __device__ float pf(float e) const {
    float d;
    return kind == VH_EOS_TABLE ? splinef(0, e, d) : e * (1.f / 3.f);
  }
  // p and dp/de in FP32 (one lookup)
  __device__ float p_dpdef(float e, float &dp) const {
    if (kind == VH_EOS_TABLE) return splinef(0, e, dp);
    dp = 1.f / 3.f;
    return e * (1.f / 3.f);
  }
  __device__ float cs2f(float e) const {
    return kind == VH_EOS_TABLE ? (float)cs2e((double)e) : 1.f / 3.f;
  }
  __device__ float Tf(float e) const {
    float d;
    if (kind == VH_EOS_TABLE) return splinef(1, e, d);
    return e > 0.f ? __builtin_amdgcn_sqrtf(__builtin_amdgcn_sqrtf(e * (float)(1. / cnst))) : 0.f;
  }
  __device__ float sf(float e, float T) const { return T > 0.f ? (e + pf(e)) / T : 0.f; }
};

// ----------------------------------------------------- con2prim (FP64)
struct Prim {
  double e, p, nb, nq, ns, vx, vy, vz;
};

// The EoS of a kernel compiled for one EoS kind EK (VH_EOS_IDEAL or
// VH_EOS_TABLE; EK < 0: the kind given at run time, for the charge-dependent
// EoSs).  The kind is fixed at kernel entry, so that an ideal-gas kernel
// carries none of the spline code inlined at every EoS call (the mixed
// corrector's row kernel shrinks from 77 to 29 KB of code).
template <int EK>
__device__ __forceinline__ DevEos eos_of_kind(const DevEos &a) {
  DevEos r = a;
  if (EK >= 0) r.kind = EK;
  return r;
}
// the kernel instance of an EoS (host side): EK of eos_of_kind
inline int eos_kind_class(const DevEos &e) {
  return e.charged() ? -1 : (e.kind == VH_EOS_TABLE ? VH_EOS_TABLE : VH_EOS_IDEAL);
}

// transformPV (src/rmn.cpp:28-139), FP64, reference operation order.  The
// input q is rescaled in place when |q_vec| > q_t (src/rmn.cpp:65-70).
// Returns the number of loop iterations executed.  Written on scalar locals
// with one exit so that nothing is spilled to scratch memory.  For an EoS
// that depends on the charges, n = N sqrt(1 - v^2) is refreshed with every
// pressure evaluation, and the final pressure uses the last iterate's n, as
// the reference does (src/rmn.cpp:74-78,99-103,122-126); for the others p(e)
// alone is evaluated (n does not enter).  CH: the class of the EoS (DevEos).
template <bool CH>
__device__ __forceinline__ int c2p_ref(const DevEos &eos, double q[7], Prim &r) {
  const double dpe = 1. / 3.;
  const double corrf = 0.9999;
  const double Qt = q[T_], Qb = q[NB_], Qq = q[NQ_], Qs = q[NS_];
  double Qx = q[X_], Qy = q[Y_], Qz = q[Z_];
  double M = sqrt(Qx * Qx + Qy * Qy + Qz * Qz);
  double e = 0., p = 0., vx = 0., vy = 0., vz = 0., nb = 0., nq = 0., ns = 0.;
  int it = 0;
  if (Qt <= 0.) {
    // vacuum: everything zero
  } else if (M == 0.) {
    e = Qt;
    nb = Qb;
    nq = Qq;
    ns = Qs;
    p = eos.p<CH>(e, nb, nq, ns);
  } else {
    if (M > Qt) {
      Qx *= corrf * Qt / M;
      Qy *= corrf * Qt / M;
      Qz *= corrf * Qt / M;
      M = Qt * corrf;
    }
    constexpr bool chg = CH;
    double vl = 0., vh = 1.;
    double v = 0.5 * (vl + vh);
    double vn = v;  // the iterate of the last pressure evaluation
    e = Qt - M * v;
    if (e < 0.) e = 0.;
    if (chg) {
      const double w = sqrt(1 - v * v);
      p = eos.p<CH>(e, Qb * w, Qq * w, Qs * w);
    } else {
      p = eos.p(e);
    }
    double f = (Qt + p) * v - M;
    double df = (Qt + p) - M * v * dpe;
    double dvold = vh - vl;
    double dv = dvold;
    for (; it < 100;) {
      const bool bis = (f + df * (vh - v)) * (f + df * (vl - v)) >= 0. ||
                       fabs(2. * f) > fabs(dvold * df);
      dvold = dv;
      if (bis) {
        dv = 0.5 * (vh - vl);
        v = vl + dv;
      } else {
        dv = f / df;
        v -= dv;
      }
      it++;
      if (fabs(dv) < 0.00001) break;
      e = Qt - M * v;
      if (e < 0.) e = 0.;
      if (chg) {
        const double w = sqrt(1 - v * v);
        p = eos.p<CH>(e, Qb * w, Qq * w, Qs * w);
        vn = v;
      } else {
        p = eos.p(e);
      }
      f = (Qt + p) * v - M;
      df = (Qt + p) - M * v * dpe;
      if (f > 0.)
        vh = v;
      else
        vl = v;
    }
    vx = v * Qx / M;
    vy = v * Qy / M;
    vz = v * Qz / M;
    e = Qt - M * v;
    if (chg) {
      const double w = sqrt(1 - vn * vn);
      p = eos.p<CH>(e, Qb * w, Qq * w, Qs * w);
    } else {
      p = eos.p(e);
    }
    const double s2 = sqrt(1 - vx * vx - vy * vy - vz * vz);
    nb = Qb * s2;
    nq = Qq * s2;
    ns = Qs * s2;
  }
  q[X_] = Qx;
  q[Y_] = Qy;
  q[Z_] = Qz;
  r.e = e;
  r.p = p;
  r.nb = nb;
  r.nq = nq;
  r.ns = ns;
  r.vx = vx;
  r.vy = vy;
  r.vz = vz;
  return it;
}

// transformCV (src/rmn.cpp:253-263)
__device__ __forceinline__ void p2c(const Prim &a, double Q[7]) {
  const double g2 = 1. / (1 - a.vx * a.vx - a.vy * a.vy - a.vz * a.vz);
  Q[T_] = (a.e + a.p * (a.vx * a.vx + a.vy * a.vy + a.vz * a.vz)) * g2;
  Q[X_] = (a.e + a.p) * a.vx * g2;
  Q[Y_] = (a.e + a.p) * a.vy * g2;
  Q[Z_] = (a.e + a.p) * a.vz * g2;
  Q[NB_] = a.nb * sqrt(g2);
  Q[NQ_] = a.nq * sqrt(g2);
  Q[NS_] = a.ns * sqrt(g2);
}

// ------------------------------------------------- transport coefficients
// TransportCoeff (src/trancoeff.h:6-45, src/trancoeff.cpp:19-100)
struct DevTrc {
  vh_trcoeff_t c;
  __device__ double zetaS(const DevEos &eos, double e, double T, double s, double P) const {
    switch (c.zetaSparam) {
      case 0:
        return c.zetaS0 * (1. / 3. - eos.cs2e(e)) / (exp((0.16 - T) / 0.001) + 1.);
      case 1: {
        const double Tp = 0.180;
        if (T < 0.180)
          return 0.03 + (0.08 * exp(((T / Tp) - 1.) / (0.0025))) +
                 (0.22 * exp(((T / Tp) - 1) / (0.022)));
        else if (T < 0.200)
          return 27.55 * (T / Tp) - 13.45 - (13.77 * (T / Tp) * (T / Tp));
        return 0.001 + (0.9 * exp(-((T / Tp) - 1.) / (0.0025))) +
               (0.25 * exp(-((T / Tp) - 1.) / (0.13)));
      }
      case 2: {
        const double Tpk = 0.165, Tw = 0.010, Bn = 0.24, Bw = 1.5;
        if (T > Tpk)
          return Bn * ((Bw * Bw) / ((((T / Tpk) - 1.) * ((T / Tpk) - 1.)) + (Bw * Bw)));
        return Bn * (exp(-((T - Tpk) / Tw) * ((T - Tpk) / Tw)));
      }
      case 3: {
        const double Tpk2 = 0.160, Bn2 = 0.13, B1 = 0.01, B2 = 0.12;
        if (T < Tpk2) return Bn2 * exp(-((T - Tpk2) * (T - Tpk2) / (B1 * B1)));
        return Bn2 * exp(-((T - Tpk2) * (T - Tpk2) / (B2 * B2)));
      }
      case 4: {
        if (fabs(s) < 1e-10 || T < 1e-1) return 0;
        const double sig = e < c.zetaSPeakEpsilon ? c.zetaSSigmaMinus : c.zetaSSigmaPlus;
        const double d = c.zetaSScaleBeta * (pow(e, 0.25) - pow(c.zetaSPeakEpsilon, 0.25));
        return ((e + P) / (s * T)) * c.zetaS0 * exp((d * d) / 2.0 * (sig * sig));
      }
      case 5: {
        if (T <= 0.) return 0.;
        const double w = T < c.zetaSPeakT ? c.zetaSWidthMinus : c.zetaSWidthPlus;
        const double d = (T - c.zetaSPeakT) / w;
        return c.zetaS0 * exp(-0.5 * d * d);
      }
      default:
        return 0.;
    }
  }
  __device__ double etaS(double e, double rho, double T, double muB, double s, double P) const {
    switch (c.etaSparam) {
      case 0:
        return c.etaS0;
      case 1:
        return c.etaSMin + ((T > c.T0) ? c.ah * (T - c.T0) : c.al * (T - c.T0));
      case 2:
        return smax(0.0, c.etaSMin + ((e > c.etaSEpsilonMin)
                                          ? ((c.ah * (e - c.etaSEpsilonMin) + c.aRho * rho))
                                          : c.al * (e - c.etaSEpsilonMin) + c.aRho * rho));
      case 3: {
        if (fabs(s) < 1e-10 || T < 1e-1) return 0;
        const double Tc = c.T0 + c.etaSShiftMuB * (muB / 0.938);
        return ((e + P) / (s * T)) *
               smax(0.0, (c.etaSMin + ((T > Tc) ? c.ah * (T - Tc) : c.al * (T - Tc))) *
                             (1 + c.etaSScaleMuB * (muB / 0.938)));
      }
      default:
        return 0.;
    }
  }
  // FP32 forms of zetaS, etaS and getTau for the mixed-precision relaxation
  // step (the same parametrisations, mu = 0)
  __device__ float zetaSf(const DevEos &eos, float e, float T, float s, float P) const {
    switch (c.zetaSparam) {
      case 0:
        return (float)c.zetaS0 * (1.f / 3.f - eos.cs2f(e)) / (expf((0.16f - T) * 1000.f) + 1.f);
      case 1: {
        const float Tp = 0.180f, r = T * (1.f / Tp);
        if (T < 0.180f)
          return 0.03f + 0.08f * expf((r - 1.f) * 400.f) + 0.22f * expf((r - 1.f) * (1.f / 0.022f));
        else if (T < 0.200f)
          return 27.55f * r - 13.45f - 13.77f * r * r;
        return 0.001f + 0.9f * expf(-(r - 1.f) * 400.f) + 0.25f * expf(-(r - 1.f) * (1.f / 0.13f));
      }
      case 2: {
        const float Tpk = 0.165f, Tw = 0.010f, Bn = 0.24f, Bw = 1.5f;
        if (T > Tpk) {
          const float d = T * (1.f / Tpk) - 1.f;
          return Bn * (Bw * Bw) / (d * d + Bw * Bw);
        }
        const float d = (T - Tpk) * (1.f / Tw);
        return Bn * expf(-d * d);
      }
      case 3: {
        const float Tpk2 = 0.160f, Bn2 = 0.13f, B1 = 0.01f, B2 = 0.12f;
        const float d = T - Tpk2;
        if (T < Tpk2) return Bn2 * expf(-d * d * (1.f / (B1 * B1)));
        return Bn2 * expf(-d * d * (1.f / (B2 * B2)));
      }
      case 4: {
        if (fabsf(s) < 1e-10f || T < 1e-1f) return 0.f;
        const float sig = e < (float)c.zetaSPeakEpsilon ? (float)c.zetaSSigmaMinus
                                                        : (float)c.zetaSSigmaPlus;
        const float q = __builtin_amdgcn_sqrtf(__builtin_amdgcn_sqrtf(e));
        const float q0 = (float)pow(c.zetaSPeakEpsilon, 0.25);
        const float d = (float)c.zetaSScaleBeta * (q - q0);
        return ((e + P) / (s * T)) * (float)c.zetaS0 * expf(d * d * 0.5f * (sig * sig));
      }
      case 5: {
        if (T <= 0.f) return 0.f;
        const float w = T < (float)c.zetaSPeakT ? (float)c.zetaSWidthMinus : (float)c.zetaSWidthPlus;
        const float d = (T - (float)c.zetaSPeakT) / w;
        return (float)c.zetaS0 * expf(-0.5f * d * d);
      }
      default:
        return 0.f;
    }
  }
  __device__ float etaSf(float e, float rho, float T, float s, float P) const {
    switch (c.etaSparam) {
      case 0:
        return (float)c.etaS0;
      case 1:
        return (float)c.etaSMin + ((T > (float)c.T0) ? (float)c.ah * (T - (float)c.T0)
                                                     : (float)c.al * (T - (float)c.T0));
      case 2: {
        const float em = (float)c.etaSEpsilonMin;
        const float v = (float)c.etaSMin + ((e > em) ? ((float)c.ah * (e - em) + (float)c.aRho * rho)
                                                     : (float)c.al * (e - em) + (float)c.aRho * rho);
        return v > 0.f ? v : 0.f;
      }
      case 3: {
        if (fabsf(s) < 1e-10f || T < 1e-1f) return 0.f;
        const float Tc = (float)c.T0;  // mu_B = 0
        const float v = (float)c.etaSMin + ((T > Tc) ? (float)c.ah * (T - Tc) : (float)c.al * (T - Tc));
        return ((e + P) / (s * T)) * (v > 0.f ? v : 0.f);
      }
      default:
        return 0.f;
    }
  }
  __device__ void getTauf(const DevEos &eos, float e, float rho, float T, float s, float P,
                          float etaS, float zetaS, float &taupi, float &tauPi) const {
    if (T > 1e-1f && fabsf(e + P) > 1.e-10f) {
      const float r = s / (e + P);
      taupi = (float)(5. / 5.068) * etaS * r;
      tauPi = (float)(6. / 5.068) * zetaS * r;
    } else {
      taupi = tauPi = 0.f;
    }
  }
  // getTau (src/trancoeff.cpp:92-100); abs() of the reference resolves to the
  // double overload in its GSL-enabled build.
  __device__ void getTau(const DevEos &eos, double e, double rho, double T, double muB,
                         double s, double P, double &taupi, double &tauPi) const {
    if (T > 1e-1 && fabs(e + P) > 1.e-10) {
      taupi = (5. / 5.068) * s * etaS(e, rho, T, muB, s, P) / (e + P);
      tauPi = (6. / 5.068) * zetaS(eos, e, T, s, P) * s / (e + P);
    } else {
      taupi = tauPi = 0.;
    }
  }
};

// --------------------------------------------------------------- geometry
// Workgroups are dispatched to the 8 XCDs of the chip round-robin (workgroup
// b runs on XCD b % 8), each XCD with its own L2.  This renumbering gives XCD
// x one contiguous range of workgroup indices, so that the stencil neighbours
// a workgroup reads were mostly fetched into the same L2 by its neighbours.
// A bijection of [0, gridDim.x).
__device__ __forceinline__ long xcd_block() {
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u, i = b >> 3;
  return (long)x * q + (x < r ? x : r) + i;
}

// Maximum over the 64 lanes of a wave (every lane must be active); the result
// is in every lane.  Butterfly of cross-lane exchanges.
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
// Signal-speed maxima are gathered in groups of kSigSlots slots, one 128-byte
// line apart: a wave adds its maximum to the slot of its index, so the
// atomics of a launch spread over 64 lines instead of serialising on one
// address; k_advance_tau folds each group.  A slot holds a non-negative
// double as its bit pattern (for those the integer order is the numeric
// order, so an integer atomic max is a floating-point max).
constexpr int kSigSlots = 64, kSigStride = 16, kSigGroup = kSigSlots * kSigStride;
// Every lane of the wave must call it.
__device__ __forceinline__ void wave_speed_max(unsigned long long *grp, double v) {
  v = wave_max(v);
  const unsigned w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if ((threadIdx.x & 63) == 0 && v > 0.)
    atomicMax(grp + (w % kSigSlots) * kSigStride, (unsigned long long)__double_as_longlong(v));
}

// Occupancy target of a kernel: VH_OCC(n) caps it at n waves per SIMD (0: the
// compiler's default).  The stencil gathers are latency bound; a lower target
// lets the scheduler keep more loads in flight per wave, which pays more than
// the extra waves.  Values per kernel measured on the 200x200x100 step; each
// can be overridden at build time (-DVH_W_<KERNEL>=n).
#define VH_OCC_0
#define VH_OCC_1 __attribute__((amdgpu_waves_per_eu(1, 1)))
#define VH_OCC_2 __attribute__((amdgpu_waves_per_eu(1, 2)))
#define VH_OCC_3 __attribute__((amdgpu_waves_per_eu(1, 3)))
#define VH_OCC_4 __attribute__((amdgpu_waves_per_eu(1, 4)))
#define VH_OCC_5 __attribute__((amdgpu_waves_per_eu(1, 5)))
#define VH_OCC_6 __attribute__((amdgpu_waves_per_eu(1, 6)))
// M<n>: at least n waves per SIMD (registers capped accordingly, spilling
// what does not fit)
#define VH_OCC_M2 __attribute__((amdgpu_waves_per_eu(2)))
#define VH_OCC_M3 __attribute__((amdgpu_waves_per_eu(3)))
#define VH_OCC_M4 __attribute__((amdgpu_waves_per_eu(4)))
#define VH_OCC_M5 __attribute__((amdgpu_waves_per_eu(5)))
#define VH_OCC_EXP(n) VH_OCC_##n
#define VH_OCC(n) VH_OCC_EXP(n)
#ifndef VH_W_FACES
#define VH_W_FACES 0
#endif
#ifndef VH_W_UPDATE
#define VH_W_UPDATE 0
#endif
#ifndef VH_W_VPRIM
#define VH_W_VPRIM 0
#endif
#ifndef VH_W_VISC1
// FP64 relaxation: 255 VGPRs + 76 AGPRs at the default (one wave per SIMD);
// two waves with 108 registers spilled to scratch run 12 % faster
#define VH_W_VISC1 M2
#endif
#ifndef VH_W_VISC2
#define VH_W_VISC2 0
#endif
#ifndef VH_W_VISC3
#define VH_W_VISC3 0
#endif
#ifndef VH_W_MX_STAGE
#define VH_W_MX_STAGE 0
#endif
#ifndef VH_W_MX_ROWS
#define VH_W_MX_ROWS 0
#endif
#ifndef VH_W_MX_BANDS
#define VH_W_MX_BANDS 0
#endif
#ifndef VH_W_MX_VPRIM
#define VH_W_MX_VPRIM 0
#endif
#ifndef VH_W_MX_VISC1
// mixed relaxation: 131 VGPRs (three waves) at the default, 127 with no
// spill at four waves, 13 % faster
#define VH_W_MX_VISC1 M4
#endif
#ifndef VH_W_MX_VISC2
#define VH_W_MX_VISC2 2
#endif
#ifndef VH_W_MX_VISC3
#define VH_W_MX_VISC3 0
#endif
#ifndef VH_W_GUESS
#define VH_W_GUESS 0
#endif
// rows per y-strip of the per-cell kernels' cell order (0: x-fastest)
#ifndef VH_STRIP_ROWS
#define VH_STRIP_ROWS 8
#endif

// Cell (ix, iy, iz) of work item u (0 <= u < nx ny nz) in y-strip order: the
// grid's rows are taken in strips of sy rows (the last strip shorter), each
// strip swept over all eta in contiguous chunks of sy*nx cells.  With
// XCD-contiguous workgroups, the eta- and y-neighbours a workgroup gathers
// were fetched into the same L2 by the workgroups just before it (in
// x-fastest order an eta-neighbour is a whole plane away).  A bijection of
// the cells, so results do not depend on it.  32-bit arithmetic: a context
// holds fewer than 2^29 cells (checked at creation).
__device__ __forceinline__ void strip_cell(int nx_, int ny_, int nz_, int sy_, long u_, int &ix,
                                           int &iy, int &iz) {
  const unsigned nx = nx_, ny = ny_, u = (unsigned)u_;
  if (sy_ <= 0 || ny_ <= sy_) {  // x-fastest
    const unsigned r = u / nx, q = r / ny;
    ix = (int)(u - r * nx);
    iy = (int)(r - q * ny);
    iz = (int)q;
    return;
  }
  const unsigned sy = sy_, nz = nz_;
  const unsigned nfull = ny / sy, plane = sy * nx, strip = plane * nz;
  unsigned s = u / strip, r = u - s * strip, pl = plane;
  if (s >= nfull) {  // the last, shorter strip
    s = nfull;
    r = u - nfull * strip;
    pl = (ny - nfull * sy) * nx;
  }
  const unsigned z = r / pl, f = r - z * pl, yy = f / nx;
  iz = (int)z;
  iy = (int)(s * sy + yy);
  ix = (int)(f - yy * nx);
}

// Grid geometry.  N: cells of this grid; base: index of its cell 0 in the
// state arrays; P: the arrays' component pitch (cells of all grids stored
// together; base = 0 and P = N for a single grid).
struct Geom {
  int nx, ny, nz;
  long N;
  double dx, dy, dz;
  long base = 0, P = 0;
  __device__ inline long idx(int ix, int iy, int iz) const {
    return base + ix + (long)nx * (iy + (long)ny * iz);
  }
  // Fluid::getCell clamp (src/fld.h:70-78)
  __device__ inline long cidx(int ix, int iy, int iz) const {
    ix = ix > 0 ? ix : 0;
    ix = ix < nx ? ix : nx - 1;
    iy = iy > 0 ? iy : 0;
    iy = iy < ny ? iy : ny - 1;
    iz = iz > 0 ? iz : 0;
    iz = iz < nz ? iz : nz - 1;
    return idx(ix, iy, iz);
  }
};

// The ghost cells (the two outer layers of every face, src/fld.cpp:211-382)
// of a grid whose extents are all >= 5, enumerated without the interior: the
// four eta layers iz = 0, 1, nz-2, nz-1 (x fastest), then the four y rows of
// the other eta slices, then the four x columns of the remaining rows.
// ghost_count of them; item t is cell (ix, iy, iz).
__host__ __device__ inline long ghost_count(int nx, int ny, int nz) {
  return 4L * nx * ny + 4L * nx * (nz - 4) + 4L * (ny - 4) * (nz - 4);
}
__host__ __device__ inline int ghost_outer(int k, int n) { return k < 2 ? k : n - 4 + k; }
__host__ __device__ inline void ghost_cell(int nx, int ny, int nz, long t, int &ix, int &iy, int &iz) {
  const unsigned nxu = nx, nxy = (unsigned)nx * (unsigned)ny;
  const unsigned n1 = 4u * nxy, n2 = 4u * nxu * (unsigned)(nz - 4);
  unsigned u = (unsigned)t;
  if (u < n1) {
    const unsigned l = u / nxy, r = u - l * nxy, y = r / nxu;
    iz = ghost_outer((int)l, nz);
    iy = (int)y;
    ix = (int)(r - y * nxu);
  } else if (u < n1 + n2) {
    u -= n1;
    const unsigned per = 4u * nxu, z = u / per, r = u - z * per, row = r / nxu;
    iz = (int)z + 2;
    iy = ghost_outer((int)row, ny);
    ix = (int)(r - row * nxu);
  } else {
    u -= n1 + n2;
    const unsigned per = 4u * (unsigned)(ny - 4), z = u / per, r = u - z * per, y = r >> 2;
    iz = (int)z + 2;
    iy = (int)y + 2;
    ix = ghost_outer((int)(r & 3u), nx);
  }
}

}  // namespace vh
