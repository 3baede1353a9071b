// Mixed-precision kernels of the hydro step on gfx950.
//
// Precision split: the conserved variables Q and Qh, their accumulation, the
// MUSCL reconstruction, con2prim and the velocity gradients of NSquant stay
// FP64.  The HLLE interface fluxes, the viscous interface fluxes and sources,
// and the Israel-Stewart relaxation with its cross terms run in FP32.  pi^{mu nu}
// and Pi are stored in FP32.
//
// Structure: one fused kernel per predictor/corrector stage.  Each thread owns
// one cell and evaluates its six interfaces itself; the two con2prim of an
// interface are recomputed by both neighbours.  Nothing is written to HBM but
// the stage result, so a stage moves about its compulsory bytes (read Q, and
// Qh in the corrector, write Qh or Q).  The con2prim algorithm is a template
// parameter: the reference iteration (C2P_REFERENCE), the closed-form
// conformal root (C2P_CLOSED), or an FP64 Newton polish (vh_mlp.h) started
// from a per-cell guess array gv (C2P_LEARNED: the MLP guess of the cell that
// owns the state; C2P_NEWTON: v = 0.5 everywhere, the ablation).
#include "vh_kernels_mixed.h"
#include "vh_vort.h"
#include "vh_mlp.h"

#include <algorithm>
#include <vector>

namespace vh {
namespace mx {

// ------------------------------------------------------------- helpers
// cell coordinates of local index c (a context holds fewer than 2^29 cells,
// checked at creation, so 32-bit divisions suffice)
__device__ __forceinline__ void coords_of(const Geom &g, long c, int &ix, int &iy, int &iz) {
  const unsigned u = (unsigned)c, nx = (unsigned)g.nx, ny = (unsigned)g.ny;
  const unsigned r = u / nx;
  ix = (int)(u - r * nx);
  const unsigned q = r / ny;
  iy = (int)(r - q * ny);
  iz = (int)q;
}

// Grid of workgroup bk and the thread's local cell cl.  A single grid comes
// by value in the kernel arguments, so no dependent table load precedes the
// cell's own loads; a batch looks its grid up by workgroup.
__device__ __forceinline__ Geom block_grid(const Grids &G, long bk, long &cl) {
  if (G.nev == 1) {
    cl = bk * blockDim.x + threadIdx.x;
    return G.g0;
  }
  const int gi = G.bev[bk];
  cl = (bk - G.bfirst[gi]) * blockDim.x + threadIdx.x;
  return G.ev[gi];
}
// The work item of a per-cell thread: grid g, its cell c (array index) and
// coordinates; false past the grid's last cell.  With G.sy > 0 workgroups
// are XCD-contiguous and a grid's cells are visited in y-strips of sy rows,
// each swept over all eta in contiguous chunks of sy*nx cells: the eta- and
// y-neighbours a workgroup gathers were then fetched into the same L2 by the
// workgroups just before it (near-compulsory HBM traffic).  A bijection of
// the grid's cells, so results do not depend on it.
__device__ __forceinline__ bool work_cell(const Grids &G, Geom &g, long &c, int &ix, int &iy,
                                          int &iz) {
  long cl;
  g = block_grid(G, G.sy > 0 ? xcd_block() : (long)blockIdx.x, cl);
  if (cl >= g.N) return false;
  strip_cell(g.nx, g.ny, g.nz, G.sy, cl, ix, iy, iz);
  c = g.idx(ix, iy, iz);
  return true;
}

// grid of row segment sgl (over all grids) and the segment's index in it
__device__ __forceinline__ Geom seg_grid(const Grids &G, long sgl, long &seg) {
  if (G.nev == 1) {
    seg = sgl;
    return G.g0;
  }
  const int gi = G.sev[sgl];
  seg = sgl - G.sfirst[gi];
  return G.ev[gi];
}
// Element c of an array whose base is the same for the whole wave: a context
// holds fewer than 2^29 cells (checked at creation), so the byte offset fits
// 32 bits and the load takes the base in scalar registers and a 32-bit lane
// offset (no 64-bit address arithmetic per lane).
template <class T>
__device__ __forceinline__ T ldu(const T *__restrict__ a, long c) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(a) +
                                      ((unsigned)c * (unsigned)sizeof(T)));
}
template <class T>
__device__ __forceinline__ void stu(T *__restrict__ a, long c, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(a) + ((unsigned)c * (unsigned)sizeof(T))) = v;
}
// A store of a result the kernel does not read back and the next kernel
// reads long after it has left the L2 (a stage's arrays are hundreds of MB):
// nontemporal, so that it does not evict the lines the kernel's own gathers
// reuse (the stage, relaxation and advection kernels: 2-6 % faster; the
// primitive-state kernel, whose outputs the relaxation gathers next, runs 5 %
// slower with it and keeps plain stores).
template <class T>
__device__ __forceinline__ void stn(T *__restrict__ a, long c, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(a) +
                                                       ((unsigned)c * (unsigned)sizeof(T))));
}
__device__ __forceinline__ void stn(float4 *__restrict__ a, long c, float4 v) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<f4 *>(reinterpret_cast<char *>(a) +
                                                        ((unsigned)c * (unsigned)sizeof(float4))));
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fmin_(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_(float a, float b) { return (a < b) ? b : a; }
// FP64 1/x and 1/sqrt(x) to about an ulp, not correctly rounded: the hardware
// estimate refined by two Newton steps (each doubles the correct bits).  For
// the per-cell Lorentz factors and 1/u^k of the mixed relaxation kernel,
// whose results are scaled into FP32.  x must be finite and nonzero (1/inf
// and 1/0 come out NaN, not 0 and inf): the uniform factors of the grid
// spacing, where a grid of one cell along an axis has spacing inf, keep the
// division.
__device__ __forceinline__ double drcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.), y);
  return fma(y, fma(-x, y, 1.), y);
}
__device__ __forceinline__ double drsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = fma(y, fma(-h * y, y, 0.5), y);
  return fma(y, fma(-h * y, y, 0.5), y);
}

// FP32 copy of a primitive state; gamma^2 is formed in FP64 first so that
// |v| close to 1 does not round to 1 in FP32
struct PrimF {
  float e, p, nb, nq, ns, vx, vy, vz, g2;
};

__device__ __forceinline__ PrimF to_f(const Prim &a) {
  const double g2 = 1. / (1. - (a.vx * a.vx + a.vy * a.vy + a.vz * a.vz));
  return PrimF{(float)a.e, (float)a.p, (float)a.nb, (float)a.nq, (float)a.ns,
               (float)a.vx, (float)a.vy, (float)a.vz, (float)g2};
}

// Closed-form con2prim of the conformal EoS p = e/3: the root of
// M v^2 - 4 Q_t v + 3 M = 0 in its cancellation-free form, then one FP64
// Newton step on f(v) = (Q_t + p) v - M with f' = Q_t + p - M v/3.  Same
// branches as transformPV (src/rmn.cpp:47-70): vacuum, at rest, rescale of
// |Q_vec| > Q_t to 0.9999 Q_t.
__device__ __forceinline__ void c2p_closed(double q[7], Prim &r) {
  const double Qt = q[T_];
  double Qx = q[X_], Qy = q[Y_], Qz = q[Z_];
  double M = sqrt(Qx * Qx + Qy * Qy + Qz * Qz);
  double e = 0., p = 0., vx = 0., vy = 0., vz = 0., nb = 0., nq = 0., ns = 0.;
  if (Qt > 0.) {
    if (M == 0.) {
      e = Qt;
      nb = q[NB_];
      nq = q[NQ_];
      ns = q[NS_];
      p = e * (1. / 3.);
    } else {
      if (M > Qt) {
        const double s = 0.9999 * Qt / M;
        Qx *= s;
        Qy *= s;
        Qz *= s;
        M = Qt * 0.9999;
      }
      // in x = |Q|/Q_t: Q_t^2 underflows FP64 for Q_t below 1e-154
      const double x = M / Qt;
      double v = 3. * x / (2. + sqrt(4. - 3. * x * x));
      const double pp = (Qt - M * v) * (1. / 3.);
      v -= ((Qt + pp) * v - M) / ((Qt + pp) - M * v * (1. / 3.));
      e = Qt - M * v;
      p = e * (1. / 3.);
      q[X_] = Qx;  // the rescaled input, as transformPV leaves it
      q[Y_] = Qy;
      q[Z_] = Qz;
      const double iM = v / M;
      vx = Qx * iM;
      vy = Qy * iM;
      vz = Qz * iM;
      const double g = sqrt(1. - v * v);
      nb = q[NB_] * g;
      nq = q[NQ_] * g;
      ns = q[NS_] * g;
    }
  }
  r = Prim{e, p, nb, nq, ns, vx, vy, vz};
}

// FP32 closed-form con2prim of p = e/3 for an FP64 state, for the interface
// states of the FP32 flux.  Only |Q|^2 and Q_t^2 - |Q|^2 are formed in FP64
// (the latter cancels as |v| -> 1), scaled by an FP32 reciprocal of Q_t so that
// nothing overflows FP32.  With x = |Q|/Q_t and y = 1 - x:
//   s = sqrt(4 - 3x^2) = sqrt(1 + y (6 - 3y)),   v = 3x / (2 + s),
//   1 - v = 3y (3 - y + s) / ((1 + s)(2 + s)),
//   e = Q_t ((1 - v) + y v),   1/gamma^2 = (1 - v)(1 + v),
// all sums of positive terms, so v, e and gamma^2 keep FP32 relative accuracy
// (a few ulp) for every |v| < 1.  Branches as transformPV (src/rmn.cpp:47-70):
// vacuum (Q_t <= 0), at rest, and |Q| > Q_t treated as |Q| = 0.9999 Q_t.
__device__ __forceinline__ PrimF c2p_fast(const double q[7]) {
  PrimF r{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f};
  const double Qt = q[T_];
  // Q^tau below 1e-30 is vacuum at FP32 precision (and 1/(float)Q^tau would
  // overflow to inf for Q^tau below the FP32 range, giving 0 * inf velocities)
  if (!(Qt > 1e-30)) return r;
  const double M2 = q[X_] * q[X_] + q[Y_] * q[Y_] + q[Z_] * q[Z_];
  if (M2 == 0.) {
    r.e = (float)Qt;
    r.p = r.e * (1.f / 3.f);
    r.nb = (float)q[NB_];
    r.nq = (float)q[NQ_];
    r.ns = (float)q[NS_];
    return r;
  }
  const double rq = (double)frcp((float)Qt);
  const double rq2 = rq * rq;
  const float a = (float)((Qt * Qt - M2) * rq2);  // (1 - x^2) t^2
  const float t = (float)(Qt * rq);                 // 1 up to the rcp error
  const float mm = fsqrt((float)(M2 * rq2));        // x t
  float y = a * frcp(t * (t + mm)), x = mm * frcp(t);
  if (!(a > 0.f)) {
    y = 1e-4f;
    x = 0.9999f;
  }
  const float s = fsqrt(1.f + y * (6.f - 3.f * y));
  const float omv = 3.f * y * (3.f - y + s) * frcp((1.f + s) * (2.f + s));
  const float v = 3.f * x * frcp(2.f + s);
  const float ig2 = omv * (2.f - omv);
  r.e = (float)Qt * (omv + y * v);
  r.p = r.e * (1.f / 3.f);
  r.g2 = frcp(ig2);
  const float vm = v * frcp(mm);
  r.vx = (float)(q[X_] * rq) * vm;
  r.vy = (float)(q[Y_] * rq) * vm;
  r.vz = (float)(q[Z_] * rq) * vm;
  const float ig = fsqrt(ig2);
  r.nb = (float)q[NB_] * ig;
  r.nq = (float)q[NQ_] * ig;
  r.ns = (float)q[NS_] * ig;
  return r;
}

// FP32 Newton con2prim from a start v0 (the learned guess, or 0.5), for the
// FP32 consumers (interface states of the FP32 flux, the Milne source of the
// FP32-accurate stage).  Unknown w = 1 - v, so that |v| -> 1 keeps full
// relative precision; with x = |Q|/Q_t, y = 1 - x (from FP64 Q_t^2 - |Q|^2 as
// in c2p_fast) and e/Q_t = 1 - x v = y + x w:
//   f(w) = (1 + p/Q_t)(1 - w) - x = y - w + (p/Q_t)(1 - w),
//   f'(w) = -(1 + p/Q_t) + c_s^2 x (1 - w) < 0.
// f(y) = x p/Q_t >= 0 and f(1) = -x <= 0 bracket the root in [y, 1].  For
// p = e/3, f is concave, so Newton from any start in the bracket reaches the
// right of the root in one update and then decreases monotonically to it.
// With 2/3 <= |f'| <= 4/3 and |f''| <= 2/3 on the bracket, the error after an
// update dw is at most 2 dw^2, so an update below 1.5e-4 w leaves w accurate
// to 4.5e-8 w: the loop stops there.  Same vacuum, rest and
// |Q| >= Q_t branches as c2p_fast.
__device__ __forceinline__ PrimF c2p_newton_f32(const DevEos &eos, const double q[7], float v0) {
  PrimF r{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f};
  const double Qt = q[T_];
  if (!(Qt > 1e-30)) return r;
  const double M2 = q[X_] * q[X_] + q[Y_] * q[Y_] + q[Z_] * q[Z_];
  if (M2 == 0.) {
    r.e = (float)Qt;
    r.p = eos.pf(r.e);
    r.nb = (float)q[NB_];
    r.nq = (float)q[NQ_];
    r.ns = (float)q[NS_];
    return r;
  }
  const double rq = (double)frcp((float)Qt);
  const double rq2 = rq * rq;
  const float a = (float)((Qt * Qt - M2) * rq2);
  const float t = (float)(Qt * rq);
  const float mm = fsqrt((float)(M2 * rq2));
  float y = a * frcp(t * (t + mm)), x = mm * frcp(t);
  if (!(a > 0.f)) {
    y = 1e-4f;
    x = 0.9999f;
  }
  const float Qtf = (float)Qt, iQt = frcp(Qtf);
  float w = 1.f - v0;
  if (!(w >= y)) w = y;  // also a NaN guess
  if (!(w <= 1.f)) w = 1.f;
#pragma unroll 1
  for (int it = 0; it < 16; it++) {
    const float eh = y + x * w;
    float ph, dpde;
    if (eos.kind == VH_EOS_TABLE) {  // one table lookup for p and dp/de
      ph = eos.p_dpdef(eh * Qtf, dpde) * iQt;
    } else {
      ph = eos.pf(eh * Qtf) * iQt;
      dpde = eos.cs2f(eh * Qtf);
    }
    const float f = y - w + ph * (1.f - w);
    const float df = -(1.f + ph) + dpde * x * (1.f - w);
    float wn = w - f * frcp(df);
    wn = fminf(fmaxf(wn, y), 1.f);
    const float dw = fabsf(wn - w);
    w = wn;
    if (dw <= 1.5e-4f * w) break;
  }
  const float v = 1.f - w;
  const float ig2 = w * (2.f - w);  // 1 - v^2
  r.e = Qtf * (y + x * w);
  r.p = eos.pf(r.e);
  r.g2 = frcp(ig2);
  const float vm = v * frcp(mm);
  r.vx = (float)(q[X_] * rq) * vm;
  r.vy = (float)(q[Y_] * rq) * vm;
  r.vz = (float)(q[Z_] * rq) * vm;
  const float ig = fsqrt(ig2);
  r.nb = (float)q[NB_] * ig;
  r.nq = (float)q[NQ_] * ig;
  r.ns = (float)q[NS_] * ig;
  return r;
}

// con2prim by mode; v0: the learned guess (C2P_LEARNED only)
template <int CM>
__device__ __forceinline__ void c2p_mode(const DevEos &eos, double q[7], Prim &r, float v0) {
  if (CM == VH_C2P_CLOSED)
    c2p_closed(q, r);
  else if (CM == VH_C2P_LEARNED)
    ml::c2p_guess(eos, q, (double)v0, r);
  else if (CM == VH_C2P_NEWTON)
    ml::c2p_guess(eos, q, 0.5, r);
  else
    c2p_ref<false>(eos, q, r);
}
template <int CM>
__device__ __forceinline__ float guess_of(const float *__restrict__ gv, long c) {
  return CM == VH_C2P_LEARNED ? ldu(gv, c) : 0.5f;
}

// The kernels that evaluate the EoS are compiled per EoS kind EK (VH_EOS_IDEAL
// or VH_EOS_TABLE, the kinds the mixed path takes; kinds 2 and 3 are refused
// when the precision is set), see eos_of_kind.

// MUSCL face state (src/cll.cpp:106-204) in FP64, con2prim by mode, as the
// FP32 state the flux takes.  NC: conserved components carried (7, or 4 for
// a state without charges, whose N^tau components are zero and stay zero).
template <int SIDE, int CM, int NC>
__device__ __forceinline__ PrimF face_state(const DevEos &eos, const double *__restrict__ Qs,
                                            const float *__restrict__ gv, long N, long cm,
                                            long c0, long cp, double itau) {
  double q[7];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    if (k >= NC) {
      q[k] = 0.;
      continue;
    }
    const double *Qk = Qs + k * N;
    const double a = ldu(Qk, c0);
    // minmod(x/2, y/2) = minmod(x, y)/2 (the halving is exact; only products
    // x y below ~1e-308 can decide the sign test differently), and a -+ mm/2
    // as one fused operation rounds as a -+ (mm/2) does
    const double mm = minmod(ldu(Qk, cp) - a, a - ldu(Qk, cm));
    q[k] = fma(SIDE < 0 ? -0.5 : 0.5, mm, a) * itau;
  }
  if (CM == VH_C2P_CLOSED) return c2p_fast(q);
  if (CM == VH_C2P_LEARNED || CM == VH_C2P_NEWTON) return c2p_newton_f32(eos, q, guess_of<CM>(gv, c0));
  Prim r;
  c2p_mode<CM>(eos, q, r, guess_of<CM>(gv, c0));
  return to_f(r);
}

// HLLE interface flux in FP32 (Hydro::hlle_flux, src/hdo.cpp:77-305: same
// wave speeds, vacuum rules and skips).  Returns false if the reference would
// return without a flux.  pre = tauFactor * dta / dx.  A negative pressure of
// the right state (a table EoS whose spline undershoots below its clamped
// nodes) makes the reference evaluate both face states again from Q, in either
// stage, and go on with them and with El, Er of the first states
// (src/hdo.cpp:134-140): reeval(L, R) does that.  *sp receives the interface's
// signal speed max(-b_l, b_r) when a flux is formed.
template <int D, int NC, class RE>
__device__ __forceinline__ bool hlle_f32(PrimF L, PrimF R, float itau, float itf, float pre,
                                         bool mskip, float F[7], int *status, float &sp,
                                         bool table, RE &&reeval) {
  if (L.e < 0.f) L.e = L.p = 0.f;
  if (R.e < 0.f) R.e = R.p = 0.f;
  if (L.e > 1e10f) {
    atomicOr(status, VH_ST_BIG_E);
    return false;
  }
  if (L.e == 0.f && R.e == 0.f) return false;
  const float El = (L.e + L.p) * L.g2, Er = (R.e + R.p) * R.g2;
  // table EoS only: p = e/3 >= 0 for e >= 0 (a NaN fails the test as well), so
  // an ideal-gas kernel carries no re-evaluation code
  if (table && R.p < 0.f) reeval(L, R);
  if (mskip) return false;
  const float g2l = L.g2, g2r = R.g2;
  const float wl = (L.e + L.p) * g2l, wr = (R.e + R.p) * g2r;
  const float gl = fsqrt(g2l), gr = fsqrt(g2r);
  float Ul[7], Ur[7];
  Ul[X_] = wl * L.vx;
  Ul[Y_] = wl * L.vy;
  Ul[Z_] = wl * L.vz;
  Ul[T_] = wl - L.p;
  if (NC == 7) {
    Ul[NB_] = gl * L.nb;
    Ul[NQ_] = gl * L.nq;
    Ul[NS_] = gl * L.ns;
  }
  Ur[X_] = wr * R.vx;
  Ur[Y_] = wr * R.vy;
  Ur[Z_] = wr * R.vz;
  Ur[T_] = wr - R.p;
  if (NC == 7) {
    Ur[NB_] = gr * R.nb;
    Ur[NQ_] = gr * R.nq;
    Ur[NS_] = gr * R.ns;
  }
  const float vl = D == 1 ? L.vx : (D == 2 ? L.vy : L.vz);
  const float vr = D == 1 ? R.vx : (D == 2 ? R.vy : R.vz);
  float Fl[7], Fr[7];
#pragma unroll
  for (int k = 0; k < NC; k++) {
    Fl[k] = Ul[k] * vl;
    Fr[k] = Ur[k] * vr;
  }
  Fl[T_] += L.p * vl;
  Fr[T_] += R.p * vr;
  Fl[D] += L.p;
  Fr[D] += R.p;
  if (D == 3) {
#pragma unroll
    for (int k = 0; k < NC; k++) {
      Fl[k] *= itf;
      Fr[k] *= itf;
    }
  }
  const float cs = 0.57735026918962576f;
  const float sEl = fsqrt(El), sEr = fsqrt(Er);
  // both sides below the FP32 range (e ~ 1e-40 in the FP64 state, flushed
  // here): vacuum to FP32 precision, no flux (instead of 0 * inf)
  if (!(sEl + sEr > 0.f)) return false;
  const float isum = frcp(sEl + sEr);
  const float dv = vl - vr;
  const float csb = fsqrt(1.f / 3.f + 0.5f * sEl * sEr * isum * isum * dv * dv);
  const float vb = (sEl * vl + sEr * vr) * isum;
  float bl = fmin_(0.f, fmin_((vb - csb) * frcp(1.f - vb * csb), (vl - cs) * frcp(1.f - vl * cs)));
  float br = fmax_(0.f, fmax_((vb + csb) * frcp(1.f + vb * csb), (vr + cs) * frcp(1.f + vr * cs)));
  if (D == 3) {
    bl *= itau;
    br *= itau;
    if (L.e == 0.f) bl = -itau;
    if (R.e == 0.f) br = itau;
  } else {
    if (L.e == 0.f) bl = -1.f;
    if (R.e == 0.f) br = 1.f;
  }
  if (bl == 0.f && br == 0.f) return false;
  const float s = pre * frcp(br - bl);
  const float bb = bl * br;
#pragma unroll
  for (int k = 0; k < NC; k++) F[k] = s * (br * Fl[k] - bl * Fr[k] - bb * (Ul[k] - Ur[k]));
  // the reference tests the N_b flux; without charges a NaN state shows in F^tau
  if (F[NC == 7 ? NB_ : T_] != F[NC == 7 ? NB_ : T_]) {
    atomicOr(status, VH_ST_NAN_FLUX);
    return false;
  }
  sp = fmax_(sp, fmax_(-bl, br));
  return true;
}

// the two interfaces of cell c along axis D, added into acc in the order
// +F(i-1/2), -F(i+1/2).  Qs: the stage's state; Qr: Q (the states of the
// negative-pressure re-evaluation); sp: the largest signal speed so far.
template <int D, int MODE, int CM, int NC>
__device__ __forceinline__ void axis_fluxes(const Geom &g, const DevEos &eos,
                                            const double *__restrict__ Qs,
                                            const double *__restrict__ Qr,
                                            const float *__restrict__ gv,
                                            const double *__restrict__ m, int m_full, int i,
                                            int n, long c, long s, double tau, double dt,
                                            double acc[7], int *status, float &sp) {
  if (n < 2) return;
  const long N = g.P;
  const double itau = drcp(tau);  // tau > 0 and finite
  const double tauFactor = MODE == 0 ? tau + 0.25 * dt : tau + 0.5 * dt;
  const double dta = MODE == 0 ? 0.5 * dt : dt;
  const double dx = D == 1 ? g.dx : (D == 2 ? g.dy : g.dz);
  const float pre = (float)(tauFactor * dta / dx);  // dx = inf on a one-cell axis
  const float itf = (float)drcp(tauFactor), itauf = (float)itau;
  const long cm = i > 0 ? c - s : c, cp = i < n - 1 ? c + s : c;
  const PrimF L0 = face_state<-1, CM, NC>(eos, Qs, gv, N, cm, c, cp, itau);
  const PrimF R0 = face_state<+1, CM, NC>(eos, Qs, gv, N, cm, c, cp, itau);
  float F[7];
  if (i > 0) {
    const long cmm = i > 1 ? cm - s : cm;
    const PrimF Rm = face_state<+1, CM, NC>(eos, Qs, gv, N, cmm, cm, c, itau);
    const bool mk = !m_full && ldu(m + (D - 1) * N, cm) < 1. && ldu(m + (D - 1) * N, c) < 1.;
    auto re = [&](PrimF &a, PrimF &b) {
      a = face_state<+1, CM, NC>(eos, Qr, gv, N, cmm, cm, c, itau);
      b = face_state<-1, CM, NC>(eos, Qr, gv, N, cm, c, cp, itau);
    };
    if (hlle_f32<D, NC>(Rm, L0, itauf, itf, pre, mk, F, status, sp, eos.kind == VH_EOS_TABLE, re))
#pragma unroll
      for (int k = 0; k < NC; k++) acc[k] += (double)F[k];
  }
  if (i < n - 1) {
    const long cpp = i + 2 < n ? cp + s : cp;
    const PrimF Lp = face_state<-1, CM, NC>(eos, Qs, gv, N, c, cp, cpp, itau);
    const bool mk = !m_full && ldu(m + (D - 1) * N, c) < 1. && ldu(m + (D - 1) * N, cp) < 1.;
    auto re = [&](PrimF &a, PrimF &b) {
      a = face_state<+1, CM, NC>(eos, Qr, gv, N, cm, c, cp, itau);
      b = face_state<-1, CM, NC>(eos, Qr, gv, N, c, cp, cpp, itau);
    };
    if (hlle_f32<D, NC>(R0, Lp, itauf, itf, pre, mk, F, status, sp, eos.kind == VH_EOS_TABLE, re))
#pragma unroll
      for (int k = 0; k < NC; k++) acc[k] -= (double)F[k];
  }
}

// Milne source on the unreconstructed state (src/hdo.cpp:307-353) and the
// stage update of cell c with the accumulated interface fluxes acc.
template <int MODE, int CM, int NC>
__device__ __forceinline__ void finish_cell(const DevEos &eos, const double *__restrict__ Q,
                                            const double *__restrict__ Qs,
                                            const float *__restrict__ gv, long N, long c,
                                            double tau, double dt, double acc[7],
                                            double *__restrict__ out) {
  const double tau1 = MODE == 0 ? tau : tau + 0.5 * dt;
  const double sdt = MODE == 0 ? 0.5 * dt : dt;
  const double it1 = drcp(tau1);
  double q[7];
#pragma unroll
  for (int k = 0; k < 7; k++) q[k] = k < NC ? ldu(Qs + k * N, c) * it1 : 0.;
  double vz, pr;
  if (CM == VH_C2P_CLOSED) {
    const PrimF r = c2p_fast(q);
    vz = r.vz;
    pr = r.p;
  } else if (CM == VH_C2P_LEARNED || CM == VH_C2P_NEWTON) {
    const PrimF r = c2p_newton_f32(eos, q, guess_of<CM>(gv, c));
    vz = r.vz;
    pr = r.p;
  } else {
    Prim r;
    c2p_mode<CM>(eos, q, r, guess_of<CM>(gv, c));
    vz = r.vz;
    pr = r.p;
  }
  acc[T_] += (-q[T_] * vz * vz - pr * (1. + vz * vz)) * sdt;
  acc[Z_] += -q[Z_] * sdt;
  if (MODE == 0) {
#pragma unroll
    for (int k = 0; k < NC; k++) stn(out + k * N, c, ldu(Q + k * N, c) + acc[k]);
  } else {
    const bool skip = ldu(Q, c) + acc[0] < 0.;
#pragma unroll
    for (int k = 0; k < NC; k++) {
      const double q0 = ldu(Q + k * N, c);
      stn(out + k * N, c, skip ? q0 : q0 + acc[k]);
    }
  }
}

// Fused stage: interface fluxes along x, y, eta + Milne source + update.
// MODE 0 (predictor): reads Q, writes Qh = Q + flux.
// MODE 1 (corrector): reads Q and Qh, writes Qnew = Q + flux unless
// Q0 + flux0 < 0 (src/cll.cpp:70-74).
// sig: the stage's three signal-speed slot groups (x, y, eta; kSigGroup
// words apart), one atomic per wave and group.
template <int MODE, int CM, int NC, int EK>
__global__ void __launch_bounds__(256) VH_OCC(VH_W_MX_STAGE)
k_stage(Grids G, DevEos eos_a, const double *__restrict__ Q, const double *__restrict__ Qh,
        const float *__restrict__ gv, const double *__restrict__ m, int m_full,
        const double *__restrict__ tp, double dt, double *__restrict__ out,
        int *__restrict__ status, unsigned long long *__restrict__ sig) {
  const DevEos eos = eos_of_kind<EK>(eos_a);
  const double tau = *tp;
  Geom g;
  long c;
  int ix, iy, iz;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  if (work_cell(G, g, c, ix, iy, iz)) {
    const long N = g.P;
    const double *Qs = MODE == 0 ? Q : Qh;
    double acc[7];
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] = 0.;
    axis_fluxes<1, MODE, CM, NC>(g, eos, Qs, Q, gv, m, m_full, ix, g.nx, c, 1, tau, dt, acc,
                                 status, sx);
    axis_fluxes<2, MODE, CM, NC>(g, eos, Qs, Q, gv, m, m_full, iy, g.ny, c, g.nx, tau, dt, acc,
                                 status, sy);
    axis_fluxes<3, MODE, CM, NC>(g, eos, Qs, Q, gv, m, m_full, iz, g.nz, c, (long)g.nx * g.ny,
                                 tau, dt, acc, status, sz);
    finish_cell<MODE, CM, NC>(eos, Q, Qs, gv, N, c, tau, dt, acc, out);
  }
  wave_speed_max(sig, (double)sx);
  wave_speed_max(sig + kSigGroup, (double)sy);
  wave_speed_max(sig + 2 * kSigGroup, (double)sz);
}

// Interface flux between cells a (left) and b (right) along axis D: MUSCL
// states of a (neighbours amm, b) and b (neighbours a, bp), HLLE in FP32.
// Same states, mask and flux as the interfaces of axis_fluxes.
template <int D, int CM, int NC>
__device__ __forceinline__ bool iface(const DevEos &eos, const double *__restrict__ Qs,
                                      const double *__restrict__ Qr,
                                      const float *__restrict__ gv,
                                      const double *__restrict__ m, int m_full, long N, long amm,
                                      long a, long b, long bp, double itau, float itauf,
                                      float itf, float pre, float F[7], int *status, float &sp) {
  const PrimF L = face_state<+1, CM, NC>(eos, Qs, gv, N, amm, a, b, itau);
  const PrimF R = face_state<-1, CM, NC>(eos, Qs, gv, N, a, b, bp, itau);
  const bool mk = !m_full && ldu(m + (D - 1) * N, a) < 1. && ldu(m + (D - 1) * N, b) < 1.;
  auto re = [&](PrimF &l, PrimF &r) {
    l = face_state<+1, CM, NC>(eos, Qr, gv, N, amm, a, b, itau);
    r = face_state<-1, CM, NC>(eos, Qr, gv, N, a, b, bp, itau);
  };
  return hlle_f32<D, NC>(L, R, itauf, itf, pre, mk, F, status, sp, eos.kind == VH_EOS_TABLE,
                         re);
}

// Row stage: every x- and eta-interface is evaluated once, without barriers.
// A wave owns 64 consecutive cells of the flattened (x, y) plane (waves
// overlap by one cell, so 63 of 64 lanes update a cell whatever nx is) and
// marches over CZ cells in eta.  Lane l evaluates the lower x-interface of its
// cell and receives the upper one from lane l + 1 by a cross-lane shuffle; the
// upper eta-interface stays in registers as the next cell's lower one; the
// two y-interfaces are evaluated per lane.  Interface states, fluxes and the
// accumulation order are those of k_stage: identical results.
template <int MODE, int CM, int NC, int EK>
__global__ void __launch_bounds__(256) VH_OCC(VH_W_MX_ROWS)
k_stage_rows(Grids G, DevEos eos_a, const double *__restrict__ Q, const double *__restrict__ Qh,
             const float *__restrict__ gv, const double *__restrict__ m, int m_full,
             const double *__restrict__ tp, double dt, double *__restrict__ out,
             int *__restrict__ status, unsigned long long *__restrict__ sig, int CZ) {
  const DevEos eos = eos_of_kind<EK>(eos_a);
  const int lane = threadIdx.x & 63;
  // XCD-contiguous workgroups: the segments an XCD processes at once are
  // consecutive, so the y-neighbour rows of a segment (nx cells away) were
  // fetched into that XCD's L2 by the waves just before it
  const long w = xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long sgl = w % G.nseg, zc = w / G.nseg;  // row segment over all grids, eta chunk
  long seg;
  const Geom g = seg_grid(G, sgl, seg);
  const long nxy = (long)g.nx * g.ny;
  const int z0 = (int)zc * CZ;
  if (z0 >= g.nz) return;  // whole wave
  const int z1 = min(g.nz, z0 + CZ);
  const long f = seg * 63 + lane;
  const bool cx = f < nxy;
  const bool upd = cx && lane < 63;
  const unsigned fu = (unsigned)f, nxu = (unsigned)g.nx;
  const int iy = cx ? (int)(fu / nxu) : 0, ix = cx ? (int)(fu - (unsigned)iy * nxu) : 0;
  const long N = g.P, sy = g.nx, sz = nxy;
  const double tau = *tp;
  const double *Qs = MODE == 0 ? Q : Qh;
  const double itau = drcp(tau);  // as in axis_fluxes: identical interface states
  const double tauFactor = MODE == 0 ? tau + 0.25 * dt : tau + 0.5 * dt;
  const double dta = MODE == 0 ? 0.5 * dt : dt;
  const float prx = (float)(tauFactor * dta / g.dx), pry = (float)(tauFactor * dta / g.dy),
              prz = (float)(tauFactor * dta / g.dz);
  const float itf = (float)drcp(tauFactor), itauf = (float)itau;
  float fzl[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float fzu[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  bool hzl = false;
  float spx = 0.f, spy = 0.f, spz = 0.f;
  for (int iz = z0; iz < z1; iz++) {
    const long c = cx ? g.base + f + iz * sz : 0;
    float fx[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float fyl[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float fyu[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool hx = false, hyl = false, hyu = false, hzu = false;
    if (cx && ix >= 1)
      hx = iface<1, CM, NC>(eos, Qs, Q, gv, m, m_full, N, ix >= 2 ? c - 2 : c - 1, c - 1, c,
                            ix + 1 < g.nx ? c + 1 : c, itau, itauf, itf, prx, fx, status, spx);
    float fxu[7];
#pragma unroll
    for (int k = 0; k < NC; k++) fxu[k] = __shfl_down(fx[k], 1, 64);
    const bool hxu = __shfl_down(hx ? 1 : 0, 1, 64) != 0 && ix + 1 < g.nx;
    if (cx && iy >= 1)
      hyl = iface<2, CM, NC>(eos, Qs, Q, gv, m, m_full, N, iy >= 2 ? c - 2 * sy : c - sy, c - sy,
                             c, iy + 1 < g.ny ? c + sy : c, itau, itauf, itf, pry, fyl, status,
                             spy);
    if (cx && iy + 1 < g.ny)
      hyu = iface<2, CM, NC>(eos, Qs, Q, gv, m, m_full, N, iy >= 1 ? c - sy : c, c, c + sy,
                             iy + 2 < g.ny ? c + 2 * sy : c + sy, itau, itauf, itf, pry, fyu,
                             status, spy);
    if (iz == z0 && cx && iz >= 1)
      hzl = iface<3, CM, NC>(eos, Qs, Q, gv, m, m_full, N, iz >= 2 ? c - 2 * sz : c - sz, c - sz,
                             c, iz + 1 < g.nz ? c + sz : c, itau, itauf, itf, prz, fzl, status,
                             spz);
    if (cx && iz + 1 < g.nz)
      hzu = iface<3, CM, NC>(eos, Qs, Q, gv, m, m_full, N, iz >= 1 ? c - sz : c, c, c + sz,
                             iz + 2 < g.nz ? c + 2 * sz : c + sz, itau, itauf, itf, prz, fzu,
                             status, spz);
    if (upd) {
      double acc[7];
#pragma unroll
      for (int k = 0; k < 7; k++) acc[k] = 0.;
      if (hx)
#pragma unroll
        for (int k = 0; k < NC; k++) acc[k] += (double)fx[k];
      if (hxu)
#pragma unroll
        for (int k = 0; k < NC; k++) acc[k] -= (double)fxu[k];
      if (hyl)
#pragma unroll
        for (int k = 0; k < NC; k++) acc[k] += (double)fyl[k];
      if (hyu)
#pragma unroll
        for (int k = 0; k < NC; k++) acc[k] -= (double)fyu[k];
      if (hzl)
#pragma unroll
        for (int k = 0; k < NC; k++) acc[k] += (double)fzl[k];
      if (hzu)
#pragma unroll
        for (int k = 0; k < NC; k++) acc[k] -= (double)fzu[k];
      finish_cell<MODE, CM, NC>(eos, Q, Qs, gv, N, c, tau, dt, acc, out);
    }
#pragma unroll
    for (int k = 0; k < NC; k++) fzl[k] = fzu[k];
    hzl = hzu;
  }
  wave_speed_max(sig, (double)spx);
  wave_speed_max(sig + kSigGroup, (double)spy);
  wave_speed_max(sig + 2 * kSigGroup, (double)spz);
}

// Band stage: the y-interfaces shared as well.  The band plane of a grid
// lays the first rows of its bands of CY rows end to end (cell (ix, band) at
// band * nx + ix).  A wave owns 64 consecutive cells of it (waves overlap by
// one cell, so 63 of 64 lanes update cells whatever nx and CY are) and
// marches CZ cells in eta; at each eta cell every lane walks the CY rows of
// its band.  Lane l evaluates the lower x-interface of its cell and receives
// the upper one from lane l + 1 by a cross-lane shuffle; the upper
// y-interface of a row is carried in registers to the next row of the band
// (the first row of a band evaluates its lower one); the upper eta-interface
// of each row is kept in LDS, in the lane's own slots (no barrier), for the
// same row at the next eta cell.  Per cell 3 + (1/CY + 1/CZ) interfaces
// instead of 4 + 1/CZ in k_stage_rows.  Interface states, fluxes and the
// accumulation order are those of k_stage: identical results.
// bsev, bsfirst, nseg: the band segments of this CY over all grids.
template <int MODE, int CM, int NC, int EK>
__global__ void __launch_bounds__(256) VH_OCC(VH_W_MX_BANDS)
k_stage_bands(Grids G, const int *__restrict__ bsev, const int *__restrict__ bsfirst, long nseg,
              DevEos eos_a, const double *__restrict__ Q, const double *__restrict__ Qh,
              const float *__restrict__ gv, const double *__restrict__ m, int m_full,
              const double *__restrict__ tp, double dt, double *__restrict__ out,
              int *__restrict__ status, unsigned long long *__restrict__ sig, int CY, int CZ) {
  constexpr int NV = NC > 4 ? 2 : 1;  // float4 slots of a carried flux
  extern __shared__ float4 fzs[];     // [CY][NV][blockDim.x]
  const DevEos eos = eos_of_kind<EK>(eos_a);
  const int lane = threadIdx.x & 63;
  const long w = xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long sgl = w % nseg, zc = w / nseg;  // band segment over all grids, eta chunk
  long seg = sgl;
  Geom g = G.g0;
  if (G.nev > 1) {
    const int gi = bsev[sgl];
    seg = sgl - bsfirst[gi];
    g = G.ev[gi];
  }
  const int z0 = (int)zc * CZ;
  if (z0 >= g.nz) return;  // whole wave
  const int z1 = min(g.nz, z0 + CZ);
  const unsigned nxu = (unsigned)g.nx, nbu = ((unsigned)g.ny + (unsigned)CY - 1u) / (unsigned)CY;
  const long f = seg * 63 + lane;
  const bool cv = f < (long)nbu * nxu;
  const unsigned fu = (unsigned)f;
  const int band = cv ? (int)(fu / nxu) : 0, ix = cv ? (int)(fu - (unsigned)band * nxu) : 0;
  const int y0 = band * CY, nrow = cv ? min(CY, g.ny - y0) : 0;
  const long N = g.P, sy = g.nx, sz = (long)g.nx * g.ny;
  const double tau = *tp;
  const double *Qs = MODE == 0 ? Q : Qh;
  const double itau = drcp(tau);  // as in axis_fluxes: identical interface states
  const double tauFactor = MODE == 0 ? tau + 0.25 * dt : tau + 0.5 * dt;
  const double dta = MODE == 0 ? 0.5 * dt : dt;
  const float prx = (float)(tauFactor * dta / g.dx), pry = (float)(tauFactor * dta / g.dy),
              prz = (float)(tauFactor * dta / g.dz);
  const float itf = (float)drcp(tauFactor), itauf = (float)itau;
  float4 *slot = fzs + threadIdx.x;  // row j, part v: slot[(j * NV + v) * blockDim.x]
  unsigned hzm = 0u;                  // bit j: row j carries an eta flux
  float spx = 0.f, spy = 0.f, spz = 0.f;
  for (int iz = z0; iz < z1; iz++) {
    float fyc[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool hyc = false;
#pragma unroll 1
    for (int j = 0; j < CY; j++) {  // uniform trip count: every lane shuffles
      const bool cx = j < nrow;
      const int iy = y0 + j;
      const long c = cx ? g.base + ix + iy * sy + iz * sz : 0;
      float fx[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      bool hx = false;
      if (cx && ix >= 1)
        hx = iface<1, CM, NC>(eos, Qs, Q, gv, m, m_full, N, ix >= 2 ? c - 2 : c - 1, c - 1, c,
                              ix + 1 < g.nx ? c + 1 : c, itau, itauf, itf, prx, fx, status, spx);
      float fxu[7];
#pragma unroll
      for (int k = 0; k < NC; k++) fxu[k] = __shfl_down(fx[k], 1, 64);
      const bool hxu = __shfl_down(hx ? 1 : 0, 1, 64) != 0 && ix + 1 < g.nx;
      float fyl[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      bool hyl = false;
      if (j == 0) {
        if (cx && iy >= 1)
          hyl = iface<2, CM, NC>(eos, Qs, Q, gv, m, m_full, N, iy >= 2 ? c - 2 * sy : c - sy,
                                 c - sy, c, iy + 1 < g.ny ? c + sy : c, itau, itauf, itf, pry,
                                 fyl, status, spy);
      } else {
#pragma unroll
        for (int k = 0; k < NC; k++) fyl[k] = fyc[k];
        hyl = hyc;
      }
      float fyu[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      bool hyu = false;
      if (cx && iy + 1 < g.ny)
        hyu = iface<2, CM, NC>(eos, Qs, Q, gv, m, m_full, N, iy >= 1 ? c - sy : c, c, c + sy,
                               iy + 2 < g.ny ? c + 2 * sy : c + sy, itau, itauf, itf, pry, fyu,
                               status, spy);
      float fzl[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      bool hzl = false;
      if (iz == z0) {
        if (cx && iz >= 1)
          hzl = iface<3, CM, NC>(eos, Qs, Q, gv, m, m_full, N, iz >= 2 ? c - 2 * sz : c - sz,
                                 c - sz, c, iz + 1 < g.nz ? c + sz : c, itau, itauf, itf, prz,
                                 fzl, status, spz);
      } else {
        hzl = (hzm >> j) & 1u;
        if (hzl) {
#pragma unroll
          for (int v = 0; v < NV; v++) {
            const float4 t = slot[(j * NV + v) * blockDim.x];
            fzl[4 * v] = t.x;
            fzl[4 * v + 1] = t.y;
            fzl[4 * v + 2] = t.z;
            fzl[4 * v + 3] = t.w;
          }
        }
      }
      float fzu[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      bool hzu = false;
      if (cx && iz + 1 < g.nz)
        hzu = iface<3, CM, NC>(eos, Qs, Q, gv, m, m_full, N, iz >= 1 ? c - sz : c, c, c + sz,
                               iz + 2 < g.nz ? c + 2 * sz : c + sz, itau, itauf, itf, prz, fzu,
                               status, spz);
      if (hzu) {
#pragma unroll
        for (int v = 0; v < NV; v++)
          slot[(j * NV + v) * blockDim.x] =
              make_float4(fzu[4 * v], fzu[4 * v + 1], fzu[4 * v + 2], fzu[4 * v + 3]);
        hzm |= 1u << j;
      } else {
        hzm &= ~(1u << j);
      }
      if (cx && lane < 63) {
        double acc[7];
#pragma unroll
        for (int k = 0; k < 7; k++) acc[k] = 0.;
        if (hx)
#pragma unroll
          for (int k = 0; k < NC; k++) acc[k] += (double)fx[k];
        if (hxu)
#pragma unroll
          for (int k = 0; k < NC; k++) acc[k] -= (double)fxu[k];
        if (hyl)
#pragma unroll
          for (int k = 0; k < NC; k++) acc[k] += (double)fyl[k];
        if (hyu)
#pragma unroll
          for (int k = 0; k < NC; k++) acc[k] -= (double)fyu[k];
        if (hzl)
#pragma unroll
          for (int k = 0; k < NC; k++) acc[k] += (double)fzl[k];
        if (hzu)
#pragma unroll
          for (int k = 0; k < NC; k++) acc[k] -= (double)fzu[k];
        finish_cell<MODE, CM, NC>(eos, Q, Qs, gv, N, c, tau, dt, acc, out);
      }
#pragma unroll
      for (int k = 0; k < NC; k++) fyc[k] = fyu[k];
      hyc = hyu;
    }
  }
  wave_speed_max(sig, (double)spx);
  wave_speed_max(sig + kSigGroup, (double)spy);
  wave_speed_max(sig + 2 * kSigGroup, (double)spz);
}

// ---------------------------------------------------------------- viscous
template <int CM, int NC, int EK>
__global__ void __launch_bounds__(256) VH_OCC(VH_W_MX_VPRIM)
k_vprim(long N, DevEos eos_a, const double *__restrict__ Qh, const float *__restrict__ gv,
        const double *__restrict__ tp, double dt, double *__restrict__ ph,
        double *__restrict__ pt) {
  const DevEos eos = eos_of_kind<EK>(eos_a);
  const double tau = *tp;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double q[7];
  Prim r;
  const double ith = 1. / (tau - 0.5 * dt), it = 1. / tau;
  double qs[NC];  // Qh is read once for both time levels
#pragma unroll
  for (int k = 0; k < NC; k++) qs[k] = ldu(Qh + k * N, c);
#pragma unroll
  for (int k = 0; k < 7; k++) q[k] = k < NC ? qs[k] * ith : 0.;
  c2p_mode<CM>(eos, q, r, guess_of<CM>(gv, c));
  stu(ph, c, r.e);
  stu(ph + N, c, r.p);
  if (NC == 7) {  // without charges n_b, n_q, n_s are zero and not stored
    stu(ph + 2 * N, c, r.nb);
    stu(ph + 3 * N, c, r.nq);
    stu(ph + 4 * N, c, r.ns);
  }
  stu(ph + 5 * N, c, r.vx);
  stu(ph + 6 * N, c, r.vy);
  stu(ph + 7 * N, c, r.vz);
#pragma unroll
  for (int k = 0; k < 7; k++) q[k] = k < NC ? qs[k] * it : 0.;
  c2p_mode<CM>(eos, q, r, guess_of<CM>(gv, c));
  stu(pt, c, r.e);
  stu(pt + N, c, r.vx);
  stu(pt + 2 * N, c, r.vy);
  stu(pt + 3 * N, c, r.vz);
}

// one Israel-Stewart update of the packed tensor (FP32), in the algebraically
// reduced form of src/hdo.cpp:790-809 (half step) and 843-862 (full step):
//   sum_l u_l dmu[l][k]               -> a_k  (4 terms instead of 16 per pair)
//   sum_kl P_kl(phi7 P_kl - tpp s_kl) -> one scalar S for all pairs
// X: the tensor the cross terms act on (pi for the half step, piH0 for the
// full step); Pm: the tensor of the phi7 term's first factor (pi in both).
__device__ __forceinline__ void is_cross(const float *X, const float *Pm, const float *sig,
                                         const float u[4], const float a[4], float du,
                                         float Xbulk, float dpp, float tpp, float lpP,
                                         float phi7, float fac, float out[10]) {
  const float gm[4] = {1.f, -1.f, -1.f, -1.f};
  float b[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) t += X[sym(i, k)] * gm[k] * a[k];
    b[i] = t;
  }
  float S = 0.f;
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int l = 0; l < 4; l++)
      S += X[sym(k, l)] * (phi7 * X[sym(k, l)] - tpp * sig[sym(k, l)]) * gm[k] * gm[l];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      const int ij = sym(i, j);
      float cij = 0.f, dij = 0.f, dji = 0.f;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        cij += Pm[sym(i, k)] * X[sym(j, k)] * gm[k];
        dij += X[sym(i, k)] * sig[sym(j, k)] * gm[k];
        dji += X[sym(j, k)] * sig[sym(i, k)] * gm[k];
      }
      const float Dij = (i == j ? gm[i] : 0.f) - u[i] * u[j];
      out[ij] = (-dpp * X[ij] * du + lpP * Xbulk * sig[ij] + phi7 * cij - 0.5f * tpp * (dij + dji) -
                 u[j] * b[i] - u[i] * b[j] - (1.f / 3.f) * Dij * S) *
                fac;
    }
}

// With VORT, also the cell's vorticity entries and op code (vh_vort.h), formed
// in FP64 from the same states as NSquant's velocity gradients.
template <int CM, int NC, bool VORT, int EK>
__global__ void __launch_bounds__(256) VH_OCC(VH_W_MX_VISC1)
k_visc1(Grids G, DevEos eos_a, DevTrc trc, const double *__restrict__ Qprev,
        const double *__restrict__ Qn, const float *__restrict__ gv,
        const double *__restrict__ ph,
        const double *__restrict__ pt, const double *__restrict__ m, int m_full,
        const float *__restrict__ pi, const float *__restrict__ Pi,
        const double *__restrict__ tp, double dt, float4 *__restrict__ rec,
        int *__restrict__ status, double *__restrict__ db, unsigned char *__restrict__ dop) {
  const DevEos eos = eos_of_kind<EK>(eos_a);
  const double tau = *tp;
  Geom g;
  long c;
  int ix, iy, iz;
  if (!work_cell(G, g, c, ix, iy, iz)) return;
  const long N = g.P;
  const double e = ldu(ph, c), nb = NC == 7 ? ldu(ph + 2 * N, c) : 0.;
  const double vx = ldu(ph + 5 * N, c), vy = ldu(ph + 6 * N, c), vz = ldu(ph + 7 * N, c);
  if (e <= 0.) {
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int q = 0; q < 6; q++) stn(rec + q * N, c, z);
    if (VORT) vort::store(db, dop, N, c, vort::K_EMPTY, nullptr);
    return;
  }
  const double gamma = drsqrt(1.0 - vx * vx - vy * vy - vz * vz);
  const double u[4] = {gamma, gamma * vx, gamma * vy, gamma * vz};
  // ---- NSquant in FP64 (src/hdo.cpp:381-652)
  double dmu[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) dmu[i][j] = 0.;
  bool vac = false;
  if (!m_full) {
    auto maxM = [&](int a, int b) {
      const long cc = g.cidx(ix + a, iy + b, iz);
      return smax(ldu(m, cc), smax(ldu(m + N, cc), ldu(m + 2 * N, cc)));
    };
    vac = maxM(1, 0) <= 0.9 || maxM(0, 1) <= 0.9 || maxM(-1, 0) <= 0.9 || maxM(0, -1) <= 0.9 ||
          maxM(1, 1) <= 0.9 || maxM(1, -1) <= 0.9 || maxM(-1, 1) <= 0.9 || maxM(-1, -1) <= 0.9;
  }
  const double VMIN = 1e-2, UDIFF = 3.0;
  // the reference's tests |0.5 (a + b) / u| > UDIFF as |0.5 (a + b)| > UDIFF |u|,
  // and its divisions by u^k as products with 1/u^k (used only where
  // |u^k| >= VMIN, or k = 0 with u^0 >= 1)
  const double ru[4] = {drcp(u[0]), drcp(u[1]), drcp(u[2]), drcp(u[3])};
  const double ihx = 1. / g.dx, ihy = 1. / g.dy, ihz = 1. / (g.dz * (tau + 0.5 * dt));
  double vv[16];  // vorticity entries formed by this cell (VORT)
  unsigned char vflags = 0;
  if (VORT && vac) vort::store(db, dop, N, c, vort::K_STALE, nullptr);
  if (!vac) {
    // The neighbours' gradients and the time derivative (two con2prim
    // solves): with the iterative solves the gradients go first, so that
    // their loads do not wait on the solves' loops (the learned relaxation
    // 3 % faster); the closed form's straight-line solves go first and
    // overlap the loads (1 % faster in that order).
    auto space = [&] {
#pragma unroll
      for (int d = 1; d <= 3; d++) {
        const long ca = g.cidx(ix + (d == 1), iy + (d == 2), iz + (d == 3));
        const long cb = g.cidx(ix - (d == 1), iy - (d == 2), iz - (d == 3));
        // the reference takes both temperatures before its e test (src/hdo.cpp:474-481)
        const double ea = ldu(pt, ca), eb = ldu(pt, cb);
        // loaded with e (ca, cb are valid cells), so that one round trip serves
        const double avx = ldu(pt + N, ca), avy = ldu(pt + 2 * N, ca), avz = ldu(pt + 3 * N, ca);
        const double bvx = ldu(pt + N, cb), bvy = ldu(pt + 2 * N, cb), bvz = ldu(pt + 3 * N, cb);
        const double TA = VORT ? eos.T(ea) : 0., TB = VORT ? eos.T(eb) : 0.;
        if (ea > 0. && eb > 0.) {
          const double gb = drsqrt(1.0 - bvx * bvx - bvy * bvy - bvz * bvz);
          const double ga = drsqrt(1.0 - avx * avx - avy * avy - avz * avz);
          const double B[4] = {gb, gb * bvx, gb * bvy, gb * bvz};
          const double A[4] = {ga, ga * avx, ga * avy, ga * avz};
          const double ih = d == 1 ? ihx : (d == 2 ? ihy : ihz);
          if (VORT) {
            const double den = d == 1 ? g.dx : (d == 2 ? g.dy : g.dz * (tau + 0.5 * dt));
            vort::row_space(vv + 4 * d, A, B, TA, TB, den);
            vflags |= d == 1 ? vort::F_R1 : (d == 2 ? vort::F_R2 : vort::F_R3);
          }
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const bool fast = fabs(0.5 * (A[k] + B[k])) > UDIFF * fabs(u[k]);
            const bool plain = k == 0 ? fast : (fabs(u[k]) < VMIN || fast);
            dmu[d][k] = plain ? 0.5 * (A[k] - B[k]) * ih
                              : (A[k] * A[k] - B[k] * B[k]) * (0.25 * ih) * ru[k];
          }
        }
      }
    };
    auto time = [&] {
      double q[7];
      Prim p0, p1;
      const double itn = 1. / (tau - dt), it1 = 1. / tau;
#pragma unroll
      for (int k = 0; k < 7; k++) q[k] = k < NC ? ldu(Qprev + k * N, c) * itn : 0.;
      c2p_mode<CM>(eos, q, p0, guess_of<CM>(gv, c));
#pragma unroll
      for (int k = 0; k < 7; k++) q[k] = k < NC ? ldu(Qn + k * N, c) * it1 : 0.;
      c2p_mode<CM>(eos, q, p1, guess_of<CM>(gv, c));
      const double idt = 1. / dt;
      const double ut0 = drsqrt(1.0 - p0.vx * p0.vx - p0.vy * p0.vy - p0.vz * p0.vz);
      const double ut1 = drsqrt(1.0 - p1.vx * p1.vx - p1.vy * p1.vy - p1.vz * p1.vz);
      const double a0[4] = {ut0, ut0 * p0.vx, ut0 * p0.vy, ut0 * p0.vz};
      const double a1[4] = {ut1, ut1 * p1.vx, ut1 * p1.vy, ut1 * p1.vz};
      if (VORT) vort::row_time(vv, a0, a1, eos.T(p0.e), eos.T(p1.e), e, p0.e, dt);
      if (e > 0. && p0.e > 0.) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
          // d(u^2)/2u form; plain difference where u^k is small or changes fast
          const bool fast = fabs(0.5 * (a1[k] + a0[k])) > UDIFF * fabs(a1[k]);
          const bool plain = k == 0 ? fast : (fabs(u[k]) < VMIN || fast);
          dmu[0][k] = plain ? (a1[k] - a0[k]) * idt
                            : (a1[k] * a1[k] - a0[k] * a0[k]) * (0.5 * idt) * ru[k];
        }
      }
    };
    if (CM == VH_C2P_CLOSED) {
      time();
      space();
    } else {
      space();
      time();
    }
    const double itm = 1. / (tau - 0.5 * dt);
    dmu[3][0] += u[3] * itm;
    dmu[3][3] += u[0] * itm;
    if (VORT) {  // Christoffel terms (src/hdo.cpp:608-611), then Cell::setDbeta
      const double TH = eos.T(e);
      if (TH > 0.) {
        const double a30 = u[3] / (TH * (tau - 0.5 * dt)), a33 = u[0] / (TH * (tau - 0.5 * dt));
        if (vflags & vort::F_R3) {
          vv[12] += a30;
          vv[15] += a33;
        } else {
          vv[12] = a30;
          vv[15] = a33;
        }
        vflags |= vort::F_ADD;
      }
      vort::store(db, dop, N, c, vort::K_FULL | vflags, vv);
    }
  }
  const double du = dmu[0][0] + dmu[1][1] + dmu[2][2] + dmu[3][3];
  // sigma^{ij} = 1/2 (D^{ik} dmu_kj + D^{jk} dmu_ki) - 1/3 D^{ij} theta, D = g - uu
  const double gmd[4] = {1., -1., -1., -1.};
  double Dm[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) Dm[i][j] = (i == j ? gmd[i] : 0.) - u[i] * u[j];
  // thermodynamics and transport coefficients in FP32 (they scale pi, which
  // is stored in FP32)
  const float ef = (float)e, nbf = (float)nb;
  const float p = eos.pf(ef), T = eos.Tf(ef), s = eos.sf(ef, T);
  const float etaS = trc.etaSf(ef, nbf, T, s, p);
  const float zetaS = trc.zetaSf(eos, ef, T, s, p);
  const float eta = etaS * s;
  const double kpi = (double)(eta * (float)(2.0 / 5.068));
  float sig[10], piNS[10];
  bool bad = false;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      double t = 0.;
#pragma unroll
      for (int k = 0; k < 4; k++) t += 0.5 * (Dm[i][k] * dmu[k][j] + Dm[j][k] * dmu[k][i]);
      t -= Dm[i][j] * du * (1. / 3.);
      const double pns = kpi * t;
      bad |= !(fabs(pns) <= 1.7976931348623157e308);
      piNS[sym(i, j)] = (float)pns;
      sig[sym(i, j)] = eta > 0. ? (float)t : 0.f;
    }
  if (bad) atomicOr(status, VH_ST_NAN_PI);
  const float PiNS = -zetaS * s * (float)du * (float)(1. / 5.068);
  // ---- relaxation times and second-order coefficients (src/trancoeff.h:28-37)
  float taupi, tauPi;
  trc.getTauf(eos, ef, nbf, T, s, p, etaS, zetaS, taupi, tauPi);
  const float hdt = (float)(0.5 * dt), fdt = (float)dt;
  float dpp = 4.f / 3.f, tpp = 10.f / 7.f, lpP = 6.f / 5.f;
  // phi7 = 9/(70 p) / tau_pi for every p != 0, negative p included; 0 for
  // p = 0, where the reference's quotient is inf (src/hdo.cpp:744-746).  A
  // table EoS whose spline rings next to a clamped region gives |p| down to
  // 1e-40: there the FP64 reference forms cross terms phi7 pi pi of 1e20 and
  // more, which its limiter then scales back to 0.1 T^{mu nu}; in FP32 the
  // full step's phi7 pi piH would overflow to inf (and the limiter's inf * 0
  // to NaN).  |phi7| is capped at 1e10, far above the 9/(70 p tau_pi) < 1e4
  // of any p > 1e-3 GeV/fm^3 and tau_pi > 0.01 fm, so that such cells stay
  // finite and keep the sign and the dominance of the term.
  float phi7 = 0.f;
  if (p != 0.f) phi7 = fminf(fmaxf((9.f / 70.f) / (p * taupi), -1e10f), 1e10f);
  if (taupi < hdt) dpp = tpp = lpP = phi7 = 0.f;
  float dPP = 2.f / 3.f, lPp = 1.6f * (1.f / 3.f - eos.cs2f(ef));
  if (tauPi < hdt) dPP = lPp = 0.f;
  const float ig = (float)drcp(gamma);
  const float uf[4] = {(float)u[0], (float)u[1], (float)u[2], (float)u[3]};
  float af[4];
#pragma unroll
  for (int k = 0; k < 4; k++)
    af[k] = (float)(u[0] * dmu[0][k] + u[1] * dmu[1][k] + u[2] * dmu[2][k] + u[3] * dmu[3][k]);
  const float duf = (float)du, vzf = (float)vz;
  float P[10], H[10], F[10], cr[10];
#pragma unroll
  for (int k = 0; k < 10; k++) P[k] = ldu(pi + k * N, c);
  const float Pc = ldu(Pi, c);
  const float gmf[4] = {1.f, -1.f, -1.f, -1.f};
  // half step
  const bool rel = taupi > hdt, relB = tauPi > hdt;
  const float rH = rel ? hdt * ig / taupi : 0.f, rHB = relB ? hdt * ig / tauPi : 0.f;
#pragma unroll
  for (int k = 0; k < 10; k++) H[k] = rel ? P[k] - (P[k] - piNS[k]) * rH : piNS[k];
  float PH = relB ? Pc - (Pc - PiNS) * rHB : PiNS;
  {
    const float w = vzf / (float)(tau - dt * 0.75) * hdt;
    H[sym(0, 0)] += -2.f * w * P[sym(0, 3)];
    H[sym(3, 3)] += -2.f * w * P[sym(0, 3)];
    H[sym(3, 0)] += -w * (P[sym(0, 0)] + P[sym(3, 3)]);
    H[sym(1, 0)] += -w * P[sym(1, 3)];
    H[sym(2, 0)] += -w * P[sym(2, 3)];
    H[sym(3, 1)] += -w * P[sym(0, 1)];
    H[sym(3, 2)] += -w * P[sym(0, 2)];
  }
  is_cross(P, P, sig, uf, af, duf, Pc, dpp, tpp, lpP, phi7, ig * hdt, cr);
#pragma unroll
  for (int k = 0; k < 10; k++) H[k] += cr[k];
  {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int l = 0; l < 4; l++) t += P[sym(k, l)] * sig[sym(k, l)] * gmf[k] * gmf[l];
    PH += (lPp * t - dPP * Pc * duf) * ig * hdt;
  }
  // full step
  const float rF = rel ? fdt * ig / taupi : 0.f, rFB = relB ? fdt * ig / tauPi : 0.f;
#pragma unroll
  for (int k = 0; k < 10; k++) F[k] = rel ? P[k] - (H[k] - piNS[k]) * rF : piNS[k];
  float PF = relB ? Pc - (PH - PiNS) * rFB : PiNS;
  {
    const float w = vzf / (float)(tau - dt * 0.5) * fdt;
    F[sym(0, 0)] += -2.f * w * H[sym(0, 3)];
    F[sym(3, 3)] += -2.f * w * H[sym(0, 3)];
    F[sym(3, 0)] += -w * (H[sym(0, 0)] + H[sym(3, 3)]);
    F[sym(1, 0)] += -w * H[sym(1, 3)];
    F[sym(2, 0)] += -w * H[sym(2, 3)];
    F[sym(3, 1)] += -w * H[sym(0, 1)];
    F[sym(3, 2)] += -w * H[sym(0, 2)];
  }
  is_cross(H, P, sig, uf, af, duf, PH, dpp, tpp, lpP, phi7, ig * fdt, cr);
#pragma unroll
  for (int k = 0; k < 10; k++) F[k] += cr[k];
  {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int l = 0; l < 4; l++) t += H[sym(k, l)] * sig[sym(k, l)] * gmf[k] * gmf[l];
    PF += (lPp * t - dPP * PH * duf) * ig * fdt;
  }
  // six planes of float4: (pi0 0-3) (pi0 4-7) (pi0 8-9, Pi0, 0) and the same for piH0
  stn(rec, c, make_float4(F[0], F[1], F[2], F[3]));
  stn(rec + N, c, make_float4(F[4], F[5], F[6], F[7]));
  stn(rec + 2 * N, c, make_float4(F[8], F[9], PF, 0.f));
  stn(rec + 3 * N, c, make_float4(H[0], H[1], H[2], H[3]));
  stn(rec + 4 * N, c, make_float4(H[4], H[5], H[6], H[7]));
  stn(rec + 5 * N, c, make_float4(H[8], H[9], PH, 0.f));
}

// ISformal loop #2 (src/hdo.cpp:866-962) in FP32 with FP32 pi storage, for
// cell c = (ix, iy, iz) of grid g.
__device__ __forceinline__ void visc2_cell(const Geom &g, long c, int ix, int iy, int iz,
                                           const double *__restrict__ ph,
                                           const float4 *__restrict__ rec,
                                           float *__restrict__ pi, float *__restrict__ piH,
                                           float *__restrict__ Pi, float *__restrict__ PiH,
                                           double *__restrict__ vcc, float *__restrict__ vf,
                                           double tau, double dt) {
  const long N = g.P;
  // The vacuum test (e <= 0: no advection, the old pi kept) comes after the
  // gathers, so that the cell's own loads and the eight neighbours' loads
  // form two round trips to memory, not three; a vacuum cell's gathers are
  // of clamped, valid cells and their results are not stored.
  const double vxd = ph[5 * N + c], vyd = ph[6 * N + c], vzd = ph[7 * N + c];
  // the upwind directions, the signs of xm, ym, zm below, taken from the
  // velocity itself (dt/dx, dt/dy, dt/(dz (tau - dt/2)) > 0; where xm
  // underflows to 0 in FP32 the extra corner has weight 0), so that the
  // gathers are issued before the weights' divisions and overlap them
  const int sx = vxd < 0. ? 1 : (vxd > 0. ? -1 : 0);
  const int sy = vyd < 0. ? 1 : (vyd > 0. ? -1 : 0);
  const int sz = vzd < 0. ? 1 : (vzd > 0. ? -1 : 0);
  float4 R[8][6];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int jx = j >> 2, jy = (j >> 1) & 1, jz = j & 1;
    const long c1 = g.cidx(ix + jx * sx, iy + jy * sy, iz + jz * sz);
#pragma unroll
    for (int q = 0; q < 6; q++) R[j][q] = rec[q * N + c1];
  }
  __builtin_amdgcn_sched_barrier(0);
  const float e = (float)ph[c], p = (float)ph[N + c];
  const float vx = (float)vxd, vy = (float)vyd, vz = (float)vzd;
  const double v2d = vxd * vxd + vyd * vyd + vzd * vzd;
  const float v2 = (float)v2d;
  const float g2 = (float)(1. / (1. - v2d)), gamma = fsqrt(g2);
  const float u[4] = {gamma, gamma * vx, gamma * vy, gamma * vz};
  const float tf = (float)tau, dtf = (float)dt;
  float fl[4];
  {
    const float Po = Pi[c];
#pragma unroll
    for (int i = 0; i < 4; i++) fl[i] = (tf - dtf) * (pi[sym(0, i) * N + c] + Po * u[0] * u[i]);
    fl[0] -= (tf - dtf) * Po;
  }
  const float xm = (float)(-vxd * dt / g.dx);
  const float ym = (float)(-vyd * dt / g.dy);
  const float zm = (float)(-vzd * dt / g.dz / (tau - 0.5 * dt));
  const float wx[2] = {1.f - fabsf(xm), fabsf(xm)}, wy[2] = {1.f - fabsf(ym), fabsf(ym)},
              wz[2] = {1.f - fabsf(zm), fabsf(zm)};
  const float wxH[2] = {1.f - 0.5f * fabsf(xm), 0.5f * fabsf(xm)},
              wyH[2] = {1.f - 0.5f * fabsf(ym), 0.5f * fabsf(ym)},
              wzH[2] = {1.f - 0.5f * fabsf(zm), 0.5f * fabsf(zm)};
  float a[10], aH[10], A = 0.f, AH = 0.f;
#pragma unroll
  for (int k = 0; k < 10; k++) a[k] = aH[k] = 0.f;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int jx = j >> 2, jy = (j >> 1) & 1, jz = j & 1;
    const float w = wx[jx] * wy[jy] * wz[jz], wH = wxH[jx] * wyH[jy] * wzH[jz];
    const float4 r0 = R[j][0], r1 = R[j][1], r2 = R[j][2], r3 = R[j][3], r4 = R[j][4],
                 r5 = R[j][5];
    a[0] += w * r0.x; a[1] += w * r0.y; a[2] += w * r0.z; a[3] += w * r0.w;
    a[4] += w * r1.x; a[5] += w * r1.y; a[6] += w * r1.z; a[7] += w * r1.w;
    a[8] += w * r2.x; a[9] += w * r2.y; A += w * r2.z;
    aH[0] += wH * r3.x; aH[1] += wH * r3.y; aH[2] += wH * r3.z; aH[3] += wH * r3.w;
    aH[4] += wH * r4.x; aH[5] += wH * r4.y; aH[6] += wH * r4.z; aH[7] += wH * r4.w;
    aH[8] += wH * r5.x; aH[9] += wH * r5.y; AH += wH * r5.z;
  }
  const bool live = e > 0.f;
  const float maxT0 = fmax_((e + p) * g2 - p, (e + p) * v2 * g2 + p);
  float maxpi = 0.f;
#pragma unroll
  for (int k = 0; k < 10; k++) maxpi = fmaxf(maxpi, fabsf(a[k]));
  bool rescaled = false;
  if (maxT0 < maxpi) {
    const float f = 0.1f * maxT0 / maxpi;
#pragma unroll
    for (int k = 0; k < 10; k++) {
      a[k] *= f;
      aH[k] *= f;
    }
    rescaled = true;
  }
  if (fabsf(A) > p) {
    if (A != 0.f) A = copysignf(p, A);
    if (AH != 0.f) AH = copysignf(p, AH);
    rescaled = true;
  }
  if (live) {
    stn(vcc, c, rescaled ? (double)(maxT0 / maxpi) : 1.);
#pragma unroll
    for (int k = 0; k < 10; k++) {
      stn(pi + k * N, c, a[k]);
      stn(piH + k * N, c, aH[k]);
    }
    stn(Pi, c, A);
    stn(PiH, c, AH);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) fl[i] += -tf * (a[sym(0, i)] + A * u[0] * u[i]);
  fl[0] += tf * A;
#pragma unroll
  for (int i = 0; i < 4; i++) stn(vf + i * N, c, live ? fl[i] : 0.f);
}

// one thread per cell
__global__ void __launch_bounds__(256) VH_OCC(VH_W_MX_VISC2)
k_visc2(Grids G, const double *__restrict__ ph, const float4 *__restrict__ rec,
        float *__restrict__ pi, float *__restrict__ piH,
        float *__restrict__ Pi, float *__restrict__ PiH, double *__restrict__ vcc,
        float *__restrict__ vf, const double *__restrict__ tp, double dt) {
  Geom g;
  long c;
  int ix, iy, iz;
  if (!work_cell(G, g, c, ix, iy, iz)) return;
  visc2_cell(g, c, ix, iy, iz, ph, rec, pi, piH, Pi, PiH, vcc, vf, *tp, dt);
}

// viscous interface flux in FP32 (Hydro::visc_flux, src/hdo.cpp:966-1014)
// between cells l and r along axis D
template <int D>
__device__ __forceinline__ void visc_face(long N, long l, long r, const double *__restrict__ ph,
                                          const float *__restrict__ piH,
                                          const float *__restrict__ PiH, float pre, float F[4]) {
  const double dvx = 0.5 * (ph[5 * N + l] + ph[5 * N + r]);
  const double dvy = 0.5 * (ph[6 * N + l] + ph[6 * N + r]);
  const double dvz = 0.5 * (ph[7 * N + l] + ph[7 * N + r]);
  const float vx = (float)dvx, vy = (float)dvy, vz = (float)dvz;
  const float gamma = (float)(1. / sqrt(1. - (dvx * dvx + dvy * dvy + dvz * dvz)));
  const float u[4] = {gamma, gamma * vx, gamma * vy, gamma * vz};
  const float sP = 0.5f * (PiH[l] + PiH[r]);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    float f = 0.5f * (piH[sym(mu, D) * N + l] + piH[sym(mu, D) * N + r]);
    if (mu == D) f += sP;  // -sP * g^{DD} with g^{DD} = -1
    f += sP * u[mu] * u[D];
    F[mu] = f * pre;
  }
}

// The six faces are formed without branches: a face outside the grid takes
// the cell itself as its neighbour (valid memory) and is masked off, like a
// face whose two cells both have m < 1 (the reference's skip), so that all
// the loads are issued before the flux arithmetic instead of one face (and
// one m test) at a time.
__global__ void __launch_bounds__(256) VH_OCC(VH_W_MX_VISC3)
k_visc3(Grids G, const double *__restrict__ ph, const float *__restrict__ piH,
        const float *__restrict__ PiH, const float *__restrict__ vf,
        const double *__restrict__ m, int m_full, const double *__restrict__ tp, double dt,
        double *__restrict__ Q) {
  const double tau = *tp;
  Geom g;
  long c;
  int ix, iy, iz;
  if (!work_cell(G, g, c, ix, iy, iz)) return;
  const long N = g.P;
  float acc[4];
#pragma unroll
  for (int k = 0; k < 4; k++) acc[k] = vf[k * N + c];
  const double pre0 = (tau - 0.5 * dt) * dt;
  const float px = (float)(pre0 / g.dx), py = (float)(pre0 / g.dy),
              pz = (float)(pre0 / (g.dz * (tau + 0.5 * dt)));
  const long sy = g.nx, sz = (long)g.nx * g.ny;
  // face k: axis 1 + k / 2, lower (even k: neighbour, c) or upper (c, neighbour)
  const bool in[6] = {ix > 0, ix < g.nx - 1, iy > 0, iy < g.ny - 1, iz > 0, iz < g.nz - 1};
  const long step[6] = {-1, 1, -sy, sy, -sz, sz};
  bool use[6];
  long nbr[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    nbr[k] = in[k] ? c + step[k] : c;
    use[k] = in[k];
  }
  if (!m_full) {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const double *md = m + (k / 2) * N;
      const double ma = md[nbr[k]], mb = md[c];
      use[k] = use[k] & !((ma < 1.) & (mb < 1.));
    }
  }
  float F[6][4];
  visc_face<1>(N, nbr[0], c, ph, piH, PiH, px, F[0]);
  visc_face<1>(N, c, nbr[1], ph, piH, PiH, px, F[1]);
  visc_face<2>(N, nbr[2], c, ph, piH, PiH, py, F[2]);
  visc_face<2>(N, c, nbr[3], ph, piH, PiH, py, F[3]);
  visc_face<3>(N, nbr[4], c, ph, piH, PiH, pz, F[4]);
  visc_face<3>(N, c, nbr[5], ph, piH, PiH, pz, F[5]);
#pragma unroll
  for (int k = 0; k < 6; k++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float f = use[k] ? F[k][q] : 0.f;
      if (k & 1)
        acc[q] -= f;
      else
        acc[q] += f;
    }
  if (ph[c] > 0.) {
    const double dvx = ph[5 * N + c], dvy = ph[6 * N + c], dvz = ph[7 * N + c];
    const float u0 = (float)(1. / sqrt(1. - dvx * dvx - dvy * dvy - dvz * dvz)),
                u3 = u0 * (float)dvz;
    const float dtf = (float)dt;
    acc[0] += (-piH[sym(3, 3) * N + c] - PiH[c] * (1.f + u3 * u3)) * dtf;
    acc[3] += -(piH[sym(0, 3) * N + c] + PiH[c] * u0 * u3) * dtf;
  }
  const double q0 = Q[c];
  const double a0 = (double)acc[0];
  double fac = 1.;
  if (fabs(a0) > 0.5 * q0) fac = a0 != 0. ? fabs(0.5 * q0 / a0) : 0.;
#pragma unroll
  for (int k = 0; k < 4; k++) Q[k * N + c] += fac * (double)acc[k];
}

// ghost fill of Q and the FP32 pi, Pi (src/fld.cpp:275-382)
template <int AXIS>
__global__ void __launch_bounds__(256)
k_ghost_full(Geom g, double *__restrict__ Q, float *__restrict__ pi, float *__restrict__ Pi) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long nline;
  if (AXIS == 3) nline = (long)g.nx * g.ny;
  else if (AXIS == 2) nline = (long)g.nx * g.nz;
  else nline = (long)g.ny * g.nz;
  if (t >= nline) return;
  const long N = g.P;
  const int n = AXIS == 3 ? g.nz : (AXIS == 2 ? g.ny : g.nx);
  int a, b;
  if (AXIS == 1) { a = (int)(t % g.ny); b = (int)(t / g.ny); }
  else { a = (int)(t % g.nx); b = (int)(t / g.nx); }
  auto cell = [&](int i) -> long {
    if (AXIS == 3) return g.cidx(a, b, i);
    if (AXIS == 2) return g.cidx(a, i, b);
    return g.cidx(i, a, b);
  };
  const int src[2] = {2, n - 3}, d1[2] = {1, n - 2}, d0[2] = {0, n - 1};
  for (int side = 0; side < 2; side++) {
    const long cs = cell(src[side]), c1 = cell(d1[side]), c0 = cell(d0[side]);
    double v[7];
#pragma unroll
    for (int k = 0; k < 7; k++) v[k] = Q[k * N + cs];
    const bool z = v[0] < 0.;
#pragma unroll
    for (int k = 0; k < 7; k++) {
      Q[k * N + c1] = z ? 0. : v[k];
      Q[k * N + c0] = z ? 0. : v[k];
    }
    float pv[10];
#pragma unroll
    for (int k = 0; k < 10; k++) pv[k] = pi[k * N + cs];
    const float P = Pi[cs];
#pragma unroll
    for (int k = 0; k < 10; k++) {
      pi[k * N + c0] = pv[k];
      pi[k * N + c1] = pv[k];
    }
    Pi[c0] = P;
    Pi[c1] = P;
  }
}

// Single-pass ghost fill over a set of grids (each with every extent 1 or
// >= 5): a ghost cell takes the value of the cell whose ghost coordinates
// are mapped to the nearest interior index (0,1 -> 2 and n-2,n-1 -> n-3),
// the composition of the reference's z, y, x passes (src/fld.cpp:211-382;
// setQ's Q[T_] < 0 zeroing is idempotent).  An axis with n = 1 makes every
// cell a ghost of itself, which only applies the zeroing: k_ghost_zero_b
// does it first for such grids, so k_ghost_b never writes a cell another
// thread reads.  FULL = 0: Q and Qh; FULL = 1: Q, pi, Pi (FP32).
__device__ __forceinline__ int ghost_src(int i, int n) {
  return n == 1 ? 0 : (i < 2 ? 2 : (i > n - 3 ? n - 3 : i));
}

template <int FULL>
__global__ void __launch_bounds__(256)
k_ghost_zero_b(Grids G, double *__restrict__ Q, double *__restrict__ Qh) {
  long cl;
  const Geom g = block_grid(G, blockIdx.x, cl);
  if (cl >= g.N || (g.nx > 1 && g.ny > 1 && g.nz > 1)) return;
  const long c = g.base + cl, N = g.P;
  if (Q[c] < 0.)
    for (int k = 0; k < 7; k++) Q[k * N + c] = 0.;
  if (!FULL && Qh[c] < 0.)
    for (int k = 0; k < 7; k++) Qh[k * N + c] = 0.;
}

template <int FULL>
__global__ void __launch_bounds__(256)
k_ghost_b(Grids G, double *__restrict__ Q, double *__restrict__ Qh, float *__restrict__ pi,
          float *__restrict__ Pi) {
  long cl;
  const Geom g = block_grid(G, blockIdx.x, cl);
  if (cl >= g.N) return;
  int ix, iy, iz;
  coords_of(g, cl, ix, iy, iz);
  const int sx = ghost_src(ix, g.nx), sy = ghost_src(iy, g.ny), sz = ghost_src(iz, g.nz);
  if (sx == ix && sy == iy && sz == iz) return;
  const long N = g.P, c = g.base + cl, src = g.idx(sx, sy, sz);
  double v[7];
#pragma unroll
  for (int k = 0; k < 7; k++) v[k] = Q[k * N + src];
  const bool z = v[0] < 0.;
#pragma unroll
  for (int k = 0; k < 7; k++) Q[k * N + c] = z ? 0. : v[k];
  if (!FULL) {
#pragma unroll
    for (int k = 0; k < 7; k++) v[k] = Qh[k * N + src];
    const bool zh = v[0] < 0.;
#pragma unroll
    for (int k = 0; k < 7; k++) Qh[k * N + c] = zh ? 0. : v[k];
  } else {
#pragma unroll
    for (int k = 0; k < 10; k++) pi[k * N + c] = pi[k * N + src];
    Pi[c] = Pi[src];
  }
}

// k_ghost_b of a single grid over its ghost cells only (every extent >= 5;
// ghost_cell)
template <int FULL, int NC>
__global__ void __launch_bounds__(256)
k_ghost_list(Geom g, long n, double *__restrict__ Q, double *__restrict__ Qh,
             float *__restrict__ pi, float *__restrict__ Pi) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int ix, iy, iz;
  ghost_cell(g.nx, g.ny, g.nz, t, ix, iy, iz);
  const long N = g.P, c = g.idx(ix, iy, iz);
  const long src = g.idx(ghost_src(ix, g.nx), ghost_src(iy, g.ny), ghost_src(iz, g.nz));
  double v[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) v[k] = Q[k * N + src];
  const bool z = v[0] < 0.;
#pragma unroll
  for (int k = 0; k < NC; k++) Q[k * N + c] = z ? 0. : v[k];
  if (!FULL) {
#pragma unroll
    for (int k = 0; k < NC; k++) v[k] = Qh[k * N + src];
    const bool zh = v[0] < 0.;
#pragma unroll
    for (int k = 0; k < NC; k++) Qh[k * N + c] = zh ? 0. : v[k];
  } else {
#pragma unroll
    for (int k = 0; k < 10; k++) pi[k * N + c] = pi[k * N + src];
    Pi[c] = Pi[src];
  }
}

// ---------------------------------------------------------- launchers
static inline unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }


void launch_stage(int mode, int cm, int nc, const Grids &G, const DevEos &eos, const double *Q,
                  const double *Qh, const float *gv, const double *m, int m_full,
                  const double *tp, double dt, double *out, int *status,
                  unsigned long long *sig, hipStream_t st) {
  const dim3 gr((unsigned)G.nblk), b(256);
  const int ek = eos_kind_class(eos);
#define VH_SK(MM, CC, NN, EE)                                                                  \
  if (mode == MM && cm == CC && nc == NN && ek == EE)                                          \
    hipLaunchKernelGGL((k_stage<MM, CC, NN, EE>), gr, b, 0, st, G, eos, Q, Qh, gv, m, m_full,   \
                       tp, dt, out, status, sig);
#define VH_S(MM, CC)                                                                           \
  VH_SK(MM, CC, 7, VH_EOS_IDEAL) VH_SK(MM, CC, 4, VH_EOS_IDEAL) VH_SK(MM, CC, 7, VH_EOS_TABLE) \
  VH_SK(MM, CC, 4, VH_EOS_TABLE)
  VH_S(0, VH_C2P_REFERENCE) VH_S(1, VH_C2P_REFERENCE) VH_S(0, VH_C2P_CLOSED)
  VH_S(1, VH_C2P_CLOSED) VH_S(0, VH_C2P_LEARNED) VH_S(1, VH_C2P_LEARNED)
  VH_S(0, VH_C2P_NEWTON) VH_S(1, VH_C2P_NEWTON)
#undef VH_S
#undef VH_SK
}

Tile make_rows(const Grids &G, int CZ) {
  Tile t;
  t.kind = STAGE_ROWS;
  t.CZ = CZ < 1 ? 1 : (CZ > G.maxnz ? G.maxnz : CZ);
  t.CY = 1;
  return t;
}

Tile make_bands(const Grids &G, int CY, int CZ) {
  Tile t;
  t.kind = STAGE_BANDS;
  t.CZ = CZ < 1 ? 1 : (CZ > G.maxnz ? G.maxnz : CZ);
  t.CY = CY;
  return t;
}

Tile make_cells() {
  Tile t;
  t.kind = STAGE_CELLS;
  t.CZ = 0;
  t.CY = 0;
  return t;
}

// waves of a row or band launch
static long tile_waves(const Grids &G, const Tile &t) {
  const long nseg = t.kind == STAGE_BANDS ? G.nbseg[band_table(t.CY)] : G.nseg;
  return nseg * ((G.maxnz + t.CZ - 1) / t.CZ);
}

std::vector<Tile> tile_candidates(const Grids &G) {
  std::vector<Tile> out;
  out.push_back(make_cells());
  // rows: eta chunks of 1..32 cells (distinct after clamping to the largest
  // nz), keeping at least 4096 waves (4 per SIMD of the chip) in flight
  const int czs[] = {1, 2, 4, 8, 16, 32};
  for (int cz : czs) {
    const Tile t = make_rows(G, cz);
    if (t.CZ > 1 && tile_waves(G, t) < 4096) continue;
    bool dup = false;
    for (auto &o : out) dup |= o.kind == STAGE_ROWS && o.CZ == t.CZ;
    if (!dup) out.push_back(t);
  }
  // bands: fewer, longer waves (each updates CY cells per eta cell), so at
  // least 2048 of them
  const int cys[] = {2, 4, 8};
  for (int cy : cys)
    for (int cz : czs) {
      const Tile t = make_bands(G, cy, cz);
      if ((t.CZ > 1 || cy > 2) && tile_waves(G, t) < 2048) continue;
      bool dup = false;
      for (auto &o : out) dup |= o.kind == STAGE_BANDS && o.CZ == t.CZ && o.CY == t.CY;
      if (!dup) out.push_back(t);
    }
  return out;
}

void launch_stage_tile(int mode, int cm, int nc, const Grids &G, const DevEos &eos, const double *Q,
                       const double *Qh, const float *gv, const double *m, int m_full,
                       const double *tp, double dt, double *out, int *status,
                       unsigned long long *sig, const Tile &tl, hipStream_t st) {
  if (tl.kind == STAGE_CELLS) {
    launch_stage(mode, cm, nc, G, eos, Q, Qh, gv, m, m_full, tp, dt, out, status, sig, st);
    return;
  }
  const long waves = tile_waves(G, tl);
  const dim3 gr((unsigned)((waves + 3) / 4)), b(256);
  const int ek = eos_kind_class(eos);
  if (tl.kind == STAGE_BANDS) {
    const int h = band_table(tl.CY);
    const int *bsev = G.bsev[h], *bsfirst = G.bsfirst[h];
    const long nseg = G.nbseg[h];
    const size_t lds = (size_t)tl.CY * (nc > 4 ? 2 : 1) * 256 * sizeof(float4);
#define VH_BK(MM, CC, NN, EE)                                                                  \
  if (mode == MM && cm == CC && nc == NN && ek == EE)                                          \
    hipLaunchKernelGGL((k_stage_bands<MM, CC, NN, EE>), gr, b, lds, st, G, bsev, bsfirst, nseg, \
                       eos, Q, Qh, gv, m, m_full, tp, dt, out, status, sig, tl.CY, tl.CZ);
#define VH_B(MM, CC)                                                                           \
  VH_BK(MM, CC, 7, VH_EOS_IDEAL) VH_BK(MM, CC, 4, VH_EOS_IDEAL) VH_BK(MM, CC, 7, VH_EOS_TABLE) \
  VH_BK(MM, CC, 4, VH_EOS_TABLE)
    VH_B(0, VH_C2P_REFERENCE) VH_B(1, VH_C2P_REFERENCE) VH_B(0, VH_C2P_CLOSED)
    VH_B(1, VH_C2P_CLOSED) VH_B(0, VH_C2P_LEARNED) VH_B(1, VH_C2P_LEARNED)
    VH_B(0, VH_C2P_NEWTON) VH_B(1, VH_C2P_NEWTON)
#undef VH_B
#undef VH_BK
    return;
  }
#define VH_SK(MM, CC, NN, EE)                                                                  \
  if (mode == MM && cm == CC && nc == NN && ek == EE)                                          \
    hipLaunchKernelGGL((k_stage_rows<MM, CC, NN, EE>), gr, b, 0, st, G, eos, Q, Qh, gv, m,      \
                       m_full, tp, dt, out, status, sig, tl.CZ);
#define VH_S(MM, CC)                                                                           \
  VH_SK(MM, CC, 7, VH_EOS_IDEAL) VH_SK(MM, CC, 4, VH_EOS_IDEAL) VH_SK(MM, CC, 7, VH_EOS_TABLE) \
  VH_SK(MM, CC, 4, VH_EOS_TABLE)
  VH_S(0, VH_C2P_REFERENCE) VH_S(1, VH_C2P_REFERENCE) VH_S(0, VH_C2P_CLOSED)
  VH_S(1, VH_C2P_CLOSED) VH_S(0, VH_C2P_LEARNED) VH_S(1, VH_C2P_LEARNED)
  VH_S(0, VH_C2P_NEWTON) VH_S(1, VH_C2P_NEWTON)
#undef VH_S
#undef VH_SK
}

void launch_vprim(int cm, int nc, long N, const DevEos &eos, const double *Qh, const float *gv,
                  const double *tp, double dt, double *ph, double *pt, hipStream_t st) {
  const dim3 gr(blocks(N)), b(256);
  const int ek = eos_kind_class(eos);
#define VH_VK(CC, NN, EE)                                                                      \
  if (cm == CC && nc == NN && ek == EE)                                                        \
    hipLaunchKernelGGL((k_vprim<CC, NN, EE>), gr, b, 0, st, N, eos, Qh, gv, tp, dt, ph, pt);
#define VH_V(CC)                                                                               \
  VH_VK(CC, 7, VH_EOS_IDEAL) VH_VK(CC, 4, VH_EOS_IDEAL) VH_VK(CC, 7, VH_EOS_TABLE)            \
  VH_VK(CC, 4, VH_EOS_TABLE)
  VH_V(VH_C2P_REFERENCE) VH_V(VH_C2P_CLOSED) VH_V(VH_C2P_LEARNED) VH_V(VH_C2P_NEWTON)
#undef VH_V
#undef VH_VK
}

void launch_visc1(int cm, int nc, const Grids &G, const DevEos &eos, const DevTrc &trc,
                  const double *Qprev, const double *Qn, const float *gv, const double *ph,
                  const double *pt, const double *m, int m_full, const float *pi,
                  const float *Pi, const double *tp, double dt, float4 *rec, int *status,
                  double *db, unsigned char *dop, hipStream_t st) {
  const dim3 gr((unsigned)G.nblk), b(256);
  const int ek = eos_kind_class(eos);
#define VH_VV(CC, NN, VV, EE)                                                                  \
  if (cm == CC && nc == NN && (db != nullptr) == VV && ek == EE)                               \
    hipLaunchKernelGGL((k_visc1<CC, NN, VV, EE>), gr, b, 0, st, G, eos, trc, Qprev, Qn, gv, ph, \
                       pt, m, m_full, pi, Pi, tp, dt, rec, status, db, dop);
#define VH_VE(CC, EE)                                                                          \
  VH_VV(CC, 7, false, EE) VH_VV(CC, 4, false, EE) VH_VV(CC, 7, true, EE) VH_VV(CC, 4, true, EE)
#define VH_V(CC) VH_VE(CC, VH_EOS_IDEAL) VH_VE(CC, VH_EOS_TABLE)
  VH_V(VH_C2P_REFERENCE) VH_V(VH_C2P_CLOSED) VH_V(VH_C2P_LEARNED) VH_V(VH_C2P_NEWTON)
#undef VH_V
#undef VH_VE
#undef VH_VV
}

void launch_visc2(const Grids &G, const double *ph, const float4 *rec, float *pi, float *piH,
                  float *Pi, float *PiH, double *vcc, float *vf, const double *tp, double dt,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_visc2, dim3((unsigned)G.nblk), dim3(256), 0, st, G, ph, rec, pi, piH, Pi,
                     PiH, vcc, vf, tp, dt);
}

void launch_visc3(const Grids &G, const double *ph, const float *piH, const float *PiH,
                  const float *vf, const double *m, int m_full, const double *tp, double dt,
                  double *Q, hipStream_t st) {
  hipLaunchKernelGGL(k_visc3, dim3((unsigned)G.nblk), dim3(256), 0, st, G, ph, piH, PiH, vf, m,
                     m_full, tp, dt, Q);
}

void launch_ghost_cells(int full, int nc, const Grids &G, bool degenerate, double *Q, double *Qh,
                        float *pi, float *Pi, hipStream_t st) {
  const dim3 gr((unsigned)G.nblk), b(256);
  const Geom &g = G.g0;
  if (G.nev == 1 && g.nx >= 5 && g.ny >= 5 && g.nz >= 5) {  // over the ghost cells only
    const long n = ghost_count(g.nx, g.ny, g.nz);
    const dim3 gl(blocks(n));
    if (full && nc == 4) hipLaunchKernelGGL((k_ghost_list<1, 4>), gl, b, 0, st, g, n, Q, Qh, pi, Pi);
    else if (full) hipLaunchKernelGGL((k_ghost_list<1, 7>), gl, b, 0, st, g, n, Q, Qh, pi, Pi);
    else if (nc == 4) hipLaunchKernelGGL((k_ghost_list<0, 4>), gl, b, 0, st, g, n, Q, Qh, pi, Pi);
    else hipLaunchKernelGGL((k_ghost_list<0, 7>), gl, b, 0, st, g, n, Q, Qh, pi, Pi);
    return;
  }
  if (degenerate) {
    if (full) hipLaunchKernelGGL(k_ghost_zero_b<1>, gr, b, 0, st, G, Q, Qh);
    else hipLaunchKernelGGL(k_ghost_zero_b<0>, gr, b, 0, st, G, Q, Qh);
  }
  if (full) hipLaunchKernelGGL(k_ghost_b<1>, gr, b, 0, st, G, Q, Qh, pi, Pi);
  else hipLaunchKernelGGL(k_ghost_b<0>, gr, b, 0, st, G, Q, Qh, pi, Pi);
}

void launch_ghost_full(int axis, const Geom &g, double *Q, float *pi, float *Pi,
                       hipStream_t st) {
  const long n = axis == 3 ? (long)g.nx * g.ny : (axis == 2 ? (long)g.nx * g.nz : (long)g.ny * g.nz);
  if (axis == 3)
    hipLaunchKernelGGL(k_ghost_full<3>, dim3(blocks(n)), dim3(256), 0, st, g, Q, pi, Pi);
  else if (axis == 2)
    hipLaunchKernelGGL(k_ghost_full<2>, dim3(blocks(n)), dim3(256), 0, st, g, Q, pi, Pi);
  else
    hipLaunchKernelGGL(k_ghost_full<1>, dim3(blocks(n)), dim3(256), 0, st, g, Q, pi, Pi);
}

}  // namespace mx
}  // namespace vh
