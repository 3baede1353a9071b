// FP64 faithful kernels of the vHLLE hydro step on gfx950.
//
// Layout in HBM: structure-of-arrays, component-major, a[k*N + cell] with
// cell = ix + nx*(iy + ny*iz) (x fastest), so that every kernel's loads of one
// component by consecutive lanes are consecutive doubles.  Q is double
// buffered: the corrector reads Q[cur] and writes Q[cur^1], so the previous
// Q (Cell::Qprev, src/cll.h:46) is just the other buffer and saveQprev
// (src/hdo.cpp:1028) costs nothing.
//
// The reference scatters each interface flux into its two cells
// (src/hdo.cpp:301-304).  Here every interface is evaluated once by one thread
// and stored; each cell then gathers its interfaces in the reference's
// accumulation order (+x-, -x+, +y-, -y+, +z-, -z+, then the source), which
// reproduces the reference's sums bit for bit.
#include <algorithm>

#include "vh_kernels.h"
#include "vh_vort.h"

namespace vh {

// A store of a result that the kernel does not read back and that the next
// kernel reads after it has left the L2 (the arrays are hundreds of MB):
// nontemporal, so that it does not evict the lines the kernel's own gathers
// reuse (the interface fluxes, the stage update, the relaxed tensors and the
// advected ones; the FP64 step 2-3 % faster).
template <class T>
__device__ __forceinline__ void st_nt(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}

// ------------------------------------------------------------- helpers
__device__ inline void decompose(const Geom &g, long c, int &ix, int &iy, int &iz) {
  ix = (int)(c % g.nx);
  const long r = c / g.nx;
  iy = (int)(r % g.ny);
  iz = (int)(r / g.ny);
}

template <int D>
__device__ inline int coord(int ix, int iy, int iz) {
  return D == 1 ? ix : (D == 2 ? iy : iz);
}
template <int D>
__device__ inline int extent(const Geom &g) {
  return D == 1 ? g.nx : (D == 2 ? g.ny : g.nz);
}
template <int D>
__device__ inline long stride(const Geom &g) {
  return D == 1 ? 1L : (D == 2 ? (long)g.nx : (long)g.nx * g.ny);
}

// MUSCL face state of a cell (Cell::getPrimVar{Left,Right}[H],
// src/cll.cpp:106-204): half slopes limited by minmod, state (q -+ dq)/tau,
// then con2prim.  cm/c0/cp: the cell and its clamped neighbours.
template <int SIDE, bool CH>
__device__ __forceinline__ void face_state(const DevEos &eos, const double *__restrict__ Qs, long N,
                                  long cm, long c0, long cp, const Divider &dtau, Prim &r) {
  double q[7];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    const double a = Qs[k * N + c0];
    const double dq = minmod((Qs[k * N + cp] - a) / 2., (a - Qs[k * N + cm]) / 2.);
    q[k] = SIDE < 0 ? dtau(a - dq) : dtau(a + dq);  // (a -+ dq) / tau
  }
  c2p_ref<CH>(eos, q, r);
}

// ------------------------------------------------------- interface flux
// Hydro::hlle_flux (src/hdo.cpp:77-305) for the interface between cell cl
// and its +1 neighbour cr along axis D, in two parts.
//
// flux_tail: everything after the face states are final (src/hdo.cpp:142-
// 304): the partial-vacuum skip, the HLLE wave speeds from El, Er, the flux.
// Writes F (7 x N, slot cl) and returns the interface's signal speed
// max(-b_l, b_r) (in units of dx/dtau, deta/dtau along eta); *ok = 1 when a
// flux is formed, 0 when the reference returns without one.
template <int D, int MODE>
__device__ __forceinline__ double flux_tail(const Geom &g, const double *__restrict__ m,
                                            int m_full, Prim &L, Prim &R, double El, double Er,
                                            double tau, double dt, long cl, long cr,
                                            double *__restrict__ F, unsigned char &ok,
                                            int *__restrict__ status) {
  const long N = g.N, c = cl;
  const double dta = MODE == 0 ? dt / 2. : dt;
  const double tauFactor = MODE == 0 ? tau + 0.25 * dt : tau + 0.5 * dt;
  double speed = 0.;
  ok = 0;
  do {
    if (!m_full && m[(D - 1) * N + cl] < 1. && m[(D - 1) * N + cr] < 1.) break;
    const double gl = 1. / sqrt(1 - L.vx * L.vx - L.vy * L.vy - L.vz * L.vz);
    const double gr = 1. / sqrt(1 - R.vx * R.vx - R.vy * R.vy - R.vz * R.vz);
    double Ul[7], Ur[7], Fl[7], Fr[7];
    Ul[X_] = gl * gl * (L.e + L.p) * L.vx;
    Ul[Y_] = gl * gl * (L.e + L.p) * L.vy;
    Ul[Z_] = gl * gl * (L.e + L.p) * L.vz;
    Ul[T_] = gl * gl * (L.e + L.p) - L.p;
    Ul[NB_] = gl * L.nb;
    Ul[NQ_] = gl * L.nq;
    Ul[NS_] = gl * L.ns;
    Ur[X_] = gr * gr * (R.e + R.p) * R.vx;
    Ur[Y_] = gr * gr * (R.e + R.p) * R.vy;
    Ur[Z_] = gr * gr * (R.e + R.p) * R.vz;
    Ur[T_] = gr * gr * (R.e + R.p) - R.p;
    Ur[NB_] = gr * R.nb;
    Ur[NQ_] = gr * R.nq;
    Ur[NS_] = gr * R.ns;
    const double vl = D == 1 ? L.vx : (D == 2 ? L.vy : L.vz);
    const double vr = D == 1 ? R.vx : (D == 2 ? R.vy : R.vz);
    if (D != 3) {
#pragma unroll
      for (int k = 0; k < 7; k++) {
        Fl[k] = Ul[k] * vl;
        Fr[k] = Ur[k] * vr;
      }
      Fl[T_] = Ul[T_] * vl + L.p * vl;
      Fr[T_] = Ur[T_] * vr + R.p * vr;
      Fl[D] = Ul[D] * vl + L.p;
      Fr[D] = Ur[D] * vr + R.p;
    } else {
      const double t1 = tauFactor;
#pragma unroll
      for (int k = 0; k < 7; k++) {
        Fl[k] = Ul[k] * vl / t1;
        Fr[k] = Ur[k] * vr / t1;
      }
      Fl[T_] = Ul[T_] * vl / t1 + L.p * vl / t1;
      Fr[T_] = Ur[T_] * vr / t1 + R.p * vr / t1;
      Fl[Z_] = Ul[Z_] * vl / t1 + L.p / t1;
      Fr[Z_] = Ur[Z_] * vr / t1 + R.p / t1;
    }
    // wave speeds with cs^2 = 1/3 always (src/eos.h:34-35, src/hdo.cpp:181-188)
    const double cs2 = 1. / 3.;
    const double cs = sqrt(1. / 3.);  // EoS::cs() (src/eos.h:35)
    const double sEl = sqrt(El), sEr = sqrt(Er);
    const double sum = sEl + sEr;
    const double dv = vl - vr;
    const double csb = sqrt(cs2 + 0.5 * sqrt(El * Er) / (sum * sum) * (dv * dv));
    const double vb = (sEl * vl + sEr * vr) / sum;
    double bl = smin(0., smin((vb - csb) / (1 - vb * csb), (vl - cs) / (1 - vl * cs)));
    double br = smax(0., smax((vb + csb) / (1 + vb * csb), (vr + cs) / (1 + vr * cs)));
    double dx;
    if (D == 3) {
      bl = 1. / tau * bl;
      br = 1. / tau * br;
      if (L.e == 0.) bl = -1. / tau;
      if (R.e == 0.) br = 1. / tau;
      dx = g.dz;
    } else {
      if (L.e == 0.) bl = -1.;
      if (R.e == 0.) br = 1.;
      dx = D == 1 ? g.dx : g.dy;
    }
    if (bl == 0. && br == 0.) break;
    const double pre = tauFactor * dta / dx;
    const double den = -bl + br;
    double f[7];
#pragma unroll
    for (int k = 0; k < 7; k++)
      f[k] = pre * (-bl * br * (Ul[k] - Ur[k]) + br * Fl[k] - bl * Fr[k]) / den;
    if (f[NB_] != f[NB_]) {
      atomicOr(status, VH_ST_NAN_FLUX);
      break;
    }
#pragma unroll
    for (int k = 0; k < 7; k++) st_nt(F + k * N + c, f[k]);
    ok = 1;
    speed = smax(-bl, br);
  } while (0);
  return speed;
}

// The face states of the stage's state Qs (Q in the predictor, Qh in the
// corrector), the reference's early returns, then flux_tail.  A negative
// pressure of the right state (a table EoS whose spline undershoots below
// its clamped nodes) makes the reference evaluate both face states again
// from Q -- not Qh, even in the corrector -- and go on with them, without
// clamping e again and with El, Er of the first states (src/hdo.cpp:134-140):
// such an interface is left to k_faces_fix with flag 2 and El, Er in its F
// slot, so that this kernel carries one copy of the face-state code.
template <int D, int MODE, bool CH>
__device__ __forceinline__ double face_flux(const Geom &g, const DevEos &eos,
                                            const double *__restrict__ Qs,
                                            const double *__restrict__ m, int m_full,
                                            double tau, double dt, long c, int ix, int iy,
                                            int iz, double *__restrict__ F,
                                            unsigned char *__restrict__ flag,
                                            int *__restrict__ status) {
  const int i = coord<D>(ix, iy, iz), n = extent<D>(g);
  if (i >= n - 1) return 0.;
  const long s = stride<D>(g);
  const long cl = c, cr = c + s;
  const long cll = i > 0 ? c - s : c;          // clamped prev of the left cell
  const long crr = i + 2 < n ? cr + s : cr;    // clamped next of the right cell
  const long N = g.N;
  Prim L, R;
  const Divider dtau(tau);
  face_state<+1, CH>(eos, Qs, N, cll, cl, cr, dtau, L);
  face_state<-1, CH>(eos, Qs, N, cl, cr, crr, dtau, R);
  const double El = (L.e + L.p) / (1 - L.vx * L.vx - L.vy * L.vy - L.vz * L.vz);
  const double Er = (R.e + R.p) / (1 - R.vx * R.vx - R.vy * R.vy - R.vz * R.vz);
  if (L.e < 0.) {
    L.e = 0.;
    L.p = 0.;
  }
  if (R.e < 0.) {
    R.e = 0.;
    R.p = 0.;
  }
  unsigned char ok = 0;
  double speed = 0.;
  if (L.e > 1e10) {
    atomicOr(status, VH_ST_BIG_E);
  } else if (!(L.e == 0. && R.e == 0.)) {
    if (R.p < 0.) {
      F[c] = El;
      F[N + c] = Er;
      ok = 2;
    } else {
      speed = flux_tail<D, MODE>(g, m, m_full, L, R, El, Er, tau, dt, cl, cr, F, ok, status);
    }
  }
  flag[c] = ok;
  return speed;
}

// One thread per interface (slot = left cell); the wave's largest signal
// speed goes to the slot group sig of this stage and axis.
template <int D, int MODE, bool CH, int EK>
__global__ void __launch_bounds__(256) VH_OCC(VH_W_FACES)
k_faces(Geom g, DevEos eos_a, const double *__restrict__ Qs, const double *__restrict__ m,
        int m_full, const double *__restrict__ tp, double dt, double *__restrict__ F,
        unsigned char *__restrict__ flag, int *__restrict__ status,
        unsigned long long *__restrict__ sig) {
  const DevEos eos = eos_of_kind<EK>(eos_a);
  const double tau = *tp;
  const long t = xcd_block() * blockDim.x + threadIdx.x;  // interfaces in y-strip order
  double speed = 0.;
  if (t < g.N) {
    int ix, iy, iz;
    strip_cell(g.nx, g.ny, g.nz, VH_STRIP_ROWS, t, ix, iy, iz);
    speed = face_flux<D, MODE, CH>(g, eos, Qs, m, m_full, tau, dt, g.idx(ix, iy, iz), ix, iy, iz,
                                   F, flag, status);
  }
  wave_speed_max(sig, speed);
}

// The interfaces k_faces left with flag 2: both face states again from Q
// (Qr), then flux_tail with the El, Er of the first states.  Launched after
// k_faces for table EoSs only (p = e/3 >= 0 never leaves one).
template <int D, int MODE, bool CH>
__global__ void __launch_bounds__(256)
k_faces_fix(Geom g, DevEos eos, const double *__restrict__ Qr, const double *__restrict__ m,
            int m_full, const double *__restrict__ tp, double dt, double *__restrict__ F,
            unsigned char *__restrict__ flag, int *__restrict__ status,
            unsigned long long *__restrict__ sig) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double speed = 0.;
  if (c < g.N && flag[c] == 2) {
    const double tau = *tp;
    int ix, iy, iz;
    decompose(g, c, ix, iy, iz);
    const int i = coord<D>(ix, iy, iz), n = extent<D>(g);
    const long s = stride<D>(g), N = g.N;
    const long cl = c, cr = c + s;
    const long cll = i > 0 ? c - s : c;
    const long crr = i + 2 < n ? cr + s : cr;
    const double El = F[c], Er = F[N + c];
    Prim L, R;
    const Divider dtau(tau);
    face_state<+1, CH>(eos, Qr, N, cll, cl, cr, dtau, L);
    face_state<-1, CH>(eos, Qr, N, cl, cr, crr, dtau, R);
    unsigned char ok = 0;
    speed = flux_tail<D, MODE>(g, m, m_full, L, R, El, Er, tau, dt, cl, cr, F, ok, status);
    flag[c] = ok;
  }
  wave_speed_max(sig, speed);
}

// ------------------------------------------------------ stage update
// Gathers the six interface fluxes of each cell in the reference's order,
// adds the Milne source (Hydro::source/source_step, src/hdo.cpp:307-353) and
// applies the stage update: predictor Qh = Q + flux (Cell::updateQtoQhByFlux,
// src/cll.cpp:86-88); corrector Q += flux unless Q0 + flux0 < 0
// (Cell::updateByFlux, src/cll.cpp:70-74).
template <int MODE, bool CH, int EK>
__global__ void __launch_bounds__(256) VH_OCC(VH_W_UPDATE)
k_update(Geom g, DevEos eos_a, const double *__restrict__ Q, const double *__restrict__ Qh,
         const double *__restrict__ Fx, const unsigned char *__restrict__ fx,
         const double *__restrict__ Fy, const unsigned char *__restrict__ fy,
         const double *__restrict__ Fz, const unsigned char *__restrict__ fz,
         const double *__restrict__ tp, double dt, double *__restrict__ out,
         int *__restrict__ status) {
  const DevEos eos = eos_of_kind<EK>(eos_a);
  const double tau = *tp;
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int ix, iy, iz;
  strip_cell(g.nx, g.ny, g.nz, VH_STRIP_ROWS, t, ix, iy, iz);
  const long c = g.idx(ix, iy, iz);
  const long N = g.N;
  double acc[7];
#pragma unroll
  for (int k = 0; k < 7; k++) acc[k] = 0.;
  const long sy = g.nx, sz = (long)g.nx * g.ny;
  if (ix > 0 && fx[c - 1])
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] += Fx[k * N + c - 1];
  if (ix < g.nx - 1 && fx[c])
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] += -Fx[k * N + c];
  if (iy > 0 && fy[c - sy])
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] += Fy[k * N + c - sy];
  if (iy < g.ny - 1 && fy[c])
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] += -Fy[k * N + c];
  if (iz > 0 && fz[c - sz])
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] += Fz[k * N + c - sz];
  if (iz < g.nz - 1 && fz[c])
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] += -Fz[k * N + c];
  // source on the unreconstructed state (src/hdo.cpp:335-341)
  const double *Qsrc = MODE == 0 ? Q : Qh;
  const double tau1 = MODE == 0 ? tau : tau + 0.5 * dt;
  const double sdt = MODE == 0 ? dt / 2. : dt;
  double q[7];
  const Divider d1(tau1);
#pragma unroll
  for (int k = 0; k < 7; k++) q[k] = d1(Qsrc[k * N + c]);  // Q / tau1
  Prim r;
  c2p_ref<CH>(eos, q, r);
  double S[7];
  S[T_] = -q[T_] * r.vz * r.vz - r.p * (1. + r.vz * r.vz);
  S[X_] = 0.;
  S[Y_] = 0.;
  S[Z_] = -q[Z_];
  S[NB_] = S[NQ_] = S[NS_] = 0.;
#pragma unroll
  for (int k = 0; k < 7; k++) S[k] *= sdt;
  if (S[NB_] != S[NB_]) atomicOr(status, VH_ST_NAN_SOURCE);
#pragma unroll
  for (int k = 0; k < 7; k++) acc[k] += S[k];
  if (MODE == 0) {
#pragma unroll
    for (int k = 0; k < 7; k++) st_nt(out + k * N + c, Q[k * N + c] + acc[k]);
  } else {
    const bool skip = Q[c] + acc[0] < 0.;
#pragma unroll
    for (int k = 0; k < 7; k++) st_nt(out + k * N + c, skip ? Q[k * N + c] : Q[k * N + c] + acc[k]);
  }
}

// ------------------------------------------------------------ ghost fill
// Fluid::correctImagCells / correctImagCellsFull (src/fld.cpp:211-382): per
// boundary line one thread runs the reference's copy sequence with clamped
// cell indices (so degenerate small extents behave as in the reference).
// AXIS: 3 = z pass (thread per (ix,iy)), 2 = y pass, 1 = x pass.
__device__ inline void set_clamped(double *a, long N, long dst, const double v[7]) {
  const bool z = v[0] < 0.;
#pragma unroll
  for (int k = 0; k < 7; k++) a[k * N + dst] = z ? 0. : v[k];
}

template <int AXIS, int FULL>
__global__ void __launch_bounds__(256)
k_ghost(Geom g, double *__restrict__ Q, double *__restrict__ Qh, double *__restrict__ pi,
        double *__restrict__ Pi) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int a, b;  // the two coordinates spanning the boundary face
  long nline;
  if (AXIS == 3) nline = (long)g.nx * g.ny;
  else if (AXIS == 2) nline = (long)g.nx * g.nz;
  else nline = (long)g.ny * g.nz;
  if (t >= nline) return;
  const long N = g.N;
  const int n = AXIS == 3 ? g.nz : (AXIS == 2 ? g.ny : g.nx);
  if (AXIS == 3) { a = (int)(t % g.nx); b = (int)(t / g.nx); }
  else if (AXIS == 2) { a = (int)(t % g.nx); b = (int)(t / g.nx); }
  else { a = (int)(t % g.ny); b = (int)(t / g.ny); }
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
    set_clamped(Q, N, c1, v);
    set_clamped(Q, N, c0, v);
    if (!FULL) {
#pragma unroll
      for (int k = 0; k < 7; k++) v[k] = Qh[k * N + cs];
      set_clamped(Qh, N, c1, v);
      set_clamped(Qh, N, c0, v);
    } else {
      double p[10];
#pragma unroll
      for (int k = 0; k < 10; k++) p[k] = pi[k * N + cs];
      const double P = Pi[cs];
#pragma unroll
      for (int k = 0; k < 10; k++) {
        pi[k * N + c0] = p[k];
        pi[k * N + c1] = p[k];
      }
      Pi[c0] = P;
      Pi[c1] = P;
    }
  }
}

// Single-pass form of the three ghost passes.  When every axis has n = 1 or
// n >= 5, the z, y, x copies of the reference compose to "copy from the cell
// whose ghost coordinates are mapped to the nearest interior index"
// (0,1 -> 2 and n-2,n-1 -> n-3), because those sources are never ghosts of an
// axis with n >= 5 and the Q[T_] < 0 zeroing of setQ is idempotent.  An axis
// with n = 1 makes every cell a ghost of itself: the reference then re-sets
// every cell, which only applies the zeroing; k_ghost_zero does that first, so
// that k_ghost1 never writes a cell another thread reads.  One thread per
// cell; PT: storage type of pi/Pi.
__device__ __forceinline__ int ghost_map(int i, int n) {
  return n == 1 ? 0 : (i < 2 ? 2 : (i > n - 3 ? n - 3 : i));
}

template <int FULL>
__global__ void __launch_bounds__(256)
k_ghost_zero(Geom g, double *__restrict__ Q, double *__restrict__ Qh) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  const long N = g.N;
  if (Q[c] < 0.)
    for (int k = 0; k < 7; k++) Q[k * N + c] = 0.;
  if (!FULL && Qh[c] < 0.)
    for (int k = 0; k < 7; k++) Qh[k * N + c] = 0.;
}

template <int FULL, typename PT>
__global__ void __launch_bounds__(256)
k_ghost1(Geom g, double *__restrict__ Q, double *__restrict__ Qh, PT *__restrict__ pi,
         PT *__restrict__ Pi) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  int ix, iy, iz;
  decompose(g, c, ix, iy, iz);
  const int sx = ghost_map(ix, g.nx), sy = ghost_map(iy, g.ny), sz = ghost_map(iz, g.nz);
  if (sx == ix && sy == iy && sz == iz) return;
  const long N = g.N, src = g.idx(sx, sy, sz);
  double v[7];
#pragma unroll
  for (int k = 0; k < 7; k++) v[k] = Q[k * N + src];
  set_clamped(Q, N, c, v);
  if (!FULL) {
#pragma unroll
    for (int k = 0; k < 7; k++) v[k] = Qh[k * N + src];
    set_clamped(Qh, N, c, v);
  } else {
#pragma unroll
    for (int k = 0; k < 10; k++) pi[k * N + c] = pi[k * N + src];
    Pi[c] = Pi[src];
  }
}

// k_ghost1 over the ghost cells only (every extent >= 5; ghost_cell)
template <int FULL, typename PT>
__global__ void __launch_bounds__(256)
k_ghost_list(Geom g, long n, double *__restrict__ Q, double *__restrict__ Qh,
             PT *__restrict__ pi, PT *__restrict__ Pi) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int ix, iy, iz;
  ghost_cell(g.nx, g.ny, g.nz, t, ix, iy, iz);
  const long N = g.N, c = g.idx(ix, iy, iz);
  const long src = g.idx(ghost_map(ix, g.nx), ghost_map(iy, g.ny), ghost_map(iz, g.nz));
  double v[7];
#pragma unroll
  for (int k = 0; k < 7; k++) v[k] = Q[k * N + src];
  set_clamped(Q, N, c, v);
  if (!FULL) {
#pragma unroll
    for (int k = 0; k < 7; k++) v[k] = Qh[k * N + src];
    set_clamped(Qh, N, c, v);
  } else {
#pragma unroll
    for (int k = 0; k < 10; k++) pi[k * N + c] = pi[k * N + src];
    Pi[c] = Pi[src];
  }
}

bool ghost_fusable(const Geom &g) {
  auto ok = [](int n) { return n == 1 || n >= 5; };
  return ok(g.nx) && ok(g.ny) && ok(g.nz);
}

void launch_ghost1(int full, int pi_float, const Geom &g, double *Q, double *Qh, void *pi,
                   void *Pi, hipStream_t st) {
  const dim3 gr((unsigned)((g.N + 255) / 256)), b(256);
  if (g.nx >= 5 && g.ny >= 5 && g.nz >= 5) {  // launch over the ghost cells only
    const long n = ghost_count(g.nx, g.ny, g.nz);
    const dim3 gl((unsigned)((n + 255) / 256));
    if (!full)
      hipLaunchKernelGGL((k_ghost_list<0, double>), gl, b, 0, st, g, n, Q, Qh, (double *)pi,
                         (double *)Pi);
    else if (pi_float)
      hipLaunchKernelGGL((k_ghost_list<1, float>), gl, b, 0, st, g, n, Q, Qh, (float *)pi,
                         (float *)Pi);
    else
      hipLaunchKernelGGL((k_ghost_list<1, double>), gl, b, 0, st, g, n, Q, Qh, (double *)pi,
                         (double *)Pi);
    return;
  }
  if (g.nx == 1 || g.ny == 1 || g.nz == 1) {
    if (full) hipLaunchKernelGGL((k_ghost_zero<1>), gr, b, 0, st, g, Q, Qh);
    else hipLaunchKernelGGL((k_ghost_zero<0>), gr, b, 0, st, g, Q, Qh);
  }
  if (!full)
    hipLaunchKernelGGL((k_ghost1<0, double>), gr, b, 0, st, g, Q, Qh, (double *)pi, (double *)Pi);
  else if (pi_float)
    hipLaunchKernelGGL((k_ghost1<1, float>), gr, b, 0, st, g, Q, Qh, (float *)pi, (float *)Pi);
  else
    hipLaunchKernelGGL((k_ghost1<1, double>), gr, b, 0, st, g, Q, Qh, (double *)pi, (double *)Pi);
}

// ---------------------------------------------------------------- updateM
// Fluid::updateM (src/fld.cpp:384-423), pass 1: dm from the neighbours' m.
__global__ void __launch_bounds__(256)
k_update_m_dm(Geom g, const double *__restrict__ m, double *__restrict__ dm,
              const double *__restrict__ tp, double dt) {
  const double tau = *tp;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  int ix, iy, iz;
  decompose(g, c, ix, iy, iz);
  const long N = g.N;
  auto M = [&](int d, long cc) { return m[d * N + cc]; };
  auto maxM = [&](long cc) { return smax(M(0, cc), smax(M(1, cc), M(2, cc))); };
  double d0 = 0., d1 = 0., d2 = 0.;
  if (maxM(c) < 1.) {
    if (M(0, g.cidx(ix + 1, iy, iz)) >= 1. || M(0, g.cidx(ix - 1, iy, iz)) >= 1.) d0 = dt / g.dx;
    if (M(1, g.cidx(ix, iy + 1, iz)) >= 1. || M(1, g.cidx(ix, iy - 1, iz)) >= 1.) d1 = dt / g.dy;
    if (M(2, g.cidx(ix, iy, iz + 1)) >= 1. || M(2, g.cidx(ix, iy, iz - 1)) >= 1.)
      d2 = dt / g.dz / tau;
    if (d0 == 0. && d1 == 0.) {
      if (maxM(g.cidx(ix + 1, iy + 1, iz)) >= 1. || maxM(g.cidx(ix + 1, iy - 1, iz)) >= 1. ||
          maxM(g.cidx(ix - 1, iy + 1, iz)) >= 1. || maxM(g.cidx(ix - 1, iy - 1, iz)) >= 1.) {
        d0 = 0.707 * dt / g.dx;
        d1 = 0.707 * dt / g.dy;
      }
    }
  }
  dm[c] = d0;
  dm[N + c] = d1;
  dm[2 * N + c] = d2;
}

// pass 2: Cell::addM for x, y, z in turn (src/cll.h:146-150)
__global__ void __launch_bounds__(256)
k_update_m_apply(Geom g, double *__restrict__ m, const double *__restrict__ dm) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  const long N = g.N;
  double mm[3] = {m[c], m[N + c], m[2 * N + c]};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    mm[d] += dm[d * N + c];
    if (mm[d] > 0.9) mm[0] = mm[1] = mm[2] = 1.;
  }
  m[c] = mm[0];
  m[N + c] = mm[1];
  m[2 * N + c] = mm[2];
}

// --------------------------------------------------------- viscous part
// Primitive states shared by the Israel-Stewart kernels: con2prim of
// Qh/(tau - dt/2) (the cell-centre state of ISformal, NSquant's own u,
// visc_flux and visc_source_step) -> ph[8][N]; con2prim of Qh/tau (NSquant's
// neighbour states, src/hdo.cpp:472-568) -> pt[4][N] = {e, vx, vy, vz}.
// tau here is tau_{n+1}.  Tt (may be NULL): T of the pt state where e > 0,
// and everywhere for an EoS that depends on the charges (whose T at e <= 0
// needs the cell's n) -- the neighbour temperatures of the vorticity rows.
template <bool CH, int EK>
__global__ void __launch_bounds__(256) VH_OCC(VH_W_VPRIM)
k_vprim(Geom g, DevEos eos_a, const double *__restrict__ Qh, const double *__restrict__ tp,
        double dt, double *__restrict__ ph, double *__restrict__ pt, double *__restrict__ Tt) {
  const DevEos eos = eos_of_kind<EK>(eos_a);
  const double tau = *tp;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  const long N = g.N;
  double q[7];
  Prim r;
  const double th = tau - dt / 2.;
  const Divider dth(th), dtau(tau);
#pragma unroll
  for (int k = 0; k < 7; k++) q[k] = dth(Qh[k * N + c]);  // Qh / th
  c2p_ref<CH>(eos, q, r);
  ph[c] = r.e;
  ph[N + c] = r.p;
  ph[2 * N + c] = r.nb;
  ph[3 * N + c] = r.nq;
  ph[4 * N + c] = r.ns;
  ph[5 * N + c] = r.vx;
  ph[6 * N + c] = r.vy;
  ph[7 * N + c] = r.vz;
#pragma unroll
  for (int k = 0; k < 7; k++) q[k] = dtau(Qh[k * N + c]);  // Qh / tau
  c2p_ref<CH>(eos, q, r);
  pt[c] = r.e;
  pt[N + c] = r.vx;
  pt[2 * N + c] = r.vy;
  pt[3 * N + c] = r.vz;
  if (Tt) Tt[c] = (r.e > 0. || CH) ? eos.T<CH>(r.e, r.nb, r.nq, r.ns) : 0.;
}

// ISformal loop #1 (src/hdo.cpp:691-864) with NSquant (src/hdo.cpp:381-652):
// Navier-Stokes values, relaxation and second-order/cross terms; writes
// pi0, piH0, Pi0, PiH0.  (The loop's flux contribution (tau-dt)*dQ_n is
// formed in k_visc2, which reads the same old pi/Pi before replacing them.)
// With VORT, also the cell's vorticity entries and op code (vh_vort.h).
template <bool VORT, bool CH, int EK>
__global__ void __launch_bounds__(256) VH_OCC(VH_W_VISC1)
k_visc1(Geom g, DevEos eos_a, DevTrc trc, const double *__restrict__ Qprev,
        const double *__restrict__ Qn, const double *__restrict__ ph,
        const double *__restrict__ pt, const double *__restrict__ m, int m_full,
        const double *__restrict__ pi, const double *__restrict__ Pi,
        const double *__restrict__ tp, double dt, double *__restrict__ pi0, double *__restrict__ piH0, double *__restrict__ Pi0,
        double *__restrict__ PiH0, int *__restrict__ status, double *__restrict__ db,
        unsigned char *__restrict__ dop, const double *__restrict__ Tt) {
  const DevEos eos = eos_of_kind<EK>(eos_a);
  const double tau = *tp;
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int ix, iy, iz;
  strip_cell(g.nx, g.ny, g.nz, VH_STRIP_ROWS, t, ix, iy, iz);
  const long c = g.idx(ix, iy, iz);
  const long N = g.N;
  const double e = ph[c], nb = ph[2 * N + c], nq = ph[3 * N + c], ns = ph[4 * N + c];
  const double vx = ph[5 * N + c], vy = ph[6 * N + c], vz = ph[7 * N + c];
  if (e <= 0.) {
#pragma unroll
    for (int k = 0; k < 10; k++) {
      piH0[k * N + c] = 0.0;
      pi0[k * N + c] = 0.0;
    }
    PiH0[c] = 0.0;
    Pi0[c] = 0.0;
    if (VORT) vort::store(db, dop, N, c, vort::K_EMPTY, nullptr);
    return;
  }
  const double gamma = 1.0 / sqrt(1.0 - vx * vx - vy * vy - vz * vz);
  const double u[4] = {gamma, gamma * vx, gamma * vy, gamma * vz};
  const double gm[4] = {1., -1., -1., -1.};
  // the quotients x / gamma and x / 5.068 of the IEEE division (Divider)
  const Divider dg(gamma), d5068(5.068);

  // ---- NSquant
  double dmu[4][4], piNS[10], PiNS, du;
  double vv[16];             // vorticity entries formed by this cell (VORT)
  unsigned char vflags = 0;
  bool vac = false;
  if (!m_full) {
    auto maxM = [&](int a, int b) {
      const long cc = g.cidx(ix + a, iy + b, iz);
      return smax(m[cc], smax(m[N + cc], m[2 * N + cc]));
    };
    vac = maxM(1, 0) <= 0.9 || maxM(0, 1) <= 0.9 || maxM(-1, 0) <= 0.9 || maxM(0, -1) <= 0.9 ||
          maxM(1, 1) <= 0.9 || maxM(1, -1) <= 0.9 || maxM(-1, 1) <= 0.9 || maxM(-1, -1) <= 0.9;
  }
  if (vac) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) dmu[i][j] = 0.;
#pragma unroll
    for (int k = 0; k < 10; k++) piNS[k] = 0.;
    PiNS = du = 0.;
    if (VORT) vort::store(db, dop, N, c, vort::K_STALE, nullptr);
  } else {
    const double VMIN = 1e-2, UDIFF = 3.0;
    double q[7];
    Prim p0, p1;
#pragma unroll
    for (int k = 0; k < 7; k++) q[k] = Divider(tau - dt)(Qprev[k * N + c]);
    c2p_ref<CH>(eos, q, p0);
#pragma unroll
    for (int k = 0; k < 7; k++) q[k] = Divider(tau)(Qn[k * N + c]);
    c2p_ref<CH>(eos, q, p1);
    const double e0 = p0.e, e1 = e;  // e1: the H-centre state (src/hdo.cpp:427)
    double ut0 = 1.0 / sqrt(1.0 - p0.vx * p0.vx - p0.vy * p0.vy - p0.vz * p0.vz);
    double ux0 = ut0 * p0.vx, uy0 = ut0 * p0.vy, uz0 = ut0 * p0.vz;
    double ut1 = 1.0 / sqrt(1.0 - p1.vx * p1.vx - p1.vy * p1.vy - p1.vz * p1.vz);
    double ux1 = ut1 * p1.vx, uy1 = ut1 * p1.vy, uz1 = ut1 * p1.vz;
    const Divider ddt(dt);  // x / dt
    dmu[0][0] = ddt((ut1 * ut1 - ut0 * ut0) / 2. / u[0]);
    dmu[0][1] = ddt((ux1 * ux1 - ux0 * ux0) / 2. / u[1]);
    dmu[0][2] = ddt((uy1 * uy1 - uy0 * uy0) / 2. / u[2]);
    dmu[0][3] = ddt((uz1 * uz1 - uz0 * uz0) / 2. / u[3]);
    if (fabs(0.5 * (ut1 + ut0) / ut1) > UDIFF) dmu[0][0] = ddt(ut1 - ut0);
    if (fabs(u[1]) < VMIN || fabs(0.5 * (ux1 + ux0) / ux1) > UDIFF) dmu[0][1] = ddt(ux1 - ux0);
    if (fabs(u[2]) < VMIN || fabs(0.5 * (uy1 + uy0) / uy1) > UDIFF) dmu[0][2] = ddt(uy1 - uy0);
    if (fabs(u[3]) < VMIN || fabs(0.5 * (uz1 + uz0) / uz1) > UDIFF) dmu[0][3] = ddt(uz1 - uz0);
    if (e1 <= 0. || e0 <= 0.) dmu[0][0] = dmu[0][1] = dmu[0][2] = dmu[0][3] = 0.;
    if (VORT) {  // d_tau (u^nu / T), src/hdo.cpp:419-458
      const double a0[4] = {ut0, ux0, uy0, uz0}, a1[4] = {ut1, ux1, uy1, uz1};
      vort::row_time(vv, a0, a1, eos.T<CH>(p0.e, p0.nb, p0.nq, p0.ns), eos.T<CH>(p1.e, p1.nb, p1.nq, p1.ns),
                     e1, e0, dt);
    }
    // space derivatives from the neighbours' Qh/tau_{n+1} states
#pragma unroll
    for (int d = 1; d <= 3; d++) {
      const long ca = g.cidx(ix + (d == 1), iy + (d == 2), iz + (d == 3));
      const long cb = g.cidx(ix - (d == 1), iy - (d == 2), iz - (d == 3));
      const double ea = pt[ca], eb = pt[cb];
      // the reference takes both temperatures before its e test (src/hdo.cpp:474-481);
      // k_vprim stored them where e > 0 (the same evaluation)
      const double TA = VORT ? ((ea > 0. || CH) ? Tt[ca] : eos.T(ea)) : 0.;
      const double TB = VORT ? ((eb > 0. || CH) ? Tt[cb] : eos.T(eb)) : 0.;
      if (ea > 0. && eb > 0.) {
        const double avx = pt[N + ca], avy = pt[2 * N + ca], avz = pt[3 * N + ca];
        const double bvx = pt[N + cb], bvy = pt[2 * N + cb], bvz = pt[3 * N + cb];
        ut0 = 1.0 / sqrt(1.0 - bvx * bvx - bvy * bvy - bvz * bvz);
        ux0 = ut0 * bvx;
        uy0 = ut0 * bvy;
        uz0 = ut0 * bvz;
        ut1 = 1.0 / sqrt(1.0 - avx * avx - avy * avy - avz * avz);
        ux1 = ut1 * avx;
        uy1 = ut1 * avy;
        uz1 = ut1 * avz;
        if (VORT) {  // src/hdo.cpp:495-503,541-549,587-595
          const double A[4] = {ut1, ux1, uy1, uz1}, B[4] = {ut0, ux0, uy0, uz0};
          const double den = d == 1 ? g.dx : (d == 2 ? g.dy : g.dz * (tau + 0.5 * dt));
          vort::row_space(vv + 4 * d, A, B, TA, TB, den);
          vflags |= d == 1 ? vort::F_R1 : (d == 2 ? vort::F_R2 : vort::F_R3);
        }
        if (d < 3) {
          const double h = d == 1 ? g.dx : g.dy;
          dmu[d][0] = 0.25 * (ut1 * ut1 - ut0 * ut0) / u[0] / h;
          dmu[d][1] = 0.25 * (ux1 * ux1 - ux0 * ux0) / u[1] / h;
          dmu[d][2] = 0.25 * (uy1 * uy1 - uy0 * uy0) / u[2] / h;
          dmu[d][3] = 0.25 * (uz1 * uz1 - uz0 * uz0) / u[3] / h;
          if (fabs(0.5 * (ut1 + ut0) / u[0]) > UDIFF) dmu[d][0] = 0.5 * (ut1 - ut0) / h;
          if (fabs(u[1]) < VMIN || fabs(0.5 * (ux1 + ux0) / u[1]) > UDIFF)
            dmu[d][1] = 0.5 * (ux1 - ux0) / h;
          if (fabs(u[2]) < VMIN || fabs(0.5 * (uy1 + uy0) / u[2]) > UDIFF)
            dmu[d][2] = 0.5 * (uy1 - uy0) / h;
          if (fabs(u[3]) < VMIN || fabs(0.5 * (uz1 + uz0) / u[3]) > UDIFF)
            dmu[d][3] = 0.5 * (uz1 - uz0) / h;
        } else {
          const double dz = g.dz, tz = tau + 0.5 * dt;
          dmu[3][0] = 0.25 * (ut1 * ut1 - ut0 * ut0) / u[0] / dz / tz;
          dmu[3][1] = 0.25 * (ux1 * ux1 - ux0 * ux0) / u[1] / dz / tz;
          dmu[3][2] = 0.25 * (uy1 * uy1 - uy0 * uy0) / u[2] / dz / tz;
          dmu[3][3] = 0.25 * (uz1 * uz1 - uz0 * uz0) / u[3] / dz / tz;
          if (fabs(0.5 * (ut1 + ut0) / u[0]) > UDIFF) dmu[3][0] = 0.5 * (ut1 - ut0) / dz / tz;
          if (fabs(u[1]) < VMIN || fabs(0.5 * (ux1 + ux0) / u[1]) > UDIFF)
            dmu[3][1] = 0.5 * (ux1 - ux0) / dz / tz;
          if (fabs(u[2]) < VMIN || fabs(0.5 * (uy1 + uy0) / u[2]) > UDIFF)
            dmu[3][2] = 0.5 * (uy1 - uy0) / dz / tz;
          if (fabs(u[3]) < VMIN || fabs(0.5 * (uz1 + uz0) / u[3]) > UDIFF)
            dmu[3][3] = 0.5 * (uz1 - uz0) / dz / tz;
        }
      } else {
        dmu[d][0] = dmu[d][1] = dmu[d][2] = dmu[d][3] = 0.;
      }
    }
    dmu[3][0] += u[3] / (tau - 0.5 * dt);
    dmu[3][3] += u[0] / (tau - 0.5 * dt);
    if (VORT) {  // Christoffel terms (src/hdo.cpp:608-611), then Cell::setDbeta
      const double TH = eos.T<CH>(e, nb, nq, ns);
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
    // transport coefficients at the H-centre state
    double T, mub, muq, mus, p;
    eos.eos<CH>(e, nb, nq, ns, T, mub, muq, mus, p);
    const double s = T > 0.0 ? (e + p - mub * nb - muq * nq - mus * ns) / T : 0.;  // EoS::s
    const double etaS = trc.etaS(e, nb, T, mub, s, p);
    const double zetaS = trc.zetaS(eos, e, T, s, p);
    // sigma^{mu nu} through the projector Z (src/hdo.cpp:613-637), evaluated
    // term by term; the terms where Z vanishes identically add +-0 and are
    // dropped by the compiler after unrolling (bit-identical sums).
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j <= i; j++) {
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
          for (int l = 0; l < 4; l++) {
            if (!(j == l || i == l || k == l)) continue;
            double z = 0.0;
            if (j == l) z += 0.5 * ((i == k ? gm[i] : 0.) - u[i] * u[k]);
            if (i == l) z += 0.5 * ((j == k ? gm[j] : 0.) - u[j] * u[k]);
            if (k == l) z -= ((i == j ? gm[i] : 0.) - u[i] * u[j]) / 3.0;
            a += d5068(z * dmu[k][l] * 2.0 * etaS * s);
          }
        piNS[sym(i, j)] = a;
      }
    PiNS = d5068(-zetaS * s * (dmu[0][0] + dmu[1][1] + dmu[2][2] + dmu[3][3]));
    du = dmu[0][0] + dmu[1][1] + dmu[2][2] + dmu[3][3];
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 10; k++) bad |= (isinf(piNS[k]) || isnan(piNS[k]));
    if (bad) atomicOr(status, VH_ST_NAN_PI);
  }

  // ---- relaxation + sources
  double T, mub, muq, mus, p;
  eos.eos<CH>(e, nb, nq, ns, T, mub, muq, mus, p);
  const double s = T > 0.0 ? (e + p - mub * nb - muq * nq - mus * ns) / T : 0.;  // EoS::s
  const double etaS = trc.etaS(e, nb, T, mub, s, p);
  const double eta = etaS * s;
  double sig[10];
#pragma unroll
  for (int k = 0; k < 10; k++) sig[k] = eta <= 0.0 ? 0.0 : 0.5 * piNS[k] / eta * 5.068;
  double taupi, tauPi;
  trc.getTau(eos, e, nb, T, mub, s, p, taupi, tauPi);
  double dpp = 4. / 3., tpp = 10. / 7., lpP = 6. / 5.;
  double phi7 = 9. / 70. / eos.p<CH>(e, nb, nq, ns);
  if (isinf(phi7)) phi7 = 0.0;
  phi7 = phi7 / taupi;
  if (taupi < 0.5 * dt) dpp = tpp = lpP = phi7 = 0.0;
  const Divider dtp(taupi);  // used where taupi > dt/2
  double dPP = 2. / 3., lPp = 8. / 5. * (1. / 3. - eos.cs2e(e));
  if (tauPi < 0.5 * dt) dPP = lPp = 0.0;

  double P[10], H[10], F[10];
  const double Pc = Pi[c];
#pragma unroll
  for (int k = 0; k < 10; k++) P[k] = pi[k * N + c];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      const int ij = sym(i, j);
      H[ij] = taupi > 0.5 * dt ? P[ij] - dtp(dg((P[ij] - piNS[ij]) * dt / 2.0)) : piNS[ij];
    }
  double PH = tauPi > 0.5 * dt ? Pc - dg((Pc - PiNS) * dt / 2.0) / tauPi : PiNS;
  const Divider dh(tau - dt * 0.75);  // x / tau1 of the half step
  H[sym(0, 0)] += dh(-2. * vz * P[sym(0, 3)]) * dt / 2.;
  H[sym(3, 3)] += -(dh(2. * vz) * P[sym(0, 3)]) * dt / 2.;
  H[sym(3, 0)] += -(dh(vz) * P[sym(0, 0)] + dh(vz) * P[sym(3, 3)]) * dt / 2.;
  H[sym(1, 0)] += dh(-vz) * P[sym(1, 3)] * dt / 2.;
  H[sym(2, 0)] += dh(-vz) * P[sym(2, 3)] * dt / 2.;
  H[sym(3, 1)] += -(dh(vz) * P[sym(0, 1)]) * dt / 2.;
  H[sym(3, 2)] += -(dh(vz) * P[sym(0, 2)]) * dt / 2.;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      const int ij = sym(i, j);
      double Dij = -u[i] * u[j];  // the projector g^{ij} - u^i u^j, formed where used
      if (i == j) Dij += gm[i];
      double t = H[ij];
      t += dg(-dpp * P[ij] * du + lpP * Pc * sig[ij]) * 0.5 * dt;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        t += dg(phi7 * P[sym(i, k)] * P[sym(j, k)] * gm[k] -
                tpp * 0.5 * (P[sym(i, k)] * sig[sym(j, k)] * gm[k] + P[sym(j, k)] * sig[sym(i, k)] * gm[k])) *
             0.5 * dt;
#pragma unroll
        for (int l = 0; l < 4; l++)
          t += dg((-P[sym(i, k)] * u[j] - P[sym(j, k)] * u[i]) * u[l] * dmu[l][k] * gm[k]) * 0.5 * dt -
               dg(1. / 3. * Dij * P[sym(k, l)] * (phi7 * P[sym(k, l)] - tpp * sig[sym(k, l)]) * gm[k] *
                  gm[l]) * 0.5 * dt;
      }
      H[ij] = t;
      st_nt(piH0 + ij * N + c, t);
    }
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int l = 0; l < 4; l++)
      PH += dg(lPp * P[sym(k, l)] * sig[sym(k, l)] * gm[k] * gm[l]) * 0.5 * dt;
  PH += dg(-dPP * Pc * du) * 0.5 * dt;
  // full step
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      const int ij = sym(i, j);
      F[ij] = taupi > 0.5 * dt ? P[ij] - dtp(dg((H[ij] - piNS[ij]) * dt)) : piNS[ij];
    }
  double PF = tauPi > 0.5 * dt ? Pc - dg((PH - PiNS) * dt) / tauPi : PiNS;
  const Divider df(tau - dt * 0.5);  // x / tau1 of the full step
  F[sym(0, 0)] += df(-2. * vz) * H[sym(0, 3)] * dt;
  F[sym(3, 3)] += -(df(2. * vz) * H[sym(0, 3)]) * dt;
  F[sym(3, 0)] += -(df(vz) * H[sym(0, 0)] + df(vz) * H[sym(3, 3)]) * dt;
  F[sym(1, 0)] += df(-vz) * H[sym(1, 3)] * dt;
  F[sym(2, 0)] += df(-vz) * H[sym(2, 3)] * dt;
  F[sym(3, 1)] += -(df(vz) * H[sym(0, 1)]) * dt;
  F[sym(3, 2)] += -(df(vz) * H[sym(0, 2)]) * dt;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      const int ij = sym(i, j);
      double Dij = -u[i] * u[j];
      if (i == j) Dij += gm[i];
      double t = F[ij];
      t += dg(-dpp * H[ij] * du + lpP * PH * sig[ij]) * dt;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        t += dg(phi7 * P[sym(i, k)] * H[sym(j, k)] * gm[k] -
                tpp * 0.5 * (H[sym(i, k)] * sig[sym(j, k)] * gm[k] + H[sym(j, k)] * sig[sym(i, k)] * gm[k])) *
             dt;
#pragma unroll
        for (int l = 0; l < 4; l++)
          t += dg((-H[sym(i, k)] * u[j] - H[sym(j, k)] * u[i]) * u[l] * dmu[l][k] * gm[k] -
                  1. / 3. * Dij * H[sym(k, l)] * (phi7 * H[sym(k, l)] - tpp * sig[sym(k, l)]) * gm[k] *
                      gm[l]) *
               dt;
      }
      st_nt(pi0 + ij * N + c, t);  // stored at once: F is not read again
    }
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int l = 0; l < 4; l++)
      PF += dg(lPp * H[sym(k, l)] * sig[sym(k, l)] * gm[k] * gm[l]) * dt;
  PF += dg(-dPP * PH * du) * dt;
  st_nt(PiH0 + c, PH);
  st_nt(Pi0 + c, PF);
}

// ISformal loop #2 (src/hdo.cpp:866-962) plus loop #1's flux term
// (src/hdo.cpp:717-721): upwind advection of pi0/piH0/Pi0/PiH0 from the 8
// cells toward -v, limiter, new pi/piH/Pi/PiH/viscCorrCut, and the cell's
// viscous flux accumulator vf = (tau-dt)*dQ_n - tau*dQ_{n+1} (4 x N).
__global__ void __launch_bounds__(256) VH_OCC(VH_W_VISC2)
k_visc2(Geom g, const double *__restrict__ ph, const double *__restrict__ pi0,
        const double *__restrict__ piH0, const double *__restrict__ Pi0,
        const double *__restrict__ PiH0, double *__restrict__ pi, double *__restrict__ piH,
        double *__restrict__ Pi, double *__restrict__ PiH, double *__restrict__ vcc,
        double *__restrict__ vf, const double *__restrict__ tp, double dt) {
  const double tau = *tp;
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int ix, iy, iz;
  strip_cell(g.nx, g.ny, g.nz, VH_STRIP_ROWS, t, ix, iy, iz);
  const long c = g.idx(ix, iy, iz);
  const long N = g.N;
  const double e = ph[c], p = ph[N + c];
  const double vx = ph[5 * N + c], vy = ph[6 * N + c], vz = ph[7 * N + c];
  if (e <= 0.) {
#pragma unroll
    for (int k = 0; k < 4; k++) vf[k * N + c] = 0.;
    return;
  }
  const double gamma = 1.0 / sqrt(1.0 - vx * vx - vy * vy - vz * vz);
  const double u[4] = {gamma, gamma * vx, gamma * vy, gamma * vz};
  // loop #1 flux with the old pi, Pi
  double fl[4];
  {
    const double Po = Pi[c];
#pragma unroll
    for (int i = 0; i < 4; i++) fl[i] = (tau - dt) * (pi[sym(0, i) * N + c] + Po * u[0] * u[i]);
    fl[0] += -(tau - dt) * Po;
  }
  const double xm = -vx * dt / g.dx;
  const double ym = -vy * dt / g.dy;
  const double zm = -vz * dt / g.dz / (tau - 0.5 * dt);
  const double xmH = -vx * dt / g.dx / 2.0;
  const double ymH = -vy * dt / g.dy / 2.0;
  const double zmH = -vz * dt / g.dz / (tau - 0.5 * dt) / 2.0;
  const double wx[2] = {(1. - fabs(xm)), fabs(xm)};
  const double wy[2] = {(1. - fabs(ym)), fabs(ym)};
  const double wz[2] = {(1. - fabs(zm)), fabs(zm)};
  const double wxH[2] = {(1. - fabs(xmH)), fabs(xmH)};
  const double wyH[2] = {(1. - fabs(ymH)), fabs(ymH)};
  const double wzH[2] = {(1. - fabs(zmH)), fabs(zmH)};
  const int sx = isgn(xm), sy = isgn(ym), sz = isgn(zm);
  double a[10], aH[10], A = 0.0, AH = 0.0;
#pragma unroll
  for (int k = 0; k < 10; k++) a[k] = aH[k] = 0.0;
#pragma unroll
  for (int jx = 0; jx < 2; jx++)
#pragma unroll
    for (int jy = 0; jy < 2; jy++)
#pragma unroll
      for (int jz = 0; jz < 2; jz++) {
        const long c1 = g.cidx(ix + jx * sx, iy + jy * sy, iz + jz * sz);
        const double w = wx[jx] * wy[jy] * wz[jz];
        const double wH = wxH[jx] * wyH[jy] * wzH[jz];
#pragma unroll
        for (int k = 0; k < 10; k++) {
          a[k] += w * pi0[k * N + c1];
          aH[k] += wH * piH0[k * N + c1];
        }
        A += w * Pi0[c1];
        AH += wH * PiH0[c1];
      }
  const double den = 1. - vx * vx - vy * vy - vz * vz;
  const double maxT0 = smax((e + p) / den - p, (e + p) * (vx * vx + vy * vy + vz * vz) / den + p);
  // max |pi^{mu nu}| over all 16 entries = over the 10 packed ones
  double maxpi = 0.;
#pragma unroll
  for (int k = 0; k < 10; k++)
    if (fabs(a[k]) > maxpi) maxpi = fabs(a[k]);
  bool rescaled = false;
  if (maxT0 / maxpi < 1.0) {
#pragma unroll
    for (int k = 0; k < 10; k++) {
      a[k] = 0.1 * a[k] * maxT0 / maxpi;
      aH[k] = 0.1 * aH[k] * maxT0 / maxpi;
    }
    rescaled = true;
  }
  if (fabs(A) > p) {
    if (A != 0.) A = A / fabs(A) * p;
    if (AH != 0.) AH = AH / fabs(AH) * p;
    rescaled = true;
  }
  st_nt(vcc + c, rescaled ? maxT0 / maxpi : 1.);
#pragma unroll
  for (int k = 0; k < 10; k++) {
    st_nt(pi + k * N + c, a[k]);
    st_nt(piH + k * N + c, aH[k]);
  }
  st_nt(Pi + c, A);
  st_nt(PiH + c, AH);
  // loop #2 flux with the new pi, Pi (src/hdo.cpp:950-961)
  double f2[4];
#pragma unroll
  for (int i = 0; i < 4; i++) f2[i] = -tau * (a[sym(0, i)] + A * u[0] * u[i]);
  f2[0] += tau * A;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double t = 0.;
    t += fl[i];
    t += f2[i];
    st_nt(vf + i * N + c, t);
  }
}

// viscous interface flux (Hydro::visc_flux, src/hdo.cpp:966-1014) between
// cells l (left) and r (right) along axis D; returns false if skipped.
template <int D>
__device__ __forceinline__ bool visc_face(const Geom &g, long l, long r, const double *__restrict__ ph,
                                 const double *__restrict__ piH, const double *__restrict__ PiH,
                                 const double *__restrict__ m, int m_full, double tau, double dt,
                                 double F[4]) {
  const long N = g.N;
  if (!m_full && m[(D - 1) * N + l] < 1. && m[(D - 1) * N + r] < 1.) return false;
  const double dxa = D == 1 ? g.dx : (D == 2 ? g.dy : g.dz * (tau + 0.5 * dt));
  double vx = 0.5 * (ph[5 * N + l] + ph[5 * N + r]);
  double vy = 0.5 * (ph[6 * N + l] + ph[6 * N + r]);
  double vz = 0.5 * (ph[7 * N + l] + ph[7 * N + r]);
  const double v = sqrt(vx * vx + vy * vy + vz * vz);
  if (v > 1.) {
    vx = 0.99 * vx / v;
    vy = 0.99 * vy / v;
    vz = 0.99 * vz / v;
  }
  const double gamma = 1. / sqrt(1. - v * v);
  const double u[4] = {gamma, gamma * vx, gamma * vy, gamma * vz};
  const double gm[4] = {1., -1., -1., -1.};
  const double sP = PiH[l] + PiH[r];
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    double f = 0.5 * (piH[sym(mu, D) * N + l] + piH[sym(mu, D) * N + r]);
    if (mu == D) f += -0.5 * sP * gm[mu];
    f += 0.5 * sP * u[mu] * u[D];
    F[mu] = f * (tau - 0.5 * dt) * dt / dxa;
  }
  return true;
}

// viscous interface fluxes gathered per cell in the reference order, the
// viscous source (Hydro::visc_source_step, src/hdo.cpp:355-375) and the
// limited update (Cell::updateByViscFlux, src/cll.cpp:76-84).  Q is updated
// in place (each thread touches only its own cell).
__global__ void __launch_bounds__(256) VH_OCC(VH_W_VISC3)
k_visc3(Geom g, const double *__restrict__ ph, const double *__restrict__ piH,
        const double *__restrict__ PiH, const double *__restrict__ vf,
        const double *__restrict__ m, int m_full, const double *__restrict__ tp, double dt,
        double *__restrict__ Q) {
  const double tau = *tp;
  const long t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int ix, iy, iz;
  strip_cell(g.nx, g.ny, g.nz, VH_STRIP_ROWS, t, ix, iy, iz);
  const long c = g.idx(ix, iy, iz);
  const long N = g.N;
  double acc[4];
#pragma unroll
  for (int k = 0; k < 4; k++) acc[k] = vf[k * N + c];
  double F[4];
  const long sy = g.nx, sz = (long)g.nx * g.ny;
  if (ix > 0 && visc_face<1>(g, c - 1, c, ph, piH, PiH, m, m_full, tau, dt, F))
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += F[k];
  if (ix < g.nx - 1 && visc_face<1>(g, c, c + 1, ph, piH, PiH, m, m_full, tau, dt, F))
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += -F[k];
  if (iy > 0 && visc_face<2>(g, c - sy, c, ph, piH, PiH, m, m_full, tau, dt, F))
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += F[k];
  if (iy < g.ny - 1 && visc_face<2>(g, c, c + sy, ph, piH, PiH, m, m_full, tau, dt, F))
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += -F[k];
  if (iz > 0 && visc_face<3>(g, c - sz, c, ph, piH, PiH, m, m_full, tau, dt, F))
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += F[k];
  if (iz < g.nz - 1 && visc_face<3>(g, c, c + sz, ph, piH, PiH, m, m_full, tau, dt, F))
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += -F[k];
  const double e = ph[c];
  if (e > 0.) {
    const double vx = ph[5 * N + c], vy = ph[6 * N + c], vz = ph[7 * N + c];
    double u[4];
    u[0] = 1. / sqrt(1. - vx * vx - vy * vy - vz * vz);
    u[3] = u[0] * vz;
    double k0 = -piH[sym(3, 3) * N + c] + PiH[c] * (-1.0 - u[3] * u[3]);
    double k3 = -(piH[sym(0, 3) * N + c] + PiH[c] * u[0] * u[3]);
    acc[0] += k0 * dt;
    acc[1] += 0. * dt;
    acc[2] += 0. * dt;
    acc[3] += k3 * dt;
  }
  const double q0 = Q[c];
  if (fabs(acc[0]) <= 0.5 * q0) {
#pragma unroll
    for (int k = 0; k < 4; k++) Q[k * N + c] += acc[k];
#pragma unroll
    for (int k = 4; k < 7; k++) Q[k * N + c] += 0.;
  } else if (acc[0] != 0.) {
    const double fac = fabs(0.5 * q0 / acc[0]);
#pragma unroll
    for (int k = 0; k < 4; k++) Q[k * N + c] += fac * acc[k];
#pragma unroll
    for (int k = 4; k < 7; k++) Q[k * N + c] += fac * 0.;
  }
}

// ------------------------------------------------------- IC and batches
// Cell::setPrimVar + saveQprev + setAllM (src/cll.cpp:222-241,
// src/icGubser.cpp:47-57) on all cells.  prim: 7 arrays e,nb,nq,ns,vx,vy,vz.
template <bool CH>
__global__ void __launch_bounds__(256)
k_set_prim(Geom g, DevEos eos, double tau, const double *__restrict__ pr,
           double *__restrict__ Q, double *__restrict__ Qprev, double *__restrict__ m) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.N) return;
  const long N = g.N;
  const double e = pr[c], nb = pr[N + c], nq = pr[2 * N + c], ns = pr[3 * N + c];
  const double vx = pr[4 * N + c], vy = pr[5 * N + c], vz = pr[6 * N + c];
  const double g2 = 1. / (1 - vx * vx - vy * vy - vz * vz);
  const double p = eos.p<CH>(e, nb, nq, ns);
  double q[7];
  q[T_] = tau * (e + p * (vx * vx + vy * vy + vz * vz)) * g2;
  q[X_] = tau * (e + p) * vx * g2;
  q[Y_] = tau * (e + p) * vy * g2;
  q[Z_] = tau * (e + p) * vz * g2;
  q[NB_] = tau * nb * sqrt(g2);
  q[NQ_] = tau * nq * sqrt(g2);
  q[NS_] = tau * ns * sqrt(g2);
#pragma unroll
  for (int k = 0; k < 7; k++) {
    Q[k * N + c] = q[k];
    Qprev[k * N + c] = q[k];
  }
  if (e > 0.) m[c] = m[N + c] = m[2 * N + c] = 1.;
}

template <bool CH>
__global__ void __launch_bounds__(256)
k_c2p_batch(DevEos eos, double *__restrict__ Q, long n, double *__restrict__ out,
            int *__restrict__ iters) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double q[7];
#pragma unroll
  for (int k = 0; k < 7; k++) q[k] = Q[7 * i + k];
  Prim r;
  const int it = c2p_ref<CH>(eos, q, r);
#pragma unroll
  for (int k = 0; k < 7; k++) Q[7 * i + k] = q[k];
  double *o = out + 8 * i;
  o[0] = r.e;
  o[1] = r.p;
  o[2] = r.nb;
  o[3] = r.nq;
  o[4] = r.ns;
  o[5] = r.vx;
  o[6] = r.vy;
  o[7] = r.vz;
  if (iters) iters[i] = it;
}

__global__ void __launch_bounds__(256)
k_p2c_batch(const double *__restrict__ prim, long n, double *__restrict__ Q) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *a = prim + 8 * i;
  Prim r{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]};
  double q[7];
  p2c(r, q);
#pragma unroll
  for (int k = 0; k < 7; k++) Q[7 * i + k] = q[k];
}

// tau_{n+1} = tau_n + dt on the device (Hydro::performStep, src/hdo.cpp:1095)
// one wave: tau += dt; each of the six signal-speed groups of the step
// (predictor x, y, eta; corrector x, y, eta) is folded over its slots into
// sig[6 + k] and the running maximum sig[12 + axis], and cleared
__global__ void __launch_bounds__(64) k_advance_tau(double *__restrict__ tp, double dt,
                                                    unsigned long long *__restrict__ acc,
                                                    unsigned long long *__restrict__ sig) {
  const int t = threadIdx.x;
  if (t == 0) tp[0] = tp[0] + dt;
  for (int k = 0; k < 6; k++) {
    unsigned long long *slot = acc + k * kSigGroup + t * kSigStride;
    const double v = wave_max(__longlong_as_double((long long)*slot));
    *slot = 0ull;
    if (t == 0) {
      const unsigned long long u = (unsigned long long)__double_as_longlong(v);
      sig[6 + k] = u;
      if (u > sig[12 + k % 3]) sig[12 + k % 3] = u;
    }
  }
}

__global__ void __launch_bounds__(256) k_clear_u64(unsigned long long *__restrict__ p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

__global__ void k_fill_int(int *__restrict__ p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

// ----------------------------------------------------------- launchers
static inline unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

void launch_faces(int D, int mode, const Geom &g, const DevEos &eos, const double *Qs,
                  const double *Qr, const double *m, int m_full, const double *tp, double dt,
                  double *F, unsigned char *flag, int *status, unsigned long long *sig,
                  hipStream_t st) {
  const dim3 b(256), gr(blocks(g.N));
  // k_faces_fix re-evaluates interfaces with a negative face pressure: only
  // the spline table and the hadron table (below e_min) can produce one
  const bool fix = eos.kind == VH_EOS_TABLE || eos.kind == VH_EOS_HADRON;
  const int ek = eos_kind_class(eos);
#define VH_FK(DD, MM, CC, EE)                                                                 \
  if (D == DD && mode == MM && eos.charged() == CC && ek == EE) {                             \
    hipLaunchKernelGGL((k_faces<DD, MM, CC, EE>), gr, b, 0, st, g, eos, Qs, m, m_full, tp,     \
                       dt, F, flag, status, sig);                                             \
    if (fix)                                                                                  \
      hipLaunchKernelGGL((k_faces_fix<DD, MM, CC>), gr, b, 0, st, g, eos, Qr, m, m_full, tp,  \
                         dt, F, flag, status, sig);                                           \
  }
#define VH_F(DD, MM)                                                                          \
  VH_FK(DD, MM, false, VH_EOS_IDEAL) VH_FK(DD, MM, false, VH_EOS_TABLE) VH_FK(DD, MM, true, -1)
  VH_F(1, 0) VH_F(1, 1) VH_F(2, 0) VH_F(2, 1) VH_F(3, 0) VH_F(3, 1)
#undef VH_F
#undef VH_FK
}

void launch_update(int mode, const Geom &g, const DevEos &eos, const double *Q,
                   const double *Qh, const double *const F[3],
                   const unsigned char *const fl[3], const double *tp, double dt, double *out,
                   int *status, hipStream_t st) {
  const int ek = eos_kind_class(eos);
#define VH_U(MM, CC, EE)                                                                      \
  if (mode == MM && eos.charged() == CC && ek == EE)                                          \
    hipLaunchKernelGGL((k_update<MM, CC, EE>), dim3(blocks(g.N)), dim3(256), 0, st, g, eos, Q, \
                       Qh, F[0], fl[0], F[1], fl[1], F[2], fl[2], tp, dt, out, status);
  VH_U(0, false, VH_EOS_IDEAL) VH_U(1, false, VH_EOS_IDEAL) VH_U(0, false, VH_EOS_TABLE)
  VH_U(1, false, VH_EOS_TABLE) VH_U(0, true, -1) VH_U(1, true, -1)
#undef VH_U
}

void launch_ghost(int full, int axis, const Geom &g, double *Q, double *Qh, double *pi,
                  double *Pi, hipStream_t st) {
  const long n = axis == 3 ? (long)g.nx * g.ny : (axis == 2 ? (long)g.nx * g.nz : (long)g.ny * g.nz);
  const dim3 gr(blocks(n)), b(256);
#define VH_G(AA, FF)                                                                          \
  if (axis == AA && full == FF)                                                               \
    hipLaunchKernelGGL((k_ghost<AA, FF>), gr, b, 0, st, g, Q, Qh, pi, Pi);
  VH_G(3, 0) VH_G(2, 0) VH_G(1, 0) VH_G(3, 1) VH_G(2, 1) VH_G(1, 1)
#undef VH_G
}

void launch_update_m(int pass, const Geom &g, double *m, double *dm, const double *tp, double dt,
                     hipStream_t st) {
  if (pass == 0)
    hipLaunchKernelGGL(k_update_m_dm, dim3(blocks(g.N)), dim3(256), 0, st, g, m, dm, tp, dt);
  else
    hipLaunchKernelGGL(k_update_m_apply, dim3(blocks(g.N)), dim3(256), 0, st, g, m, dm);
}

void launch_vprim(const Geom &g, const DevEos &eos, const double *Qh, const double *tp,
                  double dt, double *ph, double *pt, double *Tt, hipStream_t st) {
  const int ek = eos_kind_class(eos);
  if (eos.charged())
    hipLaunchKernelGGL((k_vprim<true, -1>), dim3(blocks(g.N)), dim3(256), 0, st, g, eos, Qh, tp,
                       dt, ph, pt, Tt);
  else if (ek == VH_EOS_TABLE)
    hipLaunchKernelGGL((k_vprim<false, VH_EOS_TABLE>), dim3(blocks(g.N)), dim3(256), 0, st, g, eos,
                       Qh, tp, dt, ph, pt, Tt);
  else
    hipLaunchKernelGGL((k_vprim<false, VH_EOS_IDEAL>), dim3(blocks(g.N)), dim3(256), 0, st, g, eos,
                       Qh, tp, dt, ph, pt, Tt);
}

void launch_visc1(const Geom &g, const DevEos &eos, const DevTrc &trc, const double *Qprev,
                  const double *Qn, const double *ph, const double *pt, const double *m,
                  int m_full, const double *pi, const double *Pi, const double *tp, double dt,
                  double *pi0, double *piH0, double *Pi0, double *PiH0, int *status,
                  double *db, unsigned char *dop, const double *Tt, hipStream_t st) {
  const int ek = eos_kind_class(eos);
#define VH_V(VV, CC, EE)                                                                      \
  if (!!db == VV && eos.charged() == CC && ek == EE)                                          \
    hipLaunchKernelGGL((k_visc1<VV, CC, EE>), dim3(blocks(g.N)), dim3(256), 0, st, g, eos, trc, \
                       Qprev, Qn, ph, pt, m, m_full, pi, Pi, tp, dt, pi0, piH0, Pi0, PiH0,     \
                       status, VV ? db : nullptr, VV ? dop : nullptr, VV ? Tt : nullptr);
  VH_V(true, false, VH_EOS_IDEAL) VH_V(false, false, VH_EOS_IDEAL) VH_V(true, false, VH_EOS_TABLE)
  VH_V(false, false, VH_EOS_TABLE) VH_V(true, true, -1) VH_V(false, true, -1)
#undef VH_V
}

void launch_visc2(const Geom &g, const double *ph, const double *pi0, const double *piH0,
                  const double *Pi0, const double *PiH0, double *pi, double *piH, double *Pi,
                  double *PiH, double *vcc, double *vf, const double *tp, double dt,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_visc2, dim3(blocks(g.N)), dim3(256), 0, st, g, ph, pi0, piH0, Pi0, PiH0,
                     pi, piH, Pi, PiH, vcc, vf, tp, dt);
}

void launch_visc3(const Geom &g, const double *ph, const double *piH, const double *PiH,
                  const double *vf, const double *m, int m_full, const double *tp, double dt,
                  double *Q, hipStream_t st) {
  hipLaunchKernelGGL(k_visc3, dim3(blocks(g.N)), dim3(256), 0, st, g, ph, piH, PiH, vf, m,
                     m_full, tp, dt, Q);
}

void launch_advance_tau(double *tp, double dt, unsigned long long *acc, unsigned long long *sig,
                        hipStream_t st) {
  static_assert(kSigSlots == 64, "one lane per slot");
  hipLaunchKernelGGL(k_advance_tau, dim3(1), dim3(64), 0, st, tp, dt, acc, sig);
}

void launch_clear_u64(unsigned long long *p, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_u64, dim3((unsigned)std::min<long>((n + 255) / 256, 64)), dim3(256),
                     0, st, p, n);
}

void launch_fill_int(int *p, int v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, p, v);
}

void launch_set_prim(const Geom &g, const DevEos &eos, double tau, const double *pr, double *Q,
                     double *Qprev, double *m, hipStream_t st) {
  if (eos.charged())
    hipLaunchKernelGGL(k_set_prim<true>, dim3(blocks(g.N)), dim3(256), 0, st, g, eos, tau, pr, Q,
                       Qprev, m);
  else
    hipLaunchKernelGGL(k_set_prim<false>, dim3(blocks(g.N)), dim3(256), 0, st, g, eos, tau, pr, Q,
                       Qprev, m);
}

// Cell::getPrimVar of every cell (src/cll.cpp:90-96): con2prim of a copy of
// Q/tau, prim[k * N + cell] = e, p, nb, nq, ns, vx, vy, vz
template <bool CH>
__global__ void __launch_bounds__(256)
k_prim_fields(long N, DevEos eos, const double *__restrict__ Q, double tau,
              double *__restrict__ prim) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double q[7];
#pragma unroll
  for (int k = 0; k < 7; k++) q[k] = Q[k * N + c] / tau;
  Prim r;
  c2p_ref<CH>(eos, q, r);
  prim[c] = r.e;
  prim[N + c] = r.p;
  prim[2 * N + c] = r.nb;
  prim[3 * N + c] = r.nq;
  prim[4 * N + c] = r.ns;
  prim[5 * N + c] = r.vx;
  prim[6 * N + c] = r.vy;
  prim[7 * N + c] = r.vz;
}

// Fluid::getCMFvariables (src/fld.cpp:50-60): vx, vy to the lab frame and
// the longitudinal rapidity Y of a cell at eta
__device__ inline void cmf(double eta, double &vx, double &vy, double vz, double &Y) {
  Y = eta + 1. / 2. * log((1. + vz) / (1. - vz));
  vx = vx * cosh(Y - eta) / cosh(Y);
  vy = vy * cosh(Y - eta) / cosh(Y);
}

// The grid totals of outputSurface / outputCorona (src/fld.cpp:545-606,
// 889-945) over the interior cells (ix, iy, iz in [2, n-3]), one partial
// sum per block (FS sums): E, Efull, S, Px, Nb1, Nb2, vt_num, vt_den,
// vxvy_num, vxvy_den, pi0x_num, pi0x_den, txxyy_num, txxyy_den, core cells,
// core cells with the viscous correction cut.  central: the v_T, vx-vy and
// T^xx-T^yy sums only over nz/2-3 < iz < nz/2+3 (outputSurface).
constexpr int FS = 16;
template <typename PT, bool CH>
__global__ void __launch_bounds__(256)
k_fluid_stats(Geom g, DevEos eos, const double *__restrict__ Q, const PT *__restrict__ pi,
              const double *__restrict__ vcc, double tau, double zmin, double ecrit,
              int viscous, int central, double *__restrict__ part) {
  __shared__ double red[256];
  const long N = g.N;
  double acc[FS];
  for (int k = 0; k < FS; k++) acc[k] = 0.;
  const int mx = g.nx - 4, my = g.ny - 4, mz = g.nz - 4;
  const long M = (mx > 0 && my > 0 && mz > 0) ? (long)mx * my * mz : 0;
  for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < M;
       j += (long)gridDim.x * blockDim.x) {
    const int ix = (int)(j % mx) + 2, iy = (int)((j / mx) % my) + 2, iz = (int)(j / ((long)mx * my)) + 2;
    const long c = ix + (long)g.nx * (iy + (long)g.ny * iz);
    double q[7];
    for (int k = 0; k < 7; k++) q[k] = Q[k * N + c] / tau;
    Prim r;
    c2p_ref<CH>(eos, q, r);
    const double e = r.e, nb = r.nb;
    double vx = r.vx, vy = r.vy, Y;
    const double eta = zmin + iz * g.dz;
    cmf(eta, vx, vy, r.vz, Y);
    double t, mub, muq, mus, p;
    eos.eos<CH>(e, nb, r.nq, r.ns, t, mub, muq, mus, p);
    double s = t > 0.0 ? (e + p - mub * nb - muq * r.nq - mus * r.ns) / t : 0.;  // EoS::s
    const double dz = g.dz;
    const double cosh_int = (sinh(eta + 0.5 * dz) - sinh(eta - 0.5 * dz)) / dz;
    const double sinh_int = (cosh(eta + 0.5 * dz) - cosh(eta - 0.5 * dz)) / dz;
    const double th = tanh(Y);
    const double den = 1. - vx * vx - vy * vy - th * th;
    acc[0] += tau * (e + p) / den * (cosh_int - th * sinh_int) - tau * p * cosh_int;
    acc[4] += Q[4 * N + c];
    acc[5] += tau * nb * (cosh_int - th * sinh_int) / sqrt(den);
    double ef = tau * (e + p) / den * (cosh(eta) - th * sinh(eta)) - tau * p * cosh(eta);
    double pic[10];
    for (int k = 0; k < 10; k++) pic[k] = (double)pi[k * N + c];
    if (viscous) ef += tau * pic[sym(0, 0)] * cosh(eta) + tau * pic[sym(0, 3)] * sinh(eta);
    acc[1] += ef;
    if (e > ecrit) {
      acc[14] += 1.;
      if (vcc[c] < 0.9) acc[15] += 1.;
    }
    const double gmm[4] = {1., -1., -1., -1.};
    double deltas = 0.;
    if (viscous)
      for (int i = 0; i < 4; i++)
        for (int k = 0; k < 4; k++) {
          const double x = pic[sym(i, k)];
          deltas += x * x * gmm[i] * gmm[k];
        }
    if (t > 0.02) {
      s += 1.5 * deltas / ((e + p) * t);
      acc[2] += tau * s * (cosh_int - th * sinh_int) / sqrt(den);
    }
    acc[3] += tau * (e + p) * vx / den;
    if (!central || (iz > g.nz / 2 - 3 && iz < g.nz / 2 + 3)) {
      acc[8] += e * (fabs(vx) - fabs(vy));
      acc[9] += e;
      acc[7] += e / sqrt(1. - vx * vx - vy * vy);
      acc[6] += e / sqrt(1. - vx * vx - vy * vy) * sqrt(vx * vx + vy * vy);
      acc[12] += (e + p) / den * (vx * vx - vy * vy);
      acc[13] += (e + p) / den * (vx * vx + vy * vy) + 2. * p;
    }
    acc[10] += e / den * fabs(pic[sym(0, 1)]);
    acc[11] += e / den;
  }
  for (int k = 0; k < FS; k++) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)k * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

__global__ void k_fluid_stats_sum(const double *__restrict__ part, int nb, double *__restrict__ out) {
  const int k = threadIdx.x;
  if (k >= FS) return;
  double a = 0.;
  for (int b = 0; b < nb; b++) a += part[(long)k * nb + b];
  out[k] = a;
}

// selected cells' getPrimVar, pi (10), Pi, viscCorrCut (21 doubles each)
template <typename PT, bool CH>
__global__ void __launch_bounds__(256)
k_cells(long N, DevEos eos, const double *__restrict__ Q, const PT *__restrict__ pi,
        const PT *__restrict__ Pi, const double *__restrict__ vcc, double tau,
        const long *__restrict__ idx, long n, double *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long c = idx[i];
  double q[7];
  for (int k = 0; k < 7; k++) q[k] = Q[k * N + c] / tau;
  Prim r;
  c2p_ref<CH>(eos, q, r);
  double *o = out + 21 * i;
  o[0] = r.e;
  o[1] = r.p;
  o[2] = r.nb;
  o[3] = r.nq;
  o[4] = r.ns;
  o[5] = r.vx;
  o[6] = r.vy;
  o[7] = r.vz;
  for (int k = 0; k < 10; k++) o[8 + k] = (double)pi[k * N + c];
  o[18] = (double)Pi[c];
  o[19] = vcc[c];
  o[20] = 0.;
}

constexpr int FS_BLOCKS = 256;

void launch_fluid_stats(const Geom &g, const DevEos &eos, const double *Q, const void *pi,
                        int pi_float, const double *vcc, double tau, double zmin, double ecrit,
                        int viscous, int central, double *part, double *out, hipStream_t st) {
  if (pi_float)  // FP32 pi: the mixed path, charge-free EoSs only
    hipLaunchKernelGGL((k_fluid_stats<float, false>), dim3(FS_BLOCKS), dim3(256), 0, st, g, eos,
                       Q, (const float *)pi, vcc, tau, zmin, ecrit, viscous, central, part);
  else if (eos.charged())
    hipLaunchKernelGGL((k_fluid_stats<double, true>), dim3(FS_BLOCKS), dim3(256), 0, st, g, eos,
                       Q, (const double *)pi, vcc, tau, zmin, ecrit, viscous, central, part);
  else
    hipLaunchKernelGGL((k_fluid_stats<double, false>), dim3(FS_BLOCKS), dim3(256), 0, st, g, eos,
                       Q, (const double *)pi, vcc, tau, zmin, ecrit, viscous, central, part);
  hipLaunchKernelGGL(k_fluid_stats_sum, dim3(1), dim3(64), 0, st, part, FS_BLOCKS, out);
}

void launch_cells(long N, const DevEos &eos, const double *Q, const void *pi, const void *Pi,
                  int pi_float, const double *vcc, double tau, const long *idx, long n,
                  double *out, hipStream_t st) {
  if (pi_float)  // FP32 pi: the mixed path, charge-free EoSs only
    hipLaunchKernelGGL((k_cells<float, false>), dim3(blocks(n)), dim3(256), 0, st, N, eos, Q,
                       (const float *)pi, (const float *)Pi, vcc, tau, idx, n, out);
  else if (eos.charged())
    hipLaunchKernelGGL((k_cells<double, true>), dim3(blocks(n)), dim3(256), 0, st, N, eos, Q,
                       (const double *)pi, (const double *)Pi, vcc, tau, idx, n, out);
  else
    hipLaunchKernelGGL((k_cells<double, false>), dim3(blocks(n)), dim3(256), 0, st, N, eos, Q,
                       (const double *)pi, (const double *)Pi, vcc, tau, idx, n, out);
}

void launch_prim_fields(long N, const DevEos &eos, const double *Q, double tau, double *prim,
                        hipStream_t st) {
  if (eos.charged())
    hipLaunchKernelGGL(k_prim_fields<true>, dim3(blocks(N)), dim3(256), 0, st, N, eos, Q, tau,
                       prim);
  else
    hipLaunchKernelGGL(k_prim_fields<false>, dim3(blocks(N)), dim3(256), 0, st, N, eos, Q, tau,
                       prim);
}

void launch_c2p_batch(const DevEos &eos, double *Q, long n, double *out, int *iters,
                      hipStream_t st) {
  if (eos.charged())
    hipLaunchKernelGGL(k_c2p_batch<true>, dim3(blocks(n)), dim3(256), 0, st, eos, Q, n, out,
                       iters);
  else
    hipLaunchKernelGGL(k_c2p_batch<false>, dim3(blocks(n)), dim3(256), 0, st, eos, Q, n, out,
                       iters);
}

void launch_p2c_batch(const double *prim, long n, double *Q, hipStream_t st) {
  hipLaunchKernelGGL(k_p2c_batch, dim3(blocks(n)), dim3(256), 0, st, prim, n, Q);
}

}  // namespace vh
