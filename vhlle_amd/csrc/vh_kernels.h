// Launch interface of the hydro-step kernels (vh_kernels.hip).  All launches
// are asynchronous on the given stream; tp points to the device-resident
// proper time tau so that one captured step graph can be replayed.
#pragma once
#include "vh_device.h"

namespace vh {

// Qs: the stage's state; Qr: Q (states of the negative-pressure
// re-evaluation, src/hdo.cpp:134-140); sig: the signal-speed slot group of
// this stage and axis (see wave_speed_max)
void launch_faces(int D, int mode, const Geom &g, const DevEos &eos, const double *Qs,
                  const double *Qr, const double *m, int m_full, const double *tp, double dt,
                  double *F, unsigned char *flag, int *status, unsigned long long *sig,
                  hipStream_t st);
void launch_update(int mode, const Geom &g, const DevEos &eos, const double *Q,
                   const double *Qh, const double *const F[3],
                   const unsigned char *const fl[3], const double *tp, double dt, double *out,
                   int *status, hipStream_t st);
void launch_ghost(int full, int axis, const Geom &g, double *Q, double *Qh, double *pi,
                  double *Pi, hipStream_t st);
// single-pass ghost fill (valid when ghost_fusable): Q and Qh (full = 0) or
// Q, pi, Pi (full = 1); pi_float selects FP32 pi storage
bool ghost_fusable(const Geom &g);
void launch_ghost1(int full, int pi_float, const Geom &g, double *Q, double *Qh, void *pi,
                   void *Pi, hipStream_t st);
void launch_update_m(int pass, const Geom &g, double *m, double *dm, const double *tp, double dt,
                     hipStream_t st);
void launch_vprim(const Geom &g, const DevEos &eos, const double *Qh, const double *tp,
                  double dt, double *ph, double *pt, double *Tt, hipStream_t st);
void launch_visc1(const Geom &g, const DevEos &eos, const DevTrc &trc, const double *Qprev,
                  const double *Qn, const double *ph, const double *pt, const double *m,
                  int m_full, const double *pi, const double *Pi, const double *tp, double dt,
                  double *pi0, double *piH0, double *Pi0, double *PiH0, int *status,
                  double *db, unsigned char *dop, const double *Tt, hipStream_t st);
void launch_visc2(const Geom &g, const double *ph, const double *pi0, const double *piH0,
                  const double *Pi0, const double *PiH0, double *pi, double *piH, double *Pi,
                  double *PiH, double *vcc, double *vf, const double *tp, double dt,
                  hipStream_t st);
void launch_visc3(const Geom &g, const double *ph, const double *piH, const double *PiH,
                  const double *vf, const double *m, int m_full, const double *tp, double dt,
                  double *Q, hipStream_t st);
// tau += dt on the device, and the step's signal speeds: the six slot
// groups of acc (kSigGroup each: predictor x, y, eta, corrector x, y, eta)
// are folded into sig[6..11] and the running maxima sig[12..14] (per axis),
// and cleared
void launch_advance_tau(double *tp, double dt, unsigned long long *acc, unsigned long long *sig,
                        hipStream_t st);
// clears n 64-bit words from p (a kernel: graph-capturable)
void launch_clear_u64(unsigned long long *p, long n, hipStream_t st);
void launch_fill_int(int *p, int v, hipStream_t st);
void launch_set_prim(const Geom &g, const DevEos &eos, double tau, const double *pr, double *Q,
                     double *Qprev, double *m, hipStream_t st);
void launch_c2p_batch(const DevEos &eos, double *Q, long n, double *out, int *iters,
                      hipStream_t st);
void launch_p2c_batch(const double *prim, long n, double *Q, hipStream_t st);
void launch_prim_fields(long N, const DevEos &eos, const double *Q, double tau, double *prim,
                        hipStream_t st);
// grid totals of the reference's output routines (16 sums, see vh_kernels.hip);
// part: 16 * 256 doubles of scratch, out: 16 doubles (device)
void launch_fluid_stats(const Geom &g, const DevEos &eos, const double *Q, const void *pi,
                        int pi_float, const double *vcc, double tau, double zmin, double ecrit,
                        int viscous, int central, double *part, double *out, hipStream_t st);
void launch_cells(long N, const DevEos &eos, const double *Q, const void *pi, const void *Pi,
                  int pi_float, const double *vcc, double tau, const long *idx, long n,
                  double *out, hipStream_t st);

}  // namespace vh
