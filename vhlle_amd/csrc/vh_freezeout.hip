// Freeze-out hypersurface of the hydro state (Fluid::outputSurface,
// src/fld.cpp:517-874, with the Cornelius construction of vh_cornelius.h).
//
// Between a lower time level (the previous step's Q, or a snapshot) and the
// current state, every interior hypercube (ix, iy, iz in [2, n-3), the
// reference's loop range) whose 16 corner energy densities straddle e_crit
// yields surface elements.  Kernels: e at both levels per cell (the
// reference con2prim); a cut flag per hypercube; ordered compaction of the
// cut hypercubes (hipcub, so that elements come out in the reference's
// ix, iy, iz loop order); one thread per cut hypercube for the geometry and
// the element record: quadrilinear Q at the centroid, con2prim, trilinear
// pi, Pi of the upper level, u^mu in the lab frame and dsigma_mu transformed
// from Milne to Cartesian components with element 0's normal (as the
// reference does for every element of a hypercube).  FP64 throughout.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/vhlle_gpu.h"
#include "vh_cornelius.h"
#include "vh_freezeout.h"
#include "vh_vort.h"

namespace vh {
namespace fo {

namespace {

// clears the error word (a kernel, so that no memset is recorded in a graph)
__global__ void k_fill_err(int *__restrict__ err) {
  if (threadIdx.x == 0) *err = 0;
}

// e of every cell at both levels: c2p of Q/tau and of Qlow/tau_low
template <bool CH>
__global__ void __launch_bounds__(256)
k_fo_energy(long N, DevEos eos, const double *__restrict__ Q, const double *__restrict__ Qlow,
            double tau, double tlow, double *__restrict__ e1, double *__restrict__ e0) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double q[7];
  Prim r;
  for (int k = 0; k < 7; k++) q[k] = Q[k * N + c] / tau;
  c2p_ref<CH>(eos, q, r);
  e1[c] = r.e;
  for (int k = 0; k < 7; k++) q[k] = Qlow[k * N + c] / tlow;
  c2p_ref<CH>(eos, q, r);
  e0[c] = r.e;
}

struct CubeGeom {
  int nx, ny, nz, mx, my, mz;  // grid, and interior hypercubes per axis
  long N;
};

__device__ __forceinline__ void cube_origin(const CubeGeom &g, long k, int &ix, int &iy, int &iz) {
  iz = (int)(k % g.mz) + 2;
  const long r = k / g.mz;
  iy = (int)(r % g.my) + 2;
  ix = (int)(r / g.my) + 2;
}

// Hypercubes in memory order (x fastest): j -> origin, and the index of the
// same hypercube in the reference's loop order (ix, iy, iz outer to inner),
// by which the cut ones are sorted afterwards.
__device__ __forceinline__ void cube_origin_mem(const CubeGeom &g, long j, int &ix, int &iy,
                                                int &iz) {
  const unsigned u = (unsigned)j, mx = (unsigned)g.mx, my = (unsigned)g.my;
  const unsigned r = u / mx, q = r / my;
  ix = (int)(u - r * mx) + 2;
  iy = (int)(r - q * my) + 2;
  iz = (int)q + 2;
}
struct RefKey {
  CubeGeom g;
  __host__ __device__ long operator()(long j) const {
    const long mx = g.mx, my = g.my, mz = g.mz;
    const long ix = j % mx, r = j / mx, iy = r % my, iz = r / my;
    return (ix * my + iy) * mz + iz;
  }
};

// cut flag of every hypercube, in memory order (coalesced corner loads)
__global__ void __launch_bounds__(256)
k_fo_flag(CubeGeom g, const double *__restrict__ e1, const double *__restrict__ e0, double E0,
          unsigned char *__restrict__ flag) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)g.mx * g.my * g.mz) return;
  int ix, iy, iz;
  cube_origin_mem(g, k, ix, iy, iz);
  int above = 0;
  for (int jx = 0; jx < 2; jx++)
    for (int jy = 0; jy < 2; jy++)
      for (int jz = 0; jz < 2; jz++) {
        const long c = (ix + jx) + (long)g.nx * ((iy + jy) + (long)g.ny * (iz + jz));
        above += (e1[c] >= E0) + (e0[c] >= E0);
      }
  flag[k] = above > 0 && above < 16;
}

__device__ __forceinline__ void transform_to_lab(double eta, double &vx, double &vy, double &vz) {
  // transformToLab (src/fld.cpp:508-513)
  const double Y = eta + 1. / 2. * log((1. + vz) / (1. - vz));
  vx = vx * cosh(Y - eta) / cosh(Y);
  vy = vy * cosh(Y - eta) / cosh(Y);
  vz = tanh(Y);
}

// pi^{mu nu} at eta from Milne to Cartesian components (src/fld.cpp:765-790)
__device__ inline void cart_pi(const double piC[10], double ch, double sh, double pc[10]) {
  pc[sym(0, 0)] = ch * ch * piC[sym(0, 0)] + 2. * ch * sh * piC[sym(0, 3)] +
                  sh * sh * piC[sym(3, 3)];
  pc[sym(0, 1)] = ch * piC[sym(0, 1)] + sh * piC[sym(3, 1)];
  pc[sym(0, 2)] = ch * piC[sym(0, 2)] + sh * piC[sym(3, 2)];
  pc[sym(0, 3)] = ch * sh * (piC[sym(0, 0)] + piC[sym(3, 3)]) +
                  (ch * ch + sh * sh) * piC[sym(0, 3)];
  pc[sym(1, 1)] = piC[sym(1, 1)];
  pc[sym(1, 2)] = piC[sym(1, 2)];
  pc[sym(1, 3)] = sh * piC[sym(0, 1)] + ch * piC[sym(3, 1)];
  pc[sym(2, 2)] = piC[sym(2, 2)];
  pc[sym(2, 3)] = sh * piC[sym(0, 2)] + ch * piC[sym(3, 2)];
  pc[sym(3, 3)] = sh * sh * piC[sym(0, 0)] + ch * ch * piC[sym(3, 3)] +
                  2. * sh * ch * piC[sym(0, 3)];
}

template <typename PT, bool CH>
__global__ void __launch_bounds__(64)
k_fo_cubes(CubeGeom g, FoParams P, DevEos eos, const long *__restrict__ cut,
           const int *__restrict__ ncut_p, const double *__restrict__ e1,
           const double *__restrict__ e0, const double *__restrict__ Q,
           const double *__restrict__ Qlow, const PT *__restrict__ pi,
           const PT *__restrict__ Pi, const double *__restrict__ db, double *__restrict__ rows,
           int *__restrict__ count, int *__restrict__ err) {
  const int ncut = *ncut_p;
  const long N = g.N;
  const double dxs[4] = {P.dt, P.dx, P.dy, P.dz};
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ncut;
       t += (long)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    cube_origin(g, cut[t], ix, iy, iz);
    double h[2][2][2][2];
    long cc[2][2][2];
    for (int jx = 0; jx < 2; jx++)
      for (int jy = 0; jy < 2; jy++)
        for (int jz = 0; jz < 2; jz++) {
          const long c = (ix + jx) + (long)g.nx * ((iy + jy) + (long)g.ny * (iz + jz));
          cc[jx][jy][jz] = c;
          h[1][jx][jy][jz] = e1[c];
          h[0][jx][jy][jz] = e0[c];
        }
    Corn C;
    double cen[MAXE][4], nrm[MAXE][4];
    const int ne = hypercube(C, h, dxs, P.ecrit, cen, nrm);
    if (C.err) atomicOr(err, C.err);
    count[t] = ne;
    for (int el = 0; el < ne; el++) {
      double *row = rows + ((long)t * MAXE + el) * P.ncol;
      const double *c4 = cen[el];
      row[0] = P.tau + c4[0];
      row[1] = P.xmin + ix * P.dx + c4[1];
      row[2] = P.ymin + iy * P.dy + c4[2];
      row[3] = P.zmin + iz * P.dz + c4[3];
      // quadrilinear Q at the centroid (src/fld.cpp:675-695)
      const double wT[2] = {1. - c4[0] / P.dt, c4[0] / P.dt};
      const double wX[2] = {1. - c4[1] / P.dx, c4[1] / P.dx};
      const double wY[2] = {1. - c4[2] / P.dy, c4[2] / P.dy};
      const double wZ[2] = {1. - c4[3] / P.dz, c4[3] / P.dz};
      double QC[7] = {0., 0., 0., 0., 0., 0., 0.};
      for (int jt = 0; jt < 2; jt++)
        for (int jx = 0; jx < 2; jx++)
          for (int jy = 0; jy < 2; jy++)
            for (int jz = 0; jz < 2; jz++) {
              const double *src = jt ? Q : Qlow;
              const long c = cc[jx][jy][jz];
              for (int i = 0; i < 7; i++)
                QC[i] += src[i * N + c] * wT[jt] * wX[jx] * wY[jy] * wZ[jz];
            }
      for (int i = 0; i < 7; i++) QC[i] = QC[i] / (P.tau + c4[0]);
      Prim pr;
      c2p_ref<CH>(eos, QC, pr);
      const double eC = pr.e, nbC = pr.nb;
      double vxC = pr.vx, vyC = pr.vy, vzC = pr.vz;
      double TC, mubC, muqC, musC, pC;
      eos.eos<CH>(eC, nbC, pr.nq, pr.ns, TC, mubC, muqC, musC, pC);  // src/fld.cpp:695-697
      // trilinear pi, Pi of the upper level (src/fld.cpp:709-728)
      double piC[10] = {0., 0., 0., 0., 0., 0., 0., 0., 0., 0.}, PiC = 0.;
      for (int jx = 0; jx < 2; jx++)
        for (int jy = 0; jy < 2; jy++)
          for (int jz = 0; jz < 2; jz++) {
            const long c = cc[jx][jy][jz];
            for (int ii = 0; ii < 10; ii++)
              piC[ii] += (double)pi[ii * N + c] * wX[jx] * wY[jy] * wZ[jz];
            PiC += (double)Pi[c] * wX[jx] * wY[jy] * wZ[jz];
          }
      double v2C = vxC * vxC + vyC * vyC + vzC * vzC;
      if (v2C > 1.) {
        vxC *= sqrt(0.99 / v2C);
        vyC *= sqrt(0.99 / v2C);
        vzC *= sqrt(0.99 / v2C);
        v2C = 0.99;
      }
      const double etaC = P.zmin + iz * P.dz + c4[3];
      transform_to_lab(etaC, vxC, vyC, vzC);
      const double gammaC = 1. / sqrt(1. - vxC * vxC - vyC * vyC - vzC * vzC);
      const double uC[4] = {gammaC, gammaC * vxC, gammaC * vyC, gammaC * vzC};
      const double tauC = P.tau + c4[0];
      const double ch = cosh(etaC), sh = sinh(etaC);
      const double *n0 = nrm[0];  // element 0's normal for every element (src/fld.cpp:746-751)
      double ds[4];
      ds[0] = tauC * (ch * n0[0] - sh / tauC * n0[3]);
      ds[3] = tauC * (-sh * n0[0] + ch / tauC * n0[3]);
      ds[1] = tauC * n0[1];
      ds[2] = tauC * n0[2];
      for (int i = 0; i < 4; i++) row[4 + i] = ds[i];
      for (int i = 0; i < 4; i++) row[8 + i] = uC[i];
      row[12] = TC;
      row[13] = mubC;
      row[14] = muqC;
      row[15] = musC;
      double pc[10];
      cart_pi(piC, ch, sh, pc);
      for (int i = 0; i < 10; i++) row[16 + i] = pc[i];
      row[26] = PiC;
      const int rec = P.ncol - AUX - P.nbeta;
      if (rec > ROW_BASE) {
        row[27] = eC;
        row[28] = nbC;
      }
      if (P.nbeta) {
        // trilinear dbeta of the upper level's cells (Cell::getDbeta: the
        // clamped copy), then the Milne-to-Cartesian Jacobian of a covariant
        // vector on both indices, the second index lowered by g
        // (src/fld.cpp:703-727,799-831)
        double dI[16];
        for (int k = 0; k < 16; k++) dI[k] = 0.;
        for (int jx = 0; jx < 2; jx++)
          for (int jy = 0; jy < 2; jy++)
            for (int jz = 0; jz < 2; jz++) {
              const long c = cc[jx][jy][jz];
              for (int k = 0; k < 16; k++)
                dI[k] += vort::clamp100(db[k * N + c]) * wX[jx] * wY[jy] * wZ[jz];
            }
        const double jac[4][4] = {{ch, 0., 0., -sh}, {0., 1., 0., 0.}, {0., 0., 1., 0.},
                                  {-sh, 0., 0., ch}};
        const double gmm[4] = {1., -1., -1., -1.};
        for (int i = 0; i < 4; i++)
          for (int j = 0; j < 4; j++) {
            double acc = 0.0;
            for (int k = 0; k < 4; k++)
              for (int l = 0; l < 4; l++) acc += jac[i][k] * jac[j][l] * dI[4 * k + l] * gmm[l];
            row[rec + 4 * i + j] = acc;
          }
      }
      // diagnostics: dVEff, EtotSurf increment, nb flux, e_C
      double dV = 0.0;
      for (int ii = 0; ii < 4; ii++) dV += ds[ii] * uC[ii];
      double dEv = 0.;
      for (int i = 0; i < 4; i++) dEv += pc[sym(0, i)] * ds[i];
      double *aux = row + P.ncol - AUX;
      aux[0] = dV;
      aux[1] = (eC + pC) * uC[0] * dV - pC * ds[0] + dEv;
      aux[2] = nbC * dV;
      aux[3] = eC;
    }
  }
}

// Fluid::outputCorona's selection (src/fld.cpp:946-985): every corner at
// tau has e <= ecrit and the centre state (the corners' mean Q) e >= 0.01
template <bool CH>
__global__ void __launch_bounds__(256)
k_corona_flag(CubeGeom g, DevEos eos, const double *__restrict__ e1,
              const double *__restrict__ Q, double tau, double ecrit,
              unsigned char *__restrict__ flag) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)g.mx * g.my * g.mz) return;
  int ix, iy, iz;
  cube_origin_mem(g, k, ix, iy, iz);
  bool corona = true;
  for (int jx = 0; jx < 2; jx++)
    for (int jy = 0; jy < 2; jy++)
      for (int jz = 0; jz < 2; jz++) {
        const long c = (ix + jx) + (long)g.nx * ((iy + jy) + (long)g.ny * (iz + jz));
        corona = corona && !(e1[c] > ecrit);
      }
  unsigned char f = 0;
  if (corona) {
    double QC[7] = {0., 0., 0., 0., 0., 0., 0.};
    for (int jx = 0; jx < 2; jx++)
      for (int jy = 0; jy < 2; jy++)
        for (int jz = 0; jz < 2; jz++) {
          const long c = (ix + jx) + (long)g.nx * ((iy + jy) + (long)g.ny * (iz + jz));
          for (int i = 0; i < 7; i++) QC[i] += Q[i * g.N + c] * 0.125;
        }
    for (int i = 0; i < 7; i++) QC[i] = QC[i] / tau;
    Prim pr;
    c2p_ref<CH>(eos, QC, pr);
    f = pr.e >= 0.01;
  }
  flag[k] = f;
}

// one corona element per selected cube (src/fld.cpp:986-1066): the record at
// the cube centre with dsigma = tau dx dy dz (cosh eta, 0, 0, -sinh eta),
// then the AUX columns
template <typename PT, bool CH>
__global__ void __launch_bounds__(64)
k_corona_rows(CubeGeom g, FoParams P, DevEos eos, const long *__restrict__ cube,
              const int *__restrict__ n_p, const double *__restrict__ Q,
              const PT *__restrict__ pi, const PT *__restrict__ Pi, double *__restrict__ rows) {
  const int n = *n_p;
  const long N = g.N;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    cube_origin(g, cube[t], ix, iy, iz);
    double QC[7] = {0., 0., 0., 0., 0., 0., 0.}, piC[10], PiC = 0.;
    for (int ii = 0; ii < 10; ii++) piC[ii] = 0.;
    for (int jx = 0; jx < 2; jx++)
      for (int jy = 0; jy < 2; jy++)
        for (int jz = 0; jz < 2; jz++) {
          const long c = (ix + jx) + (long)g.nx * ((iy + jy) + (long)g.ny * (iz + jz));
          for (int i = 0; i < 7; i++) QC[i] += Q[i * N + c] * 0.125;
        }
    for (int i = 0; i < 7; i++) QC[i] = QC[i] / P.tau;
    Prim pr;
    c2p_ref<CH>(eos, QC, pr);
    const double eC = pr.e, nbC = pr.nb;
    double TC, mubC, muqC, musC, pC;
    eos.eos<CH>(eC, nbC, pr.nq, pr.ns, TC, mubC, muqC, musC, pC);  // src/fld.cpp:979-988
    for (int jx = 0; jx < 2; jx++)
      for (int jy = 0; jy < 2; jy++)
        for (int jz = 0; jz < 2; jz++) {
          const long c = (ix + jx) + (long)g.nx * ((iy + jy) + (long)g.ny * (iz + jz));
          for (int ii = 0; ii < 10; ii++) piC[ii] += (double)pi[ii * N + c] * 0.125;
          PiC += (double)Pi[c] * 0.125;
        }
    double vxC = pr.vx, vyC = pr.vy, vzC = pr.vz;
    const double v2C = vxC * vxC + vyC * vyC + vzC * vzC;
    if (v2C > 1.) {
      vxC *= sqrt(0.99 / v2C);
      vyC *= sqrt(0.99 / v2C);
      vzC *= sqrt(0.99 / v2C);
    }
    const double etaC = P.zmin + iz * P.dz + 0.5 * P.dz;
    transform_to_lab(etaC, vxC, vyC, vzC);
    const double gammaC = 1. / sqrt(1. - vxC * vxC - vyC * vyC - vzC * vzC);
    const double uC[4] = {gammaC, gammaC * vxC, gammaC * vyC, gammaC * vzC};
    const double ch = cosh(etaC), sh = sinh(etaC);
    double ds[4];
    ds[0] = P.tau * (ch * P.dx * P.dy * P.dz);
    ds[3] = P.tau * (-sh * P.dx * P.dy * P.dz);
    ds[1] = 0.0;
    ds[2] = 0.0;
    double *row = rows + t * P.ncol;
    row[0] = P.tau;
    row[1] = P.xmin + ix * P.dx + 0.5 * P.dx;
    row[2] = P.ymin + iy * P.dy + 0.5 * P.dy;
    row[3] = P.zmin + iz * P.dz + 0.5 * P.dz;
    for (int i = 0; i < 4; i++) row[4 + i] = ds[i];
    for (int i = 0; i < 4; i++) row[8 + i] = uC[i];
    row[12] = TC;
    row[13] = mubC;
    row[14] = muqC;
    row[15] = musC;
    double pc[10];
    cart_pi(piC, ch, sh, pc);
    for (int i = 0; i < 10; i++) row[16 + i] = pc[i];
    row[26] = PiC;
    if (P.ncol - AUX > ROW_BASE) {
      row[27] = eC;
      row[28] = nbC;
    }
    double dV = 0.0;
    for (int ii = 0; ii < 4; ii++) dV += ds[ii] * uC[ii];
    double dEv = 0.;
    for (int i = 0; i < 4; i++) dEv += pc[sym(0, i)] * ds[i];
    double *aux = row + P.ncol - AUX;
    aux[0] = dV;
    aux[1] = (eC + pC) * uC[0] * dV - pC * ds[0] + dEv;
    aux[2] = nbC * dV;
    aux[3] = eC;
  }
}

// the selected records in element order, cut to max_rows: the first rec
// columns into out (may be NULL) and, with outb, the beta.dat record (the
// first 16 columns, the BETA columns, e_C)
__global__ void __launch_bounds__(256)
k_fo_compact(const int *__restrict__ ncut_p, const int *__restrict__ count,
             const int *__restrict__ off, int per, const double *__restrict__ rows, int ncol,
             int rec, long max_rows, double *__restrict__ out, double *__restrict__ outb) {
  const int ncut = *ncut_p;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ncut;
       t += (long)gridDim.x * blockDim.x) {
    const int ne = count ? count[t] : 1;
    const long o = count ? (long)off[t] : t;
    for (int el = 0; el < ne && o + el < max_rows; el++) {
      const long r = o + el;
      const double *src = rows + (t * per + el) * ncol;
      if (out)
        for (int k = 0; k < rec; k++) out[r * rec + k] = src[k];
      if (outb) {
        double *b = outb + r * BETA_ROW;
        for (int k = 0; k < 16; k++) b[k] = src[k];
        for (int k = 0; k < BETA; k++) b[16 + k] = src[rec + k];
        b[32] = src[ncol - AUX + 3];
      }
    }
  }
}

// sums over all elements in a fixed order: element count, the suspicious
// ones, dsigma_mu, dV_eff, the EtotSurf and nbSurf increments.  Pass 1: block
// b reduces the hypercubes of its chunk [b chunk, (b+1) chunk) into part[b];
// pass 2 (k_fo_stats_sum) adds the parts in block order.
constexpr int STAT_BLOCKS = 512;
__global__ void __launch_bounds__(256)
k_fo_stats(const int *__restrict__ ncut_p, const int *__restrict__ count, int slots,
           const double *__restrict__ rows, int ncol, double ecrit, double *__restrict__ part) {
  __shared__ double red[256];
  const int ncut = *ncut_p;
  const long chunk = (ncut + gridDim.x - 1) / gridDim.x;
  const long t0 = blockIdx.x * chunk, t1 = t0 + chunk < ncut ? t0 + chunk : ncut;
  double acc[NSTAT] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
  for (long t = t0 + threadIdx.x; t < t1; t += blockDim.x)
    for (int el = 0; el < (count ? count[t] : 1); el++) {
      const double *row = rows + ((long)t * slots + el) * ncol;
      const double *aux = row + ncol - AUX;
      acc[0] += 1.;
      acc[1] += (aux[3] > ecrit * 2.0 || aux[3] < ecrit * 0.5) ? 1. : 0.;
      for (int k = 0; k < 4; k++) acc[2 + k] += row[4 + k];
      acc[6] += aux[0];
      acc[7] += aux[1];
      acc[8] += aux[2];
    }
  for (int q = 0; q < NSTAT; q++) {
    red[threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)q * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

__global__ void k_fo_stats_sum(const double *__restrict__ part, int nb, double *__restrict__ out) {
  const int q = threadIdx.x;
  if (q >= NSTAT) return;
  double s = 0.;
  for (int b = 0; b < nb; b++) s += part[(long)q * nb + b];
  out[q] = s;
}

}  // namespace

void Work::release() {
  for (void *p : {(void *)e1, (void *)e0, (void *)rows, (void *)dstat, (void *)flag, (void *)cut,
                  (void *)sorted, (void *)ncut, (void *)cnt, (void *)err, tmp, (void *)off,
                  (void *)crow, (void *)cbeta})
    if (p) (void)hipFree(p);
  *this = Work();
}

namespace {
// (re)allocates *p for n elements of T when the capacity cap is smaller
template <class T>
hipError_t grow(T **p, size_t &cap, size_t n) {
  if (*p && cap >= n) return hipSuccess;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  cap = 0;
  const hipError_t e = hipMalloc(reinterpret_cast<void **>(p), (n ? n : 1) * sizeof(T));
  if (e == hipSuccess) cap = n;
  return e;
}
}  // namespace

namespace {
// the surface (corona false) or the corona (true) elements
int run_impl(const FoArgs &a, Work &w, long &nrows, int &errbits, bool corona) {
  const CubeGeom g{a.nx, a.ny, a.nz, a.nx - 4, a.ny - 4, a.nz - 4, (long)a.nx * a.ny * a.nz};
  nrows = 0;
  errbits = 0;
  for (int q = 0; q < NSTAT; q++) a.stats[q] = 0.;
  if (g.mx < 1 || g.my < 1 || g.mz < 1) return VH_OK;  // no interior hypercubes
  const long ncube = (long)g.mx * g.my * g.mz, N = g.N;
  int rc = VH_OK;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && rc == VH_OK) {
      std::fprintf(stderr, "vhlle_gpu freeze-out: %s\n", hipGetErrorString(e));
      rc = VH_ERR_HIP;
    }
    return rc == VH_OK;
  };
  hipStream_t st = a.stream;
  size_t cap_e0 = w.n_cells, cap_flag = w.n_cubes, cap_small = w.ncut ? 2 : 0,
         cap_err = w.err ? 1 : 0, cap_stat = w.dstat ? (size_t)NSTAT * (STAT_BLOCKS + 1) : 0;
  // the two energy arrays share one capacity (allocated together)
  if (w.n_cells < (size_t)N) {
    size_t c1 = 0;
    ok(grow(&w.e1, c1, N));
    ok(grow(&w.e0, cap_e0, N));
    w.n_cells = rc == VH_OK ? (size_t)N : 0;
  }
  if (w.n_cubes < (size_t)ncube) {
    size_t c2 = 0;
    ok(grow(&w.flag, cap_flag, ncube));
    ok(grow(&w.cut, c2, ncube));
    w.n_cubes = rc == VH_OK ? (size_t)ncube : 0;
  }
  if (!w.ncut) ok(grow(&w.ncut, cap_small, 2));
  if (!w.err) ok(grow(&w.err, cap_err, 1));
  if (!w.dstat) ok(grow(&w.dstat, cap_stat, (size_t)NSTAT * (STAT_BLOCKS + 1)));
  // cut hypercubes are selected in memory order as their reference-order
  // index, then sorted (radix sort over the bits that index needs)
  int endbit = 1;
  while ((1L << endbit) < ncube) endbit++;
  using KeyIt = hipcub::TransformInputIterator<long, RefKey, hipcub::CountingInputIterator<long>>;
  const KeyIt keys(hipcub::CountingInputIterator<long>(0), RefKey{g});
  if (rc == VH_OK) {
    const bool ch = a.eos.charged();
    if (ch)
      hipLaunchKernelGGL(k_fo_energy<true>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st,
                         N, a.eos, a.Q, a.Qlow, a.P.tau, a.P.tau - a.P.dt, w.e1, w.e0);
    else
      hipLaunchKernelGGL(k_fo_energy<false>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st,
                         N, a.eos, a.Q, a.Qlow, a.P.tau, a.P.tau - a.P.dt, w.e1, w.e0);
    if (corona && ch)
      hipLaunchKernelGGL(k_corona_flag<true>, dim3((unsigned)((ncube + 255) / 256)), dim3(256), 0,
                         st, g, a.eos, w.e1, a.Q, a.P.tau, a.P.ecrit, w.flag);
    else if (corona)
      hipLaunchKernelGGL(k_corona_flag<false>, dim3((unsigned)((ncube + 255) / 256)), dim3(256),
                         0, st, g, a.eos, w.e1, a.Q, a.P.tau, a.P.ecrit, w.flag);
    else
      hipLaunchKernelGGL(k_fo_flag, dim3((unsigned)((ncube + 255) / 256)), dim3(256), 0, st, g,
                         w.e1, w.e0, a.P.ecrit, w.flag);
    hipLaunchKernelGGL(k_fill_err, dim3(1), dim3(64), 0, st, w.err);
    size_t tmpb = 0, sortb = 0;
    ok(hipcub::DeviceSelect::Flagged(nullptr, tmpb, keys, w.flag, w.cut, w.ncut, ncube, st));
    ok(hipcub::DeviceRadixSort::SortKeys(nullptr, sortb, w.cut, w.cut, (int)ncube, 0, endbit, st));
    tmpb = std::max(tmpb, sortb);
    if (w.tmp_bytes < tmpb) {
      size_t cap = w.tmp_bytes;
      char *t = static_cast<char *>(w.tmp);
      ok(grow(&t, cap, tmpb));
      w.tmp = t;
      w.tmp_bytes = rc == VH_OK ? tmpb : 0;
    }
    if (rc == VH_OK)
      ok(hipcub::DeviceSelect::Flagged(w.tmp, tmpb, keys, w.flag, w.cut, w.ncut, ncube, st));
  }
  int nc = 0;
  if (rc == VH_OK) {
    ok(hipMemcpyAsync(&nc, w.ncut, sizeof(int), hipMemcpyDeviceToHost, st));
    ok(hipStreamSynchronize(st));
  }
  if (rc == VH_OK && nc > 0) {
    const int per = corona ? 1 : MAXE;  // record slots per selected cube
    const size_t slots = (size_t)nc * per;
    ok(grow(&w.rows, w.n_rows, slots * a.P.ncol));
    ok(grow(&w.cnt, w.n_cnt, (size_t)nc));
    ok(grow(&w.sorted, w.n_sorted, (size_t)nc));
    if (rc == VH_OK) {
      size_t sortb = w.tmp_bytes;
      ok(hipcub::DeviceRadixSort::SortKeys(w.tmp, sortb, w.cut, w.sorted, nc, 0, endbit, st));
    }
    if (rc == VH_OK) {
      const unsigned grid = (unsigned)std::min<long>((nc + 63) / 64, 2048);
      // FP32 pi (the mixed path) comes with a charge-free EoS only
      const bool ch = a.eos.charged();
      if (corona) {
        if (a.pi_float)
          hipLaunchKernelGGL((k_corona_rows<float, false>), dim3(grid), dim3(64), 0, st, g, a.P,
                             a.eos, w.sorted, w.ncut, a.Q, (const float *)a.pi,
                             (const float *)a.Pi, w.rows);
        else if (ch)
          hipLaunchKernelGGL((k_corona_rows<double, true>), dim3(grid), dim3(64), 0, st, g, a.P,
                             a.eos, w.sorted, w.ncut, a.Q, (const double *)a.pi,
                             (const double *)a.Pi, w.rows);
        else
          hipLaunchKernelGGL((k_corona_rows<double, false>), dim3(grid), dim3(64), 0, st, g, a.P,
                             a.eos, w.sorted, w.ncut, a.Q, (const double *)a.pi,
                             (const double *)a.Pi, w.rows);
      } else if (a.pi_float) {
        hipLaunchKernelGGL((k_fo_cubes<float, false>), dim3(grid), dim3(64), 0, st, g, a.P, a.eos,
                           w.sorted, w.ncut, w.e1, w.e0, a.Q, a.Qlow, (const float *)a.pi,
                           (const float *)a.Pi, a.db, w.rows, w.cnt, w.err);
      } else if (ch) {
        hipLaunchKernelGGL((k_fo_cubes<double, true>), dim3(grid), dim3(64), 0, st, g, a.P, a.eos,
                           w.sorted, w.ncut, w.e1, w.e0, a.Q, a.Qlow, (const double *)a.pi,
                           (const double *)a.Pi, a.db, w.rows, w.cnt, w.err);
      } else {
        hipLaunchKernelGGL((k_fo_cubes<double, false>), dim3(grid), dim3(64), 0, st, g, a.P, a.eos,
                           w.sorted, w.ncut, w.e1, w.e0, a.Q, a.Qlow, (const double *)a.pi,
                           (const double *)a.Pi, a.db, w.rows, w.cnt, w.err);
      }
      ok(hipGetLastError());
      hipLaunchKernelGGL(k_fo_stats, dim3(STAT_BLOCKS), dim3(256), 0, st, w.ncut,
                         corona ? (const int *)nullptr : w.cnt, per, w.rows, a.P.ncol,
                         a.P.ecrit, w.dstat);
      double *sums = w.dstat + (size_t)NSTAT * STAT_BLOCKS;
      hipLaunchKernelGGL(k_fo_stats_sum, dim3(1), dim3(64), 0, st, w.dstat, STAT_BLOCKS, sums);
      ok(hipMemcpyAsync(a.stats, sums, NSTAT * sizeof(double), hipMemcpyDeviceToHost, st));
      ok(hipMemcpyAsync(&errbits, w.err, sizeof(int), hipMemcpyDeviceToHost, st));
      ok(hipStreamSynchronize(st));
    }
    if (rc == VH_OK) nrows = (long)a.stats[0];
    // the requested records, compacted on the device into the caller's layout
    const bool wr = a.rows_out != nullptr, wb = a.beta_out != nullptr && a.P.nbeta;
    const long nout = std::min(nrows, a.max_rows);
    if (rc == VH_OK && (wr || wb) && nout > 0) {
      if (!corona) {  // element offsets of the cubes (exclusive sum of the counts)
        ok(grow(&w.off, w.n_off, (size_t)nc));
        size_t scanb = 0;
        ok(hipcub::DeviceScan::ExclusiveSum(nullptr, scanb, w.cnt, w.off, nc, st));
        if (rc == VH_OK && w.tmp_bytes < scanb) {
          size_t cap = w.tmp_bytes;
          char *t = static_cast<char *>(w.tmp);
          ok(grow(&t, cap, scanb));
          w.tmp = t;
          w.tmp_bytes = rc == VH_OK ? scanb : 0;
        }
        if (rc == VH_OK) {
          scanb = w.tmp_bytes;
          ok(hipcub::DeviceScan::ExclusiveSum(w.tmp, scanb, w.cnt, w.off, nc, st));
        }
      }
      if (wr) ok(grow(&w.crow, w.n_crow, (size_t)nout * a.rec));
      if (wb) ok(grow(&w.cbeta, w.n_cbeta, (size_t)nout * BETA_ROW));
      if (rc == VH_OK) {
        const unsigned grid = (unsigned)std::min<long>((nc + 255) / 256, 2048);
        hipLaunchKernelGGL(k_fo_compact, dim3(grid), dim3(256), 0, st, w.ncut,
                           corona ? (const int *)nullptr : w.cnt, w.off, per, w.rows, a.P.ncol,
                           a.rec, nout, wr ? w.crow : nullptr, wb ? w.cbeta : nullptr);
        ok(hipGetLastError());
        if (wr)
          ok(hipMemcpyAsync(a.rows_out, w.crow, nout * a.rec * sizeof(double),
                            hipMemcpyDeviceToHost, st));
        if (wb)
          ok(hipMemcpyAsync(a.beta_out, w.cbeta, nout * BETA_ROW * sizeof(double),
                            hipMemcpyDeviceToHost, st));
        ok(hipStreamSynchronize(st));
      }
    }
  }
  if (rc != VH_OK) w.release();  // leave no half-grown buffer behind
  return rc;
}
}  // namespace

int run(const FoArgs &a, Work &w, long &nrows, int &errbits) {
  return run_impl(a, w, nrows, errbits, false);
}

int run_corona(const FoArgs &a, Work &w, long &nrows) {
  int errbits = 0;
  return run_impl(a, w, nrows, errbits, true);
}

namespace {
__global__ void __launch_bounds__(64)
k_cornelius_batch(const double *__restrict__ h, long n, Dx4 dx, double E0,
                  double *__restrict__ cen, double *__restrict__ nrm, int *__restrict__ nel) {
  const double d[4] = {dx.v[0], dx.v[1], dx.v[2], dx.v[3]};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    double hc[2][2][2][2], c[MAXE][4], m[MAXE][4];
    for (int k = 0; k < 16; k++) (&hc[0][0][0][0])[k] = h[i * 16 + k];
    Corn C;
    const int ne = hypercube(C, hc, d, E0, c, m);
    nel[i] = C.err ? -C.err : ne;
    for (int e = 0; e < MAXE; e++)
      for (int k = 0; k < 4; k++) {
        cen[(i * MAXE + e) * 4 + k] = e < ne ? c[e][k] : 0.;
        nrm[(i * MAXE + e) * 4 + k] = e < ne ? m[e][k] : 0.;
      }
  }
}
}  // namespace

int cornelius_batch(const double *h, long n, const double *dx, double E0, double *cen,
                    double *nrm, int *nel) {
  double *dh = nullptr, *dc = nullptr, *dn = nullptr;
  int *de = nullptr;
  int rc = VH_OK;
  if (hipMalloc(&dh, n * 16 * sizeof(double)) != hipSuccess ||
      hipMalloc(&dc, n * MAXE * 4 * sizeof(double)) != hipSuccess ||
      hipMalloc(&dn, n * MAXE * 4 * sizeof(double)) != hipSuccess ||
      hipMalloc(&de, n * sizeof(int)) != hipSuccess ||
      hipMemcpy(dh, h, n * 16 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = VH_ERR_HIP;
  if (rc == VH_OK) {
    const Dx4 d{{dx[0], dx[1], dx[2], dx[3]}};
    hipLaunchKernelGGL(k_cornelius_batch, dim3((unsigned)std::min<long>((n + 63) / 64, 4096)),
                       dim3(64), 0, 0, dh, n, d, E0, dc, dn, de);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(cen, dc, n * MAXE * 4 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(nrm, dn, n * MAXE * 4 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(nel, de, n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      rc = VH_ERR_HIP;
  }
  if (dh) (void)hipFree(dh);
  if (dc) (void)hipFree(dc);
  if (dn) (void)hipFree(dn);
  if (de) (void)hipFree(de);
  return rc;
}

}  // namespace fo
}  // namespace vh
