// Host interface of the freeze-out kernels (vh_freezeout.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "vh_device.h"

namespace vh {
namespace fo {

// columns of one surface element: the record of the reference's
// freezeout.dat (src/fld.cpp:656-797 with OUTPI: centroid tau, x, y, eta;
// dsigma_mu (4); u^mu (4); T, mu_B, mu_Q, mu_S; pi^{mu nu} Cartesian (10);
// Pi) = 27 columns, 29 with the extended output (e, n_b), then AUX columns:
// dV_eff = dsigma_mu u^mu, the element's EtotSurf and nbSurf increments, e.
constexpr int ROW_BASE = 27, ROW_EXT = 29, AUX = 4;
// sums over the elements: count, suspicious, dsigma_mu (4), dV_eff,
// EtotSurf and nbSurf increments
constexpr int NSTAT = 9;

// With the vorticity output, BETA columns follow the record (before AUX):
// d_i beta_j in Cartesian coordinates, 4 i + j (src/fld.cpp:799-831).
constexpr int BETA = 16;

struct FoParams {
  double tau, dt, dx, dy, dz, xmin, ymin, zmin, ecrit;
  int ncol;   // record columns + nbeta + AUX
  int nbeta;  // 0, or BETA with the vorticity output
};

struct Dx4 {
  double v[4];
};

struct FoArgs {
  int nx, ny, nz;
  DevEos eos;
  FoParams P;
  const double *Q, *Qlow;     // upper (tau) and lower (tau - dt) levels, 7 x N
  const void *pi, *Pi;        // pi (10 x N), Pi (N) of the upper level
  int pi_float;               // pi, Pi stored in FP32
  const double *db;           // the cells' dbeta (16 x N, clamped on read) with nbeta
  hipStream_t stream;
  // host outputs (each may be NULL): the first max_rows records' rec columns,
  // and their beta.dat records (BETA_ROW columns, with nbeta)
  double *rows_out, *beta_out;
  long max_rows;
  int rec;
  double *stats;              // host, NSTAT sums
};
constexpr int BETA_ROW = 33;  // a beta.dat record

// Device buffers of run(), kept between calls (a context owns one) and grown
// when a call needs more; release() frees them.
struct Work {
  double *e1 = nullptr, *e0 = nullptr, *rows = nullptr, *dstat = nullptr;
  double *crow = nullptr, *cbeta = nullptr;  // compacted output records
  unsigned char *flag = nullptr;
  long *cut = nullptr, *sorted = nullptr;
  int *ncut = nullptr, *cnt = nullptr, *err = nullptr, *off = nullptr;
  void *tmp = nullptr;
  size_t n_cells = 0, n_cubes = 0, n_rows = 0, n_cnt = 0, n_sorted = 0, tmp_bytes = 0,
         n_off = 0, n_crow = 0, n_cbeta = 0;
  void release();
};

// elements in the reference's loop order; nrows = their count
int run(const FoArgs &a, Work &w, long &nrows, int &errbits);
// Fluid::outputCorona's elements at P.tau from Q (Qlow = Q, P.dt = 0, nbeta 0)
int run_corona(const FoArgs &a, Work &w, long &nrows);
// the construction alone on n hypercubes h (n x 16, [t][x][y][z]) on the GPU
int cornelius_batch(const double *h, long n, const double *dx, double E0, double *cen,
                    double *nrm, int *nel);

}  // namespace fo
}  // namespace vh
