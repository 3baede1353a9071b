// Launch interface of the mixed-precision kernels (vh_kernels_mixed.hip):
// FP64 conserved variables and con2prim, FP32 interface fluxes, viscous
// sources and pi storage.
#pragma once
#include <vector>

#include "vh_device.h"

namespace vh {
namespace mx {

// One or several grids whose cells share the state arrays (grid g's cell 0 at
// ev[g].base, component pitch P = all cells).  Per-cell launches run nblk
// blocks of 256 threads; block b belongs to grid bev[b] and is block
// b - bfirst[bev[b]] of that grid.  Row launches cover nseg segments of 63
// cells of the (x, y) planes; segment s belongs to grid sev[s] and is its
// segment s - sfirst[sev[s]].  Device arrays.
struct Grids {
  const Geom *ev;
  const int *bev, *bfirst, *sev, *sfirst;
  long P;
  int nev, nblk, nseg, maxnz;
  Geom g0;  // grid 0 by value (the kernels use it when nev == 1)
  // cell order of the per-cell kernels: y-strips of sy rows swept over all
  // eta, XCD-contiguous workgroups (0: plain x-fastest order)
  int sy;
  // Band launches (bands of CY = 2, 4, 8 rows, table h = 0, 1, 2): nbseg[h]
  // segments of 63 cells of the band planes (the first rows of the bands laid
  // end to end); segment s belongs to grid bsev[h][s] and is its segment
  // s - bsfirst[h][grid].  Device arrays.
  const int *bsev[3], *bsfirst[3];
  int nbseg[3];
};
// table index of a band height CY (2, 4, 8), -1 for any other
inline int band_table(int CY) { return CY == 2 ? 0 : (CY == 4 ? 1 : (CY == 8 ? 2 : -1)); }

// gv: per-cell con2prim guesses (used by VH_C2P_LEARNED only, may be null otherwise)
// sig: the stage's signal-speed slot groups (x, y, eta, kSigGroup words
// apart; maxima of max(-b_l, b_r) kept as bit patterns, see wave_speed_max)
void launch_stage(int mode, int cm, int nc, const Grids &G, const DevEos &eos, const double *Q,
                  const double *Qh, const float *gv, const double *m, int m_full,
                  const double *tp, double dt, double *out, int *status,
                  unsigned long long *sig, hipStream_t st);
// Stage kernel shapes (same results for all): STAGE_CELLS one thread per
// cell evaluating its six interfaces; STAGE_ROWS waves of 64 flattened (x, y)
// cells marching CZ cells in eta, every x- and eta-interface evaluated once;
// STAGE_BANDS waves of 64 cells of the band plane marching CZ cells in eta
// and walking the CY rows of their bands, every interface evaluated once but
// the y-interfaces between bands and the eta-interfaces between chunks.
enum { STAGE_CELLS = VH_STAGE_CELLS, STAGE_ROWS = VH_STAGE_ROWS, STAGE_BANDS = VH_STAGE_BANDS };
struct Tile {
  int kind, CZ, CY;
};
Tile make_rows(const Grids &G, int CZ);
Tile make_bands(const Grids &G, int CY, int CZ);
Tile make_cells();
// candidates of the autotuner: the cell kernel, rows and bands with several
// eta chunks
std::vector<Tile> tile_candidates(const Grids &G);
void launch_stage_tile(int mode, int cm, int nc, const Grids &G, const DevEos &eos, const double *Q,
                       const double *Qh, const float *gv, const double *m, int m_full,
                       const double *tp, double dt, double *out, int *status,
                       unsigned long long *sig, const Tile &tl, hipStream_t st);
// con2prim of every cell (N cells, pitch N): ph = prims of Qh at tau - dt/2,
// pt = (e, v) of Qh at tau
void launch_vprim(int cm, int nc, long N, const DevEos &eos, const double *Qh, const float *gv,
                  const double *tp, double dt, double *ph, double *pt, hipStream_t st);
void launch_visc1(int cm, int nc, const Grids &G, const DevEos &eos, const DevTrc &trc,
                  const double *Qprev, const double *Qn, const float *gv, const double *ph,
                  const double *pt, const double *m, int m_full, const float *pi,
                  const float *Pi, const double *tp, double dt, float4 *rec, int *status,
                  double *db, unsigned char *dop, hipStream_t st);
// rec: the Israel-Stewart auxiliaries pi0, Pi0, piH0, PiH0 as six planes of
// P float4: (pi0 0-3) (pi0 4-7) (pi0 8-9, Pi0, 0) and the same for piH0
void launch_visc2(const Grids &G, const double *ph, const float4 *rec, float *pi, float *piH,
                  float *Pi, float *PiH, double *vcc, float *vf, const double *tp, double dt,
                  hipStream_t st);
void launch_visc3(const Grids &G, const double *ph, const float *piH, const float *PiH,
                  const float *vf, const double *m, int m_full, const double *tp, double dt,
                  double *Q, hipStream_t st);
// single-pass ghost fill of every grid (each extent 1 or >= 5); degenerate:
// some grid has an extent 1.  full = 0: Q, Qh; full = 1: Q, pi, Pi.  nc: the
// conserved components the states carry (4: the charge components are zero
// and are not copied on a single grid)
void launch_ghost_cells(int full, int nc, const Grids &G, bool degenerate, double *Q, double *Qh,
                        float *pi, float *Pi, hipStream_t st);
// the three-pass fill of one grid (any extents): Q, pi, Pi
void launch_ghost_full(int axis, const Geom &g, double *Q, float *pi, float *Pi, hipStream_t st);

}  // namespace mx
}  // namespace vh
