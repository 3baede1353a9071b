// Launch interface of the mixed-precision kernels (vh_kernels_mixed.hip):
// FP64 conserved variables and con2prim, FP32 interface fluxes, viscous
// sources and pi storage.
#pragma once
#include <vector>

#include "vh_device.h"

namespace vh {
namespace mx {

// gv: per-cell con2prim guesses (used by VH_C2P_LEARNED only, may be null otherwise)
void launch_stage(int mode, int cm, const Geom &g, const DevEos &eos, const double *Q,
                  const double *Qh, const float *gv, const double *m, int m_full,
                  const double *tp, double dt, double *out, int *status, hipStream_t st);
// Stage kernel shapes (same results for all): STAGE_CELLS one thread per
// cell evaluating its six interfaces; STAGE_ROWS waves of 64 flattened (x, y)
// cells marching CZ cells in eta (ntx segments, ntz chunks); STAGE_TILES
// blocks of TX x TY columns sharing interfaces through LDS, marching CZ cells
// (hx/hy = 1 when the axis has interfaces; ntx, nty, ntz tiles per axis).
enum { STAGE_CELLS = VH_STAGE_CELLS, STAGE_ROWS = VH_STAGE_ROWS, STAGE_TILES = VH_STAGE_TILES };
struct Tile {
  int kind, TX, TY, CZ, hx, hy, ntx, nty, ntz;
};
Tile make_tile(const Geom &g, int TX, int TY, int CZ);
Tile make_rows(const Geom &g, int CZ);
Tile make_cells();
bool tile_ok(const Tile &t);
// candidates of the autotuner: the cell kernel, rows with several eta chunks,
// the three block tiles that waste the fewest threads
std::vector<Tile> tile_candidates(const Geom &g);
void launch_stage_tile(int mode, int cm, const Geom &g, const DevEos &eos, const double *Q,
                       const double *Qh, const float *gv, const double *m, int m_full,
                       const double *tp, double dt, double *out, int *status, const Tile &tl,
                       hipStream_t st);
void launch_vprim(int cm, const Geom &g, const DevEos &eos, const double *Qh, const float *gv,
                  const double *tp, double dt, double *ph, double *pt, hipStream_t st);
void launch_visc1(int cm, const Geom &g, const DevEos &eos, const DevTrc &trc,
                  const double *Qprev, const double *Qn, const float *gv, const double *ph,
                  const double *pt, const double *m, int m_full, const float *pi,
                  const float *Pi, const double *tp, double dt, float4 *rec, int *status,
                  hipStream_t st);
// rec: the Israel-Stewart auxiliaries pi0, Pi0, piH0, PiH0 as six planes of
// N float4: (pi0 0-3) (pi0 4-7) (pi0 8-9, Pi0, 0) and the same for piH0
void launch_visc2(const Geom &g, const double *ph, const float4 *rec, float *pi, float *piH,
                  float *Pi, float *PiH, double *vcc, float *vf, const double *tp, double dt,
                  hipStream_t st);
void launch_visc3(const Geom &g, const double *ph, const float *piH, const float *PiH,
                  const float *vf, const double *m, int m_full, const double *tp, double dt,
                  double *Q, hipStream_t st);
void launch_ghost_full(int axis, const Geom &g, double *Q, float *pi, float *Pi, hipStream_t st);

}  // namespace mx
}  // namespace vh
