// Thermal-vorticity gradient dbeta[i][j] = d_i (u^j / T) (the reference's
// vorticity option: Hydro::NSquant, src/hdo.cpp:420-611; Cell::setDbeta,
// src/cll.h:173-183).
//
// The reference forms dbeta in ONE 4x4 matrix shared by all cells of an
// ISformal pass (src/hdo.cpp:685-687, 723-726) and copies it into each cell.
// NSquant overwrites only part of it for some cells: nothing when it returns
// at the vacuum check (src/hdo.cpp:399-414), not row d when the +-d
// neighbours are not both e > 0, and it ADDS the Christoffel terms to [3][0]
// and [3][3] whatever row 3 held.  A cell's value therefore depends on the
// cells before it in the loop order (ix, iy, iz, iz fastest).  The relaxation
// kernel records per cell the fresh entries and a one-byte op code; entries
// that depend on earlier cells are resolved afterwards (vh_vort.hip) by a
// column walk in loop order per (ix, iy) column, a scan over the columns and
// the chains of ADD-only columns, which reproduces the sequential values
// exactly (the adds are applied one by one in the reference's order).
//
// Storage: db[k * N + cell], k = 4 i + j, the unclamped shared matrix after
// the cell (after resolution); readers apply setDbeta's clamp to [-100, 100].
#pragma once
#include <hip/hip_runtime.h>

#include "vh_device.h"

namespace vh {
namespace vort {

// op code of a cell (one byte)
enum : unsigned char {
  K_EMPTY = 0,  // e <= 0: Cell::resetDbeta; the shared matrix is untouched
  K_STALE = 1,  // NSquant returned at the vacuum check: the cell copies the matrix
  K_FULL = 2,   // NSquant formed row 0, and rows 1-3 where flagged
  F_R1 = 4,     // row 1 formed (the +-x neighbours have e > 0)
  F_R2 = 8,     // row 2 formed
  F_R3 = 16,    // row 3 formed
  F_ADD = 32    // the Christoffel terms were added to [3][0], [3][3] (T > 0)
};
constexpr unsigned char K_MASK = 3;

// entry groups that share one fate per cell: rows 0, 1, 2, the entries [3][1],
// [3][2], and the entries [3][0], [3][3] (which also receive the adds)
constexpr int NGROUP = 5;
enum GroupOp { G_KEEP = 0, G_SET = 1, G_ADD = 2 };

__host__ __device__ inline int group_first(int gidx) {
  return gidx < 3 ? 4 * gidx : (gidx == 3 ? 13 : 12);
}
// the entries of group gidx: n of them, k-th at entry(gidx, k)
__host__ __device__ inline int group_size(int gidx) { return gidx < 3 ? 4 : 2; }
__host__ __device__ inline int group_entry(int gidx, int k) {
  if (gidx < 3) return 4 * gidx + k;
  if (gidx == 3) return 13 + k;
  return k ? 15 : 12;
}

// what a cell with op code o does to group gidx
__host__ __device__ inline int group_op(unsigned char o, int gidx) {
  if ((o & K_MASK) != K_FULL) return G_KEEP;
  switch (gidx) {
    case 0: return G_SET;
    case 1: return (o & F_R1) ? G_SET : G_KEEP;
    case 2: return (o & F_R2) ? G_SET : G_KEEP;
    case 3: return (o & F_R3) ? G_SET : G_KEEP;
    default: return (o & F_R3) ? G_SET : ((o & F_ADD) ? G_ADD : G_KEEP);
  }
}

// true if the cell's stored matrix needs no earlier cell
__host__ __device__ inline bool self_contained(unsigned char o) {
  const int k = o & K_MASK;
  return k == K_EMPTY || (k == K_FULL && (o & (F_R1 | F_R2 | F_R3)) == (F_R1 | F_R2 | F_R3));
}

// row 0 (src/hdo.cpp:451-458): a0, a1 = u^nu of the states at tau - dt and
// tau, T0, T1 their temperatures, eH / e0 the H-centre and previous e
__device__ inline void row_time(double r[4], const double a0[4], const double a1[4], double T0,
                                double T1, double eH, double e0, double dt) {
#pragma unroll
  for (int k = 0; k < 4; k++) r[k] = (a1[k] / T1 - a0[k] / T0) / dt;
  if (eH <= 0. || e0 <= 0. || T1 <= 0. || T0 <= 0.) r[0] = r[1] = r[2] = r[3] = 0.;
}

// rows 1-3 (src/hdo.cpp:495-503,541-549,587-595), on the branch where both
// neighbours have e > 0: A, B = u^nu of the + and - neighbours, den = dx, dy
// or dz (tau + dt/2)
__device__ inline void row_space(double r[4], const double A[4], const double B[4], double TA,
                                 double TB, double den) {
#pragma unroll
  for (int k = 0; k < 4; k++) r[k] = 0.5 * (A[k] / TA - B[k] / TB) / den;
  if (TA <= 0. || TB <= 0.) r[0] = r[1] = r[2] = r[3] = 0.;
}

// writes a cell's op code and the entries its op sets (v) or adds (v[12],
// v[15] hold the adds when row 3 is not formed)
__device__ inline void store(double *__restrict__ db, unsigned char *__restrict__ op, long N,
                             long c, unsigned char o, const double v[16]) {
  op[c] = o;
  const int kind = o & K_MASK;
  if (kind == K_EMPTY) {
#pragma unroll
    for (int k = 0; k < 16; k++) db[k * N + c] = 0.;
    return;
  }
  if (kind != K_FULL) return;
#pragma unroll
  for (int gi = 0; gi < NGROUP; gi++) {
    const int gop = group_op(o, gi);
    if (gop == G_KEEP) continue;
#pragma unroll
    for (int k = 0; k < group_size(gi); k++) {
      const int e = group_entry(gi, k);
      db[e * N + c] = v[e];
    }
  }
}

// setDbeta's clamp (std::min then std::max; NaN passes)
__host__ __device__ inline double clamp100(double x) {
  x = (100. < x) ? 100. : x;
  return (x < -100.) ? -100. : x;
}

}  // namespace vort

// resolves the entries that depend on earlier cells (single grid nx*ny*nz);
// scratch: vort_scratch_bytes(g) bytes, 256-aligned
size_t vort_scratch_bytes(const Geom &g);
void vort_resolve(const Geom &g, double *db, const unsigned char *op, void *scratch,
                  hipStream_t st);
// clamps db to [-100, 100] in place (the stored values of Cell::setDbeta)
void vort_clamp(double *db, long n, hipStream_t st);

}  // namespace vh
