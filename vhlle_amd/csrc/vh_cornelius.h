// Iso-surface elements of one 4-D hypercube (2^4 corners) at the value E0:
// the marching-hypercube construction of Cornelius (src/cornelius.cpp:780-2205,
// Huovinen & Petersen), written for one GPU thread per hypercube with
// fixed-size arrays instead of the reference's object graph.
//
// Construction (same steps, same branch rules and floating-point operation
// order as the reference, so that FP64 results agree to rounding):
//  * the hypercube's 8 boundary cubes (constant axis i = 0..3, side j = 0, 1,
//    in that order) and each cube's 6 squares (second constant axis k != i,
//    ascending, side 0, 1);
//  * a square whose corners are not all on one side of E0 gets its edge
//    crossings (linear interpolation; a corner equal to E0 is moved 1e-9 of
//    an edge inside) and one line per crossing pair, with an "outside" point
//    in the low-value region; four crossings are resolved by the mean of the
//    corners (src/cornelius.cpp:1092-1298);
//  * a cube's lines form one polygon, or, when a square was ambiguous or the
//    cube has 6 lines, closed polygons chained end to start (tolerance 1e-10,
//    a line is reversed to continue the chain) (src/cornelius.cpp:1492-1589);
//  * the polygons form one polyhedron, or, when a cube was ambiguous or the
//    hypercube has 24 lines and 2 (or 14) corners below E0, polyhedra grown
//    from polygons that share a line (src/cornelius.cpp:1783-1893);
//  * a polyhedron's centroid is the volume-weighted centroid of its
//    tetrahedra (line start, line end, polygon centroid, mean of the points)
//    and its normal the sum of the tetrahedra's normals, each oriented toward
//    the outside point of its line (src/cornelius.cpp:806-989).
// Coordinates are relative to the hypercube's corner 0, in units of dx[4].
#pragma once
#include <hip/hip_runtime.h>

namespace vh {
namespace fo {

constexpr int MAXL = 96;   // lines of one hypercube: 8 cubes x 6 squares x 2
constexpr int MAXPG = 80;  // polygons (8 cubes x at most 10)
constexpr int MAXE = 10;   // polyhedra = surface elements (src/cornelius.h:247)

// error bits of a hypercube
enum { ERR_CUTS = 1, ERR_POLYGON = 2, ERR_OVERFLOW = 4 };

struct Corn {
  // lines: start/end points (after reversals), outside point
  double p0[MAXL][4], p1[MAXL][4], out[MAXL][4];
  int nl;
  // polygons: lines in chain order, constant axis, centroid
  int pg_line[MAXL], pg_first[MAXPG], pg_n[MAXPG], pg_const[MAXPG];
  double pg_c[MAXPG][4];
  int npg;
  // polyhedra: member polygons
  int ph_pg[MAXPG], ph_first[MAXE], ph_n[MAXE], ph_ntet[MAXE];
  int nph;
  int err;
};

__host__ __device__ inline double fab(double x) { return x < 0 ? -x : x; }

// one square: corner values v[a][b] (a along ax1, b along ax2), constant axes
// c0, c1 at values cv0, cv1; appends its lines (src/cornelius.cpp:1092-1298)
__host__ __device__ inline void square_lines(Corn &C, const double v[2][2], int c0, int c1,
                                             double cv0, double cv1, const double dx[4],
                                             double E0, bool &ambiguous) {
  ambiguous = false;
  int above = 0;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      if (v[i][j] >= E0) above++;
  if (above == 0 || above == 4) return;
  int x1 = -1, x2 = -1;
  for (int i = 0; i < 4; i++)
    if (i != c0 && i != c1) {
      if (x1 < 0) x1 = i;
      else x2 = i;
    }
  double cut[4][2], outp[2][2];
  int nc = 0;
  // edges in the reference's order: (00-10), (00-01), (10-11), (01-11)
  auto edge = [&](double a, double b, double fixed_x, double fixed_y, int along) {
    // a at the edge start, b at its end; along = 0: x1 varies, 1: x2 varies
    const double len = along == 0 ? dx[x1] : dx[x2];
    double t = -1.;
    if ((a - E0) * (b - E0) < 0)
      t = (a - E0) / (a - b) * len;
    else if (a == E0 && b < E0)
      t = 1e-9 * len;
    else if (b == E0 && a < E0)
      t = (1.0 - 1e-9) * len;
    if (t >= 0. || ((a - E0) * (b - E0) < 0)) {
      if (along == 0) {
        cut[nc][0] = t;
        cut[nc][1] = fixed_y;
      } else {
        cut[nc][0] = fixed_x;
        cut[nc][1] = t;
      }
      nc++;
    }
  };
  edge(v[0][0], v[1][0], 0., 0., 0);
  edge(v[0][0], v[0][1], 0., 0., 1);
  edge(v[1][0], v[1][1], dx[x1], 0., 1);
  edge(v[0][1], v[1][1], 0., dx[x2], 0);
  if (nc != 0 && nc != 2 && nc != 4) {
    C.err |= ERR_CUTS;
    return;
  }
  if (nc == 4) {  // ambiguous: decided by the mean of the corners
    ambiguous = true;
    double Emid = 0;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) Emid += 0.25 * v[i][j];
    if ((v[0][0] < E0 && Emid < E0) || (v[0][0] > E0 && Emid > E0))
      for (int i = 0; i < 2; i++) {
        const double t = cut[1][i];
        cut[1][i] = cut[2][i];
        cut[2][i] = t;
      }
    if ((Emid - E0) < 0) {
      for (int i = 0; i < 2; i++) {
        outp[i][0] = 0.5 * dx[x1];
        outp[i][1] = 0.5 * dx[x2];
      }
    } else if ((v[0][0] - E0) < 0) {
      outp[0][0] = 0;
      outp[0][1] = 0;
      outp[1][0] = dx[x1];
      outp[1][1] = dx[x2];
    } else {
      outp[0][0] = dx[x1];
      outp[0][1] = 0;
      outp[1][0] = 0;
      outp[1][1] = dx[x2];
    }
  } else if (nc > 0) {  // the mean of the corners below E0
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) outp[i][j] = 0;
    int nout = 0;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        if (v[i][j] < E0) {
          outp[0][0] += i * dx[x1];
          outp[0][1] += j * dx[x2];
          nout++;
        }
    if (nout > 0)
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) outp[i][j] = outp[i][j] / double(nout);
  }
  for (int l = 0; l < nc / 2; l++) {
    if (C.nl >= MAXL) {
      C.err |= ERR_OVERFLOW;
      return;
    }
    double *a = C.p0[C.nl], *b = C.p1[C.nl], *o = C.out[C.nl];
    a[x1] = cut[2 * l][0];
    a[x2] = cut[2 * l][1];
    a[c0] = cv0;
    a[c1] = cv1;
    b[x1] = cut[2 * l + 1][0];
    b[x2] = cut[2 * l + 1][1];
    b[c0] = cv0;
    b[c1] = cv1;
    o[x1] = outp[l][0];
    o[x2] = outp[l][1];
    o[c0] = cv0;
    o[c1] = cv1;
    C.nl++;
  }
}

__host__ __device__ inline void swap_line(Corn &C, int l) {
  for (int k = 0; k < 4; k++) {
    const double t = C.p0[l][k];
    C.p0[l][k] = C.p1[l][k];
    C.p1[l][k] = t;
  }
}

// Polygon::add_line with the chain check (src/cornelius.cpp:416-448)
__host__ __device__ inline bool chain_line(Corn &C, int pg, int l) {
  const double eps = 1e-10;
  const int n = C.pg_n[pg];
  if (n == 0) {
    C.pg_line[C.pg_first[pg] + n] = l;
    C.pg_n[pg]++;
    return true;
  }
  const double *p3 = C.p1[C.pg_line[C.pg_first[pg] + n - 1]];
  double d1 = 0, d2 = 0;
  for (int j = 0; j < 4; j++) {
    d1 += fab(C.p0[l][j] - p3[j]);
    d2 += fab(C.p1[l][j] - p3[j]);
  }
  if (d1 < eps || d2 < eps) {
    if (d2 < eps) swap_line(C, l);
    C.pg_line[C.pg_first[pg] + n] = l;
    C.pg_n[pg]++;
    return true;
  }
  return false;
}

// one cube: constant axis ci at value cv, corner values u[a][b][c] over the
// other three axes in ascending order; appends polygons; returns whether the
// cube is ambiguous and its line count (src/cornelius.cpp:1443-1589)
__host__ __device__ inline void cube_polygons(Corn &C, const double u[2][2][2], int ci,
                                              double cv, const double dx[4], double E0,
                                              bool &amb, int &nlines) {
  int ax[3], n = 0;
  for (int i = 0; i < 4; i++)
    if (i != ci) ax[n++] = i;
  const int l0 = C.nl;
  int namb = 0;
  for (int s = 0; s < 3; s++) {  // second constant axis = ax[s]
    for (int j = 0; j < 2; j++) {
      double v[2][2];
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
          if (s == 0) v[a][b] = u[j][a][b];
          else if (s == 1) v[a][b] = u[a][j][b];
          else v[a][b] = u[a][b][j];
        }
      bool sa = false;
      square_lines(C, v, ci, ax[s], cv, j * dx[ax[s]], dx, E0, sa);
      if (sa) namb++;
    }
  }
  nlines = C.nl - l0;
  amb = false;
  if (nlines == 0) return;
  if (namb == 0 && nlines == 6) namb++;
  amb = namb > 0;
  const int first = l0;  // polygons of this cube list lines from pg_line[first]
  int pos = first;
  if (amb) {
    bool used[12];
    for (int i = 0; i < nlines; i++) used[i] = false;
    int nused = 0;
    do {
      if (nlines - nused < 3 || C.npg >= MAXPG) {
        C.err |= ERR_POLYGON;
        return;
      }
      const int pg = C.npg;
      C.pg_first[pg] = pos;
      C.pg_n[pg] = 0;
      C.pg_const[pg] = ci;
      for (int i = 0; i < nlines; i++) {
        if (!used[i] && chain_line(C, pg, l0 + i)) {
          used[i] = true;
          nused++;
          i = 0;  // the reference restarts its scan at index 1
        }
      }
      pos += C.pg_n[pg];
      C.npg++;
    } while (nused < nlines);
  } else {
    if (C.npg >= MAXPG) {
      C.err |= ERR_POLYGON;
      return;
    }
    const int pg = C.npg;
    C.pg_first[pg] = pos;
    C.pg_n[pg] = nlines;
    C.pg_const[pg] = ci;
    for (int i = 0; i < nlines; i++) C.pg_line[pos + i] = l0 + i;
    C.npg++;
  }
}

// Polyhedron::lines_equal: only the second line's start point is compared
// (src/cornelius.cpp:753-769)
__host__ __device__ inline bool lines_equal(const Corn &C, int l1, int l2) {
  const double eps = 1e-10;
  double d1 = 0, d2 = 0;
  for (int i = 0; i < 4; i++) {
    d1 += fab(C.p0[l1][i] - C.p0[l2][i]);
    d2 += fab(C.p1[l1][i] - C.p0[l2][i]);
    if (d1 > eps && d2 > eps) return false;
  }
  return true;
}

__host__ __device__ inline bool add_polygon(Corn &C, int ph, int pg, bool nocheck) {
  if (nocheck || C.ph_n[ph] == 0) {
    C.ph_pg[C.ph_first[ph] + C.ph_n[ph]] = pg;
    C.ph_n[ph]++;
    C.ph_ntet[ph] += C.pg_n[pg];
    return true;
  }
  for (int i = 0; i < C.ph_n[ph]; i++) {
    const int q = C.ph_pg[C.ph_first[ph] + i];
    for (int j = 0; j < C.pg_n[pg]; j++)
      for (int k = 0; k < C.pg_n[q]; k++)
        if (lines_equal(C, C.pg_line[C.pg_first[pg] + j], C.pg_line[C.pg_first[q] + k])) {
          C.ph_pg[C.ph_first[ph] + C.ph_n[ph]] = pg;
          C.ph_n[ph]++;
          C.ph_ntet[ph] += C.pg_n[pg];
          return true;
        }
  }
  return false;
}

// Polygon centroid (src/cornelius.cpp:456-536); a, b, c: the free axes
__host__ __device__ inline void polygon_centroid(Corn &C, int pg) {
  int x[3], n = 0;
  for (int i = 0; i < 4; i++)
    if (i != C.pg_const[pg]) x[n++] = i;
  const int x1 = x[0], x2 = x[1], x3 = x[2];
  const int nl = C.pg_n[pg], f = C.pg_first[pg];
  double mean[4] = {0, 0, 0, 0};
  for (int i = 0; i < nl; i++) {
    const int l = C.pg_line[f + i];
    for (int j = 0; j < 4; j++) mean[j] += C.p0[l][j] + C.p1[l][j];
  }
  for (int j = 0; j < 4; j++) mean[j] = mean[j] / double(2.0 * nl);
  if (nl == 3) {
    for (int j = 0; j < 4; j++) C.pg_c[pg][j] = mean[j];
    return;
  }
  double up[4] = {0, 0, 0, 0}, down = 0;
  for (int i = 0; i < nl; i++) {
    const int l = C.pg_line[f + i];
    const double *p1 = C.p0[l], *p2 = C.p1[l];
    double cm[4], a[4], b[4];
    for (int j = 0; j < 4; j++) cm[j] = (p1[j] + p2[j] + mean[j]) / 3.0;
    for (int j = 0; j < 4; j++) {
      a[j] = p1[j] - mean[j];
      b[j] = p2[j] - mean[j];
    }
    const double t1 = a[x2] * b[x3] - a[x3] * b[x2];
    const double t2 = a[x1] * b[x3] - a[x3] * b[x1];
    const double t3 = a[x2] * b[x1] - a[x1] * b[x2];
    const double A = 0.5 * sqrt(t1 * t1 + t2 * t2 + t3 * t3);
    for (int j = 0; j < 4; j++) up[j] += cm[j] * A;
    down += A;
  }
  for (int j = 0; j < 4; j++) C.pg_c[pg][j] = up[j] / down;
}

// 4-D normal of the tetrahedron (0, a, b, c) (src/cornelius.cpp:780-797)
__host__ __device__ inline void tetra(const double *a, const double *b, const double *c,
                                      double *n) {
  const double bc01 = b[0] * c[1] - b[1] * c[0];
  const double bc02 = b[0] * c[2] - b[2] * c[0];
  const double bc03 = b[0] * c[3] - b[3] * c[0];
  const double bc12 = b[1] * c[2] - b[2] * c[1];
  const double bc13 = b[1] * c[3] - b[3] * c[1];
  const double bc23 = b[2] * c[3] - b[3] * c[2];
  n[0] = 1.0 / 6.0 * (a[1] * bc23 - a[2] * bc13 + a[3] * bc12);
  n[1] = -1.0 / 6.0 * (a[0] * bc23 - a[2] * bc03 + a[3] * bc02);
  n[2] = 1.0 / 6.0 * (a[0] * bc13 - a[1] * bc03 + a[3] * bc01);
  n[3] = -1.0 / 6.0 * (a[0] * bc12 - a[1] * bc02 + a[2] * bc01);
}

// centroid and normal of polyhedron ph (src/cornelius.cpp:806-989)
__host__ __device__ inline void polyhedron_element(const Corn &C, int ph, double cen[4],
                                                   double nrm[4]) {
  const int f = C.ph_first[ph], np = C.ph_n[ph];
  double mean[4] = {0, 0, 0, 0};
  for (int i = 0; i < np; i++) {
    const int pg = C.ph_pg[f + i];
    for (int j = 0; j < C.pg_n[pg]; j++) {
      const int l = C.pg_line[C.pg_first[pg] + j];
      for (int k = 0; k < 4; k++) mean[k] += C.p0[l][k] + C.p1[l][k];
    }
  }
  for (int k = 0; k < 4; k++) mean[k] = mean[k] / double(2.0 * C.ph_ntet[ph]);
  double up[4] = {0, 0, 0, 0}, down = 0;
  for (int i = 0; i < np; i++) {
    const int pg = C.ph_pg[f + i];
    const double *cent = C.pg_c[pg];
    for (int j = 0; j < C.pg_n[pg]; j++) {
      const int l = C.pg_line[C.pg_first[pg] + j];
      const double *p1 = C.p0[l], *p2 = C.p1[l];
      double cm[4], a[4], b[4], c[4], n[4];
      for (int k = 0; k < 4; k++) cm[k] = (p1[k] + p2[k] + cent[k] + mean[k]) / 4.0;
      for (int k = 0; k < 4; k++) {
        a[k] = p1[k] - mean[k];
        b[k] = p2[k] - mean[k];
        c[k] = cent[k] - mean[k];
      }
      tetra(a, b, c, n);
      double V = 0;
      for (int k = 0; k < 4; k++) V += n[k] * n[k];
      V = sqrt(V);
      for (int k = 0; k < 4; k++) up[k] += cm[k] * V;
      down += V;
    }
  }
  for (int k = 0; k < 4; k++) cen[k] = up[k] / down;
  for (int k = 0; k < 4; k++) nrm[k] = 0;
  double sum[4] = {0, 0, 0, 0};
  for (int i = 0; i < np; i++) {
    const int pg = C.ph_pg[f + i];
    const double *cent = C.pg_c[pg];
    for (int j = 0; j < C.pg_n[pg]; j++) {
      const int l = C.pg_line[C.pg_first[pg] + j];
      double a[4], b[4], c[4], n[4], vo[4];
      for (int k = 0; k < 4; k++) {
        a[k] = C.p0[l][k] - cen[k];
        b[k] = C.p1[l][k] - cen[k];
        c[k] = cent[k] - cen[k];
      }
      tetra(a, b, c, n);
      for (int k = 0; k < 4; k++) vo[k] = C.out[l][k] - cen[k];
      double dot = 0;
      for (int k = 0; k < 4; k++) dot += vo[k] * n[k];
      if (dot < 0)
        for (int k = 0; k < 4; k++) n[k] = -n[k];
      for (int k = 0; k < 4; k++) sum[k] += n[k];
    }
  }
  for (int k = 0; k < 4; k++) nrm[k] = sum[k];
}

// All surface elements of hypercube h[t][x][y][z] at E0; returns the element
// count (0..MAXE) and the centroids/normals; C.err != 0 flags a hypercube the
// construction could not resolve (the reference exits there).
__host__ __device__ inline int hypercube(Corn &C, const double h[2][2][2][2], const double dx[4],
                                         double E0, double cen[MAXE][4], double nrm[MAXE][4]) {
  C.nl = C.npg = C.nph = 0;
  C.err = 0;
  int above = 0;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int k = 0; k < 2; k++)
        for (int l = 0; l < 2; l++)
          if (h[i][j][k][l] >= E0) above++;
  if (above == 0 || above == 16) return 0;
  int namb = 0, nlines = 0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 2; j++) {
      double u[2][2][2];
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
          for (int c = 0; c < 2; c++) {
            if (i == 0) u[a][b][c] = h[j][a][b][c];
            else if (i == 1) u[a][b][c] = h[a][j][b][c];
            else if (i == 2) u[a][b][c] = h[a][b][j][c];
            else u[a][b][c] = h[a][b][c][j];
          }
      bool amb = false;
      int nl = 0;
      cube_polygons(C, u, i, j * dx[i], dx, E0, amb, nl);
      if (amb) namb++;
      nlines += nl;
      if (C.err) return 0;
    }
  if (namb == 0) {
    int pts = 0;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        for (int k = 0; k < 2; k++)
          for (int l = 0; l < 2; l++)
            if (h[i][j][k][l] < E0) pts++;
    if (pts > 8) pts = 16 - pts;
    if (nlines == 24 && pts == 2) namb++;
  }
  for (int p = 0; p < C.npg; p++) polygon_centroid(C, p);
  if (namb > 0) {
    bool used[MAXPG];
    for (int i = 0; i < C.npg; i++) used[i] = false;
    int nused = 0, pos = 0;
    do {
      if (C.nph >= MAXE) {
        C.err |= ERR_OVERFLOW;
        return 0;
      }
      const int ph = C.nph;
      C.ph_first[ph] = pos;
      C.ph_n[ph] = 0;
      C.ph_ntet[ph] = 0;
      for (int i = 0; i < C.npg; i++) {
        if (!used[i] && add_polygon(C, ph, i, false)) {
          used[i] = true;
          nused++;
          i = 0;  // the reference restarts its scan at index 1
        }
      }
      pos += C.ph_n[ph];
      C.nph++;
    } while (nused < C.npg);
  } else {
    C.ph_first[0] = 0;
    C.ph_n[0] = 0;
    C.ph_ntet[0] = 0;
    for (int i = 0; i < C.npg; i++) add_polygon(C, 0, i, true);
    C.nph = 1;
  }
  for (int e = 0; e < C.nph; e++) polyhedron_element(C, e, cen[e], nrm[e]);
  return C.nph;
}

}  // namespace fo
}  // namespace vh
