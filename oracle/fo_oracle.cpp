// CPU oracle of the freeze-out hypersurface -- test infrastructure only.
//
// A serial restatement of Fluid::outputSurface (src/fld.cpp:517-874) and of
// the Cornelius construction (src/cornelius.cpp:780-2205) in the reference's
// own object structure: lines owned by squares, polygons holding pointers to
// lines (reversing them in place while chaining), polyhedra holding pointers
// to polygons, lazily computed centroids and normals.  Same branch rules and
// floating-point operation order as the reference; built with
// -ffp-contract=off like the rest of the oracle.
//
// Pinned by tests/test_fo_oracle.py: planar level sets (exact centroid and
// normal of the cross-section), the closure of the surface of a 4-D ball
// (the elements' normals sum to zero), mirror symmetry and the ambiguous
// saddle configurations' element counts.
#include <cmath>
#include <cstring>
#include <array>
#include <utility>
#include <vector>

#include "vh_oracle.h"

namespace {

const int DIM = 4;

struct Line {
  double p[2][DIM];
  double out[DIM];
  int s = 0, e = 1;  // start / end point index (swapped by flip)
  void set(const double a[DIM], const double b[DIM], const double o[DIM]) {
    for (int j = 0; j < DIM; j++) {
      p[0][j] = a[j];
      p[1][j] = b[j];
      out[j] = o[j];
    }
    s = 0;
    e = 1;
  }
  const double *start() const { return p[s]; }
  const double *end() const { return p[e]; }
  void flip() { std::swap(s, e); }
};

struct Polygon {
  std::vector<Line *> lines;
  int cst = 0, x1 = 0, x2 = 0, x3 = 0;
  bool have_c = false;
  double c[DIM];
  void reset(int const_axis) {
    cst = const_axis;
    int k = 0, ax[3];
    for (int i = 0; i < DIM; i++)
      if (i != const_axis) ax[k++] = i;
    x1 = ax[0];
    x2 = ax[1];
    x3 = ax[2];
    lines.clear();
    have_c = false;
  }
  // Polygon::add_line (src/cornelius.cpp:416-448)
  bool add(Line *l, bool nocheck) {
    const double eps = 1e-10;
    if (nocheck || lines.empty()) {
      lines.push_back(l);
      return true;
    }
    const double *p1 = l->start(), *p2 = l->end(), *p3 = lines.back()->end();
    double d1 = 0, d2 = 0;
    for (int j = 0; j < DIM; j++) {
      d1 += std::fabs(p1[j] - p3[j]);
      d2 += std::fabs(p2[j] - p3[j]);
    }
    if (d1 < eps || d2 < eps) {
      if (d2 < eps) l->flip();
      lines.push_back(l);
      return true;
    }
    return false;
  }
  // area-weighted centroid (src/cornelius.cpp:456-536)
  const double *centroid() {
    if (have_c) return c;
    const int n = (int)lines.size();
    double mean[DIM] = {0, 0, 0, 0};
    for (Line *l : lines)
      for (int j = 0; j < DIM; j++) mean[j] += l->start()[j] + l->end()[j];
    for (int j = 0; j < DIM; j++) mean[j] = mean[j] / double(2.0 * n);
    if (n == 3) {
      for (int j = 0; j < DIM; j++) c[j] = mean[j];
      have_c = true;
      return c;
    }
    double up[DIM] = {0, 0, 0, 0}, down = 0;
    for (Line *l : lines) {
      const double *p1 = l->start(), *p2 = l->end();
      double cm[DIM], a[DIM], b[DIM];
      for (int j = 0; j < DIM; j++) cm[j] = (p1[j] + p2[j] + mean[j]) / 3.0;
      for (int j = 0; j < DIM; j++) {
        a[j] = p1[j] - mean[j];
        b[j] = p2[j] - mean[j];
      }
      const double A = 0.5 * std::sqrt(std::pow(a[x2] * b[x3] - a[x3] * b[x2], 2.0) +
                                       std::pow(a[x1] * b[x3] - a[x3] * b[x1], 2.0) +
                                       std::pow(a[x2] * b[x1] - a[x1] * b[x2], 2.0));
      for (int j = 0; j < DIM; j++) up[j] += cm[j] * A;
      down += A;
    }
    for (int j = 0; j < DIM; j++) c[j] = up[j] / down;
    have_c = true;
    return c;
  }
};

void tetravolume(const double *a, const double *b, const double *c, double *n) {
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

struct Polyhedron {
  std::vector<Polygon *> pg;
  int ntet = 0;
  // lines_equal (src/cornelius.cpp:753-769): l2's start against both ends of l1
  static bool same(const Line *l1, const Line *l2) {
    const double eps = 1e-10;
    double d1 = 0, d2 = 0;
    for (int i = 0; i < DIM; i++) {
      d1 += std::fabs(l1->start()[i] - l2->start()[i]);
      d2 += std::fabs(l1->end()[i] - l2->start()[i]);
      if (d1 > eps && d2 > eps) return false;
    }
    return true;
  }
  bool add(Polygon *p, bool nocheck) {
    if (nocheck || pg.empty()) {
      pg.push_back(p);
      ntet += (int)p->lines.size();
      return true;
    }
    for (Polygon *q : pg)
      for (Line *a : p->lines)
        for (Line *b : q->lines)
          if (same(a, b)) {
            pg.push_back(p);
            ntet += (int)p->lines.size();
            return true;
          }
    return false;
  }
  void element(double cen[DIM], double nrm[DIM]) {
    double mean[DIM] = {0, 0, 0, 0};
    for (Polygon *q : pg)
      for (Line *l : q->lines)
        for (int k = 0; k < DIM; k++) mean[k] += l->start()[k] + l->end()[k];
    for (int k = 0; k < DIM; k++) mean[k] = mean[k] / double(2.0 * ntet);
    double up[DIM] = {0, 0, 0, 0}, down = 0;
    for (Polygon *q : pg) {
      const double *ct = q->centroid();
      for (Line *l : q->lines) {
        const double *p1 = l->start(), *p2 = l->end();
        double cm[DIM], a[DIM], b[DIM], c[DIM], n[DIM];
        for (int k = 0; k < DIM; k++) cm[k] = (p1[k] + p2[k] + ct[k] + mean[k]) / 4.0;
        for (int k = 0; k < DIM; k++) {
          a[k] = p1[k] - mean[k];
          b[k] = p2[k] - mean[k];
          c[k] = ct[k] - mean[k];
        }
        tetravolume(a, b, c, n);
        double V = 0;
        for (int k = 0; k < DIM; k++) V += n[k] * n[k];
        V = std::sqrt(V);
        for (int k = 0; k < DIM; k++) up[k] += cm[k] * V;
        down += V;
      }
    }
    for (int k = 0; k < DIM; k++) cen[k] = up[k] / down;
    std::vector<std::array<double, DIM>> ns;
    for (Polygon *q : pg) {
      const double *ct = q->centroid();
      for (Line *l : q->lines) {
        double a[DIM], b[DIM], c[DIM], vo[DIM];
        std::array<double, DIM> n;
        for (int k = 0; k < DIM; k++) {
          a[k] = l->start()[k] - cen[k];
          b[k] = l->end()[k] - cen[k];
          c[k] = ct[k] - cen[k];
        }
        tetravolume(a, b, c, n.data());
        for (int k = 0; k < DIM; k++) vo[k] = l->out[k] - cen[k];
        double dot = 0;
        for (int k = 0; k < DIM; k++) dot += vo[k] * n[k];
        if (dot < 0)
          for (int k = 0; k < DIM; k++) n[k] = -n[k];
        ns.push_back(n);
      }
    }
    for (int k = 0; k < DIM; k++) nrm[k] = 0;
    for (auto &n : ns)
      for (int k = 0; k < DIM; k++) nrm[k] += n[k];
  }
};

// Square: lines of one face (src/cornelius.cpp:1056-1298)
struct Square {
  double v[2][2];
  int c0, c1, x1, x2;
  double cv0, cv1;
  const double *dx;
  Line lines[2];
  int nlines = 0;
  bool ambiguous = false;
  void init(const double sq[2][2], int ci0, int ci1, double cval0, double cval1,
            const double *d) {
    std::memcpy(v, sq, sizeof v);
    c0 = ci0;
    c1 = ci1;
    cv0 = cval0;
    cv1 = cval1;
    dx = d;
    x1 = x2 = -1;
    for (int i = 0; i < DIM; i++)
      if (i != c0 && i != c1) {
        if (x1 < 0) x1 = i;
        else x2 = i;
      }
    nlines = 0;
    ambiguous = false;
  }
  bool build(double E0) {
    int above = 0;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        if (v[i][j] >= E0) above++;
    if (above == 0 || above == 4) return true;
    double cuts[4][2], outp[2][2];
    int nc = 0;
    // ends_of_edge (src/cornelius.cpp:1148-1211)
    if ((v[0][0] - E0) * (v[1][0] - E0) < 0) {
      cuts[nc][0] = (v[0][0] - E0) / (v[0][0] - v[1][0]) * dx[x1];
      cuts[nc++][1] = 0;
    } else if (v[0][0] == E0 && v[1][0] < E0) {
      cuts[nc][0] = 1e-9 * dx[x1];
      cuts[nc++][1] = 0;
    } else if (v[1][0] == E0 && v[0][0] < E0) {
      cuts[nc][0] = (1.0 - 1e-9) * dx[x1];
      cuts[nc++][1] = 0;
    }
    if ((v[0][0] - E0) * (v[0][1] - E0) < 0) {
      cuts[nc][0] = 0;
      cuts[nc++][1] = (v[0][0] - E0) / (v[0][0] - v[0][1]) * dx[x2];
    } else if (v[0][0] == E0 && v[0][1] < E0) {
      cuts[nc][0] = 0;
      cuts[nc++][1] = 1e-9 * dx[x2];
    } else if (v[0][1] == E0 && v[0][0] < E0) {
      cuts[nc][0] = 0;
      cuts[nc++][1] = (1.0 - 1e-9) * dx[x2];
    }
    if ((v[1][0] - E0) * (v[1][1] - E0) < 0) {
      cuts[nc][0] = dx[x1];
      cuts[nc++][1] = (v[1][0] - E0) / (v[1][0] - v[1][1]) * dx[x2];
    } else if (v[1][0] == E0 && v[1][1] < E0) {
      cuts[nc][0] = dx[x1];
      cuts[nc++][1] = 1e-9 * dx[x2];
    } else if (v[1][1] == E0 && v[1][0] < E0) {
      cuts[nc][0] = dx[x1];
      cuts[nc++][1] = (1.0 - 1e-9) * dx[x2];
    }
    if ((v[0][1] - E0) * (v[1][1] - E0) < 0) {
      cuts[nc][0] = (v[0][1] - E0) / (v[0][1] - v[1][1]) * dx[x1];
      cuts[nc++][1] = dx[x2];
    } else if (v[0][1] == E0 && v[1][1] < E0) {
      cuts[nc][0] = 1e-9 * dx[x1];
      cuts[nc++][1] = dx[x2];
    } else if (v[1][1] == E0 && v[0][1] < E0) {
      cuts[nc][0] = (1.0 - 1e-9) * dx[x1];
      cuts[nc++][1] = dx[x2];
    }
    if (nc != 0 && nc != 2 && nc != 4) return false;
    // find_outside (src/cornelius.cpp:1218-1298)
    if (nc == 4) {
      ambiguous = true;
      double Emid = 0;
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) Emid += 0.25 * v[i][j];
      if ((v[0][0] < E0 && Emid < E0) || (v[0][0] > E0 && Emid > E0))
        for (int i = 0; i < 2; i++) std::swap(cuts[1][i], cuts[2][i]);
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
    } else if (nc > 0) {
      std::memset(outp, 0, sizeof outp);
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
    for (int i = 0; i < nc; i += 2) {
      double a[DIM], b[DIM], o[DIM];
      a[x1] = cuts[i][0];
      a[x2] = cuts[i][1];
      b[x1] = cuts[i + 1][0];
      b[x2] = cuts[i + 1][1];
      o[x1] = outp[i / 2][0];
      o[x2] = outp[i / 2][1];
      a[c0] = b[c0] = o[c0] = cv0;
      a[c1] = b[c1] = o[c1] = cv1;
      lines[nlines++].set(a, b, o);
    }
    return true;
  }
};

// Cube: squares -> polygons (src/cornelius.cpp:1405-1589)
struct Cube {
  double u[2][2][2];
  int ci;
  double cv;
  const double *dx;
  Square sq[6];
  std::vector<Line *> lines;
  Polygon pgs[8];
  int npg = 0;
  int ambiguous = 0;
  bool build(double E0) {
    int ax[3], n = 0;
    for (int i = 0; i < DIM; i++)
      if (i != ci) ax[n++] = i;
    int ns = 0;
    for (int s = 0; s < 3; s++)
      for (int j = 0; j < 2; j++) {
        double v[2][2];
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++)
            v[a][b] = s == 0 ? u[j][a][b] : (s == 1 ? u[a][j][b] : u[a][b][j]);
        sq[ns].init(v, ci, ax[s], cv, j * dx[ax[s]], dx);
        if (!sq[ns].build(E0)) return false;
        ns++;
      }
    lines.clear();
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < sq[i].nlines; j++) lines.push_back(&sq[i].lines[j]);
    npg = 0;
    ambiguous = 0;
    const int nl = (int)lines.size();
    if (nl == 0) return true;
    for (int i = 0; i < 6; i++)
      if (sq[i].ambiguous) ambiguous++;
    if (ambiguous == 0 && nl == 6) ambiguous++;
    if (ambiguous > 0) {
      std::vector<int> not_used(nl, 1);
      int used = 0;
      do {
        if (nl - used < 3 || npg >= 8) return false;
        pgs[npg].reset(ci);
        for (int i = 0; i < nl; i++)
          if (not_used[i] && pgs[npg].add(lines[i], false)) {
            not_used[i] = 0;
            used++;
            i = 0;  // as the reference: the scan resumes at index 1
          }
        npg++;
      } while (used < nl);
    } else {
      pgs[0].reset(ci);
      for (int i = 0; i < nl; i++) pgs[0].add(lines[i], true);
      npg = 1;
    }
    return true;
  }
};

// find_surface_4d (src/cornelius.cpp:2168-2210) with the Hypercube steps
int hypercube(const double h[2][2][2][2], const double dx[4], double E0, double cen[10][4],
              double nrm[10][4], int &err) {
  err = 0;
  int above = 0;
  for (int i = 0; i < 16; i++)
    if ((&h[0][0][0][0])[i] >= E0) above++;
  if (above == 0 || above == 16) return 0;
  static thread_local Cube cubes[8];
  int nlines = 0, amb = 0;
  for (int i = 0; i < DIM; i++)
    for (int j = 0; j < 2; j++) {
      Cube &c = cubes[2 * i + j];
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
          for (int d = 0; d < 2; d++)
            c.u[a][b][d] = i == 0 ? h[j][a][b][d]
                                  : (i == 1 ? h[a][j][b][d] : (i == 2 ? h[a][b][j][d] : h[a][b][d][j]));
      c.ci = i;
      c.cv = j * dx[i];
      c.dx = dx;
      if (!c.build(E0)) {
        err = 1;
        return 0;
      }
      nlines += (int)c.lines.size();
      if (c.ambiguous) amb++;
    }
  std::vector<Polygon *> pgs;
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < cubes[i].npg; j++) pgs.push_back(&cubes[i].pgs[j]);
  if (amb == 0) {
    int pts = 0;
    for (int i = 0; i < 16; i++)
      if ((&h[0][0][0][0])[i] < E0) pts++;
    if (pts > 8) pts = 16 - pts;
    if (nlines == 24 && pts == 2) amb++;
  }
  std::vector<Polyhedron> ph;
  if (amb > 0) {
    std::vector<int> not_used(pgs.size(), 1);
    int used = 0;
    do {
      ph.emplace_back();
      for (int i = 0; i < (int)pgs.size(); i++)
        if (not_used[i] && ph.back().add(pgs[i], false)) {
          not_used[i] = 0;
          used++;
          i = 0;
        }
    } while (used < (int)pgs.size());
  } else {
    ph.emplace_back();
    for (Polygon *p : pgs) ph.back().add(p, true);
  }
  if (ph.size() > 10) {
    err = 4;
    return 0;
  }
  for (size_t e = 0; e < ph.size(); e++) ph[e].element(cen[e], nrm[e]);
  return (int)ph.size();
}

int sym(int i, int j) { return j < i ? (i * (i + 1)) / 2 + j : (j * (j + 1)) / 2 + i; }

}  // namespace

extern "C" {

int or_cornelius(const double *h, long n, const double dx[4], double E0, double *cen,
                 double *nrm, int *nel) {
  for (long i = 0; i < n; i++) {
    double hc[2][2][2][2], c[10][4], m[10][4];
    std::memcpy(hc, h + 16 * i, sizeof hc);
    int err = 0;
    const int ne = hypercube(hc, dx, E0, c, m, err);
    nel[i] = err ? -err : ne;
    for (int e = 0; e < 10; e++)
      for (int k = 0; k < 4; k++) {
        cen[(i * 10 + e) * 4 + k] = e < ne ? c[e][k] : 0.;
        nrm[(i * 10 + e) * 4 + k] = e < ne ? m[e][k] : 0.;
      }
  }
  return 0;
}

// Fluid::outputSurface's surface elements (src/fld.cpp:544-830), SoA inputs.
// dbeta (16 x N, may be NULL): the cells' Cell::getDbeta; with it, beta
// receives the beta.dat records (33 columns, src/fld.cpp:663-834).
long or_freezeout(const or_fo_grid_t *g, const double *Q, const double *Qlow, const double *pi,
                  const double *Pi, double tau, double dt, double ecrit, int extend,
                  double *rows, long max_rows, const or_eos_t *eos, const double *dbeta,
                  double *beta) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const long N = (long)nx * ny * nz;
  const int ncol = extend ? 29 : 27;
  const double dxs[4] = {dt, g->dx, g->dy, g->dz};
  auto cell = [&](int ix, int iy, int iz) { return ix + (long)nx * (iy + (long)ny * iz); };
  // e of every cell at both levels (one con2prim pass per level)
  std::vector<double> elev[2];
  for (int lv = 0; lv < 2; lv++) {
    const double *src = lv ? Q : Qlow;
    const double t = lv ? tau : tau - dt;
    std::vector<double> q(7 * N), out(8 * N);
    for (long c = 0; c < N; c++)
      for (int k = 0; k < 7; k++) q[7 * c + k] = src[k * N + c] / t;
    or_c2p(eos, q.data(), N, out.data(), nullptr);
    elev[lv].resize(N);
    for (long c = 0; c < N; c++) elev[lv][c] = out[8 * c];
  }
  auto energy = [&](int level, long c) { return elev[level][c]; };
  long nrow = 0;
  for (int ix = 2; ix < nx - 2; ix++)
    for (int iy = 2; iy < ny - 2; iy++)
      for (int iz = 2; iz < nz - 2; iz++) {
        double h[2][2][2][2];
        for (int jx = 0; jx < 2; jx++)
          for (int jy = 0; jy < 2; jy++)
            for (int jz = 0; jz < 2; jz++) {
              const long c = cell(ix + jx, iy + jy, iz + jz);
              h[1][jx][jy][jz] = energy(1, c);
              h[0][jx][jy][jz] = energy(0, c);
            }
        double cen[10][4], nrm[10][4];
        int err = 0;
        const int ne = hypercube(h, dxs, ecrit, cen, nrm, err);
        if (err) return -1;
        for (int el = 0; el < ne; el++, nrow++) {
          if (nrow >= max_rows) continue;
          double *r = rows + nrow * ncol;
          const double *c4 = cen[el];
          r[0] = tau + c4[0];
          r[1] = g->xmin + ix * g->dx + c4[1];
          r[2] = g->ymin + iy * g->dy + c4[2];
          r[3] = g->zmin + iz * g->dz + c4[3];
          const double wT[2] = {1. - c4[0] / dt, c4[0] / dt};
          const double wX[2] = {1. - c4[1] / g->dx, c4[1] / g->dx};
          const double wY[2] = {1. - c4[2] / g->dy, c4[2] / g->dy};
          const double wZ[2] = {1. - c4[3] / g->dz, c4[3] / g->dz};
          double QC[7] = {0, 0, 0, 0, 0, 0, 0};
          for (int jt = 0; jt < 2; jt++)
            for (int jx = 0; jx < 2; jx++)
              for (int jy = 0; jy < 2; jy++)
                for (int jz = 0; jz < 2; jz++)
                  for (int i = 0; i < 7; i++)
                    QC[i] += (jt ? Q : Qlow)[i * N + cell(ix + jx, iy + jy, iz + jz)] * wT[jt] *
                             wX[jx] * wY[jy] * wZ[jz];
          for (int i = 0; i < 7; i++) QC[i] = QC[i] / (tau + c4[0]);
          double pr[8], eo[6];
          or_c2p(eos, QC, 1, pr, nullptr);
          const double eC = pr[0], nbC = pr[2];
          double vxC = pr[5], vyC = pr[6], vzC = pr[7];
          const double in4[4] = {eC, nbC, pr[3], pr[4]};  // src/fld.cpp:695-697
          or_eos_eval(eos, in4, 1, eo);
          const double TC = eo[1], mubC = eo[2], muqC = eo[3], musC = eo[4];
          double piC[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, PiC = 0.;
          for (int jx = 0; jx < 2; jx++)
            for (int jy = 0; jy < 2; jy++)
              for (int jz = 0; jz < 2; jz++) {
                const long c = cell(ix + jx, iy + jy, iz + jz);
                for (int ii = 0; ii < 10; ii++) piC[ii] += pi[ii * N + c] * wX[jx] * wY[jy] * wZ[jz];
                PiC += Pi[c] * wX[jx] * wY[jy] * wZ[jz];
              }
          double v2C = vxC * vxC + vyC * vyC + vzC * vzC;
          if (v2C > 1.) {
            vxC *= std::sqrt(0.99 / v2C);
            vyC *= std::sqrt(0.99 / v2C);
            vzC *= std::sqrt(0.99 / v2C);
          }
          const double etaC = g->zmin + iz * g->dz + c4[3];
          const double Y = etaC + 1. / 2. * std::log((1. + vzC) / (1. - vzC));
          vxC = vxC * std::cosh(Y - etaC) / std::cosh(Y);
          vyC = vyC * std::cosh(Y - etaC) / std::cosh(Y);
          vzC = std::tanh(Y);
          const double gammaC = 1. / std::sqrt(1. - vxC * vxC - vyC * vyC - vzC * vzC);
          const double uC[4] = {gammaC, gammaC * vxC, gammaC * vyC, gammaC * vzC};
          const double tauC = tau + c4[0];
          const double ch = std::cosh(etaC), sh = std::sinh(etaC);
          const double *n0 = nrm[0];
          r[4] = tauC * (ch * n0[0] - sh / tauC * n0[3]);
          r[7] = tauC * (-sh * n0[0] + ch / tauC * n0[3]);
          r[5] = tauC * n0[1];
          r[6] = tauC * n0[2];
          for (int i = 0; i < 4; i++) r[8 + i] = uC[i];
          r[12] = TC;
          r[13] = mubC;
          r[14] = muqC;
          r[15] = musC;
          double *pc = r + 16;
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
          r[26] = PiC;
          if (extend) {
            r[27] = eC;
            r[28] = nbC;
          }
          if (dbeta && beta) {
            // trilinear interpolation of the corners' dbeta (src/fld.cpp:711-727),
            // then d_i beta_j in Cartesian coordinates with the Milne-to-
            // Cartesian Jacobian of a covariant vector and g_{jj} lowering the
            // second index (src/fld.cpp:799-831)
            double dI[4][4] = {};
            for (int jx = 0; jx < 2; jx++)
              for (int jy = 0; jy < 2; jy++)
                for (int jz = 0; jz < 2; jz++) {
                  const long c = cell(ix + jx, iy + jy, iz + jz);
                  for (int a = 0; a < 4; a++)
                    for (int b = 0; b < 4; b++)
                      dI[a][b] += dbeta[(4 * a + b) * N + c] * wX[jx] * wY[jy] * wZ[jz];
                }
            const double jac[4][4] = {{ch, 0., 0., -sh}, {0., 1., 0., 0.}, {0., 0., 1., 0.},
                                      {-sh, 0., 0., ch}};
            const double gmm[4] = {1., -1., -1., -1.};
            double *b = beta + nrow * 33;
            for (int k = 0; k < 16; k++) b[k] = r[k];  // tau x y eta, dsigma, u, T mu_B mu_Q mu_S
            for (int i = 0; i < 4; i++)
              for (int j = 0; j < 4; j++) {
                double acc = 0.0;
                for (int k = 0; k < 4; k++)
                  for (int l = 0; l < 4; l++) acc += jac[i][k] * jac[j][l] * dI[k][l] * gmm[l];
                b[16 + 4 * i + j] = acc;
              }
            b[32] = eC;
          }
        }
      }
  return nrow;
}

// Fluid::outputCorona (src/fld.cpp:876-1072): at tau, every interior cube
// (loop order ix, iy, iz in [2, n-3)) whose 8 corners have e <= ecrit and
// whose centre (the corners' mean Q) has e >= 0.01 is one element at the
// cube centre with dsigma = tau dx dy dz (cosh eta, 0, 0, -sinh eta).
// Records as or_freezeout's; stats (may be NULL): element count, sum of
// dV_eff, the EtotSurf increment, the sum of n_b dV_eff.
long or_corona(const or_fo_grid_t *g, const double *Q, const double *pi, const double *Pi,
               double tau, double ecrit, int extend, double *rows, long max_rows,
               const or_eos_t *eos, double *stats) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const long N = (long)nx * ny * nz;
  const int ncol = extend ? 29 : 27;
  auto cell = [&](int ix, int iy, int iz) { return ix + (long)nx * (iy + (long)ny * iz); };
  std::vector<double> el(N);
  {
    std::vector<double> q(7 * N), out(8 * N);
    for (long c = 0; c < N; c++)
      for (int k = 0; k < 7; k++) q[7 * c + k] = Q[k * N + c] / tau;
    or_c2p(eos, q.data(), N, out.data(), nullptr);
    for (long c = 0; c < N; c++) el[c] = out[8 * c];
  }
  double acc[4] = {0., 0., 0., 0.};
  long nrow = 0;
  for (int ix = 2; ix < nx - 2; ix++)
    for (int iy = 2; iy < ny - 2; iy++)
      for (int iz = 2; iz < nz - 2; iz++) {
        bool corona = true;
        double QC[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int jx = 0; jx < 2; jx++)
          for (int jy = 0; jy < 2; jy++)
            for (int jz = 0; jz < 2; jz++) {
              const long c = cell(ix + jx, iy + jy, iz + jz);
              if (el[c] > ecrit) corona = false;
              for (int i = 0; i < 7; i++) QC[i] += Q[i * N + c] * 0.125;
            }
        for (int i = 0; i < 7; i++) QC[i] = QC[i] / tau;
        double pr[8], eo[6];
        or_c2p(eos, QC, 1, pr, nullptr);
        const double eC = pr[0], nbC = pr[2];
        if (!(corona && eC >= 0.01)) continue;
        const double in4[4] = {eC, nbC, pr[3], pr[4]};  // src/fld.cpp:979-988
        or_eos_eval(eos, in4, 1, eo);
        const double pC = eo[0], TC = eo[1], mubC = eo[2], muqC = eo[3], musC = eo[4];
        double piC[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, PiC = 0.;
        for (int jx = 0; jx < 2; jx++)
          for (int jy = 0; jy < 2; jy++)
            for (int jz = 0; jz < 2; jz++) {
              const long c = cell(ix + jx, iy + jy, iz + jz);
              for (int ii = 0; ii < 10; ii++) piC[ii] += pi[ii * N + c] * 0.125;
              PiC += Pi[c] * 0.125;
            }
        double vxC = pr[5], vyC = pr[6], vzC = pr[7];
        const double v2C = vxC * vxC + vyC * vyC + vzC * vzC;
        if (v2C > 1.) {
          vxC *= std::sqrt(0.99 / v2C);
          vyC *= std::sqrt(0.99 / v2C);
          vzC *= std::sqrt(0.99 / v2C);
        }
        const double etaC = g->zmin + iz * g->dz + 0.5 * g->dz;
        const double Y = etaC + 1. / 2. * std::log((1. + vzC) / (1. - vzC));
        vxC = vxC * std::cosh(Y - etaC) / std::cosh(Y);
        vyC = vyC * std::cosh(Y - etaC) / std::cosh(Y);
        vzC = std::tanh(Y);
        const double gammaC = 1. / std::sqrt(1. - vxC * vxC - vyC * vyC - vzC * vzC);
        const double uC[4] = {gammaC, gammaC * vxC, gammaC * vyC, gammaC * vzC};
        const double ch = std::cosh(etaC), sh = std::sinh(etaC);
        double ds[4];
        ds[0] = tau * (ch * g->dx * g->dy * g->dz);
        ds[3] = tau * (-sh * g->dx * g->dy * g->dz);
        ds[1] = 0.0;
        ds[2] = 0.0;
        double dV = 0.0;
        for (int ii = 0; ii < 4; ii++) dV += ds[ii] * uC[ii];
        double pc[10];
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
        double dEv = 0.;
        for (int i = 0; i < 4; i++) dEv += pc[sym(0, i)] * ds[i];
        acc[1] += dV;
        acc[2] += (eC + pC) * uC[0] * dV - pC * ds[0] + dEv;
        acc[3] += nbC * dV;
        if (nrow < max_rows) {
          double *r = rows + nrow * ncol;
          r[0] = tau;
          r[1] = g->xmin + ix * g->dx + 0.5 * g->dx;
          r[2] = g->ymin + iy * g->dy + 0.5 * g->dy;
          r[3] = g->zmin + iz * g->dz + 0.5 * g->dz;
          for (int i = 0; i < 4; i++) r[4 + i] = ds[i];
          for (int i = 0; i < 4; i++) r[8 + i] = uC[i];
          r[12] = TC;
          r[13] = mubC;
          r[14] = muqC;
          r[15] = musC;
          for (int i = 0; i < 10; i++) r[16 + i] = pc[i];
          r[26] = PiC;
          if (extend) {
            r[27] = eC;
            r[28] = nbC;
          }
        }
        nrow++;
      }
  acc[0] = (double)nrow;
  if (stats)
    for (int k = 0; k < 4; k++) stats[k] = acc[k];
  return nrow;
}

}  // extern "C"
