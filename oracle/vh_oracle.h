/*
 * CPU oracle for the vHLLE hydro step: TEST INFRASTRUCTURE ONLY.
 *
 * A plain, serial C++ restatement of the reference algorithm (yukarpenko/vhlle,
 * src/hdo.cpp, src/cll.cpp, src/rmn.cpp, src/fld.cpp, src/trancoeff.*,
 * src/eos.*) in the reference's own loop order and operation order, so that it
 * reproduces the reference's FP64 results to rounding.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker (or the timed CPU baseline) -- never as the product path.
 *
 * State layout used at this interface: structure-of-arrays, component-major,
 * arr[comp * ncell + cell] with cell = ix + nx*(iy + ny*iz), i.e. the same cell
 * ordering as Fluid::getCell (src/fld.h:70-78).
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

/* Transport coefficients: the fields of TransportCoeff (src/trancoeff.h:8-11)
 * plus the T-peaked bulk parametrisation (zetaSparam 5) added by this build. */
typedef struct {
  double etaS0, zetaS0, ah, al, aRho, T0, etaSMin, etaSEpsilonMin, etaSShiftMuB,
      etaSScaleMuB, zetaSPeakEpsilon, zetaSScaleBeta, zetaSSigmaMinus,
      zetaSSigmaPlus;
  int etaSparam, zetaSparam;
  double zetaSPeakT, zetaSWidthMinus, zetaSWidthPlus;
} or_trc_t;

/* EoS (src/eos.h, src/eos.cpp): kind 0 the ideal gas p=e/3 (SIMPLE branch of
 * EoSs); kind 1 the TABLE branch: n rows of e (increasing), p (< 0 clamped to
 * 0), T, mu (may be NULL), natural cubic splines (gsl_interp_cspline); kind 2
 * EoSChiral (src/eoChiral.cpp): two bilinear (e, n_b) tables and a power law;
 * kind 3 EoSHadron (src/eoHadron.cpp): a trilinear table in log(e/e0) and
 * the log-mapped n_b, n_q.  A NULL descriptor is the ideal gas.  The table
 * layouts are those of vh_eos_chiral_tab_t / vh_eos_hadron_tab_t. */
typedef struct {
  int ne, nn;
  double emin, emax, nmin, nmax;
  const double *p, *T, *mub, *mus; /* [ie * nn + in] */
} or_eos_chiral_tab_t;
typedef struct {
  int ne, nnb, nnq;
  double e0, n0, logemin, logemax, lognmax;
  const double *p, *T, *mub, *muq, *mus; /* [inq + nnq * (inb + nnb * ie)] */
  const int *status;                     /* may be NULL (all 0) */
} or_eos_hadron_tab_t;
typedef struct {
  int kind;
  int n;
  const double *e, *p, *T, *mu;
  const or_eos_chiral_tab_t *small, *big;
  const or_eos_hadron_tab_t *hadron;
} or_eos_t;

/* Whole-state view, SoA host arrays (any pointer may be NULL on set: field
 * then left unchanged; on get: not written). */
typedef struct {
  double *Q, *Qh, *Qprev;     /* 7 x N each */
  double *pi, *piH;           /* 10 x N each, packed index44 order */
  double *Pi, *PiH;           /* N each */
  double *pi0, *piH0;         /* 10 x N: auxiliary IS arrays */
  double *Pi0, *PiH0;         /* N */
  double *m;                  /* 3 x N */
  double *viscCorrCut;        /* N */
} or_state_t;

void *or_create(int nx, int ny, int nz, double minx, double maxx, double miny,
                double maxy, double minz, double maxz, const or_eos_t *eos,
                const or_trc_t *trc);
void or_destroy(void *h);
long or_ncell(void *h);
void or_set_state(void *h, const or_state_t *s);
void or_get_state(void *h, or_state_t *s);
/* Cell::setPrimVar + saveQprev + setAllM(1 if e>0): the IC path of
 * ICGubser::setIC (src/icGubser.cpp:17-62). prim arrays are N each. */
void or_set_prim(void *h, double tau, const double *e, const double *nb,
                 const double *nq, const double *ns, const double *vx,
                 const double *vy, const double *vz);
/* Hydro::performStep (src/hdo.cpp:1016-1134) nsteps times from *tau with
 * timestep dt.  Returns 0 on success, otherwise the reference's exit reason:
 * 1 = NaN in hlle_flux / source_step / NSquant, 2 = e>1e10 (exit(0) in the
 * reference). */
int or_run(void *h, double *tau, double dt, int nsteps);
/* Hydro::enableVorticity (src/hdo.cpp:60-65): from the next viscous step on,
 * NSquant also forms dbeta[i][j] = d_i (u^j / T) and stores it per cell
 * (Cell::setDbeta, src/cll.h:173-183); every cell's dbeta is reset to 0. */
void or_set_vorticity(void *h, int on);
/* Cell::getDbeta of every cell: out 16 x N, component 4 i + j = dbeta[i][j] */
void or_get_dbeta(void *h, double *out);
/* Single functions of the step on the current state, for unit tests of the
 * restatement: the ghost fill (full = 0: correctImagCells, Q and Qh; 1:
 * correctImagCellsFull, Q, pi, Pi; src/fld.cpp:211-382); Fluid::updateM
 * (src/fld.cpp:384-423); ISformal loop #2 (advection, limiter, bulk clamp,
 * src/hdo.cpp:866-962) from the auxiliary arrays pi0, piH0, Pi0, PiH0 and the
 * Qh prims at tau - dt/2; hlle_flux of one interface (src/hdo.cpp:77-305) from
 * (ix, iy, iz) to its +1 neighbour along d (1..3) in mode 0 (predictor, Q) or
 * 1 (corrector, Qh): flux[7] receives the flux added to the right cell; the
 * return value is 1 if a flux was formed, 0 if the reference returned early,
 * negative on its error exits. */
void or_ghost(void *h, int full);
void or_update_m(void *h, double tau, double dt);
void or_advect(void *h, double tau, double dt);
int or_hlle(void *h, int ix, int iy, int iz, int d, int mode, double tau, double dt,
            double *flux);
/* Branch counters since creation: out[4] = negative-pressure re-evaluations
 * of hlle_flux in the predictor and in the corrector (src/hdo.cpp:134-140),
 * cells rescaled by the ISformal limiter (src/hdo.cpp:925-932), cells with Pi
 * clamped to +-p (src/hdo.cpp:933-937). */
void or_branch_counts(void *h, long *out);
/* Largest HLLE signal speeds max(-bl, br) over the interfaces that formed a
 * flux: out[9] = the last step's predictor (x, y, eta) and corrector (x, y,
 * eta) maxima, then the per-axis maxima since creation or the last reset
 * (reset != 0 clears them after reading). */
void or_signal_speed(void *h, int reset, double *out);
/* Hydro::setDtau validity test (src/hdo.cpp:67-75): 0 if accepted, 1 if the
 * reference would exit(1). */
int or_check_dtau(void *h, double dt);

/* transformPV (src/rmn.cpp:28-139) over n inputs.  Q: n x 7 row-major (it is
 * modified in place exactly like the reference modifies its argument).
 * out: n x 8 row-major {e, p, nb, nq, ns, vx, vy, vz}.  iters (may be NULL):
 * loop iterations executed per input. */
void or_c2p(const or_eos_t *eos, double *Q, long n, double *out, int *iters);
/* transformCV (src/rmn.cpp:253-263): prim n x 8 {e,p,nb,nq,ns,vx,vy,vz} -> Q n x 7 */
void or_p2c(const double *prim, long n, double *Q);
/* EoS relations (EoS::eos, EoS::s, cs2(e)): out n x 4 {p, T, s, cs2}. */
void or_eos(const or_eos_t *eos, const double *e, long n, double *out);
/* EoS::eos and EoS::s at n points: in n x 4 {e, nb, nq, ns} ->
 * out n x 6 {p, T, mub, muq, mus, s} */
void or_eos_eval(const or_eos_t *eos, const double *in, long n, double *out);
/* the table's spline of f (0 p, 1 T, 2 mu) at n points (kind 1 only) */
void or_eos_spline(const or_eos_t *eos, int f, const double *v, long n, double *out);
/* TransportCoeff::getEta + getTau (src/trancoeff.cpp:19-100) on n states
 * {e, nb, T, muB, s, p} -> out n x 4 {etaS, zetaS, taupi, tauPi}. */
void or_trcoeff(const or_trc_t *trc, const or_eos_t *eos, const double *in, long n,
                double *out);

/* ---- freeze-out (oracle/fo_oracle.cpp) ---- */
/* Cornelius::find_surface_4d on n hypercubes h (n x 16, [t][x][y][z]);
 * cen, nrm: n x 10 x 4; nel: element counts (negative: failed) */
int or_cornelius(const double *h, long n, const double dx[4], double E0, double *cen,
                 double *nrm, int *nel);
typedef struct {
  int nx, ny, nz;
  double dx, dy, dz, xmin, ymin, zmin;
} or_fo_grid_t;
/* Fluid::outputSurface's element records (27 or 29 columns, see
 * fo_oracle.cpp) from SoA Q, Qlow (7 x N), pi (10 x N), Pi (N); returns the
 * element count (rows beyond max_rows are counted, not written), -1 on a
 * construction failure.  dbeta (16 x N) and beta (max_rows x 33) may be
 * NULL; with both, beta receives the beta.dat records of the elements. */
long or_freezeout(const or_fo_grid_t *g, const double *Q, const double *Qlow, const double *pi,
                  const double *Pi, double tau, double dt, double ecrit, int extend,
                  double *rows, long max_rows, const or_eos_t *eos, const double *dbeta,
                  double *beta);
/* Fluid::outputCorona's element records (src/fld.cpp:876-1072) at tau from SoA
 * Q, pi, Pi; stats (may be NULL, 4 doubles): count, sum dV_eff, EtotSurf and
 * sum n_b dV_eff increments.  Returns the element count. */
long or_corona(const or_fo_grid_t *g, const double *Q, const double *pi, const double *Pi,
               double tau, double ecrit, int extend, double *rows, long max_rows,
               const or_eos_t *eos, double *stats);
#ifdef __cplusplus
}
#endif
