// vhlle.h -- the reference's class API for the hydro step, backed by the GPU.
//
// Same class names, signatures, argument meaning and error behaviour as
// yukarpenko/vhlle's EoS (src/eos.h:21-38), TransportCoeff
// (src/trancoeff.h:6-45), Cell (src/cll.h:38-267), Fluid (src/fld.h:10-88),
// Hydro (src/hdo.h:10-58), ICGubser (src/icGubser.h:6-12) and the free
// functions transformPV/transformCV (src/rmn.h:8-20), so code written against
// those headers compiles against this one.  Underneath, Fluid owns a device
// context of the C-ABI (include/vhlle_gpu.h); the cells live in HBM and the
// host sees them through a lazily synchronised mirror:
//   * Fluid::getCell returns a Cell proxy; reading a cell after a step
//     downloads the whole state once, writing one marks the mirror dirty;
//   * Hydro::performStep uploads a dirty mirror, runs the step on the GPU and
//     maps device status bits to the reference's messages and exit codes.
#pragma once

#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../../include/vhlle_gpu.h"

// Q[] component indexes and stage modes (src/inc.h:4-13)
#ifndef T_
#define T_ 0
#define X_ 1
#define Y_ 2
#define Z_ 3
#define NB_ 4
#define NQ_ 5
#define NS_ 6
#define PREDICT 0
#define CORRECT 1
#endif

const double C_PI = 3.14159265358979312;  // src/inc.h:18
const double gevtofm = 5.067728853;       // src/inc.h:22

// index of the pi^{mu nu} component (i,j) in the packed 10-array
// (src/cll.cpp:22-31); exits on an out-of-range index like the reference.
int index44(const int &i, const int &j);

// ------------------------------------------------------------------ EoS
// Abstract equation of state (src/eos.h:21-38).  device() describes the EoS
// to the GPU kernels; an EoS the kernels do not implement throws.
class EoS {
 public:
  virtual ~EoS() {}
  virtual void eos(double e, double nb, double nq, double ns, double &T, double &mub,
                   double &muq, double &mus, double &p) = 0;
  virtual double p(double e, double nb, double ns, double nq) = 0;
  double s(double e, double nb, double nq, double ns);
  inline double cs2(void) { return 1. / 3.; }
  inline double cs(void) { return std::sqrt(1. / 3.); }
  virtual inline double cs2(double) { return 1. / 3.; }
  virtual vh_eos_t device() const = 0;
};

// Ideal massless gas: the SIMPLE branch of EoSs (src/eos.cpp:87-111),
// p = e/3, T = (e/cnst)^(1/4), mu = 0.
class EoSIdeal : public EoS {
 public:
  void eos(double e, double nb, double nq, double ns, double &T, double &mub, double &muq,
           double &mus, double &p) override;
  double p(double e, double nb, double ns, double nq) override { return e / 3.; }
  double t(double e);
  double cs2(double) override { return 1. / 3.; }
  vh_eos_t device() const override;
};

// Tabulated EoS: the TABLE branch of EoSs (src/eos.h:44-71, src/eos.cpp:27-120).
// EoSs(fname, ncols) reads columns e, p, T [, mu] as the reference does
// (p < 0 clamped to 0, a warning for non-increasing e); natural cubic splines
// of p(e), T(e), mu(e); cs2(e) = dpe(e) = (p(1.1e) - p(e))/(0.1e); mu_b =
// mu_q = mu_s = 0 in eos().  The splines are the GPU's (vh_eos_table).
// Outside the table the reference's GSL call aborts; here the program exits
// with a message.  EoSs(desc) wraps a descriptor's columns (copied).
class EoSs : public EoS {
 public:
  EoSs(std::string fname, int ncols);
  explicit EoSs(const vh_eos_t &desc);
  void eos(double e, double nb, double nq, double ns, double &T, double &mub, double &muq,
           double &mus, double &p) override;
  double p(double e, double nb, double ns, double nq) override { return p(e); }
  double p(double e);
  double dpe(double e);
  double t(double e);
  double mu(double e);
  double cs2(double e) override { return dpe(e); }
  vh_eos_t device() const override;
  int rows() const { return n_; }

 private:
  void build();
  double eval(int f, double e) const;
  int n_ = 0;
  std::vector<double> te_, tp_, tT_, tmu_, tab_;
};

// EoSChiral (src/eoChiral.h, src/eoChiral.cpp): bilinear lookups in two
// (e, n_b) tables -- eos/chiralsmall.dat (201 x 201) where e < 1.46 and
// n_b < 0.3, eos/chiraleos.dat (2001 x 401) where e < 146 and n_b < 6 -- and
// p = 0.2964 e, T = 0.15120476935 e^(1/4) beyond; mu_q = 0; cs2(e) = 1/3.
// EoSChiral() reads the two files as the reference does (columns T [MeV],
// mu_q [MeV], e [e_0], p [e_0], n [n_0], s [n_0], mu_S [MeV], unused; unit
// factors 1/1000, 3/1000, 0.146, 0.146, 0.15); the second constructor wraps
// tables already in GeV, GeV/fm^3, 1/fm^3 (copied).  Host evaluations use the
// GPU kernels' lookups compiled for the CPU (vh_eos_eval).  The GPU path
// evaluates it in FP64 contexts only.
class EoSChiral : public EoS {
 public:
  EoSChiral();
  EoSChiral(const vh_eos_chiral_tab_t &small, const vh_eos_chiral_tab_t &big);
  void eos(double e, double nb, double nq, double ns, double &T, double &mub, double &muq,
           double &mus, double &p) override;
  double p(double e, double nb, double nq, double ns) override;
  vh_eos_t device() const override;

 private:
  struct Tab {
    vh_eos_chiral_tab_t d{};
    std::vector<double> p, T, mub, mus;
    void read(const char *fname, int ne, int nn);
    void copy(const vh_eos_chiral_tab_t &t);
  } small_, big_;
};

// EoSHadron (src/eoHadron.h, src/eoHadron.cpp): trilinear lookups in a table
// uniform in log(e/e0) and in the log-mapped n_b, n_q; cs2(e) = 1/3.
// EoSHadron(filename) reads the reference's format (a header ne nnb nnq e0 n0
// logemin logemax lognmax, then rows e nb nq p T mub muq mus status) and runs
// its range check; the second constructor wraps a table (copied).  The
// reference uses it as the hadronic EoS of the freeze-out surface (Fluid's
// eosH); the GPU path evaluates it in FP64 contexts only.
class EoSHadron : public EoS {
 public:
  explicit EoSHadron(const char *filename);
  explicit EoSHadron(const vh_eos_hadron_tab_t &t);
  void eos(double e, double nb, double nq, double ns, double &T, double &mub, double &muq,
           double &mus, double &p) override;
  double p(double e, double nb, double nq, double ns) override;
  vh_eos_t device() const override;

 private:
  vh_eos_hadron_tab_t d_{};
  std::vector<double> hp_, hT_, hmub_, hmuq_, hmus_;
  std::vector<int> hstatus_;
  void point();
};

// con2prim / prim2con on the host (src/rmn.cpp:28-139, 253-263), used by ICs
// and by host code reading cells; the hot path runs them on the GPU.
void transformPV(EoS *eos, double Q[7], double &e, double &p, double &nb, double &nq,
                 double &ns, double &vx, double &vy, double &vz);
void transformCV(double e, double p, double nb, double nq, double ns, double vx, double vy,
                 double vz, double Q[]);

// ------------------------------------------------------- TransportCoeff
// (src/trancoeff.h:6-45, src/trancoeff.cpp:19-100).  zetaSparam 5 is the
// T-peaked bulk viscosity of this build: zetaS0*exp(-(T-Tpk)^2/(2w^2)).
class TransportCoeff {
 public:
  const double etaS0, zetaS0, ah, al, aRho, T0, etaSMin, etaSEpsilonMin, etaSShiftMuB,
      etaSScaleMuB, zetaSPeakEpsilon, zetaSScaleBeta, zetaSSigmaMinus, zetaSSigmaPlus;
  EoS *eos;
  const int etaSparam, zetaSparam;
  double zetaSPeakT = 0.170, zetaSWidthMinus = 0.010, zetaSWidthPlus = 0.030;
  double zetaS(double e, double T, double s, double P);
  double etaS(double e, double rho, double T, double muB, double s, double P);
  TransportCoeff(double _etaS0, double _zetaS0, double _ah, double _al, double _aRho,
                 double _T0, double _etaSMin, double _etaSEpsilonMin, double _etaSShiftMuB,
                 double _etaSScaleMuB, double _zetaSPeakEpsilon, double _zetaSScaleBeta,
                 double _zetaSSigmaMinus, double _zetaSSigmaPlus, EoS *_eos, int _etaSparam,
                 int _zetaSparam)
      : etaS0(_etaS0), zetaS0(_zetaS0), ah(_ah), al(_al), aRho(_aRho), T0(_T0),
        etaSMin(_etaSMin), etaSEpsilonMin(_etaSEpsilonMin), etaSShiftMuB(_etaSShiftMuB),
        etaSScaleMuB(_etaSScaleMuB), zetaSPeakEpsilon(_zetaSPeakEpsilon),
        zetaSScaleBeta(_zetaSScaleBeta), zetaSSigmaMinus(_zetaSSigmaMinus),
        zetaSSigmaPlus(_zetaSSigmaPlus), eos(_eos), etaSparam(_etaSparam),
        zetaSparam(_zetaSparam) {}
  ~TransportCoeff() {}
  void getEta(double e, double rho, double T, double muB, double s, double P, double &_etaS,
              double &_zetaS);
  void getTau(double e, double rho, double T, double muB, double s, double P, double &_taupi,
              double &_tauPi);
  void getOther(double e, double nb, double nq, double ns, double &deltapipi,
                double &taupipi, double &lambdapiPi, double &phi7);
  void getOtherBulk(double e, double nb, double nq, double ns, double &delPiPi,
                    double &lamPipi);
  bool isViscous() const {
    return etaS0 > 0. || zetaS0 > 0. || etaSparam > 0 || zetaSparam > 0;
  }
  vh_trcoeff_t device() const;
};

class Fluid;

// ----------------------------------------------------------------- Cell
// Proxy for one cell of a Fluid (src/cll.h:38-267).  Storage is the Fluid's
// host mirror of the device state; every accessor keeps the reference's
// semantics (setQ zeroes a state with Q[T_] < 0, getpi is symmetric, ...).
class Cell {
 public:
  Cell() {}
  inline int getX(void) { return ix; }
  inline int getY(void) { return iy; }
  inline int getZ(void) { return iz; }
  void setQ(double *_Q);
  void setQh(double *_Qh);
  void getQ(double *_Q);
  void getQh(double *_Qh);
  void getQprev(double *_Qp);
  void saveQprev(void);
  double getpi(const int &i, const int &j);
  double getpiH(const int &i, const int &j);
  double getPi(void);
  double getPiH(void);
  void setpi(const int &i, const int &j, const double &val);
  void setpiH(const int &i, const int &j, const double &val);
  void setPi(const double &val);
  void setPiH(const double &val);
  void setAllM(double value);
  double getM(int dir);
  double getMaxM(void);
  void importVars(Cell *c);
  void getPrimVar(EoS *eos, double tau, double &_e, double &_p, double &_nb, double &_nq,
                  double &_ns, double &_vx, double &_vy, double &_vz);
  void getPrimVarHCenter(EoS *eos, double tau, double &_e, double &_p, double &_nb,
                         double &_nq, double &_ns, double &_vx, double &_vy, double &_vz);
  void getPrimVarPrev(EoS *eos, double tau, double &_e, double &_p, double &_nb,
                      double &_nq, double &_ns, double &_vx, double &_vy, double &_vz);
  void setPrimVar(EoS *eos, double tau, double _e, double _nb, double _nq, double _ns,
                  double _vx, double _vy, double _vz);
  void setViscCorrCutFlag(double value);
  double getViscCorrCutFlag(void);
  void Dump(double tau);
  // vorticity (src/cll.h:167-204): the matrices of all cells live on the
  // device; enableVorticity switches the whole Fluid on, getDbeta reads the
  // value the last step stored (clamped to [-100, 100])
  void enableVorticity();
  double getDbeta(int i, int j);

 private:
  friend class Fluid;
  Fluid *f = nullptr;
  long n = 0;
  int ix = 0, iy = 0, iz = 0;
};

// ---------------------------------------------------------------- Fluid
// The hydro grid (src/fld.h:10-88, src/fld.cpp:62-117).  Owns the device
// context; dx = (max-min)/(n-1) as in the reference (nz = 1 gives dz = inf).
class Fluid {
 public:
  Fluid(EoS *_eos, EoS *_eosH, TransportCoeff *_trcoeff, int _nx, int _ny, int _nz,
        double _minx, double _maxx, double _miny, double _maxy, double _minz, double _maxz,
        double dt, double eCrit);
  ~Fluid();
  inline int getNX() { return nx; }
  inline int getNY() { return ny; }
  inline int getNZ() { return nz; }
  inline double getDx() { return dx; }
  inline double getDy() { return dy; }
  inline double getDz() { return dz; }
  inline double getX(int ix) { return minx + ix * dx; }
  inline double getY(int iy) { return miny + iy * dy; }
  inline double getZ(int iz) { return minz + iz * dz; }
  inline double geteCrit() { return ecrit; }
  // clamped index access (src/fld.h:70-78)
  Cell *getCell(int ix, int iy, int iz);
  // ghost-cell fills and vacuum-front update of the reference (run on the
  // device; performStep calls them itself)
  void correctImagCells(void);
  void correctImagCellsFull(void);
  void updateM(double tau, double dt);

  // Fluid::outputSurface (src/fld.cpp:517-874): the freeze-out elements at
  // e = eCrit between the previous step (the Fluid's dt) and tau, appended to
  // the file opened by initOutput in the reference's freezeout.dat format;
  // returns the element count and prints the reference's summary line
  // (tau, EtotSurf, nbSurf, elements, suspicious elements).
  int outputSurface(double tau, bool extendFO = false);
  // Fluid::outputCorona (src/fld.cpp:876-1072): the initial state's corona
  // elements into freezeout.dat (call once before the first step; sets the
  // corona count of the beta.dat header); returns the element count
  int outputCorona(double tau, bool extendFO = false);
  // vorticity (src/fld.h:44-52): every later viscous step also forms the
  // cells' dbeta, and outputSurface writes beta.dat (opened by initOutput)
  void enableVorticity();
  bool vorticity() const { return vorticityOn; }
  // beta.dat header (src/fld.cpp:174-187); throws std::runtime_error like the
  // reference when outputCorona has not run
  void printDbetaHeader();
  // Fluid::initOutput (src/fld.cpp:119-172): creates dir, opens
  // freezeout.dat.unfinished (removing old freezeout.dat files), beta.dat with
  // vorticity, and unless hsOnly the gnuplot files outx/outy/outz/outdiag/
  // out2D/out.aniz.dat (and the empty outx.visc/outy.visc/diag.visc.dat);
  // writes outputGnuplot(tau0)
  void initOutput(const char *dir, double tau0, bool hsOnly = false);
  // the hypersurface files only (initOutput(dir, tau, true))
  void initOutput(const char *dir) { initOutput(dir, 0., true); }
  // Fluid::outputGnuplot (src/fld.cpp:425-504): the x, y, diagonal and eta
  // lines through the grid centre at tau
  void outputGnuplot(double tau);
  // Fluid::renameOutput (src/fld.cpp:189-196): freezeout.dat.unfinished ->
  // freezeout.dat after a clean run
  void renameOutput(const char *dir);
  void closeOutput();
  // continue writing into another Fluid's freezeout.dat (the reference's
  // output streams are global, so a Fluid created by expandGrid2x keeps them)
  void adoptOutput(Fluid *from);
  // ---- precision / con2prim of the device step (new): recreates the device
  // context before the IC is set (VH_PREC_*, VH_C2P_*; model: the con2prim
  // MLP for VH_C2P_LEARNED, VH_MLP_NPARAM floats)
  void configure(int precision, int c2pMode, const float *model = nullptr);

  // ---- GPU residency (new): the state lives in HBM between steps
  vh_ctx_t *context() { return ctx; }
  void syncToDevice();    // uploads the host mirror if host code changed it
  void syncFromDevice();  // downloads the state if a step changed it
  void markDeviceNewer() {
    dev_newer = true;
    db_stale = true;
  }
  long ncell() const { return N; }

 private:
  friend class Cell;
  void setSurfaceEos();
  EoS *eos, *eosH;
  TransportCoeff *trcoeff;
  int nx, ny, nz;
  long N;
  double minx, maxx, miny, maxy, minz, maxz;
  double dx, dy, dz, dt, ecrit;
  vh_ctx_t *ctx = nullptr;
  // host mirror, SoA component-major (the C-ABI layout)
  std::vector<double> hQ, hQh, hQprev, hpi, hpiH, hPi, hPiH, hm, hvcc;
  std::vector<Cell> cells;
  bool host_dirty = false, dev_newer = false;
  std::FILE *ffreeze = nullptr, *fbeta = nullptr;
  // gnuplot-style outputs: x, y, z, diag, 2D, aniz, x.visc, y.visc, diag.visc
  std::FILE *fgp[9] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr};
  // prints the grid totals line of outputSurface / outputCorona
  void gridTotals(double tau, bool central, double out[16]);
  double EtotSurf = 0., nbSurf = 0., vEff = 0.;
  bool vorticityOn = false, db_stale = true;
  int num_corona_cells = -1;  // -1: outputCorona has not run
  std::vector<double> hdb;    // host copy of the cells' dbeta (16 x N)
  // record buffers of outputSurface, grown ahead of the surface so that a
  // call usually needs one pass
  std::vector<double> fo_rows, fo_beta;
  long fo_cap = 0;
  void hostRead();
  void hostWrite();
};

// ---------------------------------------------------------------- Hydro
// (src/hdo.h:10-58).  performStep runs one full predictor-corrector +
// Israel-Stewart step on the GPU.
class Hydro {
 public:
  Hydro(Fluid *_f, EoS *_eos, TransportCoeff *_trcoeff, double _t0, double _dt);
  ~Hydro();
  void setDtau(double deltaTau);
  double getDtau() { return dt; }
  void setFluid(Fluid *_f) { f = _f; }
  Fluid *getFluid() const { return f; }
  void performStep(void);
  // nsteps calls of performStep replayed as HIP graphs (no host round trip)
  void performSteps(int nsteps);
  inline double getTau() const { return tau; }
  // Hydro::enableVorticity (src/hdo.cpp:60-64)
  void enableVorticity() { f->enableVorticity(); }

 private:
  Fluid *f;
  EoS *eos;
  TransportCoeff *trcoeff;
  double dt, tau;
  void check(int rc, int status);
};

// ------------------------------------------------------------ VtkOutput
// ASCII VTK structured-points files of cell fields (src/vtk.h:16-77,
// src/vtk.cpp:1-181): one file per quantity and call, named
// <path>/<quantity>_taustepNNNNN.vtk, values "%.3f".  Quantities: eps, p,
// nb, nq, ns, T, mub, muq, mus, Pi (scalars), v (vector), pi (16 scalars
// pi00..pi33).  The primitive fields come from the device (vh_prim_fields).
class VtkOutput {
 public:
  VtkOutput(std::string path, EoS *eos, double xmin, double ymin, double etamin)
      : path_(std::move(path)), eos_(eos), xmin_(xmin), ymin_(ymin), etamin_(etamin) {}
  void write(const Hydro &h, const std::string &quantities);
  bool is_quantity_implemented(const std::string &quantity);
  std::string make_filename(const std::string &descr, int counter);

 private:
  std::string path_;
  EoS *eos_;
  double xmin_, ymin_, etamin_;
  int vtk_output_counter_ = 0;
};

// Gubser analytic IC (src/icGubser.h:6-12, src/icGubser.cpp:17-62), with an
// optional energy scale (1 = the reference's normalisation).
class ICGubser {
 public:
  explicit ICGubser(double scale = 1.0) : scale_(scale) {}
  void setIC(Fluid *f, EoS *eos, double tau);

 private:
  double scale_;
};

// learned con2prim model files (magic + VH_MLP_NPARAM floats)
namespace vhlle_learn {
bool saveC2PModel(const char *path, const float *params);
bool loadC2PModel(const char *path, float *params);
}  // namespace vhlle_learn
