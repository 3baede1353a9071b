/*
 * vhlle_gpu.h -- C-ABI of the MI355X-native vHLLE hydro step.
 *
 * The reference (yukarpenko/vhlle) exposes its hot path as a C++ class API:
 * Hydro::performStep (src/hdo.cpp:1016-1134) acting on a Fluid of Cells
 * (src/fld.h:10-88, src/cll.h:38-267) with an EoS (src/eos.h:21-38) and a
 * TransportCoeff (src/trancoeff.h:6-45).  This header is the boundary BELOW
 * that class API: plain C, POD structs, host pointers and sizes, integer
 * status codes.  The C++ facade (vhlle_amd/host/vhlle.h: Fluid, Hydro, Cell, EoS,
 * TransportCoeff) is written on top of it, and any other host language can
 * bind it directly.
 *
 * State layout crossing the boundary: structure-of-arrays, component-major,
 * arr[comp*N + cell], cell = ix + nx*(iy + ny*iz) -- the cell order of
 * Fluid::getCell (src/fld.h:70-78).  pi-like fields hold the 10 independent
 * components of the symmetric tensor in the order of index44 (src/cll.cpp:22-31).
 *
 * Every call is synchronous with respect to the host unless stated otherwise;
 * the context owns one HIP stream.  Nothing here calls exit(): failures are
 * returned as VH_* codes and the C++ facade maps them to the reference's exits.
 */
#ifndef VHLLE_GPU_H
#define VHLLE_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
  VH_OK = 0,
  VH_ERR_ARG = -1,        /* bad argument (null pointer, size, mode) */
  VH_ERR_HIP = -2,        /* a HIP runtime call failed */
  VH_ERR_NODEVICE = -3,   /* no GPU visible */
  VH_ERR_DTAU = -4        /* Hydro::setDtau would exit(1) (src/hdo.cpp:67-75) */
};
/* step status bits (vh_step / vh_run *status), the reference's exits: */
enum {
  VH_ST_NAN_FLUX = 1,     /* hlle_flux f_nb NaN -> exit(1)  (src/hdo.cpp:285-298) */
  VH_ST_BIG_E = 2,        /* e > 1e10 -> exit(0)            (src/hdo.cpp:118-130) */
  VH_ST_NAN_SOURCE = 4,   /* source_step k_nb NaN -> exit(1)(src/hdo.cpp:345-351) */
  VH_ST_NAN_PI = 8,       /* NSquant pi inf/nan -> exit(1)  (src/hdo.cpp:643-651) */
  VH_ST_EOS_RANGE = 16    /* tabulated EoS evaluated outside its table: GSL's
                             domain error aborts the reference (gsl_spline_eval,
                             src/eos.cpp:89-111); here the end interval's cubic
                             is used and this bit raised */
};

/* precision modes of vh_create */
enum {
  VH_PREC_FP64 = 0,       /* faithful FP64: matches the reference to rounding */
  VH_PREC_MIXED = 1       /* FP32 fluxes/viscous sources, FP64 Q accumulation */
};

/* con2prim modes (vh_set_c2p_mode) */
enum {
  VH_C2P_REFERENCE = 0,   /* transformPV's bisection/Newton from v=0.5 */
  VH_C2P_LEARNED = 1,     /* MLP initial guess + FP64 Newton polish to a residual
                             <= 1e-12 Q_t, fallback to the reference iteration
                             (needs vh_set_c2p_model) */
  VH_C2P_CLOSED = 2,      /* closed-form root of the conformal EoS */
  VH_C2P_NEWTON = 3       /* the same Newton polish started at v = 0.5, no MLP
                             (the ablation of VH_C2P_LEARNED) */
};

/* Fluid geometry: the Fluid ctor arguments (src/fld.cpp:62-80).
 * dx = (xmax-xmin)/(nx-1) etc.; nz = 1 gives dz = inf as in the reference. */
typedef struct {
  int nx, ny, nz;
  double xmin, xmax, ymin, ymax, zmin, zmax;
} vh_grid_t;

/* EoS descriptor replacing the virtual EoS (src/eos.h:21-38).
 * kind 0: ideal massless gas p=e/3, T=(e/cnst)^(1/4), mu=0, the SIMPLE branch
 * of EoSs (src/eos.cpp:87-111).
 * kind 1: the TABLE branch of EoSs (src/eos.cpp:27-120): n rows of e (strictly
 * increasing, GeV/fm^3), p (negative values clamped to 0 as the reference
 * does), T (GeV) and mu (GeV, may be NULL = 0); natural cubic splines of p(e),
 * T(e), mu(e) (gsl_interp_cspline); cs2(e) = dpe(e) = (p(1.1e) - p(e))/(0.1e);
 * mu_b = mu_q = mu_s = 0 in eos().
 * kind 2: EoSChiral (src/eoChiral.cpp:99-154): bilinear lookups in two
 * uniform (e, n_b) tables -- `small` where e < 1.46 and n_b < 0.3, else `big`
 * where e < 146 and n_b < 6, else p = 0.2964 e, T = 0.15120476935 e^(1/4),
 * mu_b = mu_s = 0; mu_q = 0 always; cs2(e) = 1/3 (src/eos.h:37).
 * kind 3: EoSHadron (src/eoHadron.cpp:78-208): trilinear lookups in a table
 * uniform in log(e/e0) and in the log-mapped n_b, n_q; cs2(e) = 1/3.
 * Kinds 2 and 3 depend on the charge densities: con2prim refreshes n_b, n_q,
 * n_s every iteration as transformPV does (src/rmn.cpp:74-104).  They are
 * evaluated by the FP64 path only; VH_PREC_MIXED and the learned con2prim
 * reject them with VH_ERR_ARG.  All arrays are copied (host memory). */
enum { VH_EOS_IDEAL = 0, VH_EOS_TABLE = 1, VH_EOS_CHIRAL = 2, VH_EOS_HADRON = 3 };

/* One table of EoSChiral (the auxiliary class EoSaux, src/eoChiral.cpp:38-84):
 * ne x nn nodes uniform over [emin, emax] x [nmin, nmax] (GeV/fm^3, 1/fm^3 --
 * the units after the reference's conversion of the file columns), each field
 * stored at [ie * nn + in].  ne, nn >= 2. */
typedef struct {
  int ne, nn;
  double emin, emax, nmin, nmax;
  const double *p, *T, *mub, *mus;
} vh_eos_chiral_tab_t;

/* The table of EoSHadron (src/eoHadron.cpp:13-65): ne x nnb x nnq nodes,
 * field index inq + nnq * (inb + nnb * ie) (index3, src/eoHadron.h:9-11);
 * nodes uniform in x_e = log(e/e0) over [logemin, logemax] and in the mapped
 * x_nb, x_nq over [-lognmax, lognmax].  status 1 marks a node outside the
 * EoS (lookups whose base node has it return zeros).  ne, nnb, nnq >= 2, and
 * nnb, nnq >= 3 so that the linear region near n = 0 exists. */
typedef struct {
  int ne, nnb, nnq;
  double e0, n0, logemin, logemax, lognmax;
  const double *p, *T, *mub, *muq, *mus;
  const int *status;
} vh_eos_hadron_tab_t;

typedef struct {
  int kind;
  int n;                          /* table rows (kind 1), >= 3 */
  const double *e, *p, *T, *mu;   /* table columns (kind 1) */
  const vh_eos_chiral_tab_t *small, *big;  /* kind 2 */
  const vh_eos_hadron_tab_t *hadron;       /* kind 3 */
} vh_eos_t;
/* The natural cubic splines of a table EoS (gsl_spline_init of EoSs,
 * src/eos.cpp:59-69), computed on the host (no GPU needed) into buf[13 n]:
 * the n nodes e_i, then for p, T and mu in turn the arrays y_i, b_i, c_i, d_i
 * (n each; y = y_i + dx (b_i + dx (c_i + dx d_i)) on [e_i, e_i+1], dx = e - e_i).
 * Host code (the C++ facade's EoSs) evaluates the same splines as the GPU. */
int vh_eos_table(const vh_eos_t *eos, double *buf);
/* EoS::eos and EoS::s (src/eos.h:24-30, src/eos.cpp:18-25) at n points:
 * in[4 i .. 4 i + 3] = e, n_b, n_q, n_s; out[6 i .. 6 i + 5] = p, T, mu_b,
 * mu_q, mu_s, s.  where = 1 evaluates with the GPU kernels' device EoS (needs
 * a GPU); where = 0 on the host, with the same lookups compiled for the CPU
 * (kinds 2 and 3; kinds 0 and 1 evaluate e/3 and the splines).  The C++
 * facade's EoSChiral and EoSHadron call it with where = 0. */
int vh_eos_eval(const vh_eos_t *eos, const double *in, long n, double *out, int where);

/* TransportCoeff (src/trancoeff.h:8-11, ctor :14-20) as a POD, plus the
 * T-peaked bulk parametrisation zetaSparam = 5 (this build):
 * zeta/s = zetaS0 * exp(-(T-zetaSPeakT)^2 / (2 w^2)), w = zetaSWidthMinus
 * below the peak and zetaSWidthPlus above it. */
typedef struct {
  double etaS0, zetaS0, ah, al, aRho, T0, etaSMin, etaSEpsilonMin,
      etaSShiftMuB, etaSScaleMuB, zetaSPeakEpsilon, zetaSScaleBeta,
      zetaSSigmaMinus, zetaSSigmaPlus;
  int etaSparam, zetaSparam;
  double zetaSPeakT, zetaSWidthMinus, zetaSWidthPlus;
} vh_trcoeff_t;

/* Host view of the whole per-cell state (SoA, see above).  On upload a NULL
 * field is set to the Cell ctor default (src/cll.cpp:33-52: zeros,
 * viscCorrCut = 1); on download a NULL field is skipped. */
typedef struct {
  double *Q, *Qh, *Qprev;   /* 7 x N each: Cell::Q, Qh, Qprev (src/cll.h:44-46) */
  double *pi, *piH;         /* 10 x N each: Cell::pi, piH (src/cll.h:47) */
  double *Pi, *PiH;         /* N each (src/cll.h:49) */
  double *m;                /* 3 x N: Cell::m (src/cll.h:57) */
  double *viscCorrCut;      /* N: Cell::viscCorrCut (src/cll.h:60) */
} vh_state_t;

typedef struct vh_ctx vh_ctx_t;

/* Creates a context (device buffers for one Fluid + one Hydro).  Replaces
 * Fluid::Fluid (src/fld.cpp:62-112; the neighbour links become index clamps)
 * and Hydro::Hydro (src/hdo.cpp:50-57). */
int vh_create(const vh_grid_t *grid, const vh_eos_t *eos,
              const vh_trcoeff_t *trc, int precision, vh_ctx_t **out);
/* Several independent grids in one context, for event-by-event batches (no
 * counterpart in the reference, which runs one Fluid per process): the state
 * arrays hold the cells of grid 0, then grid 1, ... (each in its own cell
 * order) with the component pitch N = the total cell count, which vh_ncell
 * returns; vh_upload/download/set_prim take arrays of that layout.  One
 * vh_step/vh_run advances every grid by the same dt, all grids in the same
 * kernel launches (one graph per step).  Mixed precision only; each extent 1
 * or >= 5; e > 0 in every cell (the batch has no vacuum fronts: steps return
 * VH_ERR_ARG otherwise). */
int vh_create_batch(int ngrid, const vh_grid_t *grids, const vh_eos_t *eos,
                    const vh_trcoeff_t *trc, int precision, vh_ctx_t **out);
/* Fluid::~Fluid (src/fld.cpp:114-117) */
int vh_destroy(vh_ctx_t *ctx);
/* number of cells nx*ny*nz */
long vh_ncell(const vh_ctx_t *ctx);

/* Writes the whole state (the cells' fields as the ICs leave them through
 * Cell::setPrimVar / saveQprev / setAllM, src/cll.cpp:222-241). */
int vh_upload(vh_ctx_t *ctx, const vh_state_t *host);
/* Reads the state back (what host code sees through Fluid::getCell after a
 * step: Cell::getQ/getQprev/getpi/getPi/getM/getViscCorrCutFlag). */
int vh_download(vh_ctx_t *ctx, vh_state_t *host);

/* IC helper on the device: Cell::setPrimVar (src/cll.cpp:222-241) on every
 * cell from primitive arrays (N each), then saveQprev and setAllM(1.) where
 * e > 0 -- the body of ICGubser::setIC (src/icGubser.cpp:47-57).  nb/nq/ns
 * may be NULL (zero). */
int vh_set_prim(vh_ctx_t *ctx, double tau, const double *e, const double *nb,
                const double *nq, const double *ns, const double *vx,
                const double *vy, const double *vz);

/* Cell::getPrimVar of every cell at tau (src/cll.cpp:90-96, transformPV of
 * Q/tau), the inputs of the reference's cell outputs (VtkOutput::write,
 * src/vtk.cpp:38-120): prim, 8 x N SoA, rows e, p, nb, nq, ns, vx, vy, vz. */
int vh_prim_fields(vh_ctx_t *ctx, double tau, double *prim);

/* The EoS of the freeze-out outputs: Fluid's eosH (src/fld.cpp:62-66), which
 * the reference swaps in for all of outputSurface and outputCorona
 * (SWAP_EOS, src/fld.cpp:39,540-542,870-872,885-887,1085-1087) -- the grid
 * totals, the con2prim at the element centroids and the T, mu columns.  It
 * applies to vh_fluid_stats, vh_freezeout, vh_freezeout_beta and vh_corona.
 * NULL restores the hydro EoS (the default).  The arrays are copied.  A
 * charge-dependent EoS (kinds 2, 3) needs a VH_PREC_FP64 context. */
int vh_set_surface_eos(vh_ctx_t *ctx, const vh_eos_t *eos);

/* The grid totals that Fluid::outputSurface and Fluid::outputCorona print
 * (src/fld.cpp:545-606, 852-857, 889-945, 1071-1084), summed over the
 * interior cells (ix, iy, iz in [2, n-3]) of the state at tau.  out[16]:
 * E, Efull, S, Px, Nb1, Nb2 (before the factor dx dy dz), vt_num, vt_den,
 * vxvy_num, vxvy_den, pi0x_num, pi0x_den, txxyy_num, txxyy_den, the cells
 * with e > e_crit, and those of them with viscCorrCut < 0.9.  central != 0:
 * the vt, vxvy and txxyy sums only over nz/2-3 < iz < nz/2+3, as
 * outputSurface; 0: over all, as outputCorona.  Single-grid contexts. */
int vh_fluid_stats(vh_ctx_t *ctx, double tau, double e_crit, int central, double *out);
/* The state of n listed cells (cell indices) at tau, VH_CELL_COLS doubles
 * each: Cell::getPrimVar (e, p, nb, nq, ns, vx, vy, vz), pi (10, index44),
 * Pi, viscCorrCut, 0 -- the inputs of Fluid::outputGnuplot
 * (src/fld.cpp:425-504). */
enum { VH_CELL_COLS = 21 };
int vh_cells(vh_ctx_t *ctx, double tau, const long *cells, long n, double *out);
/* vh_cells for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): no
 * counterpart in the reference.  Same records as vh_cells, but the call
 * allocates nothing: the cell list and the records go through one fixed
 * device array per GPU (1.4 MB, held until the process ends), so at most
 * VH_CELLS_NOALLOC_MAX cells per call (more: VH_ERR_ARG).  Calls that
 * overlap on one GPU share that array and race; the caller serialises them. */
enum { VH_CELLS_NOALLOC_MAX = 8192 };
int vh_cells_noalloc(vh_ctx_t *ctx, double tau, const long *cells, long n, double *out);

/* Hydro::performStep (src/hdo.cpp:1016-1134): one predictor-corrector step
 * (+ Israel-Stewart step when viscous) from *tau with timestep dt; *tau is
 * advanced by dt.  *status receives VH_ST_* bits (0 = the reference would
 * have continued). */
int vh_step(vh_ctx_t *ctx, double *tau, double dt, int *status);
/* nsteps of vh_step without host synchronisation in between: the step
 * sequence is replayed from HIP graphs, one graph launch per 8 steps (and
 * one per step for the rest); *status after all steps. */
int vh_run(vh_ctx_t *ctx, double *tau, double dt, int nsteps, int *status);
/* Hydro::setDtau validity test (src/hdo.cpp:67-75): VH_OK or VH_ERR_DTAU. */
int vh_check_dtau(const vh_ctx_t *ctx, double dt);
/* Largest HLLE signal speeds max(-b_l, b_r) over the interfaces that formed a
 * flux: the wave-speed estimates of Hydro::hlle_flux (src/hdo.cpp:181-188,
 * 210-227, 255-264; 1 or 1/tau at a vacuum boundary), in units of dx/dtau
 * (eta: deta/dtau, as the reference divides them by tau).  The reference
 * keeps its timestep fixed and never forms this maximum; here every flux
 * kernel reduces it per wave, so a caller can check the Courant numbers
 * dt * out[k] / dx_k without another pass over the grid.  Steps are not
 * changed by it.  out[9]: the last step's predictor (x, y, eta) and
 * corrector (x, y, eta) maxima, then the per-axis maxima over every step
 * since the last vh_set_prim / vh_upload or reset call; reset != 0 clears the
 * latter after reading. */
int vh_signal_speed(vh_ctx_t *ctx, int reset, double *out);

/* Selects the con2prim algorithm for later steps (VH_C2P_*). */
int vh_set_c2p_mode(vh_ctx_t *ctx, int mode);

/* Shape of the mixed-precision stage kernel (no counterpart in the
 * reference).  VH_STAGE_CELLS: one thread per cell evaluating its six
 * interfaces; VH_STAGE_ROWS: waves of 64 consecutive cells of an (x, y)
 * plane marching CZ cells in eta, every x- and eta-interface evaluated once;
 * VH_STAGE_BANDS: waves of 64 cells of the first rows of bands of CY rows,
 * marching CZ cells in eta and walking the CY rows of their bands, the
 * y-interfaces inside a band evaluated once as well (vh_set_stage takes
 * CY = 4 for it; vh_set_stage_bands sets CY = 2, 4 or 8).  All give identical
 * results.  VH_STAGE_AUTO (default): an autotuner times
 * the candidates (cells; rows with CZ = 1..32; bands) on the first vh_step/vh_run
 * (and after vh_set_c2p_mode) and keeps the fastest; 32 steps later it times
 * that shape against the best of the other kind again, on the developed
 * flow, and keeps the faster (within a vh_run if one is going on: the
 * results do not depend on the shape).  vh_get_stage reports
 * the kernel in use (VH_STAGE_AUTO: not tuned yet), CZ and the tuned time of
 * a predictor and a corrector stage in ms.  Mixed precision only (VH_ERR_ARG
 * otherwise). */
enum { VH_STAGE_AUTO = 0, VH_STAGE_CELLS = 1, VH_STAGE_ROWS = 2, VH_STAGE_BANDS = 3 };
int vh_set_stage(vh_ctx_t *ctx, int kind, int CZ);
int vh_get_stage(const vh_ctx_t *ctx, int *kind, int *CZ, double *ms);
/* The band shape with CY rows per band (2, 4 or 8) and CZ eta cells per wave
 * (no counterpart in the reference); vh_get_stage_bands reports CY of the
 * shape in use (0 when it is not a band shape). */
int vh_set_stage_bands(vh_ctx_t *ctx, int CY, int CZ);
int vh_get_stage_bands(const vh_ctx_t *ctx, int *CY);
/* Per-kernel timing over the next vh_run: enable=1 records HIP events around
 * every launch on the context's stream; vh_kernel_times returns, for up to
 * max kernels, the name, launch count and summed milliseconds. */
int vh_enable_timing(vh_ctx_t *ctx, int enable);
int vh_kernel_times(const vh_ctx_t *ctx, int max, char names[][32],
                    int *launches, double *ms);

/* ---- per-element entry points (unit-level parity checks, tools) ---- */

/* transformPV (src/rmn.cpp:28-139) on n inputs: Q n x 7 row-major host array
 * (rescaled in place exactly as the reference rescales its argument when
 * |Q_vec| > Q_t), out n x 8 {e,p,nb,nq,ns,vx,vy,vz}; iters (may be NULL):
 * iterations used per input. */
int vh_c2p(const vh_eos_t *eos, double *Q, long n, double *out, int *iters);

/* transformCV (src/rmn.cpp:253-263): prim n x 8 -> Q n x 7 (host arrays). */
int vh_p2c(const double *prim, long n, double *Q);

/* ---- learned con2prim: MLP 5 -> 64 -> 64 -> 2 (ReLU) ----
 * No counterpart in the reference: an initial guess for transformPV's root
 * (src/rmn.cpp:28-139).  params: VH_MLP_NPARAM floats,
 *   W1[64][5] | b1[64] | W2[64][64] | b2[64] | W3[2][64] | b3[2];
 * inputs of a state q = Q/tau: ln(q_t)/10, |q|/q_t, q_eta/q_t, n_b/q_t,
 * sqrt(1 - (|q|/q_t)^2); outputs: |v| and e/q_t.  Inference runs in BF16 MFMA
 * with FP32 accumulation; W1, b1, W2, the inputs and the first hidden layer
 * are rounded to BF16, the rest is FP32. */
#define VH_MLP_NPARAM 4674
enum { VH_START_MLP = 0, VH_START_HALF = 1 };
typedef struct {
  int batch;                  /* samples per Adam step */
  int epochs;
  double lr, lr_final;        /* cosine schedule from lr to lr_final */
  unsigned long long seed;    /* shuffling */
} vh_mlp_train_cfg_t;
/* He-uniform initial weights (zero biases) from a SplitMix64 stream */
int vh_mlp_init(unsigned long long seed, float *params);
/* MFMA inference: out n x 2 (|v|, e/q_t) for states Q (n x 7, already / tau) */
int vh_mlp_forward(const float *params, const double *Q, long n, float *out);
/* gradient of the loss (1/n) sum_i |MLP(q_i) - (|v_i|, e_i/q_t,i)|^2 over n
 * states Q (n x 7) with reference primitives prim (n x 8, vh_c2p layout);
 * grad: VH_MLP_NPARAM floats; *loss (may be NULL) the loss.  n <= 2^20. */
int vh_mlp_grad(const float *params, const double *Q, const double *prim, long n,
                float *grad, double *loss);
/* Adam training on the GPU from params (updated in place); loss_hist (may be
 * NULL): mean sampled batch loss per epoch (cfg->epochs entries). */
int vh_mlp_train(const vh_mlp_train_cfg_t *cfg, const double *Q, const double *prim,
                 long n, float *params, double *loss_hist);
/* con2prim of n states (as vh_c2p) started from the MLP guess
 * (VH_START_MLP) or from v = 0.5 (VH_START_HALF): Newton polish to
 * |f| <= 1e-12 q_t, else the reference iteration.  iters (may be NULL): polish
 * updates, or -1 - (reference iterations) where it fell back. */
int vh_c2p_learned(const vh_eos_t *eos, const float *params, double *Q, long n,
                   int start, double *out, int *iters);
/* Installs an MLP for VH_C2P_LEARNED in a context (copied to the device). */
int vh_set_c2p_model(vh_ctx_t *ctx, const float *params);
/* Training pairs from ideal Gubser runs (the 101 x 101 x 1 grid, dx = 0.1,
 * tau0 = 1, dtau = 0.025) on the FP64 path: n states q = Q/tau drawn (seeded)
 * from every cell of every step of runs at the IC amplitudes amp[0..namp),
 * the fraction augment of them replaced by transformCV images of
 * random (e, v, n_b) with |v| <= vmax; prim = vh_c2p of each state.
 * Q: n x 7, prim: n x 8. */
typedef struct {
  unsigned long long seed;
  int steps;                  /* steps per run */
  int namp;
  double amp[4];
  double augment;             /* fraction in [0, 1] */
  double vmax;
  const vh_eos_t *eos;        /* EoS of the runs and labels (NULL: ideal gas);
                                 augmented e are drawn inside a table's range */
} vh_harvest_cfg_t;
int vh_mlp_harvest(const vh_harvest_cfg_t *cfg, long n, double *Q, double *prim);
/* ---- hot-spot initial states (event-by-event batches) ----
 * No direct counterpart in the reference (its Glissando IC,
 * src/icGlissando.cpp:209-246, reads the spots from a file): per grid g,
 * N_part ~ Poisson(npart_mean) spots at 2-D Woods-Saxon positions
 * (density r dr dphi / (1 + exp((r - R)/a))) with Gamma(k = 1, theta = 1)
 * weights w, sampled from a SplitMix64 stream seeded by (seed, g); then
 *   e(x, y, eta) = amp * sum_i w_i exp(-|x - x_i|^2 / (2 sigma^2)) * H(eta),
 * H = 1 for |eta| < eta_plateau (and for nz = 1), else Gaussian tails of width
 * eta_sigma, deposited on the GPU.  e: the cells of all grids in the layout of
 * vh_create_batch; npart (may be NULL): the spot count of each grid; spots
 * (may be NULL): the first max_spots spots of all grids in order, as
 * (x, y, w) triples. */
typedef struct {
  unsigned long long seed;
  double npart_mean;          /* <= 500 */
  double R, a;                /* Woods-Saxon radius and diffuseness, fm */
  double sigma;               /* spot width, fm */
  double amp;                 /* GeV/fm^3 per unit weight */
  double eta_plateau, eta_sigma;
} vh_hotspot_cfg_t;
int vh_hotspot_ic(const vh_hotspot_cfg_t *cfg, int ngrid, const vh_grid_t *grids,
                  double *e, int *npart, double *spots, long max_spots);
/* ---- freeze-out hypersurface (Fluid::outputSurface, src/fld.cpp:517-874;
 * Cornelius, src/cornelius.cpp:780-2205) ----
 * Surface elements of e = e_crit between a lower time level and the current
 * state at tau, one or more per interior hypercube (ix, iy, iz in [2, n-3)),
 * in the reference's loop order.  lower = VH_FO_QPREV: the previous step's Q
 * at tau - dt (outputSurface after every step, dt = the Fluid's dt);
 * VH_FO_SNAPSHOT: the state saved by vh_fo_snapshot (dt ignored).  rows (may
 * be NULL): up to max_rows records of 27 doubles (29 with extend) in the
 * column order of the reference's freezeout.dat (OUTPI): tau, x, y, eta;
 * dsigma_mu (4); u^mu (4); T, mu_B, mu_Q, mu_S; pi^{mu nu} (10, Cartesian,
 * index44 order); Pi [; e, n_b].  *nrows: the element count.  stats (may be
 * NULL): sums over the elements (the reference's diagnostics).  Single-grid
 * contexts only. */
enum { VH_FO_QPREV = 0, VH_FO_SNAPSHOT = 1 };
typedef struct {
  long nelements, nsusp;       /* elements; those with e_C outside [e_crit/2, 2 e_crit] */
  double dsigma[4];            /* sum of dsigma_mu */
  double v_eff;                /* sum of dsigma_mu u^mu (vEff increment) */
  double e_surf, nb_surf;      /* EtotSurf and nbSurf increments */
  int err;                     /* hypercubes the construction could not resolve (bits) */
} vh_fo_stats_t;
int vh_fo_snapshot(vh_ctx_t *ctx, double tau);
int vh_freezeout(vh_ctx_t *ctx, double tau, double dt, double e_crit, int extend, int lower,
                 double *rows, long max_rows, long *nrows, vh_fo_stats_t *stats);

/* ---- vorticity: the thermal-vorticity gradient dbeta[i][j] = d_i (u^j / T)
 * (the reference's `vorticity` key: Hydro::enableVorticity, src/hdo.cpp:60-65;
 * formed in NSquant, src/hdo.cpp:420-611; Cell::setDbeta / getDbeta,
 * src/cll.h:167-204) ----
 * vh_set_vorticity(ctx, 1) allocates 16 doubles per cell, sets every cell's
 * matrix to zero (Cell::enableVorticity) and makes every later viscous step
 * form it as the reference does, including its sharing of one matrix across
 * the cells of a step (entries NSquant does not form keep the values of the
 * cell before in the loop order ix, iy, iz).  0 frees it.  Single-grid
 * contexts; with VH_PREC_MIXED the matrix is formed in FP64 from the mixed
 * step's states (the same states as its velocity gradients).  Ideal-fluid
 * steps leave the matrices unchanged, as in the reference (NSquant runs only
 * in the viscous step). */
int vh_set_vorticity(vh_ctx_t *ctx, int on);
/* Cell::getDbeta of every cell after the last step: dbeta, 16 x N SoA, row
 * 4 i + j = dbeta[i][j] (tau, x, y, eta; the eta row and column as the
 * reference forms them), clamped to [-100, 100] like Cell::setDbeta. */
int vh_get_dbeta(vh_ctx_t *ctx, double *dbeta);
/* vh_freezeout plus, in beta (may be NULL; needs vh_set_vorticity), up to
 * max_rows records of VH_BETA_COLS doubles in the column order of the
 * reference's beta.dat (Fluid::outputSurface with vorticity,
 * src/fld.cpp:663-834): tau, x, y, eta; dsigma_mu (4); u^mu (4); T, mu_B,
 * mu_Q, mu_S; d_i beta_j in Cartesian coordinates (16, i major: the
 * corners' dbeta interpolated to the centroid, both indices transformed with
 * the Milne-to-Cartesian Jacobian, j lowered); e. */
enum { VH_BETA_COLS = 33 };
int vh_freezeout_beta(vh_ctx_t *ctx, double tau, double dt, double e_crit, int extend,
                      int lower, double *rows, double *beta, long max_rows, long *nrows,
                      vh_fo_stats_t *stats);
/* Fluid::outputCorona (src/fld.cpp:876-1072): the elements of the initial
 * state at tau (called once, before the first step): every interior cube
 * (ix, iy, iz in [2, n-3)) whose 8 corners have e <= e_crit and whose centre
 * state, the corners' mean Q / tau, has e >= 0.01, is one element at the cube
 * centre with dsigma_mu = tau dx dy dz (cosh eta, 0, 0, -sinh eta).  rows,
 * max_rows, *nrows as vh_freezeout (27 or 29 columns, the reference's
 * element order); stats as vh_freezeout (nsusp 0).  Its element count is the
 * "Number of corona cells" of the beta.dat header. */
int vh_corona(vh_ctx_t *ctx, double tau, double e_crit, int extend, double *rows, long max_rows,
              long *nrows, vh_fo_stats_t *stats);
/* The Cornelius construction alone on the GPU (Cornelius::find_surface_4d):
 * n hypercubes h (n x 16 corner values, index [t][x][y][z]), edge lengths
 * dx[4] (t, x, y, z), level E0; cen, nrm: n x 10 x 4 (unused elements zero);
 * nel: elements per hypercube (negative: the construction failed). */
int vh_cornelius(const double *h, long n, const double *dx, double E0, double *cen,
                 double *nrm, int *nel);
/* Achieved HBM bandwidth (GB/s, read + write) of a plain 16-byte-vector copy
 * of `bytes` bytes, averaged over reps launches (no counterpart in the
 * reference: the measured bandwidth reference of the rooflines). */
int vh_copy_bandwidth(long bytes, int reps, double *gbs);
/* device/driver information line (name, CUs, clocks) into buf */
int vh_device_info(char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif /* VHLLE_GPU_H */
