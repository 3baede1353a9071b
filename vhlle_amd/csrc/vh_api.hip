// C-ABI of the vHLLE hydro step (include/vhlle_gpu.h): context, device
// memory, the step sequence of Hydro::performStep (src/hdo.cpp:1016-1134) as
// kernel launches on one HIP stream, and its replay as HIP graphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vhlle_gpu.h"
#include "vh_eos.h"
#include "vh_kernels.h"
#include "vh_kernels_mixed.h"
#include "vh_mlp_host.h"
#include "vh_freezeout.h"
#include "vh_vort.h"

#define VH_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      std::fprintf(stderr, "vhlle_gpu: %s failed: %s (%s:%d)\n", #call,      \
                   hipGetErrorString(e_), __FILE__, __LINE__);               \
      return VH_ERR_HIP;                                                     \
    }                                                                        \
  } while (0)

namespace {

// Device scratch of vh_cells_noalloc: the cell list (VH_CELLS_NOALLOC_MAX
// indices) and the records (VH_CELL_COLS doubles per cell), 1.4 MB held on
// each GPU the library runs on until the process ends.  Host code never uses
// the name `scratch` as an address (in host code it is a stand-in, not the
// device array): it uses the address hipGetSymbolAddress returns for each
// device, kept in scratch_ptr.
constexpr long kScratchDoubles = (long)VH_CELLS_NOALLOC_MAX * (1 + VH_CELL_COLS);
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
double *scratch_ptr[kMaxDevices] = {};
std::mutex scratch_mu;

// Looks up this device's scratch address once (the device current); called
// at context creation, so that vh_cells_noalloc itself allocates nothing.
int scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return VH_ERR_ARG;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_ptr[dev]) return VH_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    std::fprintf(stderr, "vhlle_gpu: hipGetSymbolAddress(scratch) failed on device %d: %s\n",
                 dev, hipGetErrorString(e));
    return VH_ERR_HIP;
  }
  scratch_ptr[dev] = static_cast<double *>(ptr);
  return VH_OK;
}

struct KTime {
  int launches = 0;
  double ms = 0.;
};

}  // namespace

struct vh_ctx {
  vh::Geom g;
  vh::DevEos eos;
  double *eos_buf = nullptr;  // tables of a VH_EOS_TABLE/CHIRAL/HADRON EoS (device)
  vh::DevEos eos_fo;          // surface EoS of the freeze-out outputs (vh_set_surface_eos)
  double *eos_fo_buf = nullptr;
  bool has_eos_fo = false;
  vh::fo::Work fo_work;       // freeze-out buffers, kept between calls
  vh::DevTrc trc;
  vh_grid_t grid;
  int precision = 0, c2p_mode = 0;
  bool viscous = false;
  int dev = 0;
  hipStream_t st = nullptr;
  // device memory (one allocation)
  char *pool = nullptr;
  double *Q[2] = {nullptr, nullptr}, *Qh = nullptr;
  double *pi = nullptr, *piH = nullptr, *Pi = nullptr, *PiH = nullptr;
  double *pi0 = nullptr, *piH0 = nullptr, *Pi0 = nullptr, *PiH0 = nullptr;
  double *m = nullptr, *dm = nullptr, *vcc = nullptr;
  double *F[3] = {nullptr, nullptr, nullptr};
  unsigned char *fl[3] = {nullptr, nullptr, nullptr};
  double *ph = nullptr, *pt = nullptr, *vf = nullptr;
  double *tp = nullptr;
  int *status = nullptr;
  // HLLE signal speeds (bit patterns of non-negative doubles): sig_acc, nine
  // slot groups (vh::kSigGroup words each) the flux kernels add to: the
  // running step's predictor and corrector (x, y, eta), then the autotuner's
  // scratch; sig: [6..11] the last completed step's maxima, [12..14] the
  // per-axis maxima since the last reset
  unsigned long long *sig = nullptr, *sig_acc = nullptr;
  // learned con2prim: MFMA fragments of the installed MLP and the per-cell
  // guesses (mixed precision only; gv reuses the FP64 path's F[0] storage)
  vh::ml::Packed *mlp = nullptr;
  float *gv = nullptr;
  // vorticity (vh_set_vorticity): per-cell dbeta entries and op codes written
  // by the relaxation kernel, resolved on demand (dirty: a step ran since)
  bool vort = false, vort_dirty = false;
  double *db = nullptr, *vT = nullptr;  // vT: T of the neighbour states (k_vprim)
  double *pf = nullptr;                 // primitive fields (vh_prim_fields)
  double *fsb = nullptr;                // grid totals: 16 x 256 partial sums + 16
  long *cidx = nullptr;                 // vh_cells: cell list and records
  double *cbuf = nullptr;
  long ccap = 0;
  unsigned char *dop = nullptr;
  void *vscr = nullptr;
  // freeze-out snapshot (the lower time level of vh_freezeout)
  double *fo_snap = nullptr;
  double fo_tau = 0.;
  // grids of the context (one, or a batch of independent grids sharing the
  // state arrays) and their device lookup tables (mixed precision kernels)
  std::vector<vh::Geom> evh;
  vh::mx::Grids G = {};
  char *gtab = nullptr;
  bool fusable = true, degenerate = false;
  // some n_b, n_q or n_s nonzero: the mixed kernels carry 7 conserved
  // components, else 4 (charges that are zero everywhere stay zero)
  bool charged = true;
  // stage kernel shape (mixed precision): kind 0 until the autotuner ran.
  // tune_stage 1: tuned on the state of the first run (tile_alt: the best
  // shape of the other kind); 2: tuned again on the developed flow
  vh::mx::Tile tile = {0, 0, 0}, tile_alt = {0, 0, 0};
  bool tile_forced = false;
  double tile_ms = 0.;
  int tune_stage = 0;
  long steps_since_tune = 0;
  int cur = 0;
  int m_full = 0;
  // graph cache: one executable per parity of cur for a single step, and
  // one per parity for kGraphSteps steps in a row
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipGraphExec_t gmulti[2] = {nullptr, nullptr};
  double g_dt = 0.;
  int g_mfull = -1;
  // timing
  bool timing = false;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spare;
  std::map<std::string, KTime> times;
};

namespace {

template <class F>
int timed(vh_ctx *c, const char *name, F &&launch) {
  if (!c->timing) {
    launch();
    return VH_OK;
  }
  std::pair<hipEvent_t, hipEvent_t> ev;
  if (!c->spare.empty()) {
    ev = c->spare.back();
    c->spare.pop_back();
  } else {
    VH_HIP(hipEventCreate(&ev.first));
    VH_HIP(hipEventCreate(&ev.second));
  }
  VH_HIP(hipEventRecord(ev.first, c->st));
  launch();
  VH_HIP(hipEventRecord(ev.second, c->st));
  c->pending.push_back({name, ev});
  return VH_OK;
}

int collect_times(vh_ctx *c) {
  for (auto &p : c->pending) {
    float ms = 0.f;
    VH_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    KTime &t = c->times[p.first];
    t.launches++;
    t.ms += ms;
    c->spare.push_back(p.second);
  }
  c->pending.clear();
  return VH_OK;
}

const char *face_name(int D, int mode) {
  static const char *n[2][3] = {{"faces_x_pred", "faces_y_pred", "faces_z_pred"},
                                {"faces_x_corr", "faces_y_corr", "faces_z_corr"}};
  return n[mode][D - 1];
}

// Enqueues one Hydro::performStep on c->st (src/hdo.cpp:1016-1134).  Reads
// tau from c->tp, advances it on the device.  Q[cur] -> Q[cur^1].
// Ghost-cell fill of the reference's z, y, x passes (src/fld.cpp:211-382):
// one fused kernel when every extent is 1 or >= 5, else the three passes.
void ghost_fill(vh_ctx *c, int full, double *Qn, hipStream_t st, int &rc) {
  const vh::Geom &g = c->g;
  const int pf = c->precision == VH_PREC_MIXED;
  const char *name = full ? "ghost_full" : "ghost_ideal";
  if (pf && c->fusable) {
    rc |= timed(c, name, [&] {
      vh::mx::launch_ghost_cells(full, c->charged ? 7 : 4, c->G, c->degenerate, Qn, c->Qh,
                                 reinterpret_cast<float *>(c->pi),
                                 reinterpret_cast<float *>(c->Pi), st);
    });
    return;
  }
  if (vh::ghost_fusable(g)) {
    rc |= timed(c, name, [&] { vh::launch_ghost1(full, pf, g, Qn, c->Qh, c->pi, c->Pi, st); });
    return;
  }
  for (int ax = 3; ax >= 1; ax--)
    rc |= timed(c, name, [&] {
      if (full && pf)
        vh::mx::launch_ghost_full(ax, g, Qn, reinterpret_cast<float *>(c->pi),
                                  reinterpret_cast<float *>(c->Pi), st);
      else
        vh::launch_ghost(full, ax, g, Qn, c->Qh, c->pi, c->Pi, st);
    });
}

// FP32 views of the pi-like arrays in mixed precision (same allocation)
inline float *fp(double *a) { return reinterpret_cast<float *>(a); }
// mixed precision keeps pi0/Pi0/piH0/PiH0 as six planes of N float4 (96 N
// bytes); they occupy the adjacent, 256-aligned pi0 and piH0 arrays (160 N)
inline float4 *rec(vh_ctx *c) { return reinterpret_cast<float4 *>(c->pi0); }

// mixed precision: FP64 Q/Qh/con2prim, FP32 fluxes, viscous sources and pi
int enqueue_step_mixed(vh_ctx *c, double dt) {
  const vh::Geom &g = c->g;
  hipStream_t st = c->st;
  double *Qo = c->Q[c->cur], *Qn = c->Q[c->cur ^ 1];
  const int cm = c->c2p_mode, nc = c->charged ? 7 : 4;
  const vh::mx::Grids &G = c->G;
  int rc = VH_OK;
  if (!c->m_full) {
    if (c->evh.size() > 1) return VH_ERR_ARG;  // batches have no vacuum fronts
    rc |= timed(c, "update_m_dm", [&] { vh::launch_update_m(0, g, c->m, c->dm, c->tp, dt, st); });
    rc |= timed(c, "update_m_apply", [&] { vh::launch_update_m(1, g, c->m, c->dm, c->tp, dt, st); });
  }
  const bool learned = cm == VH_C2P_LEARNED;
  // the MLP guess of every cell of the step's input state Q/tau_n: the Newton
  // start of the predictor, corrector and viscous con2prim of the step (the
  // states they solve differ from it by at most one step's change)
  if (learned)
    rc |= timed(c, "c2p_guess", [&] { vh::ml::launch_guess(Qo, g.N, nc, c->tp, 0., c->mlp, c->gv, st); });
  auto stage = [&](int mode, double *out) {
    vh::mx::launch_stage_tile(mode, cm, nc, G, c->eos, Qo, c->Qh, c->gv, c->m, c->m_full, c->tp,
                              dt, out, c->status, c->sig_acc + 3 * mode * vh::kSigGroup, c->tile, st);
  };
  rc |= timed(c, "stage_pred", [&] { stage(0, c->Qh); });
  rc |= timed(c, "stage_corr", [&] { stage(1, Qn); });
  rc |= timed(c, "advance_tau", [&] { vh::launch_advance_tau(c->tp, dt, c->sig_acc, c->sig, st); });
  ghost_fill(c, 0, Qn, st, rc);
  if (c->viscous) {
    rc |= timed(c, "visc_prim", [&] { vh::mx::launch_vprim(cm, nc, g.N, c->eos, c->Qh, c->gv, c->tp, dt, c->ph, c->pt, st); });
    rc |= timed(c, "visc_relax", [&] {
      vh::mx::launch_visc1(cm, nc, G, c->eos, c->trc, Qo, Qn, c->gv, c->ph, c->pt, c->m, c->m_full,
                           fp(c->pi), fp(c->Pi), c->tp, dt, rec(c), c->status,
                           c->vort ? c->db : nullptr, c->dop, st);
    });
    rc |= timed(c, "visc_advect", [&] {
      vh::mx::launch_visc2(G, c->ph, rec(c), fp(c->pi), fp(c->piH), fp(c->Pi), fp(c->PiH), c->vcc,
                           fp(c->vf), c->tp, dt, st);
    });
    rc |= timed(c, "visc_flux_update", [&] {
      vh::mx::launch_visc3(G, c->ph, fp(c->piH), fp(c->PiH), fp(c->vf), c->m, c->m_full, c->tp,
                           dt, Qn, st);
    });
  }
  ghost_fill(c, 1, Qn, st, rc);
  VH_HIP(hipGetLastError());
  c->cur ^= 1;
  return rc ? VH_ERR_HIP : VH_OK;
}

int enqueue_step_fp64(vh_ctx *c, double dt) {
  const vh::Geom &g = c->g;
  hipStream_t st = c->st;
  double *Qo = c->Q[c->cur], *Qn = c->Q[c->cur ^ 1];
  int rc = VH_OK;
  // Fluid::updateM (src/fld.cpp:384-423): the identity once every m == 1
  if (!c->m_full) {
    rc |= timed(c, "update_m_dm", [&] { vh::launch_update_m(0, g, c->m, c->dm, c->tp, dt, st); });
    rc |= timed(c, "update_m_apply", [&] { vh::launch_update_m(1, g, c->m, c->dm, c->tp, dt, st); });
  }
  const int ext[3] = {g.nx, g.ny, g.nz};
  for (int mode = 0; mode < 2; mode++) {
    const double *src = mode == 0 ? Qo : c->Qh;
    for (int D = 1; D <= 3; D++) {
      if (ext[D - 1] < 2) continue;  // no interfaces along a 1-cell axis
      rc |= timed(c, face_name(D, mode), [&] {
        vh::launch_faces(D, mode, g, c->eos, src, Qo, c->m, c->m_full, c->tp, dt, c->F[D - 1],
                         c->fl[D - 1], c->status,
                         c->sig_acc + (3 * mode + D - 1) * vh::kSigGroup, st);
      });
    }
    const double *Fs[3] = {c->F[0], c->F[1], c->F[2]};
    const unsigned char *fs[3] = {c->fl[0], c->fl[1], c->fl[2]};
    rc |= timed(c, mode == 0 ? "update_pred" : "update_corr", [&] {
      vh::launch_update(mode, g, c->eos, Qo, c->Qh, Fs, fs, c->tp, dt, mode == 0 ? c->Qh : Qn,
                        c->status, st);
    });
  }
  rc |= timed(c, "advance_tau", [&] { vh::launch_advance_tau(c->tp, dt, c->sig_acc, c->sig, st); });
  // Fluid::correctImagCells (src/fld.cpp:211-273): z, y, x
  ghost_fill(c, 0, Qn, st, rc);
  if (c->viscous) {
    rc |= timed(c, "visc_prim", [&] {
      vh::launch_vprim(g, c->eos, c->Qh, c->tp, dt, c->ph, c->pt, c->vort ? c->vT : nullptr, st);
    });
    rc |= timed(c, "visc_relax", [&] {
      vh::launch_visc1(g, c->eos, c->trc, Qo, Qn, c->ph, c->pt, c->m, c->m_full, c->pi, c->Pi,
                       c->tp, dt, c->pi0, c->piH0, c->Pi0, c->PiH0, c->status,
                       c->vort ? c->db : nullptr, c->dop, c->vT, st);
    });
    rc |= timed(c, "visc_advect", [&] {
      vh::launch_visc2(g, c->ph, c->pi0, c->piH0, c->Pi0, c->PiH0, c->pi, c->piH, c->Pi, c->PiH,
                       c->vcc, c->vf, c->tp, dt, st);
    });
    rc |= timed(c, "visc_flux_update", [&] {
      vh::launch_visc3(g, c->ph, c->piH, c->PiH, c->vf, c->m, c->m_full, c->tp, dt, Qn, st);
    });
  }
  // Fluid::correctImagCellsFull (src/fld.cpp:275-382): z, y, x
  ghost_fill(c, 1, Qn, st, rc);
  VH_HIP(hipGetLastError());
  c->cur ^= 1;
  return rc ? VH_ERR_HIP : VH_OK;
}

int enqueue_step(vh_ctx *c, double dt) {
  return c->precision == VH_PREC_MIXED ? enqueue_step_mixed(c, dt) : enqueue_step_fp64(c, dt);
}

void drop_graphs(vh_ctx *c) {
  for (int i = 0; i < 2; i++) {
    if (c->gexec[i]) {
      (void)hipGraphExecDestroy(c->gexec[i]);
      c->gexec[i] = nullptr;
    }
    if (c->gmulti[i]) {
      (void)hipGraphExecDestroy(c->gmulti[i]);
      c->gmulti[i] = nullptr;
    }
  }
  c->g_mfull = -1;
}

// Steps per multi-step graph: a run replays one graph launch per kGraphSteps
// steps (even, so that the graph ends on the parity of Q it starts from).
constexpr int kGraphSteps = 8;

// clears the signal-speed maxima and slot groups (kernels, on the stream)
void clear_sig(vh_ctx *c) {
  vh::launch_clear_u64(c->sig, 32, c->st);
  vh::launch_clear_u64(c->sig_acc, 9L * vh::kSigGroup, c->st);
}

bool any_nonzero(const double *a, long n) {
  for (long i = 0; i < n; i++)
    if (a[i] != 0.) return true;
  return false;
}

// switches the mixed kernels between 7 and 4 conserved components
void set_charged(vh_ctx *c, bool ch) {
  if (c->charged == ch) return;
  c->charged = ch;
  if (!c->tile_forced) c->tile.kind = 0;  // retune for the other kernels
  drop_graphs(c);
}

int build_graphs(vh_ctx *c, double dt) {
  if (c->gexec[0] && c->gexec[1] && c->gmulti[0] && c->gmulti[1] && c->g_dt == dt &&
      c->g_mfull == c->m_full)
    return VH_OK;
  drop_graphs(c);
  const int cur0 = c->cur;
  for (int k = 0; k < 4; k++) {
    const int par = cur0 ^ (k & 1);  // graphs for steps that start from Q[par]
    const int nst = k < 2 ? 1 : kGraphSteps;
    c->cur = par;
    hipGraph_t graph;
    VH_HIP(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    const bool t = c->timing;
    c->timing = false;
    int rc = VH_OK;
    for (int s = 0; s < nst && rc == VH_OK; s++) rc = enqueue_step(c, dt);
    c->timing = t;
    VH_HIP(hipStreamEndCapture(c->st, &graph));
    if (rc != VH_OK) {
      (void)hipGraphDestroy(graph);
      c->cur = cur0;
      return rc;
    }
    VH_HIP(hipGraphInstantiate(k < 2 ? &c->gexec[par] : &c->gmulti[par], graph, nullptr,
                               nullptr, 0));
    VH_HIP(hipGraphDestroy(graph));
  }
  c->cur = cur0;
  c->g_dt = dt;
  c->g_mfull = c->m_full;
  return VH_OK;
}

int refresh_m_full(vh_ctx *c) {
  std::vector<double> h((size_t)3 * c->g.N);
  VH_HIP(hipMemcpyAsync(h.data(), c->m, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  VH_HIP(hipStreamSynchronize(c->st));
  int full = 1;
  for (double v : h)
    if (!(v >= 1.)) {
      full = 0;
      break;
    }
  c->m_full = full;
  return VH_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// Fluid ctor geometry (src/fld.cpp:77-79)
vh::Geom geom_of(const vh_grid_t &gr) {
  vh::Geom g;
  g.nx = gr.nx;
  g.ny = gr.ny;
  g.nz = gr.nz;
  g.N = (long)g.nx * g.ny * g.nz;
  g.dx = (gr.xmax - gr.xmin) / (g.nx - 1);
  g.dy = (gr.ymax - gr.ymin) / (g.ny - 1);
  g.dz = (gr.zmax - gr.zmin) / (g.nz - 1);
  return g;
}

bool fusable_extent(int n) { return n == 1 || n >= 5; }

// device lookup tables of the grids (vh::mx::Grids)
int build_grid_tables(vh_ctx *c) {
  const int nev = (int)c->evh.size();
  std::vector<int> bev, bfirst(nev), sev, sfirst(nev);
  std::vector<int> bsev[3], bsfirst[3];  // band segments of CY = 2, 4, 8
  int maxnz = 1;
  for (int e = 0; e < nev; e++) {
    const vh::Geom &g = c->evh[e];
    bfirst[e] = (int)bev.size();
    for (long b = 0; b < (g.N + 255) / 256; b++) bev.push_back(e);
    sfirst[e] = (int)sev.size();
    const long nxy = (long)g.nx * g.ny;
    for (long q = 0; q < (nxy + 62) / 63; q++) sev.push_back(e);
    for (int h = 0; h < 3; h++) {
      const long CY = 2L << h, nbp = (g.ny + CY - 1) / CY * g.nx;
      bsfirst[h].push_back((int)bsev[h].size());
      for (long q = 0; q < (nbp + 62) / 63; q++) bsev[h].push_back(e);
    }
    maxnz = std::max(maxnz, g.nz);
  }
  const size_t al = 256;
  auto rnd = [&](size_t b) { return (b + al - 1) / al * al; };
  size_t sz = rnd(nev * sizeof(vh::Geom)) + rnd(bev.size() * 4) + rnd(nev * 4) +
              rnd(sev.size() * 4) + rnd(nev * 4);
  for (int h = 0; h < 3; h++) sz += rnd(bsev[h].size() * 4) + rnd(nev * 4);
  VH_HIP(hipMalloc(&c->gtab, sz));
  char *p = c->gtab;
  auto put = [&](const void *src, size_t bytes) -> char * {
    char *d = p;
    if (hipMemcpy(d, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    p += rnd(bytes);
    return d;
  };
  char *ev = put(c->evh.data(), nev * sizeof(vh::Geom));
  char *pb = put(bev.data(), bev.size() * 4), *pbf = put(bfirst.data(), nev * 4);
  char *ps = put(sev.data(), sev.size() * 4), *psf = put(sfirst.data(), nev * 4);
  if (!ev || !pb || !pbf || !ps || !psf) return VH_ERR_HIP;
  vh::mx::Grids &G = c->G;
  for (int h = 0; h < 3; h++) {
    char *a = put(bsev[h].data(), bsev[h].size() * 4), *b = put(bsfirst[h].data(), nev * 4);
    if (!a || !b) return VH_ERR_HIP;
    G.bsev[h] = reinterpret_cast<const int *>(a);
    G.bsfirst[h] = reinterpret_cast<const int *>(b);
    G.nbseg[h] = (int)bsev[h].size();
  }
  G.ev = reinterpret_cast<const vh::Geom *>(ev);
  G.bev = reinterpret_cast<const int *>(pb);
  G.bfirst = reinterpret_cast<const int *>(pbf);
  G.sev = reinterpret_cast<const int *>(ps);
  G.sfirst = reinterpret_cast<const int *>(psf);
  G.P = c->g.N;
  G.nev = nev;
  G.nblk = (int)bev.size();
  G.nseg = (int)sev.size();
  G.maxnz = maxnz;
  G.g0 = c->evh[0];
  G.sy = VH_STRIP_ROWS;
  return VH_OK;
}

int create_impl(const vh_grid_t *grids, int ngrid, const vh_eos_t *eos, const vh_trcoeff_t *trc,
                int precision, vh_ctx_t **out) {
  if (!grids || ngrid < 1 || !eos || !trc || !out) return VH_ERR_ARG;
  if (eos->kind == VH_EOS_TABLE) {  // validate before allocating anything
    std::vector<double> tab;
    int n = 0;
    if (vh::eos_table_host(*eos, tab, n) != VH_OK) return VH_ERR_ARG;
  } else if (eos->kind == VH_EOS_CHIRAL || eos->kind == VH_EOS_HADRON) {
    // charge-dependent EoSs: the FP64 path only
    if (precision != VH_PREC_FP64) return VH_ERR_ARG;
    vh::DevEos de = vh::DevEos();
    de.kind = eos->kind;
    if (vh::charged_tables_host(*eos, de, nullptr, nullptr) != VH_OK) return VH_ERR_ARG;
  } else if (eos->kind != VH_EOS_IDEAL) {
    return VH_ERR_ARG;
  }
  if (precision != VH_PREC_FP64 && precision != VH_PREC_MIXED) return VH_ERR_ARG;
  if (ngrid > 1 && precision != VH_PREC_MIXED) return VH_ERR_ARG;
  long ncell = 0;
  for (int e = 0; e < ngrid; e++) {
    const vh_grid_t &gr = grids[e];
    if (gr.nx < 1 || gr.ny < 1 || gr.nz < 1) return VH_ERR_ARG;
    if (ngrid > 1 && !(fusable_extent(gr.nx) && fusable_extent(gr.ny) && fusable_extent(gr.nz)))
      return VH_ERR_ARG;
    ncell += (long)gr.nx * gr.ny * gr.nz;
  }
  if (ncell > (1L << 31) / 4) return VH_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return VH_ERR_NODEVICE;
  vh_ctx *c = new vh_ctx;
  c->grid = grids[0];
  c->precision = precision;
  for (int e = 0; e < ngrid; e++) {
    vh::Geom g = geom_of(grids[e]);
    c->fusable &= fusable_extent(g.nx) && fusable_extent(g.ny) && fusable_extent(g.nz);
    c->degenerate |= g.nx == 1 || g.ny == 1 || g.nz == 1;
    c->evh.push_back(g);
  }
  long base = 0;
  for (auto &g : c->evh) {
    g.base = base;
    g.P = ncell;
    base += g.N;
  }
  // the context's own geometry: the grid itself, or (batches) only the total
  if (ngrid == 1) {
    c->g = c->evh[0];
  } else {
    c->g = vh::Geom();
    c->g.nx = c->g.ny = c->g.nz = 0;
    c->g.dx = c->g.dy = c->g.dz = 0.;
  }
  c->g.N = ncell;
  c->g.base = 0;
  c->g.P = ncell;
  const vh::Geom &g = c->g;
  c->trc.c = *trc;
  // TransportCoeff::isViscous (src/trancoeff.h:39-44)
  c->viscous = trc->etaS0 > 0. || trc->zetaS0 > 0. || trc->etaSparam > 0 || trc->zetaSparam > 0;
  if (hipGetDevice(&c->dev) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess) {
    delete c;
    return VH_ERR_HIP;
  }
  if (scratch_setup(c->dev) != VH_OK) {
    (void)hipStreamDestroy(c->st);
    delete c;
    return VH_ERR_HIP;
  }
  // one pool, every array 256-byte aligned
  const long N = g.N;
  const size_t al = 256;
  auto rnd = [&](size_t b) { return (b + al - 1) / al * al; };
  const size_t dN = rnd(N * sizeof(double)), bN = rnd(N);
  const size_t ndoubles_arrays = 14 + 7 + 20 + 2 + 20 + 2 + 3 + 3 + 1 + 21 + 8 + 4 + 4 + 1;
  const size_t total = ndoubles_arrays * dN + 3 * bN + 3 * al +
                       rnd(9 * vh::kSigGroup * sizeof(unsigned long long));
  if (hipMalloc(&c->pool, total) != hipSuccess) {
    (void)hipStreamDestroy(c->st);
    delete c;
    return VH_ERR_HIP;
  }
  char *p = c->pool;
  // component k of a field lives at base + k*N: pack components contiguously
  auto takef = [&](int ncomp) {
    double *r = reinterpret_cast<double *>(p);
    p += rnd((size_t)ncomp * N * sizeof(double));
    return r;
  };
  c->Q[0] = takef(7);
  c->Q[1] = takef(7);
  c->Qh = takef(7);
  c->pi = takef(10);
  c->piH = takef(10);
  c->Pi = takef(1);
  c->PiH = takef(1);
  c->pi0 = takef(10);
  c->piH0 = takef(10);
  c->Pi0 = takef(1);
  c->PiH0 = takef(1);
  c->m = takef(3);
  c->dm = takef(3);
  c->vcc = takef(1);
  for (int d = 0; d < 3; d++) c->F[d] = takef(7);
  c->ph = takef(8);
  c->pt = takef(4);
  c->vf = takef(4);
  c->tp = takef(1);
  c->gv = reinterpret_cast<float *>(c->F[0]);  // the fused mixed stages use no F
  c->status = reinterpret_cast<int *>(p);
  p += al;
  c->sig = reinterpret_cast<unsigned long long *>(p);
  p += al;
  c->sig_acc = reinterpret_cast<unsigned long long *>(p);
  p += rnd(9 * vh::kSigGroup * sizeof(unsigned long long));
  if (vh::make_dev_eos(eos, c->status, c->eos, &c->eos_buf) != VH_OK) {
    vh_destroy(c);
    return VH_ERR_HIP;
  }
  for (int d = 0; d < 3; d++) {
    c->fl[d] = reinterpret_cast<unsigned char *>(p);
    p += bN;
  }
  // Cell ctor defaults (src/cll.cpp:33-52): zeros, viscCorrCut = 1, m = 0
  if (hipMemsetAsync(c->pool, 0, total, c->st) != hipSuccess) {
    vh_destroy(c);
    return VH_ERR_HIP;
  }
  std::vector<double> ones(N, 1.0);
  if (hipMemcpyAsync(c->vcc, ones.data(), N * sizeof(double), hipMemcpyHostToDevice, c->st) !=
          hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess) {
    vh_destroy(c);
    return VH_ERR_HIP;
  }
  if (build_grid_tables(c) != VH_OK) {
    vh_destroy(c);
    return VH_ERR_HIP;
  }
  *out = c;
  return VH_OK;
}

}  // namespace

extern "C" {

int vh_create(const vh_grid_t *grid, const vh_eos_t *eos, const vh_trcoeff_t *trc,
              int precision, vh_ctx_t **out) {
  if (!grid) return VH_ERR_ARG;
  return create_impl(grid, 1, eos, trc, precision, out);
}

int vh_create_batch(int ngrid, const vh_grid_t *grids, const vh_eos_t *eos,
                    const vh_trcoeff_t *trc, int precision, vh_ctx_t **out) {
  return create_impl(grids, ngrid, eos, trc, precision, out);
}

int vh_destroy(vh_ctx_t *c) {
  if (!c) return VH_ERR_ARG;
  if (c->st) (void)hipStreamSynchronize(c->st);
  drop_graphs(c);
  for (auto &p : c->pending) c->spare.push_back(p.second);
  for (auto &e : c->spare) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  if (c->pool) (void)hipFree(c->pool);
  if (c->gtab) (void)hipFree(c->gtab);
  if (c->fo_snap) (void)hipFree(c->fo_snap);
  if (c->db) (void)hipFree(c->db);
  if (c->vT) (void)hipFree(c->vT);
  if (c->pf) (void)hipFree(c->pf);
  if (c->fsb) (void)hipFree(c->fsb);
  if (c->cidx) (void)hipFree(c->cidx);
  if (c->cbuf) (void)hipFree(c->cbuf);
  if (c->dop) (void)hipFree(c->dop);
  if (c->vscr) (void)hipFree(c->vscr);
  if (c->mlp) (void)hipFree(c->mlp);
  if (c->eos_buf) (void)hipFree(c->eos_buf);
  if (c->eos_fo_buf) (void)hipFree(c->eos_fo_buf);
  c->fo_work.release();
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return VH_OK;
}

long vh_ncell(const vh_ctx_t *c) { return c ? c->g.N : 0; }

int vh_upload(vh_ctx_t *c, const vh_state_t *h) {
  if (!c || !h) return VH_ERR_ARG;
  {
    bool ch = false;
    for (const double *a : {h->Q, h->Qprev, h->Qh}) ch |= a && any_nonzero(a + 4 * c->g.N, 3 * c->g.N);
    set_charged(c, ch);
  }
  const long N = c->g.N;
  const size_t d = sizeof(double);
  auto put = [&](double *dst, const double *src, int nc, double dflt) -> int {
    if (src) {
      VH_HIP(hipMemcpyAsync(dst, src, nc * N * d, hipMemcpyHostToDevice, c->st));
    } else {
      std::vector<double> z((size_t)nc * N, dflt);
      VH_HIP(hipMemcpyAsync(dst, z.data(), nc * N * d, hipMemcpyHostToDevice, c->st));
      VH_HIP(hipStreamSynchronize(c->st));
    }
    return VH_OK;
  };
  int rc = VH_OK;
  rc |= put(c->Q[c->cur], h->Q, 7, 0.);
  rc |= put(c->Q[c->cur ^ 1], h->Qprev, 7, 0.);
  rc |= put(c->Qh, h->Qh, 7, 0.);
  if (c->precision == VH_PREC_MIXED) {  // pi, Pi stored in FP32
    auto putf = [&](double *dst, const double *src, int nc) -> int {
      std::vector<float> f((size_t)nc * N, 0.f);
      if (src)
        for (size_t i = 0; i < f.size(); i++) f[i] = (float)src[i];
      VH_HIP(hipMemcpyAsync(dst, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, c->st));
      VH_HIP(hipStreamSynchronize(c->st));
      return VH_OK;
    };
    rc |= putf(c->pi, h->pi, 10);
    rc |= putf(c->piH, h->piH, 10);
    rc |= putf(c->Pi, h->Pi, 1);
    rc |= putf(c->PiH, h->PiH, 1);
  } else {
    rc |= put(c->pi, h->pi, 10, 0.);
    rc |= put(c->piH, h->piH, 10, 0.);
    rc |= put(c->Pi, h->Pi, 1, 0.);
    rc |= put(c->PiH, h->PiH, 1, 0.);
  }
  rc |= put(c->m, h->m, 3, 0.);
  rc |= put(c->vcc, h->viscCorrCut, 1, 1.);
  clear_sig(c);  // a new state: no signal speeds yet
  VH_HIP(hipStreamSynchronize(c->st));
  if (rc != VH_OK) return VH_ERR_HIP;
  return refresh_m_full(c);
}

int vh_download(vh_ctx_t *c, vh_state_t *h) {
  if (!c || !h) return VH_ERR_ARG;
  const long N = c->g.N;
  const size_t d = sizeof(double);
  auto get = [&](double *dst, const double *src, int nc) -> int {
    if (dst) VH_HIP(hipMemcpyAsync(dst, src, nc * N * d, hipMemcpyDeviceToHost, c->st));
    return VH_OK;
  };
  int rc = VH_OK;
  rc |= get(h->Q, c->Q[c->cur], 7);
  rc |= get(h->Qprev, c->Q[c->cur ^ 1], 7);
  rc |= get(h->Qh, c->Qh, 7);
  if (c->precision == VH_PREC_MIXED) {
    auto getf = [&](double *dst, const double *src, int nc) -> int {
      if (!dst) return VH_OK;
      std::vector<float> f((size_t)nc * N);
      VH_HIP(hipMemcpyAsync(f.data(), src, f.size() * sizeof(float), hipMemcpyDeviceToHost, c->st));
      VH_HIP(hipStreamSynchronize(c->st));
      for (size_t i = 0; i < f.size(); i++) dst[i] = (double)f[i];
      return VH_OK;
    };
    rc |= getf(h->pi, c->pi, 10);
    rc |= getf(h->piH, c->piH, 10);
    rc |= getf(h->Pi, c->Pi, 1);
    rc |= getf(h->PiH, c->PiH, 1);
  } else {
    rc |= get(h->pi, c->pi, 10);
    rc |= get(h->piH, c->piH, 10);
    rc |= get(h->Pi, c->Pi, 1);
    rc |= get(h->PiH, c->PiH, 1);
  }
  rc |= get(h->m, c->m, 3);
  rc |= get(h->viscCorrCut, c->vcc, 1);
  VH_HIP(hipStreamSynchronize(c->st));
  return rc ? VH_ERR_HIP : VH_OK;
}

int vh_set_prim(vh_ctx_t *c, double tau, const double *e, const double *nb, const double *nq,
                const double *ns, const double *vx, const double *vy, const double *vz) {
  if (!c || !e || !vx || !vy || !vz) return VH_ERR_ARG;
  const long N = c->g.N;
  std::vector<double> h((size_t)7 * N, 0.);
  const double *src[7] = {e, nb, nq, ns, vx, vy, vz};
  for (int k = 0; k < 7; k++)
    if (src[k]) std::memcpy(&h[(size_t)k * N], src[k], N * sizeof(double));
  double *d = nullptr;
  VH_HIP(hipMalloc(&d, h.size() * sizeof(double)));
  int rc = VH_OK;
  if (hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->st) !=
      hipSuccess)
    rc = VH_ERR_HIP;
  if (rc == VH_OK) {
    bool ch = false;
    for (const double *a : {nb, nq, ns}) ch |= a && any_nonzero(a, N);
    set_charged(c, ch);
    // Qh is not written here: clear its charge components for the
    // 4-component kernels, which leave them untouched
    if (!ch && hipMemsetAsync(c->Qh + 4 * N, 0, 3 * N * sizeof(double), c->st) != hipSuccess)
      rc = VH_ERR_HIP;
  }
  if (rc == VH_OK) {
    vh::launch_set_prim(c->g, c->eos, tau, d, c->Q[c->cur], c->Q[c->cur ^ 1], c->m, c->st);
    clear_sig(c);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->st) != hipSuccess)
      rc = VH_ERR_HIP;
  }
  (void)hipFree(d);
  if (rc != VH_OK) return rc;
  return refresh_m_full(c);
}

int vh_check_dtau(const vh_ctx_t *c, double dt) {
  if (!c) return VH_ERR_ARG;
  for (const vh::Geom &g : c->evh)
    if (dt > g.dx / 2. || dt > g.dy / 2.) return VH_ERR_DTAU;
  return VH_OK;
}

int vh_set_c2p_mode(vh_ctx_t *c, int mode) {
  if (!c) return VH_ERR_ARG;
  // the FP64 path is the faithful one: reference iteration only
  if (c->precision == VH_PREC_FP64 && mode != VH_C2P_REFERENCE) return VH_ERR_ARG;
  if (mode != VH_C2P_REFERENCE && mode != VH_C2P_CLOSED && mode != VH_C2P_LEARNED &&
      mode != VH_C2P_NEWTON)
    return VH_ERR_ARG;
  if (mode == VH_C2P_LEARNED && !c->mlp) return VH_ERR_ARG;  // no model installed
  if (mode == VH_C2P_CLOSED && c->eos.kind != VH_EOS_IDEAL) return VH_ERR_ARG;  // p = e/3 only
  c->c2p_mode = mode;
  if (!c->tile_forced) c->tile.kind = 0;  // retune for this con2prim
  drop_graphs(c);
  return VH_OK;
}

int vh_set_stage(vh_ctx_t *c, int kind, int CZ) {
  if (!c || c->precision != VH_PREC_MIXED) return VH_ERR_ARG;
  vh::mx::Tile t;
  if (kind == VH_STAGE_AUTO) {
    c->tile.kind = 0;
    c->tile_forced = false;
    drop_graphs(c);
    return VH_OK;
  } else if (kind == VH_STAGE_CELLS) {
    t = vh::mx::make_cells();
  } else if (kind == VH_STAGE_ROWS) {
    if (CZ < 1) return VH_ERR_ARG;
    t = vh::mx::make_rows(c->G, CZ);
  } else if (kind == VH_STAGE_BANDS) {
    return vh_set_stage_bands(c, 4, CZ);
  } else {
    return VH_ERR_ARG;
  }
  c->tile = t;
  c->tile_forced = true;
  c->tile_ms = 0.;
  drop_graphs(c);
  return VH_OK;
}

int vh_set_stage_bands(vh_ctx_t *c, int CY, int CZ) {
  if (!c || c->precision != VH_PREC_MIXED || CZ < 1 || vh::mx::band_table(CY) < 0)
    return VH_ERR_ARG;
  c->tile = vh::mx::make_bands(c->G, CY, CZ);
  c->tile_forced = true;
  c->tile_ms = 0.;
  drop_graphs(c);
  return VH_OK;
}

int vh_get_stage(const vh_ctx_t *c, int *kind, int *CZ, double *ms) {
  if (!c || !kind || !CZ) return VH_ERR_ARG;
  *kind = c->tile.kind;
  *CZ = c->tile.CZ;
  if (ms) *ms = c->tile_ms;
  return VH_OK;
}

int vh_get_stage_bands(const vh_ctx_t *c, int *CY) {
  if (!c || !CY) return VH_ERR_ARG;
  *CY = c->tile.kind == VH_STAGE_BANDS ? c->tile.CY : 0;
  return VH_OK;
}

int vh_set_c2p_model(vh_ctx_t *c, const float *params) {
  if (!c || !params) return VH_ERR_ARG;
  if (c->precision != VH_PREC_MIXED) return VH_ERR_ARG;
  vh::ml::Packed hp;
  vh::ml::pack(params, hp);
  VH_HIP(hipStreamSynchronize(c->st));
  if (!c->mlp) VH_HIP(hipMalloc(&c->mlp, sizeof(vh::ml::Packed)));
  VH_HIP(hipMemcpy(c->mlp, &hp, sizeof hp, hipMemcpyHostToDevice));
  return VH_OK;
}

// Time of `reps` predictor + corrector pairs of stage shape t (ms per pair)
// on the current state at the device tau, written into scratch memory (the
// FP64 path's flux buffers, a scratch status word and signal-speed slots):
// the context's state is left as it was.
static int time_tile(vh_ctx_t *c, const vh::mx::Tile &t, double dt, int reps, float &ms) {
  hipStream_t st = c->st;
  double *so = c->F[1], *sc = c->F[2];               // unused by the mixed path
  int *ss = reinterpret_cast<int *>(c->fl[0]);        // scratch status word
  const int nc = c->charged ? 7 : 4;
  unsigned long long *tune_sig = c->sig_acc + 6 * vh::kSigGroup;  // scratch slot groups
  auto go = [&] {
    vh::mx::launch_stage_tile(0, c->c2p_mode, nc, c->G, c->eos, c->Q[c->cur], c->Qh, c->gv,
                              c->m, c->m_full, c->tp, dt, so, ss, tune_sig, t, st);
    vh::mx::launch_stage_tile(1, c->c2p_mode, nc, c->G, c->eos, c->Q[c->cur], so, c->gv,
                              c->m, c->m_full, c->tp, dt, sc, ss, tune_sig, t, st);
  };
  hipEvent_t e0, e1;
  VH_HIP(hipEventCreate(&e0));
  VH_HIP(hipEventCreate(&e1));
  go();
  go();
  int rc = VH_OK;
  if (hipEventRecord(e0, st) != hipSuccess) rc = VH_ERR_HIP;
  for (int r = 0; r < reps; r++) go();
  if (hipEventRecord(e1, st) != hipSuccess) rc = VH_ERR_HIP;
  ms = 0.f;
  if (rc == VH_OK &&
      (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = VH_ERR_HIP;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != VH_OK) return rc;
  VH_HIP(hipGetLastError());
  ms /= (float)reps;
  return VH_OK;
}

// Autotuner of the stage kernel: times a predictor and the corrector from
// its state for every candidate shape (vh::mx::tile_candidates) on the
// current state, and keeps the fastest.  Every shape gives the same result.
static int ensure_tile(vh_ctx_t *c, double tau, double dt) {
  if (c->precision != VH_PREC_MIXED || c->tile.kind || c->tile_forced) return VH_OK;
  const std::vector<vh::mx::Tile> cand = vh::mx::tile_candidates(c->G);
  VH_HIP(hipMemcpyAsync(c->tp, &tau, sizeof(double), hipMemcpyHostToDevice, c->st));
  if (c->c2p_mode == VH_C2P_LEARNED)
    vh::ml::launch_guess(c->Q[c->cur], c->g.N, c->charged ? 7 : 4, c->tp, 0., c->mlp, c->gv,
                         c->st);
  float best = 3.0e38f, best_alt = 3.0e38f;
  vh::mx::Tile bt = cand[0], alt = cand[0];
  for (const vh::mx::Tile &t : cand) {
    float ms = 0.f;
    const int rc = time_tile(c, t, dt, 6, ms);
    if (rc != VH_OK) return rc;
    // the first tuning sees the initial state; as the flow develops, the row
    // kernel's waves (64 cells along x) diverge more in the con2prim
    // iterations than the cell kernel's, so a row shape is kept only when it
    // is clearly faster here (the second tuning, on the developed flow,
    // decides without this margin)
    const bool rows = t.kind != vh::mx::STAGE_CELLS, brows = bt.kind != vh::mx::STAGE_CELLS;
    const float m = rows == brows ? 1.f : (rows ? 0.95f : 1.05f);
    if (ms < best * m) {
      best = ms;
      bt = t;
    }
  }
  // the best shape of the other kind, for the second tuning
  for (const vh::mx::Tile &t : cand) {
    if (t.kind == bt.kind) continue;
    float ms = 0.f;
    const int rc = time_tile(c, t, dt, 2, ms);
    if (rc != VH_OK) return rc;
    if (ms < best_alt) {
      best_alt = ms;
      alt = t;
    }
  }
  c->tile = bt;
  c->tile_alt = alt.kind != bt.kind ? alt : bt;
  c->tile_ms = best;
  c->tune_stage = 1;
  c->steps_since_tune = 0;
  drop_graphs(c);
  return VH_OK;
}

// Steps after the first tuning at which the stage shape is tuned again, on
// the developed flow: the current shape against the best of the other kind.
constexpr long kRetuneAfter = 32;

static int retune(vh_ctx_t *c, double dt) {
  c->tune_stage = 2;
  if (c->tile_forced || c->tile_alt.kind == c->tile.kind || c->tile_alt.kind == 0) return VH_OK;
  VH_HIP(hipStreamSynchronize(c->st));
  // the learned con2prim's guesses of the current state (each step forms its
  // own before its stages, so the replayed steps are unaffected)
  if (c->c2p_mode == VH_C2P_LEARNED)
    vh::ml::launch_guess(c->Q[c->cur], c->g.N, c->charged ? 7 : 4, c->tp, 0., c->mlp, c->gv,
                         c->st);
  float a = 0.f, b = 0.f;
  int rc = time_tile(c, c->tile, dt, 4, a);
  if (rc == VH_OK) rc = time_tile(c, c->tile_alt, dt, 4, b);
  if (rc != VH_OK) return rc;
  if (b < a) {
    std::swap(c->tile, c->tile_alt);
    c->tile_ms = b;
    drop_graphs(c);
  } else {
    c->tile_ms = a;
  }
  return VH_OK;
}

static int run_impl(vh_ctx_t *c, double *tau, double dt, int nsteps, int *status, bool graph) {
  if (!c || !tau || nsteps < 0) return VH_ERR_ARG;
  {
    int rc = ensure_tile(c, *tau, dt);
    if (rc != VH_OK) return rc;
  }
  const bool use_graph = graph && !c->timing && nsteps > 1;
  if (use_graph) {  // capture before any work of this call is enqueued
    int rc = build_graphs(c, dt);
    if (rc != VH_OK) return rc;
  }
  vh::launch_fill_int(c->status, 0, c->st);
  VH_HIP(hipMemcpyAsync(c->tp, tau, sizeof(double), hipMemcpyHostToDevice, c->st));
  int done = 0;
  while (done < nsteps) {
    // the steps up to the second tuning of the stage shape, if it is due
    int chunk = nsteps - done;
    const bool due = c->tune_stage == 1;
    if (due) chunk = (int)std::min<long>(chunk, std::max<long>(1, kRetuneAfter - c->steps_since_tune));
    if (use_graph) {
      int left = chunk;
      for (; left >= kGraphSteps; left -= kGraphSteps)  // kGraphSteps is even: same parity
        VH_HIP(hipGraphLaunch(c->gmulti[c->cur], c->st));
      for (; left > 0; left--) {
        VH_HIP(hipGraphLaunch(c->gexec[c->cur], c->st));
        c->cur ^= 1;
      }
    } else {
      for (int s = 0; s < chunk; s++) {
        int rc = enqueue_step(c, dt);
        if (rc != VH_OK) return rc;
      }
    }
    done += chunk;
    c->steps_since_tune += chunk;
    if (due && c->steps_since_tune >= kRetuneAfter) {
      int rc = retune(c, dt);
      if (rc == VH_OK && use_graph) rc = build_graphs(c, dt);
      if (rc != VH_OK) return rc;
    }
  }
  if (nsteps > 0 && c->vort && c->viscous) c->vort_dirty = true;
  int st = 0;
  VH_HIP(hipMemcpyAsync(tau, c->tp, sizeof(double), hipMemcpyDeviceToHost, c->st));
  VH_HIP(hipMemcpyAsync(&st, c->status, sizeof(int), hipMemcpyDeviceToHost, c->st));
  VH_HIP(hipStreamSynchronize(c->st));
  if (c->timing) {
    int rc = collect_times(c);
    if (rc != VH_OK) return rc;
  }
  if (status) *status = st;
  if (!c->m_full) return refresh_m_full(c);
  return VH_OK;
}

int vh_step(vh_ctx_t *c, double *tau, double dt, int *status) {
  return run_impl(c, tau, dt, 1, status, false);
}

int vh_run(vh_ctx_t *c, double *tau, double dt, int nsteps, int *status) {
  return run_impl(c, tau, dt, nsteps, status, true);
}

int vh_signal_speed(vh_ctx_t *c, int reset, double *out) {
  if (!c || !out) return VH_ERR_ARG;
  unsigned long long h[9];
  VH_HIP(hipMemcpyAsync(h, c->sig + 6, sizeof h, hipMemcpyDeviceToHost, c->st));
  if (reset) vh::launch_clear_u64(c->sig + 12, 3, c->st);
  VH_HIP(hipStreamSynchronize(c->st));
  for (int k = 0; k < 9; k++) std::memcpy(&out[k], &h[k], sizeof(double));
  return VH_OK;
}

int vh_enable_timing(vh_ctx_t *c, int enable) {
  if (!c) return VH_ERR_ARG;
  c->timing = enable != 0;
  c->times.clear();
  return VH_OK;
}

int vh_kernel_times(const vh_ctx_t *c, int max, char names[][32], int *launches, double *ms) {
  if (!c) return VH_ERR_ARG;
  int i = 0;
  for (auto &kv : c->times) {
    if (i >= max) break;
    std::snprintf(names[i], 32, "%s", kv.first.c_str());
    launches[i] = kv.second.launches;
    ms[i] = kv.second.ms;
    i++;
  }
  return i;
}

int vh_c2p(const vh_eos_t *eos, double *Q, long n, double *out, int *iters) {
  if (!eos || !Q || !out || n < 0) return VH_ERR_ARG;
  if (n == 0) return VH_OK;
  vh::DevEos de;
  double *tab = nullptr;
  int *dS = nullptr;
  VH_HIP(hipMalloc(&dS, sizeof(int)));
  {
    const int rc = vh::make_dev_eos(eos, dS, de, &tab);
    if (rc != VH_OK) {
      (void)hipFree(dS);
      return rc;
    }
  }
  double *dQ = nullptr, *dO = nullptr;
  int *dI = nullptr;
  VH_HIP(hipMalloc(&dQ, n * 7 * sizeof(double)));
  VH_HIP(hipMalloc(&dO, n * 8 * sizeof(double)));
  if (iters) VH_HIP(hipMalloc(&dI, n * sizeof(int)));
  VH_HIP(hipMemcpy(dQ, Q, n * 7 * sizeof(double), hipMemcpyHostToDevice));
  vh::launch_c2p_batch(de, dQ, n, dO, dI, nullptr);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpy(out, dO, n * 8 * sizeof(double), hipMemcpyDeviceToHost));
  VH_HIP(hipMemcpy(Q, dQ, n * 7 * sizeof(double), hipMemcpyDeviceToHost));
  if (iters) VH_HIP(hipMemcpy(iters, dI, n * sizeof(int), hipMemcpyDeviceToHost));
  (void)hipFree(dQ);
  (void)hipFree(dO);
  if (dI) (void)hipFree(dI);
  if (tab) (void)hipFree(tab);
  (void)hipFree(dS);
  return VH_OK;
}

int vh_p2c(const double *prim, long n, double *Q) {
  if (!prim || !Q || n < 0) return VH_ERR_ARG;
  if (n == 0) return VH_OK;
  double *dP = nullptr, *dQ = nullptr;
  VH_HIP(hipMalloc(&dP, n * 8 * sizeof(double)));
  VH_HIP(hipMalloc(&dQ, n * 7 * sizeof(double)));
  VH_HIP(hipMemcpy(dP, prim, n * 8 * sizeof(double), hipMemcpyHostToDevice));
  vh::launch_p2c_batch(dP, n, dQ, nullptr);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpy(Q, dQ, n * 7 * sizeof(double), hipMemcpyDeviceToHost));
  (void)hipFree(dP);
  (void)hipFree(dQ);
  return VH_OK;
}

}  // extern "C"

namespace {

// plain streaming copy (16-byte vectors, four per thread, loads issued
// before stores): the HBM bandwidth reference for the rooflines and the
// scale of the traffic counters
__global__ void __launch_bounds__(256)
k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, long n) {
  const long i0 = (long)blockIdx.x * 1024 + threadIdx.x;
  uint4 v[4];
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (i0 + 256 * k < n) v[k] = src[i0 + 256 * k];
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (i0 + 256 * k < n) dst[i0 + 256 * k] = v[k];
}

}  // namespace

extern "C" {

int vh_fo_snapshot(vh_ctx_t *c, double tau) {
  if (!c || c->evh.size() != 1) return VH_ERR_ARG;
  const long N = c->g.N;
  if (!c->fo_snap) VH_HIP(hipMalloc(&c->fo_snap, 7 * N * sizeof(double)));
  VH_HIP(hipMemcpyAsync(c->fo_snap, c->Q[c->cur], 7 * N * sizeof(double),
                        hipMemcpyDeviceToDevice, c->st));
  VH_HIP(hipStreamSynchronize(c->st));
  c->fo_tau = tau;
  return VH_OK;
}

static_assert(vh::fo::BETA_ROW == VH_BETA_COLS, "beta.dat record width");

namespace {
// entries of the last step's vorticity matrix that depend on earlier cells
int vort_settle(vh_ctx_t *c) {
  if (!c->vort_dirty) return VH_OK;
  vh::vort_resolve(c->g, c->db, c->dop, c->vscr, c->st);
  VH_HIP(hipGetLastError());
  c->vort_dirty = false;
  return VH_OK;
}
}  // namespace

int vh_set_vorticity(vh_ctx_t *c, int on) {
  if (!c || c->evh.size() != 1) return VH_ERR_ARG;
  const long N = c->g.N, M = (long)c->g.nx * c->g.ny;
  if (M > (1L << 28)) return VH_ERR_ARG;  // column indices of the resolver are int
  VH_HIP(hipStreamSynchronize(c->st));
  drop_graphs(c);  // the relaxation kernel's arguments change
  if (c->db) (void)hipFree(c->db);
  if (c->vT) (void)hipFree(c->vT);
  if (c->dop) (void)hipFree(c->dop);
  if (c->vscr) (void)hipFree(c->vscr);
  c->db = c->vT = nullptr;
  c->dop = nullptr;
  c->vscr = nullptr;
  c->vort = c->vort_dirty = false;
  if (!on) return VH_OK;
  VH_HIP(hipMalloc(&c->db, 16 * N * sizeof(double)));
  VH_HIP(hipMalloc(&c->vT, N * sizeof(double)));
  VH_HIP(hipMalloc(&c->dop, N));
  VH_HIP(hipMalloc(&c->vscr, vh::vort_scratch_bytes(c->g)));
  // Cell::enableVorticity: every cell's matrix zero (src/cll.h:167-170)
  VH_HIP(hipMemsetAsync(c->db, 0, 16 * N * sizeof(double), c->st));
  VH_HIP(hipStreamSynchronize(c->st));
  c->vort = true;
  return VH_OK;
}

int vh_prim_fields(vh_ctx_t *c, double tau, double *prim) {
  if (!c || !prim || !(tau > 0.)) return VH_ERR_ARG;
  const long N = c->g.N;
  if (!c->pf) VH_HIP(hipMalloc(&c->pf, 8 * N * sizeof(double)));
  vh::launch_prim_fields(N, c->eos, c->Q[c->cur], tau, c->pf, c->st);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(prim, c->pf, 8 * N * sizeof(double), hipMemcpyDeviceToHost, c->st));
  VH_HIP(hipStreamSynchronize(c->st));
  return VH_OK;
}

namespace {
// the EoS of outputSurface / outputCorona: Fluid's eosH when set
const vh::DevEos &surface_eos(const vh_ctx_t *c) { return c->has_eos_fo ? c->eos_fo : c->eos; }
}  // namespace

int vh_set_surface_eos(vh_ctx_t *c, const vh_eos_t *eos) {
  if (!c) return VH_ERR_ARG;
  if (eos && eos->kind != VH_EOS_IDEAL && eos->kind != VH_EOS_TABLE &&
      eos->kind != VH_EOS_CHIRAL && eos->kind != VH_EOS_HADRON)
    return VH_ERR_ARG;
  // the mixed path's FP32 pi goes through the charge-free kernels only
  if (eos && (eos->kind == VH_EOS_CHIRAL || eos->kind == VH_EOS_HADRON) &&
      c->precision != VH_PREC_FP64)
    return VH_ERR_ARG;
  VH_HIP(hipStreamSynchronize(c->st));
  vh::DevEos de;
  double *buf = nullptr;
  if (eos) {
    const int rc = vh::make_dev_eos(eos, c->status, de, &buf);
    if (rc != VH_OK) return rc;
  }
  if (c->eos_fo_buf) (void)hipFree(c->eos_fo_buf);
  c->eos_fo_buf = buf;
  c->eos_fo = de;
  c->has_eos_fo = eos != nullptr;
  return VH_OK;
}

int vh_fluid_stats(vh_ctx_t *c, double tau, double e_crit, int central, double *out) {
  if (!c || !out || c->evh.size() != 1 || !(tau > 0.)) return VH_ERR_ARG;
  if (!c->fsb) VH_HIP(hipMalloc(&c->fsb, (16 * 256 + 16) * sizeof(double)));
  vh::launch_fluid_stats(c->g, surface_eos(c), c->Q[c->cur], c->pi, c->precision == VH_PREC_MIXED,
                         c->vcc, tau, c->grid.zmin, e_crit, c->viscous ? 1 : 0, central ? 1 : 0,
                         c->fsb, c->fsb + 16 * 256, c->st);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(out, c->fsb + 16 * 256, 16 * sizeof(double), hipMemcpyDeviceToHost,
                        c->st));
  VH_HIP(hipStreamSynchronize(c->st));
  return VH_OK;
}

int vh_cells(vh_ctx_t *c, double tau, const long *cells, long n, double *out) {
  if (!c || n < 0 || (n && (!cells || !out)) || !(tau > 0.)) return VH_ERR_ARG;
  const long N = vh_ncell(c);
  for (long i = 0; i < n; i++)
    if (cells[i] < 0 || cells[i] >= N) return VH_ERR_ARG;
  if (n == 0) return VH_OK;
  if (c->ccap < n) {
    if (c->cidx) (void)hipFree(c->cidx);
    if (c->cbuf) (void)hipFree(c->cbuf);
    c->cidx = nullptr;
    c->cbuf = nullptr;
    c->ccap = 0;
    VH_HIP(hipMalloc(&c->cidx, n * sizeof(long)));
    VH_HIP(hipMalloc(&c->cbuf, n * VH_CELL_COLS * sizeof(double)));
    c->ccap = n;
  }
  VH_HIP(hipMemcpyAsync(c->cidx, cells, n * sizeof(long), hipMemcpyHostToDevice, c->st));
  vh::launch_cells(N, c->eos, c->Q[c->cur], c->pi, c->Pi, c->precision == VH_PREC_MIXED, c->vcc,
                   tau, c->cidx, n, c->cbuf, c->st);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(out, c->cbuf, n * VH_CELL_COLS * sizeof(double), hipMemcpyDeviceToHost,
                        c->st));
  VH_HIP(hipStreamSynchronize(c->st));
  return VH_OK;
}

int vh_cells_noalloc(vh_ctx_t *c, double tau, const long *cells, long n, double *out) {
  if (!c || n < 0 || (n && (!cells || !out)) || !(tau > 0.)) return VH_ERR_ARG;
  if (n > VH_CELLS_NOALLOC_MAX) {
    std::fprintf(stderr,
                 "vhlle_gpu: vh_cells_noalloc: %ld cells requested, at most %d per call "
                 "(use vh_cells for more)\n",
                 n, VH_CELLS_NOALLOC_MAX);
    return VH_ERR_ARG;
  }
  const long N = vh_ncell(c);
  for (long i = 0; i < n; i++)
    if (cells[i] < 0 || cells[i] >= N) return VH_ERR_ARG;
  if (n == 0) return VH_OK;
  double *base = (c->dev >= 0 && c->dev < kMaxDevices) ? scratch_ptr[c->dev] : nullptr;
  if (!base) return VH_ERR_HIP;  // set up at vh_create
  // [0, MAX): the cell list (8-byte indices); then the records.  Every index
  // and record read in this call is written in it first, so nothing needs
  // clearing; calls that overlap on one GPU share the array and race.
  long *idx = reinterpret_cast<long *>(base);
  double *rec = base + VH_CELLS_NOALLOC_MAX;
  VH_HIP(hipMemcpyAsync(idx, cells, n * sizeof(long), hipMemcpyHostToDevice, c->st));
  vh::launch_cells(N, c->eos, c->Q[c->cur], c->pi, c->Pi, c->precision == VH_PREC_MIXED, c->vcc,
                   tau, idx, n, rec, c->st);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(out, rec, n * VH_CELL_COLS * sizeof(double), hipMemcpyDeviceToHost,
                        c->st));
  VH_HIP(hipStreamSynchronize(c->st));
  return VH_OK;
}

int vh_get_dbeta(vh_ctx_t *c, double *dbeta) {
  if (!c || !dbeta || !c->vort) return VH_ERR_ARG;
  int rc = vort_settle(c);
  if (rc != VH_OK) return rc;
  const long n = 16 * c->g.N;
  vh::vort_clamp(c->db, n, c->st);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(dbeta, c->db, n * sizeof(double), hipMemcpyDeviceToHost, c->st));
  VH_HIP(hipStreamSynchronize(c->st));
  return VH_OK;
}

int vh_freezeout(vh_ctx_t *c, double tau, double dt, double e_crit, int extend, int lower,
                 double *rows, long max_rows, long *nrows, vh_fo_stats_t *stats) {
  return vh_freezeout_beta(c, tau, dt, e_crit, extend, lower, rows, nullptr, max_rows, nrows,
                           stats);
}

int vh_freezeout_beta(vh_ctx_t *c, double tau, double dt, double e_crit, int extend, int lower,
                      double *rows, double *beta, long max_rows, long *nrows,
                      vh_fo_stats_t *stats) {
  if (!c || c->evh.size() != 1 || (lower != VH_FO_QPREV && lower != VH_FO_SNAPSHOT) ||
      ((rows || beta) && max_rows < 0) || (beta && !c->vort))
    return VH_ERR_ARG;
  if (lower == VH_FO_SNAPSHOT && !c->fo_snap) return VH_ERR_ARG;
  const double edge = lower == VH_FO_SNAPSHOT ? tau - c->fo_tau : dt;
  if (!(edge > 0.)) return VH_ERR_ARG;
  VH_HIP(hipStreamSynchronize(c->st));
  vh::fo::FoArgs a;
  a.nx = c->g.nx;
  a.ny = c->g.ny;
  a.nz = c->g.nz;
  a.eos = surface_eos(c);
  a.P.tau = tau;
  a.P.dt = edge;
  a.P.dx = c->g.dx;
  a.P.dy = c->g.dy;
  a.P.dz = c->g.dz;
  a.P.xmin = c->grid.xmin;
  a.P.ymin = c->grid.ymin;
  a.P.zmin = c->grid.zmin;
  a.P.ecrit = e_crit;
  const int rec = extend ? vh::fo::ROW_EXT : vh::fo::ROW_BASE;
  const bool wb = beta != nullptr && max_rows > 0;
  a.P.nbeta = wb ? vh::fo::BETA : 0;
  a.P.ncol = rec + a.P.nbeta + vh::fo::AUX;
  a.db = nullptr;
  if (wb) {
    const int rc = vort_settle(c);
    if (rc != VH_OK) return rc;
    a.db = c->db;
  }
  a.Q = c->Q[c->cur];
  a.Qlow = lower == VH_FO_SNAPSHOT ? c->fo_snap : c->Q[c->cur ^ 1];
  a.pi = c->pi;
  a.Pi = c->Pi;
  a.pi_float = c->precision == VH_PREC_MIXED;
  a.stream = c->st;
  a.rows_out = max_rows > 0 ? rows : nullptr;
  a.beta_out = wb ? beta : nullptr;  // beta.dat columns (src/fld.cpp:663-834)
  a.max_rows = max_rows;
  a.rec = rec;
  double sums[vh::fo::NSTAT];
  a.stats = sums;
  long n = 0;
  int err = 0;
  const int rc = vh::fo::run(a, c->fo_work, n, err);
  if (rc != VH_OK) return rc;
  if (nrows) *nrows = n;
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->nelements = (long)sums[0];
    stats->nsusp = (long)sums[1];
    for (int k = 0; k < 4; k++) stats->dsigma[k] = sums[2 + k];
    stats->v_eff = sums[6];
    stats->e_surf = sums[7];
    stats->nb_surf = sums[8];
    stats->err = err;
  }
  return VH_OK;
}

int vh_corona(vh_ctx_t *c, double tau, double e_crit, int extend, double *rows, long max_rows,
              long *nrows, vh_fo_stats_t *stats) {
  if (!c || c->evh.size() != 1 || (rows && max_rows < 0) || !(tau > 0.)) return VH_ERR_ARG;
  VH_HIP(hipStreamSynchronize(c->st));
  vh::fo::FoArgs a;
  a.nx = c->g.nx;
  a.ny = c->g.ny;
  a.nz = c->g.nz;
  a.eos = surface_eos(c);
  a.P.tau = tau;
  a.P.dt = 0.;
  a.P.dx = c->g.dx;
  a.P.dy = c->g.dy;
  a.P.dz = c->g.dz;
  a.P.xmin = c->grid.xmin;
  a.P.ymin = c->grid.ymin;
  a.P.zmin = c->grid.zmin;
  a.P.ecrit = e_crit;
  const int rec = extend ? vh::fo::ROW_EXT : vh::fo::ROW_BASE;
  a.P.nbeta = 0;
  a.P.ncol = rec + vh::fo::AUX;
  a.Q = a.Qlow = c->Q[c->cur];
  a.pi = c->pi;
  a.Pi = c->Pi;
  a.pi_float = c->precision == VH_PREC_MIXED;
  a.db = nullptr;
  a.stream = c->st;
  a.rows_out = max_rows > 0 ? rows : nullptr;
  a.beta_out = nullptr;
  a.max_rows = max_rows;
  a.rec = rec;
  double sums[vh::fo::NSTAT];
  a.stats = sums;
  long n = 0;
  const int rc = vh::fo::run_corona(a, c->fo_work, n);
  if (rc != VH_OK) return rc;
  if (nrows) *nrows = n;
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->nelements = (long)sums[0];
    stats->nsusp = 0;
    for (int k = 0; k < 4; k++) stats->dsigma[k] = sums[2 + k];
    stats->v_eff = sums[6];
    stats->e_surf = sums[7];
    stats->nb_surf = sums[8];
  }
  return VH_OK;
}

int vh_cornelius(const double *h, long n, const double *dx, double E0, double *cen,
                 double *nrm, int *nel) {
  if (!h || n < 0 || !dx || !cen || !nrm || !nel) return VH_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return VH_ERR_NODEVICE;
  if (n == 0) return VH_OK;
  return vh::fo::cornelius_batch(h, n, dx, E0, cen, nrm, nel);
}

int vh_copy_bandwidth(long bytes, int reps, double *gbs) {
  if (bytes < 16 * 256 || reps < 1 || !gbs) return VH_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return VH_ERR_NODEVICE;
  const long n = bytes / 16;
  uint4 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = VH_OK;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess ||
      hipMemset(a, 1, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = VH_ERR_HIP;
  if (rc == VH_OK) {
    const unsigned grid = (unsigned)((n + 1023) / 1024);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
        hipGetLastError() != hipSuccess)
      rc = VH_ERR_HIP;
    else
      *gbs = 2.0 * n * 16.0 * reps / (ms * 1e-3) / 1e9;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

int vh_device_info(char *buf, int len) {
  if (!buf || len <= 0) return VH_ERR_ARG;
  int dev = 0, n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return VH_ERR_NODEVICE;
  VH_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  VH_HIP(hipGetDeviceProperties(&p, dev));
  std::snprintf(buf, len, "%s (%s) CUs=%d clock=%d kHz mem=%.1f GiB memclk=%d kHz bus=%d bit",
                p.name, p.gcnArchName, p.multiProcessorCount, p.clockRate,
                p.totalGlobalMem / 1073741824.0, p.memoryClockRate, p.memoryBusWidth);
  return VH_OK;
}

}  // extern "C"
