// Training pairs for the learned con2prim guess: states harvested from ideal
// Gubser runs of the FP64 step (the 101 x 101 x 1 grid of the Gubser
// benchmark), plus transformCV images of random primitive states up to
// |v| = vmax, each labelled by the reference con2prim (transformPV,
// src/rmn.cpp:28-139, through vh_c2p).  Only the C-ABI is used.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/vhlle_gpu.h"
#include "vhlle.h"

namespace {

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double uni(uint64_t &s) { return (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// Gubser IC of ICGubser::setIC (src/icGubser.cpp:32-43): q = 1, t = 1
void gubser_ic(double amp, double x, double y, double &e, double &vx, double &vy) {
  const double t = 1.0, q = 1.0, r = std::sqrt(x * x + y * y);
  const double k = 2. * q * q * t * r / (1. + q * q * t * t + q * q * r * r);
  const double qq = q * q * (t * t - r * r);
  double ee = 4. * q * q / (1. + 2. * q * q * (t * t + r * r) + qq * qq) / t;
  if (ee < 0.) ee = 0.;
  e = amp * std::pow(ee, 4. / 3.);
  vx = x / (r + 1e-50) * k;
  vy = y / (r + 1e-50) * k;
}

}  // namespace

extern "C" int vh_mlp_harvest(const vh_harvest_cfg_t *cfg, long n, double *Q, double *prim) {
  if (!cfg || !Q || !prim || n <= 0 || cfg->steps <= 0 || cfg->namp < 1 || cfg->namp > 4 ||
      !(cfg->augment >= 0. && cfg->augment <= 1.) || !(cfg->vmax > 0. && cfg->vmax < 1.))
    return VH_ERR_ARG;
  const int nx = 101, ny = 101;
  const long N = (long)nx * ny;
  const double tau0 = 1.0, dt = 0.025;
  const long n_aug = (long)std::llround(cfg->augment * (double)n), n_h = n - n_aug;
  uint64_t s = cfg->seed;
  // harvested part: draw (run, step, cell) uniformly, bucket by (run, step)
  const long slots = (long)cfg->namp * cfg->steps;
  std::vector<long> pick(n_h), order(n_h);
  for (long i = 0; i < n_h; i++) pick[i] = (long)(splitmix64(s) % (uint64_t)(slots * N));
  for (long i = 0; i < n_h; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](long a, long b) { return pick[a] < pick[b]; });
  long cursor = 0;
  vh_grid_t grid{nx, ny, 1, -5.0, 5.0, -5.0, 5.0, -1.0, 1.0};
  vh_eos_t ideal{};
  ideal.kind = VH_EOS_IDEAL;
  const vh_eos_t &eos = cfg->eos ? *cfg->eos : ideal;
  if (eos.kind != VH_EOS_IDEAL && eos.kind != VH_EOS_TABLE) return VH_ERR_ARG;  // p(e) EoSs only
  const bool table = eos.kind == VH_EOS_TABLE;
  vh_trcoeff_t trc;
  std::memset(&trc, 0, sizeof trc);
  trc.T0 = 0.15;
  std::vector<double> e(N), vx(N), vy(N), vz(N, 0.), q(7 * N);
  const double dx = 10.0 / (nx - 1), dy = 10.0 / (ny - 1);
  for (int run = 0; run < cfg->namp && n_h > 0; run++) {
    vh_ctx_t *ctx = nullptr;
    int rc = vh_create(&grid, &eos, &trc, VH_PREC_FP64, &ctx);
    if (rc) return rc;
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++) {
        const long c = ix + (long)nx * iy;
        gubser_ic(cfg->amp[run], -5.0 + ix * dx, -5.0 + iy * dy, e[c], vx[c], vy[c]);
      }
    rc = vh_set_prim(ctx, tau0, e.data(), nullptr, nullptr, nullptr, vx.data(), vy.data(),
                     vz.data());
    double tau = tau0;
    for (int st = 0; st < cfg->steps && rc == 0; st++) {
      int status = 0;
      rc = vh_step(ctx, &tau, dt, &status);
      if (rc) break;
      const long slot = (long)run * cfg->steps + st;
      if (cursor < n_h && pick[order[cursor]] / N == slot) {
        vh_state_t hs;
        std::memset(&hs, 0, sizeof hs);
        hs.Q = q.data();
        rc = vh_download(ctx, &hs);
        while (rc == 0 && cursor < n_h && pick[order[cursor]] / N == slot) {
          const long i = order[cursor], c = pick[i] % N;
          for (int k = 0; k < 7; k++) Q[i * 7 + k] = q[k * N + c] / tau;
          cursor++;
        }
      }
    }
    vh_destroy(ctx);
    if (rc) return rc;
  }
  // augmented part: random primitives, e log-uniform in [1e-6, 1e3] (inside
  // a table's range), p = p(e) of the EoS
  if (n_aug > 0) {
    double elo = 1e-6, ehi = 1e3;
    if (table) {
      elo = std::max(elo, eos.e[0] > 0. ? eos.e[0] : eos.e[1]);
      ehi = std::min(ehi, eos.e[eos.n - 1]);
      if (!(ehi > elo)) return VH_ERR_ARG;
    }
    std::vector<double> pr(n_aug * 8);
    for (long i = 0; i < n_aug; i++) {
      double *p = &pr[i * 8];
      const double ee = std::exp(std::log(elo) + uni(s) * (std::log(ehi) - std::log(elo)));
      const double v = cfg->vmax * uni(s);
      const double cz = 2. * uni(s) - 1., ph = 2. * M_PI * uni(s), sz = std::sqrt(1. - cz * cz);
      p[0] = ee;
      p[1] = ee / 3.;
      p[2] = uni(s) < 0.5 ? 0. : 0.1 * ee * uni(s);
      p[3] = p[4] = 0.;
      p[5] = v * sz * std::cos(ph);
      p[6] = v * sz * std::sin(ph);
      p[7] = v * cz;
    }
    if (table) {
      EoSs tab(eos);
      for (long i = 0; i < n_aug; i++) pr[i * 8 + 1] = tab.p(pr[i * 8]);
    }
    const int rc = vh_p2c(pr.data(), n_aug, Q + n_h * 7);
    if (rc) return rc;
  }
  // labels: the reference con2prim of every state
  return vh_c2p(&eos, Q, n, prim, nullptr);
}

namespace vhlle_learn {

static const char kMagic[8] = {'V', 'H', 'M', 'L', 'P', '1', 0, 0};

bool saveC2PModel(const char *path, const float *params) {
  FILE *f = std::fopen(path, "wb");
  if (!f) return false;
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 &&
            std::fwrite(params, sizeof(float), VH_MLP_NPARAM, f) == VH_MLP_NPARAM;
  return std::fclose(f) == 0 && ok;
}

bool loadC2PModel(const char *path, float *params) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  char m[8];
  bool ok = std::fread(m, 1, 8, f) == 8 && std::memcmp(m, kMagic, 8) == 0 &&
            std::fread(params, sizeof(float), VH_MLP_NPARAM, f) == VH_MLP_NPARAM;
  std::fclose(f);
  return ok;
}

}  // namespace vhlle_learn
