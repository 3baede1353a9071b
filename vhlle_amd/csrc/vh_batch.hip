// Hot-spot initial states for event-by-event batches: per event, N_part ~
// Poisson(mean) spots at 2-D Woods-Saxon positions with Gamma(k = 1, theta = 1)
// weights (sampled on the host from one SplitMix64 stream per event), summed
// as Gaussians of width sigma into e(x, y) on the GPU, times an eta profile
// for 3-D grids (plateau |eta| < eta_plateau, Gaussian tails of width
// eta_sigma).  The reference's nearest counterpart is the Glissando IC
// (src/icGlissando.cpp:209-246), which reads the spots from a file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/vhlle_gpu.h"

namespace {

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double uni(uint64_t &s) { return ((splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

// Poisson variate by multiplication of uniforms (Knuth); fine for means <~ 500
int poisson(uint64_t &s, double mean) {
  const double L = std::exp(-mean);
  int k = 0;
  double p = 1.0;
  do {
    k++;
    p *= uni(s);
  } while (p > L);
  return k - 1;
}

struct Spot {
  double x, y, w;
};

// grid of one event in the batch layout
struct EvGrid {
  int nx, ny, nz;
  double xmin, dx, ymin, dy, zmin, dz;
  long base;
  int spot0, nspot;
};

__global__ void __launch_bounds__(256)
k_deposit(const EvGrid *__restrict__ ev, const int *__restrict__ bev,
          const int *__restrict__ bfirst, const Spot *__restrict__ spots, double amp,
          double sigma, double eta_plateau, double eta_sigma, double *__restrict__ e) {
  const int gi = bev[blockIdx.x];
  const EvGrid g = ev[gi];
  const long cl = (long)(blockIdx.x - bfirst[gi]) * blockDim.x + threadIdx.x;
  const long n = (long)g.nx * g.ny * g.nz;
  if (cl >= n) return;
  const int ix = (int)(cl % g.nx);
  const long r = cl / g.nx;
  const int iy = (int)(r % g.ny), iz = (int)(r / g.ny);
  const double x = g.xmin + ix * g.dx, y = g.ymin + iy * g.dy;
  const double i2s2 = 1.0 / (2.0 * sigma * sigma);
  double sum = 0.;
  for (int k = 0; k < g.nspot; k++) {
    const Spot sp = spots[g.spot0 + k];
    const double dx = x - sp.x, dy = y - sp.y;
    sum += sp.w * exp(-(dx * dx + dy * dy) * i2s2);
  }
  double H = 1.0;
  if (g.nz > 1) {
    const double a = fabs(g.zmin + iz * g.dz);
    if (a > eta_plateau) {
      const double t = a - eta_plateau;
      H = exp(-t * t / (2.0 * eta_sigma * eta_sigma));
    }
  }
  e[g.base + cl] = amp * sum * H;
}

}  // namespace

extern "C" int vh_hotspot_ic(const vh_hotspot_cfg_t *cfg, int ngrid, const vh_grid_t *grids,
                             double *e, int *npart, double *spots_out, long max_spots) {
  if (!cfg || ngrid < 1 || !grids || !e || !(cfg->sigma > 0.) || !(cfg->npart_mean > 0.) ||
      cfg->npart_mean > 500. || !(cfg->R > 0.) || !(cfg->a > 0.))
    return VH_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return VH_ERR_NODEVICE;
  std::vector<EvGrid> ev(ngrid);
  std::vector<Spot> spots;
  std::vector<int> bev, bfirst(ngrid);
  long base = 0;
  const double rmax = cfg->R + 12.0 * cfg->a;
  for (int g = 0; g < ngrid; g++) {
    const vh_grid_t &gr = grids[g];
    if (gr.nx < 1 || gr.ny < 1 || gr.nz < 1) return VH_ERR_ARG;
    EvGrid &q = ev[g];
    q.nx = gr.nx;
    q.ny = gr.ny;
    q.nz = gr.nz;
    q.xmin = gr.xmin;
    q.ymin = gr.ymin;
    q.zmin = gr.zmin;
    q.dx = gr.nx > 1 ? (gr.xmax - gr.xmin) / (gr.nx - 1) : 0.;
    q.dy = gr.ny > 1 ? (gr.ymax - gr.ymin) / (gr.ny - 1) : 0.;
    q.dz = gr.nz > 1 ? (gr.zmax - gr.zmin) / (gr.nz - 1) : 0.;
    q.base = base;
    q.spot0 = (int)spots.size();
    uint64_t s = cfg->seed * 0x100000001b3ull + (uint64_t)g;
    splitmix64(s);
    const int np = poisson(s, cfg->npart_mean);
    for (int k = 0; k < np; k++) {
      double r;
      do {  // r dr dphi / (1 + exp((r - R)/a)) by rejection from the uniform disc
        r = rmax * std::sqrt(uni(s));
      } while (uni(s) * (1.0 + std::exp((r - cfg->R) / cfg->a)) > 1.0);
      const double ph = 2.0 * M_PI * uni(s);
      const double w = -std::log(uni(s));  // Gamma(k = 1, theta = 1)
      spots.push_back(Spot{r * std::cos(ph), r * std::sin(ph), w});
    }
    q.nspot = np;
    if (npart) npart[g] = np;
    const long n = (long)gr.nx * gr.ny * gr.nz;
    bfirst[g] = (int)bev.size();
    for (long b = 0; b < (n + 255) / 256; b++) bev.push_back(g);
    base += n;
  }
  if (spots_out)
    for (long k = 0; k < (long)spots.size() && k < max_spots; k++) {
      spots_out[3 * k] = spots[k].x;
      spots_out[3 * k + 1] = spots[k].y;
      spots_out[3 * k + 2] = spots[k].w;
    }
  if (spots.empty()) spots.push_back(Spot{0., 0., 0.});
  EvGrid *dev = nullptr;
  int *dbev = nullptr, *dbf = nullptr;
  Spot *dsp = nullptr;
  double *de = nullptr;
  int rc = VH_OK;
  if (hipMalloc(&dev, ev.size() * sizeof(EvGrid)) != hipSuccess ||
      hipMalloc(&dbev, bev.size() * sizeof(int)) != hipSuccess ||
      hipMalloc(&dbf, bfirst.size() * sizeof(int)) != hipSuccess ||
      hipMalloc(&dsp, spots.size() * sizeof(Spot)) != hipSuccess ||
      hipMalloc(&de, base * sizeof(double)) != hipSuccess)
    rc = VH_ERR_HIP;
  if (rc == VH_OK &&
      (hipMemcpy(dev, ev.data(), ev.size() * sizeof(EvGrid), hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(dbev, bev.data(), bev.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(dbf, bfirst.data(), bfirst.size() * sizeof(int), hipMemcpyHostToDevice) !=
           hipSuccess ||
       hipMemcpy(dsp, spots.data(), spots.size() * sizeof(Spot), hipMemcpyHostToDevice) !=
           hipSuccess))
    rc = VH_ERR_HIP;
  if (rc == VH_OK) {
    hipLaunchKernelGGL(k_deposit, dim3((unsigned)bev.size()), dim3(256), 0, 0, dev, dbev, dbf, dsp,
                       cfg->amp, cfg->sigma, cfg->eta_plateau, cfg->eta_sigma, de);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(e, de, base * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = VH_ERR_HIP;
  }
  if (dev) (void)hipFree(dev);
  if (dbev) (void)hipFree(dbev);
  if (dbf) (void)hipFree(dbf);
  if (dsp) (void)hipFree(dsp);
  if (de) (void)hipFree(de);
  return rc;
}
