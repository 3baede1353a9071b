// Learned con2prim: MLP inference (BF16 MFMA), Newton polish with the
// reference iteration as fallback, and training on the GPU (forward/backward
// in FP32 with the BF16 rounding of the inference path, per-parameter
// gradient reduction, Adam).  Declarations: vh_mlp.h, include/vhlle_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/vhlle_gpu.h"
#include "vh_mlp.h"
#include "vh_eos.h"
#include "vh_mlp_host.h"

namespace vh {
namespace ml {

namespace {

constexpr int ACT = 264;  // per-sample activation record of the training pass
// record rows: 0..5 x (bf16, row 5 = 1) | 6..69 h1 (bf16) | 70..133 h2 |
//              134..197 dL/dz1 | 198..261 dL/dz2 | 262..263 dL/dout
constexpr int R_X = 0, R_H1 = 6, R_H2 = 70, R_DZ1 = 134, R_DZ2 = 198, R_DO = 262;

__device__ __forceinline__ float bf(float x) { return (float)(__bf16)x; }

// grid-stride over whole blocks so that every wave stays converged
#define VH_WAVE_LOOP(base, n)                                                  \
  for (long base = (long)blockIdx.x * blockDim.x; base < (n);                  \
       base += (long)gridDim.x * blockDim.x)

// features and training targets (|v|, e/Q^tau) of n states Q (n x 7, /tau
// already applied) and their reference primitives prim (n x 8)
__global__ void __launch_bounds__(256)
k_feats(const double *__restrict__ Q, const double *__restrict__ prim, long n,
        float *__restrict__ X, float *__restrict__ Y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double q[7];
#pragma unroll
  for (int k = 0; k < 7; k++) q[k] = Q[i * 7 + k];
  float f[5];
  features(q, f);
#pragma unroll
  for (int k = 0; k < 5; k++) X[i * 5 + k] = f[k];
  if (prim) {
    const double *p = prim + i * 8;
    const double v = sqrt(p[5] * p[5] + p[6] * p[6] + p[7] * p[7]);
    Y[i * 2] = (float)v;
    Y[i * 2 + 1] = q[T_] > 0. ? (float)(p[0] / q[T_]) : 0.f;
  }
}

// MFMA inference on n feature vectors
__global__ void __launch_bounds__(256)
k_infer(const Packed *__restrict__ P, const float *__restrict__ X, long n,
        float *__restrict__ out) {
  __shared__ Packed w;
  load_packed(w, P);
  __syncthreads();
  VH_WAVE_LOOP(base, n) {
    const long i = base + threadIdx.x;
    const long j = i < n ? i : n - 1;
    float f[5];
#pragma unroll
    for (int k = 0; k < 5; k++) f[k] = X[j * 5 + k];
    const float2 o = mlp_wave(w, f);
    if (i < n) {
      out[i * 2] = o.x;
      out[i * 2 + 1] = o.y;
    }
  }
}

// con2prim from the MLP guess (start 0) or from v = 0.5 (start 1) on n
// states Q (n x 7, rescaled in place as transformPV rescales its argument)
__global__ void __launch_bounds__(256)
k_c2p_learned(DevEos eos, const Packed *__restrict__ P, double *__restrict__ Q, long n,
              int start, double *__restrict__ out, int *__restrict__ iters) {
  __shared__ Packed w;
  load_packed(w, P);
  __syncthreads();
  VH_WAVE_LOOP(base, n) {
    const long i = base + threadIdx.x;
    const long j = i < n ? i : n - 1;
    double q[7];
#pragma unroll
    for (int k = 0; k < 7; k++) q[k] = Q[j * 7 + k];
    float f[5];
    features(q, f);
    const float2 o = mlp_wave(w, f);
    if (i < n) {
      const double v0 = start == 0 ? (double)o.x : 0.5;
      Prim r;
      const int it = c2p_guess(eos, q, v0, r);
#pragma unroll
      for (int k = 0; k < 7; k++) Q[i * 7 + k] = q[k];
      double *po = out + i * 8;
      po[0] = r.e;
      po[1] = r.p;
      po[2] = r.nb;
      po[3] = r.nq;
      po[4] = r.ns;
      po[5] = r.vx;
      po[6] = r.vy;
      po[7] = r.vz;
      if (iters) iters[i] = it;
    }
  }
}

// per-cell guess of |v| for the hydro step: Qs (SoA 7 x N) / (tau + toff)
__global__ void __launch_bounds__(256) VH_OCC(VH_W_GUESS)
k_guess(const double *__restrict__ Qs, long N, int nc, const double *__restrict__ tp, double toff,
        const Packed *__restrict__ P, float *__restrict__ gv) {
  __shared__ Packed w;
  load_packed(w, P);
  __syncthreads();
  const double it = 1. / (*tp + toff);
  // the next iteration's state is loaded before this one's MLP runs
  const long stride = (long)gridDim.x * blockDim.x;
  long base = (long)blockIdx.x * blockDim.x;
  // nc = 4: no charges (the N^tau components are zero and not read)
  double qn[7] = {0., 0., 0., 0., 0., 0., 0.};
  {
    const long j = base + threadIdx.x < N ? base + threadIdx.x : N - 1;
#pragma unroll
    for (int k = 0; k < 7; k++)
      if (k < nc) qn[k] = Qs[k * N + j];
  }
  for (; base < N; base += stride) {
    const long c = base + threadIdx.x;
    double q[7];
#pragma unroll
    for (int k = 0; k < 7; k++) q[k] = qn[k] * it;
    const long cn = c + stride;
    if (base + stride < N) {
      const long j = cn < N ? cn : N - 1;
#pragma unroll
      for (int k = 0; k < 7; k++)
        if (k < nc) qn[k] = Qs[k * N + j];
    }
    float f[5];
    features(q, f);
    const float2 o = mlp_wave(w, f);
    if (c < N) gv[c] = o.x;
  }
}

// W1, b1, W2 rounded to BF16 (their values in the MFMA path); the rest as is
__global__ void k_round(const float *__restrict__ P, float *__restrict__ Pr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NPARAM) Pr[i] = i < OFF_B2 ? bf(P[i]) : P[i];
}

// forward + backward of one batch, one thread per sample, FP32 arithmetic
// with the BF16 rounding points of mlp_wave (inputs, W1|b1, h1, W2).
// Loss = (1/B) sum_s sum_o (out_o - y_o)^2.
__global__ void __launch_bounds__(256)
k_fwdbwd(const float *__restrict__ Pr, const float *__restrict__ X, const float *__restrict__ Y,
         const int *__restrict__ idx, int B, float *__restrict__ act,
         float *__restrict__ lossp) {
  __shared__ float red[256];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  float lsum = 0.f;
  if (s < B) {
    const long j = idx ? idx[s] : s;
    float *a = act + (long)s * ACT;
    float xb[6];
#pragma unroll
    for (int k = 0; k < 5; k++) xb[k] = bf(X[j * 5 + k]);
    xb[5] = 1.f;
#pragma unroll
    for (int k = 0; k < 6; k++) a[R_X + k] = xb[k];
    float h1[64];
    uint64_t m1 = 0, m2 = 0;
#pragma unroll
    for (int h = 0; h < 64; h++) {
      float z = 0.f;
#pragma unroll
      for (int k = 0; k < 5; k++) z += Pr[OFF_W1 + h * 5 + k] * xb[k];
      z += Pr[OFF_B1 + h];
      h1[h] = bf(fmaxf(z, 0.f));
      if (z > 0.f) m1 |= 1ull << h;
      a[R_H1 + h] = h1[h];
    }
    float o0 = Pr[OFF_B3], o1 = Pr[OFF_B3 + 1];
    for (int h2 = 0; h2 < 64; h2++) {
      float z = Pr[OFF_B2 + h2];
#pragma unroll
      for (int h = 0; h < 64; h++) z += Pr[OFF_W2 + h2 * 64 + h] * h1[h];
      const float hv = fmaxf(z, 0.f);
      if (z > 0.f) m2 |= 1ull << h2;
      a[R_H2 + h2] = hv;
      o0 += Pr[OFF_W3 + h2] * hv;
      o1 += Pr[OFF_W3 + 64 + h2] * hv;
    }
    float d0 = o0 - Y[j * 2], d1 = o1 - Y[j * 2 + 1];
    lsum = d0 * d0 + d1 * d1;
    const float sc = 2.f / (float)B;
    d0 *= sc;
    d1 *= sc;
    a[R_DO] = d0;
    a[R_DO + 1] = d1;
    float dh1[64];
#pragma unroll
    for (int h = 0; h < 64; h++) dh1[h] = 0.f;
    for (int h2 = 0; h2 < 64; h2++) {
      const float dz = ((m2 >> h2) & 1ull) ? Pr[OFF_W3 + h2] * d0 + Pr[OFF_W3 + 64 + h2] * d1 : 0.f;
      a[R_DZ2 + h2] = dz;
#pragma unroll
      for (int h = 0; h < 64; h++) dh1[h] += Pr[OFF_W2 + h2 * 64 + h] * dz;
    }
#pragma unroll
    for (int h = 0; h < 64; h++) a[R_DZ1 + h] = ((m1 >> h) & 1ull) ? dh1[h] : 0.f;
  }
  red[threadIdx.x] = lsum;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

// the two activation rows whose product, summed over samples, is dL/dP[p]
__device__ __forceinline__ void grad_rows(int p, int &ra, int &rb) {
  if (p < OFF_B1) {
    ra = R_DZ1 + p / 5;
    rb = R_X + p % 5;
  } else if (p < OFF_W2) {
    ra = R_DZ1 + (p - OFF_B1);
    rb = R_X + 5;
  } else if (p < OFF_B2) {
    const int q = p - OFF_W2;
    ra = R_DZ2 + q / 64;
    rb = R_H1 + q % 64;
  } else if (p < OFF_W3) {
    ra = R_DZ2 + (p - OFF_B2);
    rb = R_X + 5;
  } else if (p < OFF_B3) {
    const int q = p - OFF_W3;
    ra = R_DO + q / 64;
    rb = R_H2 + q % 64;
  } else {
    ra = R_DO + (p - OFF_B3);
    rb = R_X + 5;
  }
}

// partial gradients over sample chunks: part[chunk][p], fixed summation order
__global__ void __launch_bounds__(256)
k_wgrad(const float *__restrict__ act, int B, int cs, float *__restrict__ part) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  int ra, rb;
  grad_rows(p, ra, rb);
  const int s0 = blockIdx.y * cs, s1 = min(B, s0 + cs);
  float acc = 0.f;
  for (int s = s0; s < s1; s++) acc += act[(long)s * ACT + ra] * act[(long)s * ACT + rb];
  part[(long)blockIdx.y * NPARAM + p] = acc;
}

// g = sum over chunks; optional Adam update (t >= 1) or plain gradient copy
__global__ void __launch_bounds__(256)
k_adam(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
       const float *__restrict__ part, int nchunk, float lr, float b1c, float b2c, int t,
       float *__restrict__ gout) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float g = 0.f;
  for (int c = 0; c < nchunk; c++) g += part[(long)c * NPARAM + p];
  if (gout) gout[p] = g;
  if (t <= 0) return;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const float mm = beta1 * m[p] + (1.f - beta1) * g;
  const float vv = beta2 * v[p] + (1.f - beta2) * g * g;
  m[p] = mm;
  v[p] = vv;
  P[p] -= lr * (mm / b1c) / (sqrtf(vv / b2c) + eps);
}

uint16_t bf16_bits(float x) {  // round to nearest even, as the device conversion
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t &s) { return (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// device buffer that frees itself
template <class T>
struct DBuf {
  T *p = nullptr;
  hipError_t alloc(size_t n) { return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)); }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
};

#define ML_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      std::fprintf(stderr, "vhlle_gpu: %s failed: %s (%s:%d)\n", #call,      \
                   hipGetErrorString(e_), __FILE__, __LINE__);               \
      return VH_ERR_HIP;                                                     \
    }                                                                        \
  } while (0)

int have_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
  return 1;
}

int wave_grid(long n) {
  const long blocks = (n + 255) / 256;
  return (int)std::max(1L, std::min(blocks, 2048L));
}

// features and targets on the device from host Q (n x 7) and prim (n x 8)
int upload_xy(const double *Q, const double *prim, long n, DBuf<float> &X, DBuf<float> &Y) {
  DBuf<double> dQ, dP;
  ML_HIP(dQ.alloc(n * 7));
  ML_HIP(X.alloc(n * 5));
  ML_HIP(hipMemcpy(dQ.p, Q, n * 7 * sizeof(double), hipMemcpyHostToDevice));
  if (prim) {
    ML_HIP(dP.alloc(n * 8));
    ML_HIP(Y.alloc(n * 2));
    ML_HIP(hipMemcpy(dP.p, prim, n * 8 * sizeof(double), hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(k_feats, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dQ.p,
                     prim ? dP.p : nullptr, n, X.p, prim ? Y.p : nullptr);
  ML_HIP(hipGetLastError());
  ML_HIP(hipDeviceSynchronize());
  return VH_OK;
}

// buffers of one training/gradient pass over batches of up to B samples
struct Pass {
  int B = 0, nchunk = 0, cs = 0;
  DBuf<float> P, Pr, m, v, act, part, lossp, g;
  int init(int batch) {
    B = batch;
    cs = 256;
    nchunk = (B + cs - 1) / cs;
    ML_HIP(P.alloc(NPARAM));
    ML_HIP(Pr.alloc(NPARAM));
    ML_HIP(m.alloc(NPARAM));
    ML_HIP(v.alloc(NPARAM));
    ML_HIP(g.alloc(NPARAM));
    ML_HIP(act.alloc((size_t)B * ACT));
    ML_HIP(part.alloc((size_t)nchunk * NPARAM));
    ML_HIP(lossp.alloc((B + 255) / 256));
    ML_HIP(hipMemset(m.p, 0, NPARAM * sizeof(float)));
    ML_HIP(hipMemset(v.p, 0, NPARAM * sizeof(float)));
    return VH_OK;
  }
  // one forward/backward over nb samples idx[0..nb) (idx may be null), then
  // either Adam (t >= 1) or the gradient into g
  void step(const float *X, const float *Y, const int *idx, int nb, float lr, int t) {
    const dim3 b(256);
    hipLaunchKernelGGL(k_round, dim3((NPARAM + 255) / 256), b, 0, 0, P.p, Pr.p);
    hipLaunchKernelGGL(k_fwdbwd, dim3((nb + 255) / 256), b, 0, 0, Pr.p, X, Y, idx, nb, act.p,
                       lossp.p);
    const int nc = (nb + cs - 1) / cs;
    hipLaunchKernelGGL(k_wgrad, dim3((NPARAM + 255) / 256, nc), b, 0, 0, act.p, nb, cs, part.p);
    const float b1c = t > 0 ? 1.f - std::pow(0.9f, (float)t) : 1.f;
    const float b2c = t > 0 ? 1.f - std::pow(0.999f, (float)t) : 1.f;
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), b, 0, 0, P.p, m.p, v.p, part.p, nc,
                       lr, b1c, b2c, t, t > 0 ? nullptr : g.p);
  }
};

}  // namespace

void pack(const float *P, Packed &out) {
  std::memset(&out, 0, sizeof out);
  auto put = [](bf16x8 &frag, int t, float x) {
    uint16_t *u = reinterpret_cast<uint16_t *>(&frag);
    u[t] = bf16_bits(x);
  };
  for (int mt = 0; mt < 2; mt++)
    for (int l = 0; l < 64; l++)
      for (int t = 0; t < 8; t++) {
        const int h = 32 * mt + l % 32, kk = 8 * (l / 32) + t;
        float x = 0.f;
        if (kk < 5) x = P[OFF_W1 + h * 5 + kk];
        else if (kk == 5) x = P[OFF_B1 + h];
        put(out.w1[mt][l], t, x);
      }
  for (int mt = 0; mt < 2; mt++)
    for (int kb = 0; kb < 4; kb++)
      for (int l = 0; l < 64; l++)
        for (int t = 0; t < 8; t++) {
          const int h2 = 32 * mt + l % 32, half = l / 32;
          const int h1 = 16 * kb + 8 * (t / 4) + 4 * half + (t % 4);
          put(out.w2[mt][kb][l], t, P[OFF_W2 + h2 * 64 + h1]);
        }
  for (int h = 0; h < 64; h++) {
    out.b2[h] = P[OFF_B2 + h];
    out.w3[0][h] = P[OFF_W3 + h];
    out.w3[1][h] = P[OFF_W3 + 64 + h];
  }
  out.b3[0] = P[OFF_B3];
  out.b3[1] = P[OFF_B3 + 1];
}

void launch_guess(const double *Qs, long N, int nc, const double *tp, double toff, const Packed *P,
                  float *gv, hipStream_t st) {
  hipLaunchKernelGGL(k_guess, dim3(wave_grid(N)), dim3(256), 0, st, Qs, N, nc, tp, toff, P, gv);
}

}  // namespace ml
}  // namespace vh

using namespace vh::ml;

extern "C" {

int vh_mlp_init(unsigned long long seed, float *params) {
  if (!params) return VH_ERR_ARG;
  uint64_t s = seed;
  auto fill = [&](int off, int n, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < n; i++) params[off + i] = (float)((2. * uniform01(s) - 1.) * a);
  };
  std::fill(params, params + NPARAM, 0.f);
  fill(OFF_W1, 320, NIN);
  fill(OFF_W2, 4096, NH);
  fill(OFF_W3, 128, NH);
  return VH_OK;
}

int vh_mlp_forward(const float *params, const double *Q, long n, float *out) {
  if (!params || !Q || !out || n < 0) return VH_ERR_ARG;
  if (!have_device()) return VH_ERR_NODEVICE;
  if (n == 0) return VH_OK;
  DBuf<float> X, Y, O;
  DBuf<Packed> P;
  int rc = upload_xy(Q, nullptr, n, X, Y);
  if (rc) return rc;
  Packed hp;
  pack(params, hp);
  ML_HIP(P.alloc(1));
  ML_HIP(O.alloc(n * 2));
  ML_HIP(hipMemcpy(P.p, &hp, sizeof hp, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_infer, dim3(wave_grid(n)), dim3(256), 0, 0, P.p, X.p, n, O.p);
  ML_HIP(hipGetLastError());
  ML_HIP(hipMemcpy(out, O.p, n * 2 * sizeof(float), hipMemcpyDeviceToHost));
  return VH_OK;
}

int vh_mlp_grad(const float *params, const double *Q, const double *prim, long n, float *grad,
                double *loss) {
  if (!params || !Q || !prim || !grad || n <= 0 || n > (1 << 20)) return VH_ERR_ARG;
  if (!have_device()) return VH_ERR_NODEVICE;
  DBuf<float> X, Y;
  int rc = upload_xy(Q, prim, n, X, Y);
  if (rc) return rc;
  Pass ps;
  if ((rc = ps.init((int)n))) return rc;
  ML_HIP(hipMemcpy(ps.P.p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice));
  ps.step(X.p, Y.p, nullptr, (int)n, 0.f, 0);
  ML_HIP(hipGetLastError());
  ML_HIP(hipMemcpy(grad, ps.g.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) {
    std::vector<float> lp((n + 255) / 256);
    ML_HIP(hipMemcpy(lp.data(), ps.lossp.p, lp.size() * sizeof(float), hipMemcpyDeviceToHost));
    double s = 0.;
    for (float x : lp) s += x;
    *loss = s / (double)n;
  }
  return VH_OK;
}

int vh_mlp_train(const vh_mlp_train_cfg_t *cfg, const double *Q, const double *prim, long n,
                 float *params, double *loss_hist) {
  if (!cfg || !Q || !prim || !params || n <= 0 || cfg->batch <= 0 || cfg->batch > (1 << 20) ||
      cfg->epochs <= 0 || n > 0x7fffffffL)
    return VH_ERR_ARG;
  if (!have_device()) return VH_ERR_NODEVICE;
  DBuf<float> X, Y;
  int rc = upload_xy(Q, prim, n, X, Y);
  if (rc) return rc;
  const int B = (int)std::min<long>(cfg->batch, n);
  Pass ps;
  if ((rc = ps.init(B))) return rc;
  ML_HIP(hipMemcpy(ps.P.p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice));
  DBuf<int> didx;
  ML_HIP(didx.alloc(n));
  std::vector<int> idx(n);
  for (long i = 0; i < n; i++) idx[i] = (int)i;
  uint64_t s = cfg->seed ^ 0x5eedc2b0ull;
  const long nbatch = n / B;
  const long total = nbatch * cfg->epochs;
  std::vector<float> lp((B + 255) / 256);
  long t = 0;
  for (int ep = 0; ep < cfg->epochs; ep++) {
    for (long i = n - 1; i > 0; i--) std::swap(idx[i], idx[(long)(splitmix64(s) % (uint64_t)(i + 1))]);
    ML_HIP(hipMemcpy(didx.p, idx.data(), n * sizeof(int), hipMemcpyHostToDevice));
    double lsum = 0.;
    int nl = 0;
    for (long bi = 0; bi < nbatch; bi++) {
      const double fr = (double)t / (double)std::max(1L, total - 1);
      const double lr = cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1. + std::cos(M_PI * fr));
      ps.step(X.p, Y.p, didx.p + bi * B, B, (float)lr, (int)(++t));
      if (loss_hist && (bi == nbatch - 1 || bi % 64 == 0)) {
        ML_HIP(hipMemcpy(lp.data(), ps.lossp.p, lp.size() * sizeof(float), hipMemcpyDeviceToHost));
        double sb = 0.;
        for (float x : lp) sb += x;
        lsum += sb / B;
        nl++;
      }
    }
    ML_HIP(hipGetLastError());
    if (loss_hist) loss_hist[ep] = nl ? lsum / nl : 0.;
  }
  ML_HIP(hipMemcpy(params, ps.P.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return VH_OK;
}

int vh_c2p_learned(const vh_eos_t *eos, const float *params, double *Q, long n, int start,
                   double *out, int *iters) {
  if (!eos || !params || !Q || !out || n < 0 || (start != VH_START_MLP && start != VH_START_HALF))
    return VH_ERR_ARG;
  // the polish evaluates p(e) alone: EoSs with charge dependence are refused
  if (eos->kind != VH_EOS_IDEAL && eos->kind != VH_EOS_TABLE) return VH_ERR_ARG;
  if (!have_device()) return VH_ERR_NODEVICE;
  if (n == 0) return VH_OK;
  vh::DevEos de;
  double *tab = nullptr;
  DBuf<int> dS;
  ML_HIP(dS.alloc(1));
  {
    const int rc = vh::make_dev_eos(eos, dS.p, de, &tab);
    if (rc != VH_OK) return rc;
  }
  struct FreeTab {
    double *t;
    ~FreeTab() {
      if (t) (void)hipFree(t);
    }
  } free_tab{tab};
  DBuf<double> dQ, dO;
  DBuf<int> dI;
  DBuf<Packed> P;
  Packed hp;
  pack(params, hp);
  ML_HIP(P.alloc(1));
  ML_HIP(dQ.alloc(n * 7));
  ML_HIP(dO.alloc(n * 8));
  if (iters) ML_HIP(dI.alloc(n));
  ML_HIP(hipMemcpy(P.p, &hp, sizeof hp, hipMemcpyHostToDevice));
  ML_HIP(hipMemcpy(dQ.p, Q, n * 7 * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_c2p_learned, dim3(wave_grid(n)), dim3(256), 0, 0, de, P.p, dQ.p, n, start,
                     dO.p, iters ? dI.p : nullptr);
  ML_HIP(hipGetLastError());
  ML_HIP(hipMemcpy(out, dO.p, n * 8 * sizeof(double), hipMemcpyDeviceToHost));
  ML_HIP(hipMemcpy(Q, dQ.p, n * 7 * sizeof(double), hipMemcpyDeviceToHost));
  if (iters) ML_HIP(hipMemcpy(iters, dI.p, n * sizeof(int), hipMemcpyDeviceToHost));
  return VH_OK;
}

}  // extern "C"
