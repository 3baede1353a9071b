// Learned con2prim initial guess: an MLP 5 -> 64 -> 64 -> 2 (ReLU) that maps
// features of a conserved state to guesses of |v| and e/Q^tau, evaluated with
// BF16 MFMA (v_mfma_f32_32x32x16_bf16, FP32 accumulation), followed by an FP64
// Newton polish of the con2prim equation of transformPV (src/rmn.cpp:28-139)
// to a residual |f| <= 1e-12 Q^tau, with the reference iteration as fallback.
//
// Parameter vector (FP32 master copy, VH_MLP_NPARAM = 4674 floats):
//   W1[64][5] | b1[64] | W2[64][64] | b2[64] | W3[2][64] | b3[2]
// Features of q = Q/tau (all FP32): ln(Q^tau)/10, |Q|/Q^tau, Q^eta/Q^tau,
// N_b/Q^tau, sqrt(1 - (|Q|/Q^tau)^2).  Outputs: |v|, e/Q^tau.
//
// MFMA data flow (one sample per lane, the wave's 64 samples as two batches
// of 32).  Every layer is computed transposed, H^T = W X^T, so that a
// sample is an MFMA column (lane % 32) in both the B operand and the
// accumulator.  The accumulator of one 32x32 tile holds, in lane l and
// register r, hidden row 8(r/4) + 4(l/32) + (r%4); used as the next layer's
// B operand (K index 8(l/32) + t), register t therefore stands for hidden
// unit sigma(8(l/32) + t) = 8(t/4) + 4(l/32) + (t%4) of the K block.  The
// W2 fragments are packed with the same permutation, so no data crosses
// lanes between layers.  b1 enters layer 1 through a constant-1 sixth input;
// b2 initialises the layer-2 accumulator; layer 3 (64 -> 2) is a per-lane dot
// product over the lane's 32 hidden units plus one cross-half exchange.
#pragma once
#include <hip/hip_runtime.h>

#include "vh_device.h"

namespace vh {
namespace ml {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int NIN = 5, NH = 64, NOUT = 2;
constexpr int OFF_W1 = 0, OFF_B1 = 320, OFF_W2 = 384, OFF_B2 = 4480, OFF_W3 = 4544,
              OFF_B3 = 4672, NPARAM = 4674;
static_assert(NPARAM == VH_MLP_NPARAM, "parameter count");

// MFMA operand fragments in LDS/HBM (16-byte aligned, 11024 bytes)
struct alignas(16) Packed {
  bf16x8 w1[2][64];     // [M tile][lane]: A operand of layer 1 (W1 | b1 | 0)
  bf16x8 w2[2][4][64];  // [M tile][K block][lane]: A operand of layer 2, permuted K
  float b2[64];
  float w3[2][64];
  float b3[4];
};
static_assert(sizeof(Packed) == 11024, "packed MLP size");

// features of a con2prim input (q = Q / tau); finite for every input
__host__ __device__ inline void features(const double q[7], float f[5]) {
  const double Qt = q[T_];
  if (!(Qt > 0.)) {
    for (int k = 0; k < 5; k++) f[k] = 0.f;
    return;
  }
  double M = sqrt(q[X_] * q[X_] + q[Y_] * q[Y_] + q[Z_] * q[Z_]);
  if (M > Qt) M = 0.9999 * Qt;
  const double iq = 1. / Qt;  // one division for the three ratios
  const float x = (float)(M * iq);
  f[0] = 0.1f * logf((float)Qt);
  f[1] = x;
  f[2] = (float)(q[Z_] * iq);
  f[3] = (float)(q[NB_] * iq);
  const float g = 1.f - x * x;
  f[4] = sqrtf(g > 0.f ? g : 0.f);
}

// Copies the packed fragments into LDS; every thread of the block calls it,
// followed by __syncthreads() in the caller.
__device__ inline void load_packed(Packed &dst, const Packed *__restrict__ src) {
  const uint4 *s = reinterpret_cast<const uint4 *>(src);
  uint4 *d = reinterpret_cast<uint4 *>(&dst);
  for (int i = threadIdx.x; i < (int)(sizeof(Packed) / 16); i += blockDim.x) d[i] = s[i];
}

// ReLU as an integer max on the bit pattern: negative values (sign bit set)
// become +0, the rest pass unchanged.  Equal to fmaxf(x, 0) for every finite
// x (the activations are finite: finite features and weights), in one
// instruction (fmaxf adds a NaN-quieting max before each).
__device__ __forceinline__ float relu(float x) {
  return __int_as_float(max(__float_as_int(x), 0));
}

// MLP on the 64 lanes' feature vectors; returns (|v| guess, e/Q^tau guess) of
// the calling lane's sample.  Every lane of the wave must call it.
__device__ inline float2 mlp_wave(const Packed &w, const float f[5]) {
  const int l = threadIdx.x & 63, half = l >> 5;
  float x[6] = {f[0], f[1], f[2], f[3], f[4], 1.f};
  float xp[6];
#pragma unroll
  for (int k = 0; k < 6; k++) xp[k] = __shfl_xor(x[k], 32, 64);
  // the two batches (samples of lanes 0..31 and 32..63) side by side, so
  // that four independent MFMA chains are in flight; lanes of the low half
  // carry input rows 0..7 of sample l % 32, the high half rows 8..15 (zero)
  bf16x8 b0, b1;
#pragma unroll
  for (int t = 0; t < 8; t++) {
    b0[t] = (__bf16)((half == 0 && t < 6) ? x[t] : 0.f);
    b1[t] = (__bf16)((half == 0 && t < 6) ? xp[t] : 0.f);
  }
  const floatx16 zero = {};
  const floatx16 d1a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w.w1[0][l], b0, zero, 0, 0, 0);
  const floatx16 d1b0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w.w1[1][l], b0, zero, 0, 0, 0);
  const floatx16 d1a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w.w1[0][l], b1, zero, 0, 0, 0);
  const floatx16 d1b1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w.w1[1][l], b1, zero, 0, 0, 0);
  floatx16 d2a0, d2b0, d2a1, d2b1;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int h = 8 * (r >> 2) + 4 * half + (r & 3);
    d2a0[r] = d2a1[r] = w.b2[h];
    d2b0[r] = d2b1[r] = w.b2[32 + h];
  }
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    bf16x8 h0, h1;
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const float z0 = kb < 2 ? d1a0[8 * (kb & 1) + t] : d1b0[8 * (kb & 1) + t];
      const float z1 = kb < 2 ? d1a1[8 * (kb & 1) + t] : d1b1[8 * (kb & 1) + t];
      h0[t] = (__bf16)relu(z0);
      h1[t] = (__bf16)relu(z1);
    }
    const bf16x8 wa = w.w2[0][kb][l], wb = w.w2[1][kb][l];
    d2a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, h0, d2a0, 0, 0, 0);
    d2b0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wb, h0, d2b0, 0, 0, 0);
    d2a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, h1, d2a1, 0, 0, 0);
    d2b1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wb, h1, d2b1, 0, 0, 0);
  }
  float o00 = 0.f, o10 = 0.f, o01 = 0.f, o11 = 0.f;  // o<output><batch>
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int h = 8 * (r >> 2) + 4 * half + (r & 3);
    const float wa0 = w.w3[0][h], wb0 = w.w3[0][32 + h], wa1 = w.w3[1][h], wb1 = w.w3[1][32 + h];
    const float a0 = relu(d2a0[r]), c0 = relu(d2b0[r]), a1 = relu(d2a1[r]), c1 = relu(d2b1[r]);
    o00 += wa0 * a0 + wb0 * c0;
    o10 += wa1 * a0 + wb1 * c0;
    o01 += wa0 * a1 + wb0 * c1;
    o11 += wa1 * a1 + wb1 * c1;
  }
  o00 += __shfl_xor(o00, 32, 64);
  o10 += __shfl_xor(o10, 32, 64);
  o01 += __shfl_xor(o01, 32, 64);
  o11 += __shfl_xor(o11, 32, 64);
  return half == 0 ? make_float2(o00 + w.b3[0], o10 + w.b3[1])
                   : make_float2(o01 + w.b3[0], o11 + w.b3[1]);
}

// Newton polish of the con2prim root.  Same equation f(v) = (Q_t + p) v - M,
// bracket and Newton/bisection safeguard as transformPV (src/rmn.cpp:72-117),
// but with the exact derivative f' = Q_t + p - M v dp/de (transformPV keeps
// dp/de = 1/3, which for a non-conformal table EoS makes the iteration
// linearly convergent), started from v0 and stopped on the residual
// |f| <= tol Q_t instead of |dv| < 1e-5.  Same vacuum, rest and |Q| > Q_t
// branches (src/rmn.cpp:47-70).  Returns the number of updates, or -1 if the
// residual was not reached within maxit (the caller then falls back to the
// reference iteration).  q is rescaled in place as transformPV does.
__device__ __forceinline__ int c2p_polish(const DevEos &eos, double q[7], double v0, Prim &r,
                                          double tol, int maxit) {
  const double Qt = q[T_];
  double Qx = q[X_], Qy = q[Y_], Qz = q[Z_];
  double M = sqrt(Qx * Qx + Qy * Qy + Qz * Qz);
  double dpe = 1. / 3.;  // dp/de at the current e (exact for a table EoS)
  double e = 0., p = 0., vx = 0., vy = 0., vz = 0., nb = 0., nq = 0., ns = 0.;
  int it = 0;
  if (Qt <= 0.) {
  } else if (M == 0.) {
    e = Qt;
    nb = q[NB_];
    nq = q[NQ_];
    ns = q[NS_];
    p = eos.p(e);
  } else {
    if (M > Qt) {
      Qx *= 0.9999 * Qt / M;
      Qy *= 0.9999 * Qt / M;
      Qz *= 0.9999 * Qt / M;
      M = Qt * 0.9999;
    }
    double v = v0 >= 0. ? (v0 < 0.999999 ? v0 : 0.999999) : 0.;
    double vl = 0., vh = 1.;
    e = Qt - M * v;
    if (e < 0.) e = 0.;
    p = eos.p_dpde(e, dpe);
    double f = (Qt + p) * v - M;
    double df = (Qt + p) - M * v * dpe;
    if (f > 0.)
      vh = v;
    else
      vl = v;
    double dvold = 1., dv = 1.;
    const double ftol = tol * Qt;
    bool conv = fabs(f) <= ftol;
    while (!conv && it < maxit) {
      const bool bis = (f + df * (vh - v)) * (f + df * (vl - v)) >= 0. ||
                       fabs(2. * f) > fabs(dvold * df);
      dvold = dv;
      if (bis) {
        dv = 0.5 * (vh - vl);
        v = vl + dv;
      } else {
        dv = f / df;
        v -= dv;
      }
      it++;
      e = Qt - M * v;
      if (e < 0.) e = 0.;
      p = eos.p_dpde(e, dpe);
      f = (Qt + p) * v - M;
      df = (Qt + p) - M * v * dpe;
      if (f > 0.)
        vh = v;
      else
        vl = v;
      conv = fabs(f) <= ftol;
    }
    if (!conv) it = -1;
    const double iM = v / M;
    vx = Qx * iM;
    vy = Qy * iM;
    vz = Qz * iM;
    e = Qt - M * v;
    p = eos.p(e);
    const double s2 = sqrt(1. - v * v);
    nb = q[NB_] * s2;
    nq = q[NQ_] * s2;
    ns = q[NS_] * s2;
  }
  q[X_] = Qx;
  q[Y_] = Qy;
  q[Z_] = Qz;
  r = Prim{e, p, nb, nq, ns, vx, vy, vz};
  return it;
}

constexpr double POLISH_TOL = 1e-12;
constexpr int POLISH_MAXIT = 30;

// con2prim from a guess: polish, else the reference iteration on the
// (possibly rescaled) input.  Returns polish updates, or -1 - (reference
// iterations) on fallback.
__device__ __forceinline__ int c2p_guess(const DevEos &eos, double q[7], double v0, Prim &r) {
  const int it = c2p_polish(eos, q, v0, r, POLISH_TOL, POLISH_MAXIT);
  if (it >= 0) return it;
  return -1 - c2p_ref<false>(eos, q, r);
}

}  // namespace ml
}  // namespace vh
