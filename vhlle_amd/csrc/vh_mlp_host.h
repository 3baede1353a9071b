// Host-side helpers of the learned con2prim shared by the context code.
#pragma once
#include <hip/hip_runtime.h>

#include "vh_mlp.h"

namespace vh {
namespace ml {

// EoSs::t constant of the ideal massless gas (src/eos.cpp:98-99)
inline double ideal_eos_cnst() {
  const double pi_ = 3.14159265358979312;
  return (16 + 0.5 * 21.0 * 2.5) * (pi_ * pi_) / 30.0 /
         (0.197326968 * 0.197326968 * 0.197326968);
}

// FP32 parameter vector -> MFMA fragments (BF16, round to nearest even)
void pack(const float *params, Packed &out);

// per-cell |v| guesses gv[N] of Qs (SoA 7 x N) / (*tp + toff)
void launch_guess(const double *Qs, long N, int nc, const double *tp, double toff, const Packed *P,
                  float *gv, hipStream_t st);

}  // namespace ml
}  // namespace vh
