// Host-side helpers of the learned con2prim shared by the context code.
#pragma once
#include <hip/hip_runtime.h>

#include "vh_mlp.h"

namespace vh {
namespace ml {

// FP32 parameter vector -> MFMA fragments (BF16, round to nearest even)
void pack(const float *params, Packed &out);

// per-cell |v| guesses gv[N] of Qs (SoA 7 x N) / (*tp + toff)
void launch_guess(const double *Qs, long N, int nc, const double *tp, double toff, const Packed *P,
                  float *gv, hipStream_t st);

}  // namespace ml
}  // namespace vh
