// lyo_net_internal.hpp -- network entry points used by the run handle.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/mass_lyo.h"
#include "lyo_net_fused_api.hpp"

namespace lyo {
// delta (FP32 [N][8], physical units) = network(features FP32 [N][8]) on s
hipError_t net_predict(mass_lyo_net_t n, const float* feat, int LX, int LY, float* delta,
                       hipStream_t s);
// the network's first-layer input buffer for N nodes (normalised bf16
// [N][8], allocated on first use) and its input normalisation (device:
// mean[8], inv_std[8])
hipError_t net_input(mass_lyo_net_t n, size_t N, __bf16** in, const float** norm);
// delta (FP32 [N][8], physical units) = network(the input buffer) on s;
// out_scale (device [8]) replaces the network's own output scales when given
hipError_t net_predict_input(mass_lyo_net_t n, int LX, int LY, float* delta, hipStream_t s,
                             const float* out_scale = nullptr);
// the network's output scales (host copy, 8 values)
void net_out_scale(mass_lyo_net_t n, float* out8);
// trained (normalisation known)
bool net_ready(mass_lyo_net_t n);
// the weights in the layout of the one-kernel network step (device memory,
// repacked with the other fragments after every parameter update)
NetFusedW net_fused_weights(mass_lyo_net_t n, const float* out_scale = nullptr);
}  // namespace lyo
