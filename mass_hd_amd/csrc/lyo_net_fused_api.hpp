// lyo_net_fused_api.hpp -- launch interface of the one-kernel surrogate
// corrector (lyo_net_fused.hpp, built in lyo_net_fused.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "lyo_physics.hpp"

namespace lyo {

constexpr int kNetStrip = 52;    // interior columns per wave
constexpr int kNetRepeat = 10;   // rows a chunk computes beyond its own
constexpr int kNetMinSide = 16;  // smallest LX and LY

// The network's weights in the fused kernel's layout (device memory):
// fragments (47 x 64 lanes x 8 bf16), layer biases (104), the input
// normalisation (mean[8], inv_std[8]) and the output scales (8: the
// network's own, or a run's calibrated ones).
struct NetFusedW {
  const void* wfrag = nullptr;
  const float* bias = nullptr;
  const float* norm = nullptr;
  const float* out_scale = nullptr;
};

// A state buffer set: f' (9 planes of NP values), Qxx, Qyx, phi and phi's
// compensation.
struct StatePlanes {
  float* f = nullptr;
  float* qxx = nullptr;
  float* qyx = nullptr;
  float* phi = nullptr;
  float* phic = nullptr;
};

// The network part of a surrogate step: reads the one-corrector state `in`,
// writes the corrected state to `out` (another buffer set), xc rows per wave
// chunk.  delta (optional, [LX*LY][8]): the increments applied.  dphi_part
// receives one sum of the applied phi increments per wave (net_fused_waves).
hipError_t launch_net_fused(const StatePlanes& in, const StatePlanes& out, int LX, int LY, int NP,
                            int xc, const Params<float>& P, double rho, const NetFusedW& w,
                            float* delta, double* dphi_part, hipStream_t s);
// waves of that launch
int net_fused_waves(int LX, int LY, int xc);

}  // namespace lyo
