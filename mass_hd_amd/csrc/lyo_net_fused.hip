// lyo_net_fused.hip -- the one-kernel surrogate corrector (FP32), built with
// the FP32 step's contraction rules so that its features are computed as the
// separate features kernel computes them.
#include "lyo_net_fused.hpp"
#include "lyo_net_fused_api.hpp"

namespace lyo {

static_assert(netf::kStrip == kNetStrip, "strip width of the launch interface");
static_assert(netf::kWarm == kNetRepeat, "rows a chunk repeats");

int net_fused_waves(int LX, int LY, int xc) {
  return ((LY + kNetStrip - 1) / kNetStrip) * ((LX + xc - 1) / xc);
}

hipError_t launch_net_fused(const StatePlanes& in, const StatePlanes& out, int LX, int LY, int NP,
                            int xc, const Params<float>& P, double rho, const NetFusedW& w,
                            float* delta, double* dphi_part, hipStream_t s) {
  if (LX < kNetMinSide || LY < kNetMinSide || xc < 1 || NP < LX * LY || !dphi_part || !w.wfrag || !w.bias ||
      !w.norm || !w.out_scale || !in.f || !in.qxx || !in.qyx || !in.phi || !in.phic || !out.f || !out.qxx ||
      !out.qyx || !out.phi || !out.phic)
    return hipErrorInvalidValue;
  Geo g{};
  g.LX = LX;
  g.LY = LY;
  g.N = LX * LY;
  g.NP = NP;
  g.nstrips = (LY + kNetStrip - 1) / kNetStrip;
  g.xc = xc;
  g.nchunks = (LX + xc - 1) / xc;
  g.nwaves = g.nstrips * g.nchunks;
  g.xcd = 1;
  g.alt = 0;
  if (!geo_ok(g, kNetStrip, kNetMinSide)) return hipErrorInvalidValue;
  netf::Args a{};
  a.f_in = in.f;
  a.qxx_in = in.qxx;
  a.qyx_in = in.qyx;
  a.phi_in = in.phi;
  a.phic_in = in.phic;
  a.f_out = out.f;
  a.qxx_out = out.qxx;
  a.qyx_out = out.qyx;
  a.phi_out = out.phi;
  a.phic_out = out.phic;
  a.wfrag = static_cast<const netf::bf16x8*>(w.wfrag);
  a.bias = w.bias;
  a.norm = w.norm;
  a.out_scale = w.out_scale;
  a.delta = delta;
  a.dphi_part = dphi_part;
  a.rho = rho;
  hipError_t e = hipFuncSetAttribute((const void*)netf::net_fused_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)netf::kLdsBytes);
  if (e != hipSuccess) return e;
  const int blocks = (g.nwaves + netf::kWaves - 1) / netf::kWaves;
  hipLaunchKernelGGL(netf::net_fused_kernel, dim3(blocks), dim3(64 * netf::kWaves),
                     netf::kLdsBytes, s, a, P, g);
  return hipGetLastError();
}

}  // namespace lyo
