// lyo_net_act.hpp -- operand types and the activation of the surrogate
// network, shared by its layer kernels (lyo_net.hpp) and the one-kernel
// network step (lyo_net_fused.hpp).
#pragma once

#include <hip/hip_runtime.h>

namespace lyo {
namespace net {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTaps = 9;
constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)

// tanh(u) = 1 - 2/(exp(2u) + 1) with u = sqrt(2/pi) (x + 0.044715 x^3), a
// hardware exp2 and reciprocal: exp(2u) = 2^(x (K1 + K3 x^2)) with K1 =
// 2 sqrt(2/pi) log2(e), K3 = 0.044715 K1.  gelu(x) = x/2 (1 + tanh(u)) =
// x (1 - r) with r = 1/(exp(2u) + 1).  exp overflowing to +inf gives r = 0
// (gelu = x), exp underflowing to 0 gives r = 1.
constexpr float kGeluK1 = 2.3022082f;  // 2 * sqrt(2/pi) * log2(e)
constexpr float kGeluK3 = 0.10294324f;  // 0.044715 * kGeluK1
__device__ __forceinline__ float gelu_r(float x) {
  const float w = x * __builtin_fmaf(x * x, kGeluK3, kGeluK1);
  return __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(w) + 1.f);
}
__device__ __forceinline__ float gelu(float x) { return __builtin_fmaf(-x, gelu_r(x), x); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float r = gelu_r(x);
  const float t = 1.f - 2.f * r;  // tanh(u)
  return (1.f - r) + 0.5f * x * (1.f - t * t) * kGeluC * (1.f + 3.f * 0.044715f * x * x);
}

}  // namespace net
}  // namespace lyo
