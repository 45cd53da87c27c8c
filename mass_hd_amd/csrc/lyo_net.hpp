// lyo_net.hpp -- kernels of the surrogate corrector network (gfx950).
//
// Network: four periodic 3x3 convolutions 8 -> 32 -> 32 -> 32 -> 8 channels
// with tanh-GELU after the first three.  Activations are bf16 [pixel][C]
// (pixel k = y + LY*x, C = 8 or 32 contiguous); accumulation is FP32.
//
// conv_mfma_kernel: implicit GEMM on v_mfma_f32_16x16x32_bf16.  One wave
// computes 16 consecutive pixels of a row (M) x 16 output channels (N) per
// MFMA block; K = 32 is one tap x 32 input channels (Cin = 32) or four taps x
// 8 input channels (Cin = 8).  The weights of a layer are packed on the host
// into per-lane fragments and held in registers for the whole (persistent)
// wave, so the inner loop only loads 16 bytes of activations per lane per
// K-slice.  The same kernel runs the backward data gradient, with the weights
// transposed and the taps mirrored.
//
// MFMA 16x16x32 lane layout (checked on the device): lane l = 16*q + i holds
// A[i][8q..8q+7], B[8q..8q+7][i] and D[4q..4q+3][i].
#pragma once

#include <hip/hip_runtime.h>

namespace lyo {
namespace net {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTaps = 9;
constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)

__device__ __forceinline__ float gelu(float x) {
  const float u = kGeluC * (x + 0.044715f * x * x * x);
  const float t = 1.f - 2.f / (__expf(2.f * u) + 1.f);  // tanh(u)
  return 0.5f * x * (1.f + t);
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float u = kGeluC * (x + 0.044715f * x * x * x);
  const float t = 1.f - 2.f / (__expf(2.f * u) + 1.f);
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * kGeluC * (1.f + 3.f * 0.044715f * x * x);
}

// tap t = 3*(dx+1) + (dy+1)
__device__ __forceinline__ int tap_dx(int t) { return t / 3 - 1; }
__device__ __forceinline__ int tap_dy(int t) { return t % 3 - 1; }

// number of K-slices of 32 for a layer with cin input channels
__host__ __device__ constexpr int kslices(int cin) { return cin == 32 ? 9 : 3; }

enum Epilogue {
  EPI_LINEAR = 0,      // out = acc + b                      (f32 and/or bf16 out)
  EPI_GELU = 1,        // z = acc + b, out = gelu(z)          (z saved when zout)
  EPI_GELU_GRAD = 2    // out = acc * gelu'(z_in)             (backward)
};

struct ConvArgs {
  const __bf16* in;     // [N][cin]
  const bf16x8* wfrag;  // [kslices][nblk][64] fragments
  const float* bias;    // [cout] or null
  float* zout;          // EPI_GELU: pre-activation [N][cout] (training) or null
  const float* zin;     // EPI_GELU_GRAD: pre-activation of the layer input [N][cout]
  __bf16* out_bf;       // [N][cout] or null
  float* out_f32;       // [N][cout] or null
  int cin, cout;        // 8 or 32
  int LX, LY;           // LY % 16 == 0
  int epi;
};

__global__ void __launch_bounds__(256) conv_mfma_kernel(const ConvArgs a) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, i = lane & 15;
  const int nblk = a.cout / 16 > 0 ? a.cout / 16 : 1;
  const int ks = kslices(a.cin);
  // weights of this layer in registers: at most 9 slices x 2 blocks
  bf16x8 w[9][2];
#pragma unroll
  for (int s = 0; s < 9; ++s)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
      if (s < ks && nb < nblk) w[s][nb] = a.wfrag[(s * nblk + nb) * 64 + lane];
  float bias[2] = {0.f, 0.f};
  if (a.bias)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      const int co = nb * 16 + i;
      if (nb < nblk && co < a.cout) bias[nb] = a.bias[co];
    }

  const int tiles_per_row = a.LY >> 4;
  const int ntiles = a.LX * tiles_per_row;
  const int nw = (int)(gridDim.x * (blockDim.x >> 6));
  for (int tile = __builtin_amdgcn_readfirstlane((int)blockIdx.x * ((int)blockDim.x >> 6) +
                                                  ((int)threadIdx.x >> 6));
       tile < ntiles; tile += nw) {
    const int x = tile / tiles_per_row;
    const int y0 = (tile - x * tiles_per_row) << 4;
    f32x4 acc[2];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      if (s < ks) {
        // this lane's A fragment: 8 channels of pixel (x+dx, y0+i+dy)
        int t, c0;
        if (a.cin == 32) {
          t = s;
          c0 = 8 * q;
        } else {
          t = 4 * s + q;
          c0 = 0;
        }
        bf16x8 av;
        if (t < kTaps) {
          int xx = x + tap_dx(t), yy = y0 + i + tap_dy(t);
          xx += (xx < 0) ? a.LX : 0;
          xx -= (xx >= a.LX) ? a.LX : 0;
          yy += (yy < 0) ? a.LY : 0;
          yy -= (yy >= a.LY) ? a.LY : 0;
          const unsigned k = (unsigned)(xx * a.LY + yy);
          av = *reinterpret_cast<const bf16x8*>(a.in + (size_t)k * a.cin + c0);
        } else {
          av = bf16x8{};
        }
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
          if (nb < nblk) acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, w[s][nb], acc[nb], 0, 0, 0);
      }
    }
    // epilogue: lane holds D[4q + r][i] = pixel y0 + 4q + r, channel nb*16 + i
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      const int co = nb * 16 + i;
      if (nb >= nblk || co >= a.cout) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const size_t k = (size_t)x * a.LY + y0 + 4 * q + r;
        float v = acc[nb][r];
        if (a.epi == EPI_GELU_GRAD) {
          v *= gelu_grad(a.zin[k * a.cout + co]);
        } else {
          v += bias[nb];
          if (a.epi == EPI_GELU) {
            if (a.zout) a.zout[k * a.cout + co] = v;
            v = gelu(v);
          }
        }
        if (a.out_f32) a.out_f32[k * a.cout + co] = v;
        if (a.out_bf) a.out_bf[k * a.cout + co] = (__bf16)v;
      }
    }
  }
}

// Weight gradient (FP32, LDS-staged): dW[t][co][ci] += sum_p d[p][co] *
// in[p + off(t)][ci] and db[co] += sum_p d[p][co] over the 16-pixel row tiles
// [tile0, tile0 + tiles) of one block.  One thread per (co, ci) pair.
__global__ void __launch_bounds__(256)
conv_wgrad_kernel(const __bf16* __restrict__ in, const __bf16* __restrict__ d, int cin, int cout,
                  int LX, int LY, int tiles_per_block, float* __restrict__ dW,
                  float* __restrict__ db) {
  __shared__ float sd[16][32];
  __shared__ float sin_[3][18][32];  // rows x-1..x+1, columns y0-1..y0+16
  const int tiles_per_row = LY >> 4;
  const int ntiles = LX * tiles_per_row;
  const int pairs = cin * cout;
  float accw[4][kTaps];  // up to 4 (co,ci) pairs per thread (32*32/256)
  float accb = 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int t = 0; t < kTaps; ++t) accw[j][t] = 0.f;
  const int t0 = blockIdx.x * tiles_per_block;
  for (int tile = t0; tile < min(t0 + tiles_per_block, ntiles); ++tile) {
    const int x = tile / tiles_per_row;
    const int y0 = (tile - x * tiles_per_row) << 4;
    __syncthreads();
    for (int e = threadIdx.x; e < 16 * cout; e += 256) {
      const int p = e / cout, c = e - p * cout;
      sd[p][c] = (float)d[((size_t)x * LY + y0 + p) * cout + c];
    }
    for (int e = threadIdx.x; e < 3 * 18 * cin; e += 256) {
      const int c = e % cin, r = e / cin;
      const int row = r / 18, col = r % 18;
      int xx = x + row - 1, yy = y0 + col - 1;
      xx += (xx < 0) ? LX : 0;
      xx -= (xx >= LX) ? LX : 0;
      yy += (yy < 0) ? LY : 0;
      yy -= (yy >= LY) ? LY : 0;
      sin_[row][col][c] = (float)in[((size_t)xx * LY + yy) * cin + c];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int pr = threadIdx.x + 256 * j;
      if (pr >= pairs) break;
      const int co = pr / cin, ci = pr - co * cin;
      for (int p = 0; p < 16; ++p) {
        const float dv = sd[p][co];
#pragma unroll
        for (int t = 0; t < kTaps; ++t) accw[j][t] += dv * sin_[t / 3][p + t % 3][ci];
      }
    }
    if (threadIdx.x < cout)
      for (int p = 0; p < 16; ++p) accb += sd[p][threadIdx.x];
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int pr = threadIdx.x + 256 * j;
    if (pr >= pairs) break;
    const int co = pr / cin, ci = pr - co * cin;
#pragma unroll
    for (int t = 0; t < kTaps; ++t) atomicAdd(&dW[(t * cout + co) * cin + ci], accw[j][t]);
  }
  if (threadIdx.x < cout) atomicAdd(&db[threadIdx.x], accb);
}

// Loss of the output layer: out (f32 [N][8]) against y (f32 [N][8]);
// L = sum (out - y)^2 / (2 N); d = (out - y) / N as bf16.  Per-block partial
// sums of the squared error.
__global__ void __launch_bounds__(256)
loss_grad_kernel(const float* __restrict__ out, const float* __restrict__ y, int N,
                 __bf16* __restrict__ d, double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0;
  const float invn = 1.f / (float)N;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < N * 8; e += gridDim.x * 256) {
    const float r = out[e] - y[e];
    s += (double)r * r;
    d[e] = (__bf16)(r * invn);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + (red[2] + red[3])) / (2.0 * N);
}

// Adam step on FP32 master parameters.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace net
}  // namespace lyo
