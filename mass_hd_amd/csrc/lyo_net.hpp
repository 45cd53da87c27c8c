// lyo_net.hpp -- kernels of the surrogate corrector network (gfx950).
//
// Network: four periodic 3x3 convolutions 8 -> 32 -> 32 -> 32 -> 8 channels
// with tanh-GELU after the first three.  Activations are bf16 [pixel][C]
// (pixel k = y + LY*x, C = 8 or 32 contiguous); accumulation is FP32.
//
// conv_rows_kernel: implicit GEMM on v_mfma_f32_16x16x32_bf16 with the
// weights as the A operand (M = 16 output channels) and the activations as B
// (N = 16 consecutive pixels of a row), so a lane's accumulator holds four
// consecutive output channels of one pixel and the epilogue stores 8 or 16
// bytes per lane.  K = 32 is one tap x 32 input channels (Cin = 32) or four
// taps x 8 input channels (Cin = 8).  One wave computes ROWS output rows x 16
// pixels per task: it streams the ROWS + 2 input rows once, and each row's
// three y-shifted fragments feed the up to three output rows (x-offsets)
// that use them.  The weights of a layer are packed on the host into
// per-lane fragments and held in registers for the whole (persistent) wave,
// one wave per block.
// The same kernel runs the backward data gradient, with the weights
// transposed and the taps mirrored.
//
// MFMA 16x16x32 lane layout (checked on the device): lane l = 16*q + i holds
// A[i][8q..8q+7], B[8q..8q+7][i] and D[4q..4q+3][i].
#pragma once

#include <hip/hip_runtime.h>

#include "lyo_net_act.hpp"

namespace lyo {
namespace net {

// tap t = 3*(dx+1) + (dy+1)
__host__ __device__ constexpr int tap_dx(int t) { return t / 3 - 1; }
__host__ __device__ constexpr int tap_dy(int t) { return t % 3 - 1; }

// number of K-slices of 32 for a layer with cin input channels
__host__ __device__ constexpr int kslices(int cin) { return cin == 32 ? 9 : 3; }

enum Epilogue {
  EPI_LINEAR = 0,      // out = (acc + b) * scale               (f32 and/or bf16 out)
  EPI_GELU = 1,        // z = acc + b, out = gelu(z)          (z saved when zout)
  EPI_GELU_GRAD = 2    // out = acc * gelu'(z_in)             (backward)
};

struct ConvArgs {
  const __bf16* in;        // [N][cin] bf16, or null when in_f32 is given
  const float* in_f32;     // [N][cin] FP32 (cin = 8), normalised on load
  const float* in_norm;    // in_f32: mean[cin], inv_std[cin]
  const bf16x8* wfrag;     // [kslices][nblk][64] fragments
  const float* bias;       // [cout] or null
  float* zout;             // EPI_GELU: pre-activation [N][cout] (training) or null
  const float* zin;        // EPI_GELU_GRAD: pre-activation of the layer input [N][cout]
  __bf16* out_bf;          // [N][cout]: EPI_GELU, EPI_GELU_GRAD
  float* out_f32;          // [N][cout]: EPI_LINEAR
  const float* out_scale;  // EPI_LINEAR: [cout] factors of the output, or null
  int cin, cout;           // 8 or 32
  int LX, LY;              // LY % 16 == 0
};

// Raw input fragment of one lane: 8 channels of one pixel, bf16 or (layer 0
// of inference) FP32 still to be normalised.
template <bool F32>
struct InFrag {
  bf16x8 v;
};
template <>
struct InFrag<true> {
  f32x4 lo, hi;
};

template <int CIN, bool F32>
__device__ __forceinline__ void load_in(const ConvArgs& a, size_t k, int c0, InFrag<F32>& f) {
  if constexpr (F32) {
    const f32x4* p = reinterpret_cast<const f32x4*>(a.in_f32 + k * CIN + c0);
    f.lo = p[0];
    f.hi = p[1];
  } else {
    f.v = *reinterpret_cast<const bf16x8*>(a.in + k * CIN + c0);
  }
}

template <bool F32>
__device__ __forceinline__ bf16x8 to_operand(const InFrag<F32>& f, const float (&nm)[8],
                                             const float (&ni)[8]) {
  if constexpr (F32) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = (__bf16)((f.lo[j] - nm[j]) * ni[j]);
      v[4 + j] = (__bf16)((f.hi[j] - nm[4 + j]) * ni[4 + j]);
    }
    return v;
  } else {
    return f.v;
  }
}

// x modulo LX for any x (rows of tiny grids wrap more than once)
__device__ __forceinline__ int wrap_row(int x, int LX) {
  x %= LX;
  return x < 0 ? x + LX : x;
}

// Input fragments of one task: Cin = 32: rows x0-1 .. x0+ROWS, three
// y-shifts each; Cin = 8: per output row, the three K-slices (lane quarter q
// holds tap 4s+q; taps >= 9 load the centre pixel and are zeroed at use).
template <int CIN, int ROWS>
constexpr int frag_rows() { return CIN == 32 ? ROWS + 2 : ROWS; }

template <int CIN, int ROWS, bool F32>
__device__ __forceinline__ void load_task(const ConvArgs& a, int task, int ytiles, int q, int i,
                                          InFrag<F32> (&fr)[frag_rows<CIN, ROWS>()][3]) {
  const int LX = a.LX, LY = a.LY;
  const int xt = task / ytiles;
  const int x0 = xt * ROWS, yc = ((task - xt * ytiles) << 4) + i;
  const int yv[3] = {yc == 0 ? LY - 1 : yc - 1, yc, yc + 1 == LY ? 0 : yc + 1};
  if constexpr (CIN == 32) {
#pragma unroll
    for (int j = 0; j < ROWS + 2; ++j) {
      const size_t row = (size_t)wrap_row(x0 - 1 + j, LX) * LY;
#pragma unroll
      for (int d = 0; d < 3; ++d) load_in<CIN, F32>(a, row + yv[d], 8 * q, fr[j][d]);
    }
  } else {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      const int xv[3] = {wrap_row(x0 + r - 1, LX), wrap_row(x0 + r, LX),
                         wrap_row(x0 + r + 1, LX)};
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int t = 4 * s + q < kTaps ? 4 * s + q : 4;
        const int dxi = tap_dx(t) + 1, dyi = tap_dy(t) + 1;
        const int xr = dxi == 0 ? xv[0] : (dxi == 1 ? xv[1] : xv[2]);
        const int yy = dyi == 0 ? yv[0] : (dyi == 1 ? yv[1] : yv[2]);
        load_in<CIN, F32>(a, (size_t)xr * LY + yy, 0, fr[r][s]);
      }
    }
  }
}

// conv_rows_kernel: see the file header.  Each wave loads the input
// fragments of its next task before computing the current one (two register
// buffers, the task loop unrolled by two), so a task's loads are in flight
// for a whole task of MFMAs and epilogue.  GUARD: LX % ROWS != 0, the last
// row of tasks is partial (its extra rows are computed from wrapped rows and
// not stored).
template <int CIN, int COUT, int ROWS, bool F32, bool GUARD, int EPI, bool ZOUT>
__global__ void __launch_bounds__(64) conv_rows_kernel(const ConvArgs a) {
  constexpr int KS = kslices(CIN);
  constexpr int MB = COUT >= 16 ? COUT / 16 : 1;
  constexpr int FR = frag_rows<CIN, ROWS>();
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, i = lane & 15;
  bf16x8 w[KS][MB];
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) w[s][mb] = a.wfrag[(s * MB + mb) * 64 + lane];
  // the lane's output channels: (mb*16 + 4q + r) mod COUT (COUT = 8: the
  // packed weights repeat channels 0..7 in M rows 8..15, so lanes q >= 2 hold
  // the same values as lanes q - 2 and store them to the same addresses)
  float bias[MB][4], osc[MB][4];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int c = (mb * 16 + 4 * q + r) % COUT;
      bias[mb][r] = a.bias ? a.bias[c] : 0.f;
      osc[mb][r] = a.out_scale ? a.out_scale[c] : 1.f;
    }
  float nm[8], ni[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    nm[j] = F32 ? a.in_norm[j] : 0.f;
    ni[j] = F32 ? a.in_norm[8 + j] : 1.f;
  }
  // CIN = 8: K-slice 2 of lane quarters 1..3 holds taps 9..11 (zero weights)
  const bool pad_tap = CIN == 8 && q > 0;

  const int LX = a.LX, LY = a.LY;
  const int ytiles = LY >> 4;
  const int ntasks = ((LX + ROWS - 1) / ROWS) * ytiles;
  const int nw = (int)gridDim.x;

  auto compute_store = [&](int task, const InFrag<F32> (&fr)[FR][3]) {
    f32x4 acc[ROWS][MB];
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) acc[r][mb] = f32x4{0.f, 0.f, 0.f, 0.f};
    if constexpr (CIN == 32) {
      // input row j (x0 - 1 + j) feeds output row r = j - dxi (tap x-offset dxi - 1)
#pragma unroll
      for (int j = 0; j < ROWS + 2; ++j) {
        bf16x8 b[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) b[d] = to_operand<F32>(fr[j][d], nm, ni);
#pragma unroll
        for (int dxi = 0; dxi < 3; ++dxi) {
          const int r = j - dxi;
          if (r >= 0 && r < ROWS) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
              for (int mb = 0; mb < MB; ++mb)
                acc[r][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[3 * dxi + d][mb], b[d],
                                                                      acc[r][mb], 0, 0, 0);
          }
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          bf16x8 b = to_operand<F32>(fr[r][s], nm, ni);
          if (s == 2 && pad_tap) b = bf16x8{};
#pragma unroll
          for (int mb = 0; mb < MB; ++mb)
            acc[r][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[s][mb], b, acc[r][mb], 0, 0, 0);
        }
      }
    }
    // epilogue: lane holds D[4q + r'][i] = channel mb*16 + 4q + r' of pixel
    // yc; EPI_LINEAR writes FP32 (out_f32), the others bf16 (out_bf)
    const int xt = task / ytiles;
    const int x0 = xt * ROWS, yc = ((task - xt * ytiles) << 4) + i;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (GUARD && x0 + r >= LX) continue;
      const size_t k = (size_t)(x0 + r) * LY + yc;
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const int c = (mb * 16 + 4 * q) % COUT;
        const size_t e = k * COUT + c;
        f32x4 v = acc[r][mb];
        if constexpr (EPI == EPI_GELU_GRAD) {
          const f32x4 z = *reinterpret_cast<const f32x4*>(a.zin + e);
#pragma unroll
          for (int j = 0; j < 4; ++j) v[j] *= gelu_grad(z[j]);
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) v[j] += bias[mb][j];
          if constexpr (EPI == EPI_GELU) {
            if constexpr (ZOUT) *reinterpret_cast<f32x4*>(a.zout + e) = v;
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = gelu(v[j]);
          } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] *= osc[mb][j];
          }
        }
        if constexpr (EPI == EPI_LINEAR) {
          *reinterpret_cast<f32x4*>(a.out_f32 + e) = v;
        } else {
          bf16x4 o;
#pragma unroll
          for (int j = 0; j < 4; ++j) o[j] = (__bf16)v[j];
          *reinterpret_cast<bf16x4*>(a.out_bf + e) = o;
        }
      }
    }
  };

  int ta = (int)blockIdx.x;
  if (ta >= ntasks) return;
  InFrag<F32> fa[FR][3], fb[FR][3];
  load_task<CIN, ROWS, F32>(a, ta, ytiles, q, i, fa);
  while (true) {
    // the next task's loads go out first (a valid task index past the end:
    // loaded and never used)
    const int tb = ta + nw;
    load_task<CIN, ROWS, F32>(a, tb < ntasks ? tb : ntasks - 1, ytiles, q, i, fb);
    compute_store(ta, fa);
    if (tb >= ntasks) break;
    const int tc = tb + nw;
    load_task<CIN, ROWS, F32>(a, tc < ntasks ? tc : ntasks - 1, ytiles, q, i, fa);
    compute_store(tb, fb);
    if (tc >= ntasks) break;
    ta = tc;
  }
}

// Weight and bias gradients of one layer on MFMA:
//   dW[t][co][ci] = sum_p d[p][co] * in[p + off(t)][ci],  db[co] = sum_p d[p][co]
// with v_mfma_f32_16x16x16bf16_1k, M = 16 output channels, N = 16 input
// channels and K = 16 consecutive pixels of a row.  Those operands need four
// consecutive pixels of one channel per lane (lane l = 16q + i: A[i][4q..4q+3],
// B[4q..4q+3][i]), while the activations are [pixel][channel]: each 16-pixel x
// 32-channel fragment (lane: pixel i, channels 8q..8q+7) is transposed by a
// 16x16x32 MFMA against a 0/1 matrix selecting 16 channels, which leaves lane
// (q, i) with channel i of pixels 4q..4q+3 -- exactly, one product per sum.
// Each wave accumulates a grid-stride set of 16-pixel row segments and writes
// its partial sums to part[wave][9*COUT*CIN + COUT] (the layer's parameter
// order); wgrad_reduce_kernel adds the partials of all waves in wave order.
typedef short s16x4 __attribute__((ext_vector_type(4)));

template <int CIN, int COUT>
__global__ void __launch_bounds__(64)
conv_wgrad_mfma_kernel(const __bf16* __restrict__ in, const __bf16* __restrict__ d, int LX, int LY,
                       float* __restrict__ part) {
  constexpr int MB = COUT >= 16 ? COUT / 16 : 1;  // blocks of output channels
  constexpr int NB = CIN >= 16 ? CIN / 16 : 1;    // blocks of input channels
  constexpr int NPAR = kTaps * COUT * CIN + COUT;
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, i = lane & 15;
  // selection matrices: B[k][n] = 1 where k = 16*blk + n (lane: B[8q..8q+7][i])
  bf16x8 sel[2];
#pragma unroll
  for (int blk = 0; blk < 2; ++blk)
#pragma unroll
    for (int j = 0; j < 8; ++j) sel[blk][j] = (__bf16)(8 * q + j == 16 * blk + i ? 1.f : 0.f);
  // the transpose of a natural fragment, as a bf16 operand of the K=16 MFMA
  auto transpose = [&](const bf16x8& nat, int blk) {
    const f32x4 t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nat, sel[blk], f32x4{0.f, 0.f, 0.f, 0.f},
                                                           0, 0, 0);
    s16x4 o;
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = __builtin_bit_cast(short, (__bf16)t[r]);
    return o;
  };
  f32x4 acc[kTaps][MB][NB];
#pragma unroll
  for (int t = 0; t < kTaps; ++t)
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[t][mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) bsum[mb] = 0.f;

  const int ytiles = LY >> 4;
  const int nseg = LX * ytiles;
  for (int seg = (int)blockIdx.x; seg < nseg; seg += (int)gridDim.x) {
    const int x = seg / ytiles;
    const int y0 = (seg - x * ytiles) << 4;
    // d of the segment: pixel y0+i, channels 8q..8q+7 (zero past COUT)
    bf16x8 dn = bf16x8{};
    if (8 * q < COUT)
      dn = *reinterpret_cast<const bf16x8*>(d + ((size_t)x * LY + y0 + i) * COUT + 8 * q);
    s16x4 dt[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const f32x4 t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(dn, sel[mb],
                                                             f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        dt[mb][r] = __builtin_bit_cast(short, (__bf16)t[r]);
        bsum[mb] += t[r];
      }
    }
#pragma unroll
    for (int dxi = 0; dxi < 3; ++dxi) {
      int xr = x + dxi - 1;
      xr += xr < 0 ? LX : 0;
      xr -= xr >= LX ? LX : 0;
#pragma unroll
      for (int dyi = 0; dyi < 3; ++dyi) {
        int yy = y0 + i + dyi - 1;
        yy += yy < 0 ? LY : 0;
        yy -= yy >= LY ? LY : 0;
        bf16x8 an = bf16x8{};
        if (8 * q < CIN)
          an = *reinterpret_cast<const bf16x8*>(in + ((size_t)xr * LY + yy) * CIN + 8 * q);
        const int t = 3 * dxi + dyi;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          const s16x4 at = transpose(an, nb);
#pragma unroll
          for (int mb = 0; mb < MB; ++mb)
            acc[t][mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(dt[mb], at, acc[t][mb][nb],
                                                                       0, 0, 0);
        }
      }
    }
  }
  // lane (q, i) holds dW_t[co = 16 mb + 4q + r][ci = 16 nb + i]
  float* o = part + (size_t)blockIdx.x * NPAR;
#pragma unroll
  for (int t = 0; t < kTaps; ++t)
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int co = 16 * mb + 4 * q + r, ci = 16 * nb + i;
          if (co < COUT && ci < CIN) o[(t * COUT + co) * CIN + ci] = acc[t][mb][nb][r];
        }
  // bias: channel 16 mb + i, summed over the four lane quarters
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) {
    float v = bsum[mb];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    const int co = 16 * mb + i;
    if (q == 0 && co < COUT) o[kTaps * COUT * CIN + co] = v;
  }
}

// out[e] = sum over waves w of part[w][e], e < n: a block takes 32
// consecutive elements; its eight groups of 32 threads add the waves
// w = g, g + 8, ... (coalesced rows), then the eight group sums are added in
// group order (a fixed order: the result does not depend on timing).
__global__ void __launch_bounds__(256) wgrad_reduce_kernel(const float* __restrict__ part, int nw,
                                                           int n, float* __restrict__ out) {
  __shared__ float red[8][32];
  const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int e = blockIdx.x * 32 + c;
  float s = 0.f;
  if (e < n)
    for (int w = g; w < nw; w += 8) s += part[(size_t)w * n + e];
  red[g][c] = s;
  __syncthreads();
  if (g == 0 && e < n) {
    float t = red[0][c];
#pragma unroll
    for (int k = 1; k < 8; ++k) t += red[k][c];
    out[e] = t;
  }
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
