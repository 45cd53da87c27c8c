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

namespace lyo {
namespace net {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTaps = 9;
constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)

// tanh(u) = 1 - 2/(exp(2u) + 1), with a hardware reciprocal; gelu(x) =
// x/2 (1 + tanh(u)) = x (1 - r) with r = 1/(exp(2u) + 1).  exp overflowing
// to +inf gives r = 0 (gelu = x), exp underflowing to 0 gives r = 1.
__device__ __forceinline__ float gelu_r(float x) {
  const float u = kGeluC * (x + 0.044715f * x * x * x);
  return __builtin_amdgcn_rcpf(__expf(2.f * u) + 1.f);
}
__device__ __forceinline__ float gelu(float x) { return x - x * gelu_r(x); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float r = gelu_r(x);
  const float t = 1.f - 2.f * r;  // tanh(u)
  return (1.f - r) + 0.5f * x * (1.f - t * t) * kGeluC * (1.f + 3.f * 0.044715f * x * x);
}

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
