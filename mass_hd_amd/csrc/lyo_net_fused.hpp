// lyo_net_fused.hpp -- the surrogate corrector in one kernel (FP32 state, gfx950).
//
// A surrogate step runs the predictor and one corrector (the npc = 1
// one-kernel step), then the network predicts what the remaining correctors
// add to (u, Q, phi) and the increment is applied.  Here the network part --
// features (UpdateQuantities and the stress divergence, lyotropic.cpp:168-247,
// 268-299), the four 3x3 convolutions 8 -> 32 -> 32 -> 32 -> 8 with tanh-GELU,
// and the increment -- runs as one strip march: a wave owns 64 consecutive y
// columns (52 interior ones, a halo of six: two for the features, one per
// convolution) and walks along x over a chunk of rows.  Per head row h:
//   features   UpdateQuantities on row h-1, features of row h-2 (natural
//              layout: lane = column), written as bf16 into an LDS ring
//   layer 1    row h-3 from feature rows h-4..h-2 (LDS)
//   layer 2-4  rows h-4, h-5, h-6 from register rings of the layer below
//   increment  row h-6, applied to the state read again from memory
// The state is read once per row (plus the output row's second read) and the
// new state written once; features and activations never reach HBM.  The
// output goes to a second state buffer: neighbouring waves still read the
// input rows this wave has finished.
//
// Convolutions run on v_mfma_f32_32x32x16_bf16 (M = 32 output channels,
// N = 32 pixels, K = 16; lane l = 32g + j holds A[j][8g..8g+7],
// B[8g..8g+7][j] and D[8b + 4g + r][j] in accumulator 4b + r).  A wave's 64
// columns are two interleaved tiles: tile t holds pixels 2j + t, so a y shift
// by one pixel stays in the lane (the other tile) except at the tile seam,
// where one DPP wave shift supplies the neighbour lane's value.  The output
// D of one layer is the B operand of the next without any exchange: the K
// order of each 16-channel half h is the lane's own accumulator order
// (channel 8(2h + e/4) + 4g + e%4 at K = 8g + e), and the weights are packed
// to match (mass_lyo_net's fused fragments).  Layer 1 packs two taps into
// each K = 16 slice (K = 8g + e: tap 2s + g, feature e), five slices.  Layer
// 4 (8 outputs) folds its three y taps into M (row 8s + c: output c of y tap
// s), so it takes 6 MFMAs per tile and row instead of 18; the three partial
// outputs are added after one pixel shift each.
#pragma once

#include "lyo_net_act.hpp"
#include "lyo_net_engine.hpp"
#include "lyo_sweep.hpp"

namespace lyo {
namespace netf {

using net::bf16x8;
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int kHalo = 6;                  // 2 (features) + 4 (convolutions)
constexpr int kStrip = kWave - 2 * kHalo;  // 52 interior columns
constexpr int kWarm = 10;                  // head rows before the first output row
// first iteration whose features / layer 1 / 2 / 3 rows an output row needs
// (output row x0 needs layer-3 rows from x0-1, layer-2 rows from x0-2, ...)
constexpr int kFeatFrom = 2, kL1From = 4, kL2From = 6, kL3From = 8;
// weight fragments (64 lanes x 16 bytes each) per layer, in this order
constexpr int kFr1 = 0, kFr2 = 5, kFr3 = 23, kFr4 = 41, kFrags = 47;
constexpr int kNBias = 32 + 32 + 32 + 8;
constexpr int kFeatRow = 66;  // 16-byte entries of one feature row (pixels -1..64)
constexpr int kWaves = 4;     // waves per block
constexpr size_t kLdsW = (size_t)kFrags * 64 * 16;
constexpr size_t kLdsC = (size_t)(kNBias + 8) * 4;  // biases, output scales
constexpr size_t kLdsFeat = (size_t)3 * kFeatRow * 16;
constexpr size_t kLdsDelta = (size_t)64 * 8 * 4;
constexpr size_t kLdsBytes = kLdsW + kLdsC + kWaves * (kLdsFeat + kLdsDelta);

struct Args {
  const float* f_in;  // the one-corrector state S[cur] (f' 9 planes, Q, phi, phi compensation)
  const float* qxx_in;
  const float* qyx_in;
  const float* phi_in;
  const float* phic_in;
  float* f_out;  // the corrected state (another buffer set)
  float* qxx_out;
  float* qyx_out;
  float* phi_out;
  float* phic_out;
  const bf16x8* wfrag;  // kFrags x 64 fused fragments
  const float* bias;    // kNBias: layer biases in layer order
  const float* norm;       // mean[8], inv_std[8] of the inputs
  const float* out_scale;  // [8] physical increment per network output
  float* delta;         // optional [N][8]: the increments applied (tests)
  double* dphi_part;    // [nwaves]: per-wave sums of the applied phi increments
  double rho;
};

// feature ring entry of strip pixel p (-1..64): the two tiles' pixels lie in
// separate halves so that the lanes of a tile read consecutive entries;
// pixels -1 and 64 share the zeroed pad entry 33
__device__ __forceinline__ int fidx(int p) { return (p & 1) * 33 + (p >> 1) + 1; }

// bf16x8 of lane l-1 (wave_shr; lane 0 receives zeros) and of lane l+1
// (wave_shl; lane 63 receives zeros)
__device__ __forceinline__ bf16x8 from_lm(bf16x8 v) {
  i32x4 x = __builtin_bit_cast(i32x4, v);
#pragma unroll
  for (int c = 0; c < 4; ++c) x[c] = __builtin_amdgcn_mov_dpp(x[c], 0x138, 0xF, 0xF, true);
  return __builtin_bit_cast(bf16x8, x);
}
__device__ __forceinline__ bf16x8 from_lp(bf16x8 v) {
  i32x4 x = __builtin_bit_cast(i32x4, v);
#pragma unroll
  for (int c = 0; c < 4; ++c) x[c] = __builtin_amdgcn_mov_dpp(x[c], 0x130, 0xF, 0xF, true);
  return __builtin_bit_cast(bf16x8, x);
}

__device__ __forceinline__ f32x16 mfma(const bf16x8& a, const bf16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Biases of a 32-output layer in the accumulator layout (lane g: accumulator
// 4b + r is output channel 8b + 4g + r): the initial accumulator of the
// layer's first MFMA.
__device__ __forceinline__ f32x16 bias_acc(const float* __restrict__ bias, int gq) {
  f32x16 v;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const float4 q = *reinterpret_cast<const float4*>(bias + 8 * b + 4 * gq);
    v[4 * b] = q.x;
    v[4 * b + 1] = q.y;
    v[4 * b + 2] = q.z;
    v[4 * b + 3] = q.w;
  }
  return v;
}

// Epilogue of a 32-output layer: GELU, bf16, in the next layer's K order
// (half h of tile t: accumulators 8h..8h+7).
__device__ __forceinline__ void gelu_out(const f32x16 (&acc)[2], bf16x8 (&out)[2][2]) {
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int k = 0; k < 16; ++k) out[t][k >> 3][k & 7] = (__bf16)net::gelu(acc[t][k]);
}

// One 32 -> 32 layer on the rows of ring `in` (slots sm, s0, sp: rows
// out-1, out, out+1), fragments from `w`, accumulators starting at the
// biases bv.
__device__ __forceinline__ void conv32(const bf16x8 (&in)[3][2][2], int sm, int s0, int sp,
                                       const bf16x8* __restrict__ w, int lane, const f32x16& bv,
                                       f32x16 (&acc)[2]) {
#pragma unroll
  for (int dxi = 0; dxi < 3; ++dxi) {
    const int sl = dxi == 0 ? sm : (dxi == 1 ? s0 : sp);
    bf16x8 m1[2], p1[2];  // tile 0 at y-1 (lane j-1, tile 1); tile 1 at y+1 (lane j+1, tile 0)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      m1[h] = from_lm(in[sl][1][h]);
      p1[h] = from_lp(in[sl][0][h]);
    }
#pragma unroll
    for (int dyi = 0; dyi < 3; ++dyi) {
      const int tap = 3 * dxi + dyi;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const bf16x8 a = w[(2 * tap + h) * 64 + lane];
        const bf16x8 b0 = dyi == 0 ? m1[h] : in[sl][dyi - 1][h];
        const bf16x8 b1 = dyi == 2 ? p1[h] : in[sl][dyi][h];
        acc[0] = mfma(a, b0, tap == 0 && h == 0 ? bv : acc[0]);
        acc[1] = mfma(a, b1, tap == 0 && h == 0 ? bv : acc[1]);
      }
    }
  }
}

__device__ __forceinline__ void item(const Args& a, const Params<float>& P, const Geo& g,
                                     const int wid, const bf16x8* __restrict__ wl,
                                     const float* __restrict__ cl, bf16x8* __restrict__ fl,
                                     float* __restrict__ dl) {
  const int lane = threadIdx.x & 63;
  const int strip = wid % g.nstrips;
  const int chunk = wid / g.nstrips;
  const int LX = g.LX, LY = g.LY, N = g.NP;
  const int ycol = strip * kStrip - kHalo + lane;
  int yi = ycol % LY;
  if (yi < 0) yi += LY;
  LaneOff<float> ly{(unsigned)yi * 4u};
  const bool col_ok = lane >= kHalo && lane < kWave - kHalo && ycol < LY;
  const int gq = lane >> 5, jj = lane & 31;
  const int x0 = chunk * g.xc;
  const int x1 = min(x0 + g.xc, LX);
  // rows run from x0-6 to x1+8 (LX >= 16: one wrap in either direction)
  auto roff = [&](int r) {
    r = (r < 0) ? r + LX : r;
    r = (r >= LX) ? r - LX : r;
    return r * LY;
  };

  for (int e = lane; e < 3 * kFeatRow; e += 64) fl[e] = bf16x8{};

  // layer 1 operands: slice s, tile t reads the feature entry of the lane's
  // tap 2s + g (tap 9 has zero weights and reads tap 8's entry) -- its row
  // offset dx: -1 for s = 0 and for (s = 1, g = 0), 0 for (s = 1, g = 1) and
  // s = 2, +1 for s = 3, 4; its pixel 2j + c with c = t + dy, whose entry is
  // j + (c & 1) * 33 + (c >> 1) + 1
  auto l1pix = [&](int s, int t) {
    constexpr int kDy[2][5] = {{-1, 1, 0, -1, 1}, {0, -1, 1, 0, 1}};  // [g][s]
    const int c0 = t + kDy[0][s], c1 = t + kDy[1][s];
    return jj + (gq ? (c1 & 1) * 33 + (c1 >> 1) + 1 : (c0 & 1) * 33 + (c0 >> 1) + 1);
  };
  const float* b1 = cl;
  const float* b2 = cl + 32;
  const float* b3 = cl + 64;
  const float* b4 = cl + 96;
  const float* osc = cl + kNBias;

  // ---- natural-layout state: stencil rows h-2 (qa), h-1 (qb), h (qc) of Q
  // and phi, populations of row h-1 (fb), phi compensation of row h-2 (pca);
  // UpdateQuantities of rows h-3 (um) and h-2 (uc)
  auto load_qp = [&](int r) {
    const int o = roff(r);
    QP<float> q;
    q.xx = ly.ld(a.qxx_in + o);
    q.yx = ly.ld(a.qyx_in + o);
    q.p = ly.ld(a.phi_in + o);
    return q;
  };
  auto load_f = [&](int r, float(&f)[9]) {
    const int o = roff(r);
#pragma unroll
    for (int v = 0; v < 9; ++v) f[v] = ly.ld(a.f_in + (v * N + o));
  };
  const int xs = x0 - 4;  // head row of the first iteration
  QP<float> qa = load_qp(xs - 2), qb = load_qp(xs - 1), qc = load_qp(xs);
  float fb[9];
  load_f(xs - 1, fb);
  float pca = ly.ld(a.phic_in + roff(xs - 2));
  NbrVals<float> um{}, uc{};
  bf16x8 r2[3][2][2] = {}, r3[3][2][2] = {}, r4[3][2][2] = {};

  const int nI = x1 - x0 + kWarm;  // head rows x0-4 .. x1+5
  double dphi_acc = 0;             // the lane's applied phi increments

  // One head row; U = i % 3 (ring slots), i the iteration.
  auto row = [&](const int i, auto uu) {
    constexpr int U = decltype(uu)::value;
    const int h = xs + i;
    const int xo = h - 6;                      // output row
    const bool out_row = i >= kWarm && xo < x1;  // uniform

    // ---- features: UpdateQuantities on row h-1, features of row h-2
    NbrVals<float> un;
    {
      CenterVals<float> cv;
      update_quantities<float, false>(
          fb, qb.xx, qc.xx, qa.xx, ynb<float, false>(from_yp(qb.xx), from_ym(qb.xx)), qb.yx, qc.yx,
          qa.yx, ynb<float, false>(from_yp(qb.yx), from_ym(qb.yx)), qb.p, qc.p, qa.p,
          ynb<float, false>(from_yp(qb.p), from_ym(qb.p)), P, un, cv);
    }
    if (i >= kFeatFrom) {
      float o[kFeat];
      feature_values<float>(uc, un.sxx, um.sxx, un.syx, um.syx, from_yp(uc.sxy), from_ym(uc.sxy),
                            from_yp(uc.syy), from_ym(uc.syy), qa.xx, qa.yx,
                            (double)qa.p + (double)pca, o);
      bf16x8 v;
#pragma unroll
      for (int c = 0; c < kFeat; ++c) v[c] = (__bf16)((o[c] - a.norm[c]) * a.norm[8 + c]);
      fl[((U + 1) % 3) * kFeatRow + fidx(lane)] = v;
    }

    // ---- layer 1: row h-3 from feature rows h-4 (slot U+2), h-3 (U), h-2 (U+1)
    if (i >= kL1From) {
      const int Sm = ((U + 2) % 3) * kFeatRow, S0 = U * kFeatRow, Sp = ((U + 1) % 3) * kFeatRow;
      const f32x16 bv = bias_acc(b1, gq);
      f32x16 acc[2];
#pragma unroll
      for (int s = 0; s < 5; ++s) {
        const int so = s == 0 ? Sm : (s == 1 ? (gq ? S0 : Sm) : (s == 2 ? S0 : Sp));
        const bf16x8 w = wl[(kFr1 + s) * 64 + lane];
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[t] = mfma(w, fl[so + l1pix(s, t)], s == 0 ? bv : acc[t]);
      }
      gelu_out(acc, r2[U]);
    }
    // ---- layers 2 and 3: rows h-4, h-5; inputs rows out-1 (slot U+1),
    // out (U+2), out+1 (U, written this iteration)
    if (i >= kL2From) {
      f32x16 acc[2];
      conv32(r2, (U + 1) % 3, (U + 2) % 3, U, wl + kFr2 * 64, lane, bias_acc(b2, gq), acc);
      gelu_out(acc, r3[U]);
    }
    if (i >= kL3From) {
      f32x16 acc[2];
      conv32(r3, (U + 1) % 3, (U + 2) % 3, U, wl + kFr3 * 64, lane, bias_acc(b3, gq), acc);
      gelu_out(acc, r4[U]);
    }
    // next iteration's loads (Q/phi of row h+1, f of row h, phic of row
    // h-1) and the output row's state (its second read), issued after the
    // layers that hold the most registers; the block's other waves hide
    // their latency
    const QP<float> qn = load_qp(h + 1);
    float fn[9];
    load_f(h, fn);
    const float pcn = ly.ld(a.phic_in + roff(h - 1));
    // ---- layer 4: row h-6; M row 8s + c = output c of y tap s
    if (out_row) {
      float sf[9], sq[2], sp, spc;
      {
        const int o = roff(xo);
#pragma unroll
        for (int v = 0; v < 9; ++v) sf[v] = ly.ld(a.f_in + (v * N + o));
        sq[0] = ly.ld(a.qxx_in + o);
        sq[1] = ly.ld(a.qyx_in + o);
        sp = ly.ld(a.phi_in + o);
        spc = ly.ld(a.phic_in + o);
      }
      f32x16 acc[2];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.f;
#pragma unroll
      for (int dxi = 0; dxi < 3; ++dxi) {
        const int sl = dxi == 0 ? (U + 1) % 3 : (dxi == 1 ? (U + 2) % 3 : U);
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
          const bf16x8 w = wl[(kFr4 + 2 * dxi + hh) * 64 + lane];
#pragma unroll
          for (int t = 0; t < 2; ++t) acc[t] = mfma(w, r4[sl][t][hh], acc[t]);
        }
      }
      // out(p) = D_0(p - 1) + D_1(p) + D_2(p + 1) for output c = 4g + r
      const float4 bv = *reinterpret_cast<const float4*>(b4 + 4 * gq);
      const float4 sv = *reinterpret_cast<const float4*>(osc + 4 * gq);
      const float bb[4] = {bv.x, bv.y, bv.z, bv.w}, ss[4] = {sv.x, sv.y, sv.z, sv.w};
      float4 o0, o1;
      float* p0 = &o0.x;
      float* p1 = &o1.x;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        p0[r] = ((from_ym(acc[1][r]) + acc[0][4 + r]) + acc[1][8 + r] + bb[r]) * ss[r];
        p1[r] = ((acc[0][r] + acc[1][4 + r]) + from_yp(acc[0][8 + r]) + bb[r]) * ss[r];
      }
      *reinterpret_cast<float4*>(dl + (2 * jj) * 8 + 4 * gq) = o0;
      *reinterpret_cast<float4*>(dl + (2 * jj + 1) * 8 + 4 * gq) = o1;

      // ---- increment of row h-6 (natural layout: lane = pixel)
      const float4 dv = *reinterpret_cast<const float4*>(dl + lane * 8);
      const float d[5] = {dv.x, dv.y, dv.z, dv.w, dl[lane * 8 + 4]};
      apply_node<float, true>(sf, sq[0], sq[1], sp, &spc, d, a.rho);
      // the next row's loads have arrived: no wait after the divergent stores
      wait_vmem<false>();
      if (col_ok) {
        dphi_acc += (double)d[4];
        const int o = roff(xo);
#pragma unroll
        for (int v = 0; v < 9; ++v) ly.st(a.f_out + (v * N + o), sf[v]);
        ly.st(a.qxx_out + o, sq[0]);
        ly.st(a.qyx_out + o, sq[1]);
        ly.st(a.phi_out + o, sp);
        ly.st(a.phic_out + o, spc);
        if (a.delta) {
          float4* q = reinterpret_cast<float4*>(a.delta + ((size_t)o + yi) * 8);
          q[0] = dv;
          q[1] = *reinterpret_cast<const float4*>(dl + lane * 8 + 4);
        }
      }
    }
    um = uc;
    uc = un;
    qa = qb;
    qb = qc;
    qc = qn;
#pragma unroll
    for (int v = 0; v < 9; ++v) fb[v] = fn[v];
    pca = pcn;
  };
  // rows in threes: the ring slots are compile-time; past nI the rows are
  // computed (wrapped rows, in bounds) and not stored
  for (int i = 0; i < nI; i += 3) {
    row(i, std::integral_constant<int, 0>{});
    row(i + 1, std::integral_constant<int, 1>{});
    row(i + 2, std::integral_constant<int, 2>{});
  }
  dphi_acc = wave_sum(dphi_acc);
  if (lane == 0) a.dphi_part[wid] = dphi_acc;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
net_fused_kernel(const Args a, const Params<float> P, const Geo g) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  bf16x8* wl = reinterpret_cast<bf16x8*>(lds);
  float* cl = reinterpret_cast<float*>(lds + kLdsW);
  for (int e = threadIdx.x; e < kFrags * 64; e += blockDim.x) wl[e] = a.wfrag[e];
  for (int e = threadIdx.x; e < kNBias; e += blockDim.x) cl[e] = a.bias[e];
  if (threadIdx.x < 8) cl[kNBias + threadIdx.x] = a.out_scale[threadIdx.x];
  __syncthreads();
  int b = blockIdx.x;
  if (g.xcd) {
    const int nb8 = (int)gridDim.x >> 3;
    if (b < (nb8 << 3)) b = (b & 7) * nb8 + (b >> 3);
  }
  const int wv = (int)threadIdx.x >> 6;
  const int wid = __builtin_amdgcn_readfirstlane(b * kWaves + wv);
  if (wid >= g.nwaves) return;
  unsigned char* wave_lds = lds + kLdsW + kLdsC + (size_t)wv * (kLdsFeat + kLdsDelta);
  item(a, P, g, wid, wl, cl, reinterpret_cast<bf16x8*>(wave_lds),
       reinterpret_cast<float*>(wave_lds + kLdsFeat));
}

}  // namespace netf
}  // namespace lyo
