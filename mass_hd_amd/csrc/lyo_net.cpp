// lyo_net.cpp -- the surrogate corrector network: parameters, training and
// the surrogate step (C ABI in include/mass_lyo.h).
//
// No reference counterpart: the reference always runs npc corrector sweeps.
// The network predicts, from the one-corrector state, what the remaining
// correctors add to (u, Q, phi); a plain step every judge_every steps measures
// the velocity error and hands the run back to the plain corrector for the
// next judge_every steps whenever the error exceeds the tolerance.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "lyo_handle.hpp"
#include "lyo_net.hpp"
#include "lyo_net_internal.hpp"

using lyo::net::bf16x8;
using lyo::net::ConvArgs;

namespace {

constexpr int kCh[5] = {8, 32, 32, 32, 8};
// weight gradients: waves of the partial sums, and the largest layer's
// parameter count (9 * 32 * 32 + 32)
constexpr int kWgradWaves = 512;
constexpr int kWgradMaxPar = 9 * 32 * 32 + 32;
constexpr int kOut = 5;  // used output channels (du_x, du_y, dQxx, dQyx, dphi)

int layer_woff(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += 9 * kCh[i] * kCh[i + 1] + kCh[i + 1];
  return o;
}
int nparams_total() { return layer_woff(4); }
int frag_count(int cin, int cout) {  // bf16x8 elements
  return lyo::net::kslices(cin) * (cout / 16 > 0 ? cout / 16 : 1) * 64;
}

// fragments of layer weights W[t][cout][cin] (fwd) or of the data-gradient
// weights Wd[t][cin][cout] = W[8-t][cout][cin] (bwd)
__global__ void pack_frag_kernel(const float* __restrict__ W, int cin_f, int cout_f, int bwd,
                                 __bf16* __restrict__ out) {
  const int cin = bwd ? cout_f : cin_f, cout = bwd ? cin_f : cout_f;
  const int ks = lyo::net::kslices(cin), nblk = cout / 16 > 0 ? cout / 16 : 1;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ks * nblk * 64 * 8) return;
  const int j = e & 7, lane = (e >> 3) & 63, sn = e >> 9;
  const int nb = sn % nblk, s = sn / nblk;
  const int q = lane >> 4, i = lane & 15;
  // cout = 8: M rows 8..15 repeat channels 0..7 (see conv_rows_kernel)
  const int co = cout == 8 ? (i & 7) : nb * 16 + i;
  int t, ci;
  if (cin == 32) {
    t = s;
    ci = 8 * q + j;
  } else {
    t = 4 * s + q;
    ci = j;
  }
  float v = 0.f;
  if (t < 9 && co < cout) {
    // effective weight (t, out co, in ci)
    v = bwd ? W[((8 - t) * cout_f + ci) * cin_f + co] : W[(t * cout_f + co) * cin_f + ci];
  }
  out[e] = (__bf16)v;
}

// Weights of the one-kernel network step (lyo_net_fused.hpp), A operands of
// v_mfma_f32_32x32x16_bf16 (lane l = 32g + i holds A[i][8g..8g+7]):
//   fragments 0-4 (layer 1): slice s, K = 8g + e -> tap 2s + g, input e
//   fragments 5-40 (layers 2, 3): 2 tap + h, K = 8g + e -> input channel
//     8(2h + e/4) + 4g + e%4 (the order of the layer below's accumulators)
//   fragments 41-46 (layer 4): 2 dxi + h, M row 8s + c -> output c of tap
//     3 dxi + s (rows 24..31 zero), K as layers 2, 3
// then the 104 layer biases.  W_l = params + woff[l], W_l[t][cout][cin].
constexpr int kFusedFrags = 47;
constexpr int kFusedBias = 104;
__global__ void pack_fused_kernel(const float* __restrict__ params, int woff1, int woff2,
                                  int woff3, __bf16* __restrict__ frag, float* __restrict__ bias) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int nf = kFusedFrags * 64 * 8;
  if (e < nf) {
    const int fr = e >> 9, lane = (e >> 3) & 63, el = e & 7;
    const int g = lane >> 5, i = lane & 31;
    auto perm = [&](int h) { return 8 * (2 * h + (el >> 2)) + 4 * g + (el & 3); };
    float v = 0.f;
    if (fr < 5) {
      const int tap = 2 * fr + g;
      if (tap < 9) v = params[(tap * 32 + i) * 8 + el];
    } else if (fr < 41) {
      const int f = fr < 23 ? fr - 5 : fr - 23;
      const float* W = params + (fr < 23 ? woff1 : woff2);
      const int tap = f >> 1, h = f & 1;
      v = W[(tap * 32 + i) * 32 + perm(h)];
    } else {
      const int f = fr - 41, dxi = f >> 1, h = f & 1;
      if (i < 24) v = params[woff3 + ((3 * dxi + (i >> 3)) * 8 + (i & 7)) * 32 + perm(h)];
    }
    frag[e] = (__bf16)v;
  } else if (e < nf + kFusedBias) {
    const int b = e - nf;
    const int l = b < 32 ? 0 : (b < 64 ? 1 : (b < 96 ? 2 : 3));
    const int woff[4] = {0, woff1, woff2, woff3};
    const int cin = l == 0 ? 8 : 32, cout = l == 3 ? 8 : 32;
    bias[b] = params[woff[l] + 9 * cin * cout + (b - 32 * l)];
  }
}

__global__ void normalize_kernel(const float* __restrict__ feat, const float* __restrict__ mean,
                                 const float* __restrict__ inv, int n, __bf16* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * 8) return;
  const int c = e & 7;
  out[e] = (__bf16)((feat[e] - mean[c]) * inv[c]);
}

__global__ void scale_kernel(const float* __restrict__ in, const float* __restrict__ sc, int n,
                             float* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * 8) return;
  out[e] = in[e] * sc[e & 7];
}

// per-block sums of feat, feat^2 and tgt^2 per channel: part[b][24]
__global__ void __launch_bounds__(256)
stats_kernel(const float* __restrict__ feat, const float* __restrict__ tgt, int n,
             double* __restrict__ part) {
  __shared__ double sh[24][4];
  double a[24];
  for (int c = 0; c < 24; ++c) a[c] = 0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
    for (int c = 0; c < 8; ++c) {
      const double f = feat[(size_t)k * 8 + c], t = tgt[(size_t)k * 8 + c];
      a[c] += f;
      a[8 + c] += f * f;
      a[16 + c] += t * t;
    }
  }
  for (int c = 0; c < 24; ++c) {
    double v = a[c];
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 24)
    part[blockIdx.x * 24 + threadIdx.x] =
        (sh[threadIdx.x][0] + sh[threadIdx.x][1]) + (sh[threadIdx.x][2] + sh[threadIdx.x][3]);
}

thread_local std::string g_net_error;

}  // namespace

struct mass_lyo_net_s {
  int np = 0;
  std::vector<float> host_params;
  float *d_params = nullptr, *d_grads = nullptr, *d_m = nullptr, *d_v = nullptr;
  __bf16* d_wf[4] = {};  // forward fragments per layer
  __bf16* d_wb[4] = {};  // data-gradient fragments (layers 1..3 used)
  __bf16* d_wfused = nullptr;  // the one-kernel network step's fragments and biases
  float* d_fbias = nullptr;
  int adam_t = 0;
  // normalisation (host + device: mean[8], inv_std[8], out_scale[8], inv_out_scale[8])
  float norm[32];
  float* d_norm = nullptr;
  bool have_norm = false;
  // dataset: frames of features and targets (FP32 [N][8]) of one grid size
  int dLX = 0, dLY = 0;
  std::vector<float*> feats, tgts;
  double acc[24] = {};
  long long acc_n = 0;
  // workspace
  size_t wsN = 0;
  __bf16 *a0 = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr, *dz = nullptr, *dz2 = nullptr,
         *d4 = nullptr;
  float *z1 = nullptr, *z2 = nullptr, *z3 = nullptr, *out = nullptr, *y = nullptr;
  float* wpart = nullptr;  // per-wave partial weight gradients
  double* part = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
};

namespace {

int nfail(mass_lyo_net_t n, int code, const std::string& msg) {
  if (n)
    n->err = msg;
  else
    g_net_error = msg;
  return code;
}
int nhip(mass_lyo_net_t n, hipError_t e, const char* what) {
  return nfail(n, MASS_LYO_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void free_ws(mass_lyo_net_t n) {
  for (void* p : {(void*)n->a0, (void*)n->a1, (void*)n->a2, (void*)n->a3, (void*)n->dz,
                  (void*)n->dz2, (void*)n->d4, (void*)n->z1, (void*)n->z2, (void*)n->z3,
                  (void*)n->out, (void*)n->y, (void*)n->wpart})
    if (p) (void)hipFree(p);
  n->a0 = n->a1 = n->a2 = n->a3 = n->dz = n->dz2 = n->d4 = nullptr;
  n->z1 = n->z2 = n->z3 = n->out = n->y = n->wpart = nullptr;
  n->wsN = 0;
}

hipError_t ensure_ws(mass_lyo_net_t n, size_t N) {
  if (n->wsN >= N) return hipSuccess;
  free_ws(n);
  hipError_t e = hipSuccess;
  auto get = [&](size_t bytes, void** p) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  get(N * 8 * 2, (void**)&n->a0);
  get(N * 32 * 2, (void**)&n->a1);
  get(N * 32 * 2, (void**)&n->a2);
  get(N * 32 * 2, (void**)&n->a3);
  get(N * 32 * 2, (void**)&n->dz);
  get(N * 32 * 2, (void**)&n->dz2);
  get(N * 8 * 2, (void**)&n->d4);
  get(N * 32 * 4, (void**)&n->z1);
  get(N * 32 * 4, (void**)&n->z2);
  get(N * 32 * 4, (void**)&n->z3);
  get(N * 8 * 4, (void**)&n->out);
  get(N * 8 * 4, (void**)&n->y);
  get(sizeof(float) * (size_t)kWgradWaves * kWgradMaxPar, (void**)&n->wpart);
  if (e == hipSuccess) n->wsN = N;
  return e;
}

hipError_t pack_all(mass_lyo_net_t n) {
  // the weights are repacked on the network's own stream
  for (int l = 0; l < 4; ++l) {
    const int cin = kCh[l], cout = kCh[l + 1];
    const int nf = frag_count(cin, cout) * 8;
    hipLaunchKernelGGL(pack_frag_kernel, dim3((nf + 255) / 256), dim3(256), 0, n->stream,
                       n->d_params + layer_woff(l), cin, cout, 0, n->d_wf[l]);
    if (l > 0) {
      const int nb = frag_count(cout, cin) * 8;
      hipLaunchKernelGGL(pack_frag_kernel, dim3((nb + 255) / 256), dim3(256), 0, n->stream,
                         n->d_params + layer_woff(l), cin, cout, 1, n->d_wb[l]);
    }
  }
  const int nfu = kFusedFrags * 64 * 8 + kFusedBias;
  hipLaunchKernelGGL(pack_fused_kernel, dim3((nfu + 255) / 256), dim3(256), 0, n->stream,
                     n->d_params, layer_woff(1), layer_woff(2), layer_woff(3), n->d_wfused,
                     n->d_fbias);
  return hipGetLastError();
}

// output rows per conv task: Cin = 32 streams ROWS + 2 input rows once per
// task; Cin = 8 (raw FP32 fragments in layer 0) keeps fewer rows in flight
constexpr int kRows32 = 8;
constexpr int kRows8 = 4;

int conv_waves(int LX, int LY, int rows) {
  const int tasks = ((LX + rows - 1) / rows) * (LY / 16);
  return tasks < 8192 ? tasks : 8192;
}

template <int CIN, int COUT, int ROWS, bool F32, int EPI, bool ZOUT>
void launch_conv(const lyo::net::ConvArgs& a, hipStream_t s) {
  using namespace lyo::net;
  const dim3 grid(conv_waves(a.LX, a.LY, ROWS)), block(64);
  if (a.LX % ROWS)
    hipLaunchKernelGGL((conv_rows_kernel<CIN, COUT, ROWS, F32, true, EPI, ZOUT>), grid, block, 0,
                       s, a);
  else
    hipLaunchKernelGGL((conv_rows_kernel<CIN, COUT, ROWS, F32, false, EPI, ZOUT>), grid, block,
                       0, s, a);
}

// One convolution launch.  in_f32 (layer 0 of inference): FP32 features,
// normalised on load with the network's mean / inverse deviation; out_scale
// (last layer of inference): output channels scaled to physical increments.
hipError_t conv(mass_lyo_net_t n, const __bf16* in, int l, bool bwd, int epi, float* zout,
                const float* zin, __bf16* out_bf, float* out_f32, int LX, int LY, hipStream_t s,
                const float* in_f32 = nullptr, const float* out_scale = nullptr) {
  using namespace lyo::net;
  ConvArgs a;
  a.in = in;
  a.in_f32 = in_f32;
  a.in_norm = in_f32 ? n->d_norm : nullptr;
  a.wfrag = reinterpret_cast<const bf16x8*>(bwd ? n->d_wb[l] : n->d_wf[l]);
  a.bias = bwd ? nullptr : n->d_params + layer_woff(l) + 9 * kCh[l] * kCh[l + 1];
  a.zout = zout;
  a.zin = zin;
  a.out_bf = out_bf;
  a.out_f32 = out_f32;
  a.out_scale = out_scale;
  a.cin = bwd ? kCh[l + 1] : kCh[l];
  a.cout = bwd ? kCh[l] : kCh[l + 1];
  a.LX = LX;
  a.LY = LY;
  // the layer shapes and epilogues of the network (forward: 8->32, 32->32
  // with GELU, 32->8 linear; data gradient: 32->32 and 8->32)
  const bool z = zout != nullptr;
  if (a.cin == 32 && a.cout == 32 && epi == EPI_GELU && !z && out_bf)
    launch_conv<32, 32, kRows32, false, EPI_GELU, false>(a, s);
  else if (a.cin == 32 && a.cout == 32 && epi == EPI_GELU && z && out_bf)
    launch_conv<32, 32, kRows32, false, EPI_GELU, true>(a, s);
  else if (a.cin == 32 && a.cout == 32 && epi == EPI_GELU_GRAD && zin && out_bf)
    launch_conv<32, 32, kRows32, false, EPI_GELU_GRAD, false>(a, s);
  else if (a.cin == 32 && a.cout == 8 && epi == EPI_LINEAR && out_f32)
    launch_conv<32, 8, kRows32, false, EPI_LINEAR, false>(a, s);
  else if (a.cin == 8 && a.cout == 32 && epi == EPI_GELU && in_f32 && !z && out_bf)
    launch_conv<8, 32, kRows8, true, EPI_GELU, false>(a, s);
  else if (a.cin == 8 && a.cout == 32 && epi == EPI_GELU && !in_f32 && !z && out_bf)
    launch_conv<8, 32, kRows8, false, EPI_GELU, false>(a, s);
  else if (a.cin == 8 && a.cout == 32 && epi == EPI_GELU && !in_f32 && z && out_bf)
    launch_conv<8, 32, kRows8, false, EPI_GELU, true>(a, s);
  else if (a.cin == 8 && a.cout == 32 && epi == EPI_GELU_GRAD && zin && out_bf)
    launch_conv<8, 32, kRows8, false, EPI_GELU_GRAD, false>(a, s);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// forward of normalised bf16 input a0 (or, inference, of FP32 features feat
// normalised on load) -> out (FP32 [N][8]; normalised units unless out_scale)
hipError_t forward(mass_lyo_net_t n, int LX, int LY, bool keep, hipStream_t s,
                   const float* feat = nullptr, float* out = nullptr,
                   const float* out_scale = nullptr) {
  using namespace lyo::net;
  hipError_t e = conv(n, n->a0, 0, false, EPI_GELU, keep ? n->z1 : nullptr, nullptr, n->a1,
                      nullptr, LX, LY, s, feat);
  if (e == hipSuccess) e = conv(n, n->a1, 1, false, EPI_GELU, keep ? n->z2 : nullptr, nullptr, n->a2, nullptr, LX, LY, s);
  if (e == hipSuccess) e = conv(n, n->a2, 2, false, EPI_GELU, keep ? n->z3 : nullptr, nullptr, n->a3, nullptr, LX, LY, s);
  if (e == hipSuccess)
    e = conv(n, n->a3, 3, false, EPI_LINEAR, nullptr, nullptr, nullptr, out ? out : n->out, LX,
             LY, s, nullptr, out_scale);
  return e;
}

// loss and gradients for targets y (normalised, FP32 [N][8]) into d_grads
hipError_t backward(mass_lyo_net_t n, int LX, int LY, int nparts) {
  using namespace lyo::net;
  const int N = LX * LY;
  hipError_t e = hipMemsetAsync(n->d_grads, 0, sizeof(float) * n->np, n->stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(loss_grad_kernel, dim3(nparts), dim3(256), 0, n->stream, n->out, n->y, N,
                     n->d4, n->part);
  const int segs = LX * (LY / 16);
  const int nw = segs < kWgradWaves ? segs : kWgradWaves;
  const __bf16* ins[4] = {n->a0, n->a1, n->a2, n->a3};
  const float* zs[4] = {nullptr, n->z1, n->z2, n->z3};
  const __bf16* d = n->d4;
  __bf16* bufs[2] = {n->dz, n->dz2};
  for (int l = 3; l >= 0 && e == hipSuccess; --l) {
    // weights and bias of layer l: per-wave partial sums, then their sum
    const int npar = 9 * kCh[l] * kCh[l + 1] + kCh[l + 1];
    if (l == 0)
      hipLaunchKernelGGL((conv_wgrad_mfma_kernel<8, 32>), dim3(nw), dim3(64), 0, n->stream,
                         ins[l], d, LX, LY, n->wpart);
    else if (l == 3)
      hipLaunchKernelGGL((conv_wgrad_mfma_kernel<32, 8>), dim3(nw), dim3(64), 0, n->stream,
                         ins[l], d, LX, LY, n->wpart);
    else
      hipLaunchKernelGGL((conv_wgrad_mfma_kernel<32, 32>), dim3(nw), dim3(64), 0, n->stream,
                         ins[l], d, LX, LY, n->wpart);
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((npar + 31) / 32), dim3(256), 0, n->stream,
                       n->wpart, nw, npar, n->d_grads + layer_woff(l));
    e = hipGetLastError();
    if (l > 0 && e == hipSuccess) {
      __bf16* dn = bufs[l & 1];
      e = conv(n, d, l, true, EPI_GELU_GRAD, nullptr, zs[l], dn, nullptr, LX, LY, n->stream);
      d = dn;
    }
  }
  return e;
}

}  // namespace

// ------------------------------------------------------------ internal API

namespace lyo {

hipError_t net_predict(mass_lyo_net_t n, const float* feat, int LX, int LY, float* delta,
                       hipStream_t s) {
  const int N = LX * LY;
  hipError_t e = ensure_ws(n, (size_t)N);
  if (e != hipSuccess) return e;
  // normalisation folded into layer 0's loads, output scaling into layer 3
  return forward(n, LX, LY, false, s, feat, delta, n->d_norm + 16);
}

hipError_t net_input(mass_lyo_net_t n, size_t N, __bf16** in, const float** norm) {
  const hipError_t e = ensure_ws(n, N);
  if (e != hipSuccess) return e;
  *in = n->a0;
  *norm = n->d_norm;
  return hipSuccess;
}

hipError_t net_predict_input(mass_lyo_net_t n, int LX, int LY, float* delta, hipStream_t s,
                             const float* out_scale) {
  if ((size_t)LX * LY > n->wsN) return hipErrorInvalidValue;
  return forward(n, LX, LY, false, s, nullptr, delta, out_scale ? out_scale : n->d_norm + 16);
}

void net_out_scale(mass_lyo_net_t n, float* out8) {
  for (int c = 0; c < 8; ++c) out8[c] = n->norm[16 + c];
}

bool net_ready(mass_lyo_net_t n) { return n && n->have_norm; }

NetFusedW net_fused_weights(mass_lyo_net_t n, const float* out_scale) {
  NetFusedW w;
  w.wfrag = n->d_wfused;
  w.bias = n->d_fbias;
  w.norm = n->d_norm;
  w.out_scale = out_scale ? out_scale : n->d_norm + 16;
  return w;
}

}  // namespace lyo

// ------------------------------------------------------------------ C ABI

extern "C" {

const char* mass_lyo_net_last_error(mass_lyo_net_t n) {
  return n ? n->err.c_str() : g_net_error.c_str();
}

int mass_lyo_net_create(uint32_t seed, mass_lyo_net_t* out) {
  if (!out) return nfail(nullptr, MASS_LYO_EINVAL, "null output");
  auto* n = new mass_lyo_net_s();
  n->np = nparams_total();
  n->host_params.assign(n->np, 0.f);
  std::mt19937 gen(seed);
  for (int l = 0; l < 4; ++l) {
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * kCh[l])));
    float* W = n->host_params.data() + layer_woff(l);
    for (int i = 0; i < 9 * kCh[l] * kCh[l + 1]; ++i) W[i] = (l == 3 ? 0.1f : 1.f) * nd(gen);
  }
  for (int c = 0; c < 8; ++c) {
    n->norm[c] = 0.f;
    n->norm[8 + c] = 1.f;
    n->norm[16 + c] = 1.f;
    n->norm[24 + c] = 1.f;
  }
  hipError_t e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
  auto get = [&](size_t bytes, void** p) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
    if (e == hipSuccess) e = hipMemset(*p, 0, bytes);
  };
  get(sizeof(float) * n->np, (void**)&n->d_params);
  get(sizeof(float) * n->np, (void**)&n->d_grads);
  get(sizeof(float) * n->np, (void**)&n->d_m);
  get(sizeof(float) * n->np, (void**)&n->d_v);
  get(sizeof(float) * 32, (void**)&n->d_norm);
  get(sizeof(double) * 24 * 4096, (void**)&n->part);
  for (int l = 0; l < 4; ++l) {
    get((size_t)frag_count(kCh[l], kCh[l + 1]) * 16, (void**)&n->d_wf[l]);
    get((size_t)frag_count(kCh[l + 1], kCh[l]) * 16, (void**)&n->d_wb[l]);
  }
  get((size_t)kFusedFrags * 64 * 16, (void**)&n->d_wfused);
  get(sizeof(float) * kFusedBias, (void**)&n->d_fbias);
  if (e == hipSuccess)
    e = hipMemcpy(n->d_params, n->host_params.data(), sizeof(float) * n->np, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(n->d_norm, n->norm, sizeof(n->norm), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = pack_all(n);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  if (e != hipSuccess) {
    const std::string msg = std::string("network set-up failed: ") + hipGetErrorString(e);
    mass_lyo_net_destroy(n);
    return nfail(nullptr, MASS_LYO_EDEVICE, msg);
  }
  *out = n;
  return MASS_LYO_OK;
}

void mass_lyo_net_destroy(mass_lyo_net_t n) {
  if (!n) return;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  free_ws(n);
  for (float* p : n->feats) (void)hipFree(p);
  for (float* p : n->tgts) (void)hipFree(p);
  for (void* p : {(void*)n->d_params, (void*)n->d_grads, (void*)n->d_m, (void*)n->d_v,
                  (void*)n->d_norm, (void*)n->part, (void*)n->d_wfused, (void*)n->d_fbias})
    if (p) (void)hipFree(p);
  for (int l = 0; l < 4; ++l) {
    if (n->d_wf[l]) (void)hipFree(n->d_wf[l]);
    if (n->d_wb[l]) (void)hipFree(n->d_wb[l]);
  }
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

int mass_lyo_net_num_params(mass_lyo_net_t n) { return n ? n->np : 0; }

int mass_lyo_net_get_params(mass_lyo_net_t n, float* out) {
  if (!n || !out) return nfail(n, MASS_LYO_EINVAL, "null argument");
  hipError_t e = hipMemcpy(out, n->d_params, sizeof(float) * n->np, hipMemcpyDeviceToHost);
  return e == hipSuccess ? MASS_LYO_OK : nhip(n, e, "get params");
}

int mass_lyo_net_set_params(mass_lyo_net_t n, const float* in) {
  if (!n || !in) return nfail(n, MASS_LYO_EINVAL, "null argument");
  hipError_t e = hipMemcpy(n->d_params, in, sizeof(float) * n->np, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = pack_all(n);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  return e == hipSuccess ? MASS_LYO_OK : nhip(n, e, "set params");
}

int mass_lyo_net_normalization(mass_lyo_net_t n, float* out32) {
  if (!n || !out32) return nfail(n, MASS_LYO_EINVAL, "null argument");
  std::memcpy(out32, n->norm, sizeof(n->norm));
  return MASS_LYO_OK;
}

int mass_lyo_net_forward_host(mass_lyo_net_t n, const float* x, int32_t LX, int32_t LY, float* y) {
  if (!n || !x || !y) return nfail(n, MASS_LYO_EINVAL, "null argument");
  if (LX < 1 || LY < 16 || LY % 16) return nfail(n, MASS_LYO_EINVAL, "LY must be a multiple of 16");
  const size_t N = (size_t)LX * LY;
  hipError_t e = ensure_ws(n, N);
  std::vector<__bf16> xb(N * 8);
  for (size_t i = 0; i < N * 8; ++i) xb[i] = (__bf16)x[i];
  if (e == hipSuccess) e = hipMemcpyAsync(n->a0, xb.data(), N * 16, hipMemcpyHostToDevice, n->stream);
  if (e == hipSuccess) e = forward(n, LX, LY, false, n->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(y, n->out, N * 32, hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  return e == hipSuccess ? MASS_LYO_OK : nhip(n, e, "forward");
}

int mass_lyo_net_predict_host(mass_lyo_net_t n, const float* x, int32_t LX, int32_t LY, float* y) {
  if (!n || !x || !y) return nfail(n, MASS_LYO_EINVAL, "null argument");
  if (LX < 1 || LY < 16 || LY % 16) return nfail(n, MASS_LYO_EINVAL, "LY must be a multiple of 16");
  const size_t N = (size_t)LX * LY;
  float* d = nullptr;
  hipError_t e = hipMalloc((void**)&d, N * 64);
  if (e == hipSuccess) e = hipMemcpyAsync(d, x, N * 32, hipMemcpyHostToDevice, n->stream);
  if (e == hipSuccess) e = lyo::net_predict(n, d, LX, LY, d + N * 8, n->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(y, d + N * 8, N * 32, hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  if (d) (void)hipFree(d);
  return e == hipSuccess ? MASS_LYO_OK : nhip(n, e, "predict");
}

int mass_lyo_net_grad_host(mass_lyo_net_t n, const float* x, const float* y, int32_t LX, int32_t LY,
                           double* loss, float* grads) {
  if (!n || !x || !y) return nfail(n, MASS_LYO_EINVAL, "null argument");
  if (LX < 1 || LY < 16 || LY % 16) return nfail(n, MASS_LYO_EINVAL, "LY must be a multiple of 16");
  const size_t N = (size_t)LX * LY;
  hipError_t e = ensure_ws(n, N);
  std::vector<__bf16> xb(N * 8);
  for (size_t i = 0; i < N * 8; ++i) xb[i] = (__bf16)x[i];
  const int nparts = 256;
  std::vector<double> part(nparts);
  if (e == hipSuccess) e = hipMemcpyAsync(n->a0, xb.data(), N * 16, hipMemcpyHostToDevice, n->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(n->y, y, N * 32, hipMemcpyHostToDevice, n->stream);
  if (e == hipSuccess) e = forward(n, LX, LY, true, n->stream);
  if (e == hipSuccess) e = backward(n, LX, LY, nparts);
  if (e == hipSuccess && grads)
    e = hipMemcpyAsync(grads, n->d_grads, sizeof(float) * n->np, hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(part.data(), n->part, sizeof(double) * nparts, hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  if (e != hipSuccess) return nhip(n, e, "gradient");
  double L = 0;
  for (double p : part) L += p;
  if (loss) *loss = L;
  return MASS_LYO_OK;
}

int mass_lyo_net_collect(mass_lyo_net_t n, mass_lyo_t h) {
  if (!n || !h) return nfail(n, MASS_LYO_EINVAL, "null argument");
  if (h->prm.npc < 2) return nfail(n, MASS_LYO_EINVAL, "collecting needs a run with npc >= 2");
  if (h->prm.bc != 0 || h->growth)
    return nfail(n, MASS_LYO_EINVAL, "collecting needs a plain periodic run (bc 0)");
  if (h->prm.LY % 16) return nfail(n, MASS_LYO_EINVAL, "LY must be a multiple of 16");
  if (!n->feats.empty() && (h->prm.LX != n->dLX || h->prm.LY != n->dLY))
    return nfail(n, MASS_LYO_EINVAL, "all frames of a dataset have one grid size");
  const size_t N = (size_t)h->N;
  float *f = nullptr, *t = nullptr;
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMalloc((void**)&f, N * 32);
  if (e == hipSuccess) e = hipMalloc((void**)&t, N * 32);
  // the one-corrector state of the last (plain) step and its increments
  if (e == hipSuccess) e = h->eng->net_features(1, f, nullptr, nullptr, h->stream);
  if (e == hipSuccess) e = h->eng->net_targets(t, h->stream);
  const int nb = 1024;
  std::vector<double> part(24 * nb);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(stats_kernel, dim3(nb), dim3(256), 0, h->stream, f, t, (int)N, n->part);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(part.data(), n->part, sizeof(double) * 24 * nb, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    if (f) (void)hipFree(f);
    if (t) (void)hipFree(t);
    return nhip(n, e, "collect");
  }
  for (int b = 0; b < nb; ++b)
    for (int c = 0; c < 24; ++c) n->acc[c] += part[b * 24 + c];
  n->acc_n += (long long)N;
  n->feats.push_back(f);
  n->tgts.push_back(t);
  n->dLX = h->prm.LX;
  n->dLY = h->prm.LY;
  return MASS_LYO_OK;
}

int mass_lyo_net_dataset_size(mass_lyo_net_t n) { return n ? (int)n->feats.size() : 0; }

int mass_lyo_net_train(mass_lyo_net_t n, int32_t iters, float lr, double* loss_first,
                       double* loss_last) {
  if (!n) return nfail(nullptr, MASS_LYO_EINVAL, "null network");
  if (n->feats.empty()) return nfail(n, MASS_LYO_EINVAL, "empty dataset: collect frames first");
  const int LX = n->dLX, LY = n->dLY, N = LX * LY;
  // normalisation from the whole dataset
  for (int c = 0; c < 8; ++c) {
    const double m = n->acc[c] / n->acc_n;
    const double var = n->acc[8 + c] / n->acc_n - m * m;
    const double sd = var > 1e-30 ? std::sqrt(var) : 1.0;
    const double rms = std::sqrt(n->acc[16 + c] / n->acc_n);
    n->norm[c] = (float)m;
    n->norm[8 + c] = (float)(1.0 / sd);
    n->norm[16 + c] = (float)(c < kOut && rms > 0 ? rms : 1.0);
    n->norm[24 + c] = 1.f / n->norm[16 + c];
  }
  n->have_norm = true;
  hipError_t e = hipMemcpyAsync(n->d_norm, n->norm, sizeof(n->norm), hipMemcpyHostToDevice, n->stream);
  if (e == hipSuccess) e = ensure_ws(n, (size_t)N);
  const int nparts = 256;
  std::vector<double> part(nparts);
  double first = 0, last = 0;
  for (int it = 0; it < iters && e == hipSuccess; ++it) {
    const size_t fr = (size_t)it % n->feats.size();
    hipLaunchKernelGGL(normalize_kernel, dim3((N * 8 + 255) / 256), dim3(256), 0, n->stream,
                       n->feats[fr], n->d_norm, n->d_norm + 8, N, n->a0);
    hipLaunchKernelGGL(scale_kernel, dim3((N * 8 + 255) / 256), dim3(256), 0, n->stream,
                       n->tgts[fr], n->d_norm + 24, N, n->y);
    e = forward(n, LX, LY, true, n->stream);
    if (e == hipSuccess) e = backward(n, LX, LY, nparts);
    if (e == hipSuccess) {
      ++n->adam_t;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)n->adam_t), c2 = 1.f - std::pow(b2, (float)n->adam_t);
      hipLaunchKernelGGL(lyo::net::adam_kernel, dim3((n->np + 255) / 256), dim3(256), 0, n->stream,
                         n->d_params, n->d_grads, n->d_m, n->d_v, n->np, lr, b1, b2, 1e-8f, c1, c2);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = pack_all(n);
    if (e == hipSuccess && (it == 0 || it == iters - 1)) {
      e = hipMemcpyAsync(part.data(), n->part, sizeof(double) * nparts, hipMemcpyDeviceToHost, n->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
      double L = 0;
      for (double p : part) L += p;
      if (it == 0) first = L;
      last = L;
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  if (e != hipSuccess) return nhip(n, e, "train");
  if (loss_first) *loss_first = first;
  if (loss_last) *loss_last = last;
  return MASS_LYO_OK;
}

}  // extern "C"
