// lyo_net_engine.hpp -- state <-> network kernels of the surrogate corrector.
//
// The surrogate replaces the corrector sweeps after the first (npc > 1): a
// step is the predictor and one corrector, then the network predicts what the
// remaining correctors would add to (u, Q, phi) from the one-corrector state,
// and the increment is applied.  Training pairs come for free from plain
// npc = 2 steps: the first corrector's output (buffer set C[0]) is the
// one-corrector state, the step's result the two-corrector one.
//
// Features per node (8 channels): ux, uy, Qxx, Qyx, phi, mu, and the force
// F = div(sigma) (x and y), all from UpdateQuantities
// (src/models/lyotropic.cpp:168-247) on the node and its four neighbours.
// Targets (5 channels, padded to 8): the increments of ux, uy, Qxx, Qyx, phi.
#pragma once

#include "lyo_aux.hpp"

namespace lyo {

constexpr int kFeat = 8;

template <class T, bool EXACT>
__device__ void uq_at(const T* f, const T* qxx, const T* qyx, const T* phi, const T* dphi, int NP,
                      int LX, int LY, int x, int y, const Params<T>& P, NbrVals<T>& nv,
                      CenterVals<T>& cv) {
  const int k = x * LY + y;
  const int kxp = (x + 1 == LX ? 0 : x + 1) * LY + y;
  const int kxm = (x == 0 ? LX - 1 : x - 1) * LY + y;
  const int kyp = x * LY + (y + 1 == LY ? 0 : y + 1);
  const int kym = x * LY + (y == 0 ? LY - 1 : y - 1);
  auto ph = [&](int i) { return dphi ? phi[i] + dphi[i] : phi[i]; };
  T fl[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) fl[v] = f[v * NP + k];
  update_quantities<T, EXACT>(fl, qxx[k], qxx[kxp], qxx[kxm], ynb<T, EXACT>(qxx[kyp], qxx[kym]),
                              qyx[k], qyx[kxp], qyx[kxm], ynb<T, EXACT>(qyx[kyp], qyx[kym]),
                              ph(k), ph(kxp), ph(kxm), ynb<T, EXACT>(ph(kyp), ph(kym)), P, nv,
                              cv);
}

// Features of every node: FP32 [N][8] (feat), or -- surrogate inference --
// normalised bf16 [N][8] (feat_bf, (x - mean[c]) * inv_std[c] with norm =
// mean[8], inv_std[8]), the network's first layer input.
template <class T, bool EXACT>
__global__ void features_kernel(const T* __restrict__ f, const T* __restrict__ qxx,
                                const T* __restrict__ qyx, const T* __restrict__ phi,
                                const T* __restrict__ dphi, const T* __restrict__ phic,
                                Params<T> P, int LX, int LY, int NP, float* __restrict__ feat,
                                const float* __restrict__ norm, __bf16* __restrict__ feat_bf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= LX * LY) return;
  const int x = k / LY, y = k - x * LY;
  NbrVals<T> c, xp, xm, yp, ym;
  CenterVals<T> cv;
  uq_at<T, EXACT>(f, qxx, qyx, phi, dphi, NP, LX, LY, x, y, P, c, cv);
  uq_at<T, EXACT>(f, qxx, qyx, phi, dphi, NP, LX, LY, x + 1 == LX ? 0 : x + 1, y, P, xp, cv);
  uq_at<T, EXACT>(f, qxx, qyx, phi, dphi, NP, LX, LY, x == 0 ? LX - 1 : x - 1, y, P, xm, cv);
  uq_at<T, EXACT>(f, qxx, qyx, phi, dphi, NP, LX, LY, x, y + 1 == LY ? 0 : y + 1, P, yp, cv);
  uq_at<T, EXACT>(f, qxx, qyx, phi, dphi, NP, LX, LY, x, y == 0 ? LY - 1 : y - 1, P, ym, cv);
  double p = (double)phi[k];
  if (dphi) p += (double)dphi[k];
  if (phic) p += (double)phic[k];
  float o[kFeat];
  o[0] = (float)c.ux;
  o[1] = (float)c.uy;
  o[2] = (float)qxx[k];
  o[3] = (float)qyx[k];
  o[4] = (float)p;
  o[5] = (float)c.mu;
  o[6] = (float)(T(.5) * (xp.sxx - xm.sxx) + T(.5) * (yp.sxy - ym.sxy));
  o[7] = (float)(T(.5) * (xp.syx - xm.syx) + T(.5) * (yp.syy - ym.syy));
  if (feat_bf) {
    typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
    bf16x8_t v;
#pragma unroll
    for (int j = 0; j < kFeat; ++j) v[j] = (__bf16)((o[j] - norm[j]) * norm[8 + j]);
    *reinterpret_cast<bf16x8_t*>(feat_bf + (size_t)k * kFeat) = v;
  } else {
    float* q = feat + (size_t)k * kFeat;
#pragma unroll
    for (int j = 0; j < kFeat; ++j) q[j] = o[j];
  }
}

template <class T, bool SHIFT>
__device__ __forceinline__ void velocity_at(const T* f, int NP, int k, double rho, double& n,
                                            double& ux, double& uy) {
  double fv[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) fv[v] = (double)f[v * NP + k] + (SHIFT ? rest_weight(v) * rho : 0.0);
  n = fv[0] + fv[1] + fv[2] + fv[3] + fv[4] + fv[5] + fv[6] + fv[7] + fv[8];
  ux = (fv[1] - fv[2] + fv[5] - fv[6] - fv[7] + fv[8]) / n;
  uy = (fv[3] - fv[4] + fv[5] - fv[6] + fv[7] - fv[8]) / n;
}

// increments (two-corrector state) - (one-corrector state), [N][8]
template <class T, bool SHIFT>
__global__ void targets_kernel(const T* __restrict__ f2, const T* __restrict__ q2xx,
                               const T* __restrict__ q2yx, const T* __restrict__ phi2,
                               const T* __restrict__ phic2, const T* __restrict__ f1,
                               const T* __restrict__ q1xx, const T* __restrict__ q1yx,
                               const T* __restrict__ phi1, const T* __restrict__ dphi1, int N,
                               int NP, double rho, float* __restrict__ tgt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  double n2, ux2, uy2, n1, ux1, uy1;
  velocity_at<T, SHIFT>(f2, NP, k, rho, n2, ux2, uy2);
  velocity_at<T, SHIFT>(f1, NP, k, rho, n1, ux1, uy1);
  double p2 = (double)phi2[k] + (phic2 ? (double)phic2[k] : 0.0);
  double p1 = (double)phi1[k] + (dphi1 ? (double)dphi1[k] : 0.0);
  float* o = tgt + (size_t)k * kFeat;
  o[0] = (float)(ux2 - ux1);
  o[1] = (float)(uy2 - uy1);
  o[2] = (float)((double)q2xx[k] - (double)q1xx[k]);
  o[3] = (float)((double)q2yx[k] - (double)q1yx[k]);
  o[4] = (float)(p2 - p1);
  o[5] = o[6] = o[7] = 0.f;
}

// Applies (du, dQ, dphi) to a state in place: Q += dQ, phi += dphi
// (compensated), and f += fe(n, u + du) - fe(n, u), which keeps the mass of
// every node and changes its momentum by n du.
template <class T, bool SHIFT>
__global__ void apply_increment_kernel(T* __restrict__ f, T* __restrict__ qxx, T* __restrict__ qyx,
                                       T* __restrict__ phi, T* __restrict__ phic,
                                       const float* __restrict__ delta, int N, int NP,
                                       double rho) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const float* d = delta + (size_t)k * kFeat;
  double n, ux, uy;
  velocity_at<T, SHIFT>(f, NP, k, rho, n, ux, uy);
  auto feq = [&](double vx, double vy, double* fe) {
    const double v2 = vx * vx + vy * vy;
    fe[1] = n * (1. / 9. + 1. / 3. * vx + 1. / 3. * v2 - 1. / 2. * vy * vy);
    fe[2] = n * (1. / 9. - 1. / 3. * vx + 1. / 3. * v2 - 1. / 2. * vy * vy);
    fe[3] = n * (1. / 9. + 1. / 3. * vy - 1. / 2. * vx * vx + 1. / 3. * v2);
    fe[4] = n * (1. / 9. - 1. / 3. * vy - 1. / 2. * vx * vx + 1. / 3. * v2);
    fe[5] = n * (1. / 36. + 1. / 12. * vx + 1. / 12. * vy + 1. / 12. * v2 + 1. / 4. * vx * vy);
    fe[6] = n * (1. / 36. - 1. / 12. * vx - 1. / 12. * vy + 1. / 12. * v2 + 1. / 4. * vx * vy);
    fe[7] = n * (1. / 36. - 1. / 12. * vx + 1. / 12. * vy + 1. / 12. * v2 - 1. / 4. * vx * vy);
    fe[8] = n * (1. / 36. + 1. / 12. * vx - 1. / 12. * vy + 1. / 12. * v2 - 1. / 4. * vx * vy);
    fe[0] = n - fe[1] - fe[2] - fe[3] - fe[4] - fe[5] - fe[6] - fe[7] - fe[8];
  };
  double a[9], b[9];
  feq(ux, uy, a);
  feq(ux + (double)d[0], uy + (double)d[1], b);
#pragma unroll
  for (int v = 0; v < 9; ++v) f[v * NP + k] = (T)((double)f[v * NP + k] + (b[v] - a[v]));
  qxx[k] = (T)((double)qxx[k] + (double)d[2]);
  qyx[k] = (T)((double)qyx[k] + (double)d[3]);
  if (phic) {
    const T yv = (T)d[4] + phic[k];
    const T p0 = phi[k];
    const T t = p0 + yv;
    const T bp = t - p0;
    const T ap = t - bp;
    phic[k] = (p0 - ap) + (yv - bp);
    phi[k] = t;
  } else {
    phi[k] = (T)((double)phi[k] + (double)d[4]);
  }
}

// Velocity error of a judged step: sum |u1 + du - u2|^2 and sum |u2|^2
// (per-block partials), where 1 is the one-corrector state, du the predicted
// increment and 2 the plain two-corrector state.
template <class T, bool SHIFT>
__global__ void __launch_bounds__(256)
judge_kernel(const T* __restrict__ f1, const T* __restrict__ f2, const float* __restrict__ delta,
             int N, int NP, double rho, double* __restrict__ part) {
  __shared__ double r0[4], r1[4];
  double e = 0, s = 0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < N; k += gridDim.x * 256) {
    double n1, ux1, uy1, n2, ux2, uy2;
    velocity_at<T, SHIFT>(f1, NP, k, rho, n1, ux1, uy1);
    velocity_at<T, SHIFT>(f2, NP, k, rho, n2, ux2, uy2);
    const double ex = ux1 + (double)delta[(size_t)k * kFeat] - ux2;
    const double ey = uy1 + (double)delta[(size_t)k * kFeat + 1] - uy2;
    e += ex * ex + ey * ey;
    s += ux2 * ux2 + uy2 * uy2;
  }
  e = wave_sum(e);
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) {
    r0[threadIdx.x >> 6] = e;
    r1[threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = (r0[0] + r0[1]) + (r0[2] + r0[3]);
    part[2 * blockIdx.x + 1] = (r1[0] + r1[1]) + (r1[2] + r1[3]);
  }
}

}  // namespace lyo
