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

// The eight features of one node from its UpdateQuantities values c, the
// stresses of its x+1 / x-1 rows and y+1 / y-1 columns, its Q and its
// compensated phi p: ux, uy, Qxx, Qyx, phi, mu, Fx, Fy with F = div(sigma)
// (central differences, lyotropic.cpp:268-299).
template <class T>
__device__ __forceinline__ void feature_values(const NbrVals<T>& c, T sxx_xp, T sxx_xm, T syx_xp,
                                               T syx_xm, T sxy_yp, T sxy_ym, T syy_yp, T syy_ym,
                                               T qxx, T qyx, double p, float (&o)[kFeat]) {
  o[0] = (float)c.ux;
  o[1] = (float)c.uy;
  o[2] = (float)qxx;
  o[3] = (float)qyx;
  o[4] = (float)p;
  o[5] = (float)c.mu;
  o[6] = (float)(T(.5) * (sxx_xp - sxx_xm) + T(.5) * (sxy_yp - sxy_ym));
  o[7] = (float)(T(.5) * (syx_xp - syx_xm) + T(.5) * (syy_yp - syy_ym));
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
  feature_values<T>(c, xp.sxx, xm.sxx, xp.syx, xm.syx, yp.sxy, ym.sxy, yp.syy, ym.syy, qxx[k],
                    qyx[k], p, o);
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

// Applies the increment d = (du_x, du_y, dQxx, dQyx, dphi) to the state of one
// node: Q += dQ, phi += dphi (FP32: compensated, phic the running error), and
// f += fe(n, u + du) - fe(n, u) (src/lb.hpp:24-41), which keeps the node's
// mass and changes its momentum by n du.  SHIFT: f holds deviations from the
// rest equilibrium.
template <class T, bool SHIFT>
__device__ __forceinline__ void apply_node(T (&f)[9], T& qxx, T& qyx, T& phi, T* phic,
                                           const float (&d)[5], double rho) {
  if constexpr (sizeof(T) == 4) {
    // FP32: the difference itself, from fe_v = n w_v (1 + 3 c.u + 4.5 (c.u)^2
    // - 1.5 u.u): with s = u + (u + du) and D = du.s,
    //   fe_v(u + du) - fe_v(u) = n w_v (3 c.du + 4.5 (c.du)(c.s) - 1.5 D),
    // an even part a and an odd part b for each opposite pair (no
    // cancellation between two equilibria); the rest direction takes
    // -(2/3) n D, so the nine changes sum to zero
    T n;
    if constexpr (SHIFT) {
      n = (T)rho + ((((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7]))) + f[8]);
    } else {
      n = (((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7]))) + f[8];
    }
    const T inv = fast_rcp(n);
    const T d56 = f[5] - f[6], d87 = f[8] - f[7];
    const T ux = ((f[1] - f[2]) + d56 + d87) * inv;
    const T uy = ((f[3] - f[4]) + d56 - d87) * inv;
    const T dx = d[0], dy = d[1];
    const T sx = T(2) * ux + dx, sy = T(2) * uy + dy;
    const T D = dx * sx + dy * sy;
    const T n9 = n * T(1. / 9.), n36 = n * T(1. / 36.);
    const T hD = T(1.5) * D;
    const T a1 = n9 * (T(4.5) * dx * sx - hD), b1 = n9 * (T(3) * dx);
    const T a3 = n9 * (T(4.5) * dy * sy - hD), b3 = n9 * (T(3) * dy);
    const T ep = dx + dy, em = dx - dy;
    const T a5 = n36 * (T(4.5) * ep * (sx + sy) - hD), b5 = n36 * (T(3) * ep);
    const T a8 = n36 * (T(4.5) * em * (sx - sy) - hD), b8 = n36 * (T(3) * em);
    f[0] += T(-2. / 3.) * n * D;
    f[1] += a1 + b1;
    f[2] += a1 - b1;
    f[3] += a3 + b3;
    f[4] += a3 - b3;
    f[5] += a5 + b5;
    f[6] += a5 - b5;
    f[7] += a8 - b8;
    f[8] += a8 + b8;
    qxx += (T)d[2];
    qyx += (T)d[3];
    if (phic) {
      const T yv = (T)d[4] + *phic;
      const T p0 = phi;
      const T t = p0 + yv;
      const T bp = t - p0;
      const T ap = t - bp;
      *phic = (p0 - ap) + (yv - bp);
      phi = t;
    } else {
      phi += (T)d[4];
    }
  } else {
    double fv[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) fv[v] = (double)f[v] + (SHIFT ? rest_weight(v) * rho : 0.0);
    const double n = fv[0] + fv[1] + fv[2] + fv[3] + fv[4] + fv[5] + fv[6] + fv[7] + fv[8];
    const double ux = (fv[1] - fv[2] + fv[5] - fv[6] - fv[7] + fv[8]) / n;
    const double uy = (fv[3] - fv[4] + fv[5] - fv[6] + fv[7] - fv[8]) / n;
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
    for (int v = 0; v < 9; ++v) f[v] = (T)((double)f[v] + (b[v] - a[v]));
    qxx = (T)((double)qxx + (double)d[2]);
    qyx = (T)((double)qyx + (double)d[3]);
    if (phic) {
      const T yv = (T)d[4] + *phic;
      const T p0 = phi;
      const T t = p0 + yv;
      const T bp = t - p0;
      const T ap = t - bp;
      *phic = (p0 - ap) + (yv - bp);
      phi = t;
    } else {
      phi = (T)((double)phi + (double)d[4]);
    }
  }
}

// Applies the network's increments (FP32 [N][8], physical units) to a state
// in place (apply_node per node).
template <class T, bool SHIFT>
__global__ void apply_increment_kernel(T* __restrict__ f, T* __restrict__ qxx, T* __restrict__ qyx,
                                       T* __restrict__ phi, T* __restrict__ phic,
                                       const float* __restrict__ delta, int N, int NP,
                                       double rho) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const float* dp = delta + (size_t)k * kFeat;
  const float d[5] = {dp[0], dp[1], dp[2], dp[3], dp[4]};
  T fl[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) fl[v] = f[v * NP + k];
  T qx = qxx[k], qy = qyx[k], p = phi[k], pc = phic ? phic[k] : T(0);
  apply_node<T, SHIFT>(fl, qx, qy, p, phic ? &pc : nullptr, d, rho);
#pragma unroll
  for (int v = 0; v < 9; ++v) f[v * NP + k] = fl[v];
  qxx[k] = qx;
  qyx[k] = qy;
  phi[k] = p;
  if (phic) phic[k] = pc;
}

// Conservation of sum(phi) in surrogate steps.  The plain corrector's phi
// increment is a flux divergence, so it keeps sum(phi) (to rounding); the
// network's is not.  After the increments are applied, every node's phi is
// shifted by -mean(dphi) (the smallest change that restores the sum), from
// per-block (or per-wave) partial sums of the applied dphi in double, added in
// a fixed order.
constexpr int kPhiBlocks = 1024;  // blocks of the partial-sum and shift kernels

// part[b] = sum of delta[k][4] over the nodes k of block b (grid stride,
// fixed order); C: channels per node of delta
template <int C>
__global__ void __launch_bounds__(256)
dphi_partial_kernel(const float* __restrict__ delta, int N, double* __restrict__ part) {
  __shared__ double red[4];
  double a = 0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < N; k += gridDim.x * 256)
    a += (double)delta[(size_t)k * C + 4];
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// phi -= (sum of part[0..nparts)) / N on every node; FP32 (phic given): the
// shift enters the compensated sum as the increments do
template <class T>
__global__ void __launch_bounds__(256)
phi_shift_kernel(T* __restrict__ phi, T* __restrict__ phic, int N,
                 const double* __restrict__ part, int nparts) {
  __shared__ double red[4];
  double a = 0;
  for (int i = threadIdx.x; i < nparts; i += 256) a += part[i];
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  const double mean = ((red[0] + red[1]) + (red[2] + red[3])) / (double)N;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < N; k += gridDim.x * 256) {
    if (phic) {
      const T yv = (T)(-mean) + phic[k];
      const T p0 = phi[k];
      const T t = p0 + yv;
      const T bp = t - p0;
      const T ap = t - bp;
      phic[k] = (p0 - ap) + (yv - bp);
      phi[k] = t;
    } else {
      phi[k] = (T)((double)phi[k] - mean);
    }
  }
}

// Output calibration of the surrogate: per output c (du_x, du_y, dQxx, dQyx,
// dphi) a factor a_c on the network's increments, least-squares over judged
// steps.
struct NetCalib {
  float a[5];
};
constexpr int kJudgeSums = 12;  // error, norm, then (t.p, p.p) per output

// A judged step: per-block partial sums (part[b][kJudgeSums]) of
//   |u1 + a du - u2|^2, |u2|^2, and per output c: t_c p_c, p_c p_c
// where 1 is the one-corrector state, 2 the plain two-corrector state, p the
// network's increments (delta, before calibration), a the current
// calibration and t = (state 2) - (state 1) the increments the plain
// corrector made (as targets_kernel forms them).
template <class T, bool SHIFT>
__global__ void __launch_bounds__(256)
judge_kernel(const T* __restrict__ f1, const T* __restrict__ q1xx, const T* __restrict__ q1yx,
             const T* __restrict__ phi1, const T* __restrict__ dphi1, const T* __restrict__ f2,
             const T* __restrict__ q2xx, const T* __restrict__ q2yx, const T* __restrict__ phi2,
             const T* __restrict__ phic2, const float* __restrict__ delta, NetCalib cal, int N,
             int NP, double rho, double* __restrict__ part) {
  __shared__ double red[kJudgeSums][4];
  double acc[kJudgeSums];
#pragma unroll
  for (int j = 0; j < kJudgeSums; ++j) acc[j] = 0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < N; k += gridDim.x * 256) {
    double n1, ux1, uy1, n2, ux2, uy2;
    velocity_at<T, SHIFT>(f1, NP, k, rho, n1, ux1, uy1);
    velocity_at<T, SHIFT>(f2, NP, k, rho, n2, ux2, uy2);
    const float* d = delta + (size_t)k * kFeat;
    const double ex = ux1 + (double)cal.a[0] * (double)d[0] - ux2;
    const double ey = uy1 + (double)cal.a[1] * (double)d[1] - uy2;
    acc[0] += ex * ex + ey * ey;
    acc[1] += ux2 * ux2 + uy2 * uy2;
    const double p2 = (double)phi2[k] + (phic2 ? (double)phic2[k] : 0.0);
    const double p1 = (double)phi1[k] + (dphi1 ? (double)dphi1[k] : 0.0);
    const double t[5] = {ux2 - ux1, uy2 - uy1, (double)q2xx[k] - (double)q1xx[k],
                         (double)q2yx[k] - (double)q1yx[k], p2 - p1};
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      const double pc = (double)d[c];
      acc[2 + 2 * c] += t[c] * pc;
      acc[3 + 2 * c] += pc * pc;
    }
  }
#pragma unroll
  for (int j = 0; j < kJudgeSums; ++j) {
    const double v = wave_sum(acc[j]);
    if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < kJudgeSums) {
    const int j = threadIdx.x;
    part[kJudgeSums * blockIdx.x + j] = (red[j][0] + red[j][1]) + (red[j][2] + red[j][3]);
  }
}

}  // namespace lyo
