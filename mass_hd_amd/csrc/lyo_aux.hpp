// lyo_aux.hpp -- state conversion, frame fields and FP64 reductions.
#pragma once

#include "lyo_physics.hpp"

namespace lyo {

// Field ids of the derived frame fields (same numbering as mass_lyo_field).
enum DerivedId {
  D_N = 4, D_UX = 5, D_UY = 6, D_SXX = 7, D_SXY = 8, D_SYX = 9, D_SYY = 10,
  D_BULK = 11, D_EL_XX = 12, D_EL_YX = 13, D_PF_XX = 14, D_PF_YX = 15,
  D_ACT_XX = 16, D_ACT_YX = 17, D_MU = 18, D_HXX = 19, D_HYX = 20
};

// Host FP64 AoS state -> device SoA planes of T.  SHIFT subtracts the rest
// equilibrium rest_weight(v)*rho from the populations; phic receives the part
// of phi that does not fit in T.
template <class T, bool SHIFT>
__global__ void pack_state_kernel(const double* __restrict__ ff, const double* __restrict__ qxx,
                                  const double* __restrict__ qyx, const double* __restrict__ phi,
                                  T* f, T* oqxx, T* oqyx, T* ophi, T* ophic, int N, int NP, double rho) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    const double x = ff[9 * (size_t)k + v];
    f[v * NP + k] = SHIFT ? (T)(x - rest_weight(v) * rho) : (T)x;
  }
  oqxx[k] = (T)qxx[k];
  oqyx[k] = (T)qyx[k];
  const T p = (T)phi[k];
  ophi[k] = p;
  if (ophic) ophic[k] = (T)(phi[k] - (double)p);
}

template <class T, bool SHIFT>
__global__ void unpack_f_kernel(const T* __restrict__ f, double* __restrict__ ff, int N, int NP,
                                double rho) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    const double x = (double)f[v * NP + k];
    ff[9 * (size_t)k + v] = SHIFT ? x + rest_weight(v) * rho : x;
  }
}

template <class T>
__global__ void unpack_scalar_kernel(const T* __restrict__ a, const T* __restrict__ c,
                                     double* __restrict__ out, int N) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  double x = (double)a[k];
  if (c) x += (double)c[k];
  out[k] = x;
}

// Derived fields of the last UpdateQuantities call, recomputed from that
// call's inputs (f_in, Q_in, phi_in [+ dphi_in]) with one thread per node.
template <class T, bool EXACT>
__global__ void derived_kernel(const T* __restrict__ f, const T* __restrict__ qxx,
                               const T* __restrict__ qyx, const T* __restrict__ phi,
                               const T* __restrict__ dphi, Params<T> P, int LX, int LY,
                               int NP, int field, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = LX * LY;
  if (k >= N) return;
  const int x = k / LY, y = k - x * LY;
  const int kxp = (x + 1 == LX ? 0 : x + 1) * LY + y;
  const int kxm = (x == 0 ? LX - 1 : x - 1) * LY + y;
  const int kyp = x * LY + (y + 1 == LY ? 0 : y + 1);
  const int kym = x * LY + (y == 0 ? LY - 1 : y - 1);
  auto ph = [&](int i) { return dphi ? phi[i] + dphi[i] : phi[i]; };
  T fl[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) fl[v] = f[v * NP + k];
  NbrVals<T> nv;
  CenterVals<T> cv;
  DiagVals<T> dg;
  update_quantities<T, EXACT, true>(
      fl, qxx[k], qxx[kxp], qxx[kxm], ynb<T, EXACT>(qxx[kyp], qxx[kym]), qyx[k], qyx[kxp],
      qyx[kxm], ynb<T, EXACT>(qyx[kyp], qyx[kym]), ph(k), ph(kxp), ph(kxm),
      ynb<T, EXACT>(ph(kyp), ph(kym)), P, nv, cv, &dg);
  T r = T(0);
  switch (field) {
    case D_N: r = cv.nn; break;
    case D_UX: r = nv.ux; break;
    case D_UY: r = nv.uy; break;
    case D_SXX: r = nv.sxx; break;
    case D_SXY: r = nv.sxy; break;
    case D_SYX: r = nv.syx; break;
    case D_SYY: r = nv.syy; break;
    case D_BULK: r = dg.bulk; break;
    case D_EL_XX: r = dg.el_xx; break;
    case D_EL_YX: r = dg.el_yx; break;
    case D_PF_XX: r = dg.pf_xx; break;
    case D_PF_YX: r = dg.pf_yx; break;
    case D_ACT_XX: r = dg.act_xx; break;
    case D_ACT_YX: r = dg.act_yx; break;
    case D_MU: r = nv.mu; break;
    case D_HXX: r = cv.Hxx; break;
    case D_HYX: r = cv.Hyx; break;
    default: break;
  }
  out[k] = (double)r;
}

// wave64 shuffle-tree sum
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Block b sums nodes [b*per, min((b+1)*per, N)) of sum_v f (+ nshift per
// node) and phi (+ phic) in a fixed order, leaving one partial per block.
template <class T>
__global__ void __launch_bounds__(256)
partial_sums_kernel(const T* __restrict__ f, const T* __restrict__ phi,
                    const T* __restrict__ phic, int N, int NP, int per, double* __restrict__ part_f,
                    double* __restrict__ part_p) {
  __shared__ double sf[4], sp[4];
  const int beg = blockIdx.x * per;
  const int end = min(beg + per, N);
  double af = 0, ap = 0;
  for (int k = beg + (int)threadIdx.x; k < end; k += 256) {
    double s = 0;
#pragma unroll
    for (int v = 0; v < 9; ++v) s += (double)f[v * NP + k];
    af += s;
    ap += (double)phi[k];
    if (phic) ap += (double)phic[k];
  }
  af = wave_sum(af);
  ap = wave_sum(ap);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sf[w] = af;
    sp[w] = ap;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part_f[blockIdx.x] = (sf[0] + sf[1]) + (sf[2] + sf[3]);
    part_p[blockIdx.x] = (sp[0] + sp[1]) + (sp[2] + sp[3]);
  }
}

static __global__ void __launch_bounds__(256)
final_sums_kernel(const double* __restrict__ part_f, const double* __restrict__ part_p, int n,
                  double add_f, double* __restrict__ out) {
  __shared__ double sf[4], sp[4];
  double af = 0, ap = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    af += part_f[i];
    ap += part_p[i];
  }
  af = wave_sum(af);
  ap = wave_sum(ap);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sf[w] = af;
    sp[w] = ap;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = ((sf[0] + sf[1]) + (sf[2] + sf[3])) + add_f;
    out[1] = (sp[0] + sp[1]) + (sp[2] + sp[3]);
  }
}

}  // namespace lyo
