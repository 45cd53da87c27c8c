// lyo_diag.hpp -- run diagnostics on the device: rms velocity, phase-field
// statistics and histogram, and topological defects of the nematic field.
//
// Definitions follow the reference's post-processing:
//   velocity u = sum_v c_v f_v / sum_v f_v      (POST/massPy/base_modules/flow.py:7-11)
//   director from Q: S = |Q|, nx = sqrt((1+Qxx/S)/2), ny = sign(Qyx) sqrt((1-Qxx/S)/2)
//                                                (POST/massPy/nematic/nematicPy.py:4-15)
//   charge: winding of the director around the 8-ring, each step folded into
//   (-pi/2, pi/2], divided by 2 pi            (POST/massPy/base_modules/defects.py:15-50)
//   defects: 4-connected (periodic) same-sign regions with |charge| > 0.4
//                                                (defects.py:58-110, nematicPy.py:74-95)
#pragma once

#include "lyo_aux.hpp"

namespace lyo {

constexpr int kHistBins = 64;  // phi histogram: [0,1) in 64 bins, plus under/overflow

// Per block: sum |u|^2, sum |u|, sum phi, min/max phi (FP64, fixed order), and
// the phi histogram (integer atomics, order independent).
template <class T, bool SHIFT>
__global__ void __launch_bounds__(256)
diag_partial_kernel(const T* __restrict__ f, const T* __restrict__ phi,
                    const T* __restrict__ phic, int N, int NP, int per, double rho,
                    double* __restrict__ part, unsigned long long* __restrict__ hist) {
  __shared__ double red[4][5];
  __shared__ unsigned int lh[kHistBins + 2];
  for (int i = threadIdx.x; i < kHistBins + 2; i += 256) lh[i] = 0;
  __syncthreads();
  const int beg = blockIdx.x * per;
  const int end = min(beg + per, N);
  double s2 = 0, s1 = 0, sp = 0, pmin = 1e300, pmax = -1e300;
  for (int k = beg + (int)threadIdx.x; k < end; k += 256) {
    double fv[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      fv[v] = (double)f[v * NP + k];
      if (SHIFT) fv[v] += rest_weight(v) * rho;
    }
    const double n = fv[0] + fv[1] + fv[2] + fv[3] + fv[4] + fv[5] + fv[6] + fv[7] + fv[8];
    const double ux = (fv[1] - fv[2] + fv[5] - fv[6] - fv[7] + fv[8]) / n;
    const double uy = (fv[3] - fv[4] + fv[5] - fv[6] + fv[7] - fv[8]) / n;
    const double u2 = ux * ux + uy * uy;
    s2 += u2;
    s1 += sqrt(u2);
    double p = (double)phi[k];
    if (phic) p += (double)phic[k];
    sp += p;
    pmin = fmin(pmin, p);
    pmax = fmax(pmax, p);
    // bin 0: phi < 0 (or NaN); bins 1..64: [i/64, (i+1)/64); bin 65: phi >= 1
    int b;
    if (!(p >= 0)) {
      b = 0;
    } else if (p >= 1) {
      b = kHistBins + 1;
    } else {
      b = 1 + (int)(p * kHistBins);
      if (b > kHistBins) b = kHistBins;
    }
    atomicAdd(&lh[b], 1u);
  }
  s2 = wave_sum(s2);
  s1 = wave_sum(s1);
  sp = wave_sum(sp);
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    pmin = fmin(pmin, __shfl_xor(pmin, m, 64));
    pmax = fmax(pmax, __shfl_xor(pmax, m, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w][0] = s2;
    red[w][1] = s1;
    red[w][2] = sp;
    red[w][3] = pmin;
    red[w][4] = pmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* o = part + 5 * (size_t)blockIdx.x;
    o[0] = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
    o[1] = (red[0][1] + red[1][1]) + (red[2][1] + red[3][1]);
    o[2] = (red[0][2] + red[1][2]) + (red[2][2] + red[3][2]);
    o[3] = fmin(fmin(red[0][3], red[1][3]), fmin(red[2][3], red[3][3]));
    o[4] = fmax(fmax(red[0][4], red[1][4]), fmax(red[2][4], red[3][4]));
  }
  for (int i = threadIdx.x; i < kHistBins + 2; i += 256)
    if (lh[i]) atomicAdd(&hist[i], (unsigned long long)lh[i]);
}

// out[0..4] = sum |u|^2, sum |u|, sum phi, min phi, max phi over nparts partials
static __global__ void __launch_bounds__(64)
diag_final_kernel(const double* __restrict__ part, int nparts, double* __restrict__ out) {
  double a[5] = {0, 0, 0, 1e300, -1e300};
  for (int i = threadIdx.x; i < nparts; i += 64) {
    a[0] += part[5 * i];
    a[1] += part[5 * i + 1];
    a[2] += part[5 * i + 2];
    a[3] = fmin(a[3], part[5 * i + 3]);
    a[4] = fmax(a[4], part[5 * i + 4]);
  }
  a[0] = wave_sum(a[0]);
  a[1] = wave_sum(a[1]);
  a[2] = wave_sum(a[2]);
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    a[3] = fmin(a[3], __shfl_xor(a[3], m, 64));
    a[4] = fmax(a[4], __shfl_xor(a[4], m, 64));
  }
  if (threadIdx.x == 0)
    for (int j = 0; j < 5; ++j) out[j] = a[j];
}

__device__ __forceinline__ void director(double qxx, double qyx, double& nx, double& ny) {
  const double S = sqrt(qxx * qxx + qyx * qyx);
  const double c = qxx / S;  // NaN where S = 0: no director, no defect
  nx = sqrt(.5 * (1 + c));
  const double sg = qyx > 0 ? 1.0 : (qyx < 0 ? -1.0 : 0.0);
  ny = sg * sqrt(.5 * (1 - c));
}

// angle between two directors, folded into (-pi/2, pi/2]
__device__ __forceinline__ double wang(double ax, double ay, double bx, double by) {
  const double ang = atan2(ax * by - ay * bx, ax * bx + ay * by);
  return fabs(ang) > .5 * M_PI ? ang - (ang > 0 ? 1.0 : -1.0) * M_PI : ang;
}

// charge of every node (winding over the 8-ring in the order of
// defects.py:45: offsets -(1,0), -(1,1), -(0,1), -(-1,1), ...), its sign mask
// (|w| > 0.4) and the initial union-find labels.
template <class T>
__global__ void charge_kernel(const T* __restrict__ qxx, const T* __restrict__ qyx, int LX,
                              int LY, double* __restrict__ w_out, signed char* __restrict__ sgn,
                              int* __restrict__ label) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= LX * LY) return;
  const int x = k / LY, y = k - x * LY;
  // ring positions p_i = (x, y) - s_i for s in (1,0),(1,1),(0,1),(-1,1),(-1,0),(-1,-1),(0,-1),(1,-1),(1,0)
  const int sx[9] = {1, 1, 0, -1, -1, -1, 0, 1, 1};
  const int sy[9] = {0, 1, 1, 1, 0, -1, -1, -1, 0};
  double nx[9], ny[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    int xx = x - sx[i], yy = y - sy[i];
    xx += (xx < 0) ? LX : 0;
    xx -= (xx >= LX) ? LX : 0;
    yy += (yy < 0) ? LY : 0;
    yy -= (yy >= LY) ? LY : 0;
    const int kk = xx * LY + yy;
    director((double)qxx[kk], (double)qyx[kk], nx[i], ny[i]);
  }
  double w = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) w += wang(nx[i], ny[i], nx[i + 1], ny[i + 1]);
  w = w / (2 * M_PI);
  if (w_out) w_out[k] = w;
  const signed char s = fabs(w) > .4 ? (w > 0 ? 1 : -1) : 0;
  sgn[k] = s;
  label[k] = s ? k : -1;
}

// Labels only point to smaller or equal indices, so the walk ends; the loads
// are device-coherent so that other waves' atomicMin results are seen.
__device__ __forceinline__ int uf_find(int* L, int a) {
  int b = __hip_atomic_load(&L[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (b != a) {
    a = b;
    b = __hip_atomic_load(&L[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return a;
}

// Union-find with atomicMin (labels only ever decrease, so each loop ends).
__device__ __forceinline__ void uf_unite(int* L, int a, int b) {
  bool done = false;
  while (!done) {
    a = uf_find(L, a);
    b = uf_find(L, b);
    if (a < b) {
      const int old = atomicMin(&L[b], a);
      done = (old == b);
      b = old;
    } else if (b < a) {
      const int old = atomicMin(&L[a], b);
      done = (old == a);
      a = old;
    } else {
      done = true;
    }
  }
}

// joins each defect node with its +x and +y neighbours of the same sign
// (periodic), giving 4-connected same-sign components
static __global__ void ccl_union_kernel(const signed char* __restrict__ sgn, int* label, int LX, int LY) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= LX * LY) return;
  const signed char s = sgn[k];
  if (!s) return;
  const int x = k / LY, y = k - x * LY;
  const int kx = (x + 1 == LX ? 0 : x + 1) * LY + y;
  const int ky = x * LY + (y + 1 == LY ? 0 : y + 1);
  if (sgn[kx] == s) uf_unite(label, k, kx);
  if (sgn[ky] == s) uf_unite(label, k, ky);
}

// counts component roots by sign: cnt[0] positive, cnt[1] negative
static __global__ void ccl_count_kernel(const signed char* __restrict__ sgn, const int* __restrict__ label,
                                 int N, unsigned long long* cnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const signed char s = sgn[k];
  if (s && label[k] == k) atomicAdd(&cnt[s > 0 ? 0 : 1], 1ull);
}

}  // namespace lyo
