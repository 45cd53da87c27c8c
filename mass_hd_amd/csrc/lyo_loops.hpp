// lyo_loops.hpp -- the reference's loop structure of one step, one kernel per
// node loop, FP64 in the reference's expression order.
//
// Lyotropic::Step (src/models/lyotropic.cpp:500-525) runs UpdateQuantities
// (stores 23 derived fields per node, :168-247), UpdateFields(true) (reads
// them back through the stencils, :249-330), Move (push streaming of f and fn,
// :371-386), then npc times UpdateQuantities + UpdateFields(false)
// (:331-340).  These kernels keep that structure and that data flow -- the
// derived fields go through HBM, streaming is a pass of its own -- with one
// thread per node over the SoA planes.  They compute exactly what the fused
// sweeps compute (the same device functions, the same rounding): the kernel
// set exists to measure the reference's layout of the work on this GPU, next
// to the fused kernels that avoid most of its HBM traffic.
#pragma once

#include <hip/hip_runtime.h>

#include "lyo_physics.hpp"

namespace lyo {

// The 23 planes UpdateQuantities stores (lyotropic.cpp:223-246).
enum LoopPlane {
  LP_N, LP_UX, LP_UY, LP_UXP, LP_UYP, LP_HXX, LP_HYX, LP_MU,
  LP_DXQXX, LP_DXQYX, LP_DYQXX, LP_DYQYX,
  LP_SXX, LP_SYY, LP_SXY, LP_SYX,
  LP_BULK, LP_ELXX, LP_ELYX, LP_PFXX, LP_PFYX, LP_ACTXX, LP_ACTYX,
  kLoopPlanes
};

// Storage index of a node and of its four axis neighbours.
struct LoopNbr {
  int k, kxp, kxm, kyp, kym;
};

__device__ __forceinline__ LoopNbr loop_nbr(int k, int LX, int LY) {
  const int x = k / LY, y = k - x * LY;
  LoopNbr n;
  n.k = k;
  n.kxp = (x + 1 == LX ? 0 : x + 1) * LY + y;
  n.kxm = (x == 0 ? LX - 1 : x - 1) * LY + y;
  n.kyp = x * LY + (y + 1 == LY ? 0 : y + 1);
  n.kym = x * LY + (y == 0 ? LY - 1 : y - 1);
  return n;
}

// Node t = y + LY*x of the domain is stored at t on the periodic grid
// (src/geometry.cpp:100-110) ...
struct PeriodicGrid {
  __device__ __forceinline__ static LoopNbr at(int t, int LX, int LY) {
    return loop_nbr(t, LX, LY);
  }
};

// ... and at (y+1) + (LY+2)*(x+1) on the grid with a ghost layer
// (src/geometry.cpp:140-183), whose ghost nodes hold the walls' values: a
// domain node's neighbours are never wrapped there.
struct LayerGrid {
  __device__ __forceinline__ static LoopNbr at(int t, int LX, int LY) {
    (void)LX;
    const int x = t / LY, y = t - x * LY, W = LY + 2;
    LoopNbr n;
    n.k = (y + 1) + W * (x + 1);
    n.kxp = n.k + W;
    n.kxm = n.k - W;
    n.kyp = n.k + 1;
    n.kym = n.k - 1;
    return n;
  }
};

// The ghost-layer layout with periodic neighbours taken inside the domain (a
// domain edge's neighbour is the node across the grid, not a ghost): a
// periodic run on that layout reads and writes no ghost node.
struct LayerWrapGrid {
  __device__ __forceinline__ static LoopNbr at(int t, int LX, int LY) {
    const int x = t / LY, y = t - x * LY, W = LY + 2;
    LoopNbr n;
    n.k = (y + 1) + W * (x + 1);
    n.kxp = x + 1 < LX ? n.k + W : (y + 1) + W;
    n.kxm = x > 0 ? n.k - W : (y + 1) + W * LX;
    n.kyp = y + 1 < LY ? n.k + 1 : 1 + W * (x + 1);
    n.kym = y > 0 ? n.k - 1 : LY + W * (x + 1);
    return n;
  }
};

// UpdateQuantities (lyotropic.cpp:168-247, loop :343-360): all 23 derived
// planes of every domain node, plane i at D + i*NP.
template <class T, bool EXACT, class G = PeriodicGrid>
__global__ void __launch_bounds__(256) loop_uq_kernel(const T* __restrict__ f,
                                                      const T* __restrict__ qxx,
                                                      const T* __restrict__ qyx,
                                                      const T* __restrict__ phi, Params<T> P,
                                                      int LX, int LY, int NP, T* __restrict__ D) {
  static_assert(EXACT, "the loop kernels run the reference's FP64 arithmetic");
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= LX * LY) return;
  const LoopNbr n = G::at(t, LX, LY);
  const int k = n.k;
  T fl[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) fl[v] = f[v * NP + k];
  NbrVals<T> nv;
  CenterVals<T> cv;
  DiagVals<T> dg;
  update_quantities<T, EXACT, true>(
      fl, qxx[k], qxx[n.kxp], qxx[n.kxm], ynb<T, EXACT>(qxx[n.kyp], qxx[n.kym]), qyx[k],
      qyx[n.kxp], qyx[n.kxm], ynb<T, EXACT>(qyx[n.kyp], qyx[n.kym]), phi[k], phi[n.kxp],
      phi[n.kxm], ynb<T, EXACT>(phi[n.kyp], phi[n.kym]), P, nv, cv, &dg);
  const T out[kLoopPlanes] = {cv.nn,     nv.ux,     nv.uy,     nv.uxp,    nv.uyp,    cv.Hxx,
                              cv.Hyx,    nv.mu,     cv.dxQxx,  cv.dxQyx,  cv.dyQxx,  cv.dyQyx,
                              nv.sxx,    nv.syy,    nv.sxy,    nv.syx,    dg.bulk,   dg.el_xx,
                              dg.el_yx,  dg.pf_xx,  dg.pf_yx,  dg.act_xx, dg.act_yx};
#pragma unroll
  for (int i = 0; i < kLoopPlanes; ++i) D[(size_t)i * NP + k] = out[i];
}

// Inputs and outputs of one UpdateFields loop.  FIRST (the predictor,
// lyotropic.cpp:310-330): reads f, Q, phi and writes QN = Q + D/2, Q + D,
// phn = phi + Dp/2, phi + Dp, fn = f + Om/2 and f + Om (unstreamed).  Not
// FIRST (a corrector, :331-340): reads f, phi (relaxation time) and fn, QN,
// phn, and writes QN + D/2, phn + Dp/2 and fn + Om/2.
template <class T>
struct LoopUF {
  const T* D;
  const T* f;
  const T* qxx;
  const T* qyx;
  const T* phi;
  const T* fn;    // corrector
  const T* qnxx;  // corrector
  const T* qnyx;  // corrector
  const T* phn;   // corrector
  T* f_out;
  T* fn_out;      // predictor
  T* qxx_out;
  T* qyx_out;
  T* qnxx_out;    // predictor
  T* qnyx_out;    // predictor
  T* phi_out;
  T* phn_out;     // predictor
};

template <class T, bool EXACT, bool FIRST, class G = PeriodicGrid>
__global__ void __launch_bounds__(256) loop_uf_kernel(LoopUF<T> a, Params<T> P, int LX, int LY,
                                                      int NP) {
  static_assert(EXACT, "the loop kernels run the reference's FP64 arithmetic");
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= LX * LY) return;
  const LoopNbr n = G::at(t, LX, LY);
  const int k = n.k;
  auto pl = [&](int i) { return a.D + (size_t)i * NP; };
  CenterVals<T> c;
  c.nn = pl(LP_N)[k];
  c.dn = T(0);
  c.vx = pl(LP_UX)[k];
  c.vy = pl(LP_UY)[k];
  c.Hxx = pl(LP_HXX)[k];
  c.Hyx = pl(LP_HYX)[k];
  c.dxQxx = pl(LP_DXQXX)[k];
  c.dyQxx = pl(LP_DYQXX)[k];
  c.dxQyx = pl(LP_DXQYX)[k];
  c.dyQyx = pl(LP_DYQYX)[k];
  UFStencil<T, EXACT> s;
  s.mu_c = pl(LP_MU)[k];
  s.ux_xp = pl(LP_UX)[n.kxp];
  s.ux_xm = pl(LP_UX)[n.kxm];
  s.uy_xp = pl(LP_UY)[n.kxp];
  s.uy_xm = pl(LP_UY)[n.kxm];
  s.uxp_xp = pl(LP_UXP)[n.kxp];
  s.uxp_xm = pl(LP_UXP)[n.kxm];
  s.sxx_xp = pl(LP_SXX)[n.kxp];
  s.sxx_xm = pl(LP_SXX)[n.kxm];
  s.syx_xp = pl(LP_SYX)[n.kxp];
  s.syx_xm = pl(LP_SYX)[n.kxm];
  s.mu_xp = pl(LP_MU)[n.kxp];
  s.mu_xm = pl(LP_MU)[n.kxm];
  s.ux_y = ynb<T, EXACT>(pl(LP_UX)[n.kyp], pl(LP_UX)[n.kym]);
  s.uy_y = ynb<T, EXACT>(pl(LP_UY)[n.kyp], pl(LP_UY)[n.kym]);
  s.uyp_y = ynb<T, EXACT>(pl(LP_UYP)[n.kyp], pl(LP_UYP)[n.kym]);
  s.sxy_y = ynb<T, EXACT>(pl(LP_SXY)[n.kyp], pl(LP_SXY)[n.kym]);
  s.syy_y = ynb<T, EXACT>(pl(LP_SYY)[n.kyp], pl(LP_SYY)[n.kym]);
  s.mu_y = ynb<T, EXACT>(pl(LP_MU)[n.kyp], pl(LP_MU)[n.kym]);
  T fl[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) fl[v] = a.f[v * NP + k];
  const T Qxx = a.qxx[k], Qyx = a.qyx[k], p = a.phi[k];
  T Dxx, Dyx, Dp, om[9];
  update_fields<T, EXACT>(c, Qxx, Qyx, p, fl, s, P, Dxx, Dyx, Dp, om);
  if constexpr (FIRST) {
    a.qnxx_out[k] = Qxx + T(.5) * Dxx;
    a.qnyx_out[k] = Qyx + T(.5) * Dyx;
    a.phn_out[k] = p + T(.5) * Dp;
    a.qxx_out[k] = Qxx + Dxx;
    a.qyx_out[k] = Qyx + Dyx;
    a.phi_out[k] = p + Dp;
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      a.fn_out[v * NP + k] = fl[v] + T(.5) * om[v];
      a.f_out[v * NP + k] = fl[v] + om[v];
    }
  } else {
    a.qxx_out[k] = a.qnxx[k] + T(.5) * Dxx;
    a.qyx_out[k] = a.qnyx[k] + T(.5) * Dyx;
    a.phi_out[k] = a.phn[k] + T(.5) * Dp;
#pragma unroll
    for (int v = 0; v < 9; ++v) a.f_out[v * NP + k] = a.fn[v * NP + k] + T(.5) * om[v];
  }
}

// Move (lyotropic.cpp:371-386): f_to[next(k, v)][v] = f[k][v] for f and fn,
// directions of src/geometry.hpp:17-27.
template <class T>
__global__ void __launch_bounds__(256) loop_move_kernel(const T* __restrict__ f,
                                                        const T* __restrict__ fn,
                                                        T* __restrict__ f_to,
                                                        T* __restrict__ fn_to, int LX, int LY,
                                                        int NP) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= LX * LY) return;
  const LoopNbr n = loop_nbr(k, LX, LY);
  const int x = k / LY, y = k - x * LY;
  const int xp = n.kxp - y, xm = n.kxm - y;  // row offsets of x+1, x-1
  const int yp = n.kyp - x * LY, ym = n.kym - x * LY;
  const int row[9] = {x * LY, xp, xm, x * LY, x * LY, xp, xm, xm, xp};
  const int col[9] = {y, y, y, yp, ym, yp, ym, yp, ym};
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    const int d = v * NP + row[v] + col[v];
    f_to[d] = f[v * NP + k];
    fn_to[d] = fn[v * NP + k];
  }
}

}  // namespace lyo
