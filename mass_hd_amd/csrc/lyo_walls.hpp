// lyo_walls.hpp -- walled grids (bc 1-4): free-slip / no-slip channels and
// boxes on a ghost layer, FP64 in the reference's expression order.
//
// Lyotropic with bc != 0 (src/models/lyotropic.cpp:15-17) stores its fields on
// a grid with one layer of ghost nodes around the domain
// (src/geometry.cpp:140-183): the (LX+2)*(LY+2) torus, whose ghost nodes the
// boundary conditions (lyotropic.cpp:388-498, src/fields.hpp:619-918) fill
// from the domain before each use, and through which Move streams the
// populations like any other node (lyotropic.cpp:371-386 runs over the whole
// grid).  Here every plane is that torus, node (X, Y) at Y + (LY+2)*X with
// X, Y in [0, LX+1] x [0, LY+2); domain node (x, y) is (x+1, y+1).  The step
// runs either as three fused tile sweeps (wall_sweep_kernel: field conditions
// applied in LDS, the LB conditions as kernels, Move in the correctors'
// reads) or as the reference's loop structure -- the node loops of
// lyo_loops.hpp on LayerGrid -- with the boundary conditions as kernels in
// between; both are bitwise the oracle's restatement.  go-or-grow on a
// periodic grid (bc 0) keeps this layout with LayerWrapGrid neighbours and
// no ghost conditions.
//
// The boundary conditions run in three launches per call, in the
// reference's order: Left/Right walls, then Front/Back walls, then the four
// corners.  Within a launch no ghost a thread writes is read by another
// thread (walls read the domain, the previous launch's ghosts, or -- the
// no-slip rule at a wall's ends -- ghosts of walls that come later, before
// those are written, as in the reference), so each launch is one thread per
// ghost node.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <utility>
#include <vector>

#include "lyo_aux.hpp"
#include "lyo_engine.hpp"
#include "lyo_growth.hpp"
#include "lyo_loops.hpp"

namespace lyo {

// src/geometry.hpp:148-199: walls and corners carry the number of the
// direction that points out of the domain through them.
enum WallId { kWallRight = 1, kWallLeft = 2, kWallBack = 3, kWallFront = 4 };
enum CornerId { kCornerRB = 5, kCornerLF = 6, kCornerLB = 7, kCornerRF = 8 };

// src/geometry.hpp:17-27 (directions), 40-59 (reflection about the plane of
// normal i; row 0: bounce back), 65-84 (projection onto that plane)
__constant__ int kWallCX[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
__constant__ int kWallCY[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
__constant__ int kRefl[9][9] = {
    {0, 2, 1, 4, 3, 6, 5, 8, 7}, {0, 2, 1, 3, 4, 7, 8, 5, 6}, {0, 2, 1, 3, 4, 7, 8, 5, 6},
    {0, 1, 2, 4, 3, 8, 7, 6, 5}, {0, 1, 2, 4, 3, 8, 7, 6, 5}, {0, 4, 3, 2, 1, 6, 5, 7, 8},
    {0, 4, 3, 2, 1, 6, 5, 7, 8}, {0, 3, 4, 1, 2, 5, 6, 8, 7}, {0, 3, 4, 1, 2, 5, 6, 8, 7}};
__constant__ int kProj[9][9] = {
    {0, 0, 0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 3, 4, 3, 4, 3, 4}, {0, 0, 0, 3, 4, 3, 4, 3, 4},
    {0, 1, 2, 0, 0, 1, 2, 2, 1}, {0, 1, 2, 0, 0, 1, 2, 2, 1}, {0, 5, 6, 5, 6, 5, 6, 0, 0},
    {0, 1, 2, 4, 3, 8, 7, 6, 5}, {0, 8, 7, 7, 8, 0, 0, 7, 8}, {0, 8, 7, 7, 8, 0, 0, 7, 8}};

// Geometry of the ghost-layer torus.
struct WallGeo {
  int LX, LY;
  int W;   // LY + 2: stride of X
  int NP;  // plane stride (>= (LX+2)*(LY+2))
  int bc;  // 1 free-slip channel, 2 no-slip channel, 3 free-slip box, 4 no-slip box;
           // 0: periodic grid (go-or-grow): the node kernels wrap inside the
           // domain (LayerWrapGrid) and no condition runs
};

__device__ __forceinline__ int wopp(int d) { return kRefl[0][d]; }

// node index of (X, Y) and its neighbour in direction d on the torus
__device__ __forceinline__ int wnode(const WallGeo& g, int X, int Y) { return Y + g.W * X; }
__device__ __forceinline__ int wnext(const WallGeo& g, int p, int d) {
  int X = p / g.W, Y = p - X * g.W;
  X += kWallCX[d];
  Y += kWallCY[d];
  X = X < 0 ? X + g.LX + 2 : (X > g.LX + 1 ? X - (g.LX + 2) : X);
  Y = Y < 0 ? Y + g.W : (Y >= g.W ? Y - g.W : Y);
  return wnode(g, X, Y);
}

// The ghost node of thread i in a launch (phase 0: Left/Right walls, i <
// 2*LY; phase 1: Front/Back walls, i < 2*LX; phase 2: the four corners in
// the order the box calls them, LeftFront, RightFront, LeftBack, RightBack)
// and its wall (phases 0, 1) or corner (phase 2).  False past the end.
__device__ __forceinline__ bool wall_ghost(const WallGeo& g, int phase, int i, int& p, int& id) {
  if (phase == 0) {
    if (i >= 2 * g.LY) return false;
    const bool right = i >= g.LY;
    const int Y = (right ? i - g.LY : i) + 1;
    p = wnode(g, right ? g.LX + 1 : 0, Y);
    id = right ? kWallRight : kWallLeft;
  } else if (phase == 1) {
    if (i >= 2 * g.LX) return false;
    const bool back = i >= g.LX;
    const int X = (back ? i - g.LX : i) + 1;
    p = wnode(g, X, back ? g.LY + 1 : 0);
    id = back ? kWallBack : kWallFront;
  } else {
    if (i >= 4) return false;
    const int c[4] = {kCornerLF, kCornerRF, kCornerLB, kCornerRB};
    id = c[i];
    const int X = (id == kCornerLF || id == kCornerLB) ? 0 : g.LX + 1;
    const int Y = (id == kCornerLF || id == kCornerRF) ? 0 : g.LY + 1;
    p = wnode(g, X, Y);
  }
  return true;
}

// the ghost node of thread i in a launch of the given phase, and its phase
__device__ __forceinline__ bool any_ghost(const WallGeo& g, int phase, int i, int& p, int& id,
                                          int& ph) {
  ph = phase;
  return wall_ghost(g, phase, i, p, id);
}

// channel corners: the closed wall next to corner c (the other wall is
// periodic; fields.hpp:844-885)
__device__ __forceinline__ int channel_wall(int c) {
  return (c == kCornerLF || c == kCornerRF) ? kWallFront : kWallBack;
}

// Scalar rules.  kind 0: Neumann (fields.hpp:704-723), 1: Dirichlet with
// value 0 (:725-747), 2: copy derivative (:749-768); PBC walls of a channel
// copy the opposite domain column (:619-625).
template <class T>
__device__ __forceinline__ void scalar_ghost(const WallGeo& g, T* a, int kind, int phase, int p,
                                             int id) {
  const bool channel = g.bc <= 2;
  if (phase == 0 && channel) {  // ApplyPBC(Left), ApplyPBC(Right)
    const int Y = p % g.W;
    const int src = id == kWallLeft ? wnode(g, g.LX, Y) : wnode(g, 1, Y);
    a[p] = a[src];
    return;
  }
  // direction towards the domain: away from the wall, or for a corner away
  // from the corner (box) or from its closed wall (channel)
  const int d = phase < 2 ? wopp(id) : wopp(channel ? channel_wall(id) : id);
  const int j = wnext(g, p, d);
  if (kind == 0) {
    a[p] = a[j];
  } else if (kind == 1) {
    a[p] = T(2) * T(0) - a[j];
  } else {
    a[p] = T(2) * a[j] - a[wnext(g, j, d)];
  }
}

// src/geometry.hpp:312-316
__device__ __forceinline__ bool incoming(int v, int w) {
  return kWallCX[w] * kWallCX[v] == -1 || kWallCY[w] * kWallCY[v] == -1;
}

// LB rules on one population field (9 planes): PBC walls copy the whole
// node (fields.hpp:619-625), free slip reflects the incoming directions
// about the wall (:647-661), no slip bounces them back (:663-674); corners
// bounce back (:689-701) except the free-slip channel's (:675-687).
template <class T>
__device__ __forceinline__ void lb_ghost(const WallGeo& g, T* f, int phase, int p, int id) {
  const int NP = g.NP;
  const bool channel = g.bc <= 2, slip = g.bc == 1 || g.bc == 3;
  if (phase == 0 && channel) {
    const int Y = p % g.W;
    const int src = id == kWallLeft ? wnode(g, g.LX, Y) : wnode(g, 1, Y);
#pragma unroll
    for (int v = 0; v < 9; ++v) f[v * NP + p] = f[v * NP + src];
    return;
  }
  if (phase < 2) {
    const int w = id;
#pragma unroll
    for (int v = 1; v < 9; ++v) {
      if (!incoming(v, w)) continue;
      if (slip)
        f[v * NP + p] = f[kRefl[w][v] * NP + wnext(g, p, wopp(w))];
      else
        f[v * NP + p] = f[wopp(v) * NP + wnext(g, p, v)];
    }
    return;
  }
  const int c = id;
  if (g.bc == 1) {
    const int w = channel_wall(c);
    const int pr = kProj[w][c];
    int X = 0, Y = 0;
    const int rc = kRefl[pr][c];  // the other corner along the closed wall
    X = (rc == kCornerLF || rc == kCornerLB) ? 0 : g.LX + 1;
    Y = (rc == kCornerLF || rc == kCornerRF) ? 0 : g.LY + 1;
    const int l = wnext(g, wnode(g, X, Y), pr);
    f[wopp(c) * NP + p] = f[wopp(c) * NP + l];
  } else {
    f[wopp(c) * NP + p] = f[c * NP + wnext(g, p, wopp(c))];
  }
}

// BoundaryConditionsFields (lyotropic.cpp:422-449): Neumann on Qxx, Qyx,
// phi, and for go-or-grow on m and chi (go-or-grow.cpp:365-385).
template <class T>
__global__ void __launch_bounds__(256) wall_fields_kernel(WallGeo g, int phase, T* qxx, T* qyx,
                                                          T* phi, T* m, T* chi) {
  int p, id, ph;
  if (!any_ghost(g, phase, blockIdx.x * blockDim.x + threadIdx.x, p, id, ph)) return;
  scalar_ghost(g, qxx, 0, ph, p, id);
  scalar_ghost(g, qyx, 0, ph, p, id);
  scalar_ghost(g, phi, 0, ph, p, id);
  if (m) {
    scalar_ghost(g, m, 0, ph, p, id);
    scalar_ghost(g, chi, 0, ph, p, id);
  }
}

// BoundaryConditionsFields2 (lyotropic.cpp:451-498) on the derived planes:
// copy derivative on u, Dirichlet 0 on u*phi, Neumann on MU and the stresses.
template <class T>
__global__ void __launch_bounds__(256) wall_fields2_kernel(WallGeo g, int phase, T* D) {
  int p, id, ph;
  if (!any_ghost(g, phase, blockIdx.x * blockDim.x + threadIdx.x, p, id, ph)) return;
  auto pl = [&](int i) { return D + (size_t)i * g.NP; };
  scalar_ghost(g, pl(LP_UY), 2, ph, p, id);
  scalar_ghost(g, pl(LP_UX), 2, ph, p, id);
  scalar_ghost(g, pl(LP_UYP), 1, ph, p, id);
  scalar_ghost(g, pl(LP_UXP), 1, ph, p, id);
  scalar_ghost(g, pl(LP_MU), 0, ph, p, id);
  scalar_ghost(g, pl(LP_SXX), 0, ph, p, id);
  scalar_ghost(g, pl(LP_SYY), 0, ph, p, id);
  scalar_ghost(g, pl(LP_SYX), 0, ph, p, id);
  scalar_ghost(g, pl(LP_SXY), 0, ph, p, id);
}

// Neumann on up to two more scalar planes: m and chi (go-or-grow.cpp:
// 365-385), or MU of the dry relaxation (:70-83).
template <class T>
__global__ void __launch_bounds__(256) wall_scalars_kernel(WallGeo g, int phase, T* a, T* b) {
  int p, id, ph;
  if (!any_ghost(g, phase, blockIdx.x * blockDim.x + threadIdx.x, p, id, ph)) return;
  scalar_ghost(g, a, 0, ph, p, id);
  if (b) scalar_ghost(g, b, 0, ph, p, id);
}

// BoundaryConditionsLB (lyotropic.cpp:388-420) on ff and fn.
template <class T>
__global__ void __launch_bounds__(256) wall_lb_kernel(WallGeo g, int phase, T* f, T* fn) {
  int p, id, ph;
  if (!any_ghost(g, phase, blockIdx.x * blockDim.x + threadIdx.x, p, id, ph)) return;
  lb_ghost(g, f, ph, p, id);
  lb_ghost(g, fn, ph, p, id);
}

// One sweep of the walled step in one kernel: UpdateQuantities,
// BoundaryConditionsFields2 and UpdateFields (lyotropic.cpp:168-341,451-498)
// for a tile of domain nodes.  The block stages Q and phi of the tile
// plus two nodes of halo in LDS (ghost nodes take their field condition --
// the adjacent domain node, or the periodic image across a channel's open
// ends -- as they are read), runs UpdateQuantities on the tile plus one node
// of halo into LDS, fills the walled halo nodes' derived values by the
// fields2 rules from the two nodes inside, and runs UpdateFields on the tile.
// The same device functions as the node loops, so results are bitwise theirs.
// Inputs and outputs are distinct buffers (a tile reads its neighbours'
// state); the diagonal halo nodes are never read by the 5-point stencils.
struct WallSweep {
  const double* f;     // f (predictor) or streamed f (corrector)
  const double* qxx;
  const double* qyx;
  const double* phi;
  const double* fn;    // corrector: streamed fn
  const double* qnxx;  // corrector
  const double* qnyx;  // corrector
  const double* phn;   // corrector
  double* f_out;
  double* fn_out;      // predictor
  double* qxx_out;
  double* qyx_out;
  double* qnxx_out;    // predictor
  double* qnyx_out;    // predictor
  double* phi_out;
  double* phn_out;     // predictor
  double* D;           // derive: the 23 derived planes of the tile's nodes
};

// Tile of domain nodes per block: 10 rows (x) by 62 columns (y).  The block's
// 768 threads are the 12 x 64 nodes of the tile plus one ring, one wave per
// row (contiguous in y), and each runs UpdateQuantities on its node; the
// tile's threads keep their own node's values in registers for UpdateFields.
// 77.5 KB of LDS per block: two blocks per CU, held at 6 waves per SIMD by
// an 80-register cap (the correctors spill a few values to scratch).
constexpr int kWallTX = 10, kWallTY = 62, kWallThreads = (kWallTX + 2) * (kWallTY + 2);
// go-or-grow sweeps (growth_sweep_kernel): tiles of kGrowTX rows (their
// kernels hold 128 VGPRs, four waves per SIMD; 6-row tiles, two 512-thread
// blocks per CU, ran the 500^2 step as fast as 10-row ones, within noise)
constexpr int kGrowTX = 10, kGrowThreads = (kGrowTX + 2) * (kWallTY + 2);

// Domain coordinate a stencil reads for coordinate v of an axis of length L
// (periodic: wrapped; walled: the ghost at -1 / L reads its Neumann source
// 0 / L-1), or -1 if no stencil reads v.
__device__ __forceinline__ int wall_axis(int v, int L, bool periodic) {
  if (v >= 0 && v < L) return v;
  if (periodic) return ((v % L) + L) % L;
  if (v == -1) return 0;
  if (v == L) return L - 1;
  return -1;
}

enum WallSweepMode { kSweepPredictor = 0, kSweepCorrector = 1, kSweepDerive = 2 };

// MODE kSweepDerive runs UpdateQuantities only and stores the 23 derived
// planes of the tile (the frame fields of the last UpdateQuantities,
// recomputed from its inputs when a frame asks for them).  PULL_F / PULL_FN:
// f / fn are the predictor's post-collision populations with the LB
// conditions applied, and Move (lyotropic.cpp:371-386) happens as they are
// read -- node p takes direction v from p - c_v on the torus.  Every
// corrector pulls fn from those planes, which no sweep overwrites.
template <int MODE, bool PULL_F = false, bool PULL_FN = false>
__global__ void __launch_bounds__(kWallThreads) __attribute__((amdgpu_waves_per_eu(6)))
wall_sweep_kernel(WallSweep a, Params<double> P, WallGeo g) {
  using T = double;
  constexpr int TX = kWallTX, TY = kWallTY;
  constexpr int PW = TY + 4, PN = (TX + 4) * PW;  // staged Q, phi: tile + 2
  constexpr int DW = TY + 2, DN = (TX + 2) * DW;  // derived: tile + 1 (= threads)
  __shared__ T sq[3][PN];
  __shared__ T sd[9][DN];  // ux uy uxp uyp mu sxx syy sxy syx
  // XCD-aware order: the GPU deals consecutive blocks to its 8 XCDs in turn,
  // so block b runs on XCD b % 8; give each XCD a contiguous run of tiles
  // (neighbours in y, then the next tile row) so the halo lines two tiles
  // share are fetched into one L2
  int x0, y0;
  {
    const int nb = gridDim.x * gridDim.y, b = blockIdx.x + gridDim.x * blockIdx.y;
    const int q = nb / 8, rem = nb % 8, xcd = b % 8, idx = b / 8;
    const int t = (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + idx;
    x0 = (t / gridDim.x) * TX;
    y0 = (t % gridDim.x) * TY;
  }
  const int tid = threadIdx.x;
  const bool perx = g.bc <= 2;  // channels are periodic in x
  const int NP = g.NP;
  auto node = [&](int x, int y) { return (y + 1) + g.W * (x + 1); };

  // Q and phi of the tile + 2 (BoundaryConditionsFields, lyotropic.cpp:422-449)
  for (int i = tid; i < PN; i += kWallThreads) {
    const int r = i / PW, c = i - r * PW;
    const int x = wall_axis(x0 - 2 + r, g.LX, perx);
    const int y = wall_axis(y0 - 2 + c, g.LY, false);
    T q0 = T(0), q1 = T(0), ph = T(0);
    if (x >= 0 && y >= 0) {
      const int k = node(x, y);
      q0 = a.qxx[k];
      q1 = a.qyx[k];
      ph = a.phi[k];
    }
    sq[0][i] = q0;
    sq[1][i] = q1;
    sq[2][i] = ph;
  }
  __syncthreads();

  // this thread's node of the tile + 1
  const int r = tid / DW, c = tid - r * DW;
  const int i = tid;
  const bool tile = r >= 1 && r <= TX && c >= 1 && c <= TY;
  const int xv = x0 - 1 + r, yv = y0 - 1 + c;

  // UpdateQuantities (lyotropic.cpp:168-247) on the domain nodes of the tile + 1
  CenterVals<T> cv{};
  T fl[9];  // this node's populations, kept for UpdateFields
  {
    const bool xin = (xv >= 0 && xv < g.LX) || (perx && (xv == -1 || xv == g.LX));
    const bool yin = yv >= 0 && yv < g.LY;
    T vals[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
    if (xin && yin) {
      const int k = node(wall_axis(xv, g.LX, perx), yv);
#pragma unroll
      for (int v = 0; v < 9; ++v) fl[v] = a.f[v * NP + (PULL_F ? wnext(g, k, wopp(v)) : k)];
      const int q = (r + 1) * PW + (c + 1);
      NbrVals<T> nv;
      DiagVals<T> dg;
      update_quantities<T, true, true>(
          fl, sq[0][q], sq[0][q + PW], sq[0][q - PW], ynb<T, true>(sq[0][q + 1], sq[0][q - 1]),
          sq[1][q], sq[1][q + PW], sq[1][q - PW], ynb<T, true>(sq[1][q + 1], sq[1][q - 1]),
          sq[2][q], sq[2][q + PW], sq[2][q - PW], ynb<T, true>(sq[2][q + 1], sq[2][q - 1]), P, nv,
          cv, &dg);
      vals[0] = nv.ux;
      vals[1] = nv.uy;
      vals[2] = nv.uxp;
      vals[3] = nv.uyp;
      vals[4] = nv.mu;
      vals[5] = nv.sxx;
      vals[6] = nv.syy;
      vals[7] = nv.sxy;
      vals[8] = nv.syx;
      if constexpr (MODE == kSweepDerive) {
        if (tile && xv < g.LX) {
          const T out[kLoopPlanes] = {cv.nn,     nv.ux,     nv.uy,     nv.uxp,    nv.uyp,
                                      cv.Hxx,    cv.Hyx,    nv.mu,     cv.dxQxx,  cv.dxQyx,
                                      cv.dyQxx,  cv.dyQyx,  nv.sxx,    nv.syy,    nv.sxy,
                                      nv.syx,    dg.bulk,   dg.el_xx,  dg.el_yx,  dg.pf_xx,
                                      dg.pf_yx,  dg.act_xx, dg.act_yx};
#pragma unroll
          for (int pl = 0; pl < kLoopPlanes; ++pl) a.D[(size_t)pl * NP + k] = out[pl];
        }
      }
    }
#pragma unroll
    for (int pl = 0; pl < 9; ++pl) sd[pl][i] = vals[pl];
  }
  if constexpr (MODE == kSweepDerive) return;
  __syncthreads();

  // BoundaryConditionsFields2 at the walled nodes of the tile + 1 (the ring,
  // or past the domain's end inside a partial tile; fields.hpp:704-768): copy
  // derivative on u, Dirichlet 0 on u*phi, Neumann on MU and sigma.  Ghosts
  // are written and only domain nodes read.
  {
    const bool xdom = xv >= 0 && xv < g.LX, ydom = yv >= 0 && yv < g.LY;
    int d = 0;  // offset towards the domain
    if (xdom && (yv == -1 || yv == g.LY))
      d = yv == -1 ? 1 : -1;
    else if (!perx && ydom && (xv == -1 || xv == g.LX))
      d = xv == -1 ? DW : -DW;
    if (d != 0) {
      const int j1 = i + d, j2 = i + 2 * d;
      sd[0][i] = T(2) * sd[0][j1] - sd[0][j2];
      sd[1][i] = T(2) * sd[1][j1] - sd[1][j2];
      sd[2][i] = T(2) * T(0) - sd[2][j1];
      sd[3][i] = T(2) * T(0) - sd[3][j1];
#pragma unroll
      for (int pl = 4; pl < 9; ++pl) sd[pl][i] = sd[pl][j1];
    }
  }
  __syncthreads();

  // UpdateFields (lyotropic.cpp:249-341) on the tile
  if (!tile || xv >= g.LX || yv >= g.LY) return;
  const int k = node(xv, yv);
  const int p = i;
  cv.dn = T(0);
  UFStencil<T, true> st;
  st.mu_c = sd[4][p];
  st.ux_xp = sd[0][p + DW];
  st.ux_xm = sd[0][p - DW];
  st.uy_xp = sd[1][p + DW];
  st.uy_xm = sd[1][p - DW];
  st.uxp_xp = sd[2][p + DW];
  st.uxp_xm = sd[2][p - DW];
  st.sxx_xp = sd[5][p + DW];
  st.sxx_xm = sd[5][p - DW];
  st.syx_xp = sd[8][p + DW];
  st.syx_xm = sd[8][p - DW];
  st.mu_xp = sd[4][p + DW];
  st.mu_xm = sd[4][p - DW];
  st.ux_y = ynb<T, true>(sd[0][p + 1], sd[0][p - 1]);
  st.uy_y = ynb<T, true>(sd[1][p + 1], sd[1][p - 1]);
  st.uyp_y = ynb<T, true>(sd[3][p + 1], sd[3][p - 1]);
  st.sxy_y = ynb<T, true>(sd[7][p + 1], sd[7][p - 1]);
  st.syy_y = ynb<T, true>(sd[6][p + 1], sd[6][p - 1]);
  st.mu_y = ynb<T, true>(sd[4][p + 1], sd[4][p - 1]);
  const int q = (r + 1) * PW + (c + 1);
  const T Qxx = sq[0][q], Qyx = sq[1][q], ph = sq[2][q];
  T Dxx, Dyx, Dp, om[9];
  update_fields<T, true>(cv, Qxx, Qyx, ph, fl, st, P, Dxx, Dyx, Dp, om);
  if constexpr (MODE == kSweepPredictor) {
    a.qnxx_out[k] = Qxx + T(.5) * Dxx;
    a.qnyx_out[k] = Qyx + T(.5) * Dyx;
    a.phn_out[k] = ph + T(.5) * Dp;
    a.qxx_out[k] = Qxx + Dxx;
    a.qyx_out[k] = Qyx + Dyx;
    a.phi_out[k] = ph + Dp;
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
    for (int v = 0; v < 9; ++v)
      a.f_out[v * NP + k] = a.fn[v * NP + (PULL_FN ? wnext(g, k, wopp(v)) : k)] + T(.5) * om[v];
  }
}

// Inputs and outputs of one go-or-grow sweep on a periodic grid: the
// Lyotropic planes of a wall sweep plus the phenotype density and the masks.
struct GrowthSweep {
  WallSweep w;
  const double* m;    // go-or-grow: m of the inputs
  const double* chi;  // go-or-grow: chi = m/phi of the inputs
  const double* mn;   // corrector: m + Dm/2 of the predictor
  const unsigned char* division;
  const unsigned char* death;
  double* m_out;
  double* mn_out;     // predictor
  double* chi_out;    // chi = m/phi of the outputs
  const double* countphi;  // conserve_phi: device sum of phi of the inputs
  double totalphi;
  int conserve;
};

// A go-or-grow (DIV: lyotropic-with-division) sweep on a periodic grid as
// one tile kernel: the tile's Q and phi plus two nodes of halo staged in LDS
// with periodic wrap on both axes, UpdateQuantities (go-or-grow's, with chi
// at the node) on the tile plus one node into LDS, UpdateFields with the
// phenotype update on the tile.  chi and m at the neighbours are read from
// memory (chi is m/phi of the inputs: UpdatePhenotypeQuantities follows every
// UpdateQuantities and UpdateFields of the reference, go-or-grow.cpp:201-209,
// 305-362, so the sweep before wrote it with its outputs).  The node
// arithmetic is the loop kernels' device functions (growth_quantities,
// growth_dm, update_fields), so the results are theirs bit for bit.  Modes
// and pulls as wall_sweep_kernel's; kSweepDerive stores the 23 planes.
template <int MODE, bool PULL_F = false, bool PULL_FN = false, bool DIV = false>
__global__ void __launch_bounds__(kGrowThreads) __attribute__((amdgpu_waves_per_eu(4)))
growth_sweep_kernel(GrowthSweep a, Params<double> P, GrowthDev gd, WallGeo g) {
  using T = double;
  constexpr int TX = kGrowTX, TY = kWallTY;
  constexpr int PW = TY + 4, PN = (TX + 4) * PW;  // staged Q, phi: tile + 2
  constexpr int DW = TY + 2, DN = (TX + 2) * DW;  // derived: tile + 1 (= threads)
  __shared__ T sq[3][PN];
  __shared__ T sd[9][DN];  // ux uy uxp uyp mu sxx syy sxy syx
  int x0, y0;
  {
    const int nb = gridDim.x * gridDim.y, b = blockIdx.x + gridDim.x * blockIdx.y;
    const int q = nb / 8, rem = nb % 8, xcd = b % 8, idx = b / 8;
    const int t = (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + idx;
    x0 = (t / gridDim.x) * TX;
    y0 = (t % gridDim.x) * TY;
  }
  const int tid = threadIdx.x;
  const int NP = g.NP;
  auto wrap = [](int v, int L) { return v < 0 ? v + L : (v >= L ? v - L : v); };
  auto node = [&](int x, int y) { return (y + 1) + g.W * (x + 1); };
  const WallSweep& w = a.w;

  for (int i = tid; i < PN; i += kGrowThreads) {
    const int r = i / PW, c = i - r * PW;
    const int k = node(wall_axis(x0 - 2 + r, g.LX, true), wall_axis(y0 - 2 + c, g.LY, true));
    sq[0][i] = w.qxx[k];
    sq[1][i] = w.qyx[k];
    sq[2][i] = w.phi[k];
  }
  __syncthreads();

  // this thread's node of the tile + 1 (periodic image of the position)
  const int r = tid / DW, c = tid - r * DW;
  const bool tile = r >= 1 && r <= TX && c >= 1 && c <= TY;
  const int xv = x0 - 1 + r, yv = y0 - 1 + c;
  const int xw = wall_axis(xv, g.LX, true), yw = wall_axis(yv, g.LY, true);
  const int k = node(xw, yw);
  const int q = (r + 1) * PW + (c + 1);
  auto pulled = [&](const T* pl, int v, bool pull) {
    if (!pull) return pl[v * NP + k];
    return pl[v * NP + node(wrap(xw - kWallCX[v], g.LX), wrap(yw - kWallCY[v], g.LY))];
  };
  T fl[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) fl[v] = pulled(w.f, v, PULL_F);
  CenterVals<T> cv{};
  if constexpr (DIV) {
    NbrVals<T> nv;
    DiagVals<T> dg;
    update_quantities<T, true, true>(
        fl, sq[0][q], sq[0][q + PW], sq[0][q - PW], ynb<T, true>(sq[0][q + 1], sq[0][q - 1]),
        sq[1][q], sq[1][q + PW], sq[1][q - PW], ynb<T, true>(sq[1][q + 1], sq[1][q - 1]),
        sq[2][q], sq[2][q + PW], sq[2][q - PW], ynb<T, true>(sq[2][q + 1], sq[2][q - 1]), P, nv,
        cv, &dg);
    if constexpr (MODE == kSweepDerive) {
      if (tile && xv < g.LX && yv < g.LY) {
        const T out[kLoopPlanes] = {cv.nn,    nv.ux,    nv.uy,     nv.uxp,    nv.uyp,  cv.Hxx,
                                    cv.Hyx,   nv.mu,    cv.dxQxx,  cv.dxQyx,  cv.dyQxx, cv.dyQyx,
                                    nv.sxx,   nv.syy,   nv.sxy,    nv.syx,    dg.bulk, dg.el_xx,
                                    dg.el_yx, dg.pf_xx, dg.pf_yx, dg.act_xx, dg.act_yx};
#pragma unroll
        for (int pl = 0; pl < kLoopPlanes; ++pl) w.D[(size_t)pl * NP + k] = out[pl];
      }
      return;
    }
    const T vals[9] = {nv.ux, nv.uy, nv.uxp, nv.uyp, nv.mu, nv.sxx, nv.syy, nv.sxy, nv.syx};
#pragma unroll
    for (int pl = 0; pl < 9; ++pl) sd[pl][tid] = vals[pl];
  } else {
    T out[kLoopPlanes];
    growth_quantities(fl, sq[0][q], sq[0][q + PW], sq[0][q - PW], sq[0][q + 1], sq[0][q - 1],
                      sq[1][q], sq[1][q + PW], sq[1][q - PW], sq[1][q + 1], sq[1][q - 1],
                      sq[2][q], sq[2][q + PW], sq[2][q - PW], sq[2][q + 1], sq[2][q - 1],
                      a.chi[k], P, gd, out);
    if constexpr (MODE == kSweepDerive) {
      if (tile && xv < g.LX && yv < g.LY) {
#pragma unroll
        for (int pl = 0; pl < kLoopPlanes; ++pl) w.D[(size_t)pl * NP + k] = out[pl];
      }
      return;
    }
    cv.nn = out[LP_N];
    cv.vx = out[LP_UX];
    cv.vy = out[LP_UY];
    cv.Hxx = out[LP_HXX];
    cv.Hyx = out[LP_HYX];
    cv.dxQxx = out[LP_DXQXX];
    cv.dyQxx = out[LP_DYQXX];
    cv.dxQyx = out[LP_DXQYX];
    cv.dyQyx = out[LP_DYQYX];
    const T vals[9] = {out[LP_UX],  out[LP_UY],  out[LP_UXP], out[LP_UYP], out[LP_MU],
                       out[LP_SXX], out[LP_SYY], out[LP_SXY], out[LP_SYX]};
#pragma unroll
    for (int pl = 0; pl < 9; ++pl) sd[pl][tid] = vals[pl];
  }
  __syncthreads();

  // UpdateFields (lyotropic.cpp:249-341, go-or-grow.cpp:322-362) on the tile
  if (!tile || xv >= g.LX || yv >= g.LY) return;
  const int p = tid;
  cv.dn = T(0);
  UFStencil<T, true> st;
  st.mu_c = sd[4][p];
  st.ux_xp = sd[0][p + DW];
  st.ux_xm = sd[0][p - DW];
  st.uy_xp = sd[1][p + DW];
  st.uy_xm = sd[1][p - DW];
  st.uxp_xp = sd[2][p + DW];
  st.uxp_xm = sd[2][p - DW];
  st.sxx_xp = sd[5][p + DW];
  st.sxx_xm = sd[5][p - DW];
  st.syx_xp = sd[8][p + DW];
  st.syx_xm = sd[8][p - DW];
  st.mu_xp = sd[4][p + DW];
  st.mu_xm = sd[4][p - DW];
  st.ux_y = ynb<T, true>(sd[0][p + 1], sd[0][p - 1]);
  st.uy_y = ynb<T, true>(sd[1][p + 1], sd[1][p - 1]);
  st.uyp_y = ynb<T, true>(sd[3][p + 1], sd[3][p - 1]);
  st.sxy_y = ynb<T, true>(sd[7][p + 1], sd[7][p - 1]);
  st.syy_y = ynb<T, true>(sd[6][p + 1], sd[6][p - 1]);
  st.mu_y = ynb<T, true>(sd[4][p + 1], sd[4][p - 1]);
  const T Qxx = sq[0][q], Qyx = sq[1][q], ph = sq[2][q];

  // the growth source and the phenotype increment (go-or-grow.cpp:324-344;
  // lyotropic-with-division.cpp:110-126 with DIV)
  const T division_m = a.division[k] ? T(1.) : T(0.);
  const T death_m = a.death[k] ? T(1.) : T(0.);
  const T growth_rate = gd.alpha * division_m - gd.beta * death_m;
  T R, Dm = T(0);
  if constexpr (DIV) {
    R = ph * growth_rate;
  } else {
    const int kxp = node(wrap(xw + 1, g.LX), yw), kxm = node(wrap(xw - 1, g.LX), yw);
    const int kyp = node(xw, wrap(yw + 1, g.LY)), kym = node(xw, wrap(yw - 1, g.LY));
    const T chi_eff = a.chi[k];
    R = chi_eff * ph * growth_rate;
    Dm = growth_dm(chi_eff, a.chi[kxp], a.chi[kxm], a.chi[kyp], a.chi[kym], ph, sq[2][q + PW],
                   sq[2][q - PW], sq[2][q + 1], sq[2][q - 1], a.m[kxp], a.m[kxm], a.m[kyp],
                   a.m[kym], sd[0][p + DW], sd[0][p - DW], sd[1][p + 1], sd[1][p - 1], sd[4][p],
                   sd[4][p + DW], sd[4][p - DW], sd[4][p + 1], sd[4][p - 1], R, P, gd);
  }

  T Dxx, Dyx, Dp, om[9];
  update_fields<T, true>(cv, Qxx, Qyx, ph, fl, st, P, Dxx, Dyx, Dp, om);
  Dp = Dp - (a.conserve ? (*a.countphi - a.totalphi) / T(g.LX * g.LY) : T(0));
  T phi_new;
  if constexpr (MODE == kSweepPredictor) {
    w.qnxx_out[k] = Qxx + T(.5) * Dxx;
    w.qnyx_out[k] = Qyx + T(.5) * Dyx;
    w.phn_out[k] = ph + T(.5) * Dp;
    w.qxx_out[k] = Qxx + Dxx;
    w.qyx_out[k] = Qyx + Dyx;
    phi_new = ph + Dp;
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      w.fn_out[v * NP + k] = fl[v] + T(.5) * om[v];
      w.f_out[v * NP + k] = fl[v] + om[v];
    }
  } else {
    w.qxx_out[k] = w.qnxx[k] + T(.5) * Dxx;
    w.qyx_out[k] = w.qnyx[k] + T(.5) * Dyx;
    phi_new = w.phn[k] + T(.5) * Dp;
#pragma unroll
    for (int v = 0; v < 9; ++v) w.f_out[v * NP + k] = pulled(w.fn, v, PULL_FN) + T(.5) * om[v];
  }
  phi_new += R;
  w.phi_out[k] = phi_new;
  if constexpr (DIV) return;
  T mt;
  if constexpr (MODE == kSweepPredictor) {
    const T mk = a.m[k];
    a.mn_out[k] = mk + T(.5) * Dm;
    mt = mk + Dm;
  } else {
    mt = a.mn[k] + T(.5) * Dm;
  }
  const T upper = phi_new > 0 ? phi_new : T(0.);
  if (mt < 0)
    mt = 0;
  else if (mt > upper)
    mt = upper;
  a.m_out[k] = mt;
  a.chi_out[k] = phi_new > 1e-12 ? mt / phi_new : 0.;
}

// Move over every node of the torus (lyotropic.cpp:371-386), as a pull:
// node p takes direction v from p - c_v, the node whose push lands on p.
template <class T>
__global__ void __launch_bounds__(256) wall_move_kernel(WallGeo g, const T* __restrict__ f,
                                                        const T* __restrict__ fn,
                                                        T* __restrict__ f_to,
                                                        T* __restrict__ fn_to) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (g.LX + 2) * g.W) return;
  f_to[p] = f[p];
  fn_to[p] = fn[p];
#pragma unroll
  for (int v = 1; v < 9; ++v) {
    const int src = wnext(g, p, wopp(v));
    f_to[v * g.NP + p] = f[v * g.NP + src];
    fn_to[v * g.NP + p] = fn[v * g.NP + src];
  }
}

// Move on a periodic grid over the ghost-layer layout (no ghosts): every
// domain node pulls direction v from its periodic upstream neighbour.
template <class T>
__global__ void __launch_bounds__(256) wall_move_wrap_kernel(WallGeo g, const T* __restrict__ f,
                                                             const T* __restrict__ fn,
                                                             T* __restrict__ f_to,
                                                             T* __restrict__ fn_to) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.LX * g.LY) return;
  const int x = t / g.LY, y = t - x * g.LY;
  const int p = wnode(g, x + 1, y + 1);
  f_to[p] = f[p];
  fn_to[p] = fn[p];
#pragma unroll
  for (int v = 1; v < 9; ++v) {
    int xs = x - kWallCX[v], ys = y - kWallCY[v];
    xs = xs < 0 ? xs + g.LX : (xs >= g.LX ? xs - g.LX : xs);
    ys = ys < 0 ? ys + g.LY : (ys >= g.LY ? ys - g.LY : ys);
    const int src = wnode(g, xs + 1, ys + 1);
    f_to[v * g.NP + p] = f[v * g.NP + src];
    fn_to[v * g.NP + p] = fn[v * g.NP + src];
  }
}

// The domain nodes with phi >= 1/2, in node order (the nodes the mask draws
// visit, lyotropic-with-division.cpp:42-50): select_count_kernel counts each
// block's; select_write_kernel adds up the earlier blocks' counts and writes,
// at that offset and in order within the block, each node's domain index
// and the probabilities its draw is compared with: the division rate
// rate_div*phi*(1 - phi/phi_critical) into adj[pos], the death rate
// rate_death*phi into adj[n + pos] (:46, :82; FP64 without contraction, as
// the host would compute them).
constexpr int kSelBlock = 1024;
__device__ __forceinline__ bool select_node(const WallGeo& g, const double* phi, int t) {
  return t < g.LX * g.LY && phi[LayerGrid::at(t, g.LX, g.LY).k] >= .5;
}
static __global__ void __launch_bounds__(kSelBlock) select_count_kernel(WallGeo g,
                                                                        const double* __restrict__ phi,
                                                                        int* __restrict__ counts) {
  __shared__ int ws[kSelBlock / 64];
  const bool e = select_node(g, phi, blockIdx.x * kSelBlock + threadIdx.x);
  const unsigned long long b = __ballot(e);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int w = 0; w < kSelBlock / 64; ++w) c += ws[w];
    counts[blockIdx.x] = c;
  }
}
static __global__ void __launch_bounds__(kSelBlock) select_write_kernel(
    WallGeo g, const double* __restrict__ phi, const int* __restrict__ counts,
    int* __restrict__ idx, double* __restrict__ adj, double rate_div, double rate_death,
    double phi_critical, int* __restrict__ total) {
  __shared__ int ws[kSelBlock / 64];
  __shared__ int part[kSelBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // offset of this block: the counts of blocks 0 .. blockIdx.x - 1
  int o = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += kSelBlock) o += counts[b];
  for (int d = 32; d > 0; d >>= 1) o += __shfl_xor(o, d);
  if (lane == 0) part[w] = o;
  const int t = blockIdx.x * kSelBlock + threadIdx.x;
  const bool e = select_node(g, phi, t);
  const unsigned long long b = __ballot(e);
  if (lane == 0) ws[w] = __popcll(b);
  __syncthreads();
  int base = 0, before = 0, all = 0;
  for (int i = 0; i < kSelBlock / 64; ++i) {
    base += part[i];
    before += i < w ? ws[i] : 0;
    all += ws[i];
  }
  if (e) {
    const unsigned long long lower = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int pos = base + before + __popcll(b & lower);
    idx[pos] = t;
    const double p = phi[LayerGrid::at(t, g.LX, g.LY).k];
    adj[pos] = rate_div * p * (1 - p / phi_critical);
    adj[g.LX * g.LY + pos] = rate_death * p;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = base + all;
}

// Host FP64 domain array -> the domain of a torus plane.
static __global__ void wall_scatter_kernel(WallGeo g, const double* __restrict__ in,
                                           double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.LX * g.LY) return;
  out[LayerGrid::at(t, g.LX, g.LY).k] = in[t];
}

// Domain of a 0/1 byte mask -> FP64.
static __global__ void wall_unpack_mask_kernel(WallGeo g, const unsigned char* __restrict__ mask,
                                               double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.LX * g.LY) return;
  out[t] = mask[LayerGrid::at(t, g.LX, g.LY).k] ? 1.0 : 0.0;
}

// Host FP64 AoS domain state -> torus planes (ghost nodes stay as they are).
template <class T>
__global__ void wall_pack_kernel(WallGeo g, const double* __restrict__ ff,
                                 const double* __restrict__ qxx, const double* __restrict__ qyx,
                                 const double* __restrict__ phi, T* f, T* oqxx, T* oqyx, T* ophi) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.LX * g.LY) return;
  const int p = LayerGrid::at(t, g.LX, g.LY).k;
#pragma unroll
  for (int v = 0; v < 9; ++v) f[v * g.NP + p] = (T)ff[9 * (size_t)t + v];
  oqxx[p] = (T)qxx[t];
  oqyx[p] = (T)qyx[t];
  ophi[p] = (T)phi[t];
}

// Domain of torus planes -> FP64 (nplanes = 9: AoS populations; 1: scalar).
template <class T>
__global__ void wall_unpack_kernel(WallGeo g, const T* __restrict__ a, int nplanes,
                                   double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.LX * g.LY) return;
  const int p = LayerGrid::at(t, g.LX, g.LY).k;
  for (int v = 0; v < nplanes; ++v) out[nplanes * (size_t)t + v] = (double)a[v * g.NP + p];
}

// Torus planes <-> contiguous (LX+2)*(LY+2) host-order planes (checkpoints).
template <class T>
__global__ void wall_copy_planes_kernel(const T* __restrict__ src, int src_stride,
                                        T* __restrict__ dst, int dst_stride, int n, int nplanes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int v = 0; v < nplanes; ++v) dst[(size_t)v * dst_stride + i] = src[(size_t)v * src_stride + i];
}

// Block b sums domain nodes [b*per, min((b+1)*per, N)) of sum_v f and phi in
// a fixed order (the partial sums of lyo_aux.hpp on the torus layout).
template <class T>
__global__ void __launch_bounds__(256)
wall_partial_sums_kernel(WallGeo g, const T* __restrict__ f, const T* __restrict__ phi, int per,
                         double* __restrict__ part_f, double* __restrict__ part_p) {
  __shared__ double sf[4], sp[4];
  const int N = g.LX * g.LY;
  const int beg = blockIdx.x * per;
  const int end = min(beg + per, N);
  double af = 0, ap = 0;
  for (int t = beg + (int)threadIdx.x; t < end; t += 256) {
    const int p = LayerGrid::at(t, g.LX, g.LY).k;
    double s = 0;
#pragma unroll
    for (int v = 0; v < 9; ++v) s += (double)f[v * g.NP + p];
    af += s;
    ap += (double)phi[p];
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

// The run of a walled grid.  FP64 only (the reference's arithmetic); the
// fused strip-march kernels, batches, the surrogate and the diagnostics
// assume a periodic grid and are not offered here.
class WallEngine final : public Engine {
 public:
  using T = double;
  explicit WallEngine(int bc) : bc_(bc) {}
  ~WallEngine() override { release(); }

  hipError_t alloc(int LX, int LY, int NP, int npc) override {
    (void)NP;
    LX_ = LX;
    LY_ = LY;
    npc_ = npc;
    TP_ = (LX + 2) * (LY + 2);
    // planes on 256-byte boundaries, staggered by 256 values like the
    // periodic layout
    NPw_ = ((TP_ + 256 + 63) / 64) * 64;
    const size_t n1 = (size_t)NPw_, n9 = 9 * n1;
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      if (e == hipSuccess) e = get_(n9, &F_[i]);
      if (e == hipSuccess) e = get_(n9, &Fn_[i]);
      if (e == hipSuccess) e = get_(n1, &phi_[i]);
    }
    if (e == hipSuccess) e = get_(n1, &qxx_);
    if (e == hipSuccess) e = get_(n1, &qyx_);
    if (e == hipSuccess) e = get_(n1, &qnxx_);
    if (e == hipSuccess) e = get_(n1, &qnyx_);
    if (e == hipSuccess) e = get_(n1, &phn_);
    if (e == hipSuccess) e = get_(n1, &qxx2_);
    if (e == hipSuccess) e = get_(n1, &qyx2_);
    if (e == hipSuccess) e = get_((size_t)kLoopPlanes * n1, &D_);
    return e;
  }

  void release() override {
    drop_graphs();
    for (T* p : ptrs_) (void)hipFree(p);
    ptrs_.clear();
    sizes_.clear();
    bytes_ = 0;
    F_[0] = F_[1] = Fn_[0] = Fn_[1] = phi_[0] = phi_[1] = nullptr;
    qxx_ = qyx_ = qnxx_ = qnyx_ = phn_ = D_ = qxx2_ = qyx2_ = nullptr;
    m_[0] = m_[1] = mn_ = chi_ = chi2_ = part_ = stage_ = fe_ = nullptr;
    if (div_) (void)hipFree(div_);
    if (death_) (void)hipFree(death_);
    // the centre copies of the last draws may still be in flight
    if (hcent_[0] || hcent_[1]) (void)hipDeviceSynchronize();
    for (int i = 0; i < 2; ++i) {
      if (cent_[i]) (void)hipFree(cent_[i]);
      if (hcent_[i]) (void)hipHostFree(hcent_[i]);
      cent_[i] = nullptr;
      hcent_[i] = nullptr;
      cent_cap_[i] = 0;
    }
    if (sel_idx_) (void)hipFree(sel_idx_);
    if (sel_val_) (void)hipFree(sel_val_);
    if (hsel_idx_) (void)hipHostFree(hsel_idx_);
    if (hsel_val_) (void)hipHostFree(hsel_val_);
    if (hpend_) (void)hipHostFree(hpend_);
    sel_idx_ = hsel_idx_ = nullptr;
    sel_val_ = hsel_val_ = nullptr;
    hpend_ = nullptr;
    npending_ = 0;
    sel_last_ = 0;
    div_ = death_ = nullptr;
    growth_ = false;
  }

  void set_params(const double* v) override {
    drop_graphs();
    P_.GammaQ = v[0];
    P_.GammaP = v[1];
    P_.xi = v[2];
    P_.tauNem = v[3];
    P_.tauIso = v[4];
    P_.rho = v[5];
    P_.friction = v[6];
    P_.AA = v[7];
    P_.CC = v[8];
    P_.LL = v[9];
    P_.KK = v[10];
    P_.zeta = v[11];
    P_.zetaI = v[12];
    P_.conc = v[13];
    finish_params(P_);
  }

  // Lyotropic::Step (lyotropic.cpp:500-525); go-or-grow redraws its masks
  // first (lyotropic-with-division.cpp:150-156)
  // The device part of a step is replayed from a HIP graph (captured on first
  // use for each parity of the ping-pong buffers; dropped whenever buffers,
  // parameters or the state change): a step is ~25 launches, which at small
  // grids take longer to issue than to run.  The graph holds kernels only.
  hipError_t steps(int nsteps, const LaunchCfg& L, hipStream_t s, SweepTimer* tm) override {
    (void)tm;
    // periodic go-or-grow runs as fused sweeps (growth_step); on walled
    // grids it keeps the node loops
    loops_ = L.loops != 0 || (growth_ && bc_ != 0);
    for (int st = 0; st < nsteps; ++st) {
      if (growth_) {
        const hipError_t e = set_masks(s);
        if (e != hipSuccess) return e;
      }
      const hipError_t e = graph_step(s);
      if (e != hipSuccess) return e;
      d_stale_ = !loops_;  // the node loops store the derived planes as they go
    }
    return hipSuccess;
  }

  // the walled step as three fused sweeps (wall_sweep_kernel) with the LB
  // conditions and Move between the predictor and the correctors
  hipError_t fused_step(hipStream_t s) {
    const dim3 grid = sweep_grid();
    const int bs = 256;
    WallSweep a{};
    a.f = F_[cur_];
    a.qxx = qxx_;
    a.qyx = qyx_;
    a.phi = phi_[pc_];
    a.f_out = F_[cur_ ^ 1];
    a.fn_out = Fn_[cur_ ^ 1];
    a.qxx_out = qxx2_;
    a.qyx_out = qyx2_;
    a.qnxx_out = qnxx_;
    a.qnyx_out = qnyx_;
    a.phi_out = phi_[pc_ ^ 1];
    a.phn_out = phn_;
    hipLaunchKernelGGL((wall_sweep_kernel<kSweepPredictor>), grid, dim3(kWallThreads), 0, s, a,
                       P_, geo());
    swap_q();
    pc_ ^= 1;
    // the LB conditions on the post-collision populations; the first
    // corrector streams f as it reads (writing its f to F[cur], which no
    // block pulls from), the later ones read plain f; all stream fn
    for (int ph : phases())
      hipLaunchKernelGGL((wall_lb_kernel<T>), dim3(phase_blocks(ph)), dim3(bs), 0, s, geo(), ph,
                         F_[cur_ ^ 1], Fn_[cur_ ^ 1]);
    int fc = cur_ ^ 1;  // the corrector's f input
    for (int c = 1; c <= npc_; ++c) {
      WallSweep b{};
      b.f = F_[fc];
      b.qxx = qxx_;
      b.qyx = qyx_;
      b.phi = phi_[pc_];
      b.fn = Fn_[cur_ ^ 1];
      b.qnxx = qnxx_;
      b.qnyx = qnyx_;
      b.phn = phn_;
      b.f_out = F_[fc ^ 1];
      b.qxx_out = qxx2_;
      b.qyx_out = qyx2_;
      b.phi_out = phi_[pc_ ^ 1];
      if (c == 1)
        hipLaunchKernelGGL((wall_sweep_kernel<kSweepCorrector, true, true>), grid,
                           dim3(kWallThreads), 0, s, b, P_, geo());
      else
        hipLaunchKernelGGL((wall_sweep_kernel<kSweepCorrector, false, true>), grid,
                           dim3(kWallThreads), 0, s, b, P_, geo());
      swap_q();
      pc_ ^= 1;
      fc ^= 1;
    }
    cur_ = fc;
    d_stale_ = true;  // the frame fields are recomputed from the last corrector's inputs
    return hipGetLastError();
  }
  // go-or-grow (or lyotropic-with-division) on a periodic grid as fused
  // sweeps: the predictor writes the post-collision planes, every corrector
  // pulls the streamed fn from them (the first one f too), so no Move runs;
  // conserve_phi sums phi before each sweep
  hipError_t growth_step(hipStream_t s) {
    const dim3 grid = growth_grid();
    const bool div_only = gp_.division_only != 0;
    auto countphi = [&]() {
      if (!conserve_) return;
      const int N = LX_ * LY_;
      int nparts = (N + 4095) / 4096;
      if (nparts > kGrowthParts) nparts = kGrowthParts;
      const int per = (N + nparts - 1) / nparts;
      hipLaunchKernelGGL((wall_partial_sums_kernel<T>), dim3(nparts), dim3(256), 0, s, geo(),
                         (const T*)F_[cur_], (const T*)phi_[pc_], per, part_,
                         part_ + kGrowthParts);
      hipLaunchKernelGGL(final_sums_kernel, dim3(1), dim3(256), 0, s, part_,
                         part_ + kGrowthParts, nparts, 0.0, part_ + 2 * kGrowthParts);
    };
    auto common = [&](GrowthSweep& a) {
      a.division = div_;
      a.death = death_;
      a.countphi = part_ + 2 * kGrowthParts + 1;
      a.totalphi = totalphi_;
      a.conserve = conserve_ ? 1 : 0;
      if (!div_only) {
        a.m = m_[mc_];
        a.chi = chi_;
        a.m_out = m_[mc_ ^ 1];
        a.chi_out = chi2_;
      }
    };
    auto advance = [&]() {
      swap_q();
      pc_ ^= 1;
      if (!div_only) {
        mc_ ^= 1;
        swap_chi();
      }
    };
    countphi();
    GrowthSweep a{};
    common(a);
    a.w.f = F_[cur_];
    a.w.qxx = qxx_;
    a.w.qyx = qyx_;
    a.w.phi = phi_[pc_];
    a.w.f_out = F_[cur_ ^ 1];
    a.w.fn_out = Fn_[cur_ ^ 1];
    a.w.qxx_out = qxx2_;
    a.w.qyx_out = qyx2_;
    a.w.qnxx_out = qnxx_;
    a.w.qnyx_out = qnyx_;
    a.w.phi_out = phi_[pc_ ^ 1];
    a.w.phn_out = phn_;
    a.mn_out = div_only ? nullptr : mn_;
    if (div_only)
      hipLaunchKernelGGL((growth_sweep_kernel<kSweepPredictor, false, false, true>), grid,
                         dim3(kGrowThreads), 0, s, a, P_, gd_, geo());
    else
      hipLaunchKernelGGL((growth_sweep_kernel<kSweepPredictor>), grid, dim3(kGrowThreads), 0, s,
                         a, P_, gd_, geo());
    advance();
    int fc = cur_ ^ 1;  // the corrector's f input
    for (int c = 1; c <= npc_; ++c) {
      countphi();
      GrowthSweep b{};
      common(b);
      b.w.f = F_[fc];
      b.w.qxx = qxx_;
      b.w.qyx = qyx_;
      b.w.phi = phi_[pc_];
      b.w.fn = Fn_[cur_ ^ 1];
      b.w.qnxx = qnxx_;
      b.w.qnyx = qnyx_;
      b.w.phn = phn_;
      b.w.f_out = F_[fc ^ 1];
      b.w.qxx_out = qxx2_;
      b.w.qyx_out = qyx2_;
      b.w.phi_out = phi_[pc_ ^ 1];
      b.mn = div_only ? nullptr : mn_;
      if (div_only) {
        if (c == 1)
          hipLaunchKernelGGL((growth_sweep_kernel<kSweepCorrector, true, true, true>), grid,
                             dim3(kGrowThreads), 0, s, b, P_, gd_, geo());
        else
          hipLaunchKernelGGL((growth_sweep_kernel<kSweepCorrector, false, true, true>), grid,
                             dim3(kGrowThreads), 0, s, b, P_, gd_, geo());
      } else if (c == 1) {
        hipLaunchKernelGGL((growth_sweep_kernel<kSweepCorrector, true, true>), grid,
                           dim3(kGrowThreads), 0, s, b, P_, gd_, geo());
      } else {
        hipLaunchKernelGGL((growth_sweep_kernel<kSweepCorrector, false, true>), grid,
                           dim3(kGrowThreads), 0, s, b, P_, gd_, geo());
      }
      advance();
      fc ^= 1;
    }
    cur_ = fc;
    d_stale_ = true;  // the frame fields are recomputed from the last corrector's inputs
    return hipGetLastError();
  }
  dim3 sweep_grid() const {
    return dim3((unsigned)((LY_ + kWallTY - 1) / kWallTY), (unsigned)((LX_ + kWallTX - 1) / kWallTX));
  }
  dim3 growth_grid() const {
    return dim3((unsigned)((LY_ + kWallTY - 1) / kWallTY), (unsigned)((LX_ + kGrowTX - 1) / kGrowTX));
  }
  // The derived planes of the last UpdateQuantities of a fused step, from its
  // inputs (f, Q and phi before the last corrector, left intact by it)
  hipError_t materialize(hipStream_t s) {
    if (!d_stale_) return hipSuccess;
    if (growth_) {
      // the last corrector's inputs: its f (the streamed post-collision
      // planes with npc = 1), Q, phi, m and chi, all left intact by it
      GrowthSweep a{};
      a.w.f = F_[cur_ ^ 1];
      a.w.qxx = qxx2_;
      a.w.qyx = qyx2_;
      a.w.phi = phi_[pc_ ^ 1];
      a.w.D = D_;
      if (gp_.division_only) {
        if (npc_ == 1)
          hipLaunchKernelGGL((growth_sweep_kernel<kSweepDerive, true, false, true>), growth_grid(),
                             dim3(kGrowThreads), 0, s, a, P_, gd_, geo());
        else
          hipLaunchKernelGGL((growth_sweep_kernel<kSweepDerive, false, false, true>),
                             growth_grid(), dim3(kGrowThreads), 0, s, a, P_, gd_, geo());
      } else {
        a.m = m_[mc_ ^ 1];
        a.chi = chi2_;
        if (npc_ == 1)
          hipLaunchKernelGGL((growth_sweep_kernel<kSweepDerive, true>), growth_grid(),
                             dim3(kGrowThreads), 0, s, a, P_, gd_, geo());
        else
          hipLaunchKernelGGL((growth_sweep_kernel<kSweepDerive, false>), growth_grid(),
                             dim3(kGrowThreads), 0, s, a, P_, gd_, geo());
      }
      d_stale_ = false;
      return hipGetLastError();
    }
    WallSweep a{};
    a.f = F_[cur_ ^ 1];  // with npc = 1: the post-collision planes, streamed as read
    a.qxx = qxx2_;
    a.qyx = qyx2_;
    a.phi = phi_[pc_ ^ 1];
    a.D = D_;
    if (npc_ == 1)
      hipLaunchKernelGGL((wall_sweep_kernel<kSweepDerive, true>), sweep_grid(), dim3(kWallThreads),
                         0, s, a, P_, geo());
    else
      hipLaunchKernelGGL((wall_sweep_kernel<kSweepDerive, false>), sweep_grid(),
                         dim3(kWallThreads), 0, s, a, P_, geo());
    d_stale_ = false;
    return hipGetLastError();
  }
  void swap_q() {
    std::swap(qxx_, qxx2_);
    std::swap(qyx_, qyx2_);
    qa_ ^= 1;
  }
  void swap_chi() {
    std::swap(chi_, chi2_);
    ca_ ^= 1;
  }

  hipError_t step_body(hipStream_t s) {
    if (!loops_) return growth_ ? growth_step(s) : fused_step(s);
    const int bs = 256, nbt = (TP_ + bs - 1) / bs;
    {
      fields_bc(s);
      update_quantities(s);
      fields2_bc(s);
      update_fields(true, s);
      for (int ph : phases())
        hipLaunchKernelGGL((wall_lb_kernel<T>), dim3(phase_blocks(ph)), dim3(bs), 0, s, geo(),
                           ph, F_[cur_], Fn_[cur_]);
      if (bc_ == 0)
        hipLaunchKernelGGL((wall_move_wrap_kernel<T>), dim3(node_blocks()), dim3(bs), 0, s,
                           geo(), (const T*)F_[cur_], (const T*)Fn_[cur_], F_[cur_ ^ 1],
                           Fn_[cur_ ^ 1]);
      else
        hipLaunchKernelGGL((wall_move_kernel<T>), dim3(nbt), dim3(bs), 0, s, geo(),
                           (const T*)F_[cur_], (const T*)Fn_[cur_], F_[cur_ ^ 1],
                           Fn_[cur_ ^ 1]);
      cur_ ^= 1;
      // go-or-grow: UpdateFields ended with UpdatePhenotypeQuantities (chi =
      // m/phi, pointwise) and the field conditions, and Move touches only f,
      // so a corrector's leading field conditions and its UpdateQuantities'
      // phenotype pass would rewrite the values already there: skipped
      const bool gog = growth_ && !gp_.division_only;
      for (int c = 1; c <= npc_; ++c) {
        if (!gog) fields_bc(s);
        update_quantities(s, !gog);
        fields2_bc(s);
        update_fields(false, s);
      }
    }
    return hipGetLastError();
  }

  hipError_t graph_step(hipStream_t s) {
    for (const StepGraph& g : graphs_) {
      if (g.cur == cur_ && g.pc == pc_ && g.mc == mc_ && g.qa == qa_ && g.ca == ca_ &&
          g.loops == loops_) {
        cur_ = g.cur_after;
        pc_ = g.pc_after;
        mc_ = g.mc_after;
        if (g.qa_after != qa_) swap_q();
        if (g.ca_after != ca_) swap_chi();
        return hipGraphLaunch(g.exec, s);
      }
    }
    StepGraph g{cur_, pc_, mc_, qa_, ca_, loops_, 0, 0, 0, 0, 0, nullptr};
    const int cur0 = cur_, pc0 = pc_, mc0 = mc_, qa0 = qa_, ca0 = ca_;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const hipError_t e1 = step_body(s);
    e = hipStreamEndCapture(s, &graph);
    if (e == hipSuccess) e = e1;
    if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      cur_ = cur0;
      pc_ = pc0;
      mc_ = mc0;
      if (qa_ != qa0) swap_q();
      if (ca_ != ca0) swap_chi();
      return e;
    }
    g.cur_after = cur_;
    g.pc_after = pc_;
    g.mc_after = mc_;
    g.qa_after = qa_;
    g.ca_after = ca_;
    graphs_.push_back(g);
    return hipGraphLaunch(g.exec, s);
  }

  // Every buffer restarts at zero (ghost layer included, as after
  // Field::SetSize), then the domain takes the given state.
  hipError_t pack(const double* d_ff, const double* d_qxx, const double* d_qyx,
                  const double* d_phi, hipStream_t s) override {
    drop_graphs();
    hipError_t e = zero_all(s);
    if (e != hipSuccess) return e;
    cur_ = 0;
    pc_ = 0;
    mc_ = 0;
    d_stale_ = false;
    const int N = LX_ * LY_, bs = 256;
    hipLaunchKernelGGL((wall_pack_kernel<T>), dim3((N + bs - 1) / bs), dim3(bs), 0, s, geo(), d_ff,
                       d_qxx, d_qyx, d_phi, F_[0], qxx_, qyx_, phi_[0]);
    return hipGetLastError();
  }

  // mass_lyo_field numbering; derived fields are the planes of the last
  // UpdateQuantities call (zero before the first step)
  hipError_t field(int id, double* d_out, hipStream_t s) override {
    if (id >= D_N && id <= D_HYX) {
      const hipError_t e = materialize(s);
      if (e != hipSuccess) return e;
    }
    const T* a = nullptr;
    int np = 1;
    switch (id) {
      case 0: a = F_[cur_]; np = 9; break;
      case 1: a = qxx_; break;
      case 2: a = qyx_; break;
      case 3: a = phi_[pc_]; break;
      case D_N: a = plane(LP_N); break;
      case D_UX: a = plane(LP_UX); break;
      case D_UY: a = plane(LP_UY); break;
      case D_SXX: a = plane(LP_SXX); break;
      case D_SXY: a = plane(LP_SXY); break;
      case D_SYX: a = plane(LP_SYX); break;
      case D_SYY: a = plane(LP_SYY); break;
      case D_BULK: a = plane(LP_BULK); break;
      case D_EL_XX: a = plane(LP_ELXX); break;
      case D_EL_YX: a = plane(LP_ELYX); break;
      case D_PF_XX: a = plane(LP_PFXX); break;
      case D_PF_YX: a = plane(LP_PFYX); break;
      case D_ACT_XX: a = plane(LP_ACTXX); break;
      case D_ACT_YX: a = plane(LP_ACTYX); break;
      case D_MU: a = plane(LP_MU); break;
      case D_HXX: a = plane(LP_HXX); break;
      case D_HYX: a = plane(LP_HYX); break;
      case MASS_LYO_M: a = growth_ && !gp_.division_only ? m_[mc_] : nullptr; break;
      case MASS_LYO_CHI: a = growth_ && !gp_.division_only ? chi_ : nullptr; break;
      case MASS_LYO_DIVISION_MASK:
      case MASS_LYO_DEATH_MASK: {
        if (!growth_) return hipErrorInvalidValue;
        const int N = LX_ * LY_;
        hipLaunchKernelGGL(wall_unpack_mask_kernel, dim3((N + 255) / 256), dim3(256), 0, s, geo(),
                           (const unsigned char*)(id == MASS_LYO_DIVISION_MASK ? div_ : death_),
                           d_out);
        return hipGetLastError();
      }
      default: return hipErrorInvalidValue;
    }
    if (!a) return hipErrorInvalidValue;
    const int N = LX_ * LY_, bs = 256;
    hipLaunchKernelGGL((wall_unpack_kernel<T>), dim3((N + bs - 1) / bs), dim3(bs), 0, s, geo(), a,
                       np, d_out);
    return hipGetLastError();
  }

  hipError_t sums(double* d_part, int max_part, double* d_out, hipStream_t s) override {
    const int N = LX_ * LY_;
    int nparts = (N + 4095) / 4096;
    if (nparts > max_part) nparts = max_part;
    const int per = (N + nparts - 1) / nparts;
    hipLaunchKernelGGL((wall_partial_sums_kernel<T>), dim3(nparts), dim3(256), 0, s, geo(),
                       F_[cur_], phi_[pc_], per, d_part, d_part + max_part);
    hipLaunchKernelGGL(final_sums_kernel, dim3(1), dim3(256), 0, s, d_part, d_part + max_part,
                       nparts, 0.0, d_out);
    return hipGetLastError();
  }

  hipError_t diagnostics(DiagScratch&, double*, hipStream_t) override {
    return hipErrorNotSupported;
  }
  size_t bytes() const override { return bytes_; }

  std::vector<int> plan(const LaunchCfg&, int) const override { return {}; }
  size_t job_bytes() const override { return 0; }
  int job_items(int, const LaunchCfg&) const override { return 0; }
  void fill_job(int, int, int, const LaunchCfg&, void*) const override {}
  void after_batch_steps(int, const std::vector<int>&) override {}
  int parity() const override { return cur_; }
  hipError_t launch_batch(const void*, int, int, int, bool, int, int,
                          hipStream_t) const override {
    return hipErrorNotSupported;
  }
  hipError_t net_features(int, float*, const float*, __bf16*, hipStream_t) override {
    return hipErrorNotSupported;
  }
  hipError_t net_targets(float*, hipStream_t) override { return hipErrorNotSupported; }
  hipError_t net_apply(const float*, hipStream_t) override { return hipErrorNotSupported; }
  hipError_t net_judge(const float*, const NetCalib&, double*, int, hipStream_t) override {
    return hipErrorNotSupported;
  }
  hipError_t steps_with(int, int, const LaunchCfg&, hipStream_t) override {
    return hipErrorNotSupported;
  }
  int npc() const override { return npc_; }
  hipError_t alloc_loops() override { return hipSuccess; }

  // checkpoint state: f (9), Qxx, Qyx, phi on the whole torus, ghost layer
  // included, each plane (LX+2)*(LY+2) values
  int state_planes() const override { return 12; }
  size_t state_bytes() const override { return 12 * (size_t)TP_ * sizeof(T); }
  hipError_t save_state(void* host, hipStream_t s) override {
    return copy_state(host, nullptr, s);
  }
  hipError_t load_state(const void* host, hipStream_t s) override {
    drop_graphs();
    hipError_t e = zero_all(s);
    if (e != hipSuccess) return e;
    cur_ = 0;
    pc_ = 0;
    d_stale_ = false;
    return copy_state(nullptr, host, s);
  }

 private:
  WallGeo geo() const { return WallGeo{LX_, LY_, LY_ + 2, NPw_, bc_}; }
  const T* plane(int i) const { return D_ + (size_t)i * NPw_; }
  // the launches of one boundary-condition call: the three phases in order;
  // none on a periodic grid (bc 0), whose kernels take the periodic
  // neighbours inside the domain (LayerWrapGrid) and leave the ghosts unused
  std::vector<int> phases() const { return bc_ == 0 ? std::vector<int>{} : std::vector<int>{0, 1, 2}; }
  // runs f with the node-index policy of this grid
  template <class F>
  void by_grid(F&& f) const {
    if (bc_ == 0)
      f(LayerWrapGrid{});
    else
      f(LayerGrid{});
  }
  unsigned phase_blocks(int phase) const {
    const int n = phase == 0 ? 2 * LY_ : (phase == 1 ? 2 * LX_ : 4);
    return (unsigned)((n + 255) / 256);
  }
  // BoundaryConditionsFields (lyotropic.cpp:422-449; go-or-grow.cpp:365-385
  // adds m and chi)
  void fields_bc(hipStream_t s, bool with_growth = true) {
    const bool g = growth_ && !gp_.division_only && with_growth;
    for (int ph : phases())
      hipLaunchKernelGGL((wall_fields_kernel<T>), dim3(phase_blocks(ph)), dim3(256), 0, s, geo(),
                         ph, qxx_, qyx_, phi_[pc_], g ? m_[mc_] : (T*)nullptr,
                         g ? chi_ : (T*)nullptr);
  }
  unsigned node_blocks() const { return (unsigned)((LX_ * LY_ + 255) / 256); }
  // UpdatePhenotypeQuantities (go-or-grow.cpp:201-209)
  void phenotype(hipStream_t s) {
    by_grid([&](auto gg) {
      using G = decltype(gg);
      hipLaunchKernelGGL((phenotype_kernel<G>), dim3(node_blocks()), dim3(256), 0, s,
                         (const T*)m_[mc_], (const T*)phi_[pc_], chi_, LX_, LY_);
    });
    fields_bc(s);
  }
  // UpdateQuantities (lyotropic.cpp:343-360; go-or-grow.cpp:305-320);
  // with_phenotype = false leaves out the closing phenotype pass where chi is
  // already m/phi of the current fields
  void update_quantities(hipStream_t s, bool with_phenotype = true) {
    const T* f = F_[cur_];
    const bool gog = growth_ && !gp_.division_only;
    by_grid([&](auto gg) {
      using G = decltype(gg);
      if (!gog)
        hipLaunchKernelGGL((loop_uq_kernel<T, true, G>), dim3(node_blocks()), dim3(256), 0, s, f,
                           qxx_, qyx_, phi_[pc_], P_, LX_, LY_, NPw_, D_);
      else
        hipLaunchKernelGGL((growth_uq_kernel<G>), dim3(node_blocks()), dim3(256), 0, s, f,
                           (const T*)qxx_, (const T*)qyx_, (const T*)phi_[pc_], (const T*)chi_,
                           P_, gd_, LX_, LY_, NPw_, D_);
    });
    if (growth_ && conserve_) {  // countphi: the sum of phi this UpdateQuantities sees
      const int N = LX_ * LY_;
      int nparts = (N + 4095) / 4096;
      if (nparts > kGrowthParts) nparts = kGrowthParts;
      const int per = (N + nparts - 1) / nparts;
      hipLaunchKernelGGL((wall_partial_sums_kernel<T>), dim3(nparts), dim3(256), 0, s, geo(),
                         (const T*)F_[cur_], (const T*)phi_[pc_], per, part_,
                         part_ + kGrowthParts);
      hipLaunchKernelGGL(final_sums_kernel, dim3(1), dim3(256), 0, s, part_,
                         part_ + kGrowthParts, nparts, 0.0, part_ + 2 * kGrowthParts);
    }
    if (gog && with_phenotype) phenotype(s);
  }
  // UpdateFields (lyotropic.cpp:362-369; go-or-grow.cpp:322-362)
  void update_fields(bool first, hipStream_t s) {
    T* f = F_[cur_];
    LoopUF<T> u{};
    u.D = D_;
    u.f = f;
    u.qxx = qxx_;
    u.qyx = qyx_;
    u.phi = phi_[pc_];
    u.f_out = f;  // in place, as UpdateFieldsAtNode writes ff[k]
    u.qxx_out = qxx_;
    u.qyx_out = qyx_;
    u.phi_out = phi_[pc_ ^ 1];
    if (first) {
      u.fn_out = Fn_[cur_];
      u.qnxx_out = qnxx_;
      u.qnyx_out = qnyx_;
      u.phn_out = phn_;
    } else {
      u.fn = Fn_[cur_];
      u.qnxx = qnxx_;
      u.qnyx = qnyx_;
      u.phn = phn_;
    }
    if (!growth_) {
      by_grid([&](auto gg) {
        using G = decltype(gg);
        if (first)
          hipLaunchKernelGGL((loop_uf_kernel<T, true, true, G>), dim3(node_blocks()), dim3(256),
                             0, s, u, P_, LX_, LY_, NPw_);
        else
          hipLaunchKernelGGL((loop_uf_kernel<T, true, false, G>), dim3(node_blocks()),
                             dim3(256), 0, s, u, P_, LX_, LY_, NPw_);
      });
      pc_ ^= 1;
      return;
    }
    GrowthUF a{};
    a.l = u;
    a.m = m_[mc_];
    a.chi = chi_;
    a.mn = mn_;
    a.division = div_;
    a.death = death_;
    a.m_out = m_[mc_ ^ 1];
    a.mn_out = mn_;
    a.countphi = part_ + 2 * kGrowthParts + 1;
    a.totalphi = totalphi_;
    a.conserve = conserve_ ? 1 : 0;
    const bool div_only = gp_.division_only != 0;
    // on a periodic grid no ghost condition follows the phenotype pass, so
    // the kernel writes chi = m/phi itself, into the other chi buffer (it
    // reads chi at the neighbours)
    const bool fold = bc_ == 0 && !div_only;
    a.chi_out = fold ? chi2_ : nullptr;
    by_grid([&](auto gg) {
      using G = decltype(gg);
      if (div_only) {
        if (first)
          hipLaunchKernelGGL((growth_uf_kernel<true, G, true>), dim3(node_blocks()), dim3(256), 0,
                             s, a, P_, gd_, LX_, LY_, NPw_);
        else
          hipLaunchKernelGGL((growth_uf_kernel<false, G, true>), dim3(node_blocks()), dim3(256),
                             0, s, a, P_, gd_, LX_, LY_, NPw_);
      } else if (first) {
        hipLaunchKernelGGL((growth_uf_kernel<true, G>), dim3(node_blocks()), dim3(256), 0, s, a,
                           P_, gd_, LX_, LY_, NPw_);
      } else {
        hipLaunchKernelGGL((growth_uf_kernel<false, G>), dim3(node_blocks()), dim3(256), 0, s,
                           a, P_, gd_, LX_, LY_, NPw_);
      }
    });
    if (div_only) {
      pc_ ^= 1;
      return;
    }
    pc_ ^= 1;
    mc_ ^= 1;
    if (fold)
      swap_chi();
    else
      phenotype(s);
  }
  void fields2_bc(hipStream_t s) {
    for (int ph : phases())
      hipLaunchKernelGGL((wall_fields2_kernel<T>), dim3(phase_blocks(ph)), dim3(256), 0, s, geo(),
                         ph, D_);
  }
  hipError_t zero_all(hipStream_t s) {
    for (size_t i = 0; i < ptrs_.size(); ++i) {
      const hipError_t e = hipMemsetAsync(ptrs_[i], 0, sizes_[i] * sizeof(T), s);
      if (e != hipSuccess) return e;
    }
    hipError_t e = hipSuccess;
    if (div_) e = hipMemsetAsync(div_, 0, (size_t)TP_, s);
    if (e == hipSuccess && death_) e = hipMemsetAsync(death_, 0, (size_t)TP_, s);
    return e;
  }
  // to host (out) or from host (in), through a device staging copy
  hipError_t copy_state(void* out, const void* in, hipStream_t s) {
    const size_t bytes = state_bytes();
    T* stage = nullptr;
    hipError_t e = hipMalloc((void**)&stage, bytes);
    if (e != hipSuccess) return e;
    T* planes[4] = {F_[cur_], qxx_, qyx_, phi_[pc_]};
    const int counts[4] = {9, 1, 1, 1};
    const int bs = 256, nb = (TP_ + bs - 1) / bs;
    if (in) e = hipMemcpyAsync(stage, in, bytes, hipMemcpyHostToDevice, s);
    size_t off = 0;
    for (int i = 0; i < 4 && e == hipSuccess; ++i) {
      if (in)
        hipLaunchKernelGGL((wall_copy_planes_kernel<T>), dim3(nb), dim3(bs), 0, s,
                           (const T*)(stage + off), TP_, planes[i], NPw_, TP_, counts[i]);
      else
        hipLaunchKernelGGL((wall_copy_planes_kernel<T>), dim3(nb), dim3(bs), 0, s,
                           (const T*)planes[i], NPw_, stage + off, TP_, TP_, counts[i]);
      off += (size_t)counts[i] * TP_;
      e = hipGetLastError();
    }
    if (e == hipSuccess && out) e = hipMemcpyAsync(out, stage, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(stage);
    return e;
  }
  hipError_t get_(size_t n, T** p) {
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e == hipSuccess) {
      ptrs_.push_back(*p);
      sizes_.push_back(n);
      bytes_ += n * sizeof(T);
      e = hipMemset(*p, 0, n * sizeof(T));
    }
    return e;
  }

  // ---------------------------------------------------------- go-or-grow
  static constexpr int kGrowthParts = 1024;
  // domain of a torus plane -> host (synchronous), host -> domain
  hipError_t download(const T* plane, std::vector<double>& out, hipStream_t s) {
    const int N = LX_ * LY_;
    out.resize(N);
    hipLaunchKernelGGL((wall_unpack_kernel<T>), dim3((N + 255) / 256), dim3(256), 0, s, geo(),
                       plane, 1, stage_);
    hipError_t e = hipMemcpyAsync(out.data(), stage_, sizeof(double) * N, hipMemcpyDeviceToHost,
                                  s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
  }
  hipError_t upload(const std::vector<double>& in, T* plane, hipStream_t s) {
    const int N = LX_ * LY_;
    hipError_t e = hipMemcpyAsync(stage_, in.data(), sizeof(double) * N, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(wall_scatter_kernel, dim3((N + 255) / 256), dim3(256), 0, s, geo(),
                       (const double*)stage_, plane);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // `in` may go away
    return e;
  }
  // One mask of LyotropicWithDivision::SetMasks (lyotropic-with-division.cpp:
  // 34-108): due when its counter was 0; the centres are drawn on the host
  // over the nodes with phi >= 1/2 in node order (the device's list), the
  // disks marked on the device.  The centres go through a pinned buffer per
  // mask, so nothing waits for the copy: the next draw into that buffer comes
  // after set_masks' next synchronous download.
  hipError_t redraw(unsigned char* mask, const int* sel_idx, const double* adjusted, int nsel,
                    double radius, bool division, hipStream_t s) {
    const int slot = division ? 0 : 1;
    std::vector<int>& centres = centres_;
    centres.clear();
    for (int j = 0; j < nsel; ++j) {
      const int k = sel_idx[j];
      if (next_uniform() < adjusted[j]) {
        centres.push_back(k / LY_);
        centres.push_back(k % LY_);
      }
    }
    hipError_t e = hipMemsetAsync(mask, 0, (size_t)TP_, s);
    const int nc = (int)centres.size() / 2;
    if (e != hipSuccess || nc == 0) return e;
    const size_t n = 2 * (size_t)nc;
    if (n > cent_cap_[slot]) {
      e = hipStreamSynchronize(s);  // nothing queued still uses the old buffers
      if (e != hipSuccess) return e;
      if (cent_[slot]) (void)hipFree(cent_[slot]);
      if (hcent_[slot]) (void)hipHostFree(hcent_[slot]);
      cent_[slot] = nullptr;
      hcent_[slot] = nullptr;
      cent_cap_[slot] = 0;
      e = hipMalloc((void**)&cent_[slot], sizeof(int) * n);
      if (e == hipSuccess)
        e = hipHostMalloc((void**)&hcent_[slot], sizeof(int) * n, hipHostMallocDefault);
      if (e != hipSuccess) return e;
      cent_cap_[slot] = n;
    }
    std::memcpy(hcent_[slot], centres.data(), sizeof(int) * n);
    e = hipMemcpyAsync(cent_[slot], hcent_[slot], sizeof(int) * n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    const double r2 = std::ceil(radius * radius);
    int R = 0;
    while ((double)(R + 1) * (double)(R + 1) <= r2) ++R;
    const long long span = 2 * R + 1, threads = (long long)nc * span * span;
    hipLaunchKernelGGL(mask_disk_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       s, (const int*)cent_[slot], nc, R, r2, LX_, LY_, mask);
    return hipGetLastError();
  }
  // the node list buffers of the mask draws (device and pinned host), kept
  // for the run
  hipError_t ensure_node_lists() {
    const int N = LX_ * LY_, nsb = (N + kSelBlock - 1) / kSelBlock;
    hipError_t e = hipSuccess;
    // (each buffer on its own: a failed call leaves the others for the next)
    if (!sel_idx_) e = hipMalloc((void**)&sel_idx_, sizeof(int) * ((size_t)N + nsb + 1));
    if (e == hipSuccess && !sel_val_)
      e = hipMalloc((void**)&sel_val_, sizeof(double) * 2 * (size_t)N);
    if (e == hipSuccess && !hsel_idx_)
      e = hipHostMalloc((void**)&hsel_idx_, sizeof(int) * ((size_t)N + 1), hipHostMallocDefault);
    if (e == hipSuccess && !hsel_val_)
      e = hipHostMalloc((void**)&hsel_val_, sizeof(double) * 2 * (size_t)N, hipHostMallocDefault);
    return e;
  }
  // The list length of a step whose due redraws all have rate zero, read
  // back without waiting; their masks are cleared and their counters reset.
  hipError_t count_only_redraw(bool div_due, bool death_due, hipStream_t s) {
    const int N = LX_ * LY_, nsb = (N + kSelBlock - 1) / kSelBlock;
    hipError_t e = hipSuccess;
    e = ensure_node_lists();
    if (e != hipSuccess) return e;
    if (!hpend_) {
      e = hipHostMalloc((void**)&hpend_, sizeof(int) * kMaxPending, hipHostMallocDefault);
      if (e != hipSuccess) return e;
    }
    if (npending_ == kMaxPending) {  // bound the lengths in flight
      e = hipStreamSynchronize(s);
      if (e != hipSuccess) return e;
      consume_pending();
    }
    int* counts = sel_idx_ + N;
    int* total = counts + nsb;
    hipLaunchKernelGGL(select_count_kernel, dim3(nsb), dim3(kSelBlock), 0, s, geo(),
                       (const T*)phi_[pc_], counts);
    hipLaunchKernelGGL(select_write_kernel, dim3(nsb), dim3(kSelBlock), 0, s, geo(),
                       (const T*)phi_[pc_], (const int*)counts, sel_idx_, sel_val_,
                       gp_.division_rate, gp_.death_rate, gp_.phi_critical, total);
    e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(hpend_ + npending_, total, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    pend_times_[npending_++] = (div_due ? 1 : 0) + (death_due ? 1 : 0);
    if (div_due) {
      div_count_ = (uint32_t)gp_.division_time;
      e = hipMemsetAsync(div_, 0, (size_t)TP_, s);
    }
    if (e == hipSuccess && death_due) {
      death_count_ = (uint32_t)gp_.death_time;
      e = hipMemsetAsync(death_, 0, (size_t)TP_, s);
    }
    return e;
  }
  // The draws' uniform numbers, taken from gen_ ahead of use and consumed in
  // order: the sequence each draw sees is the one drawing on demand gives.
  double next_uniform() {
    ++udraws_;
    if (ufront_ < ufifo_.size()) return ufifo_[ufront_++];
    return std::uniform_real_distribution<>(0., 1.)(*gen_);
  }
  void clear_uniforms() {
    ufifo_.clear();
    ufront_ = 0;
    uexpect_ = 0;
    npending_ = 0;
  }
  // The draws of redraws that could accept no centre, consumed now: their
  // list lengths have arrived (the stream was synchronised after they were
  // queued), and they came before any redraw that is about to draw.
  void consume_pending() {
    for (int i = 0; i < npending_; ++i) {
      const long long n = (long long)hpend_[i] * pend_times_[i];
      for (long long j = 0; j < n; ++j) (void)next_uniform();
    }
    npending_ = 0;
  }
  hipError_t set_masks(hipStream_t s) {
    const uint32_t div_old = div_count_--;
    const uint32_t death_old = death_count_--;
    if (div_old != 0 && death_old != 0) return hipSuccess;
    // A due redraw with rate zero accepts no centre (random_real() < 0 never
    // holds): its mask is cleared and it only draws one number per listed
    // node (lyotropic-with-division.cpp:44-50, 80-86).  When no due redraw
    // can accept, the step needs the list's length only, not the list: it is
    // read back without waiting and the numbers are drawn at the next
    // redraw that accepts (before its own), so the step queues without a
    // host synchronisation.  (The reference's proliferation runcards have
    // death-rate = 0: half their redraws are these.)
    const bool accept = (div_old == 0 && gp_.division_rate != 0) ||
                        (death_old == 0 && gp_.death_rate != 0);
    if (!accept) return count_only_redraw(div_old == 0, death_old == 0, s);
    // draw about as many numbers as the last redraw used while the queued
    // steps still run, before waiting for phi
    if (ufront_ > 0) {
      ufifo_.erase(ufifo_.begin(), ufifo_.begin() + (std::ptrdiff_t)ufront_);
      ufront_ = 0;
    }
    {
      const size_t want = uexpect_ + uexpect_ / 4 + 256;
      std::uniform_real_distribution<> u(0., 1.);
      while (ufifo_.size() < want) ufifo_.push_back(u(*gen_));
    }
    const size_t used0 = udraws_;
    // the nodes with phi >= 1/2 listed on the device and read back into
    // pinned buffers (kept for the run): the count and as many entries as
    // the last list held plus a margin, the rest only if the list grew more
    const int N = LX_ * LY_, nsb = (N + kSelBlock - 1) / kSelBlock;
    hipError_t e = hipSuccess;
    e = ensure_node_lists();
    if (e != hipSuccess) return e;
    int* counts = sel_idx_ + N;
    int* total = counts + nsb;
    hipLaunchKernelGGL(select_count_kernel, dim3(nsb), dim3(kSelBlock), 0, s, geo(),
                       (const T*)phi_[pc_], counts);
    hipLaunchKernelGGL(select_write_kernel, dim3(nsb), dim3(kSelBlock), 0, s, geo(),
                       (const T*)phi_[pc_], (const int*)counts, sel_idx_, sel_val_,
                       gp_.division_rate, gp_.death_rate, gp_.phi_critical, total);
    e = hipGetLastError();
    const int guess = std::min(N, sel_last_ + sel_last_ / 4 + 1024);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hsel_idx_ + N, total, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hsel_idx_, sel_idx_, sizeof(int) * guess, hipMemcpyDeviceToHost, s);
    // the rates of the due redraws: division at [0, N), death at [N, 2N)
    auto copy_rates = [&](size_t from, size_t count) {
      hipError_t r = hipSuccess;
      if (div_old == 0)
        r = hipMemcpyAsync(hsel_val_ + from, sel_val_ + from, sizeof(double) * count,
                           hipMemcpyDeviceToHost, s);
      if (r == hipSuccess && death_old == 0)
        r = hipMemcpyAsync(hsel_val_ + N + from, sel_val_ + N + from, sizeof(double) * count,
                           hipMemcpyDeviceToHost, s);
      return r;
    };
    if (e == hipSuccess) e = copy_rates(0, (size_t)guess);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    consume_pending();
    const int nsel = hsel_idx_[N];
    if (nsel < 0 || nsel > N) return hipErrorUnknown;
    if (nsel > guess) {
      const size_t rest = (size_t)(nsel - guess);
      e = hipMemcpyAsync(hsel_idx_ + guess, sel_idx_ + guess, sizeof(int) * rest,
                         hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = copy_rates((size_t)guess, rest);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) return e;
    }
    sel_last_ = nsel;
    if (div_old == 0) {
      div_count_ = (uint32_t)gp_.division_time;
      e = redraw(div_, hsel_idx_, hsel_val_, nsel, gp_.division_radius, true, s);
    }
    if (e == hipSuccess && death_old == 0) {
      death_count_ = (uint32_t)gp_.death_time;
      e = redraw(death_, hsel_idx_, hsel_val_ + N, nsel, gp_.death_radius, false, s);
    }
    uexpect_ = udraws_ - used0;
    return e;
  }

 public:
  hipError_t set_growth(const mass_lyo_growth& g, std::mt19937* gen) override {
    drop_graphs();
    const size_t n1 = (size_t)NPw_;
    hipError_t e = hipSuccess;
    if (!growth_) {
      for (int i = 0; i < 2 && e == hipSuccess; ++i) e = get_(n1, &m_[i]);
      if (e == hipSuccess) e = get_(n1, &mn_);
      if (e == hipSuccess) e = get_(n1, &chi_);
      if (e == hipSuccess) e = get_(n1, &chi2_);
      if (e == hipSuccess) e = get_(2 * kGrowthParts + 2, &part_);
      if (e == hipSuccess) e = get_((size_t)LX_ * LY_, &stage_);
      if (e == hipSuccess) e = get_(9, &fe_);
      if (e == hipSuccess) e = hipMalloc((void**)&div_, (size_t)TP_);
      if (e == hipSuccess) e = hipMalloc((void**)&death_, (size_t)TP_);
      if (e == hipSuccess) e = hipMemset(div_, 0, (size_t)TP_);
      if (e == hipSuccess) e = hipMemset(death_, 0, (size_t)TP_);
      if (e != hipSuccess) return e;
    }
    growth_ = true;
    gp_ = g;
    gen_ = gen;
    gd_.B = g.B;
    gd_.Snem = g.Snem;
    gd_.Dchi = g.Dchi;
    gd_.Achi = g.Achi;
    gd_.Ochi = g.Ochi;
    gd_.phiswitch = g.phiswitch;
    gd_.alpha = g.alpha;
    gd_.beta = g.beta;
    gd_.phi_critical = g.phi_critical;
    gd_.growTogether = g.growTogether;
    // lyotropic-with-division.cpp:23-31
    div_count_ = death_count_ = 0u;
    conserve_ = (g.alpha == 0 || g.division_rate == 0) && (g.beta == 0 || g.death_rate == 0);
    return hipSuccess;
  }

  // GoOrGrow::Configure after Lyotropic::Configure (go-or-grow.cpp:34-40)
  hipError_t growth_configure(double* ftot, double* ptot, hipStream_t s) override {
    if (!growth_) return hipErrorNotSupported;
    drop_graphs();  // relaxation and smoothing swap buffers; totalphi changes
    clear_uniforms();  // configure has just reseeded the generator
    // a configured run starts afresh: the masks were cleared with the state,
    // and both are due at the first step (lyotropic-with-division.cpp:23-31)
    div_count_ = death_count_ = 0u;
    totalphi_ = *ptot;
    if (gp_.division_only) return hipSuccess;  // Lyotropic's Configure only
    const int N = LX_ * LY_;
    const unsigned nb = node_blocks();
    hipError_t e = hipSuccess;
    // RelaxFreeEnergy (:42-119)
    if (gp_.relax_steps != 0 && (gp_.relax_phi || gp_.relax_Q)) {
      T* hxx = D_ + (size_t)LP_HXX * NPw_;
      T* hyx = D_ + (size_t)LP_HYX * NPw_;
      T* mu = D_ + (size_t)LP_MU * NPw_;
      for (uint32_t st = 0; st < gp_.relax_steps; ++st) {
        fields_bc(s, false);
        by_grid([&](auto gg) {
          using G = decltype(gg);
          hipLaunchKernelGGL((relax_mu_kernel<G>), dim3(nb), dim3(256), 0, s, (const T*)qxx_,
                             (const T*)qyx_, (const T*)phi_[pc_], P_, gd_, LX_, LY_, hxx, hyx, mu);
        });
        for (int ph : phases())
          hipLaunchKernelGGL((wall_scalars_kernel<T>), dim3(phase_blocks(ph)), dim3(256), 0, s,
                             geo(), ph, mu, (T*)nullptr);
        by_grid([&](auto gg) {
          using G = decltype(gg);
          hipLaunchKernelGGL((relax_update_kernel<G>), dim3(nb), dim3(256), 0, s, (const T*)qxx_,
                             (const T*)qyx_, (const T*)phi_[pc_], (const T*)hxx, (const T*)hyx,
                             (const T*)mu, P_, gp_.relax_dt, gp_.relax_phi, gp_.relax_Q, LX_, LY_,
                             phi_[pc_ ^ 1], qnxx_, qnyx_);
        });
        pc_ ^= 1;
        std::swap(qxx_, qnxx_);
        std::swap(qyx_, qnyx_);
      }
      fields_bc(s, false);
      std::vector<double> phi;
      e = download(phi_[pc_], phi, s);
      if (e != hipSuccess) return e;
      double relaxed_phi = 0.;
      for (int k = 0; k < N; ++k) relaxed_phi += phi[k];
      totalphi_ = relaxed_phi;
      *ptot = relaxed_phi;
    }
    // ResetHydrodynamics (:121-141): fe(0, 0, rho) of src/lb.hpp:24-41
    {
      const double nn = P_.rho, vx = 0., vy = 0.;
      const double v2 = vx * vx + vy * vy;
      double fe[9];
      fe[1] = nn * (1. / 9. + 1. / 3. * vx + 1. / 3. * v2 - 1. / 2. * vy * vy);
      fe[2] = nn * (1. / 9. - 1. / 3. * vx + 1. / 3. * v2 - 1. / 2. * vy * vy);
      fe[3] = nn * (1. / 9. + 1. / 3. * vy - 1. / 2. * vx * vx + 1. / 3. * v2);
      fe[4] = nn * (1. / 9. - 1. / 3. * vy - 1. / 2. * vx * vx + 1. / 3. * v2);
      fe[5] = nn * (1. / 36. + 1. / 12. * vx + 1. / 12. * vy + 1. / 12. * v2 + 1. / 4. * vx * vy);
      fe[6] = nn * (1. / 36. - 1. / 12. * vx - 1. / 12. * vy + 1. / 12. * v2 + 1. / 4. * vx * vy);
      fe[7] = nn * (1. / 36. - 1. / 12. * vx + 1. / 12. * vy + 1. / 12. * v2 - 1. / 4. * vx * vy);
      fe[8] = nn * (1. / 36. + 1. / 12. * vx - 1. / 12. * vy + 1. / 12. * v2 - 1. / 4. * vx * vy);
      fe[0] = nn - fe[1] - fe[2] - fe[3] - fe[4] - fe[5] - fe[6] - fe[7] - fe[8];
      e = hipMemcpy(fe_, fe, sizeof(fe), hipMemcpyHostToDevice);
      if (e != hipSuccess) return e;
      by_grid([&](auto gg) {
        using G = decltype(gg);
        hipLaunchKernelGGL((reset_hydro_kernel<G>), dim3(nb), dim3(256), 0, s, F_[0], F_[1],
                           Fn_[0], Fn_[1], D_, P_.rho, (const T*)fe_, LX_, LY_, NPw_);
      });
      double ftotal = 0.;
      for (int k = 0; k < N; ++k)
        for (int v = 0; v < 9; ++v) ftotal = ftotal + fe[v];
      *ftot = ftotal;
    }
    // ConfigurePhenotype (:143-199)
    std::vector<double> chi(N);
    if (gp_.chi_config == 0) {
      for (int k = 0; k < N; ++k) chi[k] = std::normal_distribution<>(0., 1.)(*gen_);
      e = upload(chi, chi_, s);
      if (e != hipSuccess) return e;
      fields_bc(s);
      const double smooth_rate = 0.2;
      const unsigned smooth_steps =
          gp_.chi_length > 0
              ? static_cast<unsigned>(std::ceil(gp_.chi_length * gp_.chi_length / (2 * smooth_rate)))
              : 0u;
      for (unsigned st = 0; st < smooth_steps; ++st) {
        T* tmp = m_[mc_ ^ 1];
        by_grid([&](auto gg) {
          using G = decltype(gg);
          hipLaunchKernelGGL((chi_smooth_kernel<G>), dim3(nb), dim3(256), 0, s, (const T*)chi_,
                             tmp, LX_, LY_);
        });
        m_[mc_ ^ 1] = chi_;  // swap(chi, m_tmp)
        chi_ = tmp;
        fields_bc(s);
      }
      e = download(chi_, chi, s);
      if (e != hipSuccess) return e;
      double mean = 0.;
      for (int k = 0; k < N; ++k) mean += chi[k];
      mean /= N;
      double variance = 0.;
      for (int k = 0; k < N; ++k) {
        chi[k] -= mean;
        variance += chi[k] * chi[k];
      }
      variance /= N;
      const double amplitude = std::sqrt(gp_.chi_noise);
      const double norm = variance > 0 ? amplitude / std::sqrt(variance) : 0.;
      for (int k = 0; k < N; ++k) chi[k] = gp_.chi0 + norm * chi[k];
    } else {
      const double length = gp_.chi_length > 0 ? gp_.chi_length : 1.;
      const double center = .5 * (unsigned)(LX_ - 1);
      for (int k = 0; k < N; ++k) {
        const double x = (unsigned)k / (unsigned)LY_;
        chi[k] = .5 * (1. - std::tanh((x - center) / length));
      }
    }
    e = upload(chi, chi_, s);
    if (e != hipSuccess) return e;
    by_grid([&](auto gg) {
      using G = decltype(gg);
      hipLaunchKernelGGL((phenotype_init_kernel<G>), dim3(nb), dim3(256), 0, s,
                         (const T*)phi_[pc_], (const T*)chi_, m_[mc_], LX_, LY_);
    });
    phenotype(s);
    fields_bc(s);
    return hipGetLastError();
  }

 private:
  // a captured step: the parities it starts from and leaves
  struct StepGraph {
    int cur, pc, mc, qa, ca;
    bool loops;
    int cur_after, pc_after, mc_after, qa_after, ca_after;
    hipGraphExec_t exec;
  };
  std::vector<StepGraph> graphs_;
  void drop_graphs() {
    for (StepGraph& g : graphs_) (void)hipGraphExecDestroy(g.exec);
    graphs_.clear();
  }

  bool growth_ = false, conserve_ = false;
  bool loops_ = false;  // the node loops instead of the fused sweeps
  int qa_ = 0;          // which of the two Q buffers is current (fused sweeps)
  bool d_stale_ = false;  // D_ predates the last fused step
  T *qxx2_ = nullptr, *qyx2_ = nullptr;
  mass_lyo_growth gp_{};
  GrowthDev gd_{};
  std::mt19937* gen_ = nullptr;
  uint32_t div_count_ = 0, death_count_ = 0;
  double totalphi_ = 0;
  int mc_ = 0;  // m parity (UpdateFields swaps m with m_tmp)
  T* m_[2] = {};
  T* mn_ = nullptr;
  T* chi_ = nullptr;
  T* chi2_ = nullptr;  // periodic go-or-grow: UpdateFields writes chi here
  int ca_ = 0;         // which of chi_/chi2_ holds chi (graph key)
  T* part_ = nullptr;   // fixed-tree sum partials; [2*parts+1]: sum of phi
  T* stage_ = nullptr;  // N doubles: host <-> domain staging
  T* fe_ = nullptr;
  unsigned char* div_ = nullptr;
  unsigned char* death_ = nullptr;
  int* cent_[2] = {};          // mask centres on the device: division, death
  int* hcent_[2] = {};         // their pinned host copies
  size_t cent_cap_[2] = {};
  std::vector<int> centres_;   // the draw's centres, (x, y) pairs
  std::vector<double> ufifo_;  // uniforms drawn ahead from gen_, in order
  size_t ufront_ = 0;          // the next unused one
  size_t udraws_ = 0;          // numbers the draws have consumed
  size_t uexpect_ = 0;         // used by the last redraw
  // redraws whose draws are still to be consumed: their list lengths (pinned,
  // read back without waiting) and how many redraws drew on each list
  static constexpr int kMaxPending = 64;
  int* hpend_ = nullptr;
  int pend_times_[kMaxPending] = {};
  int npending_ = 0;
  int* sel_idx_ = nullptr;     // nodes with phi >= 1/2: domain index (N), block
                               // counts, the count
  double* sel_val_ = nullptr;  // and their division and death rates (2N)
  int* hsel_idx_ = nullptr;    // pinned host copies (the count at [N])
  double* hsel_val_ = nullptr;
  int sel_last_ = 0;           // the last list's length

  int bc_;
  Params<T> P_{};
  int LX_ = 0, LY_ = 0, npc_ = 1, TP_ = 0, NPw_ = 0;
  int cur_ = 0;  // ff/fn parity (Move swaps ff with ff_tmp)
  int pc_ = 0;   // phi parity (UpdateFields swaps phi with phi_tmp)
  T* F_[2] = {};
  T* Fn_[2] = {};
  T* phi_[2] = {};
  T *qxx_ = nullptr, *qyx_ = nullptr, *qnxx_ = nullptr, *qnyx_ = nullptr, *phn_ = nullptr;
  T* D_ = nullptr;
  std::vector<T*> ptrs_;
  std::vector<size_t> sizes_;
  size_t bytes_ = 0;
};

}  // namespace lyo
