// lyo_sweep.hpp -- fused predictor / corrector sweep kernels (gfx950).
//
// One sweep = UpdateQuantities + UpdateFields (+ Move for the predictor) of
// src/models/lyotropic.cpp:343-386,500-525, fused so that the derived fields
// (u, mu, sigma, H, grad Q) never leave registers.
//
// Work decomposition ("strip march"): one 64-lane wavefront owns a strip of
// 64 consecutive y columns (y is the contiguous index, k = y + LY*x) and
// marches along x over a chunk of rows.  Lanes 0-1 and 62-63 are halo: the
// second-order stencil dependency (UpdateFields reads u/mu/sigma at the four
// neighbours, which themselves read Q/phi at their four neighbours) needs two
// columns on each side, so each wave writes 60 interior columns.  Neighbours
// in y come from the adjacent lanes through DPP wave shifts; neighbours in x
// are the previous / next row kept in registers.  Every global access is a
// contiguous 64-lane load or store of one SoA plane.
#pragma once

#include <type_traits>

#include "lyo_physics.hpp"

namespace lyo {

constexpr int kWave = 64;
constexpr int kHalo = 2;
constexpr int kStrip = kWave - 2 * kHalo;

// value held by lane-1 (the y-1 neighbour); lane 0 receives 0.  The DPP
// wave shift usually folds into the instruction that consumes it.
__device__ __forceinline__ float from_ym(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
}
// value held by lane+1 (the y+1 neighbour); lane 63 receives 0
__device__ __forceinline__ float from_yp(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xF, 0xF, true));
}
// The same shifts where the lane without a source (lane 0 for from_ym, lane
// 63 for from_yp) keeps `old` instead of receiving 0 (bound_ctrl off).
__device__ __forceinline__ float from_ym_or(float v, float old) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_yp_or(float v, float old) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ double from_ym_or(double v, double old) {
  const long long b = __double_as_longlong(v), o = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)b, 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), 0x138, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double from_yp_or(double v, double old) {
  const long long b = __double_as_longlong(v), o = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)b, 0x130, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), 0x130, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double from_ym(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x138, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double from_yp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x130, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x130, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// s_waitcnt vmcnt(0) (gfx9 encoding; expcnt and lgkmcnt at their maxima): wait
// for every vector memory access in flight.  Issued before a row's stores: the
// next row's loads (issued earlier in the row) have arrived by then, so the
// waits the compiler places after the divergent store block -- where it must
// assume the stores may still be in flight ahead of those loads -- find nothing
// left to wait for, and the next row starts without waiting for this row's
// stores.  FP32 only (two waves per SIMD): the FP64 kernels run one wave per
// SIMD, whose store block holds much of the row's remaining arithmetic, and an
// early wait there stalls the SIMD (measured 10% slower at npc = 1).
template <bool EXACT>
__device__ __forceinline__ void wait_vmem() {
  if constexpr (!EXACT) __builtin_amdgcn_s_waitcnt(0x0F70);
}

struct Geo {
  int LX, LY, N;
  int NP;  // plane stride (>= N) of every SoA field
  int nstrips, nchunks, xc;
  int nwaves;
  int xcd;  // remap block ids so each XCD gets a contiguous band of chunks
  int alt;  // 1: even chunks march downwards, odd upwards; 0: all upwards
};

// Host-side check of a launch geometry against the strip width of its kernel
// (and the rows a kernel's single wrap needs): the grid, the chunk and strip
// counts and the plane stride all have to agree with what the kernel indexes.
inline bool geo_ok(const Geo& g, int strip, int min_lx) {
  if (g.LX < min_lx || g.LY < 1 || g.xc < 1 || g.N != g.LX * g.LY || g.NP < g.N) return false;
  return g.nstrips == (g.LY + strip - 1) / strip && g.nchunks == (g.LX + g.xc - 1) / g.xc &&
         g.nwaves == g.nstrips * g.nchunks;
}

enum Stage { PRED = 0, CORR = 1, CORR_LAST = 2 };

// Buffer roles per stage (planes of N values; populations are 9 planes).
//  PRED      in : f_in, qxx_in, qyx_in, phi_in
//            out: f_out (streamed f*), fn_out (streamed fn*), qxx_out/qyx_out
//                 (Q+D), qnxx_out/qnyx_out (Q+D/2), EXACT: phi_out (phi+Dp),
//                 phn_out (phi+Dp/2); FP32: phi_out = Dp
//  CORR      in : f_in, q*_in, phi_in (+ FP32: dphi_in, stencil phi =
//                 phi_in + dphi_in), fn_in, qn*_in, phn_in (FP32: predictor Dp)
//            out: f_out, qxx_out, qyx_out, phi_out (FP32: increment of phi)
//  CORR_LAST as CORR; FP32 also reads phic_in and writes phi_out, phic_out
//            (compensated phi).  For EXACT, CORR_LAST == CORR.
// Half-step values (template flag HB, FP32 only).  fn* and Q+D/2 follow from
// the full predictor step and its source, fn = (f* + f_src)/2 with f_src the
// old population at the upstream node, QN = (Q* + Q)/2.  With HB the first
// corrector rebuilds them from its own inputs (f*, Q*) and the old state
// (fold_in, qo*_in) instead of loading fn_in/qn*_in, and the predictor stores
// them only when a later corrector loads them (HB on PRED).  EXACT always
// stores and loads them in the reference's form (f + om/2, Q + D/2).
template <class T>
struct SweepArgs {
  const T* f_in;
  const T* qxx_in;
  const T* qyx_in;
  const T* phi_in;
  const T* dphi_in;
  const T* fn_in;
  const T* qnxx_in;
  const T* qnyx_in;
  const T* phn_in;
  const T* phic_in;
  const T* fold_in;  // FP32 HB correctors: old populations, Q
  const T* qoxx_in;
  const T* qoyx_in;
  T* f_out;
  T* fn_out;
  T* qxx_out;
  T* qyx_out;
  T* qnxx_out;
  T* qnyx_out;
  T* phi_out;
  T* phn_out;
  T* phic_out;
};

// Lane part of an address: a 32-bit byte offset added to a wave-uniform row
// base, so that every access is one scalar base plus one 32-bit lane offset
// (the saddr form of global loads and stores; no 64-bit lane arithmetic).
// Each access first hides the offset's value from the optimiser (an empty asm
// that "modifies" it in place, so no instruction and no copy), which would
// otherwise fold base+offset into 64-bit per-lane pointers (hoisted out of the
// loop, or chained from plane to plane): two VGPRs and a 64-bit add each.
template <class T>
struct LaneOff {
  unsigned b;
  __device__ __forceinline__ T ld(const T* row) {
    asm volatile("" : "+v"(b));
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(row) + b);
  }
  __device__ __forceinline__ void st(T* row, T v) {
    asm volatile("" : "+v"(b));
    *reinterpret_cast<T*>(reinterpret_cast<char*>(row) + b) = v;
  }
};

template <class T>
struct QP {
  T xx, yx, p;
};

// Data one loop iteration (UpdateFields on row x) needs from memory: Q/phi of
// row x+2 and populations of row x+1 for UpdateQuantities on row x+1, and for
// the correctors the centre inputs fn, QN, phn of row x (rebuilding FP32
// correctors: the old populations at the upstream nodes and old Q instead of
// fn, QN).  The loop loads the next iteration's RowIn before computing the
// current one, so the loads of one row overlap the arithmetic of the previous
// one.
template <class T, int STAGE>
struct RowIn {
  QP<T> q;
  T dp;  // FP32 correctors: increment added to q.p when the row is used
  T f[9];
  T fn[(STAGE == PRED) ? 1 : 9];
  T qnxx, qnyx, phn;
};

// FP32 half-step value from a full step and its source (see SweepArgs); one
// expression wherever it is formed, so stored and rebuilt values agree bitwise
template <class T>
__device__ __forceinline__ T half_of(T full, T src) {
  return T(.5) * (full + src);
}

#ifndef LYO_SWEEP_WAVES_PER_EU
#define LYO_SWEEP_WAVES_PER_EU 3
#endif

// y neighbours of a value from the adjacent lanes of the wave
template <class T, bool EXACT>
__device__ __forceinline__ YNb<T, EXACT> ynb_lanes(T v) {
  return ynb<T, EXACT>(from_yp(v), from_ym(v));
}

// Neighbour values UpdateFields needs at the centre row: x neighbours from the
// rows before (um) and after (un), y neighbours from the adjacent lanes of the
// centre row (u0; uf_stencil_x leaves them to the caller).
template <class T, bool EXACT>
__device__ __forceinline__ UFStencil<T, EXACT> uf_stencil_x(const NbrVals<T>& um,
                                                            const NbrVals<T>& u0,
                                                            const NbrVals<T>& un) {
  UFStencil<T, EXACT> s;
  s.mu_c = u0.mu;
  s.ux_xp = un.ux;
  s.ux_xm = um.ux;
  s.uy_xp = un.uy;
  s.uy_xm = um.uy;
  s.uxp_xp = un.uxp;
  s.uxp_xm = um.uxp;
  s.sxx_xp = un.sxx;
  s.sxx_xm = um.sxx;
  s.syx_xp = un.syx;
  s.syx_xm = um.syx;
  s.mu_xp = un.mu;
  s.mu_xm = um.mu;
  return s;
}
template <class T, bool EXACT>
__device__ __forceinline__ UFStencil<T, EXACT> uf_stencil(const NbrVals<T>& um,
                                                          const NbrVals<T>& u0,
                                                          const NbrVals<T>& un) {
  UFStencil<T, EXACT> s = uf_stencil_x<T, EXACT>(um, u0, un);
  s.ux_y = ynb_lanes<T, EXACT>(u0.ux);
  s.uy_y = ynb_lanes<T, EXACT>(u0.uy);
  s.uyp_y = ynb_lanes<T, EXACT>(u0.uyp);
  s.sxy_y = ynb_lanes<T, EXACT>(u0.sxy);
  s.syy_y = ynb_lanes<T, EXACT>(u0.syy);
  s.mu_y = ynb_lanes<T, EXACT>(u0.mu);
  return s;
}

// One work item: wave-uniform item index wid in [0, g.nwaves) = one strip of
// 64 columns over one chunk of rows of one run.
template <class T, bool EXACT, int STAGE, bool HB>
__device__ __forceinline__ void sweep_item(const SweepArgs<T>& A, const Params<T>& P,
                                           const Geo& g, const int wid) {
  constexpr bool kStoreHalf = STAGE == PRED && (EXACT || HB);
  constexpr bool kRebuild = STAGE != PRED && !EXACT && HB;
  const int lane = threadIdx.x & 63;
  const int strip = wid % g.nstrips;
  const int chunk = wid / g.nstrips;
  const int LX = g.LX, LY = g.LY, N = g.NP;

  const int ycol = strip * kStrip - kHalo + lane;
  int yi = ycol % LY;
  if (yi < 0) yi += LY;
  // byte offsets of the columns y, y+1, y-1 (periodic) in a row
  LaneOff<T> ly{(unsigned)yi * (unsigned)sizeof(T)};
  LaneOff<T> lyp{(yi + 1 == LY ? 0u : (unsigned)(yi + 1)) * (unsigned)sizeof(T)};
  LaneOff<T> lym{(yi == 0 ? (unsigned)(LY - 1) : (unsigned)(yi - 1)) * (unsigned)sizeof(T)};
  const bool active = lane >= kHalo && lane < kWave - kHalo && ycol < LY;
  const int x0 = chunk * g.xc;
  const int x1 = min(x0 + g.xc, LX);

  // row index r in [x0-2, x1+3] (LX >= 2) -> offset of row (r mod LX)
  auto roff = [&](int r) {
    r = (r < 0) ? r + LX : r;
    r = (r < 0) ? r + LX : r;
    r = (r >= LX) ? r - LX : r;
    r = (r >= LX) ? r - LX : r;
    return r * LY;
  };
  auto load_qp = [&](int r) {
    const int o = roff(r);
    QP<T> q;
    q.xx = ly.ld(A.qxx_in + o);
    q.yx = ly.ld(A.qyx_in + o);
    T ph = ly.ld(A.phi_in + o);
    if constexpr (!EXACT && STAGE != PRED) ph += ly.ld(A.dphi_in + o);
    q.p = ph;
    return q;
  };
  auto load_f = [&](int r, T(&f)[9]) {
    const int o = roff(r);
#pragma unroll
    for (int v = 0; v < 9; ++v) f[v] = ly.ld(A.f_in + (v * N + o));
  };
  // loads of the iteration on row x of a march in direction d (+1 or -1)
  auto load_row = [&](int x, int d, RowIn<T, STAGE>& R) {
    {
      const int o = roff(x + 2 * d);
      R.q.xx = ly.ld(A.qxx_in + o);
      R.q.yx = ly.ld(A.qyx_in + o);
      R.q.p = ly.ld(A.phi_in + o);
      if constexpr (!EXACT && STAGE != PRED) R.dp = ly.ld(A.dphi_in + o);
    }
    load_f(x + d, R.f);
    if constexpr (kRebuild) {
      // old populations at the upstream node x - c_v, old Q at x
      const int o = roff(x), om = roff(x - 1), op = roff(x + 1);
      const int rowv[9] = {o, om, op, o, o, om, op, op, om};
      LaneOff<T>* const colv[9] = {&ly, &ly, &ly, &lym, &lyp, &lym, &lyp, &lym, &lyp};
#pragma unroll
      for (int v = 0; v < 9; ++v) R.fn[v] = colv[v]->ld(A.fold_in + (v * N + rowv[v]));
      R.qnxx = ly.ld(A.qoxx_in + o);
      R.qnyx = ly.ld(A.qoyx_in + o);
      R.phn = ly.ld(A.phn_in + o);
    } else if constexpr (STAGE != PRED) {
      const int o = roff(x);
#pragma unroll
      for (int v = 0; v < 9; ++v) R.fn[v] = ly.ld(A.fn_in + (v * N + o));
      R.qnxx = ly.ld(A.qnxx_in + o);
      R.qnyx = ly.ld(A.qnyx_in + o);
      R.phn = ly.ld(A.phn_in + o);
    }
  };
  auto uq = [&](const QP<T>& qm, const QP<T>& q0, const QP<T>& qp, const T(&f)[9],
                NbrVals<T>& nv, CenterVals<T>& cv) {
    update_quantities<T, EXACT>(f, q0.xx, qp.xx, qm.xx, ynb_lanes<T, EXACT>(q0.xx), q0.yx,
                                qp.yx, qm.yx, ynb_lanes<T, EXACT>(q0.yx), q0.p, qp.p, qm.p,
                                ynb_lanes<T, EXACT>(q0.p), P, nv, cv);
  };

  // The chunk is marched upwards (D = +1, rows x0 .. x1-1) or downwards
  // (D = -1).  Alternating the direction with the chunk parity makes two
  // neighbouring chunks reach their shared boundary rows at the same time
  // (both at their start or both at their end), so the rows one of them
  // loads are still in L2 for the other.  "prev"/"next" below are the rows
  // x-D / x+D; the stencils take the rows x-1 / x+1, swapped for D = -1.
  auto march = [&](auto dir) {
    constexpr int D = decltype(dir)::value;
    auto uqd = [&](const QP<T>& qprev, const QP<T>& q0, const QP<T>& qnext, const T(&f)[9],
                   NbrVals<T>& nv, CenterVals<T>& cv) {
      if constexpr (D > 0) {
        uq(qprev, q0, qnext, f, nv, cv);
      } else {
        uq(qnext, q0, qprev, f, nv, cv);
      }
    };
    const int xb = (D > 0) ? x0 : x1 - 1;
    // prologue: UpdateQuantities on rows xb-D and xb
    NbrVals<T> um, u0;
    CenterVals<T> c0;
    T fc[9];
    QP<T> qb, qc;
    RowIn<T, STAGE> cur;
    {
      const QP<T> qa = load_qp(xb - 2 * D);
      qb = load_qp(xb - D);
      qc = load_qp(xb);
      const QP<T> qd = load_qp(xb + D);
      T ft[9];
      load_f(xb - D, ft);
      load_f(xb, fc);
      load_row(xb, D, cur);
      CenterVals<T> ct;
      uqd(qa, qb, qc, ft, um, ct);
      uqd(qb, qc, qd, fc, u0, c0);
      qb = qc;
      qc = qd;
    }
    // invariant at loop top: qb = row x, qc = row x+D; um/u0 = rows x-D/x;
    // c0, fc = centre data of row x; cur = loads of this iteration
    for (int i = 0, x = xb; i < x1 - x0; ++i, x += D) {
    // prefetch unconditionally (the wrapped row past the chunk is in bounds
    // and unused): a conditional load would make the outstanding-load count
    // path dependent and force a full wait at the top of the loop
    RowIn<T, STAGE> nxt;
    load_row(x + D, D, nxt);
    const int ox = roff(x);

    // UpdateQuantities on row x+D
    if constexpr (!EXACT && STAGE != PRED) cur.q.p += cur.dp;
    NbrVals<T> un;
    CenterVals<T> cn;
    uqd(qb, qc, cur.q, cur.f, un, cn);

    // UpdateFields on row x
    const UFStencil<T, EXACT> s =
        (D > 0) ? uf_stencil<T, EXACT>(um, u0, un) : uf_stencil<T, EXACT>(un, u0, um);
    T Dxx, Dyx, Dp, om[9];
    const T Qxx = qb.xx, Qyx = qb.yx, p = qb.p;
    // FP32 predictor: om receives the post-collision populations f + Omega
    constexpr bool kPost = !EXACT && STAGE == PRED;
    update_fields<T, EXACT, kPost>(c0, Qxx, Qyx, p, fc, s, P, Dxx, Dyx, Dp, om);

    wait_vmem<EXACT>();
    if (active) {
      if constexpr (STAGE == PRED) {
        const T Qsx = Qxx + Dxx, Qsy = Qyx + Dyx;
        ly.st(A.qxx_out + ox, Qsx);
        ly.st(A.qyx_out + ox, Qsy);
        if constexpr (EXACT) {
          ly.st(A.qnxx_out + ox, Qxx + T(.5) * Dxx);
          ly.st(A.qnyx_out + ox, Qyx + T(.5) * Dyx);
          ly.st(A.phn_out + ox, p + T(.5) * Dp);
          ly.st(A.phi_out + ox, p + Dp);
        } else {
          if constexpr (kStoreHalf) {
            ly.st(A.qnxx_out + ox, half_of(Qsx, Qxx));
            ly.st(A.qnyx_out + ox, half_of(Qsy, Qyx));
          }
          ly.st(A.phi_out + ox, Dp);
        }
        // collide and push to (x+cx, y+cy) (Move, lyotropic.cpp:371-386)
        const int rm = roff(x - 1), rp = roff(x + 1);
        const int rowv[9] = {ox, rp, rm, ox, ox, rp, rm, rm, rp};
        LaneOff<T>* const colv[9] = {&ly, &ly, &ly, &lyp, &lym, &lyp, &lym, &lyp, &lym};
#pragma unroll
        for (int v = 0; v < 9; ++v) {
          const int d = v * N + rowv[v];
          const T fs = kPost ? om[v] : fc[v] + om[v];
          if constexpr (EXACT) {
            colv[v]->st(A.fn_out + d, fc[v] + T(.5) * om[v]);
          } else if constexpr (kStoreHalf) {
            colv[v]->st(A.fn_out + d, half_of(fs, fc[v]));
          }
          colv[v]->st(A.f_out + d, fs);
        }
      } else {
        // centre half-step values fn, QN of row x
        T fnv[9], qnx, qny;
        if constexpr (kRebuild) {
#pragma unroll
          for (int v = 0; v < 9; ++v) fnv[v] = half_of(fc[v], cur.fn[v]);
          qnx = half_of(Qxx, cur.qnxx);
          qny = half_of(Qyx, cur.qnyx);
        } else {
#pragma unroll
          for (int v = 0; v < 9; ++v) fnv[v] = cur.fn[v];
          qnx = cur.qnxx;
          qny = cur.qnyx;
        }
        ly.st(A.qxx_out + ox, qnx + T(.5) * Dxx);
        ly.st(A.qyx_out + ox, qny + T(.5) * Dyx);
#pragma unroll
        for (int v = 0; v < 9; ++v) ly.st(A.f_out + (v * N + ox), fnv[v] + T(.5) * om[v]);
        if constexpr (EXACT) {
          ly.st(A.phi_out + ox, cur.phn + T(.5) * Dp);
        } else if constexpr (STAGE == CORR) {
          ly.st(A.phi_out + ox, T(.5) * (cur.phn + Dp));
        } else {
          // phi_new = phi_old + phic + (Dp_pred + Dp)/2 with a TwoSum, so the
          // rounding of the addition is carried to the next step
          const T pold = ly.ld(A.phi_in + ox);
          const T yv = T(.5) * (cur.phn + Dp) + ly.ld(A.phic_in + ox);
          const T t = pold + yv;
          const T bp = t - pold;
          const T ap = t - bp;
          const T err = (pold - ap) + (yv - bp);
          ly.st(A.phi_out + ox, t);
          ly.st(A.phic_out + ox, err);
        }
      }
    }
    // rotate the rolling rows
    um = u0;
    u0 = un;
    c0 = cn;
#pragma unroll
    for (int v = 0; v < 9; ++v) fc[v] = cur.f[v];
    qb = qc;
    qc = cur.q;
    cur = nxt;
    }
  };
  if (g.alt && !(chunk & 1)) {
    march(std::integral_constant<int, -1>{});
  } else {
    march(std::integral_constant<int, 1>{});
  }
}

template <class T, bool EXACT, int STAGE, bool HB>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? LYO_SWEEP_WAVES_PER_EU : 1)))
sweep_kernel(const SweepArgs<T> A, const Params<T> P, const Geo g) {
  int b = blockIdx.x;
  if (g.xcd) {
    // blocks are dispatched round-robin over the 8 XCDs; give XCD i the
    // logical blocks [i*nb8, (i+1)*nb8) so halo rows stay in one L2
    const int nb8 = (int)gridDim.x >> 3;
    if (b < (nb8 << 3)) b = (b & 7) * nb8 + (b >> 3);
  }
  // wave-uniform: keeps the row arithmetic in scalar registers
  const int wid = __builtin_amdgcn_readfirstlane(b * ((int)blockDim.x >> 6) +
                                                 ((int)threadIdx.x >> 6));
  if (wid >= g.nwaves) return;  // whole waves exit; no block barriers below
  sweep_item<T, EXACT, STAGE, HB>(A, P, g, wid);
}

// Batched sweeps: many independent runs of any size in one launch.  Job j
// covers the work items [first, first + g.nwaves) of the global item list.
template <class T>
struct SweepJob {
  SweepArgs<T> A;
  Params<T> P;
  Geo g;
  int first;
};

// Job of a global work item: binary search over the jobs' first items (scalar
// loads; item is wave-uniform).
template <class T>
__device__ __forceinline__ const SweepJob<T>& find_job(const SweepJob<T>* __restrict__ jobs,
                                                       const int njobs, const int item) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].first <= item) lo = mid;
    else hi = mid - 1;
  }
  return jobs[lo];
}

// Persistent grid: each wave walks the global item list with a stride of the
// whole grid, finds the job of its item and runs it.  Every wave reaches the
// end of the list.
template <class T, bool EXACT, int STAGE, bool HB>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? LYO_SWEEP_WAVES_PER_EU : 1)))
sweep_batch_kernel(const SweepJob<T>* __restrict__ jobs, const int njobs, const int nitems) {
  const int nw = (int)(gridDim.x * (blockDim.x >> 6));
  int item = __builtin_amdgcn_readfirstlane((int)blockIdx.x * ((int)blockDim.x >> 6) +
                                            ((int)threadIdx.x >> 6));
  for (; item < nitems; item += nw) {
    const SweepJob<T>& J = find_job(jobs, njobs, item);
    sweep_item<T, EXACT, STAGE, HB>(J.A, J.P, J.g, item - J.first);
  }
}

// Launch helper: waves_per_block waves of 64 lanes per block.
template <class T, bool EXACT, int STAGE, bool HB>
inline hipError_t launch_sweep(const SweepArgs<T>& a, const Params<T>& p, const Geo& g,
                               int waves_per_block, hipStream_t s) {
  if (!geo_ok(g, kStrip, 2) || waves_per_block < 1) return hipErrorInvalidValue;
  const int threads = 64 * waves_per_block;
  const int blocks = (g.nwaves + waves_per_block - 1) / waves_per_block;
  hipLaunchKernelGGL((sweep_kernel<T, EXACT, STAGE, HB>), dim3(blocks), dim3(threads), 0, s, a,
                     p, g);
  return hipGetLastError();
}

}  // namespace lyo
