// lyo_full.hpp -- one whole step (predictor + one or two correctors) in one
// kernel (gfx950), FP32 and FP64.
//
// Lyotropic::Step (src/models/lyotropic.cpp:500-525) with npc = 2 is a
// predictor sweep (UpdateQuantities, UpdateFields(true), Move) followed by two
// corrector sweeps.  Here the three run as three stages of one strip march
// (upwards in x):
//   P  predictor, UpdateFields on row x (its Move is a pull, below)
//   A  corrector 1, UpdateQuantities on row x-1, UpdateFields on row x-2
//   B  corrector 2, UpdateQuantities on row x-3, UpdateFields on row x-4
// P's post-collision populations of rows x-2..x are pulled into the streamed
// populations of row x-1 (the node at (r, y) takes direction v from
// (r - cx_v, y - cy_v): rows from registers, columns through DPP lane shifts),
// so neither the predictor's outputs nor the first corrector's ever reach HBM:
// per node the kernel reads the old state and writes the new one.  Every stage
// calls the same device functions in the same order as the separate sweeps
// (lyo_sweep.hpp); FP64 forms the half-step values as there (the reference's
// f + om/2, Q + D/2, phi + Dp/2), FP32 keeps twice them (f* + f, Q* + Q: one
// addition where half_of takes two operations) and has the correctors return
// halved increments, the same values since halving is exact; so the results
// are bitwise those of the predictor sweep + fused pair.  Each stage needs its
// inputs one lane further in, so the halo is six lanes per side (52 interior
// columns), and a chunk repeats eight rows.
//
// With npc = 1 (NPC = 1) stage A is the last corrector: it stores the new
// state (FP32: phi with the TwoSum of the separate last corrector) and there
// is no stage B and no LDS ring; the halo is four lanes (56 interior columns)
// and a chunk repeats four rows.  The results are bitwise those of the
// predictor sweep + the separate corrector sweep.
#pragma once

#include <type_traits>

#include "lyo_pair.hpp"

namespace lyo {

constexpr int kFullHalo = 6;
constexpr int kFullStrip = kWave - 2 * kFullHalo;
constexpr int kFull1Halo = 4;  // npc = 1
constexpr int kFull1Strip = kWave - 2 * kFull1Halo;

// Block-cooperative one-kernel step (npc = 2): the kXWaves waves of a block
// cover kXWaves*64 contiguous columns and hand the values their stencils
// need across the wave boundaries through LDS, so only the block's outer
// edges carry the six-lane halo (244 interior columns per block instead of
// 4 x 52).  Exchange area per block: [wave][side: lane 0 / lane 63][slot]
// [quantity] values, then a scratch region the other lanes write into (every
// publication is one all-lane ds_write: lane 0 and lane 63 to their slots,
// lanes 1..62 to distinct scratch words that nobody reads, two words apart so
// that the compiler's paired writes of adjacent ids stay free of bank
// conflicts).
constexpr int kXWaves = 4;
constexpr int kXStrip = kXWaves * kWave - 2 * kFullHalo;
constexpr int kXQ = 40;     // quantities per slot
constexpr int kXSlots = 3;  // slot period = the rows-in-threes period
constexpr int kXTrash = kXWaves * 2 * kXSlots * kXQ;
constexpr int kXElems = kXTrash + 2 * kWave + kXSlots * kXQ;
// Quantity ids, published in the iteration's slot: at the iteration's start
// (P's stencil row of Q/phi, the UpdateQuantities values of the UpdateFields
// centre rows of P, A and B, A's stencil row, P's post-collision populations
// and half-step values of the two previous rows in the directions the pull
// shifts), then a barrier; after P's UpdateFields (its populations of the
// current row), then a second barrier; after A's UpdateFields (A's Q/phi
// outputs, read by B one iteration later).  Every read comes after the
// barrier that follows its write; a slot is rewritten three iterations later,
// after every wave has passed the barriers that end its reads.
enum XId {
  kXPQ = 0, kXPU = 3, kXAQ = 9, kXAU = 12, kXBU = 18, kXPCP = 24, kXHNP = 28,
  kXPC = 32, kXQA = 36, kXEnd = 39
};
static_assert(kXEnd <= kXQ, "exchange quantities fit a slot");

// Per-row outputs of P that A needs two rows later: half-step Q and the phi
// terms (FP32: Dp and old phi; FP64: phi + Dp/2).
template <class T>
struct POut {
  T qnx, qny, dp, raw;
};

// Old phi and its compensation of the output row (FP32): phic is loaded one
// iteration ahead; old phi is P's input of that row, kept in registers.
template <class T>
struct OldPhi {
  T pold, phic;
};

#ifndef LYO_FULL_WAVES_PER_EU
#define LYO_FULL_WAVES_PER_EU 2
#endif

// LDS hand-off from A to B: per wave, two rings (A's outputs f 9, Q 2, phi;
// A's half-step values fn 9, QN 2, phi term) of three rows of twelve values
// per lane, in 16-byte vectors laid out [ring][slot][vector][lane] so every
// access is a conflict-free 64-lane ds_*_b128.  B reads the rows A wrote one
// and two iterations before; the third slot is the one A writes now.
template <class T>
struct Ring {
  static constexpr int W = 16 / (int)sizeof(T);  // values per 16-byte vector
  static constexpr int G = 12 / W;               // vectors per row
  static constexpr int kVecs = 2 * 3 * G * 64;   // per wave
  typedef T V __attribute__((ext_vector_type(W)));
};
template <class T>
constexpr size_t ring_bytes_per_wave() {
  return (size_t)Ring<T>::kVecs * 16;
}

// Arguments: f_in, qxx_in, qyx_in, phi_in (and FP32 phic_in) = the state
// S[cur]; f_out, qxx_out, qyx_out, phi_out (and FP32 phic_out) = S[next].
// NPC: correctors per step, 1 or 2.
template <class T, bool EXACT, int NPC, bool XCH = false>
__device__ __forceinline__ void full_item(const SweepArgs<T>& A, const Params<T>& P,
                                          const Geo& g, const int wid,
                                          typename Ring<T>::V* ring, T* xe = nullptr) {
  using RV = typename Ring<T>::V;
  constexpr int W = Ring<T>::W, G = Ring<T>::G;
  const int lane = threadIdx.x & 63;
  auto ring_put = [&](int r, int slot, const T (&v)[12]) {
    RV* p = ring + (r * 3 + slot) * (G * 64) + lane;
#pragma unroll
    for (int q = 0; q < G; ++q) {
      RV e;
#pragma unroll
      for (int j = 0; j < W; ++j) e[j] = v[q * W + j];
      p[q * 64] = e;
    }
  };
  auto ring_get = [&](int r, int slot, T (&v)[12]) {
    const RV* p = ring + (r * 3 + slot) * (G * 64) + lane;
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const RV e = p[q * 64];
#pragma unroll
      for (int j = 0; j < W; ++j) v[q * W + j] = e[j];
    }
  };
  static_assert(NPC == 1 || NPC == 2, "one or two correctors");
  static_assert(!XCH || NPC == 2, "the block-cooperative step has two correctors");
  constexpr int HALO = NPC == 2 ? kFullHalo : kFull1Halo;
  const int strip = wid % g.nstrips;
  const int chunk = wid / g.nstrips;
  const int LX = g.LX, LY = g.LY, N = g.NP;

  // column of the lane within its strip (XCH: within the block's columns)
  const int wv = XCH ? ((int)threadIdx.x >> 6) : 0;
  const int cl = XCH ? wv * kWave + lane : lane;
  constexpr int SW = XCH ? kXWaves * kWave : kWave;
  const int ycol = strip * (SW - 2 * HALO) - HALO + cl;
  int yi = ycol % LY;
  if (yi < 0) yi += LY;
  LaneOff<T> ly{(unsigned)yi * (unsigned)sizeof(T)};
  const bool active = cl >= HALO && cl < SW - HALO && ycol < LY;

  // ---- exchange between the waves of a block (XCH)
  // write index of the lane (lane 0: its side-0 slots, lane 63: side 1,
  // others: scratch); read bases: side 1 of the left wave (the y-1
  // neighbour of lane 0), side 0 of the right wave (the y+1 neighbour of
  // lane 63).  The block's outer waves read whatever their wrapped
  // neighbours left there: those lanes are halo.
  [[maybe_unused]] const int xw =
      lane == 0 ? (wv * 2) * kXSlots * kXQ
                : (lane == kWave - 1 ? (wv * 2 + 1) * kXSlots * kXQ : kXTrash + 2 * lane);
  [[maybe_unused]] const int xr_m = (((wv + kXWaves - 1) % kXWaves) * 2 + 1) * kXSlots * kXQ;
  [[maybe_unused]] const int xr_p = (((wv + 1) % kXWaves) * 2) * kXSlots * kXQ;
  auto xpub = [&](int slot, int id, T v) {
    if constexpr (XCH) xe[xw + slot * kXQ + id] = v;
  };
  // value of the y-1 (dir < 0) or y+1 (dir > 0) neighbour; XCH: across the
  // wave boundary from the slot the neighbour published
  auto sh = [&](T v, int dir, int slot, int id) -> T {
    if constexpr (XCH) {
      const T old = xe[(dir < 0 ? xr_m : xr_p) + slot * kXQ + id];
      return dir < 0 ? from_ym_or(v, old) : from_yp_or(v, old);
    } else {
      return dir < 0 ? from_ym(v) : from_yp(v);
    }
  };
  auto xsync = [&]() {
    if constexpr (XCH) asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  };
  const int x0 = chunk * g.xc;
  const int x1 = min(x0 + g.xc, LX);

  // rows run from x0-6 to x1+6; used for LX >= 8, so one wrap in either
  // direction suffices
  auto roff = [&](int r) {
    r = (r < 0) ? r + LX : r;
    r = (r >= LX) ? r - LX : r;
    return r * LY;
  };
  auto load_qp = [&](int r) {
    const int o = roff(r);
    QP<T> q;
    q.xx = ly.ld(A.qxx_in + o);
    q.yx = ly.ld(A.qyx_in + o);
    q.p = ly.ld(A.phi_in + o);
    return q;
  };
  auto load_f = [&](int r, T(&f)[9]) {
    const int o = roff(r);
#pragma unroll
    for (int v = 0; v < 9; ++v) f[v] = ly.ld(A.f_in + (v * N + o));
  };
  auto load_old = [&](int r, OldPhi<T>& R) {
    if constexpr (!EXACT) R.phic = ly.ld(A.phic_in + roff(r));
  };
  // The last corrector's outputs of row r (Lyotropic::UpdateFieldsAtNode with
  // first = false, lyotropic.cpp:331-340) from its half-step values hv (fn 9,
  // QN 2, phi term) and increments; FP32 phi: phi_new = phi_old + phic +
  // (Dp_pred + Dp)/2 with a TwoSum, as the separate last corrector.
  // A corrector's output from its half-step value and increment.  FP64: the
  // reference's hv + inc/2.  FP32: hv holds twice the half-step value and the
  // corrector's increments come halved (update_fields HALF): s/2 + inc/2,
  // bitwise the separate sweeps' half_of(...) + inc/2, since halving is exact.
  auto hstep = [](T hv, T inc) -> T {
    if constexpr (EXACT) {
      return hv + T(.5) * inc;
    } else {
      return T(.5) * hv + inc;
    }
  };
  auto store_state = [&](int r, const T(&hv)[12], T Dxx, T Dyx, T Dp, const T(&om)[9],
                         const OldPhi<T>& old) {
    const int ox = roff(r);
    ly.st(A.qxx_out + ox, hstep(hv[9], Dxx));
    ly.st(A.qyx_out + ox, hstep(hv[10], Dyx));
#pragma unroll
    for (int v = 0; v < 9; ++v) ly.st(A.f_out + (v * N + ox), hstep(hv[v], om[v]));
    if constexpr (EXACT) {
      ly.st(A.phi_out + ox, hv[11] + T(.5) * Dp);
    } else {
      const T pold = old.pold;
      const T yv = T(.5) * (hv[11] + Dp) + old.phic;
      const T t = pold + yv;
      const T bp = t - pold;
      const T ap = t - bp;
      const T err = (pold - ap) + (yv - bp);
      ly.st(A.phi_out + ox, t);
      ly.st(A.phic_out + ox, err);
    }
  };
  // UpdateQuantities on the centre row of three stencil rows (x-1, x, x+1)
  // (slot, id: where the wave boundary values of q0 were published)
  // y neighbours of v (XCH: the wave boundary values from the slot and id
  // where the neighbour waves published v)
  auto yn = [&](T v, int slot, int id) {
    return ynb<T, EXACT>(sh(v, 1, slot, id), sh(v, -1, slot, id));
  };
  auto uq = [&](const QP<T>& qm, const QP<T>& q0, const QP<T>& qp, const T(&f)[9],
                NbrVals<T>& nv, CenterVals<T>& cv, int slot, int id) {
    update_quantities<T, EXACT>(f, q0.xx, qp.xx, qm.xx, yn(q0.xx, slot, id), q0.yx, qp.yx, qm.yx,
                                yn(q0.yx, slot, id + 1), q0.p, qp.p, qm.p,
                                yn(q0.p, slot, id + 2), P, nv, cv);
  };
  auto pub_q = [&](int slot, int id, const QP<T>& q) {
    xpub(slot, id, q.xx);
    xpub(slot, id + 1, q.yx);
    xpub(slot, id + 2, q.p);
  };
  // UpdateFields stencil of the centre row u0 (the x neighbours um, un)
  auto ufs = [&](const NbrVals<T>& um, const NbrVals<T>& u0, const NbrVals<T>& un, int slot,
                 int id) {
    if constexpr (!XCH) {
      return uf_stencil<T, EXACT>(um, u0, un);
    } else {
      UFStencil<T, EXACT> s = uf_stencil_x<T, EXACT>(um, u0, un);
      s.ux_y = yn(u0.ux, slot, id);
      s.uy_y = yn(u0.uy, slot, id + 1);
      s.uyp_y = yn(u0.uyp, slot, id + 2);
      s.sxy_y = yn(u0.sxy, slot, id + 3);
      s.syy_y = yn(u0.syy, slot, id + 4);
      s.mu_y = yn(u0.mu, slot, id + 5);
      return s;
    }
  };
  auto pub_u = [&](int slot, int id, const NbrVals<T>& u) {
    xpub(slot, id, u.ux);
    xpub(slot, id + 1, u.uy);
    xpub(slot, id + 2, u.uyp);
    xpub(slot, id + 3, u.sxy);
    xpub(slot, id + 4, u.syy);
    xpub(slot, id + 5, u.mu);
  };
  // x components of the directions (src/geometry.hpp:17-27)
  constexpr int cxv[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
  constexpr int cyv[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
  // P visits rows xs, xs+1, ...: four rows before the chunk, the chunk and
  // four rows after it; A trails by two rows and B by four (NPC = 1: two rows
  // before and after, A the last stage)
  const int xs = x0 - 2 * NPC;
  const int nI = x1 - x0 + 4 * NPC;

  // ---- stage P rolling state: stencil rows x (pqb), x+1 (pqc), x+2 (cq);
  // UpdateQuantities of rows x-1 (pum), x (pu0); populations of rows x (pfc),
  // x+1 (cf)
  NbrVals<T> pum, pu0;
  CenterVals<T> pc0;
  T pfc[9], cf[9];
  QP<T> pqb, pqc, cq;
  OldPhi<T> bold{};
  {
    const QP<T> qa = load_qp(xs - 2);
    pqb = load_qp(xs - 1);
    pqc = load_qp(xs);
    const QP<T> qd = load_qp(xs + 1);
    T ft[9];
    load_f(xs - 1, ft);
    load_f(xs, pfc);
    cq = load_qp(xs + 2);
    load_f(xs + 1, cf);
    load_old(xs - 2 * NPC, bold);
    // (XCH: slot 2 is next written in the third iteration, after every wave
    // has passed four more barriers)
    pub_q(2, kXPQ, pqb);
    pub_q(2, kXAQ, pqc);
    xsync();
    CenterVals<T> ct;
    uq(qa, pqb, pqc, ft, pum, ct, 2, kXPQ);
    uq(pqb, pqc, qd, pfc, pu0, pc0, 2, kXAQ);
    pqb = pqc;
    pqc = qd;
  }
  // post-collision populations f + om and their half-step values of the rows
  // x-1 (pcP, hnP) and x-2 (pcPP, hnPP); P outputs of rows x-1, x-2
  T pcP[9] = {}, hnP[9] = {}, pcPP[9] = {}, hnPP[9] = {};
  POut<T> po1{}, po2{};
  // FP32, NPC = 2: old phi (P's input) of rows x-3, x-4
  [[maybe_unused]] T raw3 = T(0), raw4 = T(0);
  // ---- stage A: stencil rows x-2 (aqb), x-1 (aqc) of (Q*, phi*);
  // UpdateQuantities of rows x-3 (aum), x-2 (au0); streamed populations and
  // half-step values of row x-2 (afc, ahn)
  QP<T> aqb{}, aqc{};
  NbrVals<T> aum{}, au0{};
  CenterVals<T> ac0{};
  T afc[9] = {}, ahn[9] = {};
  // ---- stage B, as in pair_item, two rows further back; A's outputs and
  // half-step values of the rows x-3, x-4 come from the LDS rings
  NbrVals<T> bum{}, bu0{};
  CenterVals<T> bc0{};

  // One row of the march (P on row x).  WARM: the first 4*NPC rows, where
  // the later stages have no inputs yet; afterwards every stage runs and the
  // rows go in threes, so that the rolling rows (three deep) and the ring
  // slots (period three) are renamed rather than copied.
  auto row = [&](const int i, const int x, const int s0, auto warm) {
    constexpr bool WARM = decltype(warm)::value;
    const int s1 = (s0 == 0) ? 2 : s0 - 1;  // slot of iteration i-1
    const int s2 = (s1 == 0) ? 2 : s1 - 1;  // slot of iteration i-2
    if constexpr (XCH) {
      // the boundary values this iteration's stencils and pull need that
      // earlier iterations produced (read after the barrier)
      pub_q(s0, kXPQ, pqc);
      pub_u(s0, kXPU, pu0);
      pub_q(s0, kXAQ, aqc);
      pub_u(s0, kXAU, au0);
      pub_u(s0, kXBU, bu0);
      xpub(s0, kXPCP, pcP[3]);
      xpub(s0, kXPCP + 1, pcP[4]);
      xpub(s0, kXPCP + 2, pcPP[5]);
      xpub(s0, kXPCP + 3, pcPP[8]);
      xpub(s0, kXHNP, hnP[3]);
      xpub(s0, kXHNP + 1, hnP[4]);
      xpub(s0, kXHNP + 2, hnPP[5]);
      xpub(s0, kXHNP + 3, hnPP[8]);
      xsync();
    }
    // ---- P: UpdateQuantities on row x+1, UpdateFields on row x
    NbrVals<T> pun;
    CenterVals<T> pcn;
    uq(pqb, pqc, cq, cf, pun, pcn, s0, kXPQ);
    T pc[9], hn[9];
    QP<T> aq_new;  // A's stencil values of row x
    POut<T> po;
    {
      const UFStencil<T, EXACT> s = ufs(pum, pu0, pun, s0, kXPU);
      T Dxx, Dyx, Dp, om[9];
      // (FP32: om receives the post-collision populations f + Omega)
      update_fields<T, EXACT, !EXACT>(pc0, pqb.xx, pqb.yx, pqb.p, pfc, s, P, Dxx, Dyx, Dp, om);
      const T Qsx = pqb.xx + Dxx, Qsy = pqb.yx + Dyx;
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        if constexpr (EXACT) {
          pc[v] = pfc[v] + om[v];
          hn[v] = pfc[v] + T(.5) * om[v];
        } else {
          pc[v] = om[v];
          hn[v] = pc[v] + pfc[v];  // twice the half-step value
        }
      }
      aq_new.xx = Qsx;
      aq_new.yx = Qsy;
      // the first corrector's stencil phi (FP64: phi + Dp as the predictor
      // stores it; FP32: old phi plus the stored increment -- the same sum)
      aq_new.p = pqb.p + Dp;
      if constexpr (EXACT) {
        po.qnx = pqb.xx + T(.5) * Dxx;
        po.qny = pqb.yx + T(.5) * Dyx;
        po.dp = pqb.p + T(.5) * Dp;
        po.raw = T(0);
      } else {
        po.qnx = Qsx + pqb.xx;
        po.qny = Qsy + pqb.yx;
        po.dp = Dp;
        po.raw = pqb.p;
      }
    }
    pum = pu0;
    pu0 = pun;
    pc0 = pcn;
#pragma unroll
    for (int v = 0; v < 9; ++v) pfc[v] = cf[v];
    pqb = pqc;
    pqc = cq;
    // P's loads of the next iteration (unconditional: past the chunk the
    // wrapped rows are in bounds and unused); in flight while A and B compute
    cq = load_qp(x + 3);
    load_f(x + 2, cf);

    // ---- pull (Move): streamed populations of row x-1
    if constexpr (XCH) {
      xpub(s0, kXPC, pc[6]);
      xpub(s0, kXPC + 1, pc[7]);
      xpub(s0, kXPC + 2, hn[6]);
      xpub(s0, kXPC + 3, hn[7]);
      xsync();
    }
    // exchange ids of the shifted directions: 3, 4 (row x-1), 5, 8 (row
    // x-2) published at the iteration's start, 6, 7 (row x) just now
    constexpr int xpc[9] = {0, 0, 0, kXPCP, kXPCP + 1, kXPCP + 2, kXPC, kXPC + 1, kXPCP + 3};
    constexpr int xhn[9] = {0, 0, 0, kXHNP, kXHNP + 1, kXHNP + 2, kXPC + 2, kXPC + 3, kXHNP + 3};
    T fst[9], hst[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      T a, h;
      if (cxv[v] == -1) {
        a = pc[v];
        h = hn[v];
      } else if (cxv[v] == 0) {
        a = pcP[v];
        h = hnP[v];
      } else {
        a = pcPP[v];
        h = hnPP[v];
      }
      if (cyv[v] != 0) {
        a = sh(a, -cyv[v], s0, xpc[v]);
        h = sh(h, -cyv[v], s0, xhn[v]);
      }
      fst[v] = a;
      hst[v] = h;
    }
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      pcPP[v] = pcP[v];
      hnPP[v] = hnP[v];
      pcP[v] = pc[v];
      hnP[v] = hn[v];
    }

    // ---- A: UpdateQuantities on row x-1, UpdateFields on row x-2
    QP<T> qa_out{};
    if (!WARM || i >= 2) {  // wave-uniform
      NbrVals<T> aun;
      CenterVals<T> acn;
      uq(aqb, aqc, aq_new, fst, aun, acn, s0, kXAQ);
      if (!WARM || i >= 4) {
        // half-step values of row x-2 (NPC = 2: kept for B two rows later)
        T cen[12];
#pragma unroll
        for (int v = 0; v < 9; ++v) cen[v] = ahn[v];
        cen[9] = po2.qnx;
        cen[10] = po2.qny;
        cen[11] = po2.dp;
        if constexpr (NPC == 2) ring_put(1, s0, cen);
        const UFStencil<T, EXACT> s = ufs(aum, au0, aun, s0, kXAU);
        T Dxx, Dyx, Dp, om[9];
        update_fields<T, EXACT, false, !EXACT>(ac0, aqb.xx, aqb.yx, aqb.p, afc, s, P, Dxx, Dyx,
                                               Dp, om);
        if constexpr (NPC == 2) {
          T ao[12];
          qa_out.xx = hstep(cen[9], Dxx);
          qa_out.yx = hstep(cen[10], Dyx);
#pragma unroll
          for (int v = 0; v < 9; ++v) ao[v] = hstep(cen[v], om[v]);
          if constexpr (EXACT) {
            qa_out.p = cen[11] + T(.5) * Dp;
          } else {
            // the separate sweep stores the increment, the next one adds it
            // to old phi: the same two roundings here
            const T inc = T(.5) * (cen[11] + Dp);
            qa_out.p = po2.raw + inc;
          }
          ao[9] = qa_out.xx;
          ao[10] = qa_out.yx;
          ao[11] = qa_out.p;
          ring_put(0, s0, ao);
          pub_q(s0, kXQA, qa_out);  // B's stencil row one iteration later
        } else {
          // the last corrector: the new state of row x-2
          wait_vmem<EXACT>();
          if (active) store_state(x - 2, cen, Dxx, Dyx, Dp, om, {po2.raw, bold.phic});
        }
      }
      aum = au0;
      au0 = aun;
      ac0 = acn;
    }
    aqb = aqc;
    aqc = aq_new;
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      afc[v] = fst[v];
      ahn[v] = hst[v];
    }
    [[maybe_unused]] const T raw2 = po2.raw;
    po2 = po1;
    po1 = po;

    // ---- B: UpdateQuantities on row x-3, UpdateFields on row x-4
    if constexpr (NPC == 2) {
      if (!WARM || i >= 6) {  // wave-uniform
        T ao1[12], ao2[12];  // A's outputs of rows x-3, x-4
        ring_get(0, s1, ao1);
        ring_get(0, s2, ao2);
        const QP<T> bqc{ao1[9], ao1[10], ao1[11]}, bqb{ao2[9], ao2[10], ao2[11]};
        T bfp[9], bfc[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) {
          bfp[v] = ao1[v];
          bfc[v] = ao2[v];
        }
        NbrVals<T> bun;
        CenterVals<T> bcn;
        uq(bqb, bqc, qa_out, bfp, bun, bcn, s1, kXQA);
        if (!WARM || i >= 8) {
          const UFStencil<T, EXACT> s = ufs(bum, bu0, bun, s0, kXBU);
          T Dxx, Dyx, Dp, om[9];
          update_fields<T, EXACT, false, !EXACT>(bc0, bqb.xx, bqb.yx, bqb.p, bfc, s, P, Dxx, Dyx,
                                                 Dp, om);
          T cmv[12];  // half-step values of row x-4
          ring_get(1, s2, cmv);
          wait_vmem<EXACT>();
          if (active) store_state(x - 4, cmv, Dxx, Dyx, Dp, om, {raw4, bold.phic});
        }
        bum = bu0;
        bu0 = bun;
        bc0 = bcn;
      }
    }
    // old phi of the next iteration's output row
    load_old(x + 1 - 2 * NPC, bold);
    if constexpr (NPC == 2 && !EXACT) {
      raw4 = raw3;
      raw3 = raw2;
    }
  };
  constexpr int kWarm = 4 * NPC;  // nI >= kWarm + 1: a chunk has a row
  int i = 0;
  for (; i < kWarm; ++i) row(i, xs + i, i % 3, std::true_type{});
  wait_vmem<EXACT>();  // the loop's entry then has no loads in flight either
  for (; i + 3 <= nI; i += 3) {
    row(i, xs + i, kWarm % 3, std::false_type{});
    row(i + 1, xs + i + 1, (kWarm + 1) % 3, std::false_type{});
    row(i + 2, xs + i + 2, (kWarm + 2) % 3, std::false_type{});
  }
  for (; i < nI; ++i) row(i, xs + i, i % 3, std::false_type{});
}

// Dynamic LDS of a launch: one pair of rings per wave of the block.
template <class T>
__device__ __forceinline__ typename Ring<T>::V* wave_ring() {
  extern __shared__ __attribute__((aligned(16))) unsigned char full_lds[];
  return reinterpret_cast<typename Ring<T>::V*>(full_lds) +
         ((int)threadIdx.x >> 6) * Ring<T>::kVecs;
}

template <class T, bool EXACT, int NPC>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(EXACT ? 1 : LYO_FULL_WAVES_PER_EU)))
full_kernel(const SweepArgs<T> A, const Params<T> P, const Geo g) {
  int b = blockIdx.x;
  if (g.xcd) {
    const int nb8 = (int)gridDim.x >> 3;
    if (b < (nb8 << 3)) b = (b & 7) * nb8 + (b >> 3);
  }
  const int wid = __builtin_amdgcn_readfirstlane(b * ((int)blockDim.x >> 6) +
                                                 ((int)threadIdx.x >> 6));
  if (wid >= g.nwaves) return;
  full_item<T, EXACT, NPC>(A, P, g, wid, wave_ring<T>());
}

// Batched one-kernel steps: persistent grid over the work items of many runs.
template <class T, bool EXACT, int NPC>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(EXACT ? 1 : LYO_FULL_WAVES_PER_EU)))
full_batch_kernel(const SweepJob<T>* __restrict__ jobs, const int njobs, const int nitems) {
  const int nw = (int)(gridDim.x * (blockDim.x >> 6));
  int item = __builtin_amdgcn_readfirstlane((int)blockIdx.x * ((int)blockDim.x >> 6) +
                                            ((int)threadIdx.x >> 6));
  for (; item < nitems; item += nw) {
    const SweepJob<T>& J = find_job(jobs, njobs, item);
    full_item<T, EXACT, NPC>(J.A, J.P, J.g, item - J.first, wave_ring<T>());
  }
}

// The block-cooperative one-kernel step (npc = 2): one block of kXWaves waves
// per item (block-uniform, so every wave of a block passes the same
// barriers).
template <class T>
__device__ __forceinline__ T* block_exchange() {
  extern __shared__ __attribute__((aligned(16))) unsigned char full_lds[];
  return reinterpret_cast<T*>(full_lds + kXWaves * ring_bytes_per_wave<T>());
}

template <class T, bool EXACT>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(EXACT ? 1 : LYO_FULL_WAVES_PER_EU)))
full_x_kernel(const SweepArgs<T> A, const Params<T> P, const Geo g) {
  int b = blockIdx.x;
  if (g.xcd) {
    const int nb8 = (int)gridDim.x >> 3;
    if (b < (nb8 << 3)) b = (b & 7) * nb8 + (b >> 3);
  }
  if (b >= g.nwaves) return;  // the whole block
  full_item<T, EXACT, 2, true>(A, P, g, __builtin_amdgcn_readfirstlane(b), wave_ring<T>(),
                               block_exchange<T>());
}

template <class T>
constexpr size_t full_x_lds_bytes() {
  return (size_t)kXWaves * ring_bytes_per_wave<T>() + (size_t)kXElems * sizeof(T);
}

template <class T, bool EXACT>
inline hipError_t launch_full_x(const SweepArgs<T>& a, const Params<T>& p, const Geo& g,
                                hipStream_t s) {
  if (!geo_ok(g, kXStrip, 8)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((full_x_kernel<T, EXACT>), dim3(g.nwaves), dim3(kXWaves * 64),
                     full_x_lds_bytes<T>(), s, a, p, g);
  return hipGetLastError();
}

// LDS of a block: the rings of its waves (NPC = 2 only)
template <class T, int NPC>
constexpr size_t full_lds_bytes(int waves_per_block) {
  return NPC == 2 ? (size_t)waves_per_block * ring_bytes_per_wave<T>() : 0;
}

template <class T, bool EXACT, int NPC>
inline hipError_t launch_full(const SweepArgs<T>& a, const Params<T>& p, const Geo& g,
                              int waves_per_block, hipStream_t s) {
  if (!geo_ok(g, NPC == 2 ? kFullStrip : kFull1Strip, 8) || waves_per_block < 1)
    return hipErrorInvalidValue;
  const int threads = 64 * waves_per_block;
  const int blocks = (g.nwaves + waves_per_block - 1) / waves_per_block;
  const size_t lds = full_lds_bytes<T, NPC>(waves_per_block);
  hipLaunchKernelGGL((full_kernel<T, EXACT, NPC>), dim3(blocks), dim3(threads), lds, s, a, p, g);
  return hipGetLastError();
}

}  // namespace lyo
