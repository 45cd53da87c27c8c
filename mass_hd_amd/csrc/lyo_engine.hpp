// lyo_engine.hpp -- buffer sets and sweep sequencing of one lyotropic run.
//
// Sequencing follows Lyotropic::Step (src/models/lyotropic.cpp:500-525): one
// predictor sweep, then npc corrector sweeps.  Each sweep reads one buffer
// set and writes another, so no sweep ever reads what it writes:
//   S[2]  model state (f, Q, phi [, phic]) -- ping-pong across steps
//   P     predictor outputs: streamed f*, fn*, Q+D, Q+D/2, phi terms
//   C[2]  intermediate corrector outputs (npc >= 2) -- ping-pong
//   X     FP32 surrogate steps: the one-kernel network step's output, which
//         then trades places with S[cur] (allocated on first use)
#pragma once

#include <random>
#include <utility>
#include <vector>

#include "../../include/mass_lyo.h"
#include "lyo_aux.hpp"
#include "lyo_diag.hpp"
#include "lyo_full.hpp"
#include "lyo_loops.hpp"
#include "lyo_net_engine.hpp"
#include "lyo_net_fused_api.hpp"
#include "lyo_pair.hpp"
#include "lyo_sweep.hpp"

namespace lyo {

template <class T>
struct Bufs {
  T* Sf[2] = {};
  T* Sqxx[2] = {};
  T* Sqyx[2] = {};
  T* Sphi[2] = {};
  T* Sphic[2] = {};  // FP32 only
  T* Pfs = nullptr;
  T* Pfns = nullptr;
  T* Pqsxx = nullptr;
  T* Pqsyx = nullptr;
  T* Pqnxx = nullptr;
  T* Pqnyx = nullptr;
  T* Pphi = nullptr;  // EXACT: phi + Dp; FP32: Dp
  T* Pphn = nullptr;  // EXACT: phi + Dp/2
  T* Cf[2] = {};
  T* Cqxx[2] = {};
  T* Cqyx[2] = {};
  T* Cphi[2] = {};  // EXACT: phi; FP32: increment of phi
  T* Xf = nullptr;
  T* Xqxx = nullptr;
  T* Xqyx = nullptr;
  T* Xphi = nullptr;
  T* Xphic = nullptr;
  int cur = 0;
  // inputs of the last UpdateQuantities (start of the last corrector sweep)
  bool have_uq = false;
  const T* uq_f = nullptr;
  const T* uq_qxx = nullptr;
  const T* uq_qyx = nullptr;
  const T* uq_phi = nullptr;
  const T* uq_dphi = nullptr;
  // >= 0: the last step ran its last two correctors fused, so the outputs of
  // corrector npc-1 (the inputs of the last UpdateQuantities) were never
  // stored; they are recomputed on demand from the predictor outputs and the
  // old state S[lazy_cur], which stay intact until the next step
  int lazy_cur = -1;
  bool lazy_full = false;  // ... and the predictor's outputs too (one-kernel step)
  // FP64 reference loop structure (lyo_loops.hpp): the 23 derived planes and
  // the predictor's post-collision f, fn before Move
  T* Ld = nullptr;
  T* Lfc = nullptr;
  T* Lfnc = nullptr;
};

struct LaunchCfg {
  int xc = 32;   // rows per wave chunk
  int wpb = 4;   // waves per block
  int xcd = 1;   // XCD-aware block order
  int fuse = 1;  // last two correctors in one kernel (npc >= 2)
  int xc2 = 16;  // rows per wave chunk of the fused pair
  int alt = 1;   // alternate the march direction between neighbouring chunks
  int full = 1;  // npc <= 2: the whole step in one kernel (2: the block-cooperative
                 // variant for npc = 2, single runs)
  int xc3 = 32;  // rows per wave chunk of the one-kernel step
  int xc3b = 0;  // ... of its one-corrector form when the run's npc is 2 (0: xc3)
  int xc4 = 32;  // rows per block chunk of the block-cooperative one-kernel step
  int loops = 0; // FP64: one kernel per loop of the reference's step instead
  int xc5 = 32;  // rows per wave chunk of the one-kernel network step (surrogate)
};

// Optional per-sweep event timing.
struct SweepTimer {
  std::vector<hipEvent_t> ev;  // pairs
  std::vector<int> kind;       // 0 predictor, 1 corrector, 2 fused pair, 3 whole step
  size_t used = 0;
};

// kind 0: sweep, 1: fused pair, 2: one-kernel step, 3: one-kernel step with
// one corrector, 4: block-cooperative one-kernel step (strip width and chunk
// rows; kind 4 counts blocks in nwaves)
constexpr int kGeoKinds = 5;
inline Geo make_geo(int LX, int LY, int NP, const LaunchCfg& L, int kind = 0) {
  Geo g;
  g.LX = LX;
  g.LY = LY;
  g.N = LX * LY;
  g.NP = NP;
  const int strip =
      kind == 4 ? kXStrip
                : (kind == 3 ? kFull1Strip
                             : (kind == 2 ? kFullStrip : (kind == 1 ? kPairStrip : kStrip)));
  g.nstrips = (LY + strip - 1) / strip;
  g.xc = kind == 4 ? L.xc4
       : kind == 3 ? (L.xc3b > 0 ? L.xc3b : L.xc3)
       : (kind == 2 ? L.xc3 : (kind == 1 ? L.xc2 : L.xc));
  g.nchunks = (LX + g.xc - 1) / g.xc;
  g.nwaves = g.nstrips * g.nchunks;
  g.xcd = L.xcd;
  g.alt = L.alt;
  return g;
}

template <class T, bool EXACT, int STAGE, bool HB>
inline hipError_t timed_launch(const SweepArgs<T>& a, const Params<T>& p, const Geo& g,
                               const LaunchCfg& L, hipStream_t s, SweepTimer* tm) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (tm && tm->used + 2 <= tm->ev.size()) {
    e0 = tm->ev[tm->used];
    e1 = tm->ev[tm->used + 1];
    tm->kind[tm->used / 2] = (STAGE == PRED) ? 0 : 1;
    tm->used += 2;
    (void)hipEventRecord(e0, s);
  }
  hipError_t err = launch_sweep<T, EXACT, STAGE, HB>(a, p, g, L.wpb, s);
  if (e1) (void)hipEventRecord(e1, s);
  return err;
}

template <class T, bool EXACT, bool HB>
inline hipError_t timed_pair(const SweepArgs<T>& a, const Params<T>& p, const Geo& g,
                             const LaunchCfg& L, hipStream_t s, SweepTimer* tm) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (tm && tm->used + 2 <= tm->ev.size()) {
    e0 = tm->ev[tm->used];
    e1 = tm->ev[tm->used + 1];
    tm->kind[tm->used / 2] = 2;
    tm->used += 2;
    (void)hipEventRecord(e0, s);
  }
  hipError_t err = launch_pair<T, EXACT, HB>(a, p, g, L.wpb, s);
  if (e1) (void)hipEventRecord(e1, s);
  return err;
}

template <class T, bool EXACT, int NPC, bool XCH = false>
inline hipError_t timed_full(const SweepArgs<T>& a, const Params<T>& p, const Geo& g,
                             const LaunchCfg& L, hipStream_t s, SweepTimer* tm) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (tm && tm->used + 2 <= tm->ev.size()) {
    e0 = tm->ev[tm->used];
    e1 = tm->ev[tm->used + 1];
    tm->kind[tm->used / 2] = 3;
    tm->used += 2;
    (void)hipEventRecord(e0, s);
  }
  hipError_t err = XCH ? launch_full_x<T, EXACT>(a, p, g, s)
                       : launch_full<T, EXACT, NPC>(a, p, g, L.wpb, s);
  if (e1) (void)hipEventRecord(e1, s);
  return err;
}

// Buffer roles of sweep c of a step whose state is S[cur]: c = 0 is the
// predictor, c = 1..npc the correctors (Lyotropic::Step,
// src/models/lyotropic.cpp:500-525).
template <class T, bool EXACT>
SweepArgs<T> stage_args(const Bufs<T>& B, int npc, int c, int cur) {
  const int nxt = cur ^ 1;
  SweepArgs<T> a{};
  if (c == 0) {
    a.f_in = B.Sf[cur];
    a.qxx_in = B.Sqxx[cur];
    a.qyx_in = B.Sqyx[cur];
    a.phi_in = B.Sphi[cur];
    a.f_out = B.Pfs;
    a.fn_out = B.Pfns;
    a.qxx_out = B.Pqsxx;
    a.qyx_out = B.Pqsyx;
    a.qnxx_out = B.Pqnxx;
    a.qnyx_out = B.Pqnyx;
    a.phi_out = B.Pphi;
    a.phn_out = B.Pphn;
    return a;
  }
  if (c == 1) {
    a.f_in = B.Pfs;
    a.qxx_in = B.Pqsxx;
    a.qyx_in = B.Pqsyx;
    if constexpr (EXACT) {
      a.phi_in = B.Pphi;
    } else {
      a.phi_in = B.Sphi[cur];
      a.dphi_in = B.Pphi;
    }
  } else {
    const int pv = (c - 2) & 1;
    a.f_in = B.Cf[pv];
    a.qxx_in = B.Cqxx[pv];
    a.qyx_in = B.Cqyx[pv];
    if constexpr (EXACT) {
      a.phi_in = B.Cphi[pv];
    } else {
      a.phi_in = B.Sphi[cur];
      a.dphi_in = B.Cphi[pv];
    }
  }
  a.fn_in = B.Pfns;
  a.qnxx_in = B.Pqnxx;
  a.qnyx_in = B.Pqnyx;
  a.phn_in = EXACT ? B.Pphn : B.Pphi;
  a.fold_in = B.Sf[cur];
  a.qoxx_in = B.Sqxx[cur];
  a.qoyx_in = B.Sqyx[cur];
  if (c == npc) {
    a.f_out = B.Sf[nxt];
    a.qxx_out = B.Sqxx[nxt];
    a.qyx_out = B.Sqyx[nxt];
    a.phi_out = B.Sphi[nxt];
    if constexpr (!EXACT) {
      a.phic_in = B.Sphic[cur];
      a.phic_out = B.Sphic[nxt];
    }
  } else {
    const int ov = (c - 1) & 1;
    a.f_out = B.Cf[ov];
    a.qxx_out = B.Cqxx[ov];
    a.qyx_out = B.Cqyx[ov];
    a.phi_out = B.Cphi[ov];
  }
  return a;
}

// The last corrector's inputs are the inputs of the last UpdateQuantities:
// the frame's derived fields are recomputed from them.
template <class T>
void note_last_uq(Bufs<T>& B, const SweepArgs<T>& a) {
  B.have_uq = true;
  B.uq_f = a.f_in;
  B.uq_qxx = a.qxx_in;
  B.uq_qyx = a.qyx_in;
  B.uq_phi = a.phi_in;
  B.uq_dphi = a.dphi_in;
}

// Arguments of the fused last pair: the inputs of corrector npc-1 and the
// outputs (and, FP32, compensation input) of corrector npc.
template <class T, bool EXACT>
SweepArgs<T> pair_args(const Bufs<T>& B, int npc, int cur) {
  SweepArgs<T> a = stage_args<T, EXACT>(B, npc, npc - 1, cur);
  const SweepArgs<T> b = stage_args<T, EXACT>(B, npc, npc, cur);
  a.f_out = b.f_out;
  a.qxx_out = b.qxx_out;
  a.qyx_out = b.qyx_out;
  a.phi_out = b.phi_out;
  a.phic_in = b.phic_in;
  a.phic_out = b.phic_out;
  return a;
}

// Launch plan of one step (Lyotropic::Step, src/models/lyotropic.cpp:500-525):
// a list of stages, each a separate sweep c (0 = predictor, 1..npc the
// correctors), the fused last corrector pair, or the whole step in one kernel
// (FP32, npc = 2).  Single runs and batches follow the same plan, so a batched
// run equals the same run alone bit for bit.
constexpr int kStagePair = 100;
constexpr int kStageFull = 200;
constexpr int kStageFull1 = 201;  // the whole step with npc = 1
constexpr int kStageFullX = 202;  // the whole step, npc = 2, block-cooperative

inline std::vector<int> step_plan(bool exact, int npc, const LaunchCfg& L, int LX) {
  const bool fuse = L.fuse && npc >= 2 && LX >= 8;  // the fused kernels wrap rows once
  std::vector<int> p;
  (void)exact;
  if (fuse && L.full && npc == 2) {
    p.push_back(L.full == 2 ? kStageFullX : kStageFull);
    return p;
  }
  if (L.full && npc == 1 && LX >= 8) {
    p.push_back(kStageFull1);
    return p;
  }
  for (int c = 0; c <= npc; ++c) {
    if (fuse && c == npc - 1) {
      p.push_back(kStagePair);
      break;
    }
    p.push_back(c);
  }
  return p;
}

// FP32 half-step shortcut (SweepArgs, flag HB): the first corrector rebuilds fn
// and QN from its own inputs; the predictor stores them only when a later
// corrector loads them (a separate sweep c >= 2, or stage A of a pair with
// npc >= 3); the fused pair rebuilds them when its stage A is corrector 1.
inline bool plan_stores_half(const std::vector<int>& p, int npc) {
  for (int sc : p)
    if ((sc >= 2 && sc <= npc) || (sc == kStagePair && npc >= 3)) return true;
  return false;
}

inline int stage_geo_kind(int sc) {
  return sc == kStageFullX
             ? 4
             : (sc == kStageFull1 ? 3 : (sc == kStageFull ? 2 : (sc == kStagePair ? 1 : 0)));
}
static_assert(kGeoKinds == 5, "stage_geo_kind returns kinds 0..4");
inline bool is_full_stage(int sc) {
  return sc == kStageFull || sc == kStageFull1 || sc == kStageFullX;
}

// Arguments of the one-kernel step: state S[cur] in, S[next] out.
template <class T, bool EXACT>
SweepArgs<T> full_args(const Bufs<T>& B, int npc, int cur) {
  SweepArgs<T> a{};
  const SweepArgs<T> b = stage_args<T, EXACT>(B, npc, npc, cur);
  a.f_in = B.Sf[cur];
  a.qxx_in = B.Sqxx[cur];
  a.qyx_in = B.Sqyx[cur];
  a.phi_in = B.Sphi[cur];
  a.phic_in = B.Sphic[cur];
  a.f_out = b.f_out;
  a.qxx_out = b.qxx_out;
  a.qyx_out = b.qyx_out;
  a.phi_out = b.phi_out;
  a.phic_out = b.phic_out;
  return a;
}

template <class T, bool EXACT>
SweepArgs<T> plan_args(const Bufs<T>& B, int npc, int sc, int cur) {
  if (is_full_stage(sc)) return full_args<T, EXACT>(B, npc, cur);
  if (sc == kStagePair) return pair_args<T, EXACT>(B, npc, cur);
  return stage_args<T, EXACT>(B, npc, sc, cur);
}

template <class T, bool EXACT>
hipError_t launch_corrector(const SweepArgs<T>& a, const Params<T>& P, const Geo& g,
                            const LaunchCfg& L, hipStream_t s, SweepTimer* tm, int c, bool last) {
  if constexpr (!EXACT) {
    if (c == 1) {
      return last ? timed_launch<T, EXACT, CORR_LAST, true>(a, P, g, L, s, tm)
                  : timed_launch<T, EXACT, CORR, true>(a, P, g, L, s, tm);
    }
  }
  return last ? timed_launch<T, EXACT, CORR_LAST, false>(a, P, g, L, s, tm)
              : timed_launch<T, EXACT, CORR, false>(a, P, g, L, s, tm);
}

// One stage of a single run.
template <class T, bool EXACT>
hipError_t launch_stage(int sc, const SweepArgs<T>& a, const Params<T>& P, const Geo& g,
                        const LaunchCfg& L, int npc, bool store_half, hipStream_t s,
                        SweepTimer* tm) {
  if (sc == kStageFull) return timed_full<T, EXACT, 2>(a, P, g, L, s, tm);
  if (sc == kStageFullX) return timed_full<T, EXACT, 2, true>(a, P, g, L, s, tm);
  if (sc == kStageFull1) return timed_full<T, EXACT, 1>(a, P, g, L, s, tm);
  if (sc == kStagePair) {
    if constexpr (!EXACT) {
      if (npc == 2) return timed_pair<T, EXACT, true>(a, P, g, L, s, tm);
    }
    return timed_pair<T, EXACT, false>(a, P, g, L, s, tm);
  }
  if (sc == 0) {
    if constexpr (!EXACT) {
      if (store_half) return timed_launch<T, EXACT, PRED, true>(a, P, g, L, s, tm);
    }
    return timed_launch<T, EXACT, PRED, false>(a, P, g, L, s, tm);
  }
  return launch_corrector<T, EXACT>(a, P, g, L, s, tm, sc, sc == npc);
}

// Bookkeeping after a step whose state was S[cur]: where the inputs of the
// last UpdateQuantities are (frame fields, surrogate), or that they have to be
// recomputed because a fused stage kept them in registers.
template <class T, bool EXACT>
void note_step(Bufs<T>& B, int npc, const std::vector<int>& plan, int cur) {
  const int last = plan.back();
  B.lazy_cur = -1;
  B.lazy_full = false;
  if (last == kStagePair || is_full_stage(last)) {
    B.have_uq = false;
    B.lazy_cur = cur;
    B.lazy_full = is_full_stage(last);
  } else {
    note_last_uq(B, stage_args<T, EXACT>(B, npc, npc, cur));
  }
}

// The reference's loop structure (lyo_loops.hpp): UpdateQuantities,
// UpdateFields(true), Move, then npc x (UpdateQuantities, UpdateFields(false)),
// each a kernel over all nodes.  The predictor's outputs land where the
// predictor sweep puts them and every corrector reads and writes the buffers
// of the separate corrector sweeps, so the state after a step, the frame
// fields and the sums are those of the fused kernels.
template <class T, bool EXACT>
hipError_t run_loop_steps(Bufs<T>& B, const Params<T>& P, int LX, int LY, int NP, int npc,
                          int nsteps, hipStream_t s) {
  if constexpr (!EXACT) {
    return hipErrorInvalidValue;
  } else {
    if (!B.Ld || !B.Lfc || !B.Lfnc) return hipErrorInvalidValue;
    const int N = LX * LY, bs = 256, nb = (N + bs - 1) / bs;
    for (int st = 0; st < nsteps; ++st) {
      const int cur = B.cur;
      const SweepArgs<T> a0 = stage_args<T, EXACT>(B, npc, 0, cur);
      hipLaunchKernelGGL((loop_uq_kernel<T, EXACT>), dim3(nb), dim3(bs), 0, s, a0.f_in, a0.qxx_in,
                         a0.qyx_in, a0.phi_in, P, LX, LY, NP, B.Ld);
      LoopUF<T> u{};
      u.D = B.Ld;
      u.f = a0.f_in;
      u.qxx = a0.qxx_in;
      u.qyx = a0.qyx_in;
      u.phi = a0.phi_in;
      u.f_out = B.Lfc;
      u.fn_out = B.Lfnc;
      u.qxx_out = a0.qxx_out;
      u.qyx_out = a0.qyx_out;
      u.qnxx_out = a0.qnxx_out;
      u.qnyx_out = a0.qnyx_out;
      u.phi_out = a0.phi_out;
      u.phn_out = a0.phn_out;
      hipLaunchKernelGGL((loop_uf_kernel<T, EXACT, true>), dim3(nb), dim3(bs), 0, s, u, P, LX, LY,
                         NP);
      hipLaunchKernelGGL((loop_move_kernel<T>), dim3(nb), dim3(bs), 0, s, (const T*)B.Lfc,
                         (const T*)B.Lfnc, a0.f_out, a0.fn_out, LX, LY, NP);
      for (int c = 1; c <= npc; ++c) {
        const SweepArgs<T> a = stage_args<T, EXACT>(B, npc, c, cur);
        hipLaunchKernelGGL((loop_uq_kernel<T, EXACT>), dim3(nb), dim3(bs), 0, s, a.f_in,
                           a.qxx_in, a.qyx_in, a.phi_in, P, LX, LY, NP, B.Ld);
        LoopUF<T> w{};
        w.D = B.Ld;
        w.f = a.f_in;
        w.qxx = a.qxx_in;
        w.qyx = a.qyx_in;
        w.phi = a.phi_in;
        w.fn = a.fn_in;
        w.qnxx = a.qnxx_in;
        w.qnyx = a.qnyx_in;
        w.phn = a.phn_in;
        w.f_out = a.f_out;
        w.qxx_out = a.qxx_out;
        w.qyx_out = a.qyx_out;
        w.phi_out = a.phi_out;
        hipLaunchKernelGGL((loop_uf_kernel<T, EXACT, false>), dim3(nb), dim3(bs), 0, s, w, P, LX,
                           LY, NP);
      }
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      note_step<T, EXACT>(B, npc, std::vector<int>{npc}, cur);
      B.cur = cur ^ 1;
    }
    return hipSuccess;
  }
}

template <class T, bool EXACT>
hipError_t run_steps(Bufs<T>& B, const Params<T>& P, int LX, int LY, int NP, int npc,
                     int nsteps, const LaunchCfg& L, hipStream_t s, SweepTimer* tm) {
  if (L.loops) return run_loop_steps<T, EXACT>(B, P, LX, LY, NP, npc, nsteps, s);
  const std::vector<int> plan = step_plan(EXACT, npc, L, LX);
  const bool store_half = !EXACT && plan_stores_half(plan, npc);  // EXACT always stores
  Geo geo[kGeoKinds];
  for (int k = 0; k < kGeoKinds; ++k) geo[k] = make_geo(LX, LY, NP, L, k);
  for (int st = 0; st < nsteps; ++st) {
    const int cur = B.cur;
    for (int sc : plan) {
      const hipError_t e = launch_stage<T, EXACT>(sc, plan_args<T, EXACT>(B, npc, sc, cur), P,
                                                  geo[stage_geo_kind(sc)], L, npc, store_half, s,
                                                  tm);
      if (e != hipSuccess) return e;
    }
    note_step<T, EXACT>(B, npc, plan, cur);
    B.cur = cur ^ 1;
  }
  return hipSuccess;
}

// Recomputes the outputs of corrector npc-1 of the last step when that step
// ran its last pair fused: the separate sweep on the same inputs, so the
// inputs of the last UpdateQuantities exist again.
template <class T, bool EXACT>
hipError_t materialize_uq(Bufs<T>& B, const Params<T>& P, int LX, int LY, int NP, int npc,
                          hipStream_t s) {
  if (B.lazy_cur < 0) return hipSuccess;
  const int cur = B.lazy_cur;
  const LaunchCfg L;
  if (B.lazy_full) {
    // the one-kernel step kept the predictor's outputs in registers as well:
    // rerun the predictor sweep first (npc <= 2: no half-step planes)
    const hipError_t e = launch_sweep<T, EXACT, PRED, false>(
        stage_args<T, EXACT>(B, npc, 0, cur), P, make_geo(LX, LY, NP, L), L.wpb, s);
    if (e != hipSuccess) return e;
  }
  if (npc >= 2) {
    const hipError_t e = launch_corrector<T, EXACT>(stage_args<T, EXACT>(B, npc, npc - 1, cur),
                                                    P, make_geo(LX, LY, NP, L), L, s, nullptr,
                                                    npc - 1, false);
    if (e != hipSuccess) return e;
  }
  note_last_uq(B, stage_args<T, EXACT>(B, npc, npc, cur));
  B.lazy_cur = -1;
  B.lazy_full = false;
  return hipSuccess;
}

template <class T, bool EXACT>
hipError_t pack_state(Bufs<T>& B, const double* d_ff, const double* d_qxx, const double* d_qyx,
                      const double* d_phi, int N, int NP, double rho, hipStream_t s) {
  const int cur = B.cur;
  const int bs = 256, nb = (N + bs - 1) / bs;
  hipLaunchKernelGGL((pack_state_kernel<T, !EXACT>), dim3(nb), dim3(bs), 0, s, d_ff, d_qxx, d_qyx,
                     d_phi, B.Sf[cur], B.Sqxx[cur], B.Sqyx[cur], B.Sphi[cur], B.Sphic[cur], N, NP,
                     rho);
  B.have_uq = false;
  B.lazy_cur = -1;
  B.lazy_full = false;
  return hipGetLastError();
}

// Downloads field `field` (mass_lyo_field numbering) into the device FP64
// buffer d_out (N*9 for populations, N otherwise).
template <class T, bool EXACT>
hipError_t field_to_f64(const Bufs<T>& B, const Params<T>& P, int LX, int LY, int NP, int field,
                        double rho, double* d_out, hipStream_t s) {
  const int N = LX * LY, cur = B.cur;
  const int bs = 256, nb = (N + bs - 1) / bs;
  switch (field) {
    case 0:
      hipLaunchKernelGGL((unpack_f_kernel<T, !EXACT>), dim3(nb), dim3(bs), 0, s, B.Sf[cur], d_out,
                         N, NP, rho);
      break;
    case 1:
      hipLaunchKernelGGL((unpack_scalar_kernel<T>), dim3(nb), dim3(bs), 0, s, B.Sqxx[cur],
                         (const T*)nullptr, d_out, N);
      break;
    case 2:
      hipLaunchKernelGGL((unpack_scalar_kernel<T>), dim3(nb), dim3(bs), 0, s, B.Sqyx[cur],
                         (const T*)nullptr, d_out, N);
      break;
    case 3:
      hipLaunchKernelGGL((unpack_scalar_kernel<T>), dim3(nb), dim3(bs), 0, s, B.Sphi[cur],
                         (const T*)B.Sphic[cur], d_out, N);
      break;
    default:
      if (!B.have_uq) {
        return hipMemsetAsync(d_out, 0, sizeof(double) * (size_t)N, s);
      }
      hipLaunchKernelGGL((derived_kernel<T, EXACT>), dim3(nb), dim3(bs), 0, s, B.uq_f, B.uq_qxx,
                         B.uq_qyx, B.uq_phi, B.uq_dphi, P, LX, LY, NP, field, d_out);
      break;
  }
  return hipGetLastError();
}

// FP64 sums of all populations and of phi; out[0], out[1] (device).
template <class T, bool EXACT>
hipError_t state_sums(const Bufs<T>& B, int N, int NP, double rho, double* d_part, int max_part,
                      double* d_out, hipStream_t s) {
  const int cur = B.cur;
  int nparts = (N + 4095) / 4096;
  if (nparts > max_part) nparts = max_part;
  const int per = (N + nparts - 1) / nparts;
  hipLaunchKernelGGL((partial_sums_kernel<T>), dim3(nparts), dim3(256), 0, s, B.Sf[cur],
                     B.Sphi[cur], (const T*)B.Sphic[cur], N, NP, per, d_part, d_part + max_part);
  // shifted populations: sum f = sum f' + N*rho (the rest weights sum to 1)
  const double add = EXACT ? 0.0 : (double)N * rho;
  hipLaunchKernelGGL(final_sums_kernel, dim3(1), dim3(256), 0, s, d_part, d_part + max_part,
                     nparts, add, d_out);
  return hipGetLastError();
}

// Device scratch of the diagnostics (allocated on first use by the handle).
struct DiagScratch {
  double* part = nullptr;              // 5 * max_part
  int max_part = 0;
  double* out = nullptr;               // 5
  unsigned long long* hist = nullptr;  // kHistBins + 2
  unsigned long long* cnt = nullptr;   // 2
  signed char* sgn = nullptr;          // N
  int* label = nullptr;                // N
};

template <class T, bool EXACT>
hipError_t run_diagnostics(const Bufs<T>& B, int LX, int LY, int NP, double rho, DiagScratch& d,
                           double* w_out, hipStream_t s) {
  const int N = LX * LY, cur = B.cur;
  hipError_t e = hipMemsetAsync(d.hist, 0, sizeof(unsigned long long) * (kHistBins + 2), s);
  if (e == hipSuccess) e = hipMemsetAsync(d.cnt, 0, sizeof(unsigned long long) * 2, s);
  if (e != hipSuccess) return e;
  int nparts = (N + 4095) / 4096;
  if (nparts > d.max_part) nparts = d.max_part;
  const int per = (N + nparts - 1) / nparts;
  hipLaunchKernelGGL((diag_partial_kernel<T, !EXACT>), dim3(nparts), dim3(256), 0, s, B.Sf[cur],
                     B.Sphi[cur], (const T*)B.Sphic[cur], N, NP, per, rho, d.part, d.hist);
  hipLaunchKernelGGL(diag_final_kernel, dim3(1), dim3(64), 0, s, d.part, nparts, d.out);
  const int bs = 256, nb = (N + bs - 1) / bs;
  hipLaunchKernelGGL((charge_kernel<T>), dim3(nb), dim3(bs), 0, s, B.Sqxx[cur], B.Sqyx[cur], LX,
                     LY, w_out, d.sgn, d.label);
  hipLaunchKernelGGL(ccl_union_kernel, dim3(nb), dim3(bs), 0, s, d.sgn, d.label, LX, LY);
  hipLaunchKernelGGL(ccl_count_kernel, dim3(nb), dim3(bs), 0, s, d.sgn, d.label, N, d.cnt);
  return hipGetLastError();
}

// Precision-independent interface of one run's device state; the two
// implementations live in their own translation units so that the FP64 one
// is compiled without FMA contraction.
class Engine {
 public:
  virtual ~Engine() {}
  virtual hipError_t alloc(int LX, int LY, int NP, int npc) = 0;
  virtual void release() = 0;
  virtual void set_params(const double* p14) = 0;  // GammaQ..conc, 14 values
  virtual hipError_t steps(int nsteps, const LaunchCfg& L, hipStream_t s, SweepTimer* tm) = 0;
  virtual hipError_t pack(const double* d_ff, const double* d_qxx, const double* d_qyx,
                          const double* d_phi, hipStream_t s) = 0;
  virtual hipError_t field(int id, double* d_out, hipStream_t s) = 0;
  virtual hipError_t sums(double* d_part, int max_part, double* d_out, hipStream_t s) = 0;
  virtual hipError_t diagnostics(DiagScratch& d, double* w_out, hipStream_t s) = 0;
  virtual size_t bytes() const = 0;
  // batched stepping: the step plan for runs of at least LXmin rows, the
  // job of stage sc at state parity cur (host bytes of a SweepJob<T>), the
  // bookkeeping after n batched steps, and the launch of a batched stage over
  // device jobs
  virtual std::vector<int> plan(const LaunchCfg& L, int LXmin) const = 0;
  virtual size_t job_bytes() const = 0;
  virtual int job_items(int sc, const LaunchCfg& L) const = 0;
  virtual void fill_job(int sc, int cur, int first, const LaunchCfg& L, void* job) const = 0;
  virtual void after_batch_steps(int nsteps, const std::vector<int>& plan) = 0;
  virtual int parity() const = 0;
  virtual hipError_t launch_batch(const void* jobs, int njobs, int nitems, int sc,
                                  bool store_half, int grid_waves, int wpb,
                                  hipStream_t s) const = 0;
  // surrogate corrector hooks (lyo_net_engine.hpp).  src 0: current state;
  // src 1: the one-corrector state of the last plain step (needs npc >= 2).
  // Features FP32 into feat, or normalised bf16 into feat_bf when given
  virtual hipError_t net_features(int src, float* feat, const float* norm, __bf16* feat_bf,
                                  hipStream_t s) = 0;
  virtual hipError_t net_targets(float* tgt, hipStream_t s) = 0;
  virtual hipError_t net_apply(const float* delta, hipStream_t s) = 0;
  // part: nparts x kJudgeSums partial sums (judge_kernel)
  virtual hipError_t net_judge(const float* delta, const NetCalib& cal, double* part, int nparts,
                               hipStream_t s) = 0;
  // the network part of a surrogate step in one kernel (FP32, lyo_net_fused.hpp):
  // features, the network and the increment of the current state, xc rows
  // per wave chunk; delta (optional): the increments applied
  virtual hipError_t net_fused(const NetFusedW&, int, float*, hipStream_t) {
    return hipErrorNotSupported;
  }
  // nsteps steps with npc_run correctors (npc_run <= npc of the engine)
  virtual hipError_t steps_with(int npc_run, int nsteps, const LaunchCfg& L, hipStream_t s) = 0;
  virtual int npc() const = 0;
  // buffers of the reference loop structure (FP64 only): allocated on first
  // use, kept until release
  virtual hipError_t alloc_loops() = 0;
  // raw state S[cur] (f 9 planes, Qxx, Qyx, phi, FP32: phi compensation) as
  // contiguous planes of LX*LY values: its size, and copies to / from host
  // memory (synchronous; after a load the frame's derived fields restart at
  // zero)
  virtual int state_planes() const = 0;
  virtual size_t state_bytes() const = 0;
  virtual hipError_t save_state(void* host, hipStream_t s) = 0;
  virtual hipError_t load_state(const void* host, hipStream_t s) = 0;
  // go-or-grow (lyo_growth.hpp; the ghost-layer engine only): the model's
  // parameters and the run's generator (before configure), and Configure's
  // relaxation, reset and phenotype set-up after the initial state is packed
  // (ptot: in, Lyotropic's; out, the relaxed one)
  virtual hipError_t set_growth(const mass_lyo_growth&, std::mt19937*) {
    return hipErrorNotSupported;
  }
  virtual hipError_t growth_configure(double*, double*, hipStream_t) {
    return hipErrorNotSupported;
  }
};

template <class T, bool EXACT>
class EngineT final : public Engine {
 public:
  ~EngineT() override { release(); }

  hipError_t alloc(int LX, int LY, int NP, int npc) override {
    LX_ = LX;
    LY_ = LY;
    NP_ = NP;
    npc_ = npc;
    const size_t n1 = (size_t)NP, n9 = 9 * (size_t)NP;
    auto get = [&](size_t n, T** p) { return get_(n, p); };
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      if (e == hipSuccess) e = get(n9, &B_.Sf[i]);
      if (e == hipSuccess) e = get(n1, &B_.Sqxx[i]);
      if (e == hipSuccess) e = get(n1, &B_.Sqyx[i]);
      if (e == hipSuccess) e = get(n1, &B_.Sphi[i]);
      if (!EXACT && e == hipSuccess) e = get(n1, &B_.Sphic[i]);
    }
    if (e == hipSuccess) e = get(n9, &B_.Pfs);
    // fn* and Q+D/2: FP32 needs them only for correctors c >= 2
    const bool half = EXACT || npc >= 2;
    if (half && e == hipSuccess) e = get(n9, &B_.Pfns);
    if (e == hipSuccess) e = get(n1, &B_.Pqsxx);
    if (e == hipSuccess) e = get(n1, &B_.Pqsyx);
    if (half && e == hipSuccess) e = get(n1, &B_.Pqnxx);
    if (half && e == hipSuccess) e = get(n1, &B_.Pqnyx);
    if (e == hipSuccess) e = get(n1, &B_.Pphi);
    if (EXACT && e == hipSuccess) e = get(n1, &B_.Pphn);
    const int nc = npc >= 3 ? 2 : (npc == 2 ? 1 : 0);
    for (int i = 0; i < nc && e == hipSuccess; ++i) {
      if (e == hipSuccess) e = get(n9, &B_.Cf[i]);
      if (e == hipSuccess) e = get(n1, &B_.Cqxx[i]);
      if (e == hipSuccess) e = get(n1, &B_.Cqyx[i]);
      if (e == hipSuccess) e = get(n1, &B_.Cphi[i]);
    }
    return e;
  }

  void release() override {
    for (T* p : ptrs_) (void)hipFree(p);
    ptrs_.clear();
    bytes_ = 0;
    B_ = Bufs<T>();
    if (dphi_part_) (void)hipFree(dphi_part_);
    dphi_part_ = nullptr;
    dphi_cap_ = 0;
  }

  void set_params(const double* v) override {
    P_.GammaQ = (T)v[0];
    P_.GammaP = (T)v[1];
    P_.xi = (T)v[2];
    P_.tauNem = (T)v[3];
    P_.tauIso = (T)v[4];
    P_.rho = (T)v[5];
    P_.friction = (T)v[6];
    P_.AA = (T)v[7];
    P_.CC = (T)v[8];
    P_.LL = (T)v[9];
    P_.KK = (T)v[10];
    P_.zeta = (T)v[11];
    P_.zetaI = (T)v[12];
    P_.conc = (T)v[13];
    finish_params(P_);
    rho_ = v[5];
  }

  hipError_t steps(int nsteps, const LaunchCfg& L, hipStream_t s, SweepTimer* tm) override {
    return run_steps<T, EXACT>(B_, P_, LX_, LY_, NP_, npc_, nsteps, L, s, tm);
  }
  hipError_t pack(const double* d_ff, const double* d_qxx, const double* d_qyx,
                  const double* d_phi, hipStream_t s) override {
    return pack_state<T, EXACT>(B_, d_ff, d_qxx, d_qyx, d_phi, LX_ * LY_, NP_, rho_, s);
  }
  hipError_t field(int id, double* d_out, hipStream_t s) override {
    if (id > 3) {
      const hipError_t e = materialize_uq<T, EXACT>(B_, P_, LX_, LY_, NP_, npc_, s);
      if (e != hipSuccess) return e;
    }
    return field_to_f64<T, EXACT>(B_, P_, LX_, LY_, NP_, id, rho_, d_out, s);
  }
  hipError_t sums(double* d_part, int max_part, double* d_out, hipStream_t s) override {
    return state_sums<T, EXACT>(B_, LX_ * LY_, NP_, rho_, d_part, max_part, d_out, s);
  }
  hipError_t diagnostics(DiagScratch& d, double* w_out, hipStream_t s) override {
    return run_diagnostics<T, EXACT>(B_, LX_, LY_, NP_, rho_, d, w_out, s);
  }
  size_t bytes() const override { return bytes_; }

  std::vector<int> plan(const LaunchCfg& L, int LXmin) const override {
    return step_plan(EXACT, npc_, L, LXmin);
  }
  size_t job_bytes() const override { return sizeof(SweepJob<T>); }
  int job_items(int sc, const LaunchCfg& L) const override {
    return make_geo(LX_, LY_, NP_, L, stage_geo_kind(sc)).nwaves;
  }
  void fill_job(int sc, int cur, int first, const LaunchCfg& L, void* job) const override {
    SweepJob<T> j;
    j.A = plan_args<T, EXACT>(B_, npc_, sc, cur);
    j.P = P_;
    j.g = make_geo(LX_, LY_, NP_, L, stage_geo_kind(sc));
    j.g.xcd = 0;
    j.first = first;
    *static_cast<SweepJob<T>*>(job) = j;
  }
  void after_batch_steps(int nsteps, const std::vector<int>& plan) override {
    if (nsteps <= 0) return;
    const int last_cur = B_.cur ^ ((nsteps - 1) & 1);
    note_step<T, EXACT>(B_, npc_, plan, last_cur);
    B_.cur ^= (nsteps & 1);
  }
  int parity() const override { return B_.cur; }

  hipError_t net_features(int src, float* feat, const float* norm, __bf16* feat_bf,
                          hipStream_t s) override {
    if (src != 0) {
      const hipError_t e = materialize_uq<T, EXACT>(B_, P_, LX_, LY_, NP_, npc_, s);
      if (e != hipSuccess) return e;
    }
    const int N = LX_ * LY_, cur = B_.cur;
    const int bs = 128, nb = (N + bs - 1) / bs;
    const T *f, *qxx, *qyx, *phi, *dphi = nullptr, *phic = nullptr;
    if (src == 0) {
      f = B_.Sf[cur];
      qxx = B_.Sqxx[cur];
      qyx = B_.Sqyx[cur];
      phi = B_.Sphi[cur];
      phic = B_.Sphic[cur];
    } else {
      if (npc_ < 2) return hipErrorInvalidValue;
      f = B_.Cf[0];
      qxx = B_.Cqxx[0];
      qyx = B_.Cqyx[0];
      if (EXACT) {
        phi = B_.Cphi[0];
      } else {
        phi = B_.Sphi[cur ^ 1];
        dphi = B_.Cphi[0];
      }
    }
    if (feat_bf && !norm) return hipErrorInvalidValue;
    hipLaunchKernelGGL((features_kernel<T, EXACT>), dim3(nb), dim3(bs), 0, s, f, qxx, qyx, phi,
                       dphi, phic, P_, LX_, LY_, NP_, feat, norm, feat_bf);
    return hipGetLastError();
  }
  hipError_t net_targets(float* tgt, hipStream_t s) override {
    if (npc_ < 2) return hipErrorInvalidValue;
    const hipError_t e = materialize_uq<T, EXACT>(B_, P_, LX_, LY_, NP_, npc_, s);
    if (e != hipSuccess) return e;
    const int N = LX_ * LY_, cur = B_.cur;
    const int bs = 256, nb = (N + bs - 1) / bs;
    hipLaunchKernelGGL((targets_kernel<T, !EXACT>), dim3(nb), dim3(bs), 0, s, B_.Sf[cur],
                       B_.Sqxx[cur], B_.Sqyx[cur], B_.Sphi[cur], (const T*)B_.Sphic[cur], B_.Cf[0],
                       B_.Cqxx[0], B_.Cqyx[0], EXACT ? B_.Cphi[0] : B_.Sphi[cur ^ 1],
                       EXACT ? (const T*)nullptr : (const T*)B_.Cphi[0], N, NP_, rho_, tgt);
    return hipGetLastError();
  }
  hipError_t net_apply(const float* delta, hipStream_t s) override {
    const int N = LX_ * LY_, cur = B_.cur;
    const int bs = 256, nb = (N + bs - 1) / bs;
    hipError_t e = ensure_dphi_part(kPhiBlocks);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((apply_increment_kernel<T, !EXACT>), dim3(nb), dim3(bs), 0, s, B_.Sf[cur],
                       B_.Sqxx[cur], B_.Sqyx[cur], B_.Sphi[cur], B_.Sphic[cur], delta, N, NP_,
                       rho_);
    // sum(phi) as the plain corrector keeps it: the mean increment removed
    hipLaunchKernelGGL(dphi_partial_kernel<kFeat>, dim3(kPhiBlocks), dim3(256), 0, s, delta, N,
                       dphi_part_);
    hipLaunchKernelGGL((phi_shift_kernel<T>), dim3(kPhiBlocks), dim3(256), 0, s, B_.Sphi[cur],
                       B_.Sphic[cur], N, (const double*)dphi_part_, kPhiBlocks);
    return hipGetLastError();
  }
  hipError_t net_judge(const float* delta, const NetCalib& cal, double* part, int nparts,
                       hipStream_t s) override {
    if (npc_ < 2) return hipErrorInvalidValue;
    const hipError_t e = materialize_uq<T, EXACT>(B_, P_, LX_, LY_, NP_, npc_, s);
    if (e != hipSuccess) return e;
    const int cur = B_.cur;
    hipLaunchKernelGGL((judge_kernel<T, !EXACT>), dim3(nparts), dim3(256), 0, s, B_.Cf[0],
                       B_.Cqxx[0], B_.Cqyx[0], EXACT ? B_.Cphi[0] : B_.Sphi[cur ^ 1],
                       EXACT ? (const T*)nullptr : (const T*)B_.Cphi[0], B_.Sf[cur], B_.Sqxx[cur],
                       B_.Sqyx[cur], B_.Sphi[cur], (const T*)B_.Sphic[cur], delta, cal,
                       LX_ * LY_, NP_, rho_, part);
    return hipGetLastError();
  }
  hipError_t net_fused(const NetFusedW& w, int xc, float* delta, hipStream_t s) override {
    if constexpr (EXACT) {
      return hipErrorNotSupported;
    } else {
      if (!B_.Xf) {
        const size_t n1 = (size_t)NP_;
        hipError_t e = get_(9 * n1, &B_.Xf);
        if (e == hipSuccess) e = get_(n1, &B_.Xqxx);
        if (e == hipSuccess) e = get_(n1, &B_.Xqyx);
        if (e == hipSuccess) e = get_(n1, &B_.Xphi);
        if (e == hipSuccess) e = get_(n1, &B_.Xphic);
        if (e != hipSuccess) return e;
      }
      const int cur = B_.cur;
      const StatePlanes in{B_.Sf[cur], B_.Sqxx[cur], B_.Sqyx[cur], B_.Sphi[cur], B_.Sphic[cur]};
      const StatePlanes out{B_.Xf, B_.Xqxx, B_.Xqyx, B_.Xphi, B_.Xphic};
      const int nw = net_fused_waves(LX_, LY_, xc);
      hipError_t e = ensure_dphi_part(nw);
      if (e != hipSuccess) return e;
      e = launch_net_fused(in, out, LX_, LY_, NP_, xc, P_, rho_, w, delta, dphi_part_, s);
      if (e != hipSuccess) return e;
      // sum(phi) as the plain corrector keeps it: the mean increment removed
      hipLaunchKernelGGL((phi_shift_kernel<T>), dim3(kPhiBlocks), dim3(256), 0, s, B_.Xphi,
                         B_.Xphic, LX_ * LY_, (const double*)dphi_part_, nw);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
      // the corrected state becomes S[cur]; the old state S[cur ^ 1] and the
      // frame fields' bookkeeping (lazy_cur, uq_*) do not refer to it
      std::swap(B_.Sf[cur], B_.Xf);
      std::swap(B_.Sqxx[cur], B_.Xqxx);
      std::swap(B_.Sqyx[cur], B_.Xqyx);
      std::swap(B_.Sphi[cur], B_.Xphi);
      std::swap(B_.Sphic[cur], B_.Xphic);
      return hipSuccess;
    }
  }
  hipError_t steps_with(int npc_run, int nsteps, const LaunchCfg& L, hipStream_t s) override {
    if (npc_run < 1 || npc_run > npc_) return hipErrorInvalidValue;
    return run_steps<T, EXACT>(B_, P_, LX_, LY_, NP_, npc_run, nsteps, L, s, nullptr);
  }
  int npc() const override { return npc_; }
  int state_planes() const override { return EXACT ? 12 : 13; }
  size_t state_bytes() const override {
    return (size_t)state_planes() * (size_t)LX_ * (size_t)LY_ * sizeof(T);
  }
  hipError_t save_state(void* host, hipStream_t s) override {
    T* planes[13];
    const int n = state_plane_ptrs(planes);
    const size_t pb = (size_t)LX_ * LY_ * sizeof(T);
    char* h = static_cast<char*>(host);
    hipError_t e = hipSuccess;
    for (int i = 0; i < n && e == hipSuccess; ++i, h += pb)
      e = hipMemcpyAsync(h, planes[i], pb, hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  }
  hipError_t load_state(const void* host, hipStream_t s) override {
    T* planes[13];
    const int n = state_plane_ptrs(planes);
    const size_t pb = (size_t)LX_ * LY_ * sizeof(T);
    const char* h = static_cast<const char*>(host);
    hipError_t e = hipSuccess;
    for (int i = 0; i < n && e == hipSuccess; ++i, h += pb)
      e = hipMemcpyAsync(planes[i], h, pb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    B_.have_uq = false;
    B_.lazy_cur = -1;
    B_.lazy_full = false;
    return e;
  }
  hipError_t alloc_loops() override {
    if (!EXACT) return hipErrorInvalidValue;
    if (B_.Ld) return hipSuccess;
    const size_t n1 = (size_t)NP_;
    hipError_t e = get_(kLoopPlanes * n1, &B_.Ld);
    if (e == hipSuccess) e = get_(9 * n1, &B_.Lfc);
    if (e == hipSuccess) e = get_(9 * n1, &B_.Lfnc);
    return e;
  }
  hipError_t launch_batch(const void* jobs, int njobs, int nitems, int c, bool store_half,
                          int grid_waves, int wpb, hipStream_t s) const override {
    const SweepJob<T>* J = static_cast<const SweepJob<T>*>(jobs);
    const int waves = grid_waves > 0 && grid_waves < nitems ? grid_waves : nitems;
    const int blocks = (waves + wpb - 1) / wpb;
    const dim3 grid(blocks), block(64 * wpb);
    constexpr bool HB = !EXACT;  // EXACT: the flag is ignored, one instance
    if (c == kStageFull) {
      const size_t lds = full_lds_bytes<T, 2>(wpb);
      hipLaunchKernelGGL((full_batch_kernel<T, EXACT, 2>), grid, block, lds, s, J, njobs, nitems);
      return hipGetLastError();
    }
    if (c == kStageFull1) {
      hipLaunchKernelGGL((full_batch_kernel<T, EXACT, 1>), grid, block, 0, s, J, njobs, nitems);
      return hipGetLastError();
    }
    if (c == kStagePair) {
      if (HB && npc_ == 2)
        hipLaunchKernelGGL((pair_batch_kernel<T, EXACT, HB>), grid, block, 0, s, J, njobs, nitems);
      else
        hipLaunchKernelGGL((pair_batch_kernel<T, EXACT, false>), grid, block, 0, s, J, njobs,
                           nitems);
      return hipGetLastError();
    }
    if (c == 0) {
      if (HB && store_half)
        hipLaunchKernelGGL((sweep_batch_kernel<T, EXACT, PRED, HB>), grid, block, 0, s, J, njobs,
                           nitems);
      else
        hipLaunchKernelGGL((sweep_batch_kernel<T, EXACT, PRED, false>), grid, block, 0, s, J,
                           njobs, nitems);
    } else if (c == 1) {
      if (c < npc_)
        hipLaunchKernelGGL((sweep_batch_kernel<T, EXACT, CORR, HB>), grid, block, 0, s, J, njobs,
                           nitems);
      else
        hipLaunchKernelGGL((sweep_batch_kernel<T, EXACT, CORR_LAST, HB>), grid, block, 0, s, J,
                           njobs, nitems);
    } else if (c < npc_) {
      hipLaunchKernelGGL((sweep_batch_kernel<T, EXACT, CORR, false>), grid, block, 0, s, J, njobs,
                         nitems);
    } else {
      hipLaunchKernelGGL((sweep_batch_kernel<T, EXACT, CORR_LAST, false>), grid, block, 0, s, J,
                         njobs, nitems);
    }
    return hipGetLastError();
  }

 private:
  // the state planes of S[cur] in checkpoint order
  int state_plane_ptrs(T* (&planes)[13]) const {
    const int cur = B_.cur;
    int n = 0;
    for (int v = 0; v < 9; ++v) planes[n++] = B_.Sf[cur] + (size_t)v * NP_;
    planes[n++] = B_.Sqxx[cur];
    planes[n++] = B_.Sqyx[cur];
    planes[n++] = B_.Sphi[cur];
    if (!EXACT) planes[n++] = B_.Sphic[cur];
    return n;
  }
  hipError_t get_(size_t n, T** p) {
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e == hipSuccess) {
      ptrs_.push_back(*p);
      bytes_ += n * sizeof(T);
      e = hipMemset(*p, 0, n * sizeof(T));
    }
    return e;
  }

  // partial sums of the surrogate's phi increments (one per block or wave)
  hipError_t ensure_dphi_part(int n) {
    if (n <= dphi_cap_) return hipSuccess;
    if (dphi_part_) (void)hipFree(dphi_part_);
    dphi_part_ = nullptr;
    dphi_cap_ = 0;
    const hipError_t e = hipMalloc((void**)&dphi_part_, sizeof(double) * (size_t)n);
    if (e == hipSuccess) dphi_cap_ = n;
    return e;
  }

  Bufs<T> B_;
  double* dphi_part_ = nullptr;
  int dphi_cap_ = 0;
  Params<T> P_{};
  double rho_ = 0;
  int LX_ = 0, LY_ = 0, NP_ = 0, npc_ = 1;
  std::vector<T*> ptrs_;
  size_t bytes_ = 0;
};

Engine* make_engine_fp64();
Engine* make_engine_fp32();
Engine* make_engine_walls(int bc);  // FP64, bc 1-4 (lyo_walls.hpp)

}  // namespace lyo
