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
// (lyo_sweep.hpp), with the half-step values formed as there (FP32: half_of;
// FP64: the reference's f + om/2, Q + D/2, phi + Dp/2), so the results are
// bitwise those of the predictor sweep + fused pair.  Each stage needs its
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

// Per-row outputs of P that A needs two rows later: half-step Q and the phi
// terms (FP32: Dp and old phi; FP64: phi + Dp/2).
template <class T>
struct POut {
  T qnx, qny, dp, raw;
};

// Old phi and its compensation of B's output row (FP32), one iteration ahead.
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
template <class T, bool EXACT, int NPC>
__device__ __forceinline__ void full_item(const SweepArgs<T>& A, const Params<T>& P,
                                          const Geo& g, const int wid,
                                          typename Ring<T>::V* ring) {
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
  constexpr int HALO = NPC == 2 ? kFullHalo : kFull1Halo;
  const int strip = wid % g.nstrips;
  const int chunk = wid / g.nstrips;
  const int LX = g.LX, LY = g.LY, N = g.NP;

  const int ycol = strip * (kWave - 2 * HALO) - HALO + lane;
  int yi = ycol % LY;
  if (yi < 0) yi += LY;
  LaneOff<T> ly{(unsigned)yi * (unsigned)sizeof(T)};
  const bool active = lane >= HALO && lane < kWave - HALO && ycol < LY;
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
    if constexpr (!EXACT) {
      const int o = roff(r);
      R.pold = ly.ld(A.phi_in + o);
      R.phic = ly.ld(A.phic_in + o);
    }
  };
  // The last corrector's outputs of row r (Lyotropic::UpdateFieldsAtNode with
  // first = false, lyotropic.cpp:331-340) from its half-step values hv (fn 9,
  // QN 2, phi term) and increments; FP32 phi: phi_new = phi_old + phic +
  // (Dp_pred + Dp)/2 with a TwoSum, as the separate last corrector.
  auto store_state = [&](int r, const T(&hv)[12], T Dxx, T Dyx, T Dp, const T(&om)[9],
                         const OldPhi<T>& old) {
    const int ox = roff(r);
    ly.st(A.qxx_out + ox, hv[9] + T(.5) * Dxx);
    ly.st(A.qyx_out + ox, hv[10] + T(.5) * Dyx);
#pragma unroll
    for (int v = 0; v < 9; ++v) ly.st(A.f_out + (v * N + ox), hv[v] + T(.5) * om[v]);
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
  auto uq = [&](const QP<T>& qm, const QP<T>& q0, const QP<T>& qp, const T(&f)[9],
                NbrVals<T>& nv, CenterVals<T>& cv) {
    update_quantities<T, EXACT>(f, q0.xx, qp.xx, qm.xx, from_yp(q0.xx), from_ym(q0.xx),
                                q0.yx, qp.yx, qm.yx, from_yp(q0.yx), from_ym(q0.yx),
                                q0.p, qp.p, qm.p, from_yp(q0.p), from_ym(q0.p), P, nv,
                                cv);
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
    CenterVals<T> ct;
    uq(qa, pqb, pqc, ft, pum, ct);
    uq(pqb, pqc, qd, pfc, pu0, pc0);
    pqb = pqc;
    pqc = qd;
  }
  // post-collision populations f + om and their half-step values of the rows
  // x-1 (pcP, hnP) and x-2 (pcPP, hnPP); P outputs of rows x-1, x-2
  T pcP[9] = {}, hnP[9] = {}, pcPP[9] = {}, hnPP[9] = {};
  POut<T> po1{}, po2{};
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
    // ---- P: UpdateQuantities on row x+1, UpdateFields on row x
    NbrVals<T> pun;
    CenterVals<T> pcn;
    uq(pqb, pqc, cq, cf, pun, pcn);
    T pc[9], hn[9];
    QP<T> aq_new;  // A's stencil values of row x
    POut<T> po;
    {
      const UFStencil<T> s = uf_stencil(pum, pu0, pun);
      T Dxx, Dyx, Dp, om[9];
      update_fields<T, EXACT>(pc0, pqb.xx, pqb.yx, pqb.p, pfc, s, P, Dxx, Dyx, Dp, om);
      const T Qsx = pqb.xx + Dxx, Qsy = pqb.yx + Dyx;
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        pc[v] = pfc[v] + om[v];
        if constexpr (EXACT) {
          hn[v] = pfc[v] + T(.5) * om[v];
        } else {
          hn[v] = half_of(pc[v], pfc[v]);
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
        po.qnx = half_of(Qsx, pqb.xx);
        po.qny = half_of(Qsy, pqb.yx);
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
      if (cyv[v] > 0) {
        a = from_ym(a);
        h = from_ym(h);
      } else if (cyv[v] < 0) {
        a = from_yp(a);
        h = from_yp(h);
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
      uq(aqb, aqc, aq_new, fst, aun, acn);
      if (!WARM || i >= 4) {
        // half-step values of row x-2 (NPC = 2: kept for B two rows later)
        T cen[12];
#pragma unroll
        for (int v = 0; v < 9; ++v) cen[v] = ahn[v];
        cen[9] = po2.qnx;
        cen[10] = po2.qny;
        cen[11] = po2.dp;
        if constexpr (NPC == 2) ring_put(1, s0, cen);
        const UFStencil<T> s = uf_stencil(aum, au0, aun);
        T Dxx, Dyx, Dp, om[9];
        update_fields<T, EXACT>(ac0, aqb.xx, aqb.yx, aqb.p, afc, s, P, Dxx, Dyx, Dp, om);
        if constexpr (NPC == 2) {
          T ao[12];
          qa_out.xx = cen[9] + T(.5) * Dxx;
          qa_out.yx = cen[10] + T(.5) * Dyx;
#pragma unroll
          for (int v = 0; v < 9; ++v) ao[v] = cen[v] + T(.5) * om[v];
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
        } else if (active) {
          // the last corrector: the new state of row x-2
          store_state(x - 2, cen, Dxx, Dyx, Dp, om, bold);
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
        uq(bqb, bqc, qa_out, bfp, bun, bcn);
        if (!WARM || i >= 8) {
          const UFStencil<T> s = uf_stencil(bum, bu0, bun);
          T Dxx, Dyx, Dp, om[9];
          update_fields<T, EXACT>(bc0, bqb.xx, bqb.yx, bqb.p, bfc, s, P, Dxx, Dyx, Dp, om);
          T cmv[12];  // half-step values of row x-4
          ring_get(1, s2, cmv);
          if (active) store_state(x - 4, cmv, Dxx, Dyx, Dp, om, bold);
        }
        bum = bu0;
        bu0 = bun;
        bc0 = bcn;
      }
    }
    // old phi of the next iteration's output row
    load_old(x + 1 - 2 * NPC, bold);
  };
  constexpr int kWarm = 4 * NPC;  // nI >= kWarm + 1: a chunk has a row
  int i = 0;
  for (; i < kWarm; ++i) row(i, xs + i, i % 3, std::true_type{});
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
