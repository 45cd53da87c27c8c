// lyo_full.hpp -- one whole FP32 step (predictor + two correctors) in one
// kernel (gfx950).
//
// Lyotropic::Step (src/models/lyotropic.cpp:500-525) with npc = 2 is a
// predictor sweep (UpdateQuantities, UpdateFields(true), Move) followed by two
// corrector sweeps.  Here the three run as three stages of one strip march:
//   P  predictor, UpdateFields on row x (its Move is a pull, below)
//   A  corrector 1, UpdateQuantities on row x-D, UpdateFields on row x-2D
//   B  corrector 2, UpdateQuantities on row x-3D, UpdateFields on row x-4D
// (D = +-1, the march direction).  P's post-collision populations of rows
// x-2D..x are pulled into the streamed populations of row x-D (the node at
// (r, y) takes direction v from (r - cx_v, y - cy_v): rows from registers,
// columns through DPP lane shifts), so neither the predictor's outputs nor the
// first corrector's ever reach HBM: per node the kernel reads the old state
// (f', Q, phi, phic) and writes the new one.  Every stage calls the same device
// functions in the same order as the separate sweeps (lyo_sweep.hpp) with the
// half-step values rebuilt as there (half_of), so the results are bitwise
// those of the predictor sweep + fused pair.  Each stage needs its inputs one
// lane further in, so the halo is six lanes per side (52 interior columns).
#pragma once

#include "lyo_pair.hpp"

namespace lyo {

constexpr int kFullHalo = 6;
constexpr int kFullStrip = kWave - 2 * kFullHalo;

// Per-row outputs of P that A needs later: half-step Q, the predictor's
// increment of phi, and old phi.
template <class T>
struct POut {
  T qnx, qny, dp, raw;
};

// Loads of P's next iteration (row x+D): Q, phi of row x+3D and f of row
// x+2D, issued once P has used the current ones, so they are in flight while
// A and B compute.
template <class T>
struct FullIn {
  QP<T> q;
  T f[9];
};
// Old phi and its compensation of B's output row, loaded one iteration ahead.
template <class T>
struct OldPhi {
  T pold, phic;
};

#ifndef LYO_FULL_WAVES_PER_EU
#define LYO_FULL_WAVES_PER_EU 2
#endif

// LDS hand-off from A to B: per wave, two rings (A's outputs f 9, Q 2, phi;
// A's half-step values fn 9, QN 2, Dp) of three rows of twelve values per
// lane, laid out [ring][slot][group of four][lane] so every access is a
// conflict-free 64-lane ds_*_b128.  B reads the rows A wrote one and two
// iterations before; the third slot is the one A writes now.
constexpr int kRingFloat4 = 2 * 3 * 3 * 64;  // per wave (18 KiB), dynamic LDS

// Arguments: f_in, qxx_in, qyx_in, phi_in, phic_in = the state S[cur];
// f_out, qxx_out, qyx_out, phi_out, phic_out = S[next].
template <class T>
__device__ __forceinline__ void full_item(const SweepArgs<T>& A, const Params<T>& P,
                                          const Geo& g, const int wid, float4* ring) {
  static_assert(sizeof(T) == 4, "the one-kernel step is the FP32 path");
  const int lane = threadIdx.x & 63;
  auto ring_put = [&](int r, int slot, const float (&v)[12]) {
    float4* p = ring + (r * 3 + slot) * 192 + lane;
    p[0] = make_float4(v[0], v[1], v[2], v[3]);
    p[64] = make_float4(v[4], v[5], v[6], v[7]);
    p[128] = make_float4(v[8], v[9], v[10], v[11]);
  };
  auto ring_get = [&](int r, int slot, float (&v)[12]) {
    const float4* p = ring + (r * 3 + slot) * 192 + lane;
    const float4 a = p[0], b = p[64], c = p[128];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    v[8] = c.x; v[9] = c.y; v[10] = c.z; v[11] = c.w;
  };
  const int strip = wid % g.nstrips;
  const int chunk = wid / g.nstrips;
  const int LX = g.LX, LY = g.LY, N = g.NP;

  const int ycol = strip * kFullStrip - kFullHalo + lane;
  int yi = ycol % LY;
  if (yi < 0) yi += LY;
  LaneOff<T> ly{(unsigned)yi * (unsigned)sizeof(T)};
  const bool active = lane >= kFullHalo && lane < kWave - kFullHalo && ycol < LY;
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
  auto load_row = [&](int x, int d, FullIn<T>& R) {
    R.q = load_qp(x + 2 * d);
    load_f(x + d, R.f);
  };
  auto load_old = [&](int r, OldPhi<T>& R) {
    const int o = roff(r);
    R.pold = ly.ld(A.phi_in + o);
    R.phic = ly.ld(A.phic_in + o);
  };
  auto uq = [&](const QP<T>& qm, const QP<T>& q0, const QP<T>& qp, const T(&f)[9],
                NbrVals<T>& nv, CenterVals<T>& cv) {
    update_quantities<T, false>(f, q0.xx, qp.xx, qm.xx, from_yp(q0.xx), from_ym(q0.xx),
                                q0.yx, qp.yx, qm.yx, from_yp(q0.yx), from_ym(q0.yx),
                                q0.p, qp.p, qm.p, from_yp(q0.p), from_ym(q0.p), P, nv,
                                cv);
  };

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
    auto ufs = [&](const NbrVals<T>& prev, const NbrVals<T>& u0, const NbrVals<T>& next) {
      return (D > 0) ? uf_stencil(prev, u0, next) : uf_stencil(next, u0, prev);
    };
    // x components of the directions (src/geometry.hpp:17-27)
    constexpr int cxv[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
    constexpr int cyv[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
    // P visits rows xs, xs+D, ...: four rows before the chunk, the chunk and
    // four rows after it; A trails by two rows and B by four
    const int xs = ((D > 0) ? x0 : x1 - 1) - 4 * D;
    const int nI = x1 - x0 + 8;

    // ---- stage P rolling state
    NbrVals<T> pum, pu0;
    CenterVals<T> pc0;
    T pfc[9];
    QP<T> pqb, pqc;
    FullIn<T> cur;
    OldPhi<T> bold;
    {
      const QP<T> qa = load_qp(xs - 2 * D);
      pqb = load_qp(xs - D);
      pqc = load_qp(xs);
      const QP<T> qd = load_qp(xs + D);
      T ft[9];
      load_f(xs - D, ft);
      load_f(xs, pfc);
      load_row(xs, D, cur);
      load_old(xs - 4 * D, bold);
      CenterVals<T> ct;
      uqd(qa, pqb, pqc, ft, pum, ct);
      uqd(pqb, pqc, qd, pfc, pu0, pc0);
      pqb = pqc;
      pqc = qd;
    }
    // post-collision populations f + om and their half-step values of the
    // rows x-D (pcP, hnP) and x-2D (pcPP, hnPP); P outputs of rows x-D, x-2D
    T pcP[9] = {}, hnP[9] = {}, pcPP[9] = {}, hnPP[9] = {};
    POut<T> po1{}, po2{};
    // ---- stage A: stencil rows x-2D (aqb), x-D (aqc) of (Q*, phi + Dp);
    // UpdateQuantities of rows x-3D (aum), x-2D (au0); streamed populations
    // and half-step values of row x-2D (afc, ahn)
    QP<T> aqb{}, aqc{};
    NbrVals<T> aum{}, au0{};
    CenterVals<T> ac0{};
    T afc[9] = {}, ahn[9] = {};
    // ---- stage B, as in pair_item, two rows further back; A's outputs and
    // half-step values of the rows x-3D, x-4D come from the LDS rings
    NbrVals<T> bum{}, bu0{};
    CenterVals<T> bc0{};

    for (int i = 0, x = xs, s0 = 0; i < nI; ++i, x += D, s0 = (s0 == 2) ? 0 : s0 + 1) {
      const int s1 = (s0 == 0) ? 2 : s0 - 1;  // slot of iteration i-1
      const int s2 = (s1 == 0) ? 2 : s1 - 1;  // slot of iteration i-2
      // ---- P: UpdateQuantities on row x+D, UpdateFields on row x
      NbrVals<T> pun;
      CenterVals<T> pcn;
      uqd(pqb, pqc, cur.q, cur.f, pun, pcn);
      T pc[9], hn[9];
      QP<T> aq_new;  // A's stencil values of row x
      POut<T> po;
      {
        const UFStencil<T> s = ufs(pum, pu0, pun);
        T Dxx, Dyx, Dp, om[9];
        update_fields<T, false>(pc0, pqb.xx, pqb.yx, pqb.p, pfc, s, P, Dxx, Dyx, Dp, om);
        const T Qsx = pqb.xx + Dxx, Qsy = pqb.yx + Dyx;
#pragma unroll
        for (int v = 0; v < 9; ++v) {
          pc[v] = pfc[v] + om[v];
          hn[v] = half_of(pc[v], pfc[v]);
        }
        aq_new.xx = Qsx;
        aq_new.yx = Qsy;
        // the first corrector's stencil phi: old phi plus the stored increment
        aq_new.p = pqb.p + Dp;
        po.qnx = half_of(Qsx, pqb.xx);
        po.qny = half_of(Qsy, pqb.yx);
        po.dp = Dp;
        po.raw = pqb.p;
      }
      pum = pu0;
      pu0 = pun;
      pc0 = pcn;
#pragma unroll
      for (int v = 0; v < 9; ++v) pfc[v] = cur.f[v];
      pqb = pqc;
      pqc = cur.q;
      // P's loads of the next iteration (unconditional: past the chunk the
      // wrapped rows are in bounds and unused)
      load_row(x + D, D, cur);

      // ---- pull (Move): streamed populations of row x-D
      T fst[9], hst[9];
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        T a, h;
        if (cxv[v] == -D) {
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

      // ---- A: UpdateQuantities on row x-D, UpdateFields on row x-2D
      QP<T> qa_out{};
      if (i >= 2) {  // wave-uniform
        NbrVals<T> aun;
        CenterVals<T> acn;
        uqd(aqb, aqc, aq_new, fst, aun, acn);
        if (i >= 4) {
          // half-step values of row x-2D, kept for B two rows later
          float cen[12];
#pragma unroll
          for (int v = 0; v < 9; ++v) cen[v] = ahn[v];
          cen[9] = po2.qnx;
          cen[10] = po2.qny;
          cen[11] = po2.dp;
          ring_put(1, s0, cen);
          const UFStencil<T> s = ufs(aum, au0, aun);
          T Dxx, Dyx, Dp, om[9];
          update_fields<T, false>(ac0, aqb.xx, aqb.yx, aqb.p, afc, s, P, Dxx, Dyx, Dp, om);
          float ao[12];
          qa_out.xx = cen[9] + T(.5) * Dxx;
          qa_out.yx = cen[10] + T(.5) * Dyx;
#pragma unroll
          for (int v = 0; v < 9; ++v) ao[v] = cen[v] + T(.5) * om[v];
          // the separate sweep stores the increment, the next one adds it to
          // old phi: the same two roundings here
          const T inc = T(.5) * (cen[11] + Dp);
          qa_out.p = po2.raw + inc;
          ao[9] = qa_out.xx;
          ao[10] = qa_out.yx;
          ao[11] = qa_out.p;
          ring_put(0, s0, ao);
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

      // ---- B: UpdateQuantities on row x-3D, UpdateFields on row x-4D
      if (i >= 6) {  // wave-uniform
        float ao1[12], ao2[12];  // A's outputs of rows x-3D, x-4D
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
        uqd(bqb, bqc, qa_out, bfp, bun, bcn);
        if (i >= 8) {
          const UFStencil<T> s = ufs(bum, bu0, bun);
          T Dxx, Dyx, Dp, om[9];
          update_fields<T, false>(bc0, bqb.xx, bqb.yx, bqb.p, bfc, s, P, Dxx, Dyx, Dp, om);
          float cmv[12];  // half-step values of row x-4D
          ring_get(1, s2, cmv);
          CenIn<T> cm2;
#pragma unroll
          for (int v = 0; v < 9; ++v) cm2.fn[v] = cmv[v];
          cm2.qnxx = cmv[9];
          cm2.qnyx = cmv[10];
          cm2.phn = cmv[11];
          if (active) {
            const int ox = roff(x - 4 * D);
            ly.st(A.qxx_out + ox, cm2.qnxx + T(.5) * Dxx);
            ly.st(A.qyx_out + ox, cm2.qnyx + T(.5) * Dyx);
#pragma unroll
            for (int v = 0; v < 9; ++v) ly.st(A.f_out + (v * N + ox), cm2.fn[v] + T(.5) * om[v]);
            // phi_new = phi_old + phic + (Dp_pred + Dp)/2 with a TwoSum, as
            // the separate last corrector
            const T pold = bold.pold;
            const T yv = T(.5) * (cm2.phn + Dp) + bold.phic;
            const T t = pold + yv;
            const T bp = t - pold;
            const T ap = t - bp;
            const T err = (pold - ap) + (yv - bp);
            ly.st(A.phi_out + ox, t);
            ly.st(A.phic_out + ox, err);
          }
        }
        bum = bu0;
        bu0 = bun;
        bc0 = bcn;
      }
      load_old(x - 3 * D, bold);
    }
  };
  // every chunk marches upwards: the step is bound by VALU issue, and one
  // direction keeps half the code in the instruction cache (alternating
  // directions measured no faster here)
  march(std::integral_constant<int, 1>{});
}

template <class T>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(LYO_FULL_WAVES_PER_EU)))
full_kernel(const SweepArgs<T> A, const Params<T> P, const Geo g) {
  int b = blockIdx.x;
  if (g.xcd) {
    const int nb8 = (int)gridDim.x >> 3;
    if (b < (nb8 << 3)) b = (b & 7) * nb8 + (b >> 3);
  }
  const int wid = __builtin_amdgcn_readfirstlane(b * ((int)blockDim.x >> 6) +
                                                 ((int)threadIdx.x >> 6));
  if (wid >= g.nwaves) return;
  extern __shared__ float4 ring[];  // waves per block * kRingFloat4
  full_item<T>(A, P, g, wid, ring + ((int)threadIdx.x >> 6) * kRingFloat4);
}

// Batched one-kernel steps: persistent grid over the work items of many runs
// (waves per block * kRingFloat4 float4 of dynamic LDS).
template <class T>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(LYO_FULL_WAVES_PER_EU)))
full_batch_kernel(const SweepJob<T>* __restrict__ jobs, const int njobs, const int nitems) {
  extern __shared__ float4 ring[];
  float4* r = ring + ((int)threadIdx.x >> 6) * kRingFloat4;
  const int nw = (int)(gridDim.x * (blockDim.x >> 6));
  int item = __builtin_amdgcn_readfirstlane((int)blockIdx.x * ((int)blockDim.x >> 6) +
                                            ((int)threadIdx.x >> 6));
  for (; item < nitems; item += nw) {
    const SweepJob<T>& J = find_job(jobs, njobs, item);
    full_item<T>(J.A, J.P, J.g, item - J.first, r);
  }
}

template <class T>
inline hipError_t launch_full(const SweepArgs<T>& a, const Params<T>& p, const Geo& g,
                              int waves_per_block, hipStream_t s) {
  const int threads = 64 * waves_per_block;
  const int blocks = (g.nwaves + waves_per_block - 1) / waves_per_block;
  const size_t lds = (size_t)waves_per_block * kRingFloat4 * sizeof(float4);
  hipLaunchKernelGGL((full_kernel<T>), dim3(blocks), dim3(threads), lds, s, a, p, g);
  return hipGetLastError();
}

}  // namespace lyo
