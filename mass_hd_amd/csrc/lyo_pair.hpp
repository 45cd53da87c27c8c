// lyo_pair.hpp -- the last two corrector sweeps of a step fused into one
// kernel (gfx950).
//
// Lyotropic::Step (src/models/lyotropic.cpp:500-525) runs npc corrector
// sweeps after the predictor; each corrector reads the state the previous one
// wrote.  For the last pair (A = corrector npc-1, B = corrector npc) this
// kernel keeps A's outputs (f, Q, phi of one row) in registers and hands them
// to B in the same wave, so they never travel through HBM: per node the pair
// reads the predictor outputs and the old state once and writes the new state
// once.
//
// Both stages march along x in one wave.  Stage A is the strip march of
// lyo_sweep.hpp and produces row x of A's outputs; stage B then runs
// UpdateQuantities on row x-1 and UpdateFields on row x-2 from A's rows
// x-2..x.  B needs A's outputs two columns to each side, so the halo is four
// lanes per side and each wave writes 56 interior columns.  Every stage runs
// the same device functions, in the same order, as the separate sweeps, so the
// results are bitwise those of the two separate sweeps.
#pragma once

#include "lyo_sweep.hpp"

namespace lyo {

constexpr int kPairHalo = 4;
constexpr int kPairStrip = kWave - 2 * kPairHalo;

// Centre inputs of one corrector at one row: fn, QN and phn (predictor
// half-step values, or, FP32, the predictor's increment of phi).  As loaded by
// a rebuilding FP32 stage (HB) fn and QN hold the old populations at the
// upstream nodes and old Q.
template <class T>
struct CenIn {
  T fn[9];
  T qnxx, qnyx, phn;
};

// Loads of one pair iteration x: Q and phi of row x+2 and f of row x+1 for
// A's UpdateQuantities, the centre inputs of row x (A uses them at x, B two
// iterations later), and, FP32, old phi and its compensation of row x-2 for
// B's final update.
template <class T, bool EXACT>
struct PairIn {
  QP<T> q;
  T dp;
  T f[9];
  CenIn<T> c;
  T pold, phic;
};

#ifndef LYO_PAIR_WAVES_PER_EU
#define LYO_PAIR_WAVES_PER_EU 2
#endif

// A's inputs are those of the separate sweep npc-1 (f_in, q*_in, phi_in,
// dphi_in, fn_in, qn*_in or, HB, fold_in, qo*_in, phn_in); B's outputs those
// of sweep npc (f_out, q*_out, phi_out, FP32 also phic_in -> phic_out).  With
// HB (FP32, npc = 2) A rebuilds the half-step values and hands them to B.
template <class T, bool EXACT, bool HB>
__device__ __forceinline__ void pair_item(const SweepArgs<T>& A, const Params<T>& P,
                                          const Geo& g, const int wid) {
  constexpr bool kRebuild = !EXACT && HB;
  const int lane = threadIdx.x & 63;
  const int strip = wid % g.nstrips;
  const int chunk = wid / g.nstrips;
  const int LX = g.LX, LY = g.LY, N = g.NP;

  const int ycol = strip * kPairStrip - kPairHalo + lane;
  int yi = ycol % LY;
  if (yi < 0) yi += LY;
  LaneOff<T> ly{(unsigned)yi * (unsigned)sizeof(T)};
  LaneOff<T> lyp{(yi + 1 == LY ? 0u : (unsigned)(yi + 1)) * (unsigned)sizeof(T)};
  LaneOff<T> lym{(yi == 0 ? (unsigned)(LY - 1) : (unsigned)(yi - 1)) * (unsigned)sizeof(T)};
  const bool active = lane >= kPairHalo && lane < kWave - kPairHalo && ycol < LY;
  const int x0 = chunk * g.xc;
  const int x1 = min(x0 + g.xc, LX);

  // rows run from x0-4 to x1+4; the fused pair is used for LX >= 8, so one
  // wrap in either direction suffices
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
    T ph = ly.ld(A.phi_in + o);
    if constexpr (!EXACT) ph += ly.ld(A.dphi_in + o);
    q.p = ph;
    return q;
  };
  auto load_f = [&](int r, T(&f)[9]) {
    const int o = roff(r);
#pragma unroll
    for (int v = 0; v < 9; ++v) f[v] = ly.ld(A.f_in + (v * N + o));
  };
  // loads of the iteration on row x of a march in direction d
  auto load_row = [&](int x, int d, PairIn<T, EXACT>& R) {
    {
      const int o = roff(x + 2 * d);
      R.q.xx = ly.ld(A.qxx_in + o);
      R.q.yx = ly.ld(A.qyx_in + o);
      R.q.p = ly.ld(A.phi_in + o);
      if constexpr (!EXACT) R.dp = ly.ld(A.dphi_in + o);
    }
    load_f(x + d, R.f);
    if constexpr (kRebuild) {
      const int o = roff(x), om = roff(x - 1), op = roff(x + 1);
      const int rowv[9] = {o, om, op, o, o, om, op, op, om};
      LaneOff<T>* const colv[9] = {&ly, &ly, &ly, &lym, &lyp, &lym, &lyp, &lym, &lyp};
#pragma unroll
      for (int v = 0; v < 9; ++v) R.c.fn[v] = colv[v]->ld(A.fold_in + (v * N + rowv[v]));
      R.c.qnxx = ly.ld(A.qoxx_in + o);
      R.c.qnyx = ly.ld(A.qoyx_in + o);
      R.c.phn = ly.ld(A.phn_in + o);
    } else {
      const int o = roff(x);
#pragma unroll
      for (int v = 0; v < 9; ++v) R.c.fn[v] = ly.ld(A.fn_in + (v * N + o));
      R.c.qnxx = ly.ld(A.qnxx_in + o);
      R.c.qnyx = ly.ld(A.qnyx_in + o);
      R.c.phn = ly.ld(A.phn_in + o);
    }
    if constexpr (!EXACT) {
      const int o = roff(x - 2 * d);
      R.pold = ly.ld(A.phi_in + o);
      R.phic = ly.ld(A.phic_in + o);
    }
  };
  auto uq = [&](const QP<T>& qm, const QP<T>& q0, const QP<T>& qp, const T(&f)[9],
                NbrVals<T>& nv, CenterVals<T>& cv) {
    update_quantities<T, EXACT>(f, q0.xx, qp.xx, qm.xx, ynb_lanes<T, EXACT>(q0.xx), q0.yx,
                                qp.yx, qm.yx, ynb_lanes<T, EXACT>(q0.yx), q0.p, qp.p, qm.p,
                                ynb_lanes<T, EXACT>(q0.p), P, nv, cv);
  };

  // The chunk is marched upwards (D = +1) or downwards (D = -1), alternating
  // with the chunk parity as in sweep_item.  "prev"/"next" are rows x-D/x+D.
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
      return (D > 0) ? uf_stencil<T, EXACT>(prev, u0, next) : uf_stencil<T, EXACT>(next, u0, prev);
    };
    // stage A visits rows xs, xs+D, ...: two rows before the chunk, the
    // chunk, two rows after it; stage B trails A by two rows
    const int xs = ((D > 0) ? x0 : x1 - 1) - 2 * D;
    const int nA = x1 - x0 + 4;

    // stage A rolling state (as in sweep_item); rb, rc: old phi of rows x, x+D
    NbrVals<T> um, u0;
    CenterVals<T> c0;
    T fc[9];
    QP<T> qb, qc;
    T rb = T(0), rc = T(0);
    PairIn<T, EXACT> cur;
    {
      const QP<T> qa = load_qp(xs - 2 * D);
      qb = load_qp(xs - D);
      qc = load_qp(xs);
      const QP<T> qd = load_qp(xs + D);
      if constexpr (!EXACT) {
        rb = ly.ld(A.phi_in + roff(xs));
        rc = ly.ld(A.phi_in + roff(xs + D));
      }
      T ft[9];
      load_f(xs - D, ft);
      load_f(xs, fc);
      load_row(xs, D, cur);
      CenterVals<T> ct;
      uqd(qa, qb, qc, ft, um, ct);
      uqd(qb, qc, qd, fc, u0, c0);
      qb = qc;
      qc = qd;
    }
    // stage B rolling state: A's outputs of rows x-2D (bqb, bfc), x-D (bqc,
    // bfp); UpdateQuantities of rows x-3D (bum), x-2D (bu0); half-step values
    // of rows x-2D (cm2), x-D (cm1).  Zeroed: the first two iterations run
    // nothing of B, the next two only its UpdateQuantities, so no stored value
    // depends on them.
    NbrVals<T> bum{}, bu0{};
    CenterVals<T> bc0{};
    T bfc[9] = {}, bfp[9] = {};
    QP<T> bqb{}, bqc{};
    CenIn<T> cm1{}, cm2{};

    for (int i = 0, x = xs; i < nA; ++i, x += D) {
      PairIn<T, EXACT> nxt;
      load_row(x + D, D, nxt);

      // ---- stage A: UpdateQuantities on row x+D, UpdateFields on row x
      T rn = cur.q.p;
      if constexpr (!EXACT) cur.q.p += cur.dp;
      NbrVals<T> un;
      CenterVals<T> cn;
      uqd(qb, qc, cur.q, cur.f, un, cn);
      QP<T> qa_out;
      T fa[9];
      CenIn<T> cen;  // half-step values of row x (B uses them two rows later)
      if constexpr (kRebuild) {
#pragma unroll
        for (int v = 0; v < 9; ++v) cen.fn[v] = half_of(fc[v], cur.c.fn[v]);
        cen.qnxx = half_of(qb.xx, cur.c.qnxx);
        cen.qnyx = half_of(qb.yx, cur.c.qnyx);
        cen.phn = cur.c.phn;
      } else {
        cen = cur.c;
      }
      {
        const UFStencil<T, EXACT> s = ufs(um, u0, un);
        T Dxx, Dyx, Dp, om[9];
        update_fields<T, EXACT>(c0, qb.xx, qb.yx, qb.p, fc, s, P, Dxx, Dyx, Dp, om);
        qa_out.xx = cen.qnxx + T(.5) * Dxx;
        qa_out.yx = cen.qnyx + T(.5) * Dyx;
#pragma unroll
        for (int v = 0; v < 9; ++v) fa[v] = cen.fn[v] + T(.5) * om[v];
        if constexpr (EXACT) {
          qa_out.p = cen.phn + T(.5) * Dp;
        } else {
          // the separate sweep stores the increment, the next one adds it to
          // old phi: the same two roundings here
          const T inc = T(.5) * (cen.phn + Dp);
          qa_out.p = rb + inc;
        }
      }
      um = u0;
      u0 = un;
      c0 = cn;
#pragma unroll
      for (int v = 0; v < 9; ++v) fc[v] = cur.f[v];
      qb = qc;
      qc = cur.q;
      rb = rc;
      rc = rn;

      // ---- stage B: UpdateQuantities on row x-D, UpdateFields on row x-2D
      if (i >= 2) {  // wave-uniform
        NbrVals<T> bun;
        CenterVals<T> bcn;
        uqd(bqb, bqc, qa_out, bfp, bun, bcn);
        if (i >= 4) {
          const UFStencil<T, EXACT> s = ufs(bum, bu0, bun);
          T Dxx, Dyx, Dp, om[9];
          update_fields<T, EXACT>(bc0, bqb.xx, bqb.yx, bqb.p, bfc, s, P, Dxx, Dyx, Dp, om);
          wait_vmem<EXACT>();
          if (active) {
            const int ox = roff(x - 2 * D);
            ly.st(A.qxx_out + ox, cm2.qnxx + T(.5) * Dxx);
            ly.st(A.qyx_out + ox, cm2.qnyx + T(.5) * Dyx);
#pragma unroll
            for (int v = 0; v < 9; ++v)
              ly.st(A.f_out + (v * N + ox), cm2.fn[v] + T(.5) * om[v]);
            if constexpr (EXACT) {
              ly.st(A.phi_out + ox, cm2.phn + T(.5) * Dp);
            } else {
              // phi_new = phi_old + phic + (Dp_pred + Dp)/2 with a TwoSum, as
              // the separate last corrector
              const T pold = cur.pold;
              const T yv = T(.5) * (cm2.phn + Dp) + cur.phic;
              const T t = pold + yv;
              const T bp = t - pold;
              const T ap = t - bp;
              const T err = (pold - ap) + (yv - bp);
              ly.st(A.phi_out + ox, t);
              ly.st(A.phic_out + ox, err);
            }
          }
        }
        bum = bu0;
        bu0 = bun;
        bc0 = bcn;
      }
#pragma unroll
      for (int v = 0; v < 9; ++v) {
        bfc[v] = bfp[v];
        bfp[v] = fa[v];
      }
      bqb = bqc;
      bqc = qa_out;
      cm2 = cm1;
      cm1 = cen;
      cur = nxt;
    }
  };
  if (g.alt && !(chunk & 1)) {
    march(std::integral_constant<int, -1>{});
  } else {
    march(std::integral_constant<int, 1>{});
  }
}

template <class T, bool EXACT, bool HB>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? LYO_PAIR_WAVES_PER_EU : 1)))
pair_kernel(const SweepArgs<T> A, const Params<T> P, const Geo g) {
  int b = blockIdx.x;
  if (g.xcd) {
    const int nb8 = (int)gridDim.x >> 3;
    if (b < (nb8 << 3)) b = (b & 7) * nb8 + (b >> 3);
  }
  const int wid = __builtin_amdgcn_readfirstlane(b * ((int)blockDim.x >> 6) +
                                                 ((int)threadIdx.x >> 6));
  if (wid >= g.nwaves) return;
  pair_item<T, EXACT, HB>(A, P, g, wid);
}

// Batched fused pairs: persistent grid over the work items of many runs.
template <class T, bool EXACT, bool HB>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? LYO_PAIR_WAVES_PER_EU : 1)))
pair_batch_kernel(const SweepJob<T>* __restrict__ jobs, const int njobs, const int nitems) {
  const int nw = (int)(gridDim.x * (blockDim.x >> 6));
  int item = __builtin_amdgcn_readfirstlane((int)blockIdx.x * ((int)blockDim.x >> 6) +
                                            ((int)threadIdx.x >> 6));
  for (; item < nitems; item += nw) {
    const SweepJob<T>& J = find_job(jobs, njobs, item);
    pair_item<T, EXACT, HB>(J.A, J.P, J.g, item - J.first);
  }
}

template <class T, bool EXACT, bool HB>
inline hipError_t launch_pair(const SweepArgs<T>& a, const Params<T>& p, const Geo& g,
                              int waves_per_block, hipStream_t s) {
  if (!geo_ok(g, kPairStrip, 8) || waves_per_block < 1) return hipErrorInvalidValue;
  const int threads = 64 * waves_per_block;
  const int blocks = (g.nwaves + waves_per_block - 1) / waves_per_block;
  hipLaunchKernelGGL((pair_kernel<T, EXACT, HB>), dim3(blocks), dim3(threads), 0, s, a, p, g);
  return hipGetLastError();
}

}  // namespace lyo
