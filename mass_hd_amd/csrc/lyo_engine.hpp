// lyo_engine.hpp -- buffer sets and sweep sequencing of one lyotropic run.
//
// Sequencing follows Lyotropic::Step (src/models/lyotropic.cpp:500-525): one
// predictor sweep, then npc corrector sweeps.  Each sweep reads one buffer
// set and writes another, so no sweep ever reads what it writes:
//   S[2]  model state (f, Q, phi [, phic]) -- ping-pong across steps
//   P     predictor outputs: streamed f*, fn*, Q+D, Q+D/2, phi terms
//   C[2]  intermediate corrector outputs (npc >= 2) -- ping-pong
#pragma once

#include <vector>

#include "lyo_aux.hpp"
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
  int cur = 0;
  // inputs of the last UpdateQuantities (start of the last corrector sweep)
  bool have_uq = false;
  const T* uq_f = nullptr;
  const T* uq_qxx = nullptr;
  const T* uq_qyx = nullptr;
  const T* uq_phi = nullptr;
  const T* uq_dphi = nullptr;
};

struct LaunchCfg {
  int xc = 32;   // rows per wave chunk
  int wpb = 4;   // waves per block
  int xcd = 1;   // XCD-aware block order
};

// Optional per-sweep event timing.
struct SweepTimer {
  std::vector<hipEvent_t> ev;  // pairs
  std::vector<int> kind;       // 0 predictor, 1 corrector
  size_t used = 0;
};

inline Geo make_geo(int LX, int LY, int NP, const LaunchCfg& L) {
  Geo g;
  g.LX = LX;
  g.LY = LY;
  g.N = LX * LY;
  g.NP = NP;
  g.nstrips = (LY + kStrip - 1) / kStrip;
  g.xc = L.xc;
  g.nchunks = (LX + L.xc - 1) / L.xc;
  g.nwaves = g.nstrips * g.nchunks;
  g.xcd = L.xcd;
  return g;
}

template <class T, bool EXACT, int STAGE>
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
  hipError_t err = launch_sweep<T, EXACT, STAGE>(a, p, g, L.wpb, s);
  if (e1) (void)hipEventRecord(e1, s);
  return err;
}

template <class T, bool EXACT>
hipError_t run_steps(Bufs<T>& B, const Params<T>& P, int LX, int LY, int NP, int npc,
                     int nsteps, const LaunchCfg& L, hipStream_t s, SweepTimer* tm) {
  const Geo g = make_geo(LX, LY, NP, L);
  for (int st = 0; st < nsteps; ++st) {
    const int cur = B.cur, nxt = cur ^ 1;
    {
      SweepArgs<T> a{};
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
      hipError_t e = timed_launch<T, EXACT, PRED>(a, P, g, L, s, tm);
      if (e != hipSuccess) return e;
    }
    for (int c = 1; c <= npc; ++c) {
      SweepArgs<T> a{};
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
      const bool last = (c == npc);
      if (last) {
        a.f_out = B.Sf[nxt];
        a.qxx_out = B.Sqxx[nxt];
        a.qyx_out = B.Sqyx[nxt];
        a.phi_out = B.Sphi[nxt];
        if constexpr (!EXACT) {
          a.phic_in = B.Sphic[cur];
          a.phic_out = B.Sphic[nxt];
        }
        B.have_uq = true;
        B.uq_f = a.f_in;
        B.uq_qxx = a.qxx_in;
        B.uq_qyx = a.qyx_in;
        B.uq_phi = a.phi_in;
        B.uq_dphi = a.dphi_in;
      } else {
        const int ov = (c - 1) & 1;
        a.f_out = B.Cf[ov];
        a.qxx_out = B.Cqxx[ov];
        a.qyx_out = B.Cqyx[ov];
        a.phi_out = B.Cphi[ov];
      }
      hipError_t e = last ? timed_launch<T, EXACT, CORR_LAST>(a, P, g, L, s, tm)
                          : timed_launch<T, EXACT, CORR>(a, P, g, L, s, tm);
      if (e != hipSuccess) return e;
    }
    B.cur = nxt;
  }
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
  virtual size_t bytes() const = 0;
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
    auto get = [&](size_t n, T** p) -> hipError_t {
      hipError_t e = hipMalloc((void**)p, n * sizeof(T));
      if (e == hipSuccess) {
        ptrs_.push_back(*p);
        bytes_ += n * sizeof(T);
        e = hipMemset(*p, 0, n * sizeof(T));
      }
      return e;
    };
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      if (e == hipSuccess) e = get(n9, &B_.Sf[i]);
      if (e == hipSuccess) e = get(n1, &B_.Sqxx[i]);
      if (e == hipSuccess) e = get(n1, &B_.Sqyx[i]);
      if (e == hipSuccess) e = get(n1, &B_.Sphi[i]);
      if (!EXACT && e == hipSuccess) e = get(n1, &B_.Sphic[i]);
    }
    if (e == hipSuccess) e = get(n9, &B_.Pfs);
    if (e == hipSuccess) e = get(n9, &B_.Pfns);
    if (e == hipSuccess) e = get(n1, &B_.Pqsxx);
    if (e == hipSuccess) e = get(n1, &B_.Pqsyx);
    if (e == hipSuccess) e = get(n1, &B_.Pqnxx);
    if (e == hipSuccess) e = get(n1, &B_.Pqnyx);
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
    return field_to_f64<T, EXACT>(B_, P_, LX_, LY_, NP_, id, rho_, d_out, s);
  }
  hipError_t sums(double* d_part, int max_part, double* d_out, hipStream_t s) override {
    return state_sums<T, EXACT>(B_, LX_ * LY_, NP_, rho_, d_part, max_part, d_out, s);
  }
  size_t bytes() const override { return bytes_; }

 private:
  Bufs<T> B_;
  Params<T> P_{};
  double rho_ = 0;
  int LX_ = 0, LY_ = 0, NP_ = 0, npc_ = 1;
  std::vector<T*> ptrs_;
  size_t bytes_ = 0;
};

Engine* make_engine_fp64();
Engine* make_engine_fp32();

}  // namespace lyo
