// lyo_capi.cpp -- the C ABI of include/mass_lyo.h.
//
// Owns the device, the stream, the per-run engine (FP64 or FP32) and the FP64
// staging buffers used to move host AoS fields across the boundary.  No
// exception crosses the ABI: every entry point returns a status code and keeps
// the message in the handle.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mass_lyo.h"
#include "lyo_config.hpp"
#include "lyo_handle.hpp"
#include "lyo_net_internal.hpp"

namespace {
thread_local std::string g_create_error;
// largest plane stride of mass_lyo_state_sums_noalloc (int indexing of nine planes)
constexpr int64_t kNoallocMaxStride = 2147483647LL / 9;
}  // namespace

static int fail(mass_lyo_t h, int code, const std::string& msg) {
  if (h)
    h->err = msg;
  else
    g_create_error = msg;
  return code;
}

static int hip_fail(mass_lyo_t h, hipError_t e, const char* what) {
  return fail(h, MASS_LYO_EDEVICE,
              std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
}

static const double* param_block(const mass_lyo_params& p) { return &p.GammaQ; }

extern "C" {

int mass_lyo_abi_version(void) { return MASS_LYO_ABI_VERSION; }

const char* mass_lyo_last_error(mass_lyo_t h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

// Rows per chunk of a one-kernel step for an LX x LY grid on a GPU with
// `slots` resident work items (waves, or blocks of the block-cooperative
// step) of `strip` columns that repeat `repeat` rows per chunk: minimises
// ceil(items / slots) x (rows + repeated rows); the shortest such chunk.
static int chunk_rows_for(int LX, int LY, int strip, int repeat, int slots) {
  const long long nstrips = (LY + strip - 1) / strip;
  if (slots < 1) slots = 1;
  int best = 64;
  long long best_cost = -1;
  for (int xc = 8; xc <= (LX > 8 ? LX : 8); ++xc) {
    const long long waves = nstrips * ((LX + xc - 1) / xc);
    const long long cost = ((waves + slots - 1) / slots) * (xc + repeat);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = xc;
    }
  }
  return best;
}

// the network part of surrogate steps can run as one kernel (FP32 state,
// LX and LY >= 16)
static bool fused_net_ok(const mass_lyo_t h) {
  return h->prm.precision == MASS_LYO_FP32 && h->prm.LX >= lyo::kNetMinSide &&
         h->prm.LY >= lyo::kNetMinSide;
}

int mass_lyo_create(const mass_lyo_params* p, mass_lyo_t* out) {
  if (!p || !out) return fail(nullptr, MASS_LYO_EINVAL, "null argument");
  *out = nullptr;
  if (p->LX < 2 || p->LY < 2)
    return fail(nullptr, MASS_LYO_EINVAL, "LX and LY must be at least 2");
  const long long N = (long long)p->LX * (long long)p->LY;
  if (N * 9 >= (1LL << 31) - 4096)
    return fail(nullptr, MASS_LYO_EINVAL, "grid too large (LX*LY*9 must stay below 2^31)");
  if (p->npc < 1) return fail(nullptr, MASS_LYO_EINVAL, "npc must be at least 1");
  if (p->precision != MASS_LYO_FP64 && p->precision != MASS_LYO_FP32)
    return fail(nullptr, MASS_LYO_EINVAL, "precision must be MASS_LYO_FP64 or MASS_LYO_FP32");
  if (p->conserve_phi)
    return fail(nullptr, MASS_LYO_EINVAL,
                "conserve_phi is not part of the plain lyotropic model");
  if (p->bc < 0 || p->bc > 4)
    return fail(nullptr, MASS_LYO_EINVAL,
                "bc must be 0 (periodic), 1/2 (free-/no-slip channel) or 3/4 (free-/no-slip box)");
  if (p->bc != 0 && p->precision != MASS_LYO_FP64)
    return fail(nullptr, MASS_LYO_EINVAL, "walled grids (bc 1-4) run in FP64 only");
  const double* v = param_block(*p);
  for (int i = 0; i < 14; ++i)
    if (!std::isfinite(v[i])) return fail(nullptr, MASS_LYO_EINVAL, "non-finite parameter");
  if (!(p->rho > 0)) return fail(nullptr, MASS_LYO_EINVAL, "rho must be positive");

  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1)
    return fail(nullptr, MASS_LYO_EDEVICE, "no HIP device available (this library runs on gfx950)");
  int dev = 0;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDevice");
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, MASS_LYO_EDEVICE,
                std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");

  if (p->plane_pad < 0 || p->plane_pad > (1 << 20))
    return fail(nullptr, MASS_LYO_EINVAL, "plane_pad must be in [0, 2^20]");
  // planes start on 256-byte boundaries and are staggered by kPlaneStagger
  // values, so that equal offsets in the nine population planes of a
  // power-of-two grid do not share HBM channels (measured at 4096^2: ~1%)
  constexpr long long kPlaneStagger = 256;
  const long long NP = ((N + p->plane_pad + kPlaneStagger + 63) / 64) * 64;
  if (NP * 9 >= (1LL << 31) - 4096)
    return fail(nullptr, MASS_LYO_EINVAL, "grid too large with this plane_pad");
  auto h = std::make_unique<mass_lyo_s>();
  h->prm = *p;
  h->N = (int)N;
  h->NP = (int)NP;
  h->device = dev;
  e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipStreamCreate");
  if (p->bc != 0)
    h->eng.reset(lyo::make_engine_walls(p->bc));
  else
    h->eng.reset(p->precision == MASS_LYO_FP64 ? lyo::make_engine_fp64()
                                               : lyo::make_engine_fp32());
  e = h->eng->alloc(p->LX, p->LY, h->NP, p->npc);
  if (e == hipSuccess) e = hipMalloc((void**)&h->d_stage, sizeof(double) * 12 * (size_t)N);
  if (e == hipSuccess) e = hipMalloc((void**)&h->d_part, sizeof(double) * (2 * kMaxPart + 2));
  if (e != hipSuccess) {
    std::string msg = std::string("device allocation failed: ") + hipGetErrorString(e);
    mass_lyo_destroy(h.release());
    return fail(nullptr, MASS_LYO_ENOMEM, msg);
  }
  h->eng->set_params(v);
  // default chunk length: the longest chunk (least prologue rework) that
  // still gives ~32 waves per SIMD, so waiting waves always have a successor
  // (measured at 4096^2: 8 rows beat 16/32/64)
  const int nstrips = (p->LY + lyo::kStrip - 1) / lyo::kStrip;
  int xc = 64;
  while (xc > 4 && (long long)nstrips * ((p->LX + xc - 1) / xc) < 32768) xc >>= 1;
  h->launch.xc = xc;
  // the fused pair holds two stages' rows (2 waves per SIMD): ~16 waves per
  // SIMD (measured at 4096^2: 16 rows ~ 64 rows > 32 > 8)
  const int nstrips2 = (p->LY + lyo::kPairStrip - 1) / lyo::kPairStrip;
  int xc2 = 64;
  while (xc2 > 4 && (long long)nstrips2 * ((p->LX + xc2 - 1) / xc2) < 16384) xc2 >>= 1;
  h->launch.xc2 = xc2;
  // the one-kernel step: all waves of a launch take about equally long, so
  // a launch lasts (rounds of the GPU's wave slots) x (rows per chunk + the
  // rows a chunk repeats); the chunk length minimising that product (measured
  // at 4096^2 FP32 npc=2, rounds 0.96 / 1.93 / 2.47 / 2.2 / 1.58 / 1.23 for
  // 164 / 82 / 64 / 72 / 100 / 128 rows: 30.8k / 30.6k / 30.0k / 27.8k /
  // 26.5k / 25.0k MLUPS)
  const int per_simd = p->precision == MASS_LYO_FP64 ? 1 : 2;  // resident waves per SIMD
  h->launch.xc3 = chunk_rows_for(p->LX, p->LY, p->npc == 1 ? lyo::kFull1Strip : lyo::kFullStrip,
                                 p->npc == 1 ? 4 : 8, prop.multiProcessorCount * 4 * per_simd);
  // the one-corrector form (surrogate steps of npc = 2 runs) has its own
  h->launch.xc3b = chunk_rows_for(p->LX, p->LY, lyo::kFull1Strip, 4,
                                  prop.multiProcessorCount * 4 * per_simd);
  // the block-cooperative variant: one block of four waves per CU and wave
  // slot per SIMD (its LDS allows two FP32 blocks per CU, one FP64 block)
  h->launch.xc4 = chunk_rows_for(p->LX, p->LY, lyo::kXStrip, 8,
                                 prop.multiProcessorCount * per_simd);
  // the one-kernel network step of surrogate runs: two waves per SIMD
  h->launch.xc5 = chunk_rows_for(p->LX, p->LY, lyo::kNetStrip, lyo::kNetRepeat,
                                 prop.multiProcessorCount * 4 * 2);
  h->sur.fused = fused_net_ok(h.get()) ? 1 : 0;
  *out = h.release();
  return MASS_LYO_OK;
}

void mass_lyo_destroy(mass_lyo_t h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->eng) h->eng->release();
  h->eng.reset();
  if (h->d_stage) (void)hipFree(h->d_stage);
  if (h->d_part) (void)hipFree(h->d_part);
  for (void* p : h->diag_allocs) (void)hipFree(p);
  for (void* p : {(void*)h->sur.delta, (void*)h->sur.part, (void*)h->sur.d_oscale})
    if (p) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

static int upload_state(mass_lyo_t h, const double* ff, const double* qxx, const double* qyx,
                        const double* phi) {
  const size_t N = (size_t)h->N;
  double* d_ff = h->d_stage;
  double* d_qxx = d_ff + 9 * N;
  double* d_qyx = d_qxx + N;
  double* d_phi = d_qyx + N;
  hipError_t e = hipMemcpyAsync(d_ff, ff, sizeof(double) * 9 * N, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_qxx, qxx, sizeof(double) * N, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_qyx, qyx, sizeof(double) * N, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_phi, phi, sizeof(double) * N, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = h->eng->pack(d_ff, d_qxx, d_qyx, d_phi, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "state upload");
  h->has_state = true;
  return MASS_LYO_OK;
}

int mass_lyo_set_state(mass_lyo_t h, const double* ff, const double* qxx, const double* qyx,
                       const double* phi) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (!ff || !qxx || !qyx || !phi) return fail(h, MASS_LYO_EINVAL, "null field pointer");
  // a go-or-grow run's state also holds m, chi, the masks, their counters and
  // the uniforms drawn ahead, which this call cannot set consistently
  if (h->growth) return fail(h, MASS_LYO_EINVAL, "go-or-grow runs take their state from configure");
  return upload_state(h, ff, qxx, qyx, phi);
}

int mass_lyo_configure(mass_lyo_t h, const mass_lyo_init* init, double* ftot, double* ptot) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (!init) return fail(h, MASS_LYO_EINVAL, "null init");
  lyo::HostInit hi;
  hi.LX = (unsigned)h->prm.LX;
  hi.LY = (unsigned)h->prm.LY;
  hi.config = init->config;
  hi.seed = init->seed;
  hi.level = init->level;
  hi.radius = init->radius;
  hi.angle_deg = init->angle_deg;
  hi.noise = init->noise;
  hi.qamp = init->qamp;
  hi.phinoise = init->phinoise;
  hi.conc = h->prm.conc;
  hi.rho = h->prm.rho;
  hi.bc = h->prm.bc;
  std::vector<double> ff, qxx, qyx, phi;
  std::string msg;
  double ft = 0, pt = 0;
  if (lyo::configure_host(hi, h->gen, ff, qxx, qyx, phi, &ft, &pt, msg) != 0)
    return fail(h, MASS_LYO_EINVAL, msg);
  const int rc = upload_state(h, ff.data(), qxx.data(), qyx.data(), phi.data());
  if (rc == MASS_LYO_OK && h->growth) {
    // GoOrGrow::Configure (go-or-grow.cpp:34-40)
    const hipError_t e = h->eng->growth_configure(&ft, &pt, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "go-or-grow configure");
  }
  if (rc == MASS_LYO_OK) {
    if (ftot) *ftot = ft;
    if (ptot) *ptot = pt;
  }
  return rc;
}

// One step in surrogate mode: a judged plain step every judge_every steps
// (plain step, network on its one-corrector state, velocity error against the
// plain result), plain steps while a failed judgement is being served, and
// otherwise predictor + one corrector + the network's increment.
static int surrogate_step(mass_lyo_t h) {
  auto& S = h->sur;
  const int LX = h->prm.LX, LY = h->prm.LY;
  hipError_t e = hipSuccess;
  const bool judge = (S.count % S.judge_every) == 0;
  if (judge) {
    e = h->eng->steps(1, h->launch, h->stream, nullptr);
    __bf16* in = nullptr;
    const float* nrm = nullptr;
    if (e == hipSuccess) e = lyo::net_input(S.net, (size_t)LX * LY, &in, &nrm);
    if (e == hipSuccess) e = h->eng->net_features(1, nullptr, nrm, in, h->stream);
    // the network's own output scales here: the judge applies the
    // calibration itself and measures the uncalibrated increments
    if (e == hipSuccess) e = lyo::net_predict_input(S.net, LX, LY, S.delta, h->stream);
    constexpr int nparts = 256;
    lyo::NetCalib cal;
    for (int c = 0; c < 5; ++c) cal.a[c] = S.calib ? S.alpha[c] : 1.f;
    if (e == hipSuccess) e = h->eng->net_judge(S.delta, cal, S.part, nparts, h->stream);
    std::vector<double> part((size_t)lyo::kJudgeSums * nparts);
    if (e == hipSuccess)
      e = hipMemcpyAsync(part.data(), S.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                         h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "judged step");
    double sum[lyo::kJudgeSums] = {};
    for (int i = 0; i < nparts; ++i)
      for (int j = 0; j < lyo::kJudgeSums; ++j) sum[j] += part[(size_t)lyo::kJudgeSums * i + j];
    const double err = sum[1] > 0 ? std::sqrt(sum[0] / sum[1]) : 0.0;
    if (S.calib) {
      // least squares of the plain corrector's increments t on the network's
      // p, per output, over the judged steps with weight kDecay^age
      constexpr double kDecay = 0.9;
      float sc[8];
      lyo::net_out_scale(S.net, sc);
      for (int c = 0; c < 5; ++c) {
        S.ctp[c] = kDecay * S.ctp[c] + sum[2 + 2 * c];
        S.cpp[c] = kDecay * S.cpp[c] + sum[3 + 2 * c];
        double a = S.cpp[c] > 0 ? S.ctp[c] / S.cpp[c] : 1.0;
        a = a < 0.25 ? 0.25 : (a > 4.0 ? 4.0 : a);
        S.alpha[c] = (float)a;
        S.h_oscale[c] = sc[c] * S.alpha[c];
      }
      for (int c = 5; c < 8; ++c) S.h_oscale[c] = sc[c];
      e = hipMemcpyAsync(S.d_oscale, S.h_oscale, sizeof(S.h_oscale), hipMemcpyHostToDevice,
                         h->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) return hip_fail(h, e, "calibration");
    }
    S.stats.judged++;
    S.stats.plain_steps++;
    S.stats.last_error = err;
    if (err > S.stats.max_error) S.stats.max_error = err;
    if (!(err <= S.tol)) {
      S.stats.fallbacks++;
      S.plain_left = S.judge_every - 1;
    } else {
      S.plain_left = 0;
    }
  } else if (S.plain_left > 0) {
    e = h->eng->steps(1, h->launch, h->stream, nullptr);
    S.plain_left--;
    S.stats.plain_steps++;
  } else if (S.fused) {
    e = h->eng->steps_with(1, 1, h->launch, h->stream);
    if (e == hipSuccess)
      e = h->eng->net_fused(lyo::net_fused_weights(S.net, S.calib ? S.d_oscale : nullptr),
                            h->launch.xc5, nullptr, h->stream);
    S.stats.surrogate_steps++;
  } else {
    e = h->eng->steps_with(1, 1, h->launch, h->stream);
    __bf16* in = nullptr;
    const float* nrm = nullptr;
    if (e == hipSuccess) e = lyo::net_input(S.net, (size_t)LX * LY, &in, &nrm);
    if (e == hipSuccess) e = h->eng->net_features(0, nullptr, nrm, in, h->stream);
    if (e == hipSuccess)
      e = lyo::net_predict_input(S.net, LX, LY, S.delta, h->stream,
                                 S.calib ? S.d_oscale : nullptr);
    if (e == hipSuccess) e = h->eng->net_apply(S.delta, h->stream);
    S.stats.surrogate_steps++;
  }
  if (e != hipSuccess) return hip_fail(h, e, "surrogate step");
  S.count++;
  S.stats.steps++;
  return MASS_LYO_OK;
}

int mass_lyo_step(mass_lyo_t h, int32_t nsteps) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (nsteps < 0) return fail(h, MASS_LYO_EINVAL, "negative step count");
  if (h->sur.net) {
    for (int i = 0; i < nsteps; ++i) {
      const int rc = surrogate_step(h);
      if (rc) return rc;
    }
    return MASS_LYO_OK;
  }
  hipError_t e = h->eng->steps(nsteps, h->launch, h->stream, nullptr);
  if (e != hipSuccess) return hip_fail(h, e, "sweep launch");
  return MASS_LYO_OK;
}

// calibration factors back to one, the network's own output scales
static void reset_calibration(mass_lyo_t h) {
  auto& S = h->sur;
  for (int c = 0; c < 5; ++c) {
    S.ctp[c] = S.cpp[c] = 0;
    S.alpha[c] = 1.f;
  }
  lyo::net_out_scale(S.net, S.h_oscale);
}

int mass_lyo_surrogate_enable(mass_lyo_t h, mass_lyo_net_t net, int32_t judge_every, double tol) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (!net || !lyo::net_ready(net))
    return fail(h, MASS_LYO_EINVAL, "the network is not trained (no normalisation yet)");
  if (h->prm.npc < 2) return fail(h, MASS_LYO_EINVAL, "the surrogate replaces correctors 2..npc: npc must be >= 2");
  if (h->prm.bc != 0 || h->growth)
    return fail(h, MASS_LYO_EINVAL, "the surrogate needs a plain periodic run (bc 0)");
  if (h->prm.LY % 16) return fail(h, MASS_LYO_EINVAL, "the surrogate needs LY to be a multiple of 16");
  if (judge_every < 1 || !(tol > 0)) return fail(h, MASS_LYO_EINVAL, "judge_every >= 1 and tol > 0");
  auto& S = h->sur;
  hipError_t e = hipSuccess;
  const size_t N = (size_t)h->N;
  if (!S.delta) e = hipMalloc((void**)&S.delta, N * 8 * sizeof(float));
  if (e == hipSuccess && !S.part)
    e = hipMalloc((void**)&S.part, (size_t)lyo::kJudgeSums * 256 * sizeof(double));
  if (e == hipSuccess && !S.d_oscale) e = hipMalloc((void**)&S.d_oscale, 8 * sizeof(float));
  if (e != hipSuccess) return hip_fail(h, e, "surrogate buffers");
  S.net = net;
  reset_calibration(h);
  e = hipMemcpyAsync(S.d_oscale, S.h_oscale, sizeof(S.h_oscale), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "surrogate buffers");
  S.judge_every = judge_every;
  S.tol = tol;
  S.count = 0;
  S.plain_left = 0;
  S.stats = mass_lyo_surrogate_stats{};
  return MASS_LYO_OK;
}

int mass_lyo_surrogate_disable(mass_lyo_t h) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  h->sur.net = nullptr;
  return MASS_LYO_OK;
}

int mass_lyo_surrogate_set_kernels(mass_lyo_t h, int32_t fused) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (fused != 0 && fused != 1) return fail(h, MASS_LYO_EINVAL, "fused is 0 or 1");
  if (fused && !fused_net_ok(h))
    return fail(h, MASS_LYO_EINVAL, "the one-kernel network step needs FP32 and LX, LY >= 16");
  h->sur.fused = fused;
  return MASS_LYO_OK;
}

int mass_lyo_surrogate_set_calibration(mass_lyo_t h, int32_t on) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (on != 0 && on != 1) return fail(h, MASS_LYO_EINVAL, "on is 0 or 1");
  if (!h->sur.net) return fail(h, MASS_LYO_EINVAL, "enable the surrogate first");
  h->sur.calib = on != 0;
  reset_calibration(h);
  hipError_t e = hipMemcpyAsync(h->sur.d_oscale, h->sur.h_oscale, sizeof(h->sur.h_oscale),
                                hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "calibration");
  return MASS_LYO_OK;
}

int mass_lyo_surrogate_get_calibration(mass_lyo_t h, double* factors5) {
  if (!h || !factors5) return fail(h, MASS_LYO_EINVAL, "null argument");
  for (int c = 0; c < 5; ++c) factors5[c] = h->sur.alpha[c];
  return MASS_LYO_OK;
}

int mass_lyo_surrogate_get_stats(mass_lyo_t h, mass_lyo_surrogate_stats* out) {
  if (!h || !out) return fail(h, MASS_LYO_EINVAL, "null argument");
  *out = h->sur.stats;
  return MASS_LYO_OK;
}

int mass_lyo_synchronize(mass_lyo_t h) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  hipError_t e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "hipStreamSynchronize");
  return MASS_LYO_OK;
}

int mass_lyo_get_field(mass_lyo_t h, int32_t field, double* out) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (!out) return fail(h, MASS_LYO_EINVAL, "null output");
  if (field < 0 || field >= MASS_LYO_NUM_FIELDS) return fail(h, MASS_LYO_EINVAL, "unknown field id");
  if (field >= MASS_LYO_M && !h->growth)
    return fail(h, MASS_LYO_EINVAL, "m, chi and the masks belong to go-or-grow runs");
  const size_t n = (field == MASS_LYO_FF ? 9 : 1) * (size_t)h->N;
  hipError_t e = h->eng->field(field, h->d_stage, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, h->d_stage, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "field download");
  for (size_t i = 0; i < n; ++i)
    if (out[i] != out[i]) return fail(h, MASS_LYO_ENAN, "nan found while writing file.");
  return MASS_LYO_OK;
}

int mass_lyo_sums(mass_lyo_t h, double* fsum, double* phisum) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  double r[2] = {0, 0};
  double* d_out = h->d_part + 2 * kMaxPart;
  hipError_t e = h->eng->sums(h->d_part, kMaxPart, d_out, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(r, d_out, sizeof(r), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "sums");
  if (fsum) *fsum = r[0];
  if (phisum) *phisum = r[1];
  return MASS_LYO_OK;
}

}  // extern "C"

// Scratch of mass_lyo_state_sums_noalloc: the per-block partial sums of Σf and
// Σφ (kMaxPart each).  One copy per GPU, held until the process ends; host
// code reaches it only through the address hipGetSymbolAddress gives for the
// current device (the array's name is a host-side stand-in).
static __device__ __align__(256) double sums_scratch[2 * kMaxPart];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the current device's copy of sums_scratch (looked up once per device)
hipError_t scratch_of_current_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(sums_scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}
}  // namespace

extern "C" {

int mass_lyo_state_sums_noalloc(int32_t precision, const void* f, int64_t plane_stride,
                                const void* phi, const void* phic, int64_t n, double rho,
                                double* out, void* stream) {
  if (precision != MASS_LYO_FP64 && precision != MASS_LYO_FP32)
    return fail(nullptr, MASS_LYO_EINVAL, "precision must be MASS_LYO_FP64 or MASS_LYO_FP32");
  if (!f || !phi || !out) return fail(nullptr, MASS_LYO_EINVAL, "null device pointer");
  if (precision == MASS_LYO_FP64 && phic)
    return fail(nullptr, MASS_LYO_EINVAL, "FP64 states have no phi compensation plane");
  if (n < 1) return fail(nullptr, MASS_LYO_EINVAL, "n must be at least 1");
  if (plane_stride < n) return fail(nullptr, MASS_LYO_EINVAL, "plane_stride must be at least n");
  if (plane_stride > kNoallocMaxStride)
    return fail(nullptr, MASS_LYO_EINVAL,
                "plane_stride too large: the nine population planes must span at most "
                "2^31-1 elements (plane_stride <= " + std::to_string(kNoallocMaxStride) + ")");
  double* scratch = nullptr;
  hipError_t e = scratch_of_current_device(&scratch);
  if (e != hipSuccess) return hip_fail(nullptr, e, "device scratch address");
  const int N = (int)n, NP = (int)plane_stride;
  // the partition of state_sums (lyo_engine.hpp): the same tree for the same n
  int nparts = (N + 4095) / 4096;
  if (nparts > kMaxPart) nparts = kMaxPart;
  const int per = (N + nparts - 1) / nparts;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == MASS_LYO_FP64)
    hipLaunchKernelGGL((lyo::partial_sums_kernel<double>), dim3(nparts), dim3(256), 0, s,
                       static_cast<const double*>(f), static_cast<const double*>(phi),
                       (const double*)nullptr, N, NP, per, scratch, scratch + kMaxPart);
  else
    hipLaunchKernelGGL((lyo::partial_sums_kernel<float>), dim3(nparts), dim3(256), 0, s,
                       static_cast<const float*>(f), static_cast<const float*>(phi),
                       static_cast<const float*>(phic), N, NP, per, scratch,
                       scratch + kMaxPart);
  const double add = precision == MASS_LYO_FP64 ? 0.0 : (double)N * rho;
  hipLaunchKernelGGL(lyo::final_sums_kernel, dim3(1), dim3(256), 0, s, scratch,
                     scratch + kMaxPart, nparts, add, out);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(nullptr, e, "state sums");
  return MASS_LYO_OK;
}

int mass_lyo_diagnostics(mass_lyo_t h, mass_lyo_diag* out, double* charge) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (!out) return fail(h, MASS_LYO_EINVAL, "null output");
  if (h->prm.bc != 0 || h->growth)
    return fail(h, MASS_LYO_EINVAL,
                "the diagnostics (defect winding numbers) need a plain periodic run");
  hipError_t e = hipSuccess;
  lyo::DiagScratch& d = h->diag;
  const size_t N = (size_t)h->N;
  if (!d.part) {
    auto get = [&](size_t bytes, void** p) {
      if (e == hipSuccess) e = hipMalloc(p, bytes);
      if (e == hipSuccess) h->diag_allocs.push_back(*p);
    };
    d.max_part = kMaxPart;
    get(sizeof(double) * 5 * kMaxPart, (void**)&d.part);
    get(sizeof(double) * 8, (void**)&d.out);
    get(sizeof(unsigned long long) * (lyo::kHistBins + 2), (void**)&d.hist);
    get(sizeof(unsigned long long) * 2, (void**)&d.cnt);
    get(N, (void**)&d.sgn);
    get(sizeof(int) * N, (void**)&d.label);
    if (e != hipSuccess) {
      for (void* p : h->diag_allocs) (void)hipFree(p);
      h->diag_allocs.clear();
      d = lyo::DiagScratch();
      return hip_fail(h, e, "diagnostics allocation");
    }
  }
  double* w_dev = charge ? h->d_stage : nullptr;
  e = h->eng->diagnostics(d, w_dev, h->stream);
  double r[5];
  unsigned long long hist[lyo::kHistBins + 2], cnt[2];
  if (e == hipSuccess) e = hipMemcpyAsync(r, d.out, sizeof(r), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(hist, d.hist, sizeof(hist), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(cnt, d.cnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess && charge)
    e = hipMemcpyAsync(charge, w_dev, sizeof(double) * N, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "diagnostics");
  out->u_rms = std::sqrt(r[0] / (double)N);
  out->u_mean = r[1] / (double)N;
  out->phi_mean = r[2] / (double)N;
  out->phi_min = r[3];
  out->phi_max = r[4];
  out->defects_pos = (int64_t)cnt[0];
  out->defects_neg = (int64_t)cnt[1];
  out->defect_density = (double)(cnt[0] + cnt[1]) / (double)N;
  for (int i = 0; i < lyo::kHistBins + 2; ++i) out->phi_hist[i] = hist[i];
  return MASS_LYO_OK;
}

int mass_lyo_time_steps(mass_lyo_t h, int32_t nsteps, double* total_ms, double* sweep_ms) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (nsteps < 1) return fail(h, MASS_LYO_EINVAL, "nsteps must be at least 1");
  hipEvent_t t0, t1;
  hipError_t e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  if (e != hipSuccess) return hip_fail(h, e, "hipEventCreate");
  lyo::SweepTimer tm;
  if (sweep_ms) {
    const size_t pairs = (size_t)nsteps * (1 + (size_t)h->prm.npc);
    tm.ev.resize(2 * pairs);
    tm.kind.resize(pairs);
    for (auto& ev : tm.ev) {
      e = hipEventCreate(&ev);
      if (e != hipSuccess) break;
    }
  }
  int rc = MASS_LYO_OK;
  if (e == hipSuccess) e = hipEventRecord(t0, h->stream);
  if (e == hipSuccess) {
    if (h->sur.net) {
      // surrogate mode: the steps as mass_lyo_step runs them (judged steps
      // synchronise with the host); no per-sweep times
      for (int i = 0; i < nsteps && rc == MASS_LYO_OK; ++i) rc = surrogate_step(h);
    } else {
      e = h->eng->steps(nsteps, h->launch, h->stream, sweep_ms ? &tm : nullptr);
    }
  }
  if (e == hipSuccess && rc == MASS_LYO_OK) e = hipEventRecord(t1, h->stream);
  if (e == hipSuccess && rc == MASS_LYO_OK) e = hipEventSynchronize(t1);
  float ms = 0;
  if (rc != MASS_LYO_OK) {  // the surrogate step left its message in h->err
    for (auto& ev : tm.ev)
      if (ev) (void)hipEventDestroy(ev);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return rc;
  }
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
  if (e == hipSuccess && sweep_ms) {
    double acc[4] = {0, 0, 0, 0};
    int cnt[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < tm.used / 2 && e == hipSuccess; ++i) {
      float m = 0;
      e = hipEventElapsedTime(&m, tm.ev[2 * i], tm.ev[2 * i + 1]);
      acc[tm.kind[i]] += m;
      cnt[tm.kind[i]] += 1;
    }
    sweep_ms[0] = cnt[0] ? acc[0] / cnt[0] : 0;
    sweep_ms[1] = cnt[1] ? acc[1] / cnt[1] : 0;
    sweep_ms[2] = cnt[2] ? acc[2] / cnt[2] : 0;
    sweep_ms[3] = cnt[3] ? acc[3] / cnt[3] : 0;
  }
  for (auto& ev : tm.ev)
    if (ev) (void)hipEventDestroy(ev);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (e != hipSuccess) return hip_fail(h, e, "timed steps");
  if (total_ms) *total_ms = ms;
  return MASS_LYO_OK;
}

int mass_lyo_set_launch(mass_lyo_t h, int32_t chunk_rows, int32_t waves_per_block,
                        int32_t xcd_order) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (chunk_rows < 0 || chunk_rows > 4096 || waves_per_block < 0 || waves_per_block > 4 ||
      xcd_order < -1 || xcd_order > 1)
    return fail(h, MASS_LYO_EINVAL,
                "chunk_rows must be in [1,4096], waves_per_block in [1,4], xcd_order in {-1,0,1}");
  if (chunk_rows) h->launch.xc = chunk_rows;
  if (waves_per_block) h->launch.wpb = waves_per_block;
  if (xcd_order >= 0) h->launch.xcd = xcd_order;
  return MASS_LYO_OK;
}

int mass_lyo_set_schedule(mass_lyo_t h, int32_t fuse_last_pair, int32_t pair_chunk_rows,
                          int32_t alternate, int32_t one_kernel_step, int32_t step_chunk_rows) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (fuse_last_pair < -1 || fuse_last_pair > 1 || pair_chunk_rows < 0 ||
      pair_chunk_rows > 4096 || alternate < -1 || alternate > 1 || one_kernel_step < -1 ||
      one_kernel_step > 2 || step_chunk_rows < 0 || step_chunk_rows > 4096)
    return fail(h, MASS_LYO_EINVAL,
                "fuse_last_pair and alternate in {-1,0,1}, one_kernel_step in {-1,0,1,2}, "
                "chunk rows in [0,4096]");
  if (fuse_last_pair >= 0) h->launch.fuse = fuse_last_pair;
  if (pair_chunk_rows) h->launch.xc2 = pair_chunk_rows;
  if (alternate >= 0) h->launch.alt = alternate;
  if (one_kernel_step >= 0) h->launch.full = one_kernel_step;
  if (step_chunk_rows) h->launch.xc3 = h->launch.xc3b = h->launch.xc4 = step_chunk_rows;
  return MASS_LYO_OK;
}

namespace {
struct CheckpointHeader {
  char magic[8];
  uint32_t version;
  uint32_t precision;
  int32_t LX, LY, npc, planes;
  double rho;
  uint64_t state_bytes;
  int32_t bc;  // 0 in checkpoints of periodic runs
  uint8_t reserved[12];
};
static_assert(sizeof(CheckpointHeader) == 64, "64-byte checkpoint header");
constexpr char kCheckpointMagic[8] = {'M', 'A', 'S', 'S', 'L', 'Y', 'O', '1'};
}  // namespace

int mass_lyo_checkpoint_size(mass_lyo_t h, uint64_t* bytes) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (h->growth) return fail(h, MASS_LYO_EINVAL, "go-or-grow runs have no checkpoints");
  if (!bytes) return fail(h, MASS_LYO_EINVAL, "null size pointer");
  *bytes = sizeof(CheckpointHeader) + h->eng->state_bytes();
  return MASS_LYO_OK;
}

int mass_lyo_checkpoint_save(mass_lyo_t h, void* buf, uint64_t bytes) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (h->growth) return fail(h, MASS_LYO_EINVAL, "go-or-grow runs have no checkpoints");
  if (!buf) return fail(h, MASS_LYO_EINVAL, "null buffer");
  if (!h->has_state) return fail(h, MASS_LYO_EINVAL, "no state yet: configure or set_state first");
  const uint64_t need = sizeof(CheckpointHeader) + h->eng->state_bytes();
  if (bytes < need) return fail(h, MASS_LYO_EINVAL, "checkpoint buffer too small");
  CheckpointHeader hd{};
  std::memcpy(hd.magic, kCheckpointMagic, 8);
  hd.version = 1;
  hd.precision = (uint32_t)h->prm.precision;
  hd.LX = h->prm.LX;
  hd.LY = h->prm.LY;
  hd.npc = h->prm.npc;
  hd.planes = h->eng->state_planes();
  hd.rho = h->prm.rho;
  hd.state_bytes = h->eng->state_bytes();
  hd.bc = h->prm.bc;
  std::memcpy(buf, &hd, sizeof(hd));
  const hipError_t e = h->eng->save_state(static_cast<char*>(buf) + sizeof(hd), h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "checkpoint save");
  return MASS_LYO_OK;
}

int mass_lyo_checkpoint_load(mass_lyo_t h, const void* buf, uint64_t bytes) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (h->growth) return fail(h, MASS_LYO_EINVAL, "go-or-grow runs have no checkpoints");
  if (!buf) return fail(h, MASS_LYO_EINVAL, "null buffer");
  CheckpointHeader hd;
  if (bytes < sizeof(hd)) return fail(h, MASS_LYO_EINVAL, "not a checkpoint (too short)");
  std::memcpy(&hd, buf, sizeof(hd));
  if (std::memcmp(hd.magic, kCheckpointMagic, 8) != 0 || hd.version != 1)
    return fail(h, MASS_LYO_EINVAL, "not a checkpoint of this format");
  if ((int)hd.precision != h->prm.precision)
    return fail(h, MASS_LYO_EINVAL, "checkpoint precision differs from the run's");
  if (hd.LX != h->prm.LX || hd.LY != h->prm.LY)
    return fail(h, MASS_LYO_EINVAL,
                "checkpoint grid " + std::to_string(hd.LX) + "x" + std::to_string(hd.LY) +
                    " differs from the run's");
  if (hd.rho != h->prm.rho)
    return fail(h, MASS_LYO_EINVAL, "checkpoint rho differs from the run's");
  if (hd.bc != h->prm.bc)
    return fail(h, MASS_LYO_EINVAL, "checkpoint boundary conditions differ from the run's");
  if (hd.planes != h->eng->state_planes() || hd.state_bytes != h->eng->state_bytes() ||
      bytes < sizeof(hd) + hd.state_bytes)
    return fail(h, MASS_LYO_EINVAL, "checkpoint truncated or of another layout");
  const hipError_t e =
      h->eng->load_state(static_cast<const char*>(buf) + sizeof(hd), h->stream);
  if (e != hipSuccess) return hip_fail(h, e, "checkpoint load");
  h->has_state = true;
  return MASS_LYO_OK;
}

int mass_lyo_set_growth(mass_lyo_t h, const mass_lyo_growth* g) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (!g) return fail(h, MASS_LYO_EINVAL, "null growth parameters");
  if (h->prm.precision != MASS_LYO_FP64)
    return fail(h, MASS_LYO_EINVAL, "go-or-grow runs in FP64 only");
  if (h->sur.net) return fail(h, MASS_LYO_EINVAL, "disable the surrogate first");
  // go-or-grow.cpp:21-26
  if (g->chi_noise < 0) return fail(h, MASS_LYO_EINVAL, "chi-noise must be non-negative.");
  if (g->chi_length < 0) return fail(h, MASS_LYO_EINVAL, "chi-length must be non-negative.");
  if (g->relax_dt < 0) return fail(h, MASS_LYO_EINVAL, "relax-dt must be non-negative.");
  if (g->chi_config != 0 && g->chi_config != 1)
    return fail(h, MASS_LYO_EINVAL, "chi_config must be 0 (noise) or 1 (front)");
  // the counters take these as unsigned (lyotropic-with-division.cpp:41,79)
  if (!(g->division_time >= 0) || !(g->death_time >= 0) || g->division_time > 4294967295. ||
      g->death_time > 4294967295.)
    return fail(h, MASS_LYO_EINVAL, "division-time and death-time must be in [0, 2^32)");
  if (!h->growth && h->prm.bc == 0) {
    // the model runs on the ghost-layer engine, with periodic ghosts here
    std::unique_ptr<lyo::Engine> eng(lyo::make_engine_walls(0));
    hipError_t e = eng->alloc(h->prm.LX, h->prm.LY, h->NP, h->prm.npc);
    if (e != hipSuccess) return hip_fail(h, e, "go-or-grow buffers");
    eng->set_params(param_block(h->prm));
    h->eng->release();
    h->eng = std::move(eng);
  }
  const hipError_t e = h->eng->set_growth(*g, &h->gen);
  if (e != hipSuccess) return hip_fail(h, e, "go-or-grow buffers");
  h->growth = true;
  h->has_state = false;
  // periodic runs: fused tile sweeps; walled runs: the node loops
  h->launch.loops = h->prm.bc != 0 ? 1 : 0;
  return MASS_LYO_OK;
}

int mass_lyo_set_loop_kernels(mass_lyo_t h, int32_t on) {
  if (!h) return fail(nullptr, MASS_LYO_EINVAL, "null handle");
  if (on != 0 && on != 1) return fail(h, MASS_LYO_EINVAL, "on must be 0 or 1");
  if (on && h->prm.precision != MASS_LYO_FP64)
    return fail(h, MASS_LYO_EINVAL, "the loop kernels run the FP64 path only");
  if (!on && h->growth && h->prm.bc != 0)
    return fail(h, MASS_LYO_EINVAL, "walled go-or-grow runs the loop kernels only");
  if (on) {
    const hipError_t e = h->eng->alloc_loops();
    if (e != hipSuccess) return hip_fail(h, e, "loop kernel buffers");
  }
  h->launch.loops = on;
  return MASS_LYO_OK;
}

}  // extern "C"
