// lyo_batch.cpp -- batched runs (C ABI, include/mass_lyo.h).
//
// A batch holds many independent lyotropic runs (any sizes, own parameters)
// that advance together: each stage of a step's launch plan (the one-kernel
// step, or predictor + fused pair, or separate sweeps) is ONE persistent
// launch over the work items of all runs (full/pair/sweep_batch_kernel), and
// two consecutive steps (both state parities) are captured once into a HIP
// graph and replayed.
// Per-run operations (configure, fields, sums, diagnostics) reuse the single
// run handles; they first wait for the batch stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "lyo_handle.hpp"

struct mass_lyo_batch_s {
  std::vector<mass_lyo_t> runs;
  hipStream_t stream = nullptr;
  int npc = 1, precision = 0;
  lyo::LaunchCfg launch;
  int grid_waves = 0;      // persistent grid size in waves (0: one wave per item)
  std::vector<int> plan;   // stages of one step (lyo::step_plan)
  bool store_half = false;
  std::vector<char> jobs;  // [parity][stage][run] SweepJob bytes
  void* d_jobs = nullptr;
  size_t d_jobs_bytes = 0;
  size_t job_bytes = 0;
  std::vector<int> nitems;  // work items per stage
  bool jobs_valid = false;
  hipGraphExec_t graph = nullptr;
  std::string err;
};

namespace {
thread_local std::string g_batch_create_error;

int bfail(mass_lyo_batch_t b, int code, const std::string& msg) {
  if (b)
    b->err = msg;
  else
    g_batch_create_error = msg;
  return code;
}
int bhip(mass_lyo_batch_t b, hipError_t e, const char* what) {
  return bfail(b, MASS_LYO_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
int from_run(mass_lyo_batch_t b, int rc, mass_lyo_t r) {
  if (rc != MASS_LYO_OK) b->err = mass_lyo_last_error(r);
  return rc;
}

void drop_graph(mass_lyo_batch_t b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
}

// (re)builds and uploads the job table for the current launch configuration
// A run's launch: the batch's chunk lengths evened out over the run's rows
// (LX = 192 with 80-row chunks: three chunks of 64 rows, not 80 + 80 + 32),
// so every work item costs about the same and no short item pays the rows a
// chunk repeats for little work.
static lyo::LaunchCfg run_launch(const lyo::LaunchCfg& L, int LX) {
  lyo::LaunchCfg r = L;
  auto even = [LX](int xc) {
    if (xc < 1) return xc;
    const int n = (LX + xc - 1) / xc;
    return (LX + n - 1) / n;
  };
  r.xc = even(L.xc);
  r.xc2 = even(L.xc2);
  r.xc3 = even(L.xc3);
  return r;
}

int build_jobs(mass_lyo_batch_t b) {
  if (b->jobs_valid) return MASS_LYO_OK;
  drop_graph(b);
  const int R = (int)b->runs.size();
  int lxmin = b->runs[0]->prm.LX;
  for (auto* h : b->runs) lxmin = std::min(lxmin, (int)h->prm.LX);
  b->plan = b->runs[0]->eng->plan(b->launch, lxmin);
  b->store_half = b->precision == MASS_LYO_FP32 && lyo::plan_stores_half(b->plan, b->npc);
  const int S = (int)b->plan.size();
  b->job_bytes = b->runs[0]->eng->job_bytes();
  b->jobs.assign(2 * (size_t)S * R * b->job_bytes, 0);
  b->nitems.assign(S, 0);
  std::vector<int> first(R);
  for (int c = 0; c < S; ++c) {
    int items = 0;
    for (int r = 0; r < R; ++r) {
      first[r] = items;
      items += b->runs[r]->eng->job_items(b->plan[c], run_launch(b->launch, b->runs[r]->prm.LX));
    }
    b->nitems[c] = items;
    for (int p = 0; p < 2; ++p)
      for (int r = 0; r < R; ++r)
        b->runs[r]->eng->fill_job(b->plan[c], p, first[r],
                                  run_launch(b->launch, b->runs[r]->prm.LX),
                                  &b->jobs[(((size_t)p * S + c) * R + r) * b->job_bytes]);
  }
  hipError_t e = hipSuccess;
  if (b->d_jobs_bytes < b->jobs.size()) {
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    b->d_jobs = nullptr;
    e = hipMalloc(&b->d_jobs, b->jobs.size());
    b->d_jobs_bytes = e == hipSuccess ? b->jobs.size() : 0;
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_jobs, b->jobs.data(), b->jobs.size(), hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return bhip(b, e, "batch job table");
  b->jobs_valid = true;
  return MASS_LYO_OK;
}

hipError_t launch_step(mass_lyo_batch_t b, int parity) {
  const int R = (int)b->runs.size(), S = (int)b->plan.size();
  for (int c = 0; c < S; ++c) {
    const char* j = static_cast<const char*>(b->d_jobs) + ((size_t)parity * S + c) * R * b->job_bytes;
    hipError_t e = b->runs[0]->eng->launch_batch(j, R, b->nitems[c], b->plan[c], b->store_half,
                                                 b->grid_waves, b->launch.wpb, b->stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

int sync_batch(mass_lyo_batch_t b) {
  hipError_t e = hipStreamSynchronize(b->stream);
  return e == hipSuccess ? MASS_LYO_OK : bhip(b, e, "batch stream");
}

int check_run(mass_lyo_batch_t b, int32_t run) {
  if (run < 0 || run >= (int)b->runs.size()) return bfail(b, MASS_LYO_EINVAL, "run index out of range");
  return sync_batch(b);
}
}  // namespace

extern "C" {

const char* mass_lyo_batch_last_error(mass_lyo_batch_t b) {
  return b ? b->err.c_str() : g_batch_create_error.c_str();
}

int mass_lyo_batch_create(const mass_lyo_params* runs, int32_t nruns, mass_lyo_batch_t* out) {
  if (!runs || !out || nruns < 1) return bfail(nullptr, MASS_LYO_EINVAL, "need at least one run");
  *out = nullptr;
  for (int i = 1; i < nruns; ++i)
    if (runs[i].npc != runs[0].npc || runs[i].precision != runs[0].precision)
      return bfail(nullptr, MASS_LYO_EINVAL, "all runs of a batch share npc and precision");
  for (int i = 0; i < nruns; ++i)
    if (runs[i].bc != 0)
      return bfail(nullptr, MASS_LYO_EINVAL, "batched runs are periodic (bc 0)");
  auto* b = new mass_lyo_batch_s();
  b->npc = runs[0].npc;
  b->precision = runs[0].precision;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete b;
    return bhip(nullptr, e, "hipStreamCreate");
  }
  for (int i = 0; i < nruns; ++i) {
    mass_lyo_t h = nullptr;
    const int rc = mass_lyo_create(&runs[i], &h);
    if (rc != MASS_LYO_OK) {
      const std::string msg = std::string("run ") + std::to_string(i) + ": " + mass_lyo_last_error(nullptr);
      mass_lyo_batch_destroy(b);
      return bfail(nullptr, rc, msg);
    }
    b->runs.push_back(h);
  }
  // defaults (autotune picks better ones): one wave per work item
  b->launch.xc = 8;
  b->launch.xc2 = 16;
  b->launch.xc3 = 32;
  b->launch.wpb = 4;
  b->launch.xcd = 0;
  *out = b;
  return MASS_LYO_OK;
}

void mass_lyo_batch_destroy(mass_lyo_batch_t b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (auto* h : b->runs) mass_lyo_destroy(h);
  if (b->d_jobs) (void)hipFree(b->d_jobs);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int mass_lyo_batch_size(mass_lyo_batch_t b) { return b ? (int)b->runs.size() : 0; }

int mass_lyo_batch_configure(mass_lyo_batch_t b, int32_t run, const mass_lyo_init* init,
                             double* ftot, double* ptot) {
  if (!b) return bfail(nullptr, MASS_LYO_EINVAL, "null batch");
  int rc = check_run(b, run);
  if (rc) return rc;
  return from_run(b, mass_lyo_configure(b->runs[run], init, ftot, ptot), b->runs[run]);
}

int mass_lyo_batch_set_state(mass_lyo_batch_t b, int32_t run, const double* ff, const double* Qxx,
                             const double* Qyx, const double* phi) {
  if (!b) return bfail(nullptr, MASS_LYO_EINVAL, "null batch");
  int rc = check_run(b, run);
  if (rc) return rc;
  return from_run(b, mass_lyo_set_state(b->runs[run], ff, Qxx, Qyx, phi), b->runs[run]);
}

int mass_lyo_batch_get_field(mass_lyo_batch_t b, int32_t run, int32_t field, double* out) {
  if (!b) return bfail(nullptr, MASS_LYO_EINVAL, "null batch");
  int rc = check_run(b, run);
  if (rc) return rc;
  return from_run(b, mass_lyo_get_field(b->runs[run], field, out), b->runs[run]);
}

int mass_lyo_batch_sums(mass_lyo_batch_t b, int32_t run, double* fsum, double* phisum) {
  if (!b) return bfail(nullptr, MASS_LYO_EINVAL, "null batch");
  int rc = check_run(b, run);
  if (rc) return rc;
  return from_run(b, mass_lyo_sums(b->runs[run], fsum, phisum), b->runs[run]);
}

int mass_lyo_batch_diagnostics(mass_lyo_batch_t b, int32_t run, mass_lyo_diag* out, double* charge) {
  if (!b) return bfail(nullptr, MASS_LYO_EINVAL, "null batch");
  int rc = check_run(b, run);
  if (rc) return rc;
  return from_run(b, mass_lyo_diagnostics(b->runs[run], out, charge), b->runs[run]);
}

int mass_lyo_batch_set_launch(mass_lyo_batch_t b, int32_t chunk_rows, int32_t grid_waves) {
  if (!b) return bfail(nullptr, MASS_LYO_EINVAL, "null batch");
  if (chunk_rows < 0 || chunk_rows > 4096 || grid_waves < 0)
    return bfail(b, MASS_LYO_EINVAL, "chunk_rows must be in [1,4096] and grid_waves >= 0");
  int rc = sync_batch(b);
  if (rc) return rc;
  if (chunk_rows) b->launch.xc = b->launch.xc2 = b->launch.xc3 = chunk_rows;
  b->grid_waves = grid_waves;
  b->jobs_valid = false;
  return MASS_LYO_OK;
}

int mass_lyo_batch_step(mass_lyo_batch_t b, int32_t nsteps) {
  if (!b) return bfail(nullptr, MASS_LYO_EINVAL, "null batch");
  if (nsteps < 0) return bfail(b, MASS_LYO_EINVAL, "negative step count");
  if (nsteps == 0) return MASS_LYO_OK;
  int rc = build_jobs(b);
  if (rc) return rc;
  hipError_t e = hipSuccess;
  int parity = b->runs[0]->eng->parity();
  int left = nsteps;
  if (parity == 1) {  // the graph starts at parity 0
    e = launch_step(b, 1);
    parity = 0;
    --left;
  }
  if (e == hipSuccess && left >= 2 && !b->graph) {
    hipGraph_t g = nullptr;
    e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipError_t e1 = launch_step(b, 0);
      if (e1 == hipSuccess) e1 = launch_step(b, 1);
      e = hipStreamEndCapture(b->stream, &g);
      if (e == hipSuccess) e = e1;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
  }
  for (; e == hipSuccess && left >= 2; left -= 2) e = hipGraphLaunch(b->graph, b->stream);
  if (e == hipSuccess && left == 1) e = launch_step(b, 0);
  if (e != hipSuccess) return bhip(b, e, "batch step");
  for (auto* h : b->runs) h->eng->after_batch_steps(nsteps, b->plan);
  return MASS_LYO_OK;
}

int mass_lyo_batch_time_steps(mass_lyo_batch_t b, int32_t nsteps, double* total_ms) {
  if (!b) return bfail(nullptr, MASS_LYO_EINVAL, "null batch");
  int rc = build_jobs(b);
  if (rc) return rc;
  hipEvent_t t0, t1;
  hipError_t e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  if (e != hipSuccess) return bhip(b, e, "hipEventCreate");
  e = hipEventRecord(t0, b->stream);
  if (e == hipSuccess) {
    rc = mass_lyo_batch_step(b, nsteps);
    if (rc) e = hipErrorUnknown;
  }
  if (e == hipSuccess) e = hipEventRecord(t1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(t1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (rc) return rc;
  if (e != hipSuccess) return bhip(b, e, "batch timing");
  if (total_ms) *total_ms = ms;
  return MASS_LYO_OK;
}

mass_lyo_t mass_lyo_batch_run(mass_lyo_batch_t b, int32_t run) {
  if (!b || run < 0 || run >= (int)b->runs.size()) return nullptr;
  if (sync_batch(b) != MASS_LYO_OK) return nullptr;
  return b->runs[run];
}

int mass_lyo_batch_autotune(mass_lyo_batch_t b, int32_t* chunk_rows, int32_t* grid_waves) {
  if (!b) return bfail(nullptr, MASS_LYO_EINVAL, "null batch");
  for (auto* h : b->runs)
    if (h->has_state)
      return bfail(b, MASS_LYO_EINVAL,
                   "autotune runs on the zeroed state of a new batch: call it before configure");
  int rc0 = build_jobs(b);
  if (rc0) return rc0;
  // chunk rows per stage kind: the one-kernel step repeats eight rows per
  // chunk (four with one corrector), the sweeps two
  const bool full = b->plan.size() == 1 && lyo::is_full_stage(b->plan[0]);
  // (the one-kernel step: a fine grid, since the repeated rows favour long
  // chunks and the balance of a few thousand items over the wave slots short
  // ones; measured on the 64-run sweep: 80 rows beat 24-48 by 5%)
  const std::vector<int> xcs = full ? std::vector<int>{16, 24, 32, 40, 48, 56, 64, 72, 80, 88,
                                                      96, 112, 128, 160}
                                    : std::vector<int>{4, 8, 16, 32};
  const int grids[] = {0, 2048, 4096, 8192, 16384};
  double best = 1e300;
  int bx = b->launch.xc, bg = b->grid_waves;
  for (int xc : xcs)
    for (int gw : grids) {
      int rc = mass_lyo_batch_set_launch(b, xc, gw);
      if (rc == MASS_LYO_OK) rc = mass_lyo_batch_step(b, 2);  // warm-up, builds the graph
      double ms = 0;
      if (rc == MASS_LYO_OK) rc = mass_lyo_batch_time_steps(b, 6, &ms);
      if (rc) return rc;
      if (ms < best) {
        best = ms;
        bx = xc;
        bg = gw;
      }
    }
  int rc = mass_lyo_batch_set_launch(b, bx, bg);
  if (rc) return rc;
  if (chunk_rows) *chunk_rows = bx;
  if (grid_waves) *grid_waves = bg;
  return MASS_LYO_OK;
}

}  // extern "C"
