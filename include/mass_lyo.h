/*
 * mass_lyo.h -- C ABI of the MI355X (gfx950) lyotropic active-nematic step.
 *
 * The reference has no device boundary: its Lyotropic model runs the step in
 * OpenMP loops behind the C++ Model plugin interface (src/models.hpp:30-116).
 * This header is the thin layer between that interface and the HIP kernels.
 * Each entry point names the reference member it stands in for.
 *
 * Conventions
 *  - Node index k = y + LY*x (src/geometry.hpp:424-431); y is contiguous.
 *  - Population order v = 0..8 with directions (0,0),(1,0),(-1,0),(0,1),
 *    (0,-1),(1,1),(-1,-1),(-1,1),(1,-1) (src/geometry.hpp:17-27).
 *  - Host buffers are FP64.  Populations cross the boundary as AoS, N*9
 *    doubles, ff[9*k+v], the layout of the reference's LBField.
 *  - Every call returns 0 on success or a MASS_LYO_E* code; no exception
 *    crosses the boundary.  mass_lyo_last_error() gives the message.
 *  - One handle is used by one host thread.  Device work is queued on the
 *    handle's own HIP stream; calls that return data synchronise it.
 *  - Boundaries: bc 0 periodic on every path; bc 1-4 (channels and boxes on
 *    a ghost layer) in FP64, as fused tile sweeps or, on request, the
 *    reference's loop structure.  Node index and layout at the boundary are
 *    those of the domain either way.
 */
#ifndef MASS_LYO_H_
#define MASS_LYO_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MASS_LYO_ABI_VERSION 2

enum mass_lyo_status {
  MASS_LYO_OK = 0,
  MASS_LYO_EINVAL = 1,   /* bad argument or parameter */
  MASS_LYO_EDEVICE = 2,  /* HIP runtime error / no usable gfx950 device */
  MASS_LYO_ENAN = 3,     /* NaN found while downloading a field */
  MASS_LYO_ECHECK = 4,   /* conservation check failed */
  MASS_LYO_ENOMEM = 5
};

enum mass_lyo_precision {
  MASS_LYO_FP64 = 0,  /* IEEE FP64, reference expression order, no FMA:
                         bit-identical to the reference arithmetic */
  MASS_LYO_FP32 = 1   /* FP32 with shifted populations f - w_i*rho and a
                         compensated phi update */
};

/* Model parameters: the runcard options of Lyotropic::GetOptions
 * (src/models/lyotropic.cpp:551-602), plus the grid size from the factory
 * (src/models.hpp:190) and the precision of the device path. */
typedef struct mass_lyo_params {
  int32_t LX, LY;
  int32_t npc;          /* corrector sweeps, >= 1 */
  int32_t precision;    /* enum mass_lyo_precision */
  int32_t conserve_phi; /* must be 0: plain Lyotropic (lyotropic.hpp:46) */
  int32_t plane_pad;    /* extra elements between SoA planes, beyond the
                           library's own stagger of 256 (0 = none); results
                           do not depend on it */
  int32_t bc;           /* boundary conditions (Lyotropic(LX,LY,BC),
                           lyotropic.cpp:15-17,388-498): 0 periodic,
                           1 free-slip channel, 2 no-slip channel (walls at
                           y = -1/2 and LY-1/2, periodic in x), 3 free-slip
                           box, 4 no-slip box.  bc 1-4 keep a ghost layer
                           (geometry.cpp:140-183) and need FP64 (fused tile
                           sweeps, or the reference's loop structure:
                           mass_lyo_set_loop_kernels); batches, the
                           surrogate and the diagnostics need bc = 0 */
  double GammaQ, GammaP, xi, tauNem, tauIso, rho, friction;
  double AA, CC, LL, KK, zeta, zetaI, conc;
} mass_lyo_params;

/* Initial configuration (Lyotropic::ConfigureAtNode, lyotropic.cpp:69-160). */
enum mass_lyo_config {
  MASS_LYO_CONFIG_CIRCLE = 0,
  MASS_LYO_CONFIG_SQUARE = 1,
  MASS_LYO_CONFIG_STRIPE = 2,
  MASS_LYO_CONFIG_HALF = 3,
  MASS_LYO_CONFIG_RANDOM = 4,  /* nematic everywhere; phi = conc +
                                  U(-phinoise, phinoise), drawn after theta */
  MASS_LYO_CONFIG_BOXATWALL = 5,
  MASS_LYO_CONFIG_CUTTINGATWALL = 6,
  MASS_LYO_CONFIG_CIRCLEATWALL = 7,
  MASS_LYO_CONFIG_WETTEDWALL = 8
};

typedef struct mass_lyo_init {
  int32_t config;     /* enum mass_lyo_config */
  uint32_t seed;      /* mt19937 seed (src/random.cpp:43-46) */
  double level, radius, angle_deg, noise;
  double qamp;        /* Qxx = qamp*order*cos(2 theta); the reference is 1 */
  double phinoise;    /* half width of the phi draw (config RANDOM) */
} mass_lyo_init;

/* Field ids for mass_lyo_get_field.  FF is N*9 AoS; all others N doubles.
 * Derived fields (N, UX ... SIGMA_ACT_YX) are those of the last
 * UpdateQuantities call of the last step (lyotropic.cpp:168-247), i.e. the
 * start of the last corrector sweep, and are zero before the first step,
 * as in the reference. */
enum mass_lyo_field {
  MASS_LYO_FF = 0,
  MASS_LYO_QQXX = 1,
  MASS_LYO_QQYX = 2,
  MASS_LYO_PHI = 3,
  MASS_LYO_N = 4,
  MASS_LYO_UX = 5,
  MASS_LYO_UY = 6,
  MASS_LYO_SIGMAXX = 7,
  MASS_LYO_SIGMAXY = 8,
  MASS_LYO_SIGMAYX = 9,
  MASS_LYO_SIGMAYY = 10,
  MASS_LYO_SIGMA_BULK = 11,
  MASS_LYO_SIGMA_ELASTIC_XX = 12,
  MASS_LYO_SIGMA_ELASTIC_YX = 13,
  MASS_LYO_SIGMA_PHASE_FIELD_XX = 14,
  MASS_LYO_SIGMA_PHASE_FIELD_YX = 15,
  MASS_LYO_SIGMA_ACTIVE_XX = 16,
  MASS_LYO_SIGMA_ACTIVE_YX = 17,
  MASS_LYO_MU = 18,
  MASS_LYO_HHXX = 19,
  MASS_LYO_HHYX = 20,
  /* go-or-grow runs (mass_lyo_set_growth) also have: */
  MASS_LYO_M = 21,             /* phenotype density m */
  MASS_LYO_CHI = 22,           /* phenotype chi = m/phi */
  MASS_LYO_DIVISION_MASK = 23, /* 0/1 */
  MASS_LYO_DEATH_MASK = 24,    /* 0/1 */
  MASS_LYO_NUM_FIELDS = 25
};

/* go-or-grow parameters (go-or-grow.hpp:10-28, lyotropic-with-division.hpp:
 * 19-34; their runcard keys go-or-grow.cpp:387-423,
 * lyotropic-with-division.cpp:128-148). */
typedef struct mass_lyo_growth {
  int32_t chi_config;   /* 0 noise, 1 front */
  int32_t growTogether; /* m's growth source: 0 R, 1 chi*R */
  int32_t relax_phi, relax_Q;
  uint32_t relax_steps; /* dry free-energy relaxation before the dynamics */
  int32_t division_only; /* 1: lyotropic-with-division instead (Lyotropic + the
                            masks, phi += phi*(alpha*division - beta*death),
                            lyotropic-with-division.cpp:110-126; only alpha ...
                            phi_critical are read) */
  double B, Snem, Dchi, Achi, Ochi, phiswitch;
  double chi0, chi_noise, chi_length, relax_dt;
  double alpha, beta, division_rate, division_time, division_radius;
  double death_rate, death_time, death_radius, phi_critical;
} mass_lyo_growth;

typedef struct mass_lyo_s* mass_lyo_t;

/* Replaces the Lyotropic(LX,LY,BC) factory + Lyotropic::Initialize
 * (src/models.hpp:196-205, src/models/lyotropic.cpp:15-67): validates the
 * parameters, selects the device, allocates all device fields and the
 * handle's stream.  Fields start at zero (src/fields.hpp:29-33). */
int mass_lyo_create(const mass_lyo_params* params, mass_lyo_t* out);

/* Frees every device buffer, event and the stream. */
void mass_lyo_destroy(mass_lyo_t h);

/* Message of the last failed call on h (or of the last failed create when
 * h is NULL).  Valid until the next call on the same handle. */
const char* mass_lyo_last_error(mass_lyo_t h);

/* Replaces Lyotropic::Configure (src/models/lyotropic.cpp:69-166): draws the
 * initial fields on the host with mt19937 and the host libm, exactly as the
 * reference does, and uploads them.  Returns the reference's check totals
 * ftot (sum of f) and ptot (sum of phi) (lyotropic.cpp:158-159); either
 * pointer may be NULL. */
int mass_lyo_configure(mass_lyo_t h, const mass_lyo_init* init, double* ftot,
                       double* ptot);

/* Uploads an arbitrary state: ff (N*9 AoS), Qxx, Qyx, phi (N each).  Derived
 * fields are reset to zero.  Refused (MASS_LYO_EINVAL) on a go-or-grow run,
 * whose state also holds m, chi and the division/death masks. */
int mass_lyo_set_state(mass_lyo_t h, const double* ff, const double* Qxx,
                       const double* Qyx, const double* phi);

/* Replaces nsteps calls of Lyotropic::Step (src/models/lyotropic.cpp:500-525):
 * one predictor sweep (UpdateQuantities, UpdateFields(true), Move) and npc
 * corrector sweeps per step.  Asynchronous. */
int mass_lyo_step(mass_lyo_t h, int32_t nsteps);

/* Downloads one field as FP64 (the data of Lyotropic::serialize_frame,
 * src/models/lyotropic.hpp:121-138).  Synchronises.  Returns MASS_LYO_ENAN if
 * the field holds a NaN (the check of src/write.cpp:41). */
int mass_lyo_get_field(mass_lyo_t h, int32_t field, double* out);

/* FP64 device reductions used by Lyotropic::RuntimeChecks
 * (src/models/lyotropic.cpp:527-549): sum of all populations and sum of phi.
 * Fixed summation tree: the result depends only on the grid size.
 * Synchronises. */
int mass_lyo_sums(mass_lyo_t h, double* fsum, double* phisum);

/* The sums of mass_lyo_sums over a state held in the caller's device memory,
 * computed without allocating anything.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).
 *   precision   MASS_LYO_FP64: f (9 planes, populations) and phi are double;
 *               MASS_LYO_FP32: float, f holds the shifted populations
 *               f' = f - w_i rho (the library's FP32 layout), and
 *               sum f = sum f' + n rho; phic (may be NULL) is the
 *               compensation plane added to phi.  phic must be NULL in FP64.
 *   f           9 planes of plane_stride elements (plane v at f + v*stride)
 *   n           nodes per plane (1 <= n <= plane_stride <= (2^31-1)/9)
 *   out         device memory for 2 doubles: sum f, sum phi
 *   stream      a hipStream_t (NULL: the null stream); the call is
 *               asynchronous on it and enqueues two kernels, nothing else
 * The result is bitwise that of mass_lyo_sums on the same state (the same
 * fixed summation tree for the same n).  The partial sums live in one
 * 64 KiB device array per GPU, kept for the life of the process; calls that
 * overlap on one GPU (other threads, other streams) race on it.  Errors are
 * reported through mass_lyo_last_error(NULL). */
int mass_lyo_state_sums_noalloc(int32_t precision, const void* f, int64_t plane_stride,
                                const void* phi, const void* phic, int64_t n, double rho,
                                double* out, void* stream);

/* Run diagnostics of the current state, computed on the device with the
 * definitions of the reference's post-processing (POST/massPy):
 *   velocity u = sum_v c_v f_v / sum_v f_v (base_modules/flow.py:7-11);
 *   defects: director from Q (nematic/nematicPy.py:4-15), winding number over
 *   the 8-ring with each step folded into (-pi/2, pi/2] (base_modules/
 *   defects.py:15-50), |charge| > 0.4, 4-connected same-sign regions with
 *   periodic wrap counted once (defects.py:58-110, nematicPy.py:74-95). */
#define MASS_LYO_HIST_BINS 64
typedef struct mass_lyo_diag {
  double u_rms;           /* sqrt(mean(ux^2 + uy^2)) */
  double u_mean;          /* mean |u| */
  double phi_mean, phi_min, phi_max;
  double defect_density;  /* (defects_pos + defects_neg) / (LX*LY) */
  int64_t defects_pos;    /* +1/2 defects */
  int64_t defects_neg;    /* -1/2 defects */
  /* [0]: phi < 0 or NaN; [1..64]: phi in [(i-1)/64, i/64); [65]: phi >= 1 */
  uint64_t phi_hist[MASS_LYO_HIST_BINS + 2];
} mass_lyo_diag;

/* Fills *out; when charge is not NULL it also receives the charge field
 * (N doubles).  Synchronises. */
int mass_lyo_diagnostics(mass_lyo_t h, mass_lyo_diag* out, double* charge);

/* Waits for all queued device work of the handle. */
int mass_lyo_synchronize(mass_lyo_t h);

/* Timing: runs nsteps steps bracketed by HIP events on the handle's stream
 * and returns the elapsed milliseconds.  When sweep_ms is not NULL, every
 * sweep launch is also bracketed by its own event pair; sweep_ms (4 doubles)
 * gets the mean time of a predictor sweep, of a separate corrector sweep, of
 * a fused last-corrector-pair launch and of a one-kernel step (0 where there
 * was none).  In surrogate mode the steps are those of mass_lyo_step
 * (judged steps included) and sweep_ms gets zeros. */
int mass_lyo_time_steps(mass_lyo_t h, int32_t nsteps, double* total_ms,
                        double* sweep_ms);

/* Launch geometry of the sweep kernels (strip-march): rows per wave chunk,
 * waves per block (0 keeps the current value) and XCD-aware block order
 * (1 on, 0 off, -1 keep).  Results do not depend on it. */
int mass_lyo_set_launch(mass_lyo_t h, int32_t chunk_rows, int32_t waves_per_block,
                        int32_t xcd_order);

/* Sweep schedule.  fuse_last_pair: the last two corrector sweeps (npc >= 2,
 * LX >= 8) in one kernel that keeps the first one's outputs in registers, 1
 * on (the default) or 0 off (separate sweeps); pair_chunk_rows: rows per
 * wave chunk of that kernel (0 keeps); alternate: neighbouring chunks march
 * in opposite directions so that they reach their shared rows together, 1 on
 * (the default) or 0 off (every chunk upwards); one_kernel_step: with npc = 1,
 * or npc = 2 and the pair fused (LX >= 8), the whole step -- predictor and
 * correctors -- runs in one kernel, 1 on (the default) or 0 off; 2 (npc = 2):
 * the block-cooperative form of that kernel, whose four waves per block hand
 * their edge values to each other through LDS so that only the block's outer
 * columns are halo; step_chunk_rows: rows per chunk of that kernel (0 keeps).  -1 keeps a setting.  Results
 * are bitwise the same either way; derived fields of a frame are recomputed on
 * demand after a fused step.  No reference counterpart (Lyotropic::Step,
 * src/models/lyotropic.cpp:500-525, runs its sweeps one after another). */
int mass_lyo_set_schedule(mass_lyo_t h, int32_t fuse_last_pair, int32_t pair_chunk_rows,
                          int32_t alternate, int32_t one_kernel_step,
                          int32_t step_chunk_rows);

/* Binary checkpoint of a run's state, for restarting long runs exactly.  No
 * reference counterpart: the reference writes text frames only
 * (src/serialization.hpp:280-286), whose six digits cannot resume a run.  The
 * image is a 64-byte header (magic "MASSLYO1", format version 1, precision,
 * LX, LY, npc, number of planes, rho, byte count of the planes; little-endian
 * host) followed by the raw device state as planes of LX*LY values (node
 * k = y + LY*x): FP64 f (9 planes, directions of src/geometry.hpp:17-27), Qxx,
 * Qyx, phi; FP32 the shifted populations f - w_i*rho (9), Qxx, Qyx, phi and
 * phi's compensation term.  Loading needs the same precision, LX, LY and rho
 * (npc may differ); a run continued from a checkpoint is bitwise the run that
 * was not interrupted.  After a load the derived frame fields are zero until
 * the next step, as after mass_lyo_set_state.  size: bytes of the image;
 * save: bytes >= size; load: the image as saved. */
int mass_lyo_checkpoint_size(mass_lyo_t h, uint64_t* bytes);
int mass_lyo_checkpoint_save(mass_lyo_t h, void* buf, uint64_t bytes);
int mass_lyo_checkpoint_load(mass_lyo_t h, const void* buf, uint64_t bytes);

/* FP64 only: 1 runs every step as the reference's node loops, one kernel per
 * loop -- UpdateQuantities (storing its 23 derived fields), UpdateFields(true),
 * Move, then npc x (UpdateQuantities, UpdateFields(false)) -- and 0 (the
 * default) returns to the fused kernels.  Results are bitwise the same; the
 * loop kernels move the reference's ~291 values per node and step instead of
 * the fused kernels' 26-108, and exist to measure that layout.  On walled
 * runs (bc 1-4) the fused kernels are the ghost-layer tile sweeps with Move
 * folded into the correctors' reads; periodic go-or-grow and
 * lyotropic-with-division runs fuse each sweep into one tile kernel, walled
 * ones always use the loops and refuse 0.  The first call with 1 allocates
 * the 41 extra planes.  sweep_ms of mass_lyo_time_steps
 * gets zeros in this mode.  Reference: Lyotropic::Step,
 * src/models/lyotropic.cpp:343-386,500-525. */
int mass_lyo_set_loop_kernels(mass_lyo_t h, int32_t on);

/* Makes h a go-or-grow run: replaces the GoOrGrow(LX,LY,BC) factory +
 * GoOrGrow::Initialize (go-or-grow.cpp:13-32, lyotropic-with-division.cpp:
 * 14-32); with division_only, a lyotropic-with-division run
 * (LyotropicWithDivision, lyotropic-with-division.cpp:14-156: no m, chi,
 * relaxation or phenotype).  Call after mass_lyo_create and before mass_lyo_configure; FP64
 * only, any bc.  mass_lyo_configure then also runs Configure's dry
 * relaxation, hydrodynamic reset and phenotype set-up (go-or-grow.cpp:34-40)
 * and returns ftot of the reset and ptot of the relaxed phi; every
 * mass_lyo_step first redraws the division/death masks when due
 * (lyotropic-with-division.cpp:34-108, 150-156) from the run's mt19937 stream
 * on the host, reading phi back from the device for that.  phi is conserved
 * (conserve_phi, lyotropic.cpp:297-298) when neither division nor death can
 * happen; its sum is then the device's fixed-tree sum, not the reference's
 * serial one, so such runs match the reference to rounding, not bitwise.
 * The run uses the ghost-layer grid: on a periodic one each sweep is one
 * fused tile kernel (the reference's loop structure on request), on a walled
 * one the reference's loop structure; the diagnostics, surrogate, batches and
 * checkpoints are not offered for it. */
int mass_lyo_set_growth(mass_lyo_t h, const mass_lyo_growth* g);

/* ------------------------------------------------------------------ batches
 * Many independent runs (own size and parameters; shared npc and precision)
 * advanced together.  No reference counterpart: the reference runs one model
 * per process; a batch is what a parameter sweep over its runcards becomes on
 * one GPU.  Each sweep of a step is one persistent launch over the work items
 * of all runs, and two consecutive steps are replayed from a HIP graph.  Runs
 * are addressed by index 0..nruns-1; per-run calls mirror the single-run ABI. */
typedef struct mass_lyo_batch_s* mass_lyo_batch_t;

int mass_lyo_batch_create(const mass_lyo_params* runs, int32_t nruns, mass_lyo_batch_t* out);
void mass_lyo_batch_destroy(mass_lyo_batch_t b);
const char* mass_lyo_batch_last_error(mass_lyo_batch_t b);
int mass_lyo_batch_size(mass_lyo_batch_t b);
int mass_lyo_batch_configure(mass_lyo_batch_t b, int32_t run, const mass_lyo_init* init,
                             double* ftot, double* ptot);
int mass_lyo_batch_set_state(mass_lyo_batch_t b, int32_t run, const double* ff,
                             const double* Qxx, const double* Qyx, const double* phi);
/* nsteps steps of every run (asynchronous) */
int mass_lyo_batch_step(mass_lyo_batch_t b, int32_t nsteps);
int mass_lyo_batch_get_field(mass_lyo_batch_t b, int32_t run, int32_t field, double* out);
int mass_lyo_batch_sums(mass_lyo_batch_t b, int32_t run, double* fsum, double* phisum);
int mass_lyo_batch_diagnostics(mass_lyo_batch_t b, int32_t run, mass_lyo_diag* out,
                               double* charge);
/* chunk rows of the sweep work items and the persistent grid size in waves
 * (0: one wave per item).  Results do not depend on them. */
int mass_lyo_batch_set_launch(mass_lyo_batch_t b, int32_t chunk_rows, int32_t grid_waves);
/* Times candidate launch shapes on the GPU and keeps the fastest.  It steps
 * the zeroed fields of a new batch, so it must be called before any run is
 * configured; the chosen shape is returned. */
int mass_lyo_batch_autotune(mass_lyo_batch_t b, int32_t* chunk_rows, int32_t* grid_waves);
int mass_lyo_batch_time_steps(mass_lyo_batch_t b, int32_t nsteps, double* total_ms);
/* The run handle of member `run` (owned by the batch; valid until the batch
 * is destroyed), after waiting for the batch's queued steps.  Stepping it on
 * its own desynchronises it from the batch and is not allowed; reading,
 * diagnostics and collecting training frames are. */
mass_lyo_t mass_lyo_batch_run(mass_lyo_batch_t b, int32_t run);

/* ------------------------------------------------- surrogate corrector
 * A small convolutional network that stands in for the corrector sweeps after
 * the first.  No reference counterpart (the reference always runs npc
 * correctors, lyotropic.cpp:517-524).  Four periodic 3x3 convolutions
 * 8 -> 32 -> 32 -> 32 -> 8 channels, tanh-GELU, BF16 weights and activations,
 * FP32 accumulation on MFMA.  Inputs per node: ux, uy, Qxx, Qyx, phi, mu and
 * the force div(sigma) (x, y) of the one-corrector state; outputs: the
 * increments of ux, uy, Qxx, Qyx, phi that the remaining correctors add.
 * Parameters: for each layer W[9][cout][cin] (tap t = 3*(dx+1) + (dy+1)) then
 * b[cout], FP32. */
typedef struct mass_lyo_net_s* mass_lyo_net_t;

int mass_lyo_net_create(uint32_t seed, mass_lyo_net_t* out);
void mass_lyo_net_destroy(mass_lyo_net_t n);
const char* mass_lyo_net_last_error(mass_lyo_net_t n);
int mass_lyo_net_num_params(mass_lyo_net_t n);
int mass_lyo_net_get_params(mass_lyo_net_t n, float* out);
int mass_lyo_net_set_params(mass_lyo_net_t n, const float* in);
/* normalisation: mean[8], 1/std[8] of the inputs, scale[8], 1/scale[8] of
 * the outputs */
int mass_lyo_net_normalization(mass_lyo_net_t n, float* out32);
/* network on host data already in network units: x, y are [LX*LY][8] FP32
 * (node k = y + LY*x), LY a multiple of 16 */
int mass_lyo_net_forward_host(mass_lyo_net_t n, const float* x, int32_t LX, int32_t LY, float* y);
/* the surrogate's inference on host data in physical units: x [LX*LY][8]
 * features are normalised with the network's statistics, the output is
 * scaled back to increments (as surrogate mode applies them) */
int mass_lyo_net_predict_host(mass_lyo_net_t n, const float* x, int32_t LX, int32_t LY, float* y);
/* loss sum((out - y)^2)/(2N) and its gradient wrt the parameters */
int mass_lyo_net_grad_host(mass_lyo_net_t n, const float* x, const float* y, int32_t LX,
                           int32_t LY, double* loss, float* grads);
/* appends one training frame taken from run h, whose last step must be a
 * plain step with npc >= 2: the inputs of its one-corrector state and the
 * increments its remaining correctors made */
int mass_lyo_net_collect(mass_lyo_net_t n, mass_lyo_t h);
int mass_lyo_net_dataset_size(mass_lyo_net_t n);
/* Adam on the GPU over the collected frames (one frame per iteration) */
int mass_lyo_net_train(mass_lyo_net_t n, int32_t iters, float lr, double* loss_first,
                       double* loss_last);

typedef struct mass_lyo_surrogate_stats {
  int64_t steps;            /* steps taken since enabling */
  int64_t surrogate_steps;  /* predictor + 1 corrector + network */
  int64_t plain_steps;      /* predictor + npc correctors */
  int64_t judged;           /* plain steps that also measured the error */
  int64_t fallbacks;        /* judgements above tolerance */
  double last_error;        /* ||u_sur - u_plain|| / ||u_plain|| of the last judgement */
  double max_error;
} mass_lyo_surrogate_stats;

/* From now on mass_lyo_step on h uses the network (h needs npc >= 2 and LY a
 * multiple of 16; the network must be trained).  Every judge_every-th step is
 * a plain step that also measures the velocity error of the surrogate; an
 * error above tol makes the next judge_every - 1 steps plain.  The network
 * must outlive its use by h. */
int mass_lyo_surrogate_enable(mass_lyo_t h, mass_lyo_net_t net, int32_t judge_every, double tol);
int mass_lyo_surrogate_disable(mass_lyo_t h);
int mass_lyo_surrogate_get_stats(mass_lyo_t h, mass_lyo_surrogate_stats* out);
/* Kernels of the network part of a surrogate step: 1 (the default for FP32
 * runs with LX, LY >= 16) one kernel that computes the features, runs the
 * four layers with the activations on chip and applies the increment; 0
 * separate feature, convolution and increment kernels (the judged steps
 * always use these).  Both apply the same increments up to the BF16 rounding
 * of the activations.  No reference counterpart. */
int mass_lyo_surrogate_set_kernels(mass_lyo_t h, int32_t fused);
/* Output calibration (on = 1; off by default; enabling or disabling resets
 * it): every judged step also fits, per output c (du_x, du_y, dQxx, dQyx,
 * dphi), the factor a_c that best maps the network's increments p onto the
 * plain corrector's increments t in least squares, over the judged steps
 * with weight 0.9^age (a_c = sum t.p / sum p.p, kept in [0.25, 4]); the
 * following surrogate steps apply a_c p.  No reference counterpart. */
int mass_lyo_surrogate_set_calibration(mass_lyo_t h, int32_t on);
/* the current factors a_c (five values; 1 while calibration is off) */
int mass_lyo_surrogate_get_calibration(mass_lyo_t h, double* factors5);

/* Version of this ABI (MASS_LYO_ABI_VERSION). */
int mass_lyo_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
