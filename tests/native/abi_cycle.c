/* Exercises every kind of handle of the C ABI (include/mass_lyo.h) in
 * create -> use -> destroy cycles, for running under AddressSanitizer and
 * LeakSanitizer: the host side of the library must neither touch memory it
 * does not own nor keep any allocation once every handle is destroyed.
 *
 *   abi_cycle <cycles>      GPU cycles (single runs in every mode, batches,
 *                           the network and surrogate mode), then the error
 *                           paths
 *   abi_cycle --no-device   the error paths only; every create must fail
 *                           with MASS_LYO_EDEVICE (a machine without a GPU)
 *   abi_cycle --no-device --leak
 *                           the same, then drops a buffer on purpose: the
 *                           leak checker must report it (is it live?)
 *
 * Prints one line per finished part; exits 0 when every call returned what
 * it should (sanitizer reports make the process exit non-zero by
 * themselves).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mass_lyo.h"

static int failures = 0;

#define CHECK(call, want)                                                      \
  do {                                                                         \
    int rc_ = (call);                                                          \
    if (rc_ != (want)) {                                                       \
      fprintf(stderr, "%s:%d: %s returned %d, expected %d\n", __FILE__,       \
              __LINE__, #call, rc_, (int)(want));                              \
      failures++;                                                              \
    }                                                                          \
  } while (0)

static mass_lyo_params phys(int LX, int LY, int npc, int precision, int bc) {
  mass_lyo_params p;
  memset(&p, 0, sizeof p);
  p.LX = LX;
  p.LY = LY;
  p.npc = npc;
  p.precision = precision;
  p.bc = bc;
  p.GammaQ = 0.1;
  p.GammaP = 0.1;
  p.xi = 0.7;
  p.tauNem = 2.0;
  p.tauIso = 2.0;
  p.rho = 40.0;
  p.AA = 0.01;
  p.LL = 0.04;
  p.KK = 0.04;
  p.zeta = 0.02;
  p.conc = 0.5;
  return p;
}

static mass_lyo_init random_init(unsigned seed) {
  mass_lyo_init in;
  memset(&in, 0, sizeof in);
  in.config = MASS_LYO_CONFIG_RANDOM;
  in.seed = seed;
  in.angle_deg = 90.0;
  in.noise = 1.0;
  in.qamp = 0.5;
  in.phinoise = 0.05;
  return in;
}

static mass_lyo_t make(int LX, int LY, int npc, int precision, int bc) {
  mass_lyo_params p = phys(LX, LY, npc, precision, bc);
  mass_lyo_t h = NULL;
  CHECK(mass_lyo_create(&p, &h), MASS_LYO_OK);
  return h;
}

static void configure(mass_lyo_t h, unsigned seed) {
  mass_lyo_init in = random_init(seed);
  double ftot = 0, ptot = 0;
  CHECK(mass_lyo_configure(h, &in, &ftot, &ptot), MASS_LYO_OK);
}

/* every field and diagnostic of a handle, into freshly allocated buffers */
static void read_out(mass_lyo_t h, int N, int nfields) {
  double* buf = (double*)malloc(sizeof(double) * 9 * (size_t)N);
  for (int f = 0; f < nfields; ++f) CHECK(mass_lyo_get_field(h, f, buf), MASS_LYO_OK);
  double fs, ps;
  CHECK(mass_lyo_sums(h, &fs, &ps), MASS_LYO_OK);
  free(buf);
}

static void single_runs(void) {
  const int L = 128, N = L * L;
  /* FP32 run: launch settings, timing, diagnostics, frame fields, checkpoint */
  mass_lyo_t h = make(L, L, 2, MASS_LYO_FP32, 0);
  configure(h, 1);
  CHECK(mass_lyo_step(h, 3), MASS_LYO_OK);
  double ms, sweep[4];
  CHECK(mass_lyo_time_steps(h, 2, &ms, sweep), MASS_LYO_OK);
  mass_lyo_diag d;
  double* charge = (double*)malloc(sizeof(double) * N);
  CHECK(mass_lyo_diagnostics(h, &d, charge), MASS_LYO_OK);
  free(charge);
  read_out(h, N, MASS_LYO_M);
  uint64_t bytes = 0;
  CHECK(mass_lyo_checkpoint_size(h, &bytes), MASS_LYO_OK);
  void* img = malloc(bytes);
  CHECK(mass_lyo_checkpoint_save(h, img, bytes), MASS_LYO_OK);
  CHECK(mass_lyo_step(h, 1), MASS_LYO_OK);
  CHECK(mass_lyo_checkpoint_load(h, img, bytes), MASS_LYO_OK);
  free(img);
  mass_lyo_destroy(h);

  /* FP64 run, fused kernels then the reference's loop structure; state import */
  h = make(L, L, 2, MASS_LYO_FP64, 0);
  configure(h, 2);
  CHECK(mass_lyo_step(h, 2), MASS_LYO_OK);
  CHECK(mass_lyo_set_loop_kernels(h, 1), MASS_LYO_OK);
  CHECK(mass_lyo_step(h, 2), MASS_LYO_OK);
  double* ff = (double*)malloc(sizeof(double) * 9 * (size_t)N);
  double* q = (double*)malloc(sizeof(double) * 3 * (size_t)N);
  CHECK(mass_lyo_get_field(h, MASS_LYO_FF, ff), MASS_LYO_OK);
  CHECK(mass_lyo_get_field(h, MASS_LYO_QQXX, q), MASS_LYO_OK);
  CHECK(mass_lyo_get_field(h, MASS_LYO_QQYX, q + N), MASS_LYO_OK);
  CHECK(mass_lyo_get_field(h, MASS_LYO_PHI, q + 2 * N), MASS_LYO_OK);
  CHECK(mass_lyo_set_state(h, ff, q, q + N, q + 2 * N), MASS_LYO_OK);
  CHECK(mass_lyo_step(h, 1), MASS_LYO_OK);
  read_out(h, N, MASS_LYO_M);
  free(ff);
  free(q);
  mass_lyo_destroy(h);

  /* walled FP64 run (no-slip box) */
  h = make(L, 64, 1, MASS_LYO_FP64, 4);
  configure(h, 3);
  CHECK(mass_lyo_step(h, 2), MASS_LYO_OK);
  read_out(h, L * 64, MASS_LYO_M);
  mass_lyo_destroy(h);

  /* go-or-grow and lyotropic-with-division runs */
  for (int division_only = 0; division_only < 2; ++division_only) {
    h = make(64, 64, 1, MASS_LYO_FP64, 0);
    mass_lyo_growth g;
    memset(&g, 0, sizeof g);
    g.relax_phi = 1;
    g.relax_Q = 1;
    g.relax_steps = 5;
    g.division_only = division_only;
    g.B = 0.1;
    g.Snem = 1.0;
    g.chi0 = 1.0;
    g.relax_dt = 1.0;
    g.alpha = 0.001;
    g.division_rate = 0.1;
    g.division_time = 2.0;
    g.division_radius = 3.0;
    g.death_time = 1.0;
    g.phi_critical = 1.2;
    CHECK(mass_lyo_set_growth(h, &g), MASS_LYO_OK);
    configure(h, 4);
    CHECK(mass_lyo_step(h, 3), MASS_LYO_OK);
    read_out(h, 64 * 64, division_only ? MASS_LYO_M : MASS_LYO_NUM_FIELDS);
    mass_lyo_destroy(h);
  }
  printf("single runs ok\n");
  fflush(stdout);
}

static void batches(void) {
  const int sizes[3] = {64, 96, 128};
  mass_lyo_params runs[3];
  for (int i = 0; i < 3; ++i) runs[i] = phys(sizes[i], sizes[i], 2, MASS_LYO_FP32, 0);
  mass_lyo_batch_t b = NULL;
  CHECK(mass_lyo_batch_create(runs, 3, &b), MASS_LYO_OK);
  int32_t rows = 0, waves = 0;
  CHECK(mass_lyo_batch_autotune(b, &rows, &waves), MASS_LYO_OK);
  for (int i = 0; i < 3; ++i) {
    mass_lyo_init in = random_init(10 + i);
    double ftot, ptot;
    CHECK(mass_lyo_batch_configure(b, i, &in, &ftot, &ptot), MASS_LYO_OK);
  }
  CHECK(mass_lyo_batch_step(b, 4), MASS_LYO_OK);
  double ms;
  CHECK(mass_lyo_batch_time_steps(b, 2, &ms), MASS_LYO_OK);
  double* buf = (double*)malloc(sizeof(double) * 9 * 128 * 128);
  for (int i = 0; i < 3; ++i) {
    CHECK(mass_lyo_batch_get_field(b, i, MASS_LYO_FF, buf), MASS_LYO_OK);
    CHECK(mass_lyo_batch_get_field(b, i, MASS_LYO_SIGMAXY, buf), MASS_LYO_OK);
    double fs, ps;
    CHECK(mass_lyo_batch_sums(b, i, &fs, &ps), MASS_LYO_OK);
    mass_lyo_diag d;
    CHECK(mass_lyo_batch_diagnostics(b, i, &d, NULL), MASS_LYO_OK);
  }
  free(buf);
  mass_lyo_batch_destroy(b);
  printf("batch ok\n");
  fflush(stdout);
}

static void network(void) {
  mass_lyo_net_t net = NULL;
  CHECK(mass_lyo_net_create(7, &net), MASS_LYO_OK);
  int np = mass_lyo_net_num_params(net);
  float* w = (float*)malloc(sizeof(float) * (size_t)np);
  CHECK(mass_lyo_net_get_params(net, w), MASS_LYO_OK);
  CHECK(mass_lyo_net_set_params(net, w), MASS_LYO_OK);
  free(w);
  /* host-buffer forward and gradient on a 16 x 16 patch */
  const int P = 16 * 16;
  float* x = (float*)calloc((size_t)8 * P, sizeof(float));
  float* y = (float*)calloc((size_t)8 * P, sizeof(float));
  for (int i = 0; i < 8 * P; ++i) x[i] = (float)((i * 37) % 11) * 0.1f - 0.5f;
  CHECK(mass_lyo_net_forward_host(net, x, 16, 16, y), MASS_LYO_OK);
  free(x);
  free(y);

  /* frames of a plain FP32 run, training, then surrogate mode */
  mass_lyo_t src = make(64, 64, 2, MASS_LYO_FP32, 0);
  configure(src, 5);
  for (int i = 0; i < 2; ++i) {
    CHECK(mass_lyo_step(src, 2), MASS_LYO_OK);
    CHECK(mass_lyo_net_collect(net, src), MASS_LYO_OK);
  }
  mass_lyo_destroy(src);
  double l0, l1;
  CHECK(mass_lyo_net_train(net, 4, 1e-3f, &l0, &l1), MASS_LYO_OK);

  mass_lyo_t h = make(128, 128, 2, MASS_LYO_FP32, 0);
  configure(h, 6);
  CHECK(mass_lyo_surrogate_enable(h, net, 2, 1.0), MASS_LYO_OK);
  CHECK(mass_lyo_surrogate_set_calibration(h, 1), MASS_LYO_OK);
  CHECK(mass_lyo_step(h, 3), MASS_LYO_OK);
  CHECK(mass_lyo_surrogate_set_kernels(h, 0), MASS_LYO_OK);
  CHECK(mass_lyo_step(h, 2), MASS_LYO_OK);
  double cal[5];
  CHECK(mass_lyo_surrogate_get_calibration(h, cal), MASS_LYO_OK);
  mass_lyo_surrogate_stats st;
  CHECK(mass_lyo_surrogate_get_stats(h, &st), MASS_LYO_OK);
  CHECK(mass_lyo_surrogate_disable(h), MASS_LYO_OK);
  CHECK(mass_lyo_step(h, 1), MASS_LYO_OK);
  mass_lyo_destroy(h);
  mass_lyo_net_destroy(net);
  printf("network and surrogate ok\n");
  fflush(stdout);
}

/* Calls that must fail cleanly, with or without a device. */
static void error_paths(int have_device) {
  mass_lyo_params p = phys(0, 64, 1, MASS_LYO_FP64, 0); /* LX = 0 */
  mass_lyo_t h = NULL;
  int rc = mass_lyo_create(&p, &h);
  if (rc == MASS_LYO_OK || h != NULL) {
    fprintf(stderr, "create with LX = 0 succeeded\n");
    failures++;
  }
  if (!mass_lyo_last_error(NULL) || !mass_lyo_last_error(NULL)[0]) {
    fprintf(stderr, "no error text after a failed create\n");
    failures++;
  }
  p = phys(64, 64, 1, MASS_LYO_FP64, 0);
  p.npc = 0; /* parameters are checked before the device */
  CHECK(mass_lyo_create(&p, &h), MASS_LYO_EINVAL);
  mass_lyo_params runs[2] = {phys(64, 64, 2, MASS_LYO_FP32, 0), phys(64, 64, 2, MASS_LYO_FP32, 0)};
  runs[1].precision = MASS_LYO_FP64; /* mixed precision in one batch */
  mass_lyo_batch_t b = NULL;
  if (mass_lyo_batch_create(runs, 2, &b) == MASS_LYO_OK) {
    fprintf(stderr, "batch of mixed precision was accepted\n");
    failures++;
    mass_lyo_batch_destroy(b);
  }
  if (!have_device) {
    p = phys(64, 64, 1, MASS_LYO_FP64, 0);
    CHECK(mass_lyo_create(&p, &h), MASS_LYO_EDEVICE);
    mass_lyo_net_t n = NULL;
    if (mass_lyo_net_create(1, &n) == MASS_LYO_OK) {
      fprintf(stderr, "network created without a device\n");
      failures++;
      mass_lyo_net_destroy(n);
    }
  } else {
    /* a handle whose calls fail after creation */
    h = make(64, 64, 1, MASS_LYO_FP64, 0);
    mass_lyo_init in = random_init(1);
    in.config = 99;
    double ftot, ptot;
    CHECK(mass_lyo_configure(h, &in, &ftot, &ptot), MASS_LYO_EINVAL);
    double v;
    CHECK(mass_lyo_get_field(h, 99, &v), MASS_LYO_EINVAL);
    CHECK(mass_lyo_surrogate_set_kernels(h, 1), MASS_LYO_EINVAL);
    CHECK(mass_lyo_checkpoint_load(h, &v, sizeof v), MASS_LYO_EINVAL);
    mass_lyo_destroy(h);
  }
  mass_lyo_destroy(NULL);
  mass_lyo_batch_destroy(NULL);
  mass_lyo_net_destroy(NULL);
  printf("error paths ok\n");
  fflush(stdout);
}

/* allocates a buffer and loses its address on return (no caller's frame
 * or register keeps it once later calls have run) */
__attribute__((noinline)) static void drop_buffer(void) {
  volatile char* p = (volatile char*)malloc(4321);
  p[0] = 1;
}

int main(int argc, char** argv) {
  if (mass_lyo_abi_version() != MASS_LYO_ABI_VERSION) {
    fprintf(stderr, "ABI version %d, header %d\n", mass_lyo_abi_version(), MASS_LYO_ABI_VERSION);
    return 2;
  }
  if (argc > 1 && strcmp(argv[1], "--no-device") == 0) {
    if (argc > 2 && strcmp(argv[2], "--leak") == 0) drop_buffer();
    error_paths(0);
  } else {
    int cycles = argc > 1 ? atoi(argv[1]) : 2;
    for (int c = 0; c < cycles; ++c) {
      single_runs();
      batches();
      network();
    }
    error_paths(1);
  }
  if (failures) fprintf(stderr, "%d failed calls\n", failures);
  return failures ? 1 : 0;
}
