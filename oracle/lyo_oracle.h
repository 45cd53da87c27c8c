/*
 * lyo_oracle.h -- CPU oracle for the lyotropic active-nematic time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only the
 * tests, the package smoke check and the benchmark's CPU-baseline leg may load
 * it.  The shipped path (libmass_lyo.so) never calls into it.
 *
 * It is a literal FP64 restatement of the reference's Lyotropic model
 * (src/models/lyotropic.cpp, src/lb.hpp, src/geometry.cpp, src/random.cpp),
 * keeping the reference's memory layout (AoS populations ff[k][v], a periodic
 * neighbour table, separate UpdateQuantities / UpdateFields / Move loops) and
 * its expression order, so that it doubles as the CPU baseline ("port").
 */
#ifndef LYO_ORACLE_H_
#define LYO_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lyo_oracle_params {
  int32_t LX, LY;
  int32_t npc;          /* corrector sweeps per step, lyotropic.hpp:40 */
  int32_t nthreads;     /* 0 = serial (reference default, options.cpp:51-52) */
  int32_t conserve_phi; /* lyotropic.hpp:46 (false for plain Lyotropic) */
  int32_t bc;           /* boundary conditions, lyotropic.cpp:388-498: 0 periodic,
                           1 free-slip channel, 2 no-slip channel, 3 free-slip
                           box, 4 no-slip box (1-4 on a ghost layer) */
  double GammaQ, GammaP, xi, tauNem, tauIso, rho, friction;
  double AA, CC, LL, KK, zeta, zetaI, conc;
} lyo_oracle_params;

/* Initial-configuration parameters (lyotropic.cpp:69-160).  config:
 * 0 circle, 1 square, 2 stripe, 3 half, 4 random (director angle drawn per
 * node, phi = conc + U(-phinoise, phinoise) drawn right after it). */
typedef struct lyo_oracle_init {
  int32_t config;
  uint32_t seed;
  double level, radius, angle_deg, noise;
  double qamp;      /* Q amplitude: Qxx = qamp*order*cos(2 theta) */
  double phinoise;  /* half width of the phi draw for config 4 */
} lyo_oracle_init;

typedef struct lyo_oracle lyo_oracle;

/* field ids for lyo_oracle_get / lyo_oracle_set_field */
enum {
  LYO_O_FF = 0,   /* N*9, AoS (k-major) */
  LYO_O_FN = 1,   /* N*9 */
  LYO_O_QQXX, LYO_O_QQYX, LYO_O_PHI, LYO_O_QNXX, LYO_O_QNYX, LYO_O_PHN,
  LYO_O_N, LYO_O_UX, LYO_O_UY, LYO_O_UX_PHI, LYO_O_UY_PHI,
  LYO_O_HHXX, LYO_O_HHYX, LYO_O_MU,
  LYO_O_DXQQXX, LYO_O_DYQQXX, LYO_O_DXQQYX, LYO_O_DYQQYX,
  LYO_O_SIGMAXX, LYO_O_SIGMAYY, LYO_O_SIGMAYX, LYO_O_SIGMAXY,
  LYO_O_SIGMA_BULK, LYO_O_SIGMA_EL_XX, LYO_O_SIGMA_EL_YX,
  LYO_O_SIGMA_PF_XX, LYO_O_SIGMA_PF_YX, LYO_O_SIGMA_ACT_XX, LYO_O_SIGMA_ACT_YX,
  LYO_O_M, LYO_O_MN, LYO_O_CHI, /* go-or-grow: phenotype density, its half step, phenotype */
  LYO_O_NFIELDS
};

/* go-or-grow (src/models/go-or-grow.cpp, lyotropic-with-division.cpp):
 * parameters of go-or-grow.hpp:10-28 and lyotropic-with-division.hpp:19-34 */
typedef struct lyo_oracle_growth {
  int32_t chi_config;   /* 0 noise, 1 front */
  int32_t growTogether, relax_phi, relax_Q;
  uint32_t relax_steps;
  int32_t division_only; /* 1: lyotropic-with-division (Lyotropic + the masks'
                            phi growth, lyotropic-with-division.cpp:110-126);
                            only alpha ... phi_critical are used */
  double B, Snem, Dchi, Achi, Ochi, phiswitch;
  double chi0, chi_noise, chi_length, relax_dt;
  double alpha, beta, division_rate, division_time, division_radius;
  double death_rate, death_time, death_radius, phi_critical;
} lyo_oracle_growth;

/* Makes the run a go-or-grow run (call before configure): Initialize of
 * go-or-grow.cpp:17-32 and lyotropic-with-division.cpp:18-32 (phi is
 * conserved when neither division nor death can happen). */
int lyo_oracle_set_growth(lyo_oracle* o, const lyo_oracle_growth* g);
/* division and death masks (N bytes each, 0/1) of the last step */
void lyo_oracle_masks(const lyo_oracle* o, uint8_t* division, uint8_t* death);
/* std::normal_distribution<>(0, 1)(gen), a fresh distribution per draw as
 * random_normal does (random.cpp:27-30) */
void lyo_oracle_random_normal(uint32_t seed, int n, double* out);

lyo_oracle* lyo_oracle_create(const lyo_oracle_params* p);
void lyo_oracle_destroy(lyo_oracle* o);

/* Restatement of Lyotropic::Configure (lyotropic.cpp:69-166) with the
 * mt19937 / generate_canonical draw of random.cpp:21-24. */
int lyo_oracle_configure(lyo_oracle* o, const lyo_oracle_init* init);

/* Direct state upload (ff AoS N*9). */
void lyo_oracle_set_state(lyo_oracle* o, const double* ff, const double* Qxx,
                          const double* Qyx, const double* phi);
int lyo_oracle_get(const lyo_oracle* o, int field, double* out);
int lyo_oracle_set_field(lyo_oracle* o, int field, const double* in);

/* Lyotropic::Step (lyotropic.cpp:500-525), nsteps times. */
void lyo_oracle_step(lyo_oracle* o, int nsteps);
/* the three loops of the step, exposed for unit checks */
void lyo_oracle_update_quantities(lyo_oracle* o);
void lyo_oracle_update_fields(lyo_oracle* o, int first);
void lyo_oracle_move(lyo_oracle* o);

/* sums as in RuntimeChecks (lyotropic.cpp:527-549): serial, k ascending */
void lyo_oracle_sums(const lyo_oracle* o, double* fsum, double* phisum,
                     double* ftot, double* ptot, double* totalphi,
                     double* countphi);

/* neighbour table, lyo_oracle_total_size(o)*9: periodic (geometry.cpp:
 * 100-110) for bc 0, the ghost-layer grid (geometry.cpp:140-183) otherwise */
void lyo_oracle_neighbours(const lyo_oracle* o, uint32_t* out);
/* nodes including the ghost layer: LX*LY (bc 0) or (LX+2)*(LY+2) */
uint32_t lyo_oracle_total_size(const lyo_oracle* o);
/* the boundary-condition calls of the step (lyotropic.cpp:388-498), exposed
 * for unit checks: which 0 = BoundaryConditionsFields, 1 = ...Fields2,
 * 2 = ...LB */
void lyo_oracle_boundary(lyo_oracle* o, int which);
/* whole-grid access including the ghost layer (total_size values; ff/fn
 * total_size*9), index as in the reference: domain, then walls Left, Right,
 * Front, Back, then the corners LeftFront, RightFront, LeftBack, RightBack */
int lyo_oracle_get_all(const lyo_oracle* o, int field, double* out);

/* raw mt19937 outputs and the canonical double of the reference's
 * random_real(), for pinning the RNG restatement */
void lyo_oracle_mt19937(uint32_t seed, int n, uint32_t* out);
void lyo_oracle_random_real(uint32_t seed, int n, double lo, double hi,
                            double* out);

/* equilibrium distribution, lb.hpp:24-41 */
void lyo_oracle_equilibrium(double vx, double vy, double nn, double* fe9);

#ifdef __cplusplus
}
#endif
#endif
