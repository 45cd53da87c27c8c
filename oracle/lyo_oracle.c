/*
 * lyo_oracle.c -- CPU oracle for the lyotropic active-nematic time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU-baseline
 * "port".  Nothing in the shipped library links or calls this file.
 *
 * Literal FP64 restatement of the reference, following it line by line:
 *   periodic neighbour table        src/geometry.cpp:100-110, src/tools.hpp:24-28
 *   equilibrium distribution        src/lb.hpp:24-41
 *   derivX / derivY / laplacian     src/lb.hpp:73-90, 127-136
 *   UpdateQuantitiesAtNode          src/models/lyotropic.cpp:168-247
 *   UpdateFieldsAtNode              src/models/lyotropic.cpp:249-341
 *   UpdateQuantities/Fields, Move   src/models/lyotropic.cpp:343-386
 *   Step                            src/models/lyotropic.cpp:500-525
 *   RuntimeChecks sums              src/models/lyotropic.cpp:527-549
 *   ghost-layer grid (bc != 0)      src/geometry.cpp:16-30,140-183,
 *                                   src/geometry.hpp:17-89,148-199,306-316,475-501
 *   boundary conditions             src/models/lyotropic.cpp:388-498,
 *                                   src/fields.hpp:619-905
 *   go-or-grow                      src/models/go-or-grow.cpp:17-404,
 *                                   src/models/lyotropic-with-division.cpp:18-156
 *   random_normal                   src/random.cpp:27-30 (libstdc++
 *                                   normal_distribution: Marsaglia polar
 *                                   method, a fresh distribution per draw)
 *   ConfigureAtNode                 src/models/lyotropic.cpp:69-166
 *   mt19937 + random_real           src/random.cpp:12,21-24,43-46 (libstdc++
 *                                   generate_canonical<double,53>: two 32-bit
 *                                   draws, (g1 + g2*2^32)/2^64, clamped < 1;
 *                                   uniform_real: U*(b-a)+a)
 * Expression order is kept exactly (no reassociation; build without FMA
 * contraction) so that an IEEE-exact device path can match it bit for bit.
 * Layout is the reference's: AoS populations ff[k][v], k = y + LY*x.
 */
#include "lyo_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* src/lb.hpp:15-16: forcing weights */
static const double W[9] = {0, 1/3., 1/3., 1/3., 1/3.,
                            1/12., 1/12., 1/12., 1/12.};
/* src/lb.hpp:21 */
static const double SB = 1/12., SD = 1/12.;
/* src/geometry.hpp:17-27 */
static const int XDIR[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
static const int YDIR[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};

/* The generator fixed by the C++ standard ([rand.predef]: mt19937). */
typedef struct { uint32_t mt[624]; int idx; } mt_state;

struct lyo_oracle {
  lyo_oracle_params p;
  uint32_t N;   /* domain nodes */
  uint32_t T;   /* all nodes: N, or (LX+2)*(LY+2) with the ghost layer */
  uint32_t* nb; /* T*9 */
  double *ff, *fn, *ff_tmp, *fn_tmp; /* T*9 each */
  double* s[LYO_O_NFIELDS];          /* scalar fields (index >= 2) */
  double* phi_tmp;
  double totalphi, countphi, ftot, ptot;
  mt_state rng; /* the reference's global generator (random.cpp:12), seeded at configure */
  /* go-or-grow */
  int growth;
  lyo_oracle_growth g;
  double* m_tmp;
  uint8_t *div_mask, *death_mask;
  uint32_t div_count, death_count;
};

/* ---------------------------------------------------------------- helpers */

/* src/tools.hpp:24-28 (the double overload is the one the periodic table
 * reaches: long + unsigned arguments fail template deduction) */
static double modu_d(double num, double div) {
  if (num < 0) return div + fmod(num, div);
  return fmod(num, div);
}

static inline const uint32_t* get_neighbours(const lyo_oracle* o, uint32_t k) {
  /* std::vector::at bounds check of geometry.hpp:544-545 */
  if (k >= o->T) {
    fprintf(stderr, "lyo_oracle: neighbour index %u out of range\n", k);
    abort();
  }
  return o->nb + 9 * (size_t)k;
}

/* src/lb.hpp:73-90,127-136: the active (non-commented) stencils */
static inline double derivX(const double* a, const uint32_t* d, double st) {
  return .5 * (a[d[1]] - a[d[2]]) + 0 * st;
}
static inline double derivY(const double* a, const uint32_t* d, double st) {
  return .5 * (a[d[3]] - a[d[4]]) + 0 * st;
}
static inline double laplacian(const double* a, const uint32_t* d, double st) {
  return a[d[1]] + a[d[2]] + a[d[3]] + a[d[4]] - 4 * a[d[0]] + 0 * st;
}

/* src/lb.hpp:24-41 */
static inline void equilibrium(double vx, double vy, double nn, double* fe) {
  const double v2 = vx * vx + vy * vy;
  fe[1] = nn * (1. / 9. + 1. / 3. * vx + 1. / 3. * v2 - 1. / 2. * vy * vy);
  fe[2] = nn * (1. / 9. - 1. / 3. * vx + 1. / 3. * v2 - 1. / 2. * vy * vy);
  fe[3] = nn * (1. / 9. + 1. / 3. * vy - 1. / 2. * vx * vx + 1. / 3. * v2);
  fe[4] = nn * (1. / 9. - 1. / 3. * vy - 1. / 2. * vx * vx + 1. / 3. * v2);
  fe[5] = nn * (1. / 36. + 1. / 12. * vx + 1. / 12. * vy + 1. / 12. * v2 + 1. / 4. * vx * vy);
  fe[6] = nn * (1. / 36. - 1. / 12. * vx - 1. / 12. * vy + 1. / 12. * v2 + 1. / 4. * vx * vy);
  fe[7] = nn * (1. / 36. - 1. / 12. * vx + 1. / 12. * vy + 1. / 12. * v2 - 1. / 4. * vx * vy);
  fe[8] = nn * (1. / 36. + 1. / 12. * vx - 1. / 12. * vy + 1. / 12. * v2 - 1. / 4. * vx * vy);
  fe[0] = nn - fe[1] - fe[2] - fe[3] - fe[4] - fe[5] - fe[6] - fe[7] - fe[8];
}

void lyo_oracle_equilibrium(double vx, double vy, double nn, double* fe9) {
  equilibrium(vx, vy, nn, fe9);
}

/* ------------------------------------------------------------ mt19937 RNG */

static void mt_seed(mt_state* s, uint32_t seed) {
  s->mt[0] = seed;
  for (int i = 1; i < 624; ++i)
    s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
  s->idx = 624;
}

static uint32_t mt_next(mt_state* s) {
  if (s->idx >= 624) {
    for (int i = 0; i < 624; ++i) {
      uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = s->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      s->mt[i] = v;
    }
    s->idx = 0;
  }
  uint32_t y = s->mt[s->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* generate_canonical<double,53> over a 32-bit engine: k = 2 draws */
static double mt_canonical(mt_state* s) {
  double sum = 0, tmp = 1;
  for (int k = 2; k != 0; --k) {
    sum += (double)mt_next(s) * tmp;
    tmp *= 4294967296.0;
  }
  double ret = sum / tmp;
  if (ret >= 1.0) ret = nextafter(1.0, 0.0);
  return ret;
}

/* uniform_real_distribution<>(lo, hi)(gen) */
static double mt_real(mt_state* s, double lo, double hi) {
  return (mt_canonical(s) * (hi - lo)) + lo;
}

/* normal_distribution<>(0, 1)(gen) of a fresh distribution (libstdc++'s
 * polar method: pairs of 2U-1 until 0 < r2 <= 1, returns y*sqrt(-2 ln r2/r2),
 * then *stddev + mean) */
static double mt_normal(mt_state* s) {
  double x, y, r2;
  do {
    x = 2.0 * mt_canonical(s) - 1.0;
    y = 2.0 * mt_canonical(s) - 1.0;
    r2 = x * x + y * y;
  } while (r2 > 1.0 || r2 == 0.0);
  const double mult = sqrt(-2 * log(r2) / r2);
  const double ret = y * mult;
  return ret * 1.0 + 0.0;
}

void lyo_oracle_random_normal(uint32_t seed, int n, double* out) {
  mt_state s;
  mt_seed(&s, seed);
  for (int i = 0; i < n; ++i) out[i] = mt_normal(&s);
}

void lyo_oracle_mt19937(uint32_t seed, int n, uint32_t* out) {
  mt_state s;
  mt_seed(&s, seed);
  for (int i = 0; i < n; ++i) out[i] = mt_next(&s);
}

void lyo_oracle_random_real(uint32_t seed, int n, double lo, double hi, double* out) {
  mt_state s;
  mt_seed(&s, seed);
  for (int i = 0; i < n; ++i) out[i] = mt_real(&s, lo, hi);
}

/* ------------------------------------------------------------- lifecycle */

/* src/geometry.cpp:150-171: node (x, y) of the (LX+2)*(LY+2) grid -> index
 * (domain first, then walls Left, Right, Front, Back, then the corners;
 * src/geometry.hpp:475-501) */
static uint32_t layer_index(uint32_t LX, uint32_t LY, uint32_t x, uint32_t y) {
  const uint32_t N = LX * LY, T = (LX + 2) * (LY + 2);
  if (x == 0 && y == 0) return T - 4;              /* LeftFront */
  if (x == LX + 1 && y == 0) return T - 3;         /* RightFront */
  if (x == 0 && y == LY + 1) return T - 2;         /* LeftBack */
  if (x == LX + 1 && y == LY + 1) return T - 1;    /* RightBack */
  if (x == 0) return N + (y - 1);                  /* Left wall */
  if (x == LX + 1) return N + LY + (y - 1);        /* Right wall */
  if (y == 0) return N + 2 * LY + (x - 1);         /* Front wall */
  if (y == LY + 1) return N + 2 * LY + LX + (x - 1); /* Back wall */
  return (y - 1) + LY * (x - 1);
}

lyo_oracle* lyo_oracle_create(const lyo_oracle_params* p) {
  if (p->LX < 2 || p->LY < 2 || p->bc < 0 || p->bc > 4) return NULL;
  lyo_oracle* o = (lyo_oracle*)calloc(1, sizeof(lyo_oracle));
  o->p = *p;
  o->N = (uint32_t)p->LX * (uint32_t)p->LY;
  /* src/geometry.cpp:24-30: Lyotropic uses the Layer grid for bc != 0
   * (lyotropic.cpp:15-17) */
  o->T = p->bc ? ((uint32_t)p->LX + 2) * ((uint32_t)p->LY + 2) : o->N;
  const size_t T = o->T;
  o->nb = (uint32_t*)malloc(sizeof(uint32_t) * 9 * T);
  /* Field::SetSize value-initialises: every field starts at zero */
  o->ff = (double*)calloc(9 * T, sizeof(double));
  o->fn = (double*)calloc(9 * T, sizeof(double));
  o->ff_tmp = (double*)calloc(9 * T, sizeof(double));
  o->fn_tmp = (double*)calloc(9 * T, sizeof(double));
  for (int f = 2; f < LYO_O_NFIELDS; ++f) o->s[f] = (double*)calloc(T, sizeof(double));
  o->phi_tmp = (double*)calloc(T, sizeof(double));
  o->m_tmp = (double*)calloc(T, sizeof(double));
  o->div_mask = (uint8_t*)calloc(o->N, 1);
  o->death_mask = (uint8_t*)calloc(o->N, 1);
  if (!p->bc) {
    /* geometry.cpp:100-110: k = y + LY*x, neighbour v at ((x+cx) mod LX, (y+cy) mod LY) */
    for (uint32_t k = 0; k < o->N; ++k) {
      const long x = k / (uint32_t)p->LY;
      const long y = k % (uint32_t)p->LY;
      for (int v = 0; v < 9; ++v) {
        const unsigned xx = (unsigned)modu_d((double)(x + XDIR[v]), (double)(unsigned)p->LX);
        const unsigned yy = (unsigned)modu_d((double)(y + YDIR[v]), (double)(unsigned)p->LY);
        o->nb[9 * (size_t)k + v] = yy + (unsigned)p->LY * xx;
      }
    }
  } else {
    /* geometry.cpp:173-182: the periodic table of the (LX+2)*(LY+2) grid,
     * converted to the layer's indices */
    const unsigned LX = (unsigned)p->LX, LY = (unsigned)p->LY;
    for (uint32_t k = 0; k < o->T; ++k) {
      const long x = k / (LY + 2);
      const long y = k % (LY + 2);
      const uint32_t kk = layer_index(LX, LY, (uint32_t)x, (uint32_t)y);
      for (int v = 0; v < 9; ++v) {
        const unsigned xx = (unsigned)modu_d((double)(x + XDIR[v]), (double)(LX + 2));
        const unsigned yy = (unsigned)modu_d((double)(y + YDIR[v]), (double)(LY + 2));
        o->nb[9 * (size_t)kk + v] = layer_index(LX, LY, xx, yy);
      }
    }
  }
  return o;
}

void lyo_oracle_destroy(lyo_oracle* o) {
  if (!o) return;
  free(o->nb);
  free(o->ff); free(o->fn); free(o->ff_tmp); free(o->fn_tmp);
  for (int f = 2; f < LYO_O_NFIELDS; ++f) free(o->s[f]);
  free(o->phi_tmp);
  free(o->m_tmp);
  free(o->div_mask);
  free(o->death_mask);
  free(o);
}

void lyo_oracle_neighbours(const lyo_oracle* o, uint32_t* out) {
  memcpy(out, o->nb, sizeof(uint32_t) * 9 * (size_t)o->T);
}

uint32_t lyo_oracle_total_size(const lyo_oracle* o) { return o->T; }

#define F(id) (o->s[id])

static void growth_configure(lyo_oracle* o);

/* ------------------------------------------------------------- configure */

static unsigned udiff(unsigned a, unsigned b) { return a > b ? a - b : b - a; }
static unsigned uwrap(unsigned x, unsigned L) {
  unsigned a = x % L, b = L - x % L;
  return a < b ? a : b;
}

int lyo_oracle_configure(lyo_oracle* o, const lyo_oracle_init* in) {
  const unsigned LX = (unsigned)o->p.LX, LY = (unsigned)o->p.LY;
  mt_state* rng = &o->rng;
  mt_seed(rng, in->seed);
  const double angle = in->angle_deg * M_PI / 180.;
  const double level = in->level, radius = in->radius;
  o->totalphi = 0; o->ftot = 0; o->ptot = 0;
  for (unsigned k = 0; k < o->N; ++k) {
    double nematicOrder = 0;
    const unsigned x = k / LY, y = k % LY;
    switch (in->config) {
      case 0: /* circle */
        if (pow(udiff(x, LX / 2), 2) + pow(udiff(y, LY / 2), 2) <= radius * radius) nematicOrder = 1;
        break;
      case 1: /* square */
        if (udiff(LY / 2, y) < level / 2 && udiff(LX / 2, x) < level / 2) nematicOrder = 1;
        break;
      case 2: /* stripe */
        if (udiff(LY / 2, y) < level / 2) nematicOrder = 1;
        break;
      case 3: /* half */
        if (y < level) nematicOrder = 1;
        break;
      case 4: /* random: uniform order, phi drawn per node */
        nematicOrder = 1;
        break;
      case 5: /* boxatwall */
        if (o->p.bc == 4) {
          if (x < level && udiff(LY / 2, y) < level / 2) nematicOrder = 1;
        } else if (y < level && udiff(LX / 2, x) < level) {
          nematicOrder = 1;
        }
        break;
      case 6: /* cuttingatwall (unsigned arithmetic as in the reference) */
        if (o->p.bc == 4) {
          const unsigned off = (unsigned)(int)(radius * 1.5);
          if (pow(x, 2) + pow(uwrap(udiff(y, LY / 2 + off), LY), 2) <= radius * radius) nematicOrder = 1;
          if (pow(x, 2) + pow(uwrap(udiff(y, LY / 2 - off), LY), 2) <= radius * radius) nematicOrder = 1;
        } else if (y < level && udiff(LX / 2, x) < level) {
          nematicOrder = 1;
        }
        break;
      case 7: /* circleatwall, bc=0 branch */
        if (pow(uwrap(udiff(x, LX / 2), LX), 2) + pow(udiff(y, 0 * LY), 2) <= radius * radius) nematicOrder = 1;
        break;
      case 8: /* wettedwall, bc=0 branch */
        if (y < level) nematicOrder = 1;
        break;
      default:
        return -1;
    }
    const double theta = angle + in->noise * M_PI * (mt_real(rng, 0., 1.) - .5);
    F(LYO_O_QQXX)[k] = in->qamp * nematicOrder * (cos(2 * theta));
    F(LYO_O_QQYX)[k] = in->qamp * nematicOrder * (sin(2 * theta));
    if (in->config == 4)
      F(LYO_O_PHI)[k] = o->p.conc + mt_real(rng, -in->phinoise, in->phinoise);
    else
      F(LYO_O_PHI)[k] = nematicOrder * o->p.conc;
    o->totalphi += F(LYO_O_PHI)[k];
    F(LYO_O_UX)[k] = F(LYO_O_UY)[k] = F(LYO_O_UX_PHI)[k] = F(LYO_O_UY_PHI)[k] = 0;
    F(LYO_O_N)[k] = o->p.rho;
    equilibrium(F(LYO_O_UX)[k], F(LYO_O_UY)[k], F(LYO_O_N)[k], o->ff + 9 * (size_t)k);
    for (int v = 0; v < 9; ++v) o->ftot = o->ftot + o->ff[9 * (size_t)k + v];
    o->ptot += F(LYO_O_PHI)[k];
  }
  if (o->growth) {
    /* a configured run starts afresh: no masks, both due at the first step
       (the reference configures once, after lyotropic-with-division.cpp:23-31
       set the counters) */
    o->div_count = o->death_count = 0u;
    memset(o->div_mask, 0, o->N);
    memset(o->death_mask, 0, o->N);
    growth_configure(o);
  }
  return 0;
}

void lyo_oracle_set_state(lyo_oracle* o, const double* ff, const double* Qxx,
                          const double* Qyx, const double* phi) {
  const size_t N = o->N;
  memcpy(o->ff, ff, sizeof(double) * 9 * N);
  memcpy(F(LYO_O_QQXX), Qxx, sizeof(double) * N);
  memcpy(F(LYO_O_QQYX), Qyx, sizeof(double) * N);
  memcpy(F(LYO_O_PHI), phi, sizeof(double) * N);
  o->ftot = 0; o->ptot = 0; o->totalphi = 0;
  for (size_t k = 0; k < N; ++k) {
    for (int v = 0; v < 9; ++v) o->ftot = o->ftot + ff[9 * k + v];
    o->ptot += phi[k];
    o->totalphi += phi[k];
  }
}

int lyo_oracle_get(const lyo_oracle* o, int field, double* out) {
  const size_t N = o->N;
  if (field == LYO_O_FF) { memcpy(out, o->ff, sizeof(double) * 9 * N); return 0; }
  if (field == LYO_O_FN) { memcpy(out, o->fn, sizeof(double) * 9 * N); return 0; }
  if (field < 2 || field >= LYO_O_NFIELDS) return -1;
  memcpy(out, o->s[field], sizeof(double) * N);
  return 0;
}

int lyo_oracle_get_all(const lyo_oracle* o, int field, double* out) {
  const size_t T = o->T;
  if (field == LYO_O_FF) { memcpy(out, o->ff, sizeof(double) * 9 * T); return 0; }
  if (field == LYO_O_FN) { memcpy(out, o->fn, sizeof(double) * 9 * T); return 0; }
  if (field < 2 || field >= LYO_O_NFIELDS) return -1;
  memcpy(out, o->s[field], sizeof(double) * T);
  return 0;
}

int lyo_oracle_set_field(lyo_oracle* o, int field, const double* in) {
  const size_t N = o->N;
  if (field == LYO_O_FF) { memcpy(o->ff, in, sizeof(double) * 9 * N); return 0; }
  if (field == LYO_O_FN) { memcpy(o->fn, in, sizeof(double) * 9 * N); return 0; }
  if (field < 2 || field >= LYO_O_NFIELDS) return -1;
  memcpy(o->s[field], in, sizeof(double) * N);
  return 0;
}

/* -------------------------------------------------------------- the step */

/* src/models/lyotropic.cpp:168-247 */
static void update_quantities_at_node(lyo_oracle* o, unsigned k) {
  const lyo_oracle_params* P = &o->p;
  const uint32_t* d = get_neighbours(o, k);
  const double* f = o->ff + 9 * (size_t)k;
  const double* QQxx = F(LYO_O_QQXX);
  const double* QQyx = F(LYO_O_QQYX);
  const double* phi = F(LYO_O_PHI);
  const double Qxx = QQxx[k];
  const double Qyx = QQyx[k];
  const double p = phi[k];

  const double nn = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  const double vx = (f[1] - f[2] + f[5] - f[6] - f[7] + f[8]) / nn;
  const double vy = (f[3] - f[4] + f[5] - f[6] + f[7] - f[8]) / nn;

  const double del2Qxx = laplacian(QQxx, d, SD);
  const double dxQxx = derivX(QQxx, d, SB);
  const double dyQxx = derivY(QQxx, d, SB);
  const double del2Qyx = laplacian(QQyx, d, SD);
  const double dxQyx = derivX(QQyx, d, SB);
  const double dyQyx = derivY(QQyx, d, SB);
  const double del2p = laplacian(phi, d, SD);
  const double dxPhi = derivX(phi, d, SB);
  const double dyPhi = derivY(phi, d, SB);

  const double term = 0 - Qxx * Qxx - Qyx * Qyx;
  const double Hxx = P->CC * term * Qxx + P->LL * del2Qxx;
  const double Hyx = P->CC * term * Qyx + P->LL * del2Qyx;
  const double mu = P->AA * p * (1 - p) * (1 - 2 * p) + 0 * P->CC * term - P->KK * del2p;

  const double sigmaB = .5 * P->AA * p * p * (1 - p) * (1 - p)
                        + .5 * P->CC * term * term - mu * p;
  const double sigmaF = 2 * P->xi * ((Qxx * Qxx - 1) * Hxx + Qxx * Qyx * Hyx)
                        - P->zeta * Qxx * (1 - p) + .5 * P->KK * (dyPhi * dyPhi - dxPhi * dxPhi)
                        + P->LL * (dyQxx * dyQxx + dyQyx * dyQyx - dxQxx * dxQxx - dxQyx * dxQyx);
  const double sigmaS = 2 * P->xi * (Qyx * Qxx * Hxx + (Qyx * Qyx - 1) * Hyx)
                        - P->zeta * Qyx * (1 - p) - P->KK * dxPhi * dyPhi
                        - 2 * P->LL * (dxQxx * dyQxx + dxQyx * dyQyx);
  const double sigmaA = 2 * (Qxx * Hyx - Qyx * Hxx);

  F(LYO_O_UX)[k] = vx;
  F(LYO_O_UY)[k] = vy;
  F(LYO_O_UX_PHI)[k] = vx * p;
  F(LYO_O_UY_PHI)[k] = vy * p;
  F(LYO_O_N)[k] = nn;
  F(LYO_O_HHXX)[k] = Hxx;
  F(LYO_O_HHYX)[k] = Hyx;
  F(LYO_O_MU)[k] = mu;
  F(LYO_O_DXQQXX)[k] = dxQxx;
  F(LYO_O_DXQQYX)[k] = dxQyx;
  F(LYO_O_DYQQXX)[k] = dyQxx;
  F(LYO_O_DYQQYX)[k] = dyQyx;
  F(LYO_O_SIGMAXX)[k] = sigmaF + sigmaB + P->zetaI * (P->conc - p);
  F(LYO_O_SIGMAYY)[k] = -sigmaF + sigmaB + P->zetaI * (P->conc - p);
  F(LYO_O_SIGMAXY)[k] = sigmaS + sigmaA;
  F(LYO_O_SIGMAYX)[k] = sigmaS - sigmaA;
  F(LYO_O_SIGMA_BULK)[k] = sigmaB + P->zetaI * (P->conc - p);
  F(LYO_O_SIGMA_EL_XX)[k] = sigmaF - .5 * P->KK * (dyPhi * dyPhi - dxPhi * dxPhi) + P->zeta * Qxx * (1 - p);
  F(LYO_O_SIGMA_EL_YX)[k] = sigmaS + P->KK * dxPhi * dyPhi + P->zeta * Qxx * (1 - p);
  F(LYO_O_SIGMA_PF_XX)[k] = .5 * P->KK * (dyPhi * dyPhi - dxPhi * dxPhi);
  F(LYO_O_SIGMA_PF_YX)[k] = -P->KK * dxPhi * dyPhi;
  F(LYO_O_SIGMA_ACT_XX)[k] = -P->zeta * Qxx * (1 - p);
  F(LYO_O_SIGMA_ACT_YX)[k] = -P->zeta * Qyx * (1 - p);
}

/* src/models/lyotropic.cpp:249-341 */
static void update_fields_at_node(lyo_oracle* o, unsigned k, int first) {
  const lyo_oracle_params* P = &o->p;
  const uint32_t* d = get_neighbours(o, k);
  double* QQxx = F(LYO_O_QQXX);
  double* QQyx = F(LYO_O_QQYX);
  double* phi = F(LYO_O_PHI);

  const double nn = F(LYO_O_N)[k];
  const double vx = F(LYO_O_UX)[k];
  const double vy = F(LYO_O_UY)[k];
  const double Qxx = QQxx[k];
  const double Qyx = QQyx[k];
  const double p = phi[k];
  const double Hxx = F(LYO_O_HHXX)[k];
  const double Hyx = F(LYO_O_HHYX)[k];
  const double dxQxx = F(LYO_O_DXQQXX)[k];
  const double dyQxx = F(LYO_O_DYQQXX)[k];
  const double dxQyx = F(LYO_O_DXQQYX)[k];
  const double dyQyx = F(LYO_O_DYQQYX)[k];

  const double del2mu = laplacian(F(LYO_O_MU), d, SD);
  const double dxux = derivX(F(LYO_O_UX), d, SB);
  const double dyux = derivY(F(LYO_O_UX), d, SB);
  const double dxuy = derivX(F(LYO_O_UY), d, SB);
  const double dyuy = derivY(F(LYO_O_UY), d, SB);
  const double dxSxx = derivX(F(LYO_O_SIGMAXX), d, SB);
  const double dySxy = derivY(F(LYO_O_SIGMAXY), d, SB);
  const double dxSyx = derivX(F(LYO_O_SIGMAYX), d, SB);
  const double dySyy = derivY(F(LYO_O_SIGMAYY), d, SB);
  const double pFlux = derivX(F(LYO_O_UX_PHI), d, SB) + derivY(F(LYO_O_UY_PHI), d, SB);

  const double expansion = dxux + dyuy;
  const double shear = .5 * (dxuy + dyux);
  const double vorticity = .5 * (dxuy - dyux);
  const double traceQL = Qxx * (dxux - dyuy) + 2 * Qyx * shear;

  const double Dxx = P->GammaQ * Hxx - vx * dxQxx - vy * dyQxx - 2 * vorticity * Qyx
                     + P->xi * ((Qxx + 1) * (2 * dxux - traceQL) + 2 * Qyx * shear - expansion);
  const double Dyx = P->GammaQ * Hyx - vx * dxQyx - vy * dyQyx + 2 * vorticity * Qxx
                     + P->xi * (Qyx * (expansion - traceQL) + 2 * shear);
  const double Dp = P->GammaP * del2mu - pFlux
                    - (P->conserve_phi ? (o->countphi - o->totalphi) / o->N : 0);

  const double Fx = dxSxx + dySxy - P->friction * vx;
  const double Fy = dxSyx + dySyy - P->friction * vy;

  const double tau = P->tauNem * p + P->tauIso * (1 - p);
  double fe[9];
  equilibrium(vx, vy, nn, fe);

  double* ffk = o->ff + 9 * (size_t)k;
  double* fnk = o->fn + 9 * (size_t)k;
  if (first) {
    F(LYO_O_QNXX)[k] = QQxx[k] + .5 * Dxx;
    F(LYO_O_QNYX)[k] = QQyx[k] + .5 * Dyx;
    F(LYO_O_PHN)[k] = phi[k] + .5 * Dp;
    QQxx[k] = QQxx[k] + Dxx;
    QQyx[k] = QQyx[k] + Dyx;
    o->phi_tmp[k] = phi[k] + Dp;
    for (int v = 0; v < 9; ++v) {
      fnk[v] = ffk[v] + .5 * ((fe[v] - ffk[v]) / tau + W[v] * (Fx * XDIR[v] + Fy * YDIR[v]));
      ffk[v] = ffk[v] + ((fe[v] - ffk[v]) / tau + W[v] * (Fx * XDIR[v] + Fy * YDIR[v]));
    }
  } else {
    QQxx[k] = F(LYO_O_QNXX)[k] + .5 * Dxx;
    QQyx[k] = F(LYO_O_QNYX)[k] + .5 * Dyx;
    o->phi_tmp[k] = F(LYO_O_PHN)[k] + .5 * Dp;
    for (int v = 0; v < 9; ++v)
      ffk[v] = fnk[v] + .5 * ((fe[v] - ffk[v]) / tau + W[v] * (Fx * XDIR[v] + Fy * YDIR[v]));
  }
}

/* ------------------------------------------------------------- go-or-grow */

static void s_apply(const lyo_oracle* o, double* a, int kind, double value, int channel);

int lyo_oracle_set_growth(lyo_oracle* o, const lyo_oracle_growth* g) {
  /* go-or-grow.cpp:21-26 */
  if (g->chi_noise < 0 || g->chi_length < 0 || g->relax_dt < 0) return -1;
  if (g->chi_config != 0 && g->chi_config != 1) return -1;
  o->growth = 1;
  o->g = *g;
  /* lyotropic-with-division.cpp:23-31 */
  o->div_count = o->death_count = 0u;
  o->p.conserve_phi = ((g->alpha == 0 || g->division_rate == 0) &&
                       (g->beta == 0 || g->death_rate == 0));
  return 0;
}

void lyo_oracle_masks(const lyo_oracle* o, uint8_t* division, uint8_t* death) {
  memcpy(division, o->div_mask, o->N);
  memcpy(death, o->death_mask, o->N);
}

/* go-or-grow.cpp:365-385: Lyotropic's field conditions, then m and chi */
static void growth_boundary_fields(lyo_oracle* o) {
  lyo_oracle_boundary(o, 0);
  if (o->growth && !o->g.division_only && o->p.bc) {
    s_apply(o, F(LYO_O_M), 0, 0., o->p.bc <= 2);
    s_apply(o, F(LYO_O_CHI), 0, 0., o->p.bc <= 2);
  }
}

/* go-or-grow.cpp:201-209 */
static void update_phenotype(lyo_oracle* o) {
  const double phi_epsilon = 1e-12;
  for (unsigned k = 0; k < o->N; ++k)
    F(LYO_O_CHI)[k] = F(LYO_O_PHI)[k] > phi_epsilon ? F(LYO_O_M)[k] / F(LYO_O_PHI)[k] : 0.;
  growth_boundary_fields(o);
}

/* go-or-grow.cpp:211-222 */
static void project_m(lyo_oracle* o) {
  for (unsigned k = 0; k < o->N; ++k) {
    const double upper = F(LYO_O_PHI)[k] > 0 ? F(LYO_O_PHI)[k] : 0.;
    if (F(LYO_O_M)[k] < 0)
      F(LYO_O_M)[k] = 0;
    else if (F(LYO_O_M)[k] > upper)
      F(LYO_O_M)[k] = upper;
  }
}

/* go-or-grow.cpp:42-119: dry free-energy descent of phi and Q */
static void relax_free_energy(lyo_oracle* o) {
  const lyo_oracle_growth* g = &o->g;
  const lyo_oracle_params* P = &o->p;
  if (g->relax_steps == 0 || (!g->relax_phi && !g->relax_Q)) return;
  for (unsigned step = 0; step < g->relax_steps; ++step) {
    lyo_oracle_boundary(o, 0);
    for (unsigned k = 0; k < o->N; ++k) {
      const uint32_t* d = get_neighbours(o, k);
      const double Qxx = F(LYO_O_QQXX)[k];
      const double Qyx = F(LYO_O_QQYX)[k];
      const double p = F(LYO_O_PHI)[k];
      const double q2 = Qxx * Qxx + Qyx * Qyx;
      const double term = g->Snem - q2;
      const double dp_critical = p - g->phi_critical;
      const double mu_compress = dp_critical > 0 ? g->B * dp_critical : 0.;
      F(LYO_O_HHXX)[k] = 2 * P->CC * term * Qxx + P->LL * laplacian(F(LYO_O_QQXX), d, SD);
      F(LYO_O_HHYX)[k] = 2 * P->CC * term * Qyx + P->LL * laplacian(F(LYO_O_QQYX), d, SD);
      F(LYO_O_MU)[k] = P->AA * p * (1 - p) * (1 - 2 * p) + mu_compress
                       - P->KK * laplacian(F(LYO_O_PHI), d, SD);
    }
    if (P->bc) s_apply(o, F(LYO_O_MU), 0, 0., P->bc <= 2);
    for (unsigned k = 0; k < o->N; ++k) {
      const uint32_t* d = get_neighbours(o, k);
      if (g->relax_phi)
        o->phi_tmp[k] = F(LYO_O_PHI)[k] + g->relax_dt * P->GammaP * laplacian(F(LYO_O_MU), d, SD);
      else
        o->phi_tmp[k] = F(LYO_O_PHI)[k];
      if (g->relax_Q) {
        F(LYO_O_QNXX)[k] = F(LYO_O_QQXX)[k] + g->relax_dt * P->GammaQ * F(LYO_O_HHXX)[k];
        F(LYO_O_QNYX)[k] = F(LYO_O_QQYX)[k] + g->relax_dt * P->GammaQ * F(LYO_O_HHYX)[k];
      } else {
        F(LYO_O_QNXX)[k] = F(LYO_O_QQXX)[k];
        F(LYO_O_QNYX)[k] = F(LYO_O_QQYX)[k];
      }
    }
    double* t = o->s[LYO_O_PHI]; o->s[LYO_O_PHI] = o->phi_tmp; o->phi_tmp = t;
    t = o->s[LYO_O_QQXX]; o->s[LYO_O_QQXX] = o->s[LYO_O_QNXX]; o->s[LYO_O_QNXX] = t;
    t = o->s[LYO_O_QQYX]; o->s[LYO_O_QQYX] = o->s[LYO_O_QNYX]; o->s[LYO_O_QNYX] = t;
  }
  lyo_oracle_boundary(o, 0);
  double relaxed_phi = 0.;
  for (unsigned k = 0; k < o->N; ++k) relaxed_phi += F(LYO_O_PHI)[k];
  o->totalphi = relaxed_phi;
  o->countphi = relaxed_phi;
  o->ptot = relaxed_phi;
}

/* go-or-grow.cpp:121-141 */
static void reset_hydrodynamics(lyo_oracle* o) {
  double ftotal = 0.;
  double fe[9];
  for (unsigned k = 0; k < o->N; ++k) {
    F(LYO_O_UX)[k] = F(LYO_O_UY)[k] = F(LYO_O_UX_PHI)[k] = F(LYO_O_UY_PHI)[k] = 0.;
    F(LYO_O_N)[k] = o->p.rho;
    equilibrium(0., 0., o->p.rho, fe);
    for (int v = 0; v < 9; ++v) {
      o->ff[9 * (size_t)k + v] = fe[v];
      o->fn[9 * (size_t)k + v] = fe[v];
      o->ff_tmp[9 * (size_t)k + v] = fe[v];
      o->fn_tmp[9 * (size_t)k + v] = fe[v];
    }
    for (int v = 0; v < 9; ++v) ftotal = ftotal + fe[v];
  }
  o->ftot = ftotal;
}

/* go-or-grow.cpp:143-199 */
static void configure_phenotype(lyo_oracle* o) {
  const lyo_oracle_growth* g = &o->g;
  const unsigned N = o->N;
  double* chi = NULL;
  if (g->chi_config == 0) { /* noise */
    for (unsigned k = 0; k < N; ++k) F(LYO_O_CHI)[k] = mt_normal(&o->rng);
    growth_boundary_fields(o);
    const double smooth_rate = 0.2;
    const unsigned smooth_steps =
        g->chi_length > 0 ? (unsigned)ceil(g->chi_length * g->chi_length / (2 * smooth_rate)) : 0u;
    for (unsigned s = 0; s < smooth_steps; ++s) {
      chi = F(LYO_O_CHI);
      for (unsigned k = 0; k < N; ++k) {
        const uint32_t* d = get_neighbours(o, k);
        o->m_tmp[k] = chi[k] + smooth_rate * laplacian(chi, d, SD);
      }
      o->s[LYO_O_CHI] = o->m_tmp;
      o->m_tmp = chi;
      growth_boundary_fields(o);
    }
    chi = F(LYO_O_CHI);
    double mean = 0.;
    for (unsigned k = 0; k < N; ++k) mean += chi[k];
    mean /= N;
    double variance = 0.;
    for (unsigned k = 0; k < N; ++k) {
      chi[k] -= mean;
      variance += chi[k] * chi[k];
    }
    variance /= N;
    const double amplitude = sqrt(g->chi_noise);
    const double norm = variance > 0 ? amplitude / sqrt(variance) : 0.;
    for (unsigned k = 0; k < N; ++k) chi[k] = g->chi0 + norm * chi[k];
  } else { /* front */
    const double length = g->chi_length > 0 ? g->chi_length : 1.;
    const double center = .5 * (unsigned)(o->p.LX - 1);
    chi = F(LYO_O_CHI);
    for (unsigned k = 0; k < N; ++k) {
      const double x = k / (unsigned)o->p.LY;
      chi[k] = .5 * (1. - tanh((x - center) / length));
    }
  }
  for (unsigned k = 0; k < N; ++k) F(LYO_O_M)[k] = F(LYO_O_PHI)[k] * F(LYO_O_CHI)[k];
  project_m(o);
  update_phenotype(o);
  growth_boundary_fields(o);
}

static void growth_configure(lyo_oracle* o) {
  if (o->g.division_only) return; /* LyotropicWithDivision keeps Lyotropic's Configure */
  relax_free_energy(o);
  reset_hydrodynamics(o);
  configure_phenotype(o);
}

/* lyotropic-with-division.cpp:34-108: one mask */
static void draw_mask(lyo_oracle* o, uint8_t* mask, uint32_t* count, double life, double rate,
                      double radius, int division) {
  if ((*count)--) return;
  const unsigned LX = (unsigned)o->p.LX, LY = (unsigned)o->p.LY;
  memset(mask, 0, o->N);
  *count = (uint32_t)life;
  unsigned nc = 0, cap = 64;
  unsigned* cx = (unsigned*)malloc(sizeof(unsigned) * cap);
  unsigned* cy = (unsigned*)malloc(sizeof(unsigned) * cap);
  for (unsigned k = 0; k < o->N; ++k) {
    const double p = F(LYO_O_PHI)[k];
    if (p >= .5) {
      const double adjusted = division ? rate * p * (1 - p / o->g.phi_critical) : rate * p;
      if (mt_real(&o->rng, 0., 1.) < adjusted) {
        if (nc == cap) {
          cap *= 2;
          cx = (unsigned*)realloc(cx, sizeof(unsigned) * cap);
          cy = (unsigned*)realloc(cy, sizeof(unsigned) * cap);
        }
        cx[nc] = k / LY;
        cy[nc] = k % LY;
        ++nc;
      }
    }
  }
  const double r2 = ceil(radius * radius);
  for (unsigned k = 0; k < o->N; ++k) {
    const unsigned yk = k % LY, xk = k / LY;
    for (unsigned c = 0; c < nc; ++c) {
      if (pow(uwrap(udiff(yk, cy[c]), LY), 2) + pow(uwrap(udiff(xk, cx[c]), LX), 2) <= r2) {
        mask[k] = 1;
        break;
      }
    }
  }
  free(cx);
  free(cy);
}

/* lyotropic-with-division.cpp:34-108 */
static void set_masks(lyo_oracle* o) {
  const lyo_oracle_growth* g = &o->g;
  draw_mask(o, o->div_mask, &o->div_count, g->division_time, g->division_rate,
            g->division_radius, 1);
  draw_mask(o, o->death_mask, &o->death_count, g->death_time, g->death_rate, g->death_radius, 0);
}

/* go-or-grow.cpp:224-303 */
static void growth_uq_at_node(lyo_oracle* o, unsigned k) {
  const lyo_oracle_params* P = &o->p;
  const lyo_oracle_growth* g = &o->g;
  const uint32_t* d = get_neighbours(o, k);
  const double* f = o->ff + 9 * (size_t)k;
  const double* QQxx = F(LYO_O_QQXX);
  const double* QQyx = F(LYO_O_QQYX);
  const double* phi = F(LYO_O_PHI);
  const double Qxx = QQxx[k];
  const double Qyx = QQyx[k];
  const double p = phi[k];

  const double nn = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  const double vx = (f[1] - f[2] + f[5] - f[6] - f[7] + f[8]) / nn;
  const double vy = (f[3] - f[4] + f[5] - f[6] + f[7] - f[8]) / nn;

  const double del2Qxx = laplacian(QQxx, d, SD);
  const double dxQxx = derivX(QQxx, d, SB);
  const double dyQxx = derivY(QQxx, d, SB);
  const double del2Qyx = laplacian(QQyx, d, SD);
  const double dxQyx = derivX(QQyx, d, SB);
  const double dyQyx = derivY(QQyx, d, SB);
  const double del2p = laplacian(phi, d, SD);
  const double dxPhi = derivX(phi, d, SB);
  const double dyPhi = derivY(phi, d, SB);

  const double q2 = Qxx * Qxx + Qyx * Qyx;
  const double term = g->Snem - q2;
  const double Hxx = 2 * P->CC * term * Qxx + P->LL * del2Qxx;
  const double Hyx = 2 * P->CC * term * Qyx + P->LL * del2Qyx;
  const double dp_critical = p - g->phi_critical;
  const double f_compress = dp_critical > 0 ? .5 * g->B * dp_critical * dp_critical : 0.;
  const double mu_compress = dp_critical > 0 ? g->B * dp_critical : 0.;
  const double mu = P->AA * p * (1 - p) * (1 - 2 * p) + mu_compress - P->KK * del2p;

  const double sigmaB = .5 * P->AA * p * p * (1 - p) * (1 - p) + f_compress
                        + .5 * P->CC * term * term - mu * p;
  const double active_prefactor = P->zeta * p * (1 - F(LYO_O_CHI)[k]);
  const double sigmaF = 2 * P->xi * ((Qxx * Qxx - 1) * Hxx + Qxx * Qyx * Hyx)
                        - active_prefactor * Qxx + .5 * P->KK * (dyPhi * dyPhi - dxPhi * dxPhi)
                        + P->LL * (dyQxx * dyQxx + dyQyx * dyQyx - dxQxx * dxQxx - dxQyx * dxQyx);
  const double sigmaS = 2 * P->xi * (Qyx * Qxx * Hxx + (Qyx * Qyx - 1) * Hyx)
                        - active_prefactor * Qyx - P->KK * dxPhi * dyPhi
                        - 2 * P->LL * (dxQxx * dyQxx + dxQyx * dyQyx);
  const double sigmaA = 2 * (Qxx * Hyx - Qyx * Hxx);

  F(LYO_O_UX)[k] = vx;
  F(LYO_O_UY)[k] = vy;
  F(LYO_O_UX_PHI)[k] = vx * p;
  F(LYO_O_UY_PHI)[k] = vy * p;
  F(LYO_O_N)[k] = nn;
  F(LYO_O_HHXX)[k] = Hxx;
  F(LYO_O_HHYX)[k] = Hyx;
  F(LYO_O_MU)[k] = mu;
  F(LYO_O_DXQQXX)[k] = dxQxx;
  F(LYO_O_DXQQYX)[k] = dxQyx;
  F(LYO_O_DYQQXX)[k] = dyQxx;
  F(LYO_O_DYQQYX)[k] = dyQyx;
  F(LYO_O_SIGMAXX)[k] = sigmaF + sigmaB + P->zetaI * (P->conc - p);
  F(LYO_O_SIGMAYY)[k] = -sigmaF + sigmaB + P->zetaI * (P->conc - p);
  F(LYO_O_SIGMAXY)[k] = sigmaS + sigmaA;
  F(LYO_O_SIGMAYX)[k] = sigmaS - sigmaA;
  F(LYO_O_SIGMA_BULK)[k] = sigmaB + P->zetaI * (P->conc - p);
  F(LYO_O_SIGMA_EL_XX)[k] = sigmaF - .5 * P->KK * (dyPhi * dyPhi - dxPhi * dxPhi) + active_prefactor * Qxx;
  F(LYO_O_SIGMA_EL_YX)[k] = sigmaS + P->KK * dxPhi * dyPhi + active_prefactor * Qyx;
  F(LYO_O_SIGMA_PF_XX)[k] = .5 * P->KK * (dyPhi * dyPhi - dxPhi * dxPhi);
  F(LYO_O_SIGMA_PF_YX)[k] = -P->KK * dxPhi * dyPhi;
  F(LYO_O_SIGMA_ACT_XX)[k] = -active_prefactor * Qxx;
  F(LYO_O_SIGMA_ACT_YX)[k] = -active_prefactor * Qyx;
}

static void update_fields_at_node(lyo_oracle* o, unsigned k, int first);

/* go-or-grow.cpp:322-362: the phenotype density's increment, Lyotropic's
 * update, then division/death and the bounded m */
static void growth_uf_at_node(lyo_oracle* o, unsigned k, int first) {
  const lyo_oracle_params* P = &o->p;
  const lyo_oracle_growth* g = &o->g;
  const uint32_t* d = get_neighbours(o, k);
  const double* phi = F(LYO_O_PHI);
  const double* chi = F(LYO_O_CHI);
  const double* m = F(LYO_O_M);
  const double* ux = F(LYO_O_UX);
  const double* uy = F(LYO_O_UY);
  const double* MU = F(LYO_O_MU);

  const double division_m = o->div_mask[k] ? 1. : 0.;
  const double death_m = o->death_mask[k] ? 1. : 0.;
  const double growth_rate = g->alpha * division_m - g->beta * death_m;
  const double chi_eff = chi[k];
  const double R = chi_eff * phi[k] * growth_rate;

  const double mFlux = .5 * (ux[d[1]] * m[d[1]] - ux[d[2]] * m[d[2]])
                     + .5 * (uy[d[3]] * m[d[3]] - uy[d[4]] * m[d[4]]);
  const double chi_px = chi[d[1]];
  const double chi_mx = chi[d[2]];
  const double chi_py = chi[d[3]];
  const double chi_my = chi[d[4]];
  const double diffusiveMFlux = .5 * (chi_px + chi_eff) * (MU[d[1]] - MU[k])
                              - .5 * (chi_eff + chi_mx) * (MU[k] - MU[d[2]])
                              + .5 * (chi_py + chi_eff) * (MU[d[3]] - MU[k])
                              - .5 * (chi_eff + chi_my) * (MU[k] - MU[d[4]]);
  const double phenotypeDiffusion = .5 * (phi[d[1]] + phi[k]) * (chi_px - chi_eff)
                                  - .5 * (phi[k] + phi[d[2]]) * (chi_eff - chi_mx)
                                  + .5 * (phi[d[3]] + phi[k]) * (chi_py - chi_eff)
                                  - .5 * (phi[k] + phi[d[4]]) * (chi_eff - chi_my);
  const double dVdChi = 2 * g->Achi * chi_eff * (1 - chi_eff) * (1 - 2 * chi_eff)
                      + g->Ochi * (phi[k] - g->phiswitch);
  const double Sswitch = -phi[k] * dVdChi;
  const double mGrowth = g->growTogether ? chi_eff * R : R;
  const double Dm = P->GammaP * diffusiveMFlux + g->Dchi * phenotypeDiffusion - mFlux + Sswitch
                    + mGrowth;

  update_fields_at_node(o, k, first);
  o->phi_tmp[k] += R;
  if (first) {
    F(LYO_O_MN)[k] = m[k] + .5 * Dm;
    o->m_tmp[k] = m[k] + Dm;
  } else {
    o->m_tmp[k] = F(LYO_O_MN)[k] + .5 * Dm;
  }
  const double upper = o->phi_tmp[k] > 0 ? o->phi_tmp[k] : 0.;
  if (o->m_tmp[k] < 0)
    o->m_tmp[k] = 0;
  else if (o->m_tmp[k] > upper)
    o->m_tmp[k] = upper;
}

/* src/models/lyotropic.cpp:343-360 */
void lyo_oracle_update_quantities(lyo_oracle* o) {
  double sum = 0;
  const int nt = o->p.nthreads;
  const unsigned N = o->N;
  (void)nt;
#pragma omp parallel for reduction(+ : sum) num_threads(nt) if (nt)
  for (unsigned k = 0; k < N; ++k) {
    sum = sum + F(LYO_O_PHI)[k];
    if (o->growth && !o->g.division_only)
      growth_uq_at_node(o, k);
    else
      update_quantities_at_node(o, k);
  }
  o->countphi = sum;
  if (o->growth && !o->g.division_only) update_phenotype(o); /* go-or-grow.cpp:305-320 */
}

/* src/models/lyotropic.cpp:362-369 */
void lyo_oracle_update_fields(lyo_oracle* o, int first) {
  const int nt = o->p.nthreads;
  const unsigned N = o->N;
  (void)nt;
  if (o->growth && o->g.division_only) {
    /* lyotropic-with-division.cpp:110-126 */
#pragma omp parallel for num_threads(nt) if (nt)
    for (unsigned k = 0; k < N; ++k) {
      const double division_m = o->div_mask[k] ? 1. : 0.;
      const double death_m = o->death_mask[k] ? 1. : 0.;
      update_fields_at_node(o, k, first);
      o->phi_tmp[k] += F(LYO_O_PHI)[k] * (o->g.alpha * division_m - o->g.beta * death_m);
    }
  } else if (o->growth) {
#pragma omp parallel for num_threads(nt) if (nt)
    for (unsigned k = 0; k < N; ++k) growth_uf_at_node(o, k, first);
  } else {
#pragma omp parallel for num_threads(nt) if (nt)
    for (unsigned k = 0; k < N; ++k) update_fields_at_node(o, k, first);
  }
  double* t = o->s[LYO_O_PHI];
  o->s[LYO_O_PHI] = o->phi_tmp;
  o->phi_tmp = t;
  if (o->growth && !o->g.division_only) { /* go-or-grow.cpp:358-362 */
    t = o->s[LYO_O_M];
    o->s[LYO_O_M] = o->m_tmp;
    o->m_tmp = t;
    update_phenotype(o);
  }
}

/* src/models/lyotropic.cpp:371-386: every node of the grid, ghost layer
 * included */
void lyo_oracle_move(lyo_oracle* o) {
  const int nt = o->p.nthreads;
  const unsigned T = o->T;
  (void)nt;
#pragma omp parallel for num_threads(nt) if (nt)
  for (unsigned k = 0; k < T; ++k) {
    const uint32_t* d = get_neighbours(o, k);
    for (int v = 0; v < 9; ++v) {
      o->ff_tmp[9 * (size_t)d[v] + v] = o->ff[9 * (size_t)k + v];
      o->fn_tmp[9 * (size_t)d[v] + v] = o->fn[9 * (size_t)k + v];
    }
  }
  double* t = o->ff; o->ff = o->ff_tmp; o->ff_tmp = t;
  t = o->fn; o->fn = o->fn_tmp; o->fn_tmp = t;
}

/* ------------------------------------------------ boundary conditions */

/* src/geometry.hpp:148-199: walls and corners carry the number of the
 * direction that points out of the domain through them */
enum { W_RIGHT = 1, W_LEFT = 2, W_BACK = 3, W_FRONT = 4 };
enum { C_RIGHTBACK = 5, C_LEFTFRONT = 6, C_LEFTBACK = 7, C_RIGHTFRONT = 8 };

/* src/geometry.hpp:40-59 (reflection about the plane with normal i; row 0:
 * bounce back) and :65-84 (projection onto that plane) */
static const int REFL[9][9] = {
    {0, 2, 1, 4, 3, 6, 5, 8, 7}, {0, 2, 1, 3, 4, 7, 8, 5, 6}, {0, 2, 1, 3, 4, 7, 8, 5, 6},
    {0, 1, 2, 4, 3, 8, 7, 6, 5}, {0, 1, 2, 4, 3, 8, 7, 6, 5}, {0, 4, 3, 2, 1, 6, 5, 7, 8},
    {0, 4, 3, 2, 1, 6, 5, 7, 8}, {0, 3, 4, 1, 2, 5, 6, 8, 7}, {0, 3, 4, 1, 2, 5, 6, 8, 7}};
static const int PROJ[9][9] = {
    {0, 0, 0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 3, 4, 3, 4, 3, 4}, {0, 0, 0, 3, 4, 3, 4, 3, 4},
    {0, 1, 2, 0, 0, 1, 2, 2, 1}, {0, 1, 2, 0, 0, 1, 2, 2, 1}, {0, 5, 6, 5, 6, 5, 6, 0, 0},
    {0, 1, 2, 4, 3, 8, 7, 6, 5}, {0, 8, 7, 7, 8, 0, 0, 7, 8}, {0, 8, 7, 7, 8, 0, 0, 7, 8}};

static inline int opp(int d) { return REFL[0][d]; }
static inline uint32_t nxt(const lyo_oracle* o, uint32_t k, int d) {
  return get_neighbours(o, k)[d];
}
/* src/geometry.hpp:475-521 */
static uint32_t wall_size(const lyo_oracle* o, int w) {
  return (w == W_LEFT || w == W_RIGHT) ? (uint32_t)o->p.LY : (uint32_t)o->p.LX;
}
static uint32_t wall_index(const lyo_oracle* o, int w, uint32_t i) {
  const uint32_t LX = (uint32_t)o->p.LX, LY = (uint32_t)o->p.LY;
  switch (w) {
    case W_LEFT: return o->N + i;
    case W_RIGHT: return o->N + LY + i;
    case W_FRONT: return o->N + 2 * LY + i;
    default: return o->N + 2 * LY + LX + i;
  }
}
static uint32_t corner_index(const lyo_oracle* o, int c) {
  switch (c) {
    case C_LEFTFRONT: return o->T - 4;
    case C_RIGHTFRONT: return o->T - 3;
    case C_LEFTBACK: return o->T - 2;
    default: return o->T - 1;
  }
}
/* src/geometry.hpp:312-316 */
static inline int is_incoming(int v, int w) {
  return XDIR[w] * XDIR[v] == -1 || YDIR[w] * YDIR[v] == -1;
}

/* scalar fields, src/fields.hpp:619-625 (PBC wall), 704-747 (Neumann,
 * Dirichlet), 738-768 (copy derivative) */
static void s_pbc_wall(const lyo_oracle* o, double* a, int w) {
  for (uint32_t i = 0; i < wall_size(o, w); ++i)
    a[wall_index(o, w, i)] = a[nxt(o, wall_index(o, opp(w), i), w)];
}
static void s_neumann_wall(const lyo_oracle* o, double* a, int w) {
  for (uint32_t i = 0; i < wall_size(o, w); ++i) {
    const uint32_t k = wall_index(o, w, i);
    a[k] = a[nxt(o, k, opp(w))];
  }
}
static void s_dirichlet_wall(const lyo_oracle* o, double* a, int w, double value) {
  for (uint32_t i = 0; i < wall_size(o, w); ++i) {
    const uint32_t k = wall_index(o, w, i);
    a[k] = 2 * value - a[nxt(o, k, opp(w))];
  }
}
static void s_copyder_wall(const lyo_oracle* o, double* a, int w) {
  for (uint32_t i = 0; i < wall_size(o, w); ++i) {
    const uint32_t k = wall_index(o, w, i);
    const uint32_t j = nxt(o, k, opp(w));
    a[k] = 2 * a[j] - a[nxt(o, j, opp(w))];
  }
}
/* corners: towards d = opp(c) (both walls closed) or opp(w) (w closed, the
 * other wall periodic) */
static void s_neumann_corner(const lyo_oracle* o, double* a, int c, int d) {
  const uint32_t k = corner_index(o, c);
  a[k] = a[nxt(o, k, d)];
}
static void s_dirichlet_corner(const lyo_oracle* o, double* a, int c, int d, double value) {
  const uint32_t k = corner_index(o, c);
  a[k] = 2 * value - a[nxt(o, k, d)];
}
static void s_copyder_corner(const lyo_oracle* o, double* a, int c, int d) {
  const uint32_t k = corner_index(o, c);
  const uint32_t j = nxt(o, k, d);
  a[k] = 2 * a[j] - a[nxt(o, j, d)];
}

/* kind 0 Neumann, 1 Dirichlet (value), 2 copy derivative; box:
 * fields.hpp:770-780,805-829, channel: 844-885 */
static void s_apply(const lyo_oracle* o, double* a, int kind, double value, int channel) {
  static const int box_corners[4] = {C_LEFTFRONT, C_RIGHTFRONT, C_LEFTBACK, C_RIGHTBACK};
  static const int ch_corners[4][2] = {{C_RIGHTBACK, W_BACK}, {C_RIGHTFRONT, W_FRONT},
                                       {C_LEFTBACK, W_BACK}, {C_LEFTFRONT, W_FRONT}};
  if (channel) {
    s_pbc_wall(o, a, W_LEFT);
    s_pbc_wall(o, a, W_RIGHT);
  }
  const int walls[4] = {W_LEFT, W_RIGHT, W_FRONT, W_BACK};
  for (int i = channel ? 2 : 0; i < 4; ++i) {
    if (kind == 0) s_neumann_wall(o, a, walls[i]);
    else if (kind == 1) s_dirichlet_wall(o, a, walls[i], value);
    else s_copyder_wall(o, a, walls[i]);
  }
  for (int i = 0; i < 4; ++i) {
    const int c = channel ? ch_corners[i][0] : box_corners[i];
    const int d = channel ? opp(ch_corners[i][1]) : opp(c);
    if (kind == 0) s_neumann_corner(o, a, c, d);
    else if (kind == 1) s_dirichlet_corner(o, a, c, d, value);
    else s_copyder_corner(o, a, c, d);
  }
}

/* LB fields (AoS, 9 per node), src/fields.hpp:619-625 (PBC wall: the whole
 * node), 647-701 (free slip, no slip) */
static void lb_pbc_wall(const lyo_oracle* o, double* f, int w) {
  for (uint32_t i = 0; i < wall_size(o, w); ++i) {
    const uint32_t k = wall_index(o, w, i);
    const uint32_t src = nxt(o, wall_index(o, opp(w), i), w);
    for (int v = 0; v < 9; ++v) f[9 * (size_t)k + v] = f[9 * (size_t)src + v];
  }
}
static void lb_free_slip_wall(const lyo_oracle* o, double* f, int w) {
  for (uint32_t i = 0; i < wall_size(o, w); ++i) {
    const uint32_t k = wall_index(o, w, i);
    for (int v = 0; v < 9; ++v)
      if (is_incoming(v, w))
        f[9 * (size_t)k + v] = f[9 * (size_t)nxt(o, k, opp(w)) + REFL[w][v]];
  }
}
static void lb_no_slip_wall(const lyo_oracle* o, double* f, int w) {
  for (uint32_t i = 0; i < wall_size(o, w); ++i) {
    const uint32_t k = wall_index(o, w, i);
    for (int v = 0; v < 9; ++v)
      if (is_incoming(v, w)) f[9 * (size_t)k + v] = f[9 * (size_t)nxt(o, k, v) + opp(v)];
  }
}
/* fields.hpp:689-701: bounce back at a corner (no slip, or free slip on
 * both walls) */
static void lb_bounce_corner(const lyo_oracle* o, double* f, int c) {
  const uint32_t k = corner_index(o, c);
  f[9 * (size_t)k + opp(c)] = f[9 * (size_t)nxt(o, k, opp(c)) + c];
}
/* fields.hpp:675-687: free slip on wall w next to a periodic wall */
static void lb_free_slip_corner(const lyo_oracle* o, double* f, int c, int w) {
  const int p = PROJ[w][c];
  const uint32_t k = corner_index(o, c);
  const int v = opp(c);
  const uint32_t l = nxt(o, corner_index(o, REFL[p][c]), p);
  f[9 * (size_t)k + v] = f[9 * (size_t)l + v];
}

/* bc 1 FreeSlipChannel, 2 NoSlipChannel, 3 FreeSlip, 4 NoSlip
 * (fields.hpp:781-803,887-918) */
static void lb_apply(const lyo_oracle* o, double* f, int bc) {
  const int channel = bc <= 2, slip = (bc == 1 || bc == 3);
  if (channel) {
    lb_pbc_wall(o, f, W_LEFT);
    lb_pbc_wall(o, f, W_RIGHT);
  }
  const int walls[4] = {W_LEFT, W_RIGHT, W_FRONT, W_BACK};
  for (int i = channel ? 2 : 0; i < 4; ++i) {
    if (slip) lb_free_slip_wall(o, f, walls[i]);
    else lb_no_slip_wall(o, f, walls[i]);
  }
  if (bc == 1) {
    lb_free_slip_corner(o, f, C_RIGHTBACK, W_BACK);
    lb_free_slip_corner(o, f, C_RIGHTFRONT, W_FRONT);
    lb_free_slip_corner(o, f, C_LEFTBACK, W_BACK);
    lb_free_slip_corner(o, f, C_LEFTFRONT, W_FRONT);
  } else if (bc == 2) {
    lb_bounce_corner(o, f, C_RIGHTBACK);
    lb_bounce_corner(o, f, C_RIGHTFRONT);
    lb_bounce_corner(o, f, C_LEFTBACK);
    lb_bounce_corner(o, f, C_LEFTFRONT);
  } else {
    lb_bounce_corner(o, f, C_LEFTFRONT);
    lb_bounce_corner(o, f, C_RIGHTFRONT);
    lb_bounce_corner(o, f, C_LEFTBACK);
    lb_bounce_corner(o, f, C_RIGHTBACK);
  }
}

/* src/models/lyotropic.cpp:388-498 */
void lyo_oracle_boundary(lyo_oracle* o, int which) {
  const int bc = o->p.bc;
  if (!bc) return;
  const int channel = bc <= 2;
  if (which == 0) {
    s_apply(o, F(LYO_O_QQXX), 0, 0., channel);
    s_apply(o, F(LYO_O_QQYX), 0, 0., channel);
    s_apply(o, F(LYO_O_PHI), 0, 0., channel);
  } else if (which == 1) {
    s_apply(o, F(LYO_O_UY), 2, 0., channel);
    s_apply(o, F(LYO_O_UX), 2, 0., channel);
    s_apply(o, F(LYO_O_UY_PHI), 1, 0., channel);
    s_apply(o, F(LYO_O_UX_PHI), 1, 0., channel);
    s_apply(o, F(LYO_O_MU), 0, 0., channel);
    s_apply(o, F(LYO_O_SIGMAXX), 0, 0., channel);
    s_apply(o, F(LYO_O_SIGMAYY), 0, 0., channel);
    s_apply(o, F(LYO_O_SIGMAYX), 0, 0., channel);
    s_apply(o, F(LYO_O_SIGMAXY), 0, 0., channel);
  } else {
    lb_apply(o, o->ff, bc);
    lb_apply(o, o->fn, bc);
  }
}

/* src/models/lyotropic.cpp:500-525 (bc=0: the boundary-condition calls are
 * no-ops) */
void lyo_oracle_step(lyo_oracle* o, int nsteps) {
  for (int s = 0; s < nsteps; ++s) {
    if (o->growth) set_masks(o); /* lyotropic-with-division.cpp:150-156 */
    growth_boundary_fields(o);
    lyo_oracle_update_quantities(o);
    lyo_oracle_boundary(o, 1);
    lyo_oracle_update_fields(o, 1);
    lyo_oracle_boundary(o, 2);
    lyo_oracle_move(o);
    for (int n = 1; n <= o->p.npc; ++n) {
      growth_boundary_fields(o);
      lyo_oracle_update_quantities(o);
      lyo_oracle_boundary(o, 1);
      lyo_oracle_update_fields(o, 0);
    }
  }
}

/* src/models/lyotropic.cpp:527-549 */
void lyo_oracle_sums(const lyo_oracle* o, double* fsum, double* phisum,
                     double* ftot, double* ptot, double* totalphi,
                     double* countphi) {
  double fcheck = 0, pcheck = 0;
  for (unsigned k = 0; k < o->N; ++k)
    for (int v = 0; v < 9; ++v) fcheck = fcheck + o->ff[9 * (size_t)k + v];
  for (unsigned k = 0; k < o->N; ++k) pcheck += o->s[LYO_O_PHI][k];
  if (fsum) *fsum = fcheck;
  if (phisum) *phisum = pcheck;
  if (ftot) *ftot = o->ftot;
  if (ptot) *ptot = o->ptot;
  if (totalphi) *totalphi = o->totalphi;
  if (countphi) *countphi = o->countphi;
}
