// lyo_config.cpp -- initial fields on the host (Lyotropic::ConfigureAtNode,
// src/models/lyotropic.cpp:69-166).
//
// The initial state is drawn on the host, as in the reference, so that it is
// bit-identical: the same std::mt19937 stream (src/random.cpp:12,43-46), one
// uniform_real_distribution<> draw per node for the director angle
// (random_real(), src/random.cpp:21-24) in ascending node order, and the host
// libm cos/sin.  The random configuration adds a second draw per node for phi,
// right after the angle.
#include <cmath>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "lyo_config.hpp"

namespace lyo {

namespace {
unsigned udiff(unsigned a, unsigned b) { return a > b ? a - b : b - a; }  // tools.hpp:36
unsigned uwrap(unsigned x, unsigned L) {                                   // tools.hpp:60-63
  return std::min(x % L, L - x % L);
}
}  // namespace

int configure_host(const HostInit& in, std::mt19937& gen, std::vector<double>& ff,
                   std::vector<double>& qxx, std::vector<double>& qyx, std::vector<double>& phi,
                   double* ftot, double* ptot, std::string& err) {
  const unsigned LX = in.LX, LY = in.LY, N = LX * LY;
  ff.assign(9 * (size_t)N, 0.0);
  qxx.assign(N, 0.0);
  qyx.assign(N, 0.0);
  phi.assign(N, 0.0);
  gen.seed(in.seed);
  auto random_real = [&](double lo, double hi) {
    return std::uniform_real_distribution<>(lo, hi)(gen);
  };
  const double angle = in.angle_deg * M_PI / 180.;
  const double level = in.level, radius = in.radius;
  double ft = 0, pt = 0;
  for (unsigned k = 0; k < N; ++k) {
    double order = 0;
    const unsigned x = k / LY, y = k % LY;
    switch (in.config) {
      case 0:  // circle
        if (std::pow(udiff(x, LX / 2), 2) + std::pow(udiff(y, LY / 2), 2) <= radius * radius)
          order = 1;
        break;
      case 1:  // square
        if (udiff(LY / 2, y) < level / 2 && udiff(LX / 2, x) < level / 2) order = 1;
        break;
      case 2:  // stripe
        if (udiff(LY / 2, y) < level / 2) order = 1;
        break;
      case 3:  // half
        if (y < level) order = 1;
        break;
      case 4:  // random
        order = 1;
        break;
      case 5:  // boxatwall
        if (in.bc == 4) {
          if (x < level && udiff(LY / 2, y) < level / 2) order = 1;
        } else if (y < level && udiff(LX / 2, x) < level) {
          order = 1;
        }
        break;
      case 6:  // cuttingatwall (unsigned arithmetic, as the reference's)
        if (in.bc == 4) {
          const unsigned off = (unsigned)int(radius * 1.5);
          if (std::pow(x, 2) + std::pow(uwrap(udiff(y, LY / 2 + off), LY), 2) <= radius * radius)
            order = 1;
          if (std::pow(x, 2) + std::pow(uwrap(udiff(y, LY / 2 - off), LY), 2) <= radius * radius)
            order = 1;
        } else if (y < level && udiff(LX / 2, x) < level) {
          order = 1;
        }
        break;
      case 7:  // circleatwall (periodic branch)
        if (std::pow(uwrap(udiff(x, LX / 2), LX), 2) + std::pow(udiff(y, 0 * LY), 2) <=
            radius * radius)
          order = 1;
        break;
      case 8:  // wettedwall (periodic branch)
        if (y < level) order = 1;
        break;
      default:
        err = "error: initial configuration '" + std::to_string(in.config) + "' unknown.";
        return -1;
    }
    const double theta = angle + in.noise * M_PI * (random_real(0., 1.) - .5);
    qxx[k] = in.qamp * order * (std::cos(2 * theta));
    qyx[k] = in.qamp * order * (std::sin(2 * theta));
    if (in.config == 4)
      phi[k] = in.conc + random_real(-in.phinoise, in.phinoise);
    else
      phi[k] = order * in.conc;
    // equilibrium at u = 0, n = rho (src/lb.hpp:24-41 with vx = vy = 0)
    double* f = &ff[9 * (size_t)k];
    const double nn = in.rho, vx = 0, vy = 0;
    const double v2 = vx * vx + vy * vy;
    f[1] = nn * (1. / 9. + 1. / 3. * vx + 1. / 3. * v2 - 1. / 2. * vy * vy);
    f[2] = nn * (1. / 9. - 1. / 3. * vx + 1. / 3. * v2 - 1. / 2. * vy * vy);
    f[3] = nn * (1. / 9. + 1. / 3. * vy - 1. / 2. * vx * vx + 1. / 3. * v2);
    f[4] = nn * (1. / 9. - 1. / 3. * vy - 1. / 2. * vx * vx + 1. / 3. * v2);
    f[5] = nn * (1. / 36. + 1. / 12. * vx + 1. / 12. * vy + 1. / 12. * v2 + 1. / 4. * vx * vy);
    f[6] = nn * (1. / 36. - 1. / 12. * vx - 1. / 12. * vy + 1. / 12. * v2 + 1. / 4. * vx * vy);
    f[7] = nn * (1. / 36. - 1. / 12. * vx + 1. / 12. * vy + 1. / 12. * v2 - 1. / 4. * vx * vy);
    f[8] = nn * (1. / 36. + 1. / 12. * vx - 1. / 12. * vy + 1. / 12. * v2 - 1. / 4. * vx * vy);
    f[0] = nn - f[1] - f[2] - f[3] - f[4] - f[5] - f[6] - f[7] - f[8];
    for (int v = 0; v < 9; ++v) ft = ft + f[v];  // std::accumulate order
    pt += phi[k];
  }
  if (ftot) *ftot = ft;
  if (ptot) *ptot = pt;
  return 0;
}

}  // namespace lyo
