// lyo_config.hpp -- host-side initial configuration.
#pragma once

#include <cstdint>
#include <random>
#include <string>
#include <vector>

namespace lyo {

struct HostInit {
  unsigned LX, LY;
  int config;
  uint32_t seed;
  double level, radius, angle_deg, noise, qamp, phinoise;
  double conc, rho;
  int bc;  // boundary conditions: bc 4 places boxatwall / cuttingatwall at x = 0
};

// Fills ff (N*9 AoS), qxx, qyx, phi; returns the check totals ftot, ptot.
// gen is seeded with in.seed and left where the draws end (go-or-grow draws
// on from there).
int configure_host(const HostInit& in, std::mt19937& gen, std::vector<double>& ff,
                   std::vector<double>& qxx, std::vector<double>& qyx, std::vector<double>& phi,
                   double* ftot, double* ptot, std::string& err);

}  // namespace lyo
