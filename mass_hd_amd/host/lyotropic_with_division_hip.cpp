// lyotropic_with_division_hip.cpp -- see lyotropic_with_division_hip.hpp.
#include "lyotropic_with_division_hip.hpp"

#include <cmath>

#include "error_msg.hpp"

void LyotropicWithDivisionHip::growth_params(mass_lyo_growth& g) const {
  g.division_only = 1;
  g.alpha = alpha;
  g.beta = beta;
  g.division_rate = division_rate;
  g.division_time = division_time;
  g.division_radius = division_radius;
  g.death_rate = death_rate;
  g.death_time = death_time;
  g.death_radius = death_radius;
  g.phi_critical = phi_critical;
}

void LyotropicWithDivisionHip::Initialize() {
  LyotropicHip::Initialize();
  if (precision != "fp64") throw error_msg("division models run in fp64 only.");
  mass_lyo_growth g{};
  g.Snem = 1.;
  g.relax_dt = 1.;
  growth_params(g);
  if (mass_lyo_set_growth(h_, &g) != MASS_LYO_OK) throw error_msg(mass_lyo_last_error(h_));
  // lyotropic-with-division.cpp:31
  conserve_phi = (alpha == 0 || division_rate == 0) && (beta == 0 || death_rate == 0);
}

void LyotropicWithDivisionHip::RuntimeChecks() {
  // lyotropic.cpp:527-549: phi is checked when it is conserved
  LyotropicHip::RuntimeChecks();
  if (conserve_phi) {
    double fcheck = 0, pcheck = 0;
    if (mass_lyo_sums(h_, &fcheck, &pcheck) != MASS_LYO_OK)
      throw error_msg(mass_lyo_last_error(h_));
    if (std::abs(ptot - pcheck) > 1)
      throw error_msg("phi is not conserved (", ptot, "/", pcheck, ")");
  }
}

std::vector<bool> LyotropicWithDivisionHip::mask(int id) const {
  const std::vector<double> m = field(id);
  std::vector<bool> out(m.size());
  for (size_t k = 0; k < m.size(); ++k) out[k] = m[k] != 0;
  return out;
}

option_list LyotropicWithDivisionHip::GetOptions() {
  option_list options = LyotropicHip::GetOptions();
  // lyotropic-with-division.cpp:128-148
  options[0]
      .add_double("alpha", &alpha, "division strength")
      .add_double("beta", &beta, "death strength")
      .add_double("division-rate", &division_rate, "division rate")
      .add_double("division-time", &division_time, "division life time")
      .add_double("division-radius", &division_radius, "division radius")
      .add_double("death-rate", &death_rate, "death rate")
      .add_double("death-time", &death_time, "death life time")
      .add_double("death-radius", &death_radius, "death radius")
      .add_double("phi-critical", &phi_critical, "phi critical");
  return options;
}
