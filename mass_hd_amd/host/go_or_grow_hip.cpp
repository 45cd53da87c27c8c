// go_or_grow_hip.cpp -- see go_or_grow_hip.hpp.
#include "go_or_grow_hip.hpp"

#include <cmath>
#include <iostream>

#include "error_msg.hpp"

void GoOrGrowHip::Initialize() {
  LyotropicHip::Initialize();
  if (precision != "fp64") throw error_msg("go-or-grow runs in fp64 only.");
  mass_lyo_growth g{};
  if (chi_config == "noise")
    g.chi_config = 0;
  else if (chi_config == "front")
    g.chi_config = 1;
  else
    throw error_msg("error: chi configuration '", chi_config, "' unknown.");
  g.growTogether = growTogether;
  g.relax_phi = relax_phi;
  g.relax_Q = relax_Q;
  g.relax_steps = relax_steps;
  g.B = B;
  g.Snem = Snem;
  g.Dchi = Dchi;
  g.Achi = Achi;
  g.Ochi = Ochi;
  g.phiswitch = phiswitch;
  g.chi0 = chi0;
  g.chi_noise = chi_noise;
  g.chi_length = chi_length;
  g.relax_dt = relax_dt;
  g.alpha = alpha;
  g.beta = beta;
  g.division_rate = division_rate;
  g.division_time = division_time;
  g.division_radius = division_radius;
  g.death_rate = death_rate;
  g.death_time = death_time;
  g.death_radius = death_radius;
  g.phi_critical = phi_critical;
  if (mass_lyo_set_growth(h_, &g) != MASS_LYO_OK) throw error_msg(mass_lyo_last_error(h_));
  // lyotropic-with-division.cpp:31
  conserve_phi = (alpha == 0 || division_rate == 0) && (beta == 0 || death_rate == 0);
}

void GoOrGrowHip::RuntimeChecks() {
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

std::vector<bool> GoOrGrowHip::mask(int id) const {
  const std::vector<double> m = field(id);
  std::vector<bool> out(m.size());
  for (size_t k = 0; k < m.size(); ++k) out[k] = m[k] != 0;
  return out;
}

option_list GoOrGrowHip::GetOptions() {
  option_list options = LyotropicHip::GetOptions();
  OptionGroup& model = options[0];
  // lyotropic-with-division.cpp:128-148
  model.add_double("alpha", &alpha, "division strength")
      .add_double("beta", &beta, "death strength")
      .add_double("division-rate", &division_rate, "division rate")
      .add_double("division-time", &division_time, "division life time")
      .add_double("division-radius", &division_radius, "division radius")
      .add_double("death-rate", &death_rate, "death rate")
      .add_double("death-time", &death_time, "death life time")
      .add_double("death-radius", &death_radius, "death radius")
      .add_double("phi-critical", &phi_critical, "phi critical");
  // go-or-grow.cpp:387-423
  model.add_double("B", &B, "crowding/compressibility penalty strength")
      .add_double("Snem", &Snem, "preferred value of one half Tr(Q^2)")
      .add_double("Dchi", &Dchi, "phenotype diffusion coefficient")
      .add_double("Achi", &Achi, "phenotype switching double-well barrier")
      .add_double("Ochi", &Ochi, "phenotype switching bias strength")
      .add_double("phiswitch", &phiswitch, "phi threshold for phenotype switching bias")
      .add_int("growTogether", &growTogether, "m growth source: 0 uses R, 1 uses chi*R")
      .add_string("chi-config", &chi_config, "phenotype initialization mode: noise or front")
      .add_double("chi0", &chi0, "initial phenotype mean")
      .add_double("chi-noise", &chi_noise, "initial phenotype variance")
      .add_double("chi-length", &chi_length, "initial phenotype correlation length")
      .add_uint("relax-steps", &relax_steps,
                "dry free-energy relaxation steps before official dynamics")
      .add_double("relax-dt", &relax_dt, "time step multiplier for dry free-energy relaxation")
      .add_int("relax-phi", &relax_phi,
               "relax phi by Cahn-Hilliard free-energy descent before official dynamics")
      .add_int("relax-Q", &relax_Q,
               "relax Q by molecular-field free-energy descent before official dynamics");
  return options;
}
