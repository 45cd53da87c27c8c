// go_or_grow_hip.cpp -- see go_or_grow_hip.hpp.
#include "go_or_grow_hip.hpp"

#include "error_msg.hpp"

void GoOrGrowHip::growth_params(mass_lyo_growth& g) const {
  LyotropicWithDivisionHip::growth_params(g);
  g.division_only = 0;
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
}

option_list GoOrGrowHip::GetOptions() {
  option_list options = LyotropicWithDivisionHip::GetOptions();
  // go-or-grow.cpp:387-423
  options[0]
      .add_double("B", &B, "crowding/compressibility penalty strength")
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
