// go_or_grow_hip.hpp -- the go-or-grow model on the GPU, behind the Model
// interface.
//
// Drop-in for the reference's GoOrGrow (src/models/go-or-grow.{hpp,cpp}, on
// LyotropicWithDivision, src/models/lyotropic-with-division.{hpp,cpp}): the
// same runcard options (go-or-grow.cpp:387-423, lyotropic-with-division.cpp:
// 128-148), parameter and frame records (go-or-grow.hpp:58-90,
// lyotropic-with-division.hpp:57-85) and runtime checks.  The run is a
// lyotropic run of the C ABI made go-or-grow by mass_lyo_set_growth.
#pragma once

#include <vector>

#include "lyotropic_hip.hpp"

class GoOrGrowHip : public LyotropicHip {
 protected:
  // lyotropic-with-division.hpp:19-34 (no defaults in the reference: every
  // runcard sets them; 0 here)
  double alpha = 0, beta = 0, division_rate = 0, division_time = 0, division_radius = 0;
  double death_rate = 0, death_time = 0, death_radius = 0, phi_critical = 0;
  // go-or-grow.hpp:10-28
  double B = 0., Snem = 1., Dchi = 0., Achi = 0., Ochi = 0., phiswitch = 0.;
  int growTogether = 0;
  double chi0 = 0., chi_noise = 0., chi_length = 0.;
  std::string chi_config = "noise";
  unsigned relax_steps = 0;
  double relax_dt = 1.;
  int relax_phi = 0, relax_Q = 0;
  bool conserve_phi = false;

  std::vector<bool> mask(int id) const;

 public:
  GoOrGrowHip(unsigned LX, unsigned LY, unsigned BC) : LyotropicHip(LX, LY, BC) {}

  void Initialize() override;
  void RuntimeChecks() override;
  option_list GetOptions() override;

  template <class Archive>
  void serialize_params(Archive& ar) {
    LyotropicHip::serialize_params(ar);
    ar.add("alpha", alpha)
        .add("beta", beta)
        .add("division_rate", division_rate)
        .add("division_time", division_time)
        .add("division_radius", division_radius)
        .add("death_rate", death_rate)
        .add("death_time", death_time)
        .add("death_radius", death_radius)
        .add("phi_critical", phi_critical)
        .add("B", B)
        .add("Snem", Snem)
        .add("Dchi", Dchi)
        .add("Achi", Achi)
        .add("Ochi", Ochi)
        .add("phiswitch", phiswitch)
        .add("growTogether", growTogether)
        .add("chi0", chi0)
        .add("chi_noise", chi_noise)
        .add("chi_length", chi_length)
        .add("chi_config", chi_config)
        .add("relax_steps", relax_steps)
        .add("relax_dt", relax_dt)
        .add("relax_phi", relax_phi)
        .add("relax_Q", relax_Q);
  }

  template <class Archive>
  void serialize_frame(Archive& ar) {
    LyotropicHip::serialize_frame(ar);
    ar.add("division_mask", mask(MASS_LYO_DIVISION_MASK))
        .add("death_mask", mask(MASS_LYO_DEATH_MASK))
        .add("m", field(MASS_LYO_M))
        .add("chi", field(MASS_LYO_CHI));
  }
};
