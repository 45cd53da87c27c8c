// go_or_grow_hip.hpp -- the go-or-grow model on the GPU, behind the Model
// interface.
//
// Drop-in for the reference's GoOrGrow (src/models/go-or-grow.{hpp,cpp}, on
// LyotropicWithDivision): the same runcard options (go-or-grow.cpp:387-423
// and the division model's), parameter and frame records
// (go-or-grow.hpp:58-90) and runtime checks.
#pragma once

#include <string>

#include "lyotropic_with_division_hip.hpp"

class GoOrGrowHip : public LyotropicWithDivisionHip {
 protected:
  // go-or-grow.hpp:10-28
  double B = 0., Snem = 1., Dchi = 0., Achi = 0., Ochi = 0., phiswitch = 0.;
  int growTogether = 0;
  double chi0 = 0., chi_noise = 0., chi_length = 0.;
  std::string chi_config = "noise";
  unsigned relax_steps = 0;
  double relax_dt = 1.;
  int relax_phi = 0, relax_Q = 0;

  void growth_params(mass_lyo_growth& g) const override;

 public:
  GoOrGrowHip(unsigned LX, unsigned LY, unsigned BC) : LyotropicWithDivisionHip(LX, LY, BC) {}

  option_list GetOptions() override;

  template <class Archive>
  void serialize_params(Archive& ar) {
    LyotropicWithDivisionHip::serialize_params(ar);
    ar.add("B", B)
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
    LyotropicWithDivisionHip::serialize_frame(ar);
    ar.add("m", field(MASS_LYO_M)).add("chi", field(MASS_LYO_CHI));
  }
};
