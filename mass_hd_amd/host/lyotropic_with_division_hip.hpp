// lyotropic_with_division_hip.hpp -- the lyotropic model with cell division
// and death on the GPU, behind the Model interface.
//
// Drop-in for the reference's LyotropicWithDivision
// (src/models/lyotropic-with-division.{hpp,cpp}): its runcard options
// (lyotropic-with-division.cpp:128-148), parameter and frame records
// (lyotropic-with-division.hpp:57-85) and runtime checks.  The run is a
// lyotropic run of the C ABI made a division run by mass_lyo_set_growth.
#pragma once

#include <vector>

#include "lyotropic_hip.hpp"

class LyotropicWithDivisionHip : public LyotropicHip {
 protected:
  // lyotropic-with-division.hpp:19-34 (no defaults in the reference: every
  // runcard sets them; 0 here)
  double alpha = 0, beta = 0, division_rate = 0, division_time = 0, division_radius = 0;
  double death_rate = 0, death_time = 0, death_radius = 0, phi_critical = 0;
  bool conserve_phi = false;

  std::vector<bool> mask(int id) const;
  // the model's growth parameters (this model: division and death only)
  virtual void growth_params(mass_lyo_growth& g) const;

 public:
  LyotropicWithDivisionHip(unsigned LX, unsigned LY, unsigned BC) : LyotropicHip(LX, LY, BC) {}

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
        .add("phi_critical", phi_critical);
  }

  template <class Archive>
  void serialize_frame(Archive& ar) {
    LyotropicHip::serialize_frame(ar);
    ar.add("division_mask", mask(MASS_LYO_DIVISION_MASK))
        .add("death_mask", mask(MASS_LYO_DEATH_MASK));
  }
};
