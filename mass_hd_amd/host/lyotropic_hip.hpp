// lyotropic_hip.hpp -- the lyotropic model on the GPU, behind the Model
// interface.
//
// Drop-in for the reference's Lyotropic (src/models/lyotropic.{hpp,cpp}): the
// same runcard options (lyotropic.cpp:551-602), the same parameter and frame
// records (lyotropic.hpp:96-138) and the same runtime checks
// (lyotropic.cpp:527-549).  All fields live on the GPU; the model calls the
// C ABI of include/mass_lyo.h and only downloads data for frames and checks.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "mass_lyo.h"
#include "models.hpp"

class LyotropicHip : public Model {
 protected:
  // runcard options (names as in the reference)
  double level = 0, conc = 0, angle_deg = 0, angle = 0, noise = 0, radius = 0;
  double totalphi = 0;
  double rho = 40.;
  double GammaP = 0, GammaQ = 0, xi = 0, tauNem = 0, tauIso = 0, friction = 0;
  double LL = 0, KK = 0, AA = 0, CC = 0, zeta = 0, zetaI = 0;
  std::string init_config;
  unsigned npc = 1;
  // device-path options
  std::string precision = "fp64";
  double Qamp = 1.0;
  double phinoise = 0.0;
  unsigned chunk_rows = 0;
  // binary checkpoints (no reference counterpart): state loaded instead of the
  // initial configuration / written after the run
  std::string restart_file, checkpoint_file;
  // check totals (lyotropic.hpp:41-44)
  double ftot = 0, ptot = 0;

  mass_lyo_t h_ = nullptr;

  std::vector<double> field(int id) const;

 public:
  LyotropicHip(unsigned LX, unsigned LY, unsigned BC) : Model(LX, LY, BC) {}
  ~LyotropicHip() override;

  void Initialize() override;
  void Configure() override;
  void Step() override;
  void Post() override;
  void RuntimeChecks() override;
  void RuntimeStats() override;
  option_list GetOptions() override;

  // lyotropic.hpp:96-117, plus the device-path options at the end
  template <class Archive>
  void serialize_params(Archive& ar) {
    ar.add("level", level)
        .add("conc", conc)
        .add("angle", angle)
        .add("noise", noise)
        .add("totalphi", totalphi)
        .add("rho", rho)
        .add("GammaP", GammaP)
        .add("GammaQ", GammaQ)
        .add("xi", xi)
        .add("zeta", zeta)
        .add("zetaI", zetaI)
        .add("tauNem", tauNem)
        .add("tauIso", tauIso)
        .add("friction", friction)
        .add("LL", LL)
        .add("KK", KK)
        .add("init_config", init_config)
        .add("AA", AA)
        .add("CC", CC)
        .add("npc", npc)
        .add("precision", precision)
        .add("Qamp", Qamp)
        .add("phinoise", phinoise);
  }

  // lyotropic.hpp:121-138
  template <class Archive>
  void serialize_frame(Archive& ar) {
    const std::vector<double> ff = field(MASS_LYO_FF);
    std::vector<std::array<double, 9>> nodes(DomainSize);
    for (size_t k = 0; k < nodes.size(); ++k)
      for (int v = 0; v < 9; ++v) nodes[k][v] = ff[9 * k + v];
    ar.add("ff", std::move(nodes))
        .add("QQxx", field(MASS_LYO_QQXX))
        .add("QQyx", field(MASS_LYO_QQYX))
        .add("phi", field(MASS_LYO_PHI))
        .add("sigmaXX", field(MASS_LYO_SIGMAXX))
        .add("sigmaXY", field(MASS_LYO_SIGMAXY))
        .add("sigmaYX", field(MASS_LYO_SIGMAYX))
        .add("sigmaYY", field(MASS_LYO_SIGMAYY))
        .add("sigma_bulk", field(MASS_LYO_SIGMA_BULK))
        .add("sigma_elastic_xx", field(MASS_LYO_SIGMA_ELASTIC_XX))
        .add("sigma_elastic_yx", field(MASS_LYO_SIGMA_ELASTIC_YX))
        .add("sigma_phase_field_xx", field(MASS_LYO_SIGMA_PHASE_FIELD_XX))
        .add("sigma_phase_field_yx", field(MASS_LYO_SIGMA_PHASE_FIELD_YX))
        .add("sigma_active_xx", field(MASS_LYO_SIGMA_ACTIVE_XX))
        .add("sigma_active_yx", field(MASS_LYO_SIGMA_ACTIVE_YX));
  }
};
