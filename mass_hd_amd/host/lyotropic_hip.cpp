// lyotropic_hip.cpp -- see lyotropic_hip.hpp.
#include "lyotropic_hip.hpp"

#include <cmath>
#include <fstream>
#include <iostream>
#include <iterator>
#include <map>

#include "error_msg.hpp"

extern unsigned nsubsteps;

namespace {
const std::map<std::string, int>& config_ids() {
  static const std::map<std::string, int> m = {
      {"circle", MASS_LYO_CONFIG_CIRCLE},
      {"square", MASS_LYO_CONFIG_SQUARE},
      {"stripe", MASS_LYO_CONFIG_STRIPE},
      {"half", MASS_LYO_CONFIG_HALF},
      {"random", MASS_LYO_CONFIG_RANDOM},
      {"boxatwall", MASS_LYO_CONFIG_BOXATWALL},
      {"cuttingatwall", MASS_LYO_CONFIG_CUTTINGATWALL},
      {"circleatwall", MASS_LYO_CONFIG_CIRCLEATWALL},
      {"wettedwall", MASS_LYO_CONFIG_WETTEDWALL}};
  return m;
}
}  // namespace

LyotropicHip::~LyotropicHip() {
  if (h_) mass_lyo_destroy(h_);
}

void LyotropicHip::Initialize() {
  // lyotropic.cpp:23-67
  angle = angle_deg * M_PI / 180.;
  if (nsubsteps > 1)
    throw error_msg("time stepping not implemented for this model, please set nsubsteps=1.");
  if (BC > 4)
    throw error_msg("boundary conditions bc=", BC,
                    " are not available on the GPU path (bc 0 periodic, 1/2 free-/no-slip "
                    "channel, 3/4 free-/no-slip box).");
  mass_lyo_params p{};
  p.LX = (int32_t)LX;
  p.LY = (int32_t)LY;
  p.npc = (int32_t)npc;
  if (precision == "fp64")
    p.precision = MASS_LYO_FP64;
  else if (precision == "fp32")
    p.precision = MASS_LYO_FP32;
  else
    throw error_msg("precision must be fp64 or fp32 (got '", precision, "')");
  p.conserve_phi = 0;
  p.plane_pad = 0;
  p.bc = (int32_t)BC;
  p.GammaQ = GammaQ;
  p.GammaP = GammaP;
  p.xi = xi;
  p.tauNem = tauNem;
  p.tauIso = tauIso;
  p.rho = rho;
  p.friction = friction;
  p.AA = AA;
  p.CC = CC;
  p.LL = LL;
  p.KK = KK;
  p.zeta = zeta;
  p.zetaI = zetaI;
  p.conc = conc;
  if (mass_lyo_create(&p, &h_) != MASS_LYO_OK)
    throw error_msg("GPU model initialisation failed: ", mass_lyo_last_error(nullptr));
  if (chunk_rows && mass_lyo_set_launch(h_, (int32_t)chunk_rows, 0, -1) != MASS_LYO_OK)
    throw error_msg(mass_lyo_last_error(h_));
}

void LyotropicHip::Configure() {
  if (!restart_file.empty()) {
    // state from a binary checkpoint; the check totals are its sums
    std::ifstream in(restart_file, std::ios::binary);
    if (!in) throw error_msg("cannot open restart file '", restart_file, "'");
    const std::vector<char> image((std::istreambuf_iterator<char>(in)),
                                  std::istreambuf_iterator<char>());
    if (mass_lyo_checkpoint_load(h_, image.data(), image.size()) != MASS_LYO_OK)
      throw error_msg("restart file '", restart_file, "': ", mass_lyo_last_error(h_));
    if (mass_lyo_sums(h_, &ftot, &ptot) != MASS_LYO_OK) throw error_msg(mass_lyo_last_error(h_));
    totalphi = ptot;
    return;
  }
  // lyotropic.cpp:69-166
  auto it = config_ids().find(init_config);
  if (it == config_ids().end())
    throw error_msg("error: initial configuration '", init_config, "' unknown.");
  mass_lyo_init in{};
  in.config = it->second;
  extern unsigned seed;
  in.seed = seed;
  in.level = level;
  in.radius = radius;
  in.angle_deg = angle_deg;
  in.noise = noise;
  in.qamp = Qamp;
  in.phinoise = phinoise;
  if (mass_lyo_configure(h_, &in, &ftot, &ptot) != MASS_LYO_OK)
    throw error_msg(mass_lyo_last_error(h_));
  totalphi = ptot;
}

void LyotropicHip::Step() {
  // queued on the device stream; errors surface at the next synchronising call
  mass_lyo_step(h_, 1);
}

std::vector<double> LyotropicHip::field(int id) const {
  std::vector<double> out((id == MASS_LYO_FF ? 9 : 1) * (size_t)DomainSize);
  const int rc = mass_lyo_get_field(h_, id, out.data());
  // a NaN is reported by the archive while writing (write.cpp:41)
  if (rc != MASS_LYO_OK && rc != MASS_LYO_ENAN) throw error_msg(mass_lyo_last_error(h_));
  return out;
}

void LyotropicHip::RuntimeChecks() {
  // lyotropic.cpp:527-549
  double fcheck = 0, pcheck = 0;
  if (mass_lyo_sums(h_, &fcheck, &pcheck) != MASS_LYO_OK) throw error_msg(mass_lyo_last_error(h_));
  std::cout << "fcheck: " << fcheck << "/" << ftot << '\n';
  if (std::abs(ftot - fcheck) > 1) throw error_msg("f is not conserved (", ftot, "/", fcheck, ")");
  std::cout << "pcheck: " << pcheck << "/" << ptot << '\n';
}

void LyotropicHip::Post() {
  if (checkpoint_file.empty()) return;
  uint64_t n = 0;
  if (mass_lyo_checkpoint_size(h_, &n) != MASS_LYO_OK) throw error_msg(mass_lyo_last_error(h_));
  std::vector<char> image(n);
  if (mass_lyo_checkpoint_save(h_, image.data(), n) != MASS_LYO_OK)
    throw error_msg(mass_lyo_last_error(h_));
  std::ofstream out(checkpoint_file, std::ios::binary | std::ios::trunc);
  if (!out.write(image.data(), (std::streamsize)n))
    throw error_msg("cannot write checkpoint file '", checkpoint_file, "'");
}

void LyotropicHip::RuntimeStats() {}

option_list LyotropicHip::GetOptions() {
  OptionGroup model{"Model options", {}};
  model.add_double("GammaQ", &GammaQ, "Q-tensor mobility")
      .add_double("GammaP", &GammaP, "binary mobility")
      .add_double("xi", &xi, "tumbling/aligning parameter")
      .add_double("tauNem", &tauNem, "nematic viscosity")
      .add_double("tauIso", &tauIso, "isotropic viscosity")
      .add_double("rho", &rho, "fluid density")
      .default_value("40")
      .add_double("friction", &friction, "friction from confinement")
      .add_double("AA", &AA, "binary fluid bulk constant")
      .add_double("CC", &CC, "coupling constant")
      .add_double("LL", &LL, "elastic constant")
      .add_double("KK", &KK, "binary gradient constant")
      .add_double("zeta", &zeta, "activity parameter")
      .add_double("zetaI", &zetaI, "activity isotropic parameter")
      .add_uint("npc", &npc, "number of correction steps for the predictor-corrector method")
      .default_value("1");
  OptionGroup config{"Initial configuration options", {}};
  config.add_string("config", &init_config, "initial configuration")
      .add_double("level", &level, "starting thickness of the nematic region")
      .add_double("conc", &conc, "starting phi concentration of nematic region")
      .add_double("radius", &radius, "radius of the initial circle")
      .add_double("angle", &angle_deg, "initial angle to x direction (in degrees)")
      .add_double("noise", &noise, "size of initial variations")
      .add_double("Qamp", &Qamp, "Q-tensor amplitude: Qxx = Qamp*cos(2 theta) (reference: 1)")
      .default_value("1")
      .add_double("phinoise", &phinoise,
                  "half width of the per-node phi draw of config=random")
      .default_value("0");
  OptionGroup device{"GPU options", {}};
  device.add_string("precision", &precision, "fp64 (reference arithmetic) or fp32")
      .default_value("fp64")
      .add_uint("chunk-rows", &chunk_rows, "rows per wavefront chunk of the sweeps (0 = auto)")
      .default_value("0")
      .add_string("restart", &restart_file,
                  "binary checkpoint to start from instead of the initial configuration")
      .add_string("checkpoint", &checkpoint_file, "write a binary checkpoint after the run");
  return {model, config, device};
}
