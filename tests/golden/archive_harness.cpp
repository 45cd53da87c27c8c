// Writes a small archive (parameters.json, frame0.json, frame5.json) with the
// host program's serializer, as the host program lays it out: the run's
// generic parameters and model options in "parameters", the Lyotropic frame
// fields (ff as an array of 9-arrays, scalar fields as arrays) in "frame<t>".
// Usage: archive_harness <output directory>
#include <array>
#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "serialization.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = std::string(argv[1]) + "/";
  const unsigned LX = 3, LY = 2, nsteps = 10, ninfo = 5, nstart = 0;
  {
    std::ofstream f(dir + "parameters.json");
    oarchive ar(f, "parameters", 1);
    const std::string model = "lyotropic", config = "random";
    const double tau = 2.0, zeta = 0.02, xi = 0.7;
    const int bc = 0;
    const bool conserve = false;
    ar.add("model_name", model).add("LX", LX).add("LY", LY).add("nsteps", nsteps)
        .add("ninfo", ninfo).add("nstart", nstart).add("bc", bc).add("tauNem", tau)
        .add("zeta", zeta).add("xi", xi).add("init_config", config)
        .add("conserve_phi", conserve);
  }
  for (unsigned t = 0; t <= nsteps; t += ninfo) {
    std::vector<std::array<double, 9>> ff(LX * LY);
    std::vector<double> qxx(LX * LY), phi(LX * LY);
    for (unsigned k = 0; k < LX * LY; ++k) {
      for (int v = 0; v < 9; ++v) ff[k][v] = 40.0 / 9.0 + 0.01 * (k + v) - 1e-7 * t;
      qxx[k] = std::cos(0.3 * k + t) * 0.5;
      phi[k] = 0.5 + 1e-3 * k - 2.5e-5 * t;
    }
    std::ofstream f(dir + "frame" + std::to_string(t) + ".json");
    oarchive ar(f, "frame", 1);
    ar.add("ff", ff).add("QQxx", qxx).add("phi", phi);
  }
  return 0;
}
