// models.cpp -- model registry (src/models.cpp:25-60, src/declare_models.cpp).
#include "models.hpp"

#include <map>
#include <sstream>

#include "error_msg.hpp"
#include "go_or_grow_hip.hpp"
#include "lyotropic_hip.hpp"
#include "lyotropic_with_division_hip.hpp"

namespace {
struct Entry {
  std::string desc;
  Factory fact;
};
std::map<std::string, Entry>& registry() {
  static std::map<std::string, Entry> r;
  return r;
}
}  // namespace

void add_model(const std::string& name, const std::string& desc, Factory f) {
  if (registry().count(name)) throw error_msg("model '", name, "' already registered.");
  registry()[name] = {desc, std::move(f)};
}

ModelPtr NewModel(const std::string& name, unsigned LX, unsigned LY, unsigned BC) {
  auto it = registry().find(name);
  if (it == registry().end())
    throw error_msg("model '", name,
                    "' can not be found. Type --list for a list of available models.");
  return it->second.fact(LX, LY, BC);
}

std::string ListModels() {
  std::ostringstream s;
  for (auto& kv : registry()) s << kv.first << "\n    " << kv.second.desc << "\n";
  return s.str();
}

void DeclareModels() {
  if (!registry().empty()) return;
  declare_model<LyotropicHip>(
      "lyotropic",
      "Biphasic, lyotropic, active nematic (D2Q9 lattice Boltzmann + Beris-Edwards Q-tensor + "
      "Cahn-Hilliard phase field, predictor-corrector), run on an MI355X GPU. bc = 0 periodic, "
      "1/2 free-/no-slip channel, 3/4 free-/no-slip box (fp64). precision = fp64 reproduces "
      "the reference arithmetic bit for bit; precision = fp32 (bc = 0) stores shifted "
      "populations and a compensated phase field.");
  declare_model<LyotropicWithDivisionHip>(
      "lyotropic-with-division",
      "The lyotropic model with cell division and death: random circular patches where phi "
      "grows (alpha) or decays (beta), redrawn every division-time / death-time steps; on an "
      "MI355X GPU in fp64, bit for bit the reference arithmetic.");
  declare_model<GoOrGrowHip>(
      "go-or-grow",
      "Lyotropic with cell division and death and a go/grow phenotype field (chi = m/phi), "
      "run on an MI355X GPU in fp64, bit for bit the reference arithmetic.");
}
