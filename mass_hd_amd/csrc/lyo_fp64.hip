// FP64 instantiation: reference expression order, built with
// -ffp-contract=off so every operation rounds as in the reference build.
#include "lyo_engine.hpp"
#include "lyo_walls.hpp"

namespace lyo {
Engine* make_engine_fp64() { return new EngineT<double, true>(); }
Engine* make_engine_walls(int bc) { return new WallEngine(bc); }
}  // namespace lyo
