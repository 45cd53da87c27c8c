// FP64 instantiation: reference expression order, built with
// -ffp-contract=off so every operation rounds as in the reference build.
#include "lyo_engine.hpp"

namespace lyo {
Engine* make_engine_fp64() { return new EngineT<double, true>(); }
}  // namespace lyo
