// FP32 instantiation: shifted populations, compensated phi, FMA contraction.
#include "lyo_engine.hpp"

namespace lyo {
Engine* make_engine_fp32() { return new EngineT<float, false>(); }
}  // namespace lyo
