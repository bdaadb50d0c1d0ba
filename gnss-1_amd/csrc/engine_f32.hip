// FP32 form of the engine: FFTs, products and RK updates in FP32, reductions in FP64.
#include "engine_impl.h"

namespace gnss {
EngineBase* make_engine_f32(const gnss_params& p) { return new Engine<float>(p); }
EngineBase* make_engine_batch_f32(const gnss_params* p, int nb) { return new Engine<float>(p[0], nb, p); }
}  // namespace gnss
