// FP64 ("plain") form of the engine.
#include "engine_impl.h"

namespace gnss {
EngineBase* make_engine_f64(const gnss_params& p) { return new Engine<double>(p); }
EngineBase* make_engine_batch_f64(const gnss_params* p, int nb) { return new Engine<double>(p[0], nb, p); }
}  // namespace gnss
