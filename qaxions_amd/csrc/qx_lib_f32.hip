// FP32 (complex64) instantiation of the FFT pass launchers.
#include "qx_engine.hpp"

namespace qx {
std::unique_ptr<Engine> make_engine_f32(int N) { return make_engine_t<float>(N); }
}  // namespace qx
