// FP64 (complex128) instantiation of the FFT pass launchers.
#include "qx_engine.hpp"

namespace qx {
std::unique_ptr<Engine> make_engine_f64(int N) { return make_engine_t<double>(N); }
}  // namespace qx
