// FP32 instantiation of the fused FFT-pass kernels (mixed mode), tile
// variant 3: the alternative FFT plan in half tiles.
#include "qg_ops_impl.hpp"

namespace qg {
template struct OpsV<float, 3>;
}  // namespace qg
