// FP64 instantiation of the fused FFT-pass kernels (plain mode), tile
// variant 3: the alternative FFT plan in half tiles.
#include "qg_ops_impl.hpp"

namespace qg {
template struct OpsV<double, 3>;
}  // namespace qg
