// FP64 instantiation of the fused FFT-pass kernels (plain mode),
// tile variant 0.
#include "qg_ops_impl.hpp"

namespace qg {
template struct OpsV<double, 0>;
}  // namespace qg
