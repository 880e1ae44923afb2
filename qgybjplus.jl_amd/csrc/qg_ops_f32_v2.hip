// FP32 instantiation of the fused FFT-pass kernels (mixed mode: FP32
// transforms and products, FP64 state, column solves and ETD coefficients),
// tile variant 2.
#include "qg_ops_impl.hpp"

namespace qg {
template struct OpsV<float, 2>;
}  // namespace qg
