// FP32 instantiation of the fused FFT-pass kernels (mixed mode: FP32
// storage, transforms and products; FP64 column solves and ETD coefficients),
// tile variant 1 with every compiled length, part 1 of the members (the
// parts compile in parallel).
#define QG_ALL_LENGTHS 1
#define QG_PART 1
#include "qg_ops_impl.hpp"

namespace qg {
template struct OpsV<float, 1>;
}  // namespace qg
