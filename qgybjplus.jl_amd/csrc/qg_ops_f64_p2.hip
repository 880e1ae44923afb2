// FP64 instantiation of the fused FFT-pass kernels (plain mode), tile
// variant 1 with every compiled length, part 2 of the members (the parts
// compile in parallel).
#define QG_ALL_LENGTHS 1
#define QG_PART 2
#include "qg_ops_impl.hpp"

namespace qg {
template struct OpsV<double, 1>;
}  // namespace qg
