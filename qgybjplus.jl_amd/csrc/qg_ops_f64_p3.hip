// FP64 instantiation of the fused FFT-pass kernels (plain mode), tile
// variant 1 with every compiled length, part 3 of the members (the parts
// compile in parallel).
#define QG_ALL_LENGTHS 1
#define QG_PART 3
#include "qg_ops_impl.hpp"

namespace qg {
template struct OpsV<double, 1>;
}  // namespace qg
