// FP64 instantiation of the fused FFT-pass kernels (plain mode),
// tile variant 1.
#include "qg_ops_impl.hpp"

namespace qg {
bool fft_size_supported(int n) {
  switch (n) {
    case 8: case 16: case 32: case 64: case 128: case 256: case 384: case 512: return true;
    default: return false;
  }
}
template struct OpsV<double, 1>;
}  // namespace qg
