// FP64 instantiation of the fused FFT-pass kernels (plain mode),
// tile variant 1.
#define QG_ALL_LENGTHS 1
#include "qg_ops_impl.hpp"

namespace qg {
bool fft_size_supported(int n) {
  switch (n) {
    case 8: case 10: case 12: case 16: case 20: case 24: case 32: case 40: case 48: case 64: case 80:
    case 96: case 128: case 160: case 192: case 256: case 320: case 384: case 512: case 640: case 768:
    case 1024:
      return true;
    default: return false;
  }
}
bool fft_variants_supported(int n) {
  switch (n) {
    case 8: case 16: case 32: case 64: case 128: case 256: case 384: case 512: return true;
    default: return false;
  }
}
template struct OpsV<double, 1>;
}  // namespace qg
