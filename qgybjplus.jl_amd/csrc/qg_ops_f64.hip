// FP64 instantiation of the fused FFT-pass kernels (plain mode),
// tile variant 1.
#define QG_ALL_LENGTHS 1
#include "qg_ops_impl.hpp"

namespace qg {
// Every transform length with a plan in qg_fft.hpp (FftPlan), compiled for
// the base tile variant.
static const int kLengths[] = {1,   2,   3,   4,   5,   6,   8,   9,   10,  12,  16,  18,  20,  24,  25,  27,
                               32,  36,  40,  45,  48,  50,  54,  64,  72,  75,  80,  90,  96,  100, 108, 125,
                               128, 144, 150, 160, 180, 192, 200, 216, 250, 256, 288, 300, 320, 360, 384, 400,
                               432, 500, 512, 576, 600, 640, 720, 768, 800, 864, 1000, 1024};
bool fft_size_supported(int n) {
  for (int l : kLengths)
    if (l == n) return true;
  return false;
}
std::string fft_sizes_text() {
  std::string t;
  for (int l : kLengths) t += (t.empty() ? "" : ",") + std::to_string(l);
  return t;
}
bool fft_variants_supported(int n) {
  switch (n) {
    case 8: case 16: case 32: case 64: case 128: case 256: case 384: case 512: return true;
    default: return false;
  }
}
template struct OpsV<double, 1>;
}  // namespace qg
