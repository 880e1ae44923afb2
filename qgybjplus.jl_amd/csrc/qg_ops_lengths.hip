// The transform lengths the library compiles (host-side lists).
#include <string>

#include "qg_ops.hpp"

namespace qg {
// Every transform length with a plan in qg_fft.hpp (FftPlan), compiled for
// the base tile variant: the lengths 1 to 1024 with no prime factor above 5.
static const int kLengths[] = {1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 18, 20, 24, 25,
                               27, 30, 32, 36, 40, 45, 48, 50, 54, 60, 64, 72, 75, 80, 81, 90,
                               96, 100, 108, 120, 125, 128, 135, 144, 150, 160, 162, 180, 192, 200, 216, 225,
                               240, 243, 250, 256, 270, 288, 300, 320, 324, 360, 375, 384, 400, 405, 432, 450,
                               480, 486, 500, 512, 540, 576, 600, 625, 640, 648, 675, 720, 729, 750, 768, 800,
                               810, 864, 900, 960, 972, 1000, 1024};
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
}  // namespace qg
