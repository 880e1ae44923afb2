// Flag words shared by the kernels and the host step orchestration.
#pragma once

namespace qg {

// tile variants of the fused FFT passes (Tile<R, N, V>, V in [0, 4); V = 3
// runs the alternative FFT plan)
constexpr int kTileVariants = 4;

// outputs of the advection Y-pass
enum : int { OUT_U = 1, OUT_V = 2, OUT_Q = 4, OUT_Z = 8 };
// products of the B X-pass
enum : int { XB_ADV = 1, XB_REFR = 2 };
// ETD update mode
enum : int { MODE_STAGE = 0, MODE_FINAL = 1 };
// q update: advection present, fixed flow, vertical diffusion
enum : int { YQ_ADV = 1, YQ_FIXED = 2, YQ_DIFF = 4 };
// B update: advection, refraction, dispersion (A) terms present
enum : int { YB_ADV = 1, YB_REFR = 2, YB_DISP = 4 };

}  // namespace qg
