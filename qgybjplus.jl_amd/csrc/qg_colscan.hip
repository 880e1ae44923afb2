// Instantiations of the partitioned column-solve kernel (qg_colscan.hpp) for
// every operator and compiled launch shape.
#include "qg_colscan.hpp"

#pragma clang fp contract(off)

namespace qg {

namespace {

template <int OP, int M, int QC, typename T>
void launch_shape(const Geo& g, hipStream_t st, const ScanArgs& a) {
  const Quarter qr = quarter(g);
  const ScanShape s{M, QC, (g.nzc + M - 1) / M};
  const int blocks = scan_tiles(g, QC);
  hipLaunchKernelGGL((k_colscan<OP, M, QC, T>), dim3(blocks, g.members), dim3(s.threads()), s.lds_bytes(), st, g,
                     qr, a);
}

template <int OP, int I>
void launch_idx(const Geo& g, hipStream_t st, int idx, const ScanArgs& a, bool f32) {
  if constexpr (I < kScanShapeCount) {
    if (idx == I) {
      if (f32) launch_shape<OP, kScanShapes[I][0], kScanShapes[I][1], float>(g, st, a);
      else launch_shape<OP, kScanShapes[I][0], kScanShapes[I][1], double>(g, st, a);
    } else {
      launch_idx<OP, I + 1>(g, st, idx, a, f32);
    }
  }
}

template <int OP0, int OP1, int M, int QC, typename T>
void launch_shape2(const Geo& g, hipStream_t st, const ScanArgs& a0, const ScanArgs& a1) {
  const Quarter qr = quarter(g);
  const ScanShape s{M, QC, (g.nzc + M - 1) / M};
  const int blocks = scan_tiles(g, QC);
  hipLaunchKernelGGL((k_colscan2<OP0, OP1, M, QC, T>), dim3(2 * blocks, g.members), dim3(s.threads()),
                     s.lds_bytes(), st, g, qr, a0, a1, blocks);
}

template <int OP0, int OP1, int I>
void launch_idx2(const Geo& g, hipStream_t st, int idx, const ScanArgs& a0, const ScanArgs& a1, bool f32) {
  if constexpr (I < kScanShapeCount) {
    if (idx == I) {
      if (f32) launch_shape2<OP0, OP1, kScanShapes[I][0], kScanShapes[I][1], float>(g, st, a0, a1);
      else launch_shape2<OP0, OP1, kScanShapes[I][0], kScanShapes[I][1], double>(g, st, a0, a1);
    } else {
      launch_idx2<OP0, OP1, I + 1>(g, st, idx, a0, a1, f32);
    }
  }
}

template <int OP, int I>
bool setup_idx() {
  if constexpr (I < kScanShapeCount) {
    const void* f64 = reinterpret_cast<const void*>(k_colscan<OP, kScanShapes[I][0], kScanShapes[I][1], double>);
    const void* f32 = reinterpret_cast<const void*>(k_colscan<OP, kScanShapes[I][0], kScanShapes[I][1], float>);
    for (const void* f : {f64, f32})
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kScanMaxLds) != hipSuccess)
        return false;
    return setup_idx<OP, I + 1>();
  }
  return true;
}

template <int OP0, int OP1, int I>
bool setup_idx2() {
  if constexpr (I < kScanShapeCount) {
    const void* f64 =
        reinterpret_cast<const void*>(k_colscan2<OP0, OP1, kScanShapes[I][0], kScanShapes[I][1], double>);
    const void* f32 =
        reinterpret_cast<const void*>(k_colscan2<OP0, OP1, kScanShapes[I][0], kScanShapes[I][1], float>);
    for (const void* f : {f64, f32})
      if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kScanMaxLds) != hipSuccess)
        return false;
    return setup_idx2<OP0, OP1, I + 1>();
  }
  return true;
}

}  // namespace

template <int OP>
void launch_colscan(const Geo& g, hipStream_t st, int idx, const ScanArgs& a, bool f32) {
  launch_idx<OP, 0>(g, st, idx, a, f32);
}

template void launch_colscan<SCAN_PSI>(const Geo&, hipStream_t, int, const ScanArgs&, bool);
template void launch_colscan<SCAN_A>(const Geo&, hipStream_t, int, const ScanArgs&, bool);
template void launch_colscan<SCAN_AC>(const Geo&, hipStream_t, int, const ScanArgs&, bool);
template void launch_colscan<SCAN_OMEGA>(const Geo&, hipStream_t, int, const ScanArgs&, bool);

template <int OP0, int OP1>
void launch_colscan2(const Geo& g, hipStream_t st, int idx, const ScanArgs& a0, const ScanArgs& a1, bool f32) {
  launch_idx2<OP0, OP1, 0>(g, st, idx, a0, a1, f32);
}

template void launch_colscan2<SCAN_PSI, SCAN_A>(const Geo&, hipStream_t, int, const ScanArgs&, const ScanArgs&,
                                                bool);
template void launch_colscan2<SCAN_PSI, SCAN_AC>(const Geo&, hipStream_t, int, const ScanArgs&, const ScanArgs&,
                                                 bool);

bool colscan_setup() {
  static const bool ok = setup_idx<SCAN_PSI, 0>() && setup_idx<SCAN_A, 0>() && setup_idx<SCAN_AC, 0>() &&
                         setup_idx<SCAN_OMEGA, 0>() && setup_idx2<SCAN_PSI, SCAN_A, 0>() &&
                         setup_idx2<SCAN_PSI, SCAN_AC, 0>();
  return ok;
}

}  // namespace qg
