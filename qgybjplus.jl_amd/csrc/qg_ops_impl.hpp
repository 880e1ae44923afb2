// Definitions of OpsV<R, V>: runtime dispatch on the transform length to the
// compile-time kernel instantiations of tile variant V.  Included once per
// (precision, variant) translation unit.
#pragma once

#include <stdexcept>
#include <string>

#include "qg_kernels.hpp"
#include "qg_ops.hpp"

// A translation unit that defines QG_PART = k defines only part k of the
// members below (explicit instantiation of OpsV then instantiates those), so
// the base variant's many lengths compile in parallel units.
#ifndef QG_PART
#define QG_PART -1
#endif
#define QG_PART_ON(k) (QG_PART < 0 || QG_PART == (k))

namespace qg {

// The transform lengths compiled.  Every tile variant has the lengths whose
// variants the autotuner explores (8, 16, 32, 64, 128, 256, 384, 512); the
// base variant (the translation units that define QG_ALL_LENGTHS) also has
// every other length from 1 to 1024 with no prime factor above 5 (a plan in
// qg_fft.hpp each), which run at the base variant only (fft_size_supported,
// fft_variants_supported).
#ifdef QG_ALL_LENGTHS
#define QG_MORE_LENGTHS(CALL)                                                 \
    case 10: CALL(10); break;                                                 \
    case 12: CALL(12); break;                                                 \
    case 20: CALL(20); break;                                                 \
    case 24: CALL(24); break;                                                 \
    case 40: CALL(40); break;                                                 \
    case 48: CALL(48); break;                                                 \
    case 80: CALL(80); break;                                                 \
    case 96: CALL(96); break;                                                 \
    case 160: CALL(160); break;                                               \
    case 192: CALL(192); break;                                               \
    case 320: CALL(320); break;                                               \
    case 640: CALL(640); break;                                               \
    case 768: CALL(768); break;                                               \
    case 1024: CALL(1024); break;                                             \
    case 1: CALL(1); break;                                                   \
    case 2: CALL(2); break;                                                   \
    case 3: CALL(3); break;                                                   \
    case 4: CALL(4); break;                                                   \
    case 5: CALL(5); break;                                                   \
    case 6: CALL(6); break;                                                   \
    case 9: CALL(9); break;                                                   \
    case 18: CALL(18); break;                                                 \
    case 25: CALL(25); break;                                                 \
    case 27: CALL(27); break;                                                 \
    case 36: CALL(36); break;                                                 \
    case 45: CALL(45); break;                                                 \
    case 50: CALL(50); break;                                                 \
    case 54: CALL(54); break;                                                 \
    case 72: CALL(72); break;                                                 \
    case 75: CALL(75); break;                                                 \
    case 90: CALL(90); break;                                                 \
    case 100: CALL(100); break;                                               \
    case 108: CALL(108); break;                                               \
    case 125: CALL(125); break;                                               \
    case 144: CALL(144); break;                                               \
    case 150: CALL(150); break;                                               \
    case 180: CALL(180); break;                                               \
    case 200: CALL(200); break;                                               \
    case 216: CALL(216); break;                                               \
    case 250: CALL(250); break;                                               \
    case 288: CALL(288); break;                                               \
    case 300: CALL(300); break;                                               \
    case 360: CALL(360); break;                                               \
    case 400: CALL(400); break;                                               \
    case 432: CALL(432); break;                                               \
    case 500: CALL(500); break;                                               \
    case 576: CALL(576); break;                                               \
    case 600: CALL(600); break;                                               \
    case 720: CALL(720); break;                                               \
    case 800: CALL(800); break;                                               \
    case 864: CALL(864); break;                                               \
    case 1000: CALL(1000); break;                                             \
    case 15: CALL(15); break;                                                 \
    case 30: CALL(30); break;                                                 \
    case 60: CALL(60); break;                                                 \
    case 120: CALL(120); break;                                               \
    case 135: CALL(135); break;                                               \
    case 225: CALL(225); break;                                               \
    case 240: CALL(240); break;                                               \
    case 375: CALL(375); break;                                               \
    case 480: CALL(480); break;                                               \
    case 540: CALL(540); break;                                               \
    case 750: CALL(750); break;                                               \
    case 900: CALL(900); break;                                               \
    case 960: CALL(960); break;                                               \
    case 81: CALL(81); break;                                                 \
    case 162: CALL(162); break;                                               \
    case 243: CALL(243); break;                                               \
    case 270: CALL(270); break;                                               \
    case 324: CALL(324); break;                                               \
    case 405: CALL(405); break;                                               \
    case 450: CALL(450); break;                                               \
    case 486: CALL(486); break;                                               \
    case 625: CALL(625); break;                                               \
    case 648: CALL(648); break;                                               \
    case 675: CALL(675); break;                                               \
    case 729: CALL(729); break;                                               \
    case 810: CALL(810); break;                                               \
    case 972: CALL(972); break;
#else
#define QG_MORE_LENGTHS(CALL)
#endif
#define QG_SWITCH(n, CALL)                                                    \
  switch (n) {                                                                \
    case 8: CALL(8); break;                                                   \
    case 16: CALL(16); break;                                                 \
    case 32: CALL(32); break;                                                 \
    case 64: CALL(64); break;                                                 \
    case 128: CALL(128); break;                                               \
    case 256: CALL(256); break;                                               \
    case 384: CALL(384); break;                                               \
    case 512: CALL(512); break;                                               \
    QG_MORE_LENGTHS(CALL)                                                     \
    default: throw std::runtime_error("transform length " + std::to_string(n) + \
                                      " not compiled for this tile variant"); \
  }

template <int N>
static dim3 ygrid(int ncols, int nz, int S) { return dim3((ncols + S - 1) / S, nz); }
static inline dim3 xgrid(int rows, int S) { return dim3((rows + S - 1) / S); }

#define YL(K, N, ncols, ...)                                                                 \
  do {                                                                                       \
    using TL = Tile<R, N, V>;                                                                   \
    hipLaunchKernelGGL(K, dim3(((ncols) + TL::S - 1) / TL::S, g.nz), dim3(TL::THREADS), 0, st, \
                       __VA_ARGS__);                                                         \
  } while (0)
// full-width Y-pass over the columns the pass carries (fw_tiles)
#define YLF(K, N, ...)                                                                       \
  do {                                                                                       \
    using TL = Tile<R, N, V>;                                                                   \
    hipLaunchKernelGGL(K, dim3(fw_tiles(g, TL::S), g.nz), dim3(TL::THREADS), 0, st, __VA_ARGS__); \
  } while (0)
#define XL(K, N, ...)                                                                        \
  do {                                                                                       \
    using TL = Tile<R, N, V>;                                                                   \
    hipLaunchKernelGGL(K, xgrid(g.nz * g.ny, TL::S), dim3(TL::THREADS), 0, st, __VA_ARGS__); \
  } while (0)

#if QG_PART_ON(0)
template <typename R, int V>
void OpsV<R, V>::ya_qw(const Geo& g, hipStream_t st, const C* B, C* T1, C* T3, const C* twy) {
#define CALL(N) YLF((k_ya_qw<R, N, V>), N, g, B, T1, T3, twy)
  QG_SWITCH(g.ny, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(2)
template <typename R, int V>
void OpsV<R, V>::xa_qw(const Geo& g, hipStream_t st, const C* T1, const C* T3, C* Qa, C* Qb, const C* twx) {
#define CALL(N) XL((k_xa_qw<R, N, V>), N, g, T1, T3, Qa, Qb, twx)
  QG_SWITCH(g.nx, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(0)
template <typename R, int V>
void OpsV<R, V>::ya2_qeff(const Geo& g, hipStream_t st, const C* Qa, const C* Qb, const C* q, C* qeff, C* qw,
                      double inv_f, const C* twy) {
#define CALL(N) YL((k_ya2_qeff<R, N, V>), N, g.Kx, g, Qa, Qb, q, qeff, qw, inv_f, twy)
  QG_SWITCH(g.ny, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(0)
template <typename R, int V>
void OpsV<R, V>::yb_adv(const Geo& g, hipStream_t st, int outs, const C* psi, const C* q, C* Pu, C* Pv, C* Pq,
                    C* Pz, const C* twy) {
  if (outs == (OUT_U | OUT_V)) {
#define CALL(N) YL((k_yb_adv<R, N, V, OUT_U | OUT_V>), N, g.Kx, g, outs, psi, q, Pu, Pv, Pq, Pz, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  } else if (outs == (OUT_Q | OUT_Z)) {
#define CALL(N) YL((k_yb_adv<R, N, V, OUT_Q | OUT_Z>), N, g.Kx, g, outs, psi, q, Pu, Pv, Pq, Pz, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  } else if (outs == (OUT_U | OUT_V | OUT_Q | OUT_Z)) {
#define CALL(N) YL((k_yb_adv<R, N, V, OUT_U | OUT_V | OUT_Q | OUT_Z>), N, g.Kx, g, outs, psi, q, Pu, Pv, Pq, Pz, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  } else {
#define CALL(N) YL((k_yb_adv<R, N, V, 0>), N, g.Kx, g, outs, psi, q, Pu, Pv, Pq, Pz, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  }
}
#endif

#if QG_PART_ON(2)
template <typename R, int V>
void OpsV<R, V>::x_c2r(const Geo& g, hipStream_t st, int nfields, const C* P0, const C* P1, R* o0, R* o1,
                   const C* twx) {
#define CALL(N) XL((k_x_c2r<R, N, V>), N, g, nfields, P0, P1, o0, o1, twx)
  QG_SWITCH(g.nx, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(2)
template <typename R, int V>
void OpsV<R, V>::xb_q(const Geo& g, hipStream_t st, const C* Pq, const C* Pu, const C* Pv, R* u, R* v, C* G1,
                  C* G2, const C* twx) {
  if (Pu) {
#define CALL(N) XL((k_xb_q<R, N, V, 1>), N, g, Pq, Pu, Pv, u, v, G1, G2, twx)
    QG_SWITCH(g.nx, CALL)
#undef CALL
  } else {
#define CALL(N) XL((k_xb_q<R, N, V, 0>), N, g, Pq, Pu, Pv, u, v, G1, G2, twx)
    QG_SWITCH(g.nx, CALL)
#undef CALL
  }
}
#endif

#if QG_PART_ON(2)
template <typename R, int V>
void OpsV<R, V>::xb_B(const Geo& g, hipStream_t st, int flags, const C* T1, const C* Pz, const R* u, const R* v,
                  C* G3, C* G4, const C* twx) {
#define CALL(N) XL((k_xb_B<R, N, V>), N, g, flags, T1, Pz, u, v, G3, G4, twx)
  QG_SWITCH(g.nx, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(2)
template <typename R, int V>
void OpsV<R, V>::xb_all(const Geo& g, hipStream_t st, const C* Pq, const C* Pz, const C* T1, const C* Pu,
                        const C* Pv, const R* u, const R* v, C* G1, C* G2, C* G3, C* G4, const C* twx) {
  if (Pu) {
#define CALL(N) XL((k_xb_all<R, N, V, 1>), N, g, Pq, Pz, T1, Pu, Pv, u, v, G1, G2, G3, G4, twx)
    QG_SWITCH(g.nx, CALL)
#undef CALL
  } else {
#define CALL(N) XL((k_xb_all<R, N, V, 0>), N, g, Pq, Pz, T1, Pu, Pv, u, v, G1, G2, G3, G4, twx)
    QG_SWITCH(g.nx, CALL)
#undef CALL
  }
}
#endif

#if QG_PART_ON(3)
template <typename R, int V>
void OpsV<R, V>::yb2_q(const Geo& g, hipStream_t st, int mode, int flags, const C* G1, const C* G2, const C* qd,
                   const C* qn,
                   C* rhs0, C* qout, EtdView etd, double nu, double inv_dz2, int* flag, const C* twy) {
  Etd e{etd.p};
  if (flags == YQ_ADV) {  // the evolving flow without vertical diffusion
#define CALL(N) YL((k_yb2_q<R, N, V, YQ_ADV>), N, g.Kx, g, mode, flags, G1, G2, qd, qn, rhs0, qout, e, nu, inv_dz2, flag, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  } else {
#define CALL(N) YL((k_yb2_q<R, N, V>), N, g.Kx, g, mode, flags, G1, G2, qd, qn, rhs0, qout, e, nu, inv_dz2, flag, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  }
}
#endif

#if QG_PART_ON(3)
template <typename R, int V>
void OpsV<R, V>::yb2_B(const Geo& g, hipStream_t st, int mode, int flags, const C* G3, const C* G4,
                   const C* A, const C* Bn, C* rhs0, C* Bout, C* T1, C* T3, EtdView etd, double alpha,
                   int* flag, const C* twy) {
  Etd e{etd.p};
#define CALL(N) YLF((k_yb2_B<R, N, V>), N, g, mode, flags, G3, G4, A, Bn, rhs0, Bout, T1, T3, e, alpha, flag, twy)
  QG_SWITCH(g.ny, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(3)
template <typename R, int V>
void OpsV<R, V>::yb2_qB(const Geo& g, hipStream_t st, int mode, const C* G1, const C* G2, const C* qn, C* rq0,
                        C* qout, int bflags, const C* G3, const C* G4, const C* A, const C* Bn, C* rB0, C* Bout,
                        C* T1, C* T3, EtdView etd, double alpha, int* flag, const C* twy) {
  Etd e{etd.p};
#define CALL(N)                                                                                              \
  do {                                                                                                       \
    using TL = Tile<R, N, V>;                                                                                \
    const int nq = (g.Kx + TL::S - 1) / TL::S;                                                               \
    hipLaunchKernelGGL((k_yb2_qB<R, N, V, YQ_ADV>), dim3(nq + fw_tiles(g, TL::S), g.nz), dim3(TL::THREADS), 0, \
                       st, g, nq, mode, YQ_ADV, G1, G2, qn, qn, rq0, qout, 0.0, 0.0, bflags, G3, G4, A, Bn,   \
                       rB0, Bout, T1, T3, alpha, e, flag, twy);                                              \
  } while (0)
  QG_SWITCH(g.ny, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(1)
template <typename R, int V>
void OpsV<R, V>::yf_omega(const Geo& g, hipStream_t st, const C* psi, C* Pbx, C* Pby, C* Pxx, C* Pxy, C* Pu,
                      C* Pv, C* Ppsi, int* flag, double* lsum, const C* twy) {
  // one tile column more than the half width: the Nyquist column's kx outputs
  // (none when pruned, the column is not carried, or for odd nx, which has none)
  if ((Ppsi || lsum) && !Pu) throw std::runtime_error("yf_omega: the check needs the velocity outputs");
  if (Ppsi && lsum) throw std::runtime_error("yf_omega: the exact check or its bound, not both");
  if (Pu && lsum) {
#define CALL(N) YL((k_yf_omega<R, N, V, 1, 2>), N, (g.Kx + (g.pruned() || (g.nx & 1) ? 0 : Tile<R, N, V>::S)), g, psi, Pbx, Pby, Pxx, Pxy, Pu, Pv, Ppsi, flag, lsum, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  } else if (Pu && Ppsi) {
#define CALL(N) YL((k_yf_omega<R, N, V, 1, 1>), N, (g.Kx + (g.pruned() || (g.nx & 1) ? 0 : Tile<R, N, V>::S)), g, psi, Pbx, Pby, Pxx, Pxy, Pu, Pv, Ppsi, flag, lsum, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  } else if (Pu) {
#define CALL(N) YL((k_yf_omega<R, N, V, 1, 0>), N, (g.Kx + (g.pruned() || (g.nx & 1) ? 0 : Tile<R, N, V>::S)), g, psi, Pbx, Pby, Pxx, Pxy, Pu, Pv, Ppsi, flag, lsum, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  } else {
#define CALL(N) YL((k_yf_omega<R, N, V, 0, 0>), N, (g.Kx + (g.pruned() || (g.nx & 1) ? 0 : Tile<R, N, V>::S)), g, psi, Pbx, Pby, Pxx, Pxy, Pu, Pv, Ppsi, flag, lsum, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  }
}
#endif

#if QG_PART_ON(1)
template <typename R, int V>
void OpsV<R, V>::xf(const Geo& g, hipStream_t st, int want_omega, const C* Pu, const C* Pv, const C* Pbx,
                const C* Pby, const C* Pxx, const C* Pxy, R* u, R* v, C* Rw, const C* Ppsi,
                unsigned long long* absmax, double* lsum, const C* twx) {
  if (Ppsi || sizeof(R) == 4) {
#define CALL(N) XL((k_xf<R, N, V, 1>), N, g, want_omega, Pu, Pv, Pbx, Pby, Pxx, Pxy, u, v, Rw, Ppsi, absmax, lsum, g.nz, twx)
    QG_SWITCH(g.nx, CALL)
#undef CALL
  } else {
#define CALL(N) XL((k_xf<R, N, V, 0>), N, g, want_omega, Pu, Pv, Pbx, Pby, Pxx, Pxy, u, v, Rw, Ppsi, absmax, lsum, g.nz, twx)
    QG_SWITCH(g.nx, CALL)
#undef CALL
  }
}
#endif

#if QG_PART_ON(0)
template <typename R, int V>
void OpsV<R, V>::y_r2c_mask(const Geo& g, hipStream_t st, const C* Rw, C* out, const C* twy) {
#define CALL(N) YL((k_y_r2c_mask<R, N, V>), N, g.Kx, g, Rw, out, twy)
  QG_SWITCH(g.ny, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(0)
template <typename R, int V>
void OpsV<R, V>::y_herm(const Geo& g, hipStream_t st, const C* in, C* P, const C* twy) {
#define CALL(N) YL((k_y_herm<R, N, V>), N, g.Kx, g, in, P, twy)
  QG_SWITCH(g.ny, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(3)
template <typename R, int V>
void OpsV<R, V>::y_c2c(const Geo& g, hipStream_t st, int dir, const D* in, D* out, const C* twy) {
  if (dir < 0) {
#define CALL(N) YL((k_y_c2c<R, N, -1, V>), N, g.nx, g, in, out, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  } else {
#define CALL(N) YL((k_y_c2c<R, N, +1, V>), N, g.nx, g, in, out, twy)
    QG_SWITCH(g.ny, CALL)
#undef CALL
  }
}
#endif

#if QG_PART_ON(3)
template <typename R, int V>
void OpsV<R, V>::x_c2c(const Geo& g, hipStream_t st, int dir, const D* in, D* out, const C* twx) {
  if (dir < 0) {
#define CALL(N) XL((k_x_c2c<R, N, -1, V>), N, g, in, out, twx)
    QG_SWITCH(g.nx, CALL)
#undef CALL
  } else {
#define CALL(N) XL((k_x_c2c<R, N, +1, V>), N, g, in, out, twx)
    QG_SWITCH(g.nx, CALL)
#undef CALL
  }
}
#endif

#if QG_PART_ON(1)
template <typename R, int V>
void OpsV<R, V>::check_y(const Geo& g, hipStream_t st, const C* psi, const C* q, const C* B, C* P, int* flag,
                         const C* twy) {
#define CALL(N) YL((k_y_check<R, N, V>), N, g.W, g, psi, q, B, P, flag, twy)
  QG_SWITCH(g.ny, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(1)
template <typename R, int V>
void OpsV<R, V>::check_x(const Geo& g, hipStream_t st, const C* P, unsigned long long* absmax, const C* twx) {
#define CALL(N) XL((k_x_psi_max<R, N, V>), N, g, P, absmax, twx)
  QG_SWITCH(g.nx, CALL)
#undef CALL
}
#endif

#if QG_PART_ON(3)
template <typename R, int V>
void OpsV<R, V>::la_max(const Geo& g, hipStream_t st, const C* B, const C* A, int ybj_plus, C* tmp,
                        unsigned long long* absmax, const C* twx, const C* twy) {
#define CALL(N) YL((k_y_la<R, N, V>), N, g.nx, g, B, A, ybj_plus, tmp, twy)
  QG_SWITCH(g.ny, CALL)
#undef CALL
#define CALL(N) XL((k_x_check<R, N, V>), N, g, tmp, absmax, twx)
  QG_SWITCH(g.nx, CALL)
#undef CALL
}
#endif

}  // namespace qg
