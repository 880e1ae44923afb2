// Host launchers of the fused FFT-pass kernels: OpsV<R, V> launches tile
// variant V at intermediate precision R (double: plain FP64 mode, float:
// mixed mode); Ops<R> picks the variant at run time.
#pragma once

#include <string>

#include "qg_common.hpp"
#include "qg_flags.hpp"

namespace qg {

// The ETD-RK2 parameter block the update passes compute their coefficients
// from (Etd in qg_kernels.hpp).
struct EtdView {
  const EtdParams* p;
};

bool fft_size_supported(int n);
// the supported lengths, comma-separated (for error messages)
std::string fft_sizes_text();
// lengths compiled for every tile variant (the others run the base variant)
bool fft_variants_supported(int n);

template <typename R, int V>
struct OpsV {
  using C = cplx<R>;
  using D = cplx<double>;
  static void ya_qw(const Geo& g, hipStream_t st, const C* B, C* T1, C* T3, const C* twy);
  static void xa_qw(const Geo& g, hipStream_t st, const C* T1, const C* T3, C* Qa, C* Qb, const C* twx);
  static void ya2_qeff(const Geo& g, hipStream_t st, const C* Qa, const C* Qb, const C* q, C* qeff, C* qw,
                       double inv_f, const C* twy);
  static void yb_adv(const Geo& g, hipStream_t st, int outs, const C* psi, const C* q, C* Pu, C* Pv, C* Pq,
                     C* Pz, const C* twy);
  static void x_c2r(const Geo& g, hipStream_t st, int nfields, const C* P0, const C* P1, R* o0, R* o1,
                    const C* twx);
  // Pu == nullptr: u, v are read; otherwise they are made from Pu, Pv and written
  static void xb_q(const Geo& g, hipStream_t st, const C* Pq, const C* Pu, const C* Pv, R* u, R* v, C* G1,
                   C* G2, const C* twx);
  static void xb_B(const Geo& g, hipStream_t st, int flags, const C* T1, const C* Pz, const R* u, const R* v,
                   C* G3, C* G4, const C* twx);
  // xb_q and xb_B (advection and refraction both on) in one pass; Pu != nullptr:
  // u, v from Pu, Pv
  static void xb_all(const Geo& g, hipStream_t st, const C* Pq, const C* Pz, const C* T1, const C* Pu,
                     const C* Pv, const R* u, const R* v, C* G1, C* G2, C* G3, C* G4, const C* twx);
  // flag != nullptr: the new q / B are tested for non-finite values (the
  // termination check fused into the step)
  static void yb2_q(const Geo& g, hipStream_t st, int mode, int flags, const C* G1, const C* G2, const C* qd,
                    const C* qn,
                    C* rhs0, C* qout, EtdView etd, double nu, double inv_dz2, int* flag, const C* twy);
  static void yb2_B(const Geo& g, hipStream_t st, int mode, int flags, const C* G3, const C* G4,
                    const C* A, const C* Bn, C* rhs0, C* Bout, C* T1, C* T3, EtdView etd, double alpha,
                    int* flag, const C* twy);
  // yb2_q (flags YQ_ADV only) and yb2_B in one launch (k_yb2_qB): the same
  // values as the two launches
  static void yb2_qB(const Geo& g, hipStream_t st, int mode, const C* G1, const C* G2, const C* qn, C* rq0,
                     C* qout, int bflags, const C* G3, const C* G4, const C* A, const C* Bn, C* rB0, C* Bout,
                     C* T1, C* T3, EtdView etd, double alpha, int* flag, const C* twy);
  // Pu != nullptr: also the y-transformed half rows of u and v (as yb_adv);
  // Ppsi != nullptr: also those of psi, and psi's non-finite test into flag;
  // lsum != nullptr (instead of Ppsi): psi's non-finite test and the per-level
  // bound of the check's max |ifft2(psi)| added into lsum[level]
  static void yf_omega(const Geo& g, hipStream_t st, const C* psi, C* Pbx, C* Pby, C* Pxx, C* Pxy, C* Pu,
                       C* Pv, C* Ppsi, int* flag, double* lsum, const C* twy);
  // Ppsi != nullptr: max |ifft2(psi)| from yf_omega's psi half rows into absmax;
  // lsum != nullptr: the largest of yf_omega's level bounds into absmax, and
  // lsum cleared
  static void xf(const Geo& g, hipStream_t st, int want_omega, const C* Pu, const C* Pv, const C* Pbx,
                 const C* Pby, const C* Pxx, const C* Pxy, R* u, R* v, C* Rw, const C* Ppsi,
                 unsigned long long* absmax, double* lsum, const C* twx);
  static void y_r2c_mask(const Geo& g, hipStream_t st, const C* Rw, C* out, const C* twy);
  static void y_herm(const Geo& g, hipStream_t st, const C* in, C* P, const C* twy);
  // plain complex 2-D transforms of double fields (dir = -1 forward, +1 inverse)
  static void y_c2c(const Geo& g, hipStream_t st, int dir, const D* in, D* out, const C* twy);
  static void x_c2c(const Geo& g, hipStream_t st, int dir, const D* in, D* out, const C* twx);
  // termination check on its own, the Y-pass and the X-pass (run at the tile
  // variants of yf_omega and xf, whose fused check they equal): non-finite
  // q, B, psi -> flag; max |ifft2(psi)| -> absmax (P: one half-width field
  // of scratch)
  static void check_y(const Geo& g, hipStream_t st, const C* psi, const C* q, const C* B, C* P, int* flag,
                      const C* twy);
  static void check_x(const Geo& g, hipStream_t st, const C* P, unsigned long long* absmax, const C* twx);
  // progress maxima: max |ifft2(B + kh2 A / 4)| -> absmax (B alone unless
  // ybj_plus; tmp: one full-width field of scratch)
  static void la_max(const Geo& g, hipStream_t st, const C* B, const C* A, int ybj_plus, C* tmp,
                     unsigned long long* absmax, const C* twx, const C* twy);
};

// Run-time choice of the tile variant: Ops<R>::name(v, args...) calls
// OpsV<R, v>::name(args...) (v outside [0, 4) selects the base variant 1).
template <typename R>
struct Ops {
#define QG_OPS_FWD(name)                                        \
  template <class... A>                                         \
  static void name(int v, A&&... a) {                           \
    switch (v) {                                                \
      case 0: OpsV<R, 0>::name(static_cast<A&&>(a)...); break;  \
      case 2: OpsV<R, 2>::name(static_cast<A&&>(a)...); break;  \
      case 3: OpsV<R, 3>::name(static_cast<A&&>(a)...); break;  \
      default: OpsV<R, 1>::name(static_cast<A&&>(a)...); break; \
    }                                                           \
  }
  QG_OPS_FWD(ya_qw)
  QG_OPS_FWD(xa_qw)
  QG_OPS_FWD(ya2_qeff)
  QG_OPS_FWD(yb_adv)
  QG_OPS_FWD(x_c2r)
  QG_OPS_FWD(xb_q)
  QG_OPS_FWD(xb_B)
  QG_OPS_FWD(xb_all)
  QG_OPS_FWD(yb2_q)
  QG_OPS_FWD(yb2_B)
  QG_OPS_FWD(yb2_qB)
  QG_OPS_FWD(yf_omega)
  QG_OPS_FWD(xf)
  QG_OPS_FWD(y_r2c_mask)
  QG_OPS_FWD(y_herm)
  QG_OPS_FWD(y_c2c)
  QG_OPS_FWD(x_c2c)
  QG_OPS_FWD(check_y)
  QG_OPS_FWD(check_x)
  QG_OPS_FWD(la_max)
#undef QG_OPS_FWD
};

}  // namespace qg
