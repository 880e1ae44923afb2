// Mirror tiles for the per-(kx, ky) vertical column solves.
//
// Every column operator of the step depends on (kx, ky) only through kh2, so
// a column's elimination is recorded once per model for the quarter of
// wavenumbers |ix| <= nx/2, |iy| <= ny/2 ("quarter columns", index
// rc = ry * (nx/2 + 1) + rx), level-indexed [k][rc].  A solve tile is QC
// consecutive quarter columns and their mirror images (+-kx, +-ky): NC = 4*QC
// grid columns that share QC table columns.  A block
//   1. stages the tile's right-hand sides [kpad][NC] (complex) and two
//      tables [kpad][QC] into LDS: every thread keeps one column slot, all
//      loads are unconditional (clamped addresses) and issued before the
//      first LDS store, padding levels (nz rounded up to kMirrorLevels) and
//      self-mirror duplicates (kx or ky equal to 0 or Nyquist) hold zeros;
//   2. runs the 2*NC recurrences (one lane per real or imaginary part, as the
//      reference solves them separately) out of LDS with one FMA per level on
//      the loop-carried chain;
//   3. writes back.
#pragma once

#include "qg_common.hpp"

namespace qg {

constexpr int kMirrorThreads = 256;
constexpr int kMirrorLevels = 16;              // sweep batch; nz is padded to a multiple
constexpr size_t kMirrorLdsBytes = 80 * 1024;  // two blocks per CU
constexpr int kMirrorLoads = 16;               // right-hand-side loads per thread and pass

struct Quarter {
  int Wr, Hr;  // nx/2 + 1, ny/2 + 1
  __host__ __device__ int size() const { return Wr * Hr; }
};

inline Quarter quarter(const Geo& g) { return Quarter{g.nx / 2 + 1, g.ny / 2 + 1}; }

__host__ __device__ inline int mirror_pad(int nz) {
  return (nz + kMirrorLevels - 1) / kMirrorLevels * kMirrorLevels;
}

// LDS of one tile: right-hand sides and two tables.
__host__ __device__ inline size_t mirror_lds_bytes(int kpad, int qc) {
  return (size_t)kpad * (4 * qc * sizeof(cplx<double>) + 2 * qc * sizeof(double));
}

// Quarter columns per tile for depth nz (0 when a column does not fit).
inline int mirror_qc(int nz) {
  for (int qc = 8; qc >= 1; qc /= 2)
    if (mirror_lds_bytes(mirror_pad(nz), qc) <= kMirrorLdsBytes) return qc;
  return 0;
}

struct QuadCol {
  int x, y;
  bool active;
};

// Grid column of mirror image m (bit 0: -kx, bit 1: -ky) of quarter column
// rc; inactive past the quarter or when it duplicates a lower image.
__device__ inline QuadCol quad_col(const Geo& g, const Quarter& qr, int rc, int m) {
  QuadCol q{0, 0, false};
  if (rc >= qr.size()) return q;
  const int rx = rc % qr.Wr, ry = rc / qr.Wr;
  q.x = (m & 1) ? (g.nx - rx) % g.nx : rx;
  q.y = (m & 2) ? (g.ny - ry) % g.ny : ry;
  q.active = !(((m & 1) && q.x == rx) || ((m & 2) && q.y == ry));
  return q;
}

// The staging/writeback slot of this thread: column slot c, levels
// kf, kf + kstep, ...
template <int QC>
struct MirrorSlot {
  static constexpr int NC = 4 * QC;
  static constexpr int kstep = kMirrorThreads / NC;
  int c, kf;
  QuadCol q;
  size_t off;  // y * nx + x
  __device__ MirrorSlot(const Geo& g, const Quarter& qr, int rc0) {
    c = threadIdx.x % NC;
    kf = threadIdx.x / NC;
    q = quad_col(g, qr, rc0 + c % QC, c / QC);
    off = q.active ? (size_t)q.y * g.nx + q.x : 0;
  }
};

// Phase 1.  xs: [kpad][NC] complex, t0/t1: [kpad][QC].  rhs and the tables
// point at the member's first level; the column has g.nzc levels.  The
// right-hand side of level k is multiplied by scale(k), which is 0 where the
// operator has no row.
template <int QC, class Scale>
__device__ inline void mirror_stage(const Geo& g, const Quarter& qr, int rc0, const MirrorSlot<QC>& sl,
                                    const cplx<double>* __restrict__ rhs, const double* __restrict__ tab0,
                                    const double* __restrict__ tab1, cplx<double>* xs, double* t0, double* t1,
                                    int kpad, Scale scale) {
  constexpr int NC = MirrorSlot<QC>::NC, kstep = MirrorSlot<QC>::kstep;
  constexpr int TU = kMirrorLoads * QC / NC;  // table loads per pass (= 4)
  const int nz = g.nzc;
  const size_t plane = (size_t)g.nx * g.ny;
  const int rplane = qr.size();
  const int j = threadIdx.x % QC, kt = threadIdx.x / QC, ktstep = kMirrorThreads / QC;
  const bool jok = rc0 + j < rplane;
  const size_t toff = (size_t)min(rc0 + j, rplane - 1);
  for (int L0 = 0; L0 < kpad; L0 += kMirrorLoads * kstep) {
    cplx<double> v[kMirrorLoads];
    double sc[kMirrorLoads], a0[TU], a1[TU];
#pragma unroll
    for (int u = 0; u < kMirrorLoads; ++u) {
      const int kc = min(L0 + sl.kf + u * kstep, nz - 1);
      v[u] = rhs[(size_t)kc * plane + sl.off];
      sc[u] = scale(kc);
    }
#pragma unroll
    for (int u = 0; u < TU; ++u) {
      const size_t at = (size_t)min(L0 + kt + u * ktstep, nz - 1) * rplane + toff;
      a0[u] = tab0[at];
      a1[u] = tab1[at];
    }
#pragma unroll
    for (int u = 0; u < kMirrorLoads; ++u) {
      const int k = L0 + sl.kf + u * kstep;
      if (k < kpad) {
        const double s = (k < nz && sl.q.active) ? sc[u] : 0.0;
        xs[k * NC + sl.c] = mk(s * v[u].re, s * v[u].im);
      }
    }
#pragma unroll
    for (int u = 0; u < TU; ++u) {
      const int k = L0 + kt + u * ktstep;
      if (k < kpad) {
        const bool ok = k < nz && jok;
        t0[k * QC + j] = ok ? a0[u] : 0.0;
        t1[k * QC + j] = ok ? a1[u] : 0.0;
      }
    }
  }
}

// Phase 2 sweeps of one column part over [0, kpad): b0 holds level k at
// b0[k * st].
//   forward:  x'_k = fma(-G_k, x'_{k-1}, H_k)   (x'_{-1} = 0)
//   backward: x_k  = fma(-P_k, x_{k+1}, Q_k)    (x_kpad = 0)
// fwd(k, b_k) returns {G_k, H_k}; bwd(k, x'_k) returns {P_k, Q_k}.  Padding
// levels must give H = 0 and Q = 0.
template <class Fwd, class Bwd>
__device__ inline void mirror_sweeps(double* b0, int st, int kpad, Fwd fwd, Bwd bwd) {
  constexpr int P = kMirrorLevels;
  double prev = 0.0;
  for (int k0 = 0; k0 < kpad; k0 += P) {
    double gk[P], hk[P];
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const double2 gh = fwd(k0 + u, b0[(k0 + u) * st]);
      gk[u] = gh.x;
      hk[u] = gh.y;
    }
#pragma unroll
    for (int u = 0; u < P; ++u) {
      prev = fma(-gk[u], prev, hk[u]);
      hk[u] = prev;
    }
#pragma unroll
    for (int u = 0; u < P; ++u) b0[(k0 + u) * st] = hk[u];
  }
  double above = 0.0;
  for (int k0 = kpad - P; k0 >= 0; k0 -= P) {
    double pk[P], qk[P];
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const int k = k0 + P - 1 - u;
      const double2 pq = bwd(k, b0[k * st]);
      pk[u] = pq.x;
      qk[u] = pq.y;
    }
#pragma unroll
    for (int u = 0; u < P; ++u) {
      above = fma(-pk[u], above, qk[u]);
      qk[u] = above;
    }
#pragma unroll
    for (int u = 0; u < P; ++u) b0[(k0 + P - 1 - u) * st] = qk[u];
  }
}

}  // namespace qg
