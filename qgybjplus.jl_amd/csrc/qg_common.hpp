// Shared device/host definitions for the QG-YBJ+ HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace qg {

// Interleaved complex value; 8/16-byte aligned so a load is one vector load.
template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R re, im;
};

template <typename R>
__host__ __device__ inline cplx<R> mk(R re, R im) {
  cplx<R> c;
  c.re = re;
  c.im = im;
  return c;
}

template <typename R>
__host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename R>
__host__ __device__ inline cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__host__ __device__ inline cplx<R> conj(cplx<R> a) { return mk(a.re, -a.im); }
template <typename R>
__host__ __device__ inline cplx<R> scale(cplx<R> a, R s) { return mk(a.re * s, a.im * s); }
// i*k*a for real k, with the operation order of `(im*k)*a` (a complex product
// whose zero real part contributes exactly zero for finite inputs).
template <typename R>
__host__ __device__ inline cplx<R> ik_mul(R k, cplx<R> a) { return mk(-(k * a.im), k * a.re); }

template <typename To, typename From>
__host__ __device__ inline cplx<To> cvt(cplx<From> a) { return mk((To)a.re, (To)a.im); }

// Grid description passed by value to every kernel.  Device layout of every
// 3-D field is planar: index (z*ny + y)*stride + x, x fastest.
// nz counts all stacked levels: an ensemble of `members` models of one grid
// class keeps member m in levels [m*nzc, (m+1)*nzc), so every horizontal pass
// covers all members in one launch; column operators work per member.
//
// Dealiased-disk pruning (kd2 >= 0).  The 2/3 mask keeps the disk
// ix^2 + iy^2 <= kmax^2 (ix, iy the signed Fourier indices), and a state that
// has been through one masked update is exactly zero outside it: q, B, psi,
// A, C and every right-hand side.  A pruned pass then touches spectral data
// only inside the disk (loads outside it read as zero, stores outside it are
// skipped and the arrays keep their zeros there), runs its y transforms only
// over the columns |kx| <= kmax, and carries only those columns in its
// x-transformed intermediates.  The x transforms still run over whole rows.
// Every value the pruned passes compute equals the unpruned one: the work
// left out is transforms of zeros and products that the mask discards.
// Unpruned (kd2 < 0): Kx = W, every column and point is touched.
struct Geo {
  int nx, ny, nz;
  int nzc;      // levels per member (the column depth)
  int members;  // nz == members * nzc
  int W;    // nx/2 + 1 : columns kept by half-spectrum (real-field) arrays
  int Wp;   // row stride of half-spectrum arrays (W rounded up to 16)
  int Kx;   // columns [0, Kx) of half-spectrum arrays carried (kmax + 1 when pruned, else W)
  int kd2;  // pruned: kmax^2; unpruned: -1
  int km2;  // kmax^2 of the 2/3 mask, kmax = min(nx, ny) / 3 (src/physics.jl:36-43)
  const double* kx;            // [nx]
  const double* ky;            // [ny]
  const unsigned char* mask;   // [ny][nx] radial 2/3 mask, 1 = keep
  double inv_n;                // 1/(nx*ny)
  double dz;                   // vertical spacing used by the column operators

  __host__ __device__ bool pruned() const { return kd2 >= 0; }
  // column x carried by a pruned pass (|kx| < Kx); every column otherwise
  __host__ __device__ bool col_kept(int x) const {
    const int ix = x <= nx - x ? x : nx - x;
    return ix < Kx;
  }
  // the 2/3 mask at (x, y): ix^2 + iy^2 <= kmax^2 with the signed Fourier
  // indices ix, iy (what the mask array holds, computed instead of loaded)
  __host__ __device__ bool keep(int x, int y) const {
    const int ix = x <= nx - x ? x : nx - x, iy = y <= ny - y ? y : ny - y;
    return ix * ix + iy * iy <= km2;
  }
  // point (x, y) touched by a pruned pass: inside the disk; every point
  // otherwise
  __host__ __device__ bool in_disk(int x, int y) const { return kd2 < 0 || keep(x, y); }
};

__host__ __device__ inline int mirror(int i, int n) { return i == 0 ? 0 : n - i; }

// The ETD-RK2 step's parameters in device memory: dt and the closure terms
// (coefficient, order / 2) of the flow (index 0, for q) and the waves (1, for
// B), none when inviscid.  Rewritten in place when dt changes, so kernels in
// a captured graph read the current values.
constexpr int kEtdTerms = 8;
struct EtdParams {
  double dt;
  int n[2];
  int half[2][kEtdTerms];
  double coef[2][kEtdTerms];
};

// Tiles of S adjacent columns over the full width: every column, or, when
// pruned, the columns |kx| <= kmax as tiles over [0, S*nlo) and
// [nx - S*nhi, nx) (the two ranges carried, each rounded up to whole tiles)
// when those do not cover the row.  fw_tiles is the number of tiles, fw_x the
// column of slot s of tile b.
__host__ __device__ inline void fw_split(const Geo& g, int S, int& nlo, int& nhi) {
  nlo = (g.Kx + S - 1) / S;
  nhi = (g.Kx - 1 + S - 1) / S;
  if (!g.pruned() || (nlo + nhi) * S >= g.nx) nlo = nhi = -1;
}
__host__ __device__ inline int fw_tiles(const Geo& g, int S) {
  int lo, hi;
  fw_split(g, S, lo, hi);
  return lo < 0 ? (g.nx + S - 1) / S : lo + hi;
}
__host__ __device__ inline int fw_x(const Geo& g, int S, int b, int s) {
  int lo, hi;
  fw_split(g, S, lo, hi);
  if (lo < 0 || b < lo) return b * S + s;
  return g.nx - (lo + hi - b) * S + s;
}

// Workgroups reach the 8 XCDs of the chip (each with its own L2) round-robin
// by linear workgroup id.  xcd_block() renumbers a 2-D grid so that every XCD
// takes a contiguous range of logical blocks (x fastest) in the order it
// receives them: neighbouring tiles, whose edges share cache lines, then run
// on one XCD close together in time and the second reads hit its L2.  The
// renumbering is a permutation of the grid, so results do not depend on it.
constexpr int kXcds = 8;
struct Blk {
  int x, y;
};
__device__ inline Blk xcd_block() {
  const int gx = gridDim.x, total = gx * gridDim.y;
  const int l = blockIdx.x + gx * blockIdx.y;
  const int xcd = l % kXcds, loc = l / kXcds, q = total / kXcds, r = total % kXcds;
  const int L = xcd * q + (xcd < r ? xcd : r) + loc;
  return Blk{L % gx, L / gx};
}

}  // namespace qg
