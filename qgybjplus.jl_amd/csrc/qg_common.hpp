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
struct Geo {
  int nx, ny, nz;
  int nzc;      // levels per member (the column depth)
  int members;  // nz == members * nzc
  int W;    // nx/2 + 1 : columns kept by half-spectrum (real-field) arrays
  int Wp;   // row stride of half-spectrum arrays (W rounded up to 16)
  const double* kx;            // [nx]
  const double* ky;            // [ny]
  const unsigned char* mask;   // [ny][nx] radial 2/3 mask, 1 = keep
  double inv_n;                // 1/(nx*ny)
  double dz;                   // vertical spacing used by the column operators
};

__host__ __device__ inline int mirror(int i, int n) { return i == 0 ? 0 : n - i; }

}  // namespace qg
