// Per-(kx,ky) vertical column operators in FP64, one thread per column.
// Planar layout makes every level access coalesced across the wave.  FMA
// contraction is disabled so the elimination rounds exactly like the
// reference's scalar loops.
#include "qg_columns.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace qg {

namespace {

constexpr double kSqrtEps = 1.4901161193847656e-08;  // sqrt(eps(Float64))

__device__ inline void flag_error(int* err, int code, int level, int col) {
  if (atomicCAS(err, 0, code) == 0) {
    err[1] = level;
    err[2] = col;
  }
}

// Thomas algorithm of thomas_solve! (elliptic.jl:1168-1202) applied to the
// Neumann operator of invert_q_to_psi! (elliptic.jl:272-307, kdiv = 1) or of
// invert_B_to_A! (elliptic.jl:949-979, kdiv = 4).  Real and imaginary parts
// share the pivots, as the reference's two solves do.
__global__ void k_thomas(Geo g, const cplx<double>* __restrict__ rhs, cplx<double>* __restrict__ out,
                         cplx<double>* __restrict__ Cout, double* __restrict__ cscr,
                         const double* __restrict__ a, int kdiv, int* __restrict__ err,
                         unsigned long long* __restrict__ warn) {
  const int plane = g.nx * g.ny;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane) return;
  const int x = col % g.nx, y = col / g.nx;
  const int nz = g.nz;
  const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
  const double dz = g.dz;
  const double dz2 = dz * dz;
  auto at = [&](int k) { return (size_t)k * plane + col; };

  if (kh2 == 0.0) {
    if (kdiv == 1) {
      // psi = 0 for the horizontal mean; warn when q has a significant mean.
      double m = 0.0;
      for (int k = 0; k < nz; ++k) {
        cplx<double> v = rhs[at(k)];
        double av = hypot(v.re, v.im);
        m = av > m ? av : m;
        out[at(k)] = mk(0.0, 0.0);
      }
      if (warn && m > 1e-10) atomicMax(warn, (unsigned long long)__double_as_longlong(m));
      return;
    }
    // B -> A gauge solve A[1] = 0 then mean removal (elliptic.jl:892-945).
    if (nz == 1) {
      out[at(0)] = mk(0.0, 0.0);
      if (Cout) Cout[at(0)] = mk(0.0, 0.0);
      return;
    }
    double c = 0.0 / 1.0;
    cplx<double> xv = mk(0.0 / 1.0, 0.0 / 1.0);
    cscr[at(0)] = c;
    out[at(0)] = xv;
    for (int i = 1; i < nz; ++i) {
      const double dl = a[i - 1];
      const double d = (i < nz - 1) ? -(a[i] + a[i - 1]) : -a[nz - 2];
      const double denom = d - dl * c;
      if (fabs(denom) < kSqrtEps) flag_error(err, 1, i, col);
      const double du = (i < nz - 1) ? a[i] : 0.0;
      c = du / denom;
      cplx<double> r = rhs[at(i)];
      xv = mk((dz2 * r.re - dl * xv.re) / denom, (dz2 * r.im - dl * xv.im) / denom);
      cscr[at(i)] = c;
      out[at(i)] = xv;
    }
    for (int i = nz - 2; i >= 0; --i) {
      cplx<double> xi = out[at(i)];
      const double ci = cscr[at(i)];
      xi = mk(xi.re - ci * xv.re, xi.im - ci * xv.im);
      out[at(i)] = xi;
      xv = xi;
    }
    cplx<double> mean = mk(0.0, 0.0);
    for (int k = 0; k < nz; ++k) mean = mean + out[at(k)];
    mean = mk(mean.re / nz, mean.im / nz);
    cplx<double> prev = out[at(0)] - mean;
    out[at(0)] = prev;
    for (int k = 1; k < nz; ++k) {
      cplx<double> cur = out[at(k)] - mean;
      out[at(k)] = cur;
      if (Cout) Cout[at(k - 1)] = mk((cur.re - prev.re) / dz, (cur.im - prev.im) / dz);
      prev = cur;
    }
    if (Cout) Cout[at(nz - 1)] = mk(0.0, 0.0);
    return;
  }

  if (nz == 1) {
    cplx<double> v = rhs[at(0)];
    if (kdiv == 1) {
      out[at(0)] = mk(-v.re / kh2, -v.im / kh2);
    } else {
      out[at(0)] = mk((-4 * v.re) / kh2, (-4 * v.im) / kh2);
      if (Cout) Cout[at(0)] = mk(0.0, 0.0);
    }
    return;
  }

  const double t = (kdiv == 1) ? kh2 * dz2 : (kh2 * dz2) / 4;
  // forward sweep
  double d0 = -(a[0] + t);
  if (fabs(d0) < kSqrtEps) flag_error(err, 1, 0, col);
  double c = a[0] / d0;
  cplx<double> r0 = rhs[at(0)];
  cplx<double> xv = mk((dz2 * r0.re) / d0, (dz2 * r0.im) / d0);
  cscr[at(0)] = c;
  out[at(0)] = xv;
  for (int i = 1; i < nz; ++i) {
    const double dl = a[i - 1];
    const double d = (i < nz - 1) ? -(a[i] + a[i - 1] + t) : -(a[nz - 2] + t);
    const double denom = d - dl * c;
    if (fabs(denom) < kSqrtEps) flag_error(err, 1, i, col);
    const double du = (i < nz - 1) ? a[i] : 0.0;
    c = du / denom;
    cplx<double> r = rhs[at(i)];
    xv = mk((dz2 * r.re - dl * xv.re) / denom, (dz2 * r.im - dl * xv.im) / denom);
    if (i < nz - 1) cscr[at(i)] = c;
    out[at(i)] = xv;
  }
  // back substitution (and C = dA/dz for the B -> A operator)
  cplx<double> above = xv;
  if (Cout) Cout[at(nz - 1)] = mk(0.0, 0.0);
  for (int i = nz - 2; i >= 0; --i) {
    cplx<double> xi = out[at(i)];
    const double ci = cscr[at(i)];
    xi = mk(xi.re - ci * above.re, xi.im - ci * above.im);
    out[at(i)] = xi;
    if (Cout) Cout[at(i)] = mk((above.re - xi.re) / dz, (above.im - xi.im) / dz);
    above = xi;
  }
}

// Interior omega-equation solve (operators.jl:415-470): LAPACK dgtsv
// elimination with partial pivoting, restated; w = 0 on both end levels and
// for kh = 0.  cz[k] = (f^2/N2[k])/dz^2, fN2[k] = f/N2[k].
__global__ void k_omega(Geo g, const cplx<double>* __restrict__ rhsw, cplx<double>* __restrict__ wk,
                        double* __restrict__ sd, double* __restrict__ sdu, double* __restrict__ sfill,
                        const double* __restrict__ cz, const double* __restrict__ fN2,
                        int* __restrict__ err) {
  const int plane = g.nx * g.ny;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane) return;
  const int x = col % g.nx, y = col / g.nx;
  const int nz = g.nz;
  const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
  auto at = [&](int k) { return (size_t)k * plane + col; };
  if (!(kh2 > 0.0 && nz > 2)) {
    for (int k = 0; k < nz; ++k) wk[at(k)] = mk(0.0, 0.0);
    return;
  }
  const int n = nz - 2;  // rows iz = 0..n-1  <->  levels k = iz + 1
  auto diag = [&](int iz) { return -2 * cz[iz + 1] - kh2; };
  auto b_of = [&](int iz) {
    cplx<double> r = rhsw[at(iz + 1)];
    const double s = fN2[iz + 1];
    return mk(s * r.re, s * r.im);
  };
  wk[at(0)] = mk(0.0, 0.0);
  wk[at(nz - 1)] = mk(0.0, 0.0);
  // current row state
  double d = diag(0);
  double du = (n > 1) ? cz[1] : 0.0;   // super-diagonal of row 0 = c_0
  cplx<double> b = b_of(0);
  // next row original values
  for (int i = 0; i < n - 1; ++i) {
    const double dl = cz[i + 2];        // A[i+1, i] = c_{i+1}
    double dn = diag(i + 1);
    double dun = (i + 1 < n - 1) ? cz[i + 2] : 0.0;  // A[i+1, i+2] = c_{i+1}
    cplx<double> bn = b_of(i + 1);
    double fill = 0.0;
    if (fabs(d) >= fabs(dl)) {
      if (d == 0.0) flag_error(err, 2, i, col);
      const double fact = dl / d;
      dn = dn - fact * du;
      bn = mk(bn.re - fact * b.re, bn.im - fact * b.im);
    } else {
      const double fact = d / dl;
      const double dnew = dl;
      const double temp = dn;
      dn = du - fact * temp;
      if (i < n - 2) {
        fill = dun;
        dun = -fact * fill;
      }
      du = temp;
      d = dnew;
      cplx<double> tb = b;
      b = bn;
      bn = mk(tb.re - fact * bn.re, tb.im - fact * bn.im);
    }
    sd[at(i)] = d;
    sdu[at(i)] = du;
    sfill[at(i)] = fill;
    wk[at(i + 1)] = b;
    d = dn;
    du = dun;
    b = bn;
  }
  if (d == 0.0) flag_error(err, 2, n - 1, col);
  // back substitution
  cplx<double> x1 = mk(b.re / d, b.im / d);  // row n-1
  wk[at(n)] = x1;
  if (n > 1) {
    const int i = n - 2;
    cplx<double> bi = wk[at(i + 1)];
    const double di = sd[at(i)], dui = sdu[at(i)];
    cplx<double> x0 = mk((bi.re - dui * x1.re) / di, (bi.im - dui * x1.im) / di);
    wk[at(i + 1)] = x0;
    cplx<double> x2 = x1;
    x1 = x0;
    for (int ii = n - 3; ii >= 0; --ii) {
      cplx<double> bb = wk[at(ii + 1)];
      const double dd = sd[at(ii)], uu = sdu[at(ii)], ff = sfill[at(ii)];
      cplx<double> xv = mk((bb.re - uu * x1.re - ff * x2.re) / dd, (bb.im - uu * x1.im - ff * x2.im) / dd);
      wk[at(ii + 1)] = xv;
      x2 = x1;
      x1 = xv;
    }
  }
}

// q = L psi (compute_q_from_psi!, initialization.jl:597-644) or the barotropic
// q = -kh2 psi (initialization.jl:653-668).
__global__ void k_q_from_psi(Geo g, const cplx<double>* __restrict__ psi, cplx<double>* __restrict__ q,
                             const double* __restrict__ a, int barotropic) {
  const int plane = g.nx * g.ny;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane) return;
  const int x = col % g.nx, y = col / g.nx;
  const int nz = g.nz;
  const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
  auto at = [&](int k) { return (size_t)k * plane + col; };
  if (barotropic || nz == 1) {
    for (int k = 0; k < nz; ++k) {
      cplx<double> p = psi[at(k)];
      q[at(k)] = mk(-kh2 * p.re, -kh2 * p.im);
    }
    return;
  }
  const double dz = g.dz, dz2 = dz * dz;
  for (int k = 1; k < nz - 1; ++k) {
    cplx<double> pu = psi[at(k + 1)], pc = psi[at(k)], pd = psi[at(k - 1)];
    const double au = a[k], ad = a[k - 1], as = a[k] + a[k - 1];
    double vr = au * pu.re - as * pc.re + ad * pd.re;
    double vi = au * pu.im - as * pc.im + ad * pd.im;
    q[at(k)] = mk(-kh2 * pc.re + vr / dz2, -kh2 * pc.im + vi / dz2);
  }
  {
    cplx<double> p1 = psi[at(1)], p0 = psi[at(0)];
    double vr = a[0] * (p1.re - p0.re), vi = a[0] * (p1.im - p0.im);
    q[at(0)] = mk(-kh2 * p0.re + vr / dz2, -kh2 * p0.im + vi / dz2);
  }
  {
    cplx<double> pm = psi[at(nz - 2)], pt = psi[at(nz - 1)];
    double vr = a[nz - 2] * (pm.re - pt.re), vi = a[nz - 2] * (pm.im - pt.im);
    q[at(nz - 1)] = mk(-kh2 * pt.re + vr / dz2, -kh2 * pt.im + vi / dz2);
  }
}

// Reference (z,x,y) order (z fastest) <-> planar [z][y][x], one y-plane per
// blockIdx.z; a 32x32 LDS tile transposes the (x, z) block.
template <typename T>
__global__ void k_transpose(const T* __restrict__ in, T* __restrict__ out, int nx, int ny, int nz,
                            int to_planar) {
  __shared__ T tile[32][33];
  const int y = blockIdx.z;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  if (to_planar) {
    // in: idx = z + nz*(x + nx*y); out: (z*ny + y)*nx + x
    const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32;
    for (int j = ty; j < 32; j += 8) {
      int z = z0 + tx, x = x0 + j;
      if (z < nz && x < nx) tile[j][tx] = in[z + (size_t)nz * (x + (size_t)nx * y)];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
      int x = x0 + tx, z = z0 + j;
      if (z < nz && x < nx) out[((size_t)z * ny + y) * nx + x] = tile[tx][j];
    }
  } else {
    const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32;
    for (int j = ty; j < 32; j += 8) {
      int x = x0 + tx, z = z0 + j;
      if (z < nz && x < nx) tile[j][tx] = in[((size_t)z * ny + y) * nx + x];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
      int z = z0 + tx, x = x0 + j;
      if (z < nz && x < nx) out[z + (size_t)nz * (x + (size_t)nx * y)] = tile[tx][j];
    }
  }
}

template <typename From, typename To>
__global__ void k_convert(const From* __restrict__ in, To* __restrict__ out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (To)in[i];
}

// Any non-finite value among n complex entries -> flag.
__global__ void k_nonfinite(const cplx<double>* __restrict__ a, size_t n, int* __restrict__ flag) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  bool bad = false;
  for (; i < n; i += stride) {
    cplx<double> v = a[i];
    bad |= !(isfinite(v.re) && isfinite(v.im));
  }
  if (bad) atomicOr(flag, 1);
}

__global__ void k_zero(unsigned int* __restrict__ p, size_t nwords) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < nwords; i += stride) p[i] = 0u;
}

}  // namespace

// Zero fill by a kernel (not hipMemsetAsync) so the step stays capturable.
void launch_zero_bytes(void* p, size_t bytes, hipStream_t st) {
  size_t words = bytes / 4;
  unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 4096);
  if (blocks == 0) return;
  hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, st, static_cast<unsigned int*>(p), words);
}

void launch_zero(cplx<double>* p, size_t n, hipStream_t st) { launch_zero_bytes(p, n * sizeof(cplx<double>), st); }

void launch_thomas(const Geo& g, hipStream_t st, const cplx<double>* rhs, cplx<double>* out,
                   cplx<double>* Cout, double* cscr, const double* a, int kdiv, int* err,
                   unsigned long long* warn) {
  const int plane = g.nx * g.ny;
  const int bs = 64;
  hipLaunchKernelGGL(k_thomas, dim3((plane + bs - 1) / bs), dim3(bs), 0, st, g, rhs, out, Cout, cscr, a,
                     kdiv, err, warn);
}

void launch_omega(const Geo& g, hipStream_t st, const cplx<double>* rhsw, cplx<double>* wk, double* sd,
                  double* sdu, double* sfill, const double* cz, const double* fN2, int* err) {
  const int plane = g.nx * g.ny;
  const int bs = 64;
  hipLaunchKernelGGL(k_omega, dim3((plane + bs - 1) / bs), dim3(bs), 0, st, g, rhsw, wk, sd, sdu, sfill, cz,
                     fN2, err);
}

void launch_q_from_psi(const Geo& g, hipStream_t st, const cplx<double>* psi, cplx<double>* q,
                       const double* a, int barotropic) {
  const int plane = g.nx * g.ny;
  const int bs = 64;
  hipLaunchKernelGGL(k_q_from_psi, dim3((plane + bs - 1) / bs), dim3(bs), 0, st, g, psi, q, a, barotropic);
}

template <typename T>
void launch_transpose(const T* in, T* out, int nx, int ny, int nz, bool to_planar, hipStream_t st) {
  dim3 grid((nz + 31) / 32, (nx + 31) / 32, ny);
  hipLaunchKernelGGL(k_transpose<T>, grid, dim3(32, 8), 0, st, in, out, nx, ny, nz, to_planar ? 1 : 0);
}
template void launch_transpose<cplx<double>>(const cplx<double>*, cplx<double>*, int, int, int, bool,
                                             hipStream_t);
template void launch_transpose<double>(const double*, double*, int, int, int, bool, hipStream_t);

template <typename From, typename To>
void launch_convert(const From* in, To* out, size_t n, hipStream_t st) {
  const int bs = 256;
  hipLaunchKernelGGL((k_convert<From, To>), dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st, in, out, n);
}
template void launch_convert<float, double>(const float*, double*, size_t, hipStream_t);
template void launch_convert<double, float>(const double*, float*, size_t, hipStream_t);

void launch_nonfinite(const cplx<double>* a, size_t n, int* flag, hipStream_t st) {
  hipLaunchKernelGGL(k_nonfinite, dim3(1024), dim3(256), 0, st, a, n, flag);
}

}  // namespace qg
