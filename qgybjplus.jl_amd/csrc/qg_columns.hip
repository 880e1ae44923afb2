// Per-(kx,ky) vertical column operators in FP64, one thread per column.
// Planar layout makes every level access coalesced across the wave.  FMA
// contraction is disabled so the elimination rounds exactly like the
// reference's scalar loops.
#include "qg_columns.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace qg {

namespace {

// q = L psi (compute_q_from_psi!, initialization.jl:597-644) or the barotropic
// q = -kh2 psi (initialization.jl:653-668).
template <typename TI, typename TO>
__global__ void k_q_from_psi(Geo g, const cplx<TI>* __restrict__ psi_s, cplx<TO>* __restrict__ q_s,
                             const double* __restrict__ a, int barotropic) {
  const int plane = g.nx * g.ny;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane) return;
  const int x = col % g.nx, y = col / g.nx;
  const int nz = g.nzc;
  const size_t m = blockIdx.y;  // member
  psi_s += m * nz * plane;
  q_s += m * nz * plane;
  a += m * nz;
  const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
  auto at = [&](int k) { return (size_t)k * plane + col; };
  struct Ld {
    const cplx<TI>* p;
    __device__ cplx<double> operator[](size_t i) const { return cvt<double>(p[i]); }
  } psi{psi_s};
  struct St {
    cplx<TO>* p;
    struct Ref {
      cplx<TO>* e;
      __device__ void operator=(cplx<double> v) { *e = cvt<TO>(v); }
    };
    __device__ Ref operator[](size_t i) const { return Ref{p + i}; }
  } q{q_s};
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

__global__ void k_zero(unsigned int* __restrict__ p, size_t nwords) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < nwords; i += stride) p[i] = 0u;
}

// Zeroes the points of a spectral field outside the dealiased disk of g (all
// g.nz levels): each thread owns one (x, y) of the plane.
template <typename T>
__global__ void k_zero_outside_disk(Geo g, cplx<T>* __restrict__ f) {
  const int plane = g.nx * g.ny;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= plane || g.in_disk(i % g.nx, i / g.nx)) return;
  for (int z = blockIdx.y; z < g.nz; z += gridDim.y) f[(size_t)z * plane + i] = mk<T>((T)0, (T)0);
}

}  // namespace

void launch_zero_outside_disk(const Geo& g, hipStream_t st, void* field, bool f32) {
  const int plane = g.nx * g.ny, bs = 256;
  const dim3 grid((plane + bs - 1) / bs, std::min(g.nz, 1024));
  if (f32)
    hipLaunchKernelGGL(k_zero_outside_disk<float>, grid, dim3(bs), 0, st, g, static_cast<cplx<float>*>(field));
  else
    hipLaunchKernelGGL(k_zero_outside_disk<double>, grid, dim3(bs), 0, st, g, static_cast<cplx<double>*>(field));
}

// Zero fill by a kernel (not hipMemsetAsync) so the step stays capturable.
void launch_zero_bytes(void* p, size_t bytes, hipStream_t st) {
  size_t words = bytes / 4;
  unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 4096);
  if (blocks == 0) return;
  hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, st, static_cast<unsigned int*>(p), words);
}

void launch_zero(cplx<double>* p, size_t n, hipStream_t st) { launch_zero_bytes(p, n * sizeof(cplx<double>), st); }

void launch_q_from_psi(const Geo& g, hipStream_t st, const void* psi, void* q, bool f32, const double* a,
                       int barotropic, bool psi_f64) {
  const int plane = g.nx * g.ny;
  const int bs = 64;
  const dim3 grid((plane + bs - 1) / bs, g.members);
  if (f32 && psi_f64)
    hipLaunchKernelGGL((k_q_from_psi<double, float>), grid, dim3(bs), 0, st, g,
                       static_cast<const cplx<double>*>(psi), static_cast<cplx<float>*>(q), a, barotropic);
  else if (f32)
    hipLaunchKernelGGL((k_q_from_psi<float, float>), grid, dim3(bs), 0, st, g, static_cast<const cplx<float>*>(psi),
                       static_cast<cplx<float>*>(q), a, barotropic);
  else
    hipLaunchKernelGGL((k_q_from_psi<double, double>), grid, dim3(bs), 0, st, g,
                       static_cast<const cplx<double>*>(psi), static_cast<cplx<double>*>(q), a, barotropic);
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

}  // namespace qg
