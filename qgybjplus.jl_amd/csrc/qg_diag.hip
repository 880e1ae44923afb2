// Energy diagnostics on the device: _local_energy_components
// (src/core/io.jl:331-395) as a deterministic two-stage reduction.  A thread
// walks one (kx, ky) column of one member through every level; a block sums
// its columns in a fixed tree order into one partial per component, and a
// second kernel adds the blocks' partials in block order.  The caller passes
// psi, A and C recomputed from q and B the way _record_energy_diagnostics!
// (src/core/io.jl:405-419) does.
#include "qg_columns.hpp"

#pragma clang fp contract(off)

namespace qg {

namespace {

constexpr int kEnergyThreads = 256;
constexpr int kEnergyTerms = 5;  // wave KE, PE, CE, mean-flow KE, PE

__device__ inline double abs2(cplx<double> v) { return v.re * v.re + v.im * v.im; }

// blockIdx.y: member (m0 + blockIdx.y); partials [member][block][5].
template <typename T>
__global__ void __launch_bounds__(kEnergyThreads)
k_energy_partial(Geo g, int m0, const cplx<T>* __restrict__ psi, const cplx<T>* __restrict__ A,
                 const cplx<T>* __restrict__ C, const double* __restrict__ a, double* __restrict__ partial) {
  const int plane = g.nx * g.ny;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int nz = g.nzc;
  const size_t mem = (size_t)m0 + blockIdx.y;
  double acc[kEnergyTerms] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (col < plane && g.mask[col]) {
    const int x = col % g.nx, y = col / g.nx;
    psi += mem * nz * plane;
    A += mem * nz * plane;
    C += mem * nz * plane;
    a += mem * nz;
    const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
    const double w = (x == 0 && y == 0) ? 0.5 : 1.0;
    const double dz = g.dz;
    auto at = [&](int k) { return (size_t)k * plane + col; };
    for (int k = 0; k < nz; ++k) {
      const double ak = a[k];
      const cplx<double> p = cvt<double>(psi[at(k)]);
      acc[3] += w * kh2 * abs2(p);
      if (k < nz - 1) {
        const cplx<double> pu = cvt<double>(psi[at(k + 1)]);
        const cplx<double> pz = mk((pu.re - p.re) / dz, (pu.im - p.im) / dz);
        acc[4] += w * ak * abs2(pz);
      }
      const cplx<double> ck = cvt<double>(C[at(k)]);
      cplx<double> la = mk(0.0, 0.0);
      if (nz > 1) {
        if (k == 0) {
          la = mk((ak * ck.re) / dz, (ak * ck.im) / dz);
        } else {
          const cplx<double> cm = cvt<double>(C[at(k - 1)]);
          const double am = a[k - 1];
          la = (k == nz - 1) ? mk((-am * cm.re) / dz, (-am * cm.im) / dz)
                             : mk((ak * ck.re - am * cm.re) / dz, (ak * ck.im - am * cm.im) / dz);
        }
      }
      acc[0] += w * abs2(la);
      acc[1] += w * (0.5 / ak) * kh2 * abs2(ck);
      acc[2] += w * (1 / (8 * (ak * ak))) * (kh2 * kh2) * abs2(cvt<double>(A[at(k)]));
    }
  }
  __shared__ double red[kEnergyTerms][kEnergyThreads];
  for (int i = 0; i < kEnergyTerms; ++i) red[i][threadIdx.x] = acc[i];
  __syncthreads();
  for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int i = 0; i < kEnergyTerms; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < kEnergyTerms)
    partial[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kEnergyTerms + threadIdx.x] = red[threadIdx.x][0];
}

// out[m][i] = normalization * sum over blocks, in block order.
__global__ void k_energy_final(int nblocks, int nmem, double norm, const double* __restrict__ partial,
                               double* __restrict__ out) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nmem * kEnergyTerms) return;
  const int m = id / kEnergyTerms, i = id % kEnergyTerms;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += partial[((size_t)m * nblocks + b) * kEnergyTerms + i];
  out[id] = s * norm;
}

// max hypot(u, v) over n points (_progress_maxima's flow speed,
// src/simulation.jl:71-76): grid-stride, one atomic per block, skipped when
// the accumulator already holds more (non-negative doubles compare like their
// bit patterns; NaN counts as +inf).
template <typename T>
__global__ void __launch_bounds__(kEnergyThreads)
k_max_hypot(const T* __restrict__ u, const T* __restrict__ v, size_t n, unsigned long long* __restrict__ out) {
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double h = hypot((double)u[i], (double)v[i]);
    m = (h > m || h != h) ? h : m;
  }
  if (m != m) m = __builtin_huge_val();
  __shared__ double red[kEnergyThreads];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x + s] > red[threadIdx.x] ? red[threadIdx.x + s]
                                                                                       : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] > 0.0) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(red[0]);
    if (bits > __hip_atomic_load(out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(out, bits);
  }
}

}  // namespace

void launch_max_hypot(const void* u, const void* v, bool f32, size_t n, unsigned long long* out, hipStream_t st) {
  const int blocks = 1024;
  if (f32)
    hipLaunchKernelGGL(k_max_hypot<float>, dim3(blocks), dim3(kEnergyThreads), 0, st, static_cast<const float*>(u),
                       static_cast<const float*>(v), n, out);
  else
    hipLaunchKernelGGL(k_max_hypot<double>, dim3(blocks), dim3(kEnergyThreads), 0, st, static_cast<const double*>(u),
                       static_cast<const double*>(v), n, out);
}

size_t energy_partial_size(const Geo& g, int nmem) {
  const int plane = g.nx * g.ny;
  return (size_t)nmem * ((plane + kEnergyThreads - 1) / kEnergyThreads) * kEnergyTerms;
}

void launch_energies(const Geo& g, hipStream_t st, int m0, int nmem, const void* psi, const void* A,
                     const void* C, bool f32, const double* a, double* partial, double* out) {
  const int plane = g.nx * g.ny;
  const int nblocks = (plane + kEnergyThreads - 1) / kEnergyThreads;
  if (f32)
    hipLaunchKernelGGL(k_energy_partial<float>, dim3(nblocks, nmem), dim3(kEnergyThreads), 0, st, g, m0,
                       static_cast<const cplx<float>*>(psi), static_cast<const cplx<float>*>(A),
                       static_cast<const cplx<float>*>(C), a, partial);
  else
    hipLaunchKernelGGL(k_energy_partial<double>, dim3(nblocks, nmem), dim3(kEnergyThreads), 0, st, g, m0,
                       static_cast<const cplx<double>*>(psi), static_cast<const cplx<double>*>(A),
                       static_cast<const cplx<double>*>(C), a, partial);
  // normalization 0.5 / ((nx*ny)^2 * nz), as the reference forms it
  const double norm = 0.5 / ((double)((long long)plane * plane) * g.nzc);
  const int n = nmem * kEnergyTerms;
  hipLaunchKernelGGL(k_energy_final, dim3((n + 63) / 64), dim3(64), 0, st, nblocks, nmem, norm, partial, out);
}

}  // namespace qg
