// Device-side initial conditions of set!: the seeded random streamfunction of
// set_mean_flow!(method = :random) and the surface wave of
// set_surface_waves!, written straight into the spectral fields so that a run
// (or an ensemble member) is set up without a host-side field or an upload.
#include "qg_columns.hpp"

#pragma clang fp contract(off)

namespace qg {

namespace {

// _mix64 / _hash_u01 / _hash_randn (src/parallel_mpi.jl:1115-1137): a
// SplitMix64 finalizer applied to the seed xor each index, then the tag.
__device__ inline unsigned long long mix64(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

__device__ inline double hash_u01(unsigned long long seed, int a, int b, int c, unsigned long long tag) {
  unsigned long long x = seed;
  x = mix64(x ^ (unsigned long long)a);
  x = mix64(x ^ (unsigned long long)b);
  x = mix64(x ^ (unsigned long long)c);
  x = mix64(x ^ tag);
  return (double)((x >> 11) + 1) * 0x1.0p-53;
}

__device__ inline double hash_randn(unsigned long long seed, int a, int b, int c, unsigned long long tag) {
  const double u1 = hash_u01(seed, a, b, c, tag);
  const double u2 = hash_u01(seed, a, b, c, tag + 1);
  const double two_pi = 6.283185307179586;  // Float64(2 * pi)
  return sqrt(-2 * log(u1)) * cos(two_pi * u2);
}

// init_mpi_random_psi! (src/parallel_mpi.jl:1159-1216) for members
// [m0, m0 + gridDim.y / nzc): psi-hat with energy amplitude * |k|^slope per
// mode, Hermitian by construction, then the 2/3 mask.  Indices are the
// reference's 1-based global (level, x, y) of one model.
template <typename T>
__global__ void k_random_psi(Geo g, int m0, unsigned long long base_seed, double amplitude, double slope,
                             cplx<T>* __restrict__ psi) {
  const int plane = g.nx * g.ny;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane) return;
  const int x = col % g.nx, y = col / g.nx;
  const int lvl = blockIdx.y;  // level among the set members
  const int member = m0 + lvl / g.nzc, k = lvl % g.nzc;
  const size_t at = ((size_t)member * g.nzc + k) * plane + col;
  const int ig = x + 1, jg = y + 1, kg = k + 1;
  const int kx = ig <= g.nx / 2 ? ig - 1 : ig - 1 - g.nx;
  const int ky = jg <= g.ny / 2 ? jg - 1 : jg - 1 - g.ny;
  cplx<double> v = mk(0.0, 0.0);
  if (!(kx == 0 && ky == 0) && g.mask[col]) {
    const double k_total = sqrt((double)(kx * kx + ky * ky));
    const double energy = amplitude * pow(k_total, slope);
    const int ic = ig == 1 ? 1 : g.nx - ig + 2;
    const int jc = jg == 1 ? 1 : g.ny - jg + 2;
    if (ig == ic && jg == jc) {
      v.re = sqrt(2 * energy) * hash_randn(base_seed, kg, ig, jg, 0x10);
    } else {
      const bool canonical = (ig < ic) || (ig == ic && jg <= jc);
      const int ican = canonical ? ig : ic, jcan = canonical ? jg : jc;
      const double amp = sqrt(2 * energy) * hash_randn(base_seed, kg, ican, jcan, 0x20);
      const double phase = 6.283185307179586 * hash_u01(base_seed, kg, ican, jcan, 0x30);
      v = mk(amp * cos(phase), amp * sin(phase));
      if (!canonical) v.im = -v.im;
    }
  }
  psi[at] = cvt<T>(v);
}

// set_surface_waves! (src/simulation.jl:292-329): a horizontally uniform
// physical B(z) = profile[k] whose forward transform is nx*ny*profile[k] on the
// kh = 0 mode and zero elsewhere.  profile has nzc entries per set member.
template <typename T>
__global__ void k_surface_wave(Geo g, int m0, const double* __restrict__ profile, cplx<T>* __restrict__ B) {
  const int plane = g.nx * g.ny;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane) return;
  const int lvl = blockIdx.y;
  const size_t at = ((size_t)m0 * g.nzc + lvl) * plane + col;
  B[at] = cvt<T>(col == 0 ? mk((double)plane * profile[lvl], 0.0) : mk(0.0, 0.0));
}

}  // namespace

unsigned long long random_psi_base_seed(unsigned long long seed) {
  // _mix64(UInt64(seed) xor UInt64(seed_offset)) with seed_offset = 0
  unsigned long long x = seed + 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

void launch_random_psi(const Geo& g, hipStream_t st, int m0, int nm, unsigned long long seed, double amplitude,
                       double slope, void* psi, bool f32) {
  const int plane = g.nx * g.ny, bs = 256;
  const dim3 grid((plane + bs - 1) / bs, nm * g.nzc);
  const unsigned long long base = random_psi_base_seed(seed);
  if (f32)
    hipLaunchKernelGGL(k_random_psi<float>, grid, dim3(bs), 0, st, g, m0, base, amplitude, slope,
                       static_cast<cplx<float>*>(psi));
  else
    hipLaunchKernelGGL(k_random_psi<double>, grid, dim3(bs), 0, st, g, m0, base, amplitude, slope,
                       static_cast<cplx<double>*>(psi));
}

void launch_surface_wave(const Geo& g, hipStream_t st, int m0, int nm, const double* profile, void* B, bool f32) {
  const int plane = g.nx * g.ny, bs = 256;
  const dim3 grid((plane + bs - 1) / bs, nm * g.nzc);
  if (f32)
    hipLaunchKernelGGL(k_surface_wave<float>, grid, dim3(bs), 0, st, g, m0, profile, static_cast<cplx<float>*>(B));
  else
    hipLaunchKernelGGL(k_surface_wave<double>, grid, dim3(bs), 0, st, g, m0, profile, static_cast<cplx<double>*>(B));
}

}  // namespace qg
