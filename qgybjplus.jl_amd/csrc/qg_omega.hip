// Interior solve of the QG omega equation (operators.jl:415-470), batched
// over columns.
//
// The reference calls LAPACK dgtsv (Gaussian elimination with partial
// pivoting) once for the real and once for the imaginary part of each column.
// The elimination depends only on (kh2, f^2/N^2, dz), so k_omega_factor runs
// it once per model and records, per interior row, the multiplier and the
// eliminated diagonal; a column whose elimination interchanged rows is marked
// and solved by the general path.  Per step, k_omega_tile stages a tile of
// right-hand sides in LDS, applies the recorded elimination to them and back
// substitutes: the same operations on the right-hand side, in the same order,
// as dgtsv.  FMA contraction is off.
#include "qg_columns.hpp"

#pragma clang fp contract(off)

namespace qg {

namespace {

constexpr int kThreads = 256;
constexpr int kPF = 16;
constexpr int kBatch = 16;

// n / d given y = RN(1/d): q0 = n*y, then one residual correction.
__device__ inline double quot(double n, double d, double y) {
  const double q0 = n * y;
  const double e = fma(-d, q0, n);
  return fma(e, y, q0);
}

__device__ inline void flag_error(int* err, int code, int level, int col) {
  if (atomicCAS(err, 0, code) == 0) {
    err[1] = level;
    err[2] = col;
  }
}

// Per column, rows iz = 0..n-1 (levels k = iz+1), n = nz-2:
//   fact[iz] (iz < n-1): multiplier applied to row iz+1,
//   dp[iz]: eliminated diagonal,  pivoted[col]: any row interchange.
__global__ void k_omega_factor(Geo g, const double* __restrict__ cz, double* __restrict__ fact,
                               double* __restrict__ dp, double* __restrict__ dinv,
                               unsigned char* __restrict__ pivoted, int* __restrict__ err) {
  const int plane = g.nx * g.ny;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane) return;
  const int x = col % g.nx, y = col / g.nx;
  const int nz = g.nz;
  const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
  pivoted[col] = 0;
  if (!(kh2 > 0.0 && nz > 2)) return;
  const int n = nz - 2;
  auto at = [&](int iz) { return (size_t)iz * plane + col; };
  double d = -2 * cz[1] - kh2;
  double du = (n > 1) ? cz[1] : 0.0;
  bool piv = false;
  for (int i = 0; i < n - 1; ++i) {
    const double dl = cz[i + 2];
    double dn = -2 * cz[i + 2] - kh2;
    double dun = (i + 1 < n - 1) ? cz[i + 2] : 0.0;
    if (fabs(d) >= fabs(dl)) {
      if (d == 0.0) flag_error(err, 2, i, col);
      const double f = dl / d;
      dn = dn - f * du;
      fact[at(i)] = f;
    } else {
      piv = true;
      const double f = d / dl;
      const double temp = dn;
      dn = du - f * temp;
      if (i < n - 2) dun = -f * dun;
      du = temp;
      d = dl;
      fact[at(i)] = f;
    }
    dp[at(i)] = d;
    dinv[at(i)] = 1.0 / d;
    d = dn;
    du = dun;
  }
  if (d == 0.0) flag_error(err, 2, n - 1, col);
  dp[at(n - 1)] = d;
  dinv[at(n - 1)] = 1.0 / d;
  pivoted[col] = piv ? 1 : 0;
}

// General dgtsv path (row interchanges) for one column; b (complex) in LDS
// with row stride TC2 doubles, eliminated rows kept in global scratch.
__device__ void gtsv_general(double* b0, int TC2, int n, double kh2, const double* cz, double* sd, double* sdu,
                             double* sfill, int plane, int col) {
  auto at = [&](int iz) { return (size_t)iz * plane + col; };
  auto Br = [&](int iz) -> double& { return b0[(size_t)iz * TC2]; };
  auto Bi = [&](int iz) -> double& { return b0[(size_t)iz * TC2 + 1]; };
  double d = -2 * cz[1] - kh2;
  double du = (n > 1) ? cz[1] : 0.0;
  double br = Br(0), bi = Bi(0);
  for (int i = 0; i < n - 1; ++i) {
    const double dl = cz[i + 2];
    double dn = -2 * cz[i + 2] - kh2;
    double dun = (i + 1 < n - 1) ? cz[i + 2] : 0.0;
    double bnr = Br(i + 1), bni = Bi(i + 1);
    double fill = 0.0;
    if (fabs(d) >= fabs(dl)) {
      const double f = dl / d;
      dn = dn - f * du;
      bnr = bnr - f * br;
      bni = bni - f * bi;
    } else {
      const double f = d / dl;
      const double temp = dn;
      dn = du - f * temp;
      if (i < n - 2) {
        fill = dun;
        dun = -f * fill;
      }
      du = temp;
      d = dl;
      const double tr = br, ti = bi;
      br = bnr;
      bi = bni;
      bnr = tr - f * bnr;
      bni = ti - f * bni;
    }
    sd[at(i)] = d;
    sdu[at(i)] = du;
    sfill[at(i)] = fill;
    Br(i) = br;
    Bi(i) = bi;
    d = dn;
    du = dun;
    br = bnr;
    bi = bni;
  }
  double xr = br / d, xi = bi / d;
  Br(n - 1) = xr;
  Bi(n - 1) = xi;
  if (n > 1) {
    const int i = n - 2;
    double x0r = (Br(i) - sdu[at(i)] * xr) / sd[at(i)];
    double x0i = (Bi(i) - sdu[at(i)] * xi) / sd[at(i)];
    Br(i) = x0r;
    Bi(i) = x0i;
    double x2r = xr, x2i = xi;
    xr = x0r;
    xi = x0i;
    for (int ii = n - 3; ii >= 0; --ii) {
      const double dd = sd[at(ii)], uu = sdu[at(ii)], ff = sfill[at(ii)];
      double vr = (Br(ii) - uu * xr - ff * x2r) / dd;
      double vi = (Bi(ii) - uu * xi - ff * x2i) / dd;
      Br(ii) = vr;
      Bi(ii) = vi;
      x2r = xr;
      x2i = xi;
      xr = vr;
      xi = vi;
    }
  }
}

__global__ void __launch_bounds__(kThreads)
k_omega_tile(Geo g, const cplx<double>* __restrict__ rhsw, cplx<double>* __restrict__ wk,
             const double* __restrict__ fact, const double* __restrict__ dp,
             const double* __restrict__ dinv, const unsigned char* __restrict__ pivoted, const double* __restrict__ cz,
             const double* __restrict__ fN2, double* __restrict__ sd, double* __restrict__ sdu,
             double* __restrict__ sfill, int TC) {
  extern __shared__ cplx<double> bs[];  // [nz][TC] then cz[nz]
  const int plane = g.nx * g.ny;
  const int nz = g.nz;
  const int n = nz - 2;
  const int col0 = blockIdx.x * TC;
  const int ncol = min(TC, plane - col0);
  double* czl = reinterpret_cast<double*>(bs + (size_t)nz * TC);

  // phase 1: b = (f/N2[k]) * RHS[k] on interior levels (operators.jl:439)
  for (int base = 0; base < nz * TC; base += kThreads * kBatch) {
    cplx<double> v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int e = base + u * kThreads + threadIdx.x;
      const int k = e / TC, c = e % TC;
      v[u] = (e < nz * TC && c < ncol && k >= 1 && k <= nz - 2) ? rhsw[(size_t)k * plane + col0 + c]
                                                                 : mk(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int e = base + u * kThreads + threadIdx.x;
      if (e < nz * TC) {
        const double s = fN2[e / TC];
        bs[e] = mk(s * v[u].re, s * v[u].im);
      }
    }
  }
  for (int k = threadIdx.x; k < nz; k += kThreads) czl[k] = cz[k];
  __syncthreads();

  const int lane = threadIdx.x;
  if (lane < 2 * TC && n >= 1) {
    const int c = lane % TC, part = lane / TC;
    const int col = col0 + c;
    if (c < ncol) {
      const int x = col % g.nx, y = col / g.nx;
      const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
      double* b0 = reinterpret_cast<double*>(bs + (size_t)1 * TC + c) + part;  // row iz -> b0[iz*2TC]
      const int TC2 = 2 * TC;
      auto B = [&](int iz) -> double& { return b0[(size_t)iz * TC2]; };
      auto at = [&](int iz) { return (size_t)iz * plane + col; };
      if (!(kh2 > 0.0)) {
        for (int iz = 0; iz < n; ++iz) B(iz) = 0.0;
      } else if (pivoted[col]) {
        if (part == 0) gtsv_general(b0, TC2, n, kh2, czl, sd, sdu, sfill, plane, col);
      } else {
        // forward: b[i+1] -= fact[i] * b[i]; kPF rows between LDS and
        // registers at a time
        double bprev = B(0);
        for (int i0 = 0; i0 < n - 1; i0 += kPF) {
          double ff[kPF], bb[kPF];
#pragma unroll
          for (int j = 0; j < kPF; ++j) {
            const bool in = i0 + j < n - 1;
            ff[j] = in ? fact[at(i0 + j)] : 0.0;
            bb[j] = in ? B(i0 + j + 1) : 0.0;
          }
#pragma unroll
          for (int j = 0; j < kPF; ++j) {
            const double bn = bb[j] - ff[j] * bprev;
            if (i0 + j < n - 1) bprev = bn;
            bb[j] = bn;
          }
#pragma unroll
          for (int j = 0; j < kPF; ++j)
            if (i0 + j < n - 1) B(i0 + j + 1) = bb[j];
        }
        // back substitution: du[i] = c_i = cz[i+1] (no interchange)
        double x1 = quot(bprev, dp[at(n - 1)], dinv[at(n - 1)]);
        B(n - 1) = x1;
        for (int i0 = n - 2; i0 >= 0; i0 -= kPF) {
          double dd[kPF], yy[kPF], bb[kPF], uu[kPF];
#pragma unroll
          for (int j = 0; j < kPF; ++j) {
            const int i = i0 - j;
            const bool in = i >= 0;
            dd[j] = in ? dp[at(i)] : 1.0;
            yy[j] = in ? dinv[at(i)] : 1.0;
            bb[j] = in ? B(i) : 0.0;
            uu[j] = in ? czl[i + 1] : 0.0;
          }
#pragma unroll
          for (int j = 0; j < kPF; ++j) {
            const double xv = quot(bb[j] - uu[j] * x1, dd[j], yy[j]);
            if (i0 - j >= 0) x1 = xv;
            bb[j] = xv;
          }
#pragma unroll
          for (int j = 0; j < kPF; ++j)
            if (i0 - j >= 0) B(i0 - j) = bb[j];
        }
      }
    }
  }
  __syncthreads();

  // phase 3: w-hat; zero on the end levels and on kh = 0
  for (int e = threadIdx.x; e < nz * TC; e += kThreads) {
    const int k = e / TC, c = e % TC;
    if (c >= ncol) continue;
    wk[(size_t)k * plane + col0 + c] = bs[e];
  }
}

}  // namespace

void launch_omega_factor(const Geo& g, hipStream_t st, const double* cz, double* fact, double* dp, double* dinv,
                         unsigned char* pivoted, int* err) {
  const int plane = g.nx * g.ny;
  const int bs = 64;
  hipLaunchKernelGGL(k_omega_factor, dim3((plane + bs - 1) / bs), dim3(bs), 0, st, g, cz, fact, dp, dinv,
                     pivoted, err);
}

bool omega_lds_setup() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(k_omega_tile),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)kThomasLdsBytes) == hipSuccess;
}

void launch_omega_tile(const Geo& g, hipStream_t st, const cplx<double>* rhsw, cplx<double>* wk,
                       const double* fact, const double* dp, const double* dinv, const unsigned char* pivoted,
                       const double* cz, const double* fN2, double* sd, double* sdu, double* sfill) {
  const int plane = g.nx * g.ny;
  const int tc = thomas_tile_columns(g.nz);
  const size_t lds = (size_t)tc * g.nz * sizeof(cplx<double>) + g.nz * sizeof(double);
  hipLaunchKernelGGL(k_omega_tile, dim3((plane + tc - 1) / tc), dim3(kThreads), lds, st, g, rhsw, wk, fact, dp,
                     dinv, pivoted, cz, fN2, sd, sdu, sfill, tc);
}

}  // namespace qg
