// Batched Thomas solves of the q->psi and B->A column operators.
//
// The elimination of each column depends only on (kh2, a, dz), so
// k_thomas_factor runs it once per model and records per level the pivot
// denom_k, its correctly rounded reciprocal and c'_k (the reference's
// thomas_solve! recomputes them on every call, src/elliptic.jl:1168-1202).
// Columns (kx, ky) and (-kx, -ky), (-kx, ky), (kx, -ky) share kh2, so the
// tables are kept for the quarter (|ix| <= nx/2, |iy| <= ny/2) only.
//
// Per solve, a block owns a tile of TC adjacent columns: all 256 threads stage
// the right-hand sides [nz][TC] into LDS with coalesced loads, 2*TC lanes run
// the recurrences in place (one lane for the real and one for the imaginary
// part, which the reference solves separately), and all threads write back.
// Each forward step's quotient (b_k - dl_k x_{k-1}) / denom_k is formed from
// the stored reciprocal with one FMA correction, which is the correctly
// rounded quotient, so the division leaves the loop-carried chain.
#include "qg_columns.hpp"

#pragma clang fp contract(off)

namespace qg {

namespace {

constexpr int kThreads = 256;
constexpr int kPF = 16;     // table entries loaded ahead in the sweeps
constexpr int kBatch = 16;  // staged loads in flight per thread
constexpr double kSqrtEps = 1.4901161193847656e-08;  // sqrt(eps(Float64))

__device__ inline void flag_error(int* err, int code, int level, int col) {
  if (atomicCAS(err, 0, code) == 0) {
    err[1] = level;
    err[2] = col;
  }
}

// n / d given y = RN(1/d): q0 = n*y, then one residual correction.
__device__ inline double quot(double n, double d, double y) {
  const double q0 = n * y;
  const double e = fma(-d, q0, n);
  return fma(e, y, q0);
}

struct Quarter {
  int Wr, Hr;  // reduced widths nx/2+1, ny/2+1
  __device__ inline int rcol(int x, int y, int nx, int ny) const {
    const int rx = x <= nx / 2 ? x : nx - x;
    const int ry = y <= ny / 2 ? y : ny - y;
    return ry * Wr + rx;
  }
};

// Tables [k][rcol]: den (pivot), inv = 1/den, cp = c'.  Pivots below
// sqrt(eps) are recorded in err (code 1) where thomas_solve! raises.
__global__ void k_thomas_factor(Geo g, Quarter qr, const double* __restrict__ a, int kdiv,
                                double* __restrict__ den, double* __restrict__ inv, double* __restrict__ cp,
                                int* __restrict__ err) {
  const int rplane = qr.Wr * qr.Hr;
  const int rc = blockIdx.x * blockDim.x + threadIdx.x;
  if (rc >= rplane) return;
  const int x = rc % qr.Wr, y = rc / qr.Wr;
  const int nz = g.nz;
  const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
  const double dz2 = g.dz * g.dz;
  auto at = [&](int k) { return (size_t)k * rplane + rc; };
  if (nz == 1 || (kh2 == 0.0 && kdiv == 1)) {
    for (int k = 0; k < nz; ++k) {
      den[at(k)] = 1.0;
      inv[at(k)] = 1.0;
      cp[at(k)] = 0.0;
    }
    return;
  }
  const bool gauge = kh2 == 0.0;  // B -> A mean column: A[1] = 0 row
  const double t = gauge ? 0.0 : ((kdiv == 1) ? kh2 * dz2 : (kh2 * dz2) / 4);
  const double d0 = gauge ? 1.0 : -(a[0] + t);
  const double du0 = gauge ? 0.0 : a[0];
  if (fabs(d0) < kSqrtEps) flag_error(err, 1, 0, y * g.nx + x);
  double c = du0 / d0;
  den[at(0)] = d0;
  inv[at(0)] = 1.0 / d0;
  cp[at(0)] = c;
  for (int i = 1; i < nz; ++i) {
    const double dl = a[i - 1];
    double d;
    if (gauge) d = (i < nz - 1) ? -(a[i] + a[i - 1]) : -a[nz - 2];
    else d = (i < nz - 1) ? -(a[i] + a[i - 1] + t) : -(a[nz - 2] + t);
    const double denom = d - dl * c;
    if (fabs(denom) < kSqrtEps) flag_error(err, 1, i, y * g.nx + x);
    const double du = (i < nz - 1) ? a[i] : 0.0;
    c = du / denom;
    den[at(i)] = denom;
    inv[at(i)] = 1.0 / denom;
    cp[at(i)] = c;
  }
}

__device__ inline double& part_ref(cplx<double>* xs, int k, int TC, int c, int part) {
  cplx<double>& v = xs[(size_t)k * TC + c];
  return part ? v.im : v.re;
}

__global__ void __launch_bounds__(kThreads)
k_thomas_tile(Geo g, Quarter qr, const cplx<double>* __restrict__ rhs, cplx<double>* __restrict__ out,
              cplx<double>* __restrict__ Cout, const double* __restrict__ den, const double* __restrict__ inv,
              const double* __restrict__ cp, const double* __restrict__ a, int kdiv, int TC,
              unsigned long long* __restrict__ warn) {
  extern __shared__ cplx<double> xs[];  // [nz][TC] then a[nz]
  const int plane = g.nx * g.ny;
  const int rplane = qr.Wr * qr.Hr;
  const int nz = g.nz;
  const int col0 = blockIdx.x * TC;
  const int ncol = min(TC, plane - col0);
  const double dz = g.dz;
  const double dz2 = dz * dz;
  double* al = reinterpret_cast<double*>(xs + (size_t)nz * TC);

  // phase 1: stage the right-hand sides (coalesced along x) and a(z)
  for (int base = 0; base < nz * TC; base += kThreads * kBatch) {
    cplx<double> v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int e = base + u * kThreads + threadIdx.x;
      const int k = e / TC, c = e % TC;
      v[u] = (e < nz * TC && c < ncol) ? rhs[(size_t)k * plane + col0 + c] : mk(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int e = base + u * kThreads + threadIdx.x;
      if (e < nz * TC) xs[e] = v[u];
    }
  }
  for (int k = threadIdx.x; k < nz; k += kThreads) al[k] = a[k];
  __syncthreads();

  // phase 2: recurrences, lanes [0, 2*TC)
  const int lane = threadIdx.x;
  if (lane < 2 * TC) {
    const int c = lane % TC, part = lane / TC;
    const int col = col0 + c;
    if (c < ncol) {
      const int x = col % g.nx, y = col / g.nx;
      const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
      const int rc = qr.rcol(x, y, g.nx, g.ny);
      auto at = [&](int k) { return (size_t)k * rplane + rc; };
      if (nz == 1) {
        double& v = part_ref(xs, 0, TC, c, part);
        if (kh2 == 0.0) v = 0.0;
        else if (kdiv == 1) v = -v / kh2;
        else v = (-4 * v) / kh2;
      } else if (kh2 == 0.0 && kdiv == 1) {
        // psi = 0 on the horizontal mean (src/elliptic.jl:257-270); warn on a
        // significant mean of q.
        if (part == 0 && warn) {
          double m = 0.0;
          for (int k = 0; k < nz; ++k) {
            cplx<double> v = xs[(size_t)k * TC + c];
            double av = hypot(v.re, v.im);
            m = av > m ? av : m;
          }
          if (m > 1e-10) atomicMax(warn, (unsigned long long)__double_as_longlong(m));
        }
      } else {
        const bool gauge = kh2 == 0.0;
        // forward sweep: x'_k = (dz2 b_k - a_{k-1} x'_{k-1}) / denom_k, in
        // place; kPF levels move between LDS and registers at a time so the
        // LDS latency is paid once per batch, not once per level.
        double xprev = 0.0;
        for (int i0 = 0; i0 < nz; i0 += kPF) {
          double dd[kPF], yy[kPF], bb[kPF], aa[kPF];
#pragma unroll
          for (int j = 0; j < kPF; ++j) {
            const int i = i0 + j;
            const bool in = i < nz;
            dd[j] = in ? den[at(i)] : 1.0;
            yy[j] = in ? inv[at(i)] : 1.0;
            bb[j] = in ? part_ref(xs, i, TC, c, part) : 0.0;
            aa[j] = (in && i > 0) ? al[i - 1] : 0.0;
          }
#pragma unroll
          for (int j = 0; j < kPF; ++j) {
            const int i = i0 + j;
            double xv;
            if (i == 0) xv = gauge ? 0.0 / 1.0 : quot(dz2 * bb[j], dd[j], yy[j]);
            else xv = quot(dz2 * bb[j] - aa[j] * xprev, dd[j], yy[j]);
            if (i < nz) xprev = xv;
            bb[j] = xv;
          }
#pragma unroll
          for (int j = 0; j < kPF; ++j)
            if (i0 + j < nz) part_ref(xs, i0 + j, TC, c, part) = bb[j];
        }
        // back substitution
        double above = xprev;
        for (int i0 = nz - 2; i0 >= 0; i0 -= kPF) {
          double cc[kPF], xx[kPF];
#pragma unroll
          for (int j = 0; j < kPF; ++j) {
            const int i = i0 - j;
            cc[j] = i >= 0 ? cp[at(i)] : 0.0;
            xx[j] = i >= 0 ? part_ref(xs, i, TC, c, part) : 0.0;
          }
#pragma unroll
          for (int j = 0; j < kPF; ++j) {
            const double xi = xx[j] - cc[j] * above;
            if (i0 - j >= 0) above = xi;
            xx[j] = xi;
          }
#pragma unroll
          for (int j = 0; j < kPF; ++j)
            if (i0 - j >= 0) part_ref(xs, i0 - j, TC, c, part) = xx[j];
        }
        if (gauge) {  // remove the vertical mean (src/elliptic.jl:930-945)
          double mean = 0.0;
          for (int k = 0; k < nz; ++k) mean = mean + part_ref(xs, k, TC, c, part);
          mean = mean / nz;
          for (int k = 0; k < nz; ++k) {
            double& slot = part_ref(xs, k, TC, c, part);
            slot = slot - mean;
          }
        }
      }
    }
  }
  __syncthreads();

  // phase 3: write back (psi = 0 on the mean column), and C = dA/dz
  for (int e = threadIdx.x; e < nz * TC; e += kThreads) {
    const int k = e / TC, c = e % TC;
    if (c >= ncol) continue;
    const int col = col0 + c;
    const int x = col % g.nx, y = col / g.nx;
    const bool zero_col = kdiv == 1 && nz > 1 && g.kx[x] == 0.0 && g.ky[y] == 0.0;
    cplx<double> v = zero_col ? mk(0.0, 0.0) : xs[e];
    out[(size_t)k * plane + col] = v;
    if (Cout) {
      cplx<double> cv = mk(0.0, 0.0);
      if (k < nz - 1) {
        cplx<double> up = xs[e + TC];
        cv = mk((up.re - v.re) / dz, (up.im - v.im) / dz);
      }
      Cout[(size_t)k * plane + col] = cv;
    }
  }
}

Quarter quarter(const Geo& g) { return Quarter{g.nx / 2 + 1, g.ny / 2 + 1}; }

}  // namespace

size_t thomas_table_size(const Geo& g) { return (size_t)(g.nx / 2 + 1) * (g.ny / 2 + 1) * g.nz; }

void launch_thomas_factor(const Geo& g, hipStream_t st, const double* a, int kdiv, double* den, double* inv,
                          double* cp, int* err) {
  const Quarter qr = quarter(g);
  const int rplane = qr.Wr * qr.Hr;
  const int bs = 64;
  hipLaunchKernelGGL(k_thomas_factor, dim3((rplane + bs - 1) / bs), dim3(bs), 0, st, g, qr, a, kdiv, den, inv,
                     cp, err);
}

int thomas_tile_columns(int nz) {
  int tc = 64;
  while (tc > 1 && (size_t)tc * nz * sizeof(cplx<double>) + nz * sizeof(double) > kThomasLdsBytes) tc /= 2;
  return tc;
}

bool thomas_lds_setup() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(k_thomas_tile),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)kThomasLdsBytes) == hipSuccess;
}

void launch_thomas_lds(const Geo& g, hipStream_t st, const cplx<double>* rhs, cplx<double>* out,
                       cplx<double>* Cout, const double* den, const double* inv, const double* cp,
                       const double* a, int kdiv, unsigned long long* warn) {
  const int plane = g.nx * g.ny;
  const int tc = thomas_tile_columns(g.nz);
  const size_t lds = (size_t)tc * g.nz * sizeof(cplx<double>) + g.nz * sizeof(double);
  hipLaunchKernelGGL(k_thomas_tile, dim3((plane + tc - 1) / tc), dim3(kThreads), lds, st, g, quarter(g), rhs,
                     out, Cout, den, inv, cp, a, kdiv, tc, warn);
}

}  // namespace qg
