// Interior solve of the QG omega equation (compute_vertical_velocity!,
// src/operators.jl:415-470), batched over columns.
//
// The reference calls LAPACK dgtsv (Gaussian elimination with partial
// pivoting) once for the real and once for the imaginary part of each column.
// The elimination depends only on (kh2, f^2/N^2, dz), so k_omega_factor runs
// it once per model for the quarter columns and records, level-indexed, the
// multiplier applied to each row and the reciprocal of its eliminated pivot;
// a column whose elimination interchanged rows is marked and solved by the
// general path, a restatement of dgtsv.  Per step, k_omega is a mirror-tile
// kernel (qg_mirror.hpp): it applies the recorded elimination to the
// right-hand sides, b_k -= m_k b_{k-1}, and back substitutes
// x_k = (b_k - u_k x_{k+1}) / d_k as fma(-(u_k / d_k), x_{k+1}, b_k / d_k),
// with the quotients formed from the stored reciprocals.
#include "qg_columns.hpp"
#include "qg_mirror.hpp"

#pragma clang fp contract(off)

namespace qg {

namespace {

__device__ inline void flag_error(int* err, int code, int level, int col) {
  if (atomicCAS(err, 0, code) == 0) {
    err[1] = level;
    err[2] = col;
  }
}

// Tables [k][rc] over levels k = 0..nz-1 (interior row iz = k-1):
//   M[k]: multiplier dgtsv applies to row k-1 from row k-2 (0 unless
//         2 <= k <= nz-2),
//   D[k]: 1 / eliminated pivot of row k-1 (0 unless 1 <= k <= nz-2 and
//         kh2 > 0),
//   pivoted[rc]: any row interchange (dgtsv INFO > 0 is recorded in err,
//   code 2).
__global__ void k_omega_factor(Geo g, Quarter qr, const double* __restrict__ cz, double* __restrict__ M,
                               double* __restrict__ D, unsigned char* __restrict__ pivoted,
                               int* __restrict__ err) {
  const int rplane = qr.size();
  const int rc = blockIdx.x * blockDim.x + threadIdx.x;
  if (rc >= rplane) return;
  const int x = rc % qr.Wr, y = rc / qr.Wr;
  const int nz = g.nzc, m = blockIdx.y;  // member m: its own N2(z) and tables
  cz += (size_t)m * nz;
  M += (size_t)m * nz * rplane;
  D += (size_t)m * nz * rplane;
  pivoted += (size_t)m * rplane;
  auto at = [&](int k) { return (size_t)k * rplane + rc; };
  for (int k = 0; k < nz; ++k) {
    M[at(k)] = 0.0;
    D[at(k)] = 0.0;
  }
  pivoted[rc] = 0;
  const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
  if (!(kh2 > 0.0 && nz > 2)) return;
  const int n = nz - 2;
  double d = -2 * cz[1] - kh2;
  double du = (n > 1) ? cz[1] : 0.0;
  bool piv = false;
  for (int i = 0; i < n - 1; ++i) {
    const double dl = cz[i + 2];
    double dn = -2 * cz[i + 2] - kh2;
    double dun = (i + 1 < n - 1) ? cz[i + 2] : 0.0;
    if (fabs(d) >= fabs(dl)) {
      if (d == 0.0) flag_error(err, 2, i, y * g.nx + x);
      const double f = dl / d;
      dn = dn - f * du;
      M[at(i + 2)] = f;
    } else {
      piv = true;
      const double f = d / dl;
      const double temp = dn;
      dn = du - f * temp;
      if (i < n - 2) dun = -f * dun;
      du = temp;
      d = dl;
    }
    D[at(i + 1)] = 1.0 / d;
    d = dn;
    du = dun;
  }
  if (d == 0.0) flag_error(err, 2, n - 1, y * g.nx + x);
  D[at(n)] = 1.0 / d;
  pivoted[rc] = piv ? 1 : 0;
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

// rhsw: the omega right-hand side (complex, all levels); wk: w-hat, zero on
// the end levels and where kh = 0.  blockIdx.y is the member.  cu[k] = cz[k]
// for 1 <= k <= nzc-3 (the upper diagonal of row k-1), zero elsewhere, padded
// to mirror_pad(nzc) per member; cz and fN2 have nzc entries per member.
template <int QC>
__global__ void __launch_bounds__(kMirrorThreads)
k_omega(Geo g, Quarter qr, const cplx<double>* __restrict__ rhsw, cplx<double>* __restrict__ wk,
        const double* __restrict__ M, const double* __restrict__ D, const unsigned char* __restrict__ pivoted,
        const double* __restrict__ cz, const double* __restrict__ cu, const double* __restrict__ fN2,
        double* __restrict__ sd, double* __restrict__ sdu, double* __restrict__ sfill) {
  constexpr int NC = 4 * QC;
  extern __shared__ cplx<double> xs[];  // [kpad][NC], then tm[kpad][QC], td[kpad][QC]
  const int nz = g.nzc, kpad = mirror_pad(nz);
  const int rc0 = blockIdx.x * QC;
  const size_t plane = (size_t)g.nx * g.ny;
  {
    const size_t m = blockIdx.y;
    rhsw += m * nz * plane;
    wk += m * nz * plane;
    M += m * nz * qr.size();
    D += m * nz * qr.size();
    pivoted += m * qr.size();
    cz += m * nz;
    cu += m * kpad;
    fN2 += m * nz;
    sd += m * nz * plane;
    sdu += m * nz * plane;
    sfill += m * nz * plane;
  }
  double* tm = reinterpret_cast<double*>(xs + (size_t)kpad * NC);
  double* td = tm + (size_t)kpad * QC;
  const MirrorSlot<QC> sl(g, qr, rc0);
  // b = (f / N2[k]) * RHS[k] on the interior levels (src/operators.jl:439)
  mirror_stage<QC>(g, qr, rc0, sl, rhsw, M, D, xs, tm, td, kpad,
                   [&](int k) { return (k >= 1 && k <= nz - 2) ? fN2[k] : 0.0; });
  __syncthreads();

  const int lane = threadIdx.x;
  if (lane < 2 * NC) {
    const int cs = lane >> 1, part = lane & 1;
    const int rc = rc0 + cs % QC;
    const QuadCol q = quad_col(g, qr, rc, cs / QC);
    if (q.active && nz > 2) {
      double* b0 = reinterpret_cast<double*>(xs + cs) + part;  // level k at b0[k * 2NC]
      const int st = 2 * NC;
      if (pivoted[rc]) {
        const double kh2 = g.kx[q.x] * g.kx[q.x] + g.ky[q.y] * g.ky[q.y];
        if (part == 0) gtsv_general(b0 + st, st, nz - 2, kh2, cz, sd, sdu, sfill, g.nx * g.ny, q.y * g.nx + q.x);
      } else {
        const double* tM = tm + cs % QC;
        const double* tD = td + cs % QC;
        mirror_sweeps(
            b0, st, kpad, [&](int k, double b) { return make_double2(tM[k * QC], b); },
            [&](int k, double bp) {
              const double r = tD[k * QC];
              return make_double2(cu[k] * r, bp * r);
            });
      }
    }
  }
  __syncthreads();

  if (!sl.q.active) return;
  for (int k = sl.kf; k < nz; k += MirrorSlot<QC>::kstep) wk[(size_t)k * plane + sl.off] = xs[k * NC + sl.c];
}

template <int QC>
void launch_qc(const Geo& g, hipStream_t st, const cplx<double>* rhsw, cplx<double>* wk, const double* M,
               const double* D, const unsigned char* pivoted, const double* cz, const double* cu, const double* fN2,
               double* sd, double* sdu, double* sfill) {
  const Quarter qr = quarter(g);
  const int blocks = (qr.size() + QC - 1) / QC;
  hipLaunchKernelGGL((k_omega<QC>), dim3(blocks, g.members), dim3(kMirrorThreads),
                     mirror_lds_bytes(mirror_pad(g.nzc), QC), st, g, qr, rhsw, wk, M, D, pivoted, cz, cu, fN2, sd, sdu,
                     sfill);
}

template <int QC>
bool set_lds_attr() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(k_omega<QC>), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kMirrorLdsBytes) == hipSuccess;
}

}  // namespace

void launch_omega_factor(const Geo& g, hipStream_t st, const double* cz, double* M, double* D,
                         unsigned char* pivoted, int* err) {
  const Quarter qr = quarter(g);
  const int bs = 64;
  hipLaunchKernelGGL(k_omega_factor, dim3((qr.size() + bs - 1) / bs, g.members), dim3(bs), 0, st, g, qr, cz, M, D,
                     pivoted, err);
}

bool omega_setup() { return set_lds_attr<8>() && set_lds_attr<4>() && set_lds_attr<2>() && set_lds_attr<1>(); }

void launch_omega(const Geo& g, hipStream_t st, const cplx<double>* rhsw, cplx<double>* wk, const double* M,
                  const double* D, const unsigned char* pivoted, const double* cz, const double* cu,
                  const double* fN2, double* sd, double* sdu, double* sfill) {
  switch (mirror_qc(g.nzc)) {
    case 8: launch_qc<8>(g, st, rhsw, wk, M, D, pivoted, cz, cu, fN2, sd, sdu, sfill); return;
    case 4: launch_qc<4>(g, st, rhsw, wk, M, D, pivoted, cz, cu, fN2, sd, sdu, sfill); return;
    case 2: launch_qc<2>(g, st, rhsw, wk, M, D, pivoted, cz, cu, fN2, sd, sdu, sfill); return;
    case 1: launch_qc<1>(g, st, rhsw, wk, M, D, pivoted, cz, cu, fN2, sd, sdu, sfill); return;
    default: return;  // rejected by column_depth_supported at model creation
  }
}

}  // namespace qg
