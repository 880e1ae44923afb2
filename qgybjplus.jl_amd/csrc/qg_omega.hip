// Interior solve of the QG omega equation (compute_vertical_velocity!,
// src/operators.jl:415-470), batched over columns.
//
// The reference calls LAPACK dgtsv (Gaussian elimination with partial
// pivoting) once for the real and once for the imaginary part of each column.
// The elimination depends only on (kh2, f^2/N^2, dz), so k_omega_factor runs
// it once per model for the quarter columns and records, level-indexed, the
// multiplier applied to each row and the reciprocal of its eliminated pivot.
// Per step the partitioned column kernel (qg_colscan.hpp) applies the
// recorded elimination to the right-hand sides, b_k -= m_k b_{k-1}, and back
// substitutes x_k = (b_k - u_k x_{k+1}) / d_k as
// fma(-(u_k / d_k), x_{k+1}, b_k / d_k), with the quotients formed from the
// stored reciprocals.  A column whose elimination interchanged rows is marked
// and solved by k_omega_pivoted, a restatement of dgtsv.
#include "qg_columns.hpp"
#include "qg_colscan.hpp"

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
//   code 2; err[3] is set when any column interchanged rows).
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
  if (piv) atomicOr(err + 3, 1);
}

// dgtsv with row interchanges for the columns k_omega_factor marked, one
// thread per grid column (both parts).  b = (f/N2) * RHS on the interior
// levels is formed in wk and solved in place; the eliminated rows are kept in
// the scratch arrays sd, sdu, sfill (per grid column and level).
template <typename T>
__global__ void k_omega_pivoted(Geo g, Quarter qr, const cplx<T>* __restrict__ rhsw, cplx<T>* __restrict__ wk,
                                const unsigned char* __restrict__ pivoted, const double* __restrict__ cz,
                                const double* __restrict__ fN2, double* __restrict__ sd, double* __restrict__ sdu,
                                double* __restrict__ sfill) {
  const int rplane = qr.size();
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= 4 * rplane) return;
  const int rc = id >> 2, mir = id & 3;
  const int nz = g.nzc;
  const size_t plane = (size_t)g.nx * g.ny;
  const size_t mem = blockIdx.y;
  if (!pivoted[mem * rplane + rc] || nz <= 2) return;
  const QuadCol q = quad_col(g, qr, rc, mir);
  if (!q.active) return;
  rhsw += mem * nz * plane;
  wk += mem * nz * plane;
  cz += mem * nz;
  fN2 += mem * nz;
  sd += mem * nz * plane;
  sdu += mem * nz * plane;
  sfill += mem * nz * plane;
  const size_t col = (size_t)q.y * g.nx + q.x;
  const double kh2 = g.kx[q.x] * g.kx[q.x] + g.ky[q.y] * g.ky[q.y];
  const int n = nz - 2;
  auto at = [&](int iz) { return (size_t)iz * plane + col; };
  // interior row iz is level iz + 1; wk holds the eliminated right-hand
  // sides between the sweeps (in the field's storage precision)
  auto ld = [&](int iz) { return cvt<double>(wk[(size_t)(iz + 1) * plane + col]); };
  auto st = [&](int iz, cplx<double> v) { wk[(size_t)(iz + 1) * plane + col] = cvt<T>(v); };
  for (int iz = 0; iz < n; ++iz) {
    const cplx<double> r = cvt<double>(rhsw[(size_t)(iz + 1) * plane + col]);
    const double s = fN2[iz + 1];
    st(iz, mk(s * r.re, s * r.im));
  }
  wk[col] = cvt<T>(mk(0.0, 0.0));
  wk[(size_t)(nz - 1) * plane + col] = cvt<T>(mk(0.0, 0.0));
  double d = -2 * cz[1] - kh2;
  double du = (n > 1) ? cz[1] : 0.0;
  double br = ld(0).re, bi = ld(0).im;
  for (int i = 0; i < n - 1; ++i) {
    const double dl = cz[i + 2];
    double dn = -2 * cz[i + 2] - kh2;
    double dun = (i + 1 < n - 1) ? cz[i + 2] : 0.0;
    double bnr = ld(i + 1).re, bni = ld(i + 1).im;
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
    st(i, mk(br, bi));
    d = dn;
    du = dun;
    br = bnr;
    bi = bni;
  }
  double xr = br / d, xi = bi / d;
  st(n - 1, mk(xr, xi));
  if (n > 1) {
    const int i = n - 2;
    const double x0r = (ld(i).re - sdu[at(i)] * xr) / sd[at(i)];
    const double x0i = (ld(i).im - sdu[at(i)] * xi) / sd[at(i)];
    st(i, mk(x0r, x0i));
    double x2r = xr, x2i = xi;
    xr = x0r;
    xi = x0i;
    for (int ii = n - 3; ii >= 0; --ii) {
      const double dd = sd[at(ii)], uu = sdu[at(ii)], ff = sfill[at(ii)];
      const cplx<double> b = ld(ii);
      const double vr = (b.re - uu * xr - ff * x2r) / dd;
      const double vi = (b.im - uu * xi - ff * x2i) / dd;
      st(ii, mk(vr, vi));
      x2r = xr;
      x2i = xi;
      xr = vr;
      xi = vi;
    }
  }
}

}  // namespace

void launch_omega_factor(const Geo& g, hipStream_t st, const double* cz, double* M, double* D,
                         unsigned char* pivoted, int* err) {
  const Quarter qr = quarter(g);
  const int bs = 64;
  hipLaunchKernelGGL(k_omega_factor, dim3((qr.size() + bs - 1) / bs, g.members), dim3(bs), 0, st, g, qr, cz, M, D,
                     pivoted, err);
}

bool omega_setup() { return colscan_setup(); }

void launch_omega(const Geo& g, hipStream_t st, const void* rhsw, void* wk, bool f32, const double* M,
                  const double* D, const unsigned char* pivoted, bool any_pivoted, const double* cz,
                  const double* cu, const double* fN2, double* sd, double* sdu, double* sfill, int shape) {
  if (shape < 0) shape = scan_default_shape(g.nzc);
  ScanArgs a{};
  a.rhs = rhsw;
  a.out = wk;
  a.t0 = M;
  a.t1 = D;
  a.lv0 = cu;
  a.lv0_stride = thomas_shifted_a_size(g.nzc);
  a.lv1 = fN2;
  a.lv1_stride = g.nzc;
  a.pivoted = pivoted;
  launch_colscan<SCAN_OMEGA>(g, st, shape, a, f32);
  if (any_pivoted) {
    const Quarter qr = quarter(g);
    const int bs = 64;
    const dim3 grid((4 * qr.size() + bs - 1) / bs, g.members);
    if (f32)
      hipLaunchKernelGGL(k_omega_pivoted<float>, grid, dim3(bs), 0, st, g, qr, static_cast<const cplx<float>*>(rhsw),
                         static_cast<cplx<float>*>(wk), pivoted, cz, fN2, sd, sdu, sfill);
    else
      hipLaunchKernelGGL(k_omega_pivoted<double>, grid, dim3(bs), 0, st, g, qr,
                         static_cast<const cplx<double>*>(rhsw), static_cast<cplx<double>*>(wk), pivoted, cz, fN2,
                         sd, sdu, sfill);
  }
}

}  // namespace qg
