// Batched Thomas solves of the q->psi and B->A column operators
// (thomas_solve!, src/elliptic.jl:1168-1202, as called by invert_q_to_psi!,
// src/elliptic.jl:189-314, and invert_B_to_A!, src/elliptic.jl:836-990).
//
// The elimination of a column depends only on (kh2, a, dz), so
// k_thomas_factor runs it once per model and records per level the pivot
// denom_k, its reciprocal and c'_k for the quarter columns (the reference
// recomputes them on every call).  A solve is a mirror-tile kernel: the
// forward sweep x'_k = (dz2 b_k - a_{k-1} x'_{k-1}) / denom_k is evaluated as
// fma(-(a_{k-1} inv_k), x'_{k-1}, (dz2 b_k) inv_k) and the back substitution
// as fma(-c'_k, x_{k+1}, x'_k), one FMA per level on the serial chain; the
// results differ from the reference's division by a few ulp.
#include "qg_columns.hpp"
#include "qg_mirror.hpp"

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

// Tables [k][rc]: inv = 1/denom, cp = c'.  Pivots below sqrt(eps) are
// recorded in err (code 1) where thomas_solve! raises.
__global__ void k_thomas_factor(Geo g, Quarter qr, const double* __restrict__ a, int kdiv,
                                double* __restrict__ inv, double* __restrict__ cp, int* __restrict__ err) {
  const int rplane = qr.size();
  const int rc = blockIdx.x * blockDim.x + threadIdx.x;
  if (rc >= rplane) return;
  const int x = rc % qr.Wr, y = rc / qr.Wr;
  const int nz = g.nzc, m = blockIdx.y;  // member m: its own a(z) and tables
  a += (size_t)m * nz;
  inv += (size_t)m * nz * rplane;
  cp += (size_t)m * nz * rplane;
  const double kh2 = g.kx[x] * g.kx[x] + g.ky[y] * g.ky[y];
  const double dz2 = g.dz * g.dz;
  auto at = [&](int k) { return (size_t)k * rplane + rc; };
  if (nz == 1 || (kh2 == 0.0 && kdiv == 1)) {
    for (int k = 0; k < nz; ++k) {
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
  inv[at(0)] = gauge ? 0.0 : 1.0 / d0;  // gauge row: x'_0 = 0 * rhs
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
    inv[at(i)] = 1.0 / denom;
    cp[at(i)] = c;
  }
}

// KIND 0: q -> psi; 1: B -> A; 2: B -> A and C = dA/dz.  blockIdx.y is the
// member.  al[k] = a[k-1] (al[0] = 0), zero-padded to mirror_pad(nzc)
// entries per member.
template <int KIND, int QC>
__global__ void __launch_bounds__(kMirrorThreads)
k_thomas(Geo g, Quarter qr, const cplx<double>* __restrict__ rhs, cplx<double>* __restrict__ out,
         cplx<double>* __restrict__ Cout, const double* __restrict__ inv, const double* __restrict__ cp,
         const double* __restrict__ al, unsigned long long* __restrict__ warn) {
  constexpr int NC = 4 * QC;
  extern __shared__ cplx<double> xs[];  // [kpad][NC], then ti[kpad][QC], tc[kpad][QC]
  const int nz = g.nzc, kpad = mirror_pad(nz);
  const int rc0 = blockIdx.x * QC;
  const size_t plane = (size_t)g.nx * g.ny;
  {
    const size_t m = blockIdx.y;
    rhs += m * nz * plane;
    out += m * nz * plane;
    if (KIND == 2) Cout += m * nz * plane;
    inv += m * nz * qr.size();
    cp += m * nz * qr.size();
    al += m * kpad;
  }
  double* ti = reinterpret_cast<double*>(xs + (size_t)kpad * NC);
  double* tc = ti + (size_t)kpad * QC;
  const MirrorSlot<QC> sl(g, qr, rc0);
  mirror_stage<QC>(g, qr, rc0, sl, rhs, inv, cp, xs, ti, tc, kpad, [](int) { return 1.0; });
  __syncthreads();

  const int lane = threadIdx.x;
  if (lane < 2 * NC) {
    const int cs = lane >> 1, part = lane & 1;
    const QuadCol q = quad_col(g, qr, rc0 + cs % QC, cs / QC);
    if (q.active) {
      double* b0 = reinterpret_cast<double*>(xs + cs) + part;  // level k at b0[k * 2NC]
      const int st = 2 * NC;
      const double* tI = ti + cs % QC;
      const double* tC = tc + cs % QC;
      const double kh2 = g.kx[q.x] * g.kx[q.x] + g.ky[q.y] * g.ky[q.y];
      const double dz2 = g.dz * g.dz;
      if (nz == 1) {
        double& v = b0[0];
        if (kh2 == 0.0) v = 0.0;
        else if (KIND == 0) v = -v / kh2;
        else v = (-4 * v) / kh2;
      } else if (kh2 == 0.0 && KIND == 0) {
        // psi = 0 on the horizontal mean (src/elliptic.jl:257-270); warn on a
        // significant mean of q.
        if (part == 0 && warn) {
          double m = 0.0;
          for (int k = 0; k < nz; ++k) {
            const double av = hypot(b0[k * st], b0[k * st + 1]);
            m = av > m ? av : m;
          }
          if (m > 1e-10) atomicMax(warn, (unsigned long long)__double_as_longlong(m));
        }
      } else {
        // For the B -> A mean column the stored reciprocal of the gauge row
        // A[1] = 0 is 0, which gives x'_0 = 0.
        mirror_sweeps(
            b0, st, kpad,
            [&](int k, double b) {
              const double y = tI[k * QC];
              return make_double2(al[k] * y, (dz2 * b) * y);
            },
            [&](int k, double xp) { return make_double2(tC[k * QC], xp); });
        if (kh2 == 0.0) {  // remove the vertical mean (src/elliptic.jl:930-945)
          double mean = 0.0;
          for (int k = 0; k < nz; ++k) mean = mean + b0[k * st];
          mean = mean / nz;
          for (int k = 0; k < nz; ++k) b0[k * st] = b0[k * st] - mean;
        }
      }
    }
  }
  __syncthreads();

  // write back (psi = 0 on the mean column), and C = dA/dz
  if (!sl.q.active) return;
  const bool zero_col = KIND == 0 && nz > 1 && sl.q.x == 0 && sl.q.y == 0;
  for (int k = sl.kf; k < nz; k += MirrorSlot<QC>::kstep) {
    const size_t at = (size_t)k * plane + sl.off;
    const cplx<double> r = zero_col ? mk(0.0, 0.0) : xs[k * NC + sl.c];
    out[at] = r;
    if (KIND == 2) {
      const cplx<double> up = xs[min(k + 1, kpad - 1) * NC + sl.c];
      Cout[at] = k < nz - 1 ? mk((up.re - r.re) / g.dz, (up.im - r.im) / g.dz) : mk(0.0, 0.0);
    }
  }
}

template <int KIND, int QC>
void launch_k(const Geo& g, hipStream_t st, const cplx<double>* rhs, cplx<double>* out, cplx<double>* Cout,
              const double* inv, const double* cp, const double* al, unsigned long long* warn) {
  const Quarter qr = quarter(g);
  const int blocks = (qr.size() + QC - 1) / QC;
  hipLaunchKernelGGL((k_thomas<KIND, QC>), dim3(blocks, g.members), dim3(kMirrorThreads),
                     mirror_lds_bytes(mirror_pad(g.nzc), QC), st, g, qr, rhs, out, Cout, inv, cp, al, warn);
}

template <int QC>
void launch_qc(const Geo& g, hipStream_t st, const cplx<double>* rhs, cplx<double>* out, cplx<double>* Cout,
               const double* inv, const double* cp, const double* al, int kdiv, unsigned long long* warn) {
  if (kdiv == 1) launch_k<0, QC>(g, st, rhs, out, nullptr, inv, cp, al, warn);
  else if (Cout) launch_k<2, QC>(g, st, rhs, out, Cout, inv, cp, al, warn);
  else launch_k<1, QC>(g, st, rhs, out, nullptr, inv, cp, al, warn);
}

template <int QC>
bool set_lds_attr() {
  const void* fns[] = {reinterpret_cast<const void*>(k_thomas<0, QC>), reinterpret_cast<const void*>(k_thomas<1, QC>),
                       reinterpret_cast<const void*>(k_thomas<2, QC>)};
  for (const void* f : fns)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMirrorLdsBytes) != hipSuccess)
      return false;
  return true;
}

}  // namespace

size_t thomas_table_size(const Geo& g) { return (size_t)quarter(g).size() * g.nz; }

int thomas_shifted_a_size(int nzc) { return mirror_pad(nzc); }

bool column_depth_supported(int nz) { return nz >= 1 && mirror_qc(nz) > 0; }

void launch_thomas_factor(const Geo& g, hipStream_t st, const double* a, int kdiv, double* inv, double* cp,
                          int* err) {
  const Quarter qr = quarter(g);
  const int bs = 64;
  hipLaunchKernelGGL(k_thomas_factor, dim3((qr.size() + bs - 1) / bs, g.members), dim3(bs), 0, st, g, qr, a, kdiv,
                     inv, cp, err);
}

bool thomas_setup() { return set_lds_attr<8>() && set_lds_attr<4>() && set_lds_attr<2>() && set_lds_attr<1>(); }

void launch_thomas(const Geo& g, hipStream_t st, const cplx<double>* rhs, cplx<double>* out, cplx<double>* Cout,
                   const double* inv, const double* cp, const double* al, int kdiv, unsigned long long* warn) {
  switch (mirror_qc(g.nzc)) {
    case 8: launch_qc<8>(g, st, rhs, out, Cout, inv, cp, al, kdiv, warn); return;
    case 4: launch_qc<4>(g, st, rhs, out, Cout, inv, cp, al, kdiv, warn); return;
    case 2: launch_qc<2>(g, st, rhs, out, Cout, inv, cp, al, kdiv, warn); return;
    case 1: launch_qc<1>(g, st, rhs, out, Cout, inv, cp, al, kdiv, warn); return;
    default: return;  // rejected by column_depth_supported at model creation
  }
}

}  // namespace qg
