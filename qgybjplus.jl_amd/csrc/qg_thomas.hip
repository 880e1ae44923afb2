// Batched Thomas solves of the q->psi and B->A column operators
// (thomas_solve!, src/elliptic.jl:1168-1202, as called by invert_q_to_psi!,
// src/elliptic.jl:189-314, and invert_B_to_A!, src/elliptic.jl:836-990).
//
// The elimination of a column depends only on (kh2, a, dz), so
// k_thomas_factor runs it once per model and records per level the pivot
// denom_k, its reciprocal and c'_k for the quarter columns (the reference
// recomputes them on every call).  A solve applies them with the partitioned
// column kernel (qg_colscan.hpp): the forward sweep
// x'_k = (dz2 b_k - a_{k-1} x'_{k-1}) / denom_k is evaluated as
// fma(-(a_{k-1} inv_k), x'_{k-1}, (dz2 b_k) inv_k) and the back substitution
// as fma(-c'_k, x_{k+1}, x'_k); the results differ from the reference's
// division by a few ulp.
#include "qg_columns.hpp"
#include "qg_colscan.hpp"

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
  if (nz == 1) {
    // one level: psi = -q / kh2, A = -4 B / kh2 (0 where kh = 0); the solve
    // multiplies by this reciprocal (dz = 1 for one level)
    inv[at(0)] = kh2 == 0.0 ? 0.0 : (kdiv == 1 ? -1.0 : -4.0) / kh2;
    cp[at(0)] = 0.0;
    return;
  }
  if (kh2 == 0.0 && kdiv == 1) {  // psi = 0 on the horizontal mean, written by the solve
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

}  // namespace

size_t thomas_table_size(const Geo& g) { return (size_t)quarter(g).size() * g.nz; }

int thomas_shifted_a_size(int nzc) { return (nzc + 15) / 16 * 16; }

bool column_depth_supported(int nz) { return nz >= 1 && nz <= 1024 && scan_default_shape(nz) >= 0; }

void launch_thomas_factor(const Geo& g, hipStream_t st, const double* a, int kdiv, double* inv, double* cp,
                          int* err) {
  const Quarter qr = quarter(g);
  const int bs = 64;
  hipLaunchKernelGGL(k_thomas_factor, dim3((qr.size() + bs - 1) / bs, g.members), dim3(bs), 0, st, g, qr, a, kdiv,
                     inv, cp, err);
}

bool thomas_setup() { return colscan_setup(); }

void launch_thomas(const Geo& g, hipStream_t st, const void* rhs, void* out, void* Cout, bool f32,
                   const double* inv, const double* cp, const double* al, int kdiv, unsigned long long* warn,
                   int shape) {
  if (shape < 0) shape = scan_default_shape(g.nzc);
  ScanArgs a{};
  a.rhs = rhs;
  a.out = out;
  a.Cout = Cout;
  a.t0 = inv;
  a.t1 = cp;
  a.lv0 = al;
  a.lv0_stride = thomas_shifted_a_size(g.nzc);
  if (kdiv == 1) {
    a.warn = warn;
    launch_colscan<SCAN_PSI>(g, st, shape, a, f32);
  } else if (Cout) {
    launch_colscan<SCAN_AC>(g, st, shape, a, f32);
  } else {
    launch_colscan<SCAN_A>(g, st, shape, a, f32);
  }
}

void launch_thomas_pair(const Geo& g, hipStream_t st, bool f32, const double* al, int shape, const void* q,
                        void* psi, const double* inv_q, const double* cp_q, unsigned long long* warn,
                        const void* B, void* A, void* Cout, const double* inv_b, const double* cp_b) {
  if (shape < 0) shape = scan_pair_default_shape(g.nzc);
  ScanArgs a0{}, a1{};
  a0.rhs = q;
  a0.out = psi;
  a0.t0 = inv_q;
  a0.t1 = cp_q;
  a0.lv0 = al;
  a0.lv0_stride = thomas_shifted_a_size(g.nzc);
  a0.warn = warn;
  a1.rhs = B;
  a1.out = A;
  a1.Cout = Cout;
  a1.t0 = inv_b;
  a1.t1 = cp_b;
  a1.lv0 = al;
  a1.lv0_stride = a0.lv0_stride;
  if (Cout) launch_colscan2<SCAN_PSI, SCAN_AC>(g, st, shape, a0, a1, f32);
  else launch_colscan2<SCAN_PSI, SCAN_A>(g, st, shape, a0, a1, f32);
}

}  // namespace qg
