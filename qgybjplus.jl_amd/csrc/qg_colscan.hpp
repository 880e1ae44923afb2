// Partitioned column solves for the per-(kx, ky) vertical operators.
//
// Every column operator of the step is a tridiagonal elimination whose
// factors depend on (kx, ky) only through kh2, so the factors are recorded
// once per model for the quarter of wavenumbers |ix| <= nx/2, |iy| <= ny/2
// ("quarter columns", index rc = ry * (nx/2 + 1) + rx), level-indexed
// [k][rc].  Applying a recorded elimination to a right-hand side b is two
// first-order linear recurrences:
//   forward   x'_k = -G_k x'_{k-1} + H_k(b_k)     (x'_{-1} = 0)
//   backward  x_k  = -P_k x_{k+1}  + Q_k(x'_k)    (x_{kpad} = 0)
//
// A tile is QC consecutive quarter columns and their mirror images (+-kx,
// +-ky): 4*QC grid columns.  A block splits the depth into P partitions of M
// levels; each thread owns one grid column of one partition and keeps its M
// complex values in registers.  The real and imaginary parts are solved
// separately, as the reference does, with the factors read once for both.
// A recurrence is evaluated in three steps:
//   1. every partition runs it locally from a zero carry, and the product of
//      its multipliers (-G or -P) over the partition;
//   2. the partitions' end values and products go through LDS, and each
//      thread folds those of the partitions before (after) its own into the
//      exact carry entering its partition;
//   3. the carry times the running product of multipliers is added to each
//      level.
// This is the same elimination with the same recorded factors as one serial
// sweep, evaluated with a different association of its sums (a few ulp).
// All threads of a block do sweep work, the column state lives in registers
// (no column tile in LDS), and the loads of a thread are independent of its
// sweeps, so many columns per CU are in flight at once.
#pragma once

#include <type_traits>

#include "qg_common.hpp"

namespace qg {

struct Quarter {
  int Wr, Hr;  // nx/2 + 1, ny/2 + 1
  int kd2;     // pruned (Geo::kd2 >= 0): only the quarter columns rx^2 + ry^2 <= kd2 are solved
  __host__ __device__ int size() const { return Wr * Hr; }
};

inline Quarter quarter(const Geo& g) { return Quarter{g.nx / 2 + 1, g.ny / 2 + 1, g.kd2}; }

struct QuadCol {
  int x, y;
  bool active;
};

// Grid column of mirror image m (bit 0: -kx, bit 1: -ky) of quarter column
// rc; inactive past the quarter, outside the disk of a pruned solve, off the
// quarter row `row` of its tile (row >= 0), or when it duplicates a lower
// image.
__device__ inline QuadCol quad_col(const Geo& g, const Quarter& qr, int rc, int m, int row = -1) {
  QuadCol q{0, 0, false};
  if (rc >= qr.size()) return q;
  const int rx = rc % qr.Wr, ry = rc / qr.Wr;
  if ((row >= 0 && ry != row) || (qr.kd2 >= 0 && rx * rx + ry * ry > qr.kd2)) return q;
  q.x = (m & 1) ? (g.nx - rx) % g.nx : rx;
  q.y = (m & 2) ? (g.ny - ry) % g.ny : ry;
  q.active = !(((m & 1) && q.x == rx) || ((m & 2) && q.y == ry));
  return q;
}

// Column operators applied by k_colscan.
enum : int { SCAN_PSI = 0, SCAN_A = 1, SCAN_AC = 2, SCAN_OMEGA = 3 };

// Launch shape: M levels per thread, QC quarter columns per tile, P
// partitions (P * M >= nz).  A block has P * 4 * QC threads.  LDS: the two
// tables of the tile and the two level arrays (kpad levels each) and the
// forward and backward exchange rows (re, im, product per thread).
struct ScanShape {
  int M, QC, P;
  int threads() const { return P * 4 * QC; }
  int kpad() const { return P * M; }
  size_t lds_bytes() const {
    return sizeof(double) * ((size_t)2 * kpad() * QC + 2 * (size_t)kpad() + 6 * (size_t)P * 4 * QC);
  }
};

// Blocks of at most 512 threads: two waves per SIMD, so a thread may keep up
// to 256 VGPRs (2 M doubles of column values plus its staged table entries).
constexpr int kScanMaxThreads = 512;
// Word 3 of a raw buffer resource on the gfx9 family (gfx950): dword
// access, no format conversion.
constexpr int kRawBufferWord3 = 0x00020000;
constexpr size_t kScanMaxLds = 96 * 1024;

// The compiled (M, QC) pairs.
constexpr int kScanShapes[][2] = {{8, 8}, {16, 8}, {32, 8}, {16, 4}, {32, 4}, {64, 4}, {16, 16}, {32, 16}};
constexpr int kScanShapeCount = sizeof(kScanShapes) / sizeof(kScanShapes[0]);

inline bool scan_shape_fits(const ScanShape& s) {
  return s.P >= 1 && s.threads() <= kScanMaxThreads && s.lds_bytes() <= kScanMaxLds;
}

inline ScanShape scan_shape_of(int idx, int nz) {
  ScanShape s{kScanShapes[idx][0], kScanShapes[idx][1], 0};
  s.P = (nz + s.M - 1) / s.M;
  return s;
}

// Default shape for depth nz: among the 8-quarter-column tiles (32 columns
// per partition), M = 8 for nz <= 8, else the smallest of
// M = 16, 32 that fits (at most 16 partitions); else the deepest
// 4-quarter-column shape that fits.  Short partitions keep fewer registers
// per thread and more columns in flight.
inline int scan_default_shape(int nz) {
  for (int i = 0; i < kScanShapeCount; ++i) {
    ScanShape s = scan_shape_of(i, nz);
    if (s.QC == 8 && scan_shape_fits(s) && (s.M >= 16 || nz <= 8)) return i;
  }
  for (int i = kScanShapeCount - 1; i >= 0; --i) {
    ScanShape s = scan_shape_of(i, nz);
    if (s.QC == 4 && scan_shape_fits(s)) return i;
  }
  return -1;
}

// Field pointers are cplx<T>: T = double, or float for fields stored in FP32
// (every recurrence runs in FP64 either way).
struct ScanArgs {
  const void* rhs;          // member-0 level-0 of the right-hand side
  void* out;
  void* Cout;               // SCAN_AC: C = dA/dz
  const double* t0;         // Thomas: 1/pivot; omega: row multiplier
  const double* t1;         // Thomas: c'; omega: 1/pivot
  const double* lv0;        // Thomas: a_{k-1}; omega: upper diagonal by level
  const double* lv1;        // omega: f/N2 by level
  int lv0_stride;           // per-member entries of lv0
  int lv1_stride;           // per-member entries of lv1
  const unsigned char* pivoted;  // omega: per quarter column, solved elsewhere
  unsigned long long* warn;      // SCAN_PSI: max |q| of the kh = 0 column
};

// Keeps the scheduler from hoisting the LDS reads of a whole unrolled sweep
// ahead of its first levels (which would hold every level's factors in
// registers at once): sweeps are scheduled in chunks of 8 levels.
__device__ inline void sweep_chunk_fence(int u) {
  if ((u & 7) == 7) __builtin_amdgcn_sched_barrier(0);
}

template <int M, int QC>
struct ScanBounds {
  static constexpr int TP = 4 * QC;
  static constexpr int MAXP = kScanMaxThreads / TP < 16 ? kScanMaxThreads / TP : 16;
  static constexpr int MAXT = MAXP * TP;
};

// Loads and stores of one column's complex value (both parts) of storage
// type T as FP64.
template <typename T>
__device__ inline void scan_load2(__amdgpu_buffer_rsrc_t r, unsigned off, int soff, double& re, double& im) {
  if constexpr (sizeof(T) == 8) {
    const double2 v = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0));
    re = v.x;
    im = v.y;
  } else {
    const float2 v = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
    re = (double)v.x;
    im = (double)v.y;
  }
}
template <typename T>
__device__ inline void scan_store2(double re, double im, __amdgpu_buffer_rsrc_t r, unsigned off, int soff) {
  if constexpr (sizeof(T) == 8) {
    using V4 = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, make_double2(re, im)), r, off, soff, 0);
  } else {
    using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, make_float2((float)re, (float)im)), r, off, soff, 0);
  }
}

// Tiles of a pruned solve: QC consecutive quarter columns of one quarter row
// ry <= kmax (tiles per row: scan_row_tiles), so that the disk's columns of a
// row are covered by its first tiles; tiles wholly outside the disk exit.
__host__ __device__ inline int scan_row_tiles(const Geo& g, int QC) { return (g.Kx + QC - 1) / QC; }
inline int scan_tiles(const Geo& g, int QC) {
  return g.pruned() ? g.Kx * scan_row_tiles(g, QC) : (quarter(g).size() + QC - 1) / QC;
}

// The solve of operator OP by one block: blk.x is the tile of QC quarter
// columns, blk.y the member; lds is the block's dynamic LDS (lds_bytes()).
template <int OP, int M, int QC, typename T>
__device__ inline void colscan_block(const Geo& g, const Quarter& qr, const ScanArgs& a, const Blk blk,
                                     double* lds) {
  constexpr int TP = 4 * QC;
  constexpr int R = (M * QC + TP - 1) / TP;  // table entries staged per thread
  constexpr int LR = (M + TP - 1) / TP;      // level entries staged per thread
  const int nz = g.nzc;
  const int P = blockDim.x / TP;
  const int kpad = P * M;
  const int w = threadIdx.x / TP, t = threadIdx.x % TP;
  const int j = t % QC, mir = t / QC;
  const int rplane = qr.size();
  int rc0 = blk.x * QC, row = -1;
  if (g.pruned()) {  // uniform
    const int tpr = scan_row_tiles(g, QC), ry = blk.x / tpr, rx0 = (blk.x % tpr) * QC;
    if (rx0 * rx0 + ry * ry > g.kd2) return;  // the whole tile is outside the disk
    rc0 = ry * qr.Wr + rx0;
    row = ry;
  }
  const size_t plane = (size_t)g.nx * g.ny;
  const size_t mem = blk.y;
  const T* rhs = static_cast<const T*>(a.rhs) + 2 * mem * nz * plane;
  T* out = static_cast<T*>(a.out) + 2 * mem * nz * plane;
  const double* t0 = a.t0 + mem * nz * rplane;
  const double* t1 = a.t1 + mem * nz * rplane;
  const double* lv0 = a.lv0 + mem * a.lv0_stride;
  const double* lv1 = (OP == SCAN_OMEGA) ? a.lv1 + mem * a.lv1_stride : nullptr;
  double* T0 = lds;               // [kpad][QC]
  double* T1 = T0 + kpad * QC;    // [kpad][QC]
  double* L0 = T1 + kpad * QC;    // [kpad]
  double* L1 = L0 + kpad;         // [kpad]
  double* XF = L1 + kpad;         // [3][P][TP]: forward end values (re, im), products
  double* XB = XF + 3 * P * TP;   // [3][P][TP]: backward end values (re, im), products

  const QuadCol q = quad_col(g, qr, rc0 + j, mir, row);
  // This thread's column within a level, in bytes.  With 64 columns per
  // partition a partition is one wave, so its first level is wave-uniform
  // and every level's address is a scalar base plus this offset.
  constexpr bool UNI = TP % 64 == 0;
  const unsigned boff = (unsigned)sizeof(T) * 2u * (q.active ? (unsigned)(q.y * g.nx + q.x) : 0u);
  int k0 = w * M;
  if constexpr (UNI) k0 = __builtin_amdgcn_readfirstlane(k0);
  const size_t lvlb = plane * 2 * sizeof(T);  // bytes per level
  // Wave-uniform partitions address their levels through a buffer resource
  // over the partition's levels inside the field: one 32-bit lane offset and
  // a scalar offset per level; levels past nz read as zero and are not
  // stored.  Otherwise each level's address is formed per lane, clamped.
  const int nlev = min(M, nz - k0);
  auto rsrc = [&](const void* field) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(static_cast<const char*>(field)) + (size_t)k0 * lvlb,
                                             0, (int)(nlev * lvlb), kRawBufferWord3);
  };
  auto at_level = [&](auto* base, int k) {
    using U = std::remove_pointer_t<decltype(base)>;
    using Byte = std::conditional_t<std::is_const_v<U>, const char, char>;
    return reinterpret_cast<U*>(reinterpret_cast<Byte*>(base + (size_t)k * (2 * plane)) + boff);
  };

  // Loads: the tables of the tile first, then this thread's right-hand side
  // levels; addresses are clamped so that no load is conditional.
  double v0[R], v1[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = threadIdx.x + r * blockDim.x;
    const int k = min(e / QC, nz - 1), jj = min(rc0 + e % QC, rplane - 1);
    v0[r] = t0[(size_t)k * rplane + jj];
    v1[r] = t1[(size_t)k * rplane + jj];
  }
  double l0[LR], l1[LR];
#pragma unroll
  for (int r = 0; r < LR; ++r) {
    const int e = min((int)threadIdx.x + r * (int)blockDim.x, nz - 1);
    l0[r] = lv0[e];
    l1[r] = (OP == SCAN_OMEGA) ? lv1[e] : 0.0;
  }
  // the real and imaginary parts of each level (solved with the same
  // factors, each as the reference solves its parts)
  double yr[M], yi[M];
  if constexpr (UNI) {
    const auto rs = rsrc(rhs);
#pragma unroll
    for (int u = 0; u < M; ++u) scan_load2<T>(rs, boff, u * (int)lvlb, yr[u], yi[u]);
  } else {
#pragma unroll
    for (int u = 0; u < M; ++u) {
      const T* p = at_level(rhs, min(k0 + u, nz - 1));
      yr[u] = (double)p[0];
      yi[u] = (double)p[1];
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int e = threadIdx.x + r * blockDim.x;
    if (e < kpad * QC) {
      const bool ok = e / QC < nz && rc0 + e % QC < rplane;
      T0[e] = ok ? v0[r] : 0.0;
      T1[e] = ok ? v1[r] : 0.0;
    }
  }
#pragma unroll
  for (int r = 0; r < LR; ++r) {
    const int e = threadIdx.x + r * blockDim.x;
    if (e < kpad) {
      L0[e] = e < nz ? l0[r] : 0.0;
      if (OP == SCAN_OMEGA) L1[e] = (e >= 1 && e <= nz - 2) ? l1[r] : 0.0;  // f/N2 on interior rows
    }
  }
  __syncthreads();

  const double dz2 = g.dz * g.dz;
  auto G = [&](int k) -> double {
    if constexpr (OP == SCAN_OMEGA) return T0[k * QC + j];
    else return L0[k] * T0[k * QC + j];
  };
  auto H = [&](int k, double b) -> double {
    if constexpr (OP == SCAN_OMEGA) return L1[k] * b;
    else return (dz2 * b) * T0[k * QC + j];
  };
  auto Pm = [&](int k) -> double {
    if constexpr (OP == SCAN_OMEGA) return L0[k] * T1[k * QC + j];
    else return T1[k * QC + j];
  };
  auto Q = [&](int k, double xp) -> double {
    if constexpr (OP == SCAN_OMEGA) return xp * T1[k * QC + j];
    else return xp;
  };

  const bool mean_col = q.active && q.x == 0 && q.y == 0;

  // q -> psi: psi = 0 on the horizontal mean; warn on a significant mean of
  // q (src/elliptic.jl:257-270).  Block 0 holds that column.
  if (OP == SCAN_PSI && a.warn && nz > 1 && blk.x == 0) {
    double m = 0.0;
#pragma unroll
    for (int u = 0; u < M; ++u) {
      const double av = hypot(yr[u], yi[u]);
      if (k0 + u < nz) m = av > m ? av : m;
    }
    if (mean_col && m > 1e-10) atomicMax(a.warn, (unsigned long long)__double_as_longlong(m));
  }

  // One-level columns need no special case: their recorded factor is the
  // reciprocal of the operator's only entry (0 where kh = 0) and the omega
  // equation has no interior row.
  double Yr = 0.0, Yi = 0.0;  // final value of the level above this partition
  {
    // forward recurrence, local
    double xr = 0.0, xi = 0.0, phi = 1.0;
#pragma unroll
    for (int u = 0; u < M; ++u) {
      const int k = k0 + u;
      const double gk = G(k);
      xr = fma(-gk, xr, H(k, yr[u]));
      xi = fma(-gk, xi, H(k, yi[u]));
      yr[u] = xr;
      yi[u] = xi;
      phi = phi * -gk;
      sweep_chunk_fence(u);
    }
    XF[w * TP + t] = xr;
    XF[(P + w) * TP + t] = xi;
    XF[(2 * P + w) * TP + t] = phi;
    __syncthreads();
    if (w > 0) {
      double Xr = 0.0, Xi = 0.0;  // x' of the level below this partition
      for (int p = 0; p < w; ++p) {
        const double f = XF[(2 * P + p) * TP + t];
        Xr = fma(f, Xr, XF[p * TP + t]);
        Xi = fma(f, Xi, XF[(P + p) * TP + t]);
      }
      double ph = 1.0;
#pragma unroll
      for (int u = 0; u < M; ++u) {
        ph = ph * -G(k0 + u);
        yr[u] = fma(ph, Xr, yr[u]);
        yi[u] = fma(ph, Xi, yi[u]);
        sweep_chunk_fence(u);
      }
    }
    // backward recurrence, local
    double nr = 0.0, ni = 0.0, psi = 1.0;
#pragma unroll
    for (int u = M - 1; u >= 0; --u) {
      const int k = k0 + u;
      const double pk = Pm(k);
      nr = fma(-pk, nr, Q(k, yr[u]));
      ni = fma(-pk, ni, Q(k, yi[u]));
      yr[u] = nr;
      yi[u] = ni;
      psi = psi * -pk;
      sweep_chunk_fence(M - 1 - u);
    }
    XB[w * TP + t] = nr;
    XB[(P + w) * TP + t] = ni;
    XB[(2 * P + w) * TP + t] = psi;
    __syncthreads();
    if (w < P - 1) {
      for (int p = P - 1; p > w; --p) {
        const double f = XB[(2 * P + p) * TP + t];
        Yr = fma(f, Yr, XB[p * TP + t]);
        Yi = fma(f, Yi, XB[(P + p) * TP + t]);
      }
      double ps = 1.0;
#pragma unroll
      for (int u = M - 1; u >= 0; --u) {
        ps = ps * -Pm(k0 + u);
        yr[u] = fma(ps, Yr, yr[u]);
        yi[u] = fma(ps, Yi, yi[u]);
        sweep_chunk_fence(M - 1 - u);
      }
    }
    // B -> A: remove the vertical mean of the kh = 0 column
    // (src/elliptic.jl:930-945).  XF is free after the second barrier.
    if ((OP == SCAN_A || OP == SCAN_AC) && blk.x == 0) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int u = 0; u < M; ++u)
        if (k0 + u < nz) {
          sr = sr + yr[u];
          si = si + yi[u];
        }
      XF[w * TP + t] = sr;
      XF[(P + w) * TP + t] = si;
      __syncthreads();
      if (mean_col) {
        double tr = 0.0, ti = 0.0;
        for (int p = 0; p < P; ++p) {
          tr = tr + XF[p * TP + t];
          ti = ti + XF[(P + p) * TP + t];
        }
        const double mr = tr / nz, mi = ti / nz;
#pragma unroll
        for (int u = 0; u < M; ++u) {
          yr[u] = yr[u] - mr;
          yi[u] = yi[u] - mi;
        }
        Yr = Yr - mr;
        Yi = Yi - mi;
      }
    }
  }

  if (!q.active) return;
  if (OP == SCAN_OMEGA && a.pivoted[mem * rplane + rc0 + j]) return;  // k_omega_pivoted
  const bool zero_col = OP == SCAN_PSI && nz > 1 && mean_col;
  T* Co = (OP == SCAN_AC) ? static_cast<T*>(a.Cout) + 2 * mem * nz * plane : nullptr;
  const auto ro = rsrc(out);
  const auto rcz = rsrc((OP == SCAN_AC) ? Co : out);
  const double idz = 1.0 / g.dz;  // C = dA/dz by a multiplication per level
#pragma unroll
  for (int u = 0; u < M; ++u) {
    const int k = k0 + u;
    if (k < nz) {
      const double rr = zero_col ? 0.0 : yr[u], ri = zero_col ? 0.0 : yi[u];
      double cr = 0.0, ci = 0.0;
      if constexpr (OP == SCAN_AC) {
        const double ur = u + 1 < M ? yr[u + 1] : Yr, ui = u + 1 < M ? yi[u + 1] : Yi;
        cr = k < nz - 1 ? (ur - rr) * idz : 0.0;
        ci = k < nz - 1 ? (ui - ri) * idz : 0.0;
      }
      if constexpr (UNI) {
        scan_store2<T>(rr, ri, ro, boff, u * (int)lvlb);
        if constexpr (OP == SCAN_AC) scan_store2<T>(cr, ci, rcz, boff, u * (int)lvlb);
      } else {
        T* p = at_level(out, k);
        p[0] = (T)rr;
        p[1] = (T)ri;
        if constexpr (OP == SCAN_AC) {
          T* pc = at_level(Co, k);
          pc[0] = (T)cr;
          pc[1] = (T)ci;
        }
      }
    }
  }
}

// Logical block x: tile of QC quarter columns; y: member (xcd_block()).
template <int OP, int M, int QC, typename T>
__global__ void __launch_bounds__((ScanBounds<M, QC>::MAXT))
k_colscan(Geo g, Quarter qr, ScanArgs a) {
  extern __shared__ double lds[];
  colscan_block<OP, M, QC, T>(g, qr, a, xcd_block(), lds);
}

// Two independent solves in one launch: logical blocks x < n0 solve OP0 with
// a0, the others OP1 with a1 (tile x - n0).  The blocks of both are
// interleaved over the grid, so the two solves run side by side, as two
// launches on two streams would, without the streams' synchronization.
template <int OP0, int OP1, int M, int QC, typename T>
__global__ void __launch_bounds__((ScanBounds<M, QC>::MAXT))
k_colscan2(Geo g, Quarter qr, ScanArgs a0, ScanArgs a1, int n0) {
  extern __shared__ double lds[];
  Blk b = xcd_block();
  if (b.x < n0) {
    colscan_block<OP0, M, QC, T>(g, qr, a0, b, lds);
  } else {
    b.x -= n0;
    colscan_block<OP1, M, QC, T>(g, qr, a1, b, lds);
  }
}

// Default shape of the paired solve (k_colscan2, twice the blocks of one
// solve): the 16-quarter-column tiles (one wave per partition, so each
// level's address is a scalar base plus a lane offset) at the smallest M
// that fits; else the single solve's default.
inline int scan_pair_default_shape(int nz) {
  for (int i = 0; i < kScanShapeCount; ++i) {
    ScanShape s = scan_shape_of(i, nz);
    if (s.QC == 16 && scan_shape_fits(s)) return i;
  }
  return scan_default_shape(nz);
}

// Host side: launch OP with compiled shape `idx` (see kScanShapes) on fields
// stored in FP32 (f32) or FP64.
template <int OP>
void launch_colscan(const Geo& g, hipStream_t st, int idx, const ScanArgs& a, bool f32);
// OP0 with a0 and OP1 with a1 in one launch (k_colscan2), both at shape idx.
template <int OP0, int OP1>
void launch_colscan2(const Geo& g, hipStream_t st, int idx, const ScanArgs& a0, const ScanArgs& a1, bool f32);
// Raise the dynamic LDS limit of every instantiation (once per process).
bool colscan_setup();

}  // namespace qg
