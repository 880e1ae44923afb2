// Fused FFT-pass kernels of the ETD-RK2 step.
//
// A 2-D transform of one z-plane is two passes: the Y-pass transforms columns
// (stride nx) for a tile of S adjacent columns of one level, the X-pass
// transforms rows (unit stride).  Spectral-space arithmetic (masks, i*k
// multipliers, Hermitian parts, unpacking, the RHS combine and the ETD update)
// is fused into the Y-pass that touches the spectral data; physical-space
// products are fused into the X-pass.  Real physical fields travel as
// half-spectrum intermediates (columns 0..nx/2 of the y-transformed data).
//
// Sign convention: forward = exp(-i k.x), unnormalised; inverse = exp(+i k.x)
// scaled by 1/(nx*ny) (applied when the spectral input is loaded).
#pragma once

#include "qg_fft.hpp"
#include "qg_flags.hpp"

namespace qg {

// Tile of a pass: S sequences of length N per block, T lanes per sequence.
// The tile variant V sets the values a block holds: half (V = 0), the base
// of 256 FP64 / 512 FP32 values (V = 1) or twice that (V = 2), within 64
// sequences, 512 threads (larger blocks cap a kernel at 128 VGPRs, which the
// multi-field passes exceed) and 64 KB of LDS.  Every sequence's transform is
// the same in every variant, so variants give bitwise identical results
// (kTileVariants of them, qg_flags.hpp).

// LDS stride of a sequence, in values: at least the padded length `base`
// plus one, and congruent to 16/S modulo 16 when S divides 16.  A 64-bit LDS
// access serves 16 lanes per cycle from the banks; the Y-passes put
// consecutive threads on consecutive sequences, so a group of 16 lanes holds
// 16/S elements of each of S sequences, and with this stride the sequences'
// copies of one element pattern fall on distinct banks (S = 8 at stride 289,
// 1 mod 16, runs 64-bit accesses 1.6x slower than at 290).
constexpr int lds_stride(int base, int S) {
  const int want = (S >= 2 && S < 16 && 16 % S == 0) ? 16 / S : 1;
  int ld = base + 1;
  while (ld % 16 != want) ++ld;
  return ld;
}

template <typename R, int N, int V = 1>
struct Tile {
  // V = 3: the alternative FFT plan (FftPlan<N, 1>) in blocks of the half
  // tile (as V = 0).
  static constexpr int PL = V == 3 ? 1 : 0;
  static constexpr int T = FftPlan<N, PL>::T;
  static constexpr int BASE = sizeof(R) == 8 ? 256 : 512;
  static constexpr int TARGET = (V == 0 || V == 3) ? BASE / 2 : (V == 1 ? BASE : 2 * BASE);
  static constexpr int S0 = TARGET / T;
  static constexpr int S1 = S0 < 1 ? 1 : (S0 > 64 ? 64 : S0);
  static constexpr int S2 = S1 * T > 512 ? (512 / T < 1 ? 1 : 512 / T) : S1;
  static constexpr int LD = lds_stride(N + N / FftPlan<N, PL>::r0, S2);  // padded sequence stride in LDS
  // sequences that fit 64 KB of LDS beside the staged twiddle table and a
  // staged wavenumber table (stage_k)
  static constexpr int SL = (int)((65536 - TwTab<R, N, PL>::bytes() - N * (int)sizeof(double)) /
                                  (LD * sizeof(cplx<R>)));
  static constexpr int S = S2 > SL ? (SL < 1 ? 1 : SL) : S2;
  static constexpr int THREADS = S * T;
  static constexpr int NN = N;
  // the X-passes put each sequence on T consecutive threads: within one
  // wave when T divides 64 (their transforms then synchronize per wave)
  static constexpr bool XWAVE = 64 % T == 0;
};

template <typename R>
__device__ inline cplx<R> czero() { return mk<R>((R)0, (R)0); }

// Hermitian part (X(k) + conj X(-k))/2, pre-scaled by `s`, at precision T.
template <typename T>
__device__ inline cplx<T> herm(cplx<T> a, cplx<T> am, T s) {
  return mk<T>((T)0.5 * s * (a.re + am.re), (T)0.5 * s * (a.im - am.im));
}

// Bounds-guarded loads, issued unconditionally: a lane outside the field
// (a partial tile, a row past the last, an unused register slot) reads
// element 0, and its value is never stored.  The compiler turns a load that
// feeds only a per-lane select into a branch around the load, and the
// arithmetic it sinks into that branch waits for the load, so a pass's
// input loads would no longer all be in flight together.
template <typename T>
__device__ inline T ldu(const T* __restrict__ p, size_t i, bool in) {
  return p[in ? i : 0];
}

// v where k holds, else +0, as a bit mask: a select on a loaded value would
// again become a branch around the load.
__device__ inline float keep_bits(float v, bool k) { return __uint_as_float(__float_as_uint(v) & (k ? ~0u : 0u)); }
__device__ inline double keep_bits(double v, bool k) {
  return __longlong_as_double(__double_as_longlong(v) & (k ? ~0ll : 0ll));
}
template <typename R>
__device__ inline cplx<R> keep_if(cplx<R> v, bool k) {
  return mk<R>(keep_bits(v.re, k), keep_bits(v.im, k));
}

// p with its value hidden from the optimizer: loads through the result are
// neither merged with earlier loads of the same addresses nor hoisted above
// the point where it is taken (the empty asm produces no instruction).
template <typename T>
__device__ inline const T* opaque(const T* p) {
  asm volatile("" : "+s"(p));
  return p;
}
// The same for a pointer into LDS, kept in the LDS address space so that
// reads through it stay LDS reads.
using lds_cdouble = __attribute__((address_space(3))) const double;
__device__ inline lds_cdouble* opaque_lds(const double* p) {
  lds_cdouble* q = (lds_cdouble*)p;
  asm volatile("" : "+s"(q));
  return q;
}

// A wavenumber table of length N (ky for the Y-passes) copied into LDS by the
// block, so the epilogues' per-point reads wait for LDS rather than for a
// global load each.  Every thread calls it, and it ends with a barrier.
template <int N>
__device__ inline const double* stage_k(const double* __restrict__ k) {
  __shared__ double s[N];
  for (int i = threadIdx.x; i < N; i += blockDim.x) s[i] = k[i];
  __syncthreads();
  return s;
}

template <typename R>
__device__ inline bool finite_c(cplx<R> a) {
  return __builtin_isfinite(a.re) && __builtin_isfinite(a.im);
}

// Anti-Hermitian part (X(k) - conj X(-k))/2, pre-scaled by `s`.
template <typename T>
__device__ inline cplx<T> aherm(cplx<T> a, cplx<T> am, T s) {
  return mk<T>((T)0.5 * s * (a.re - am.re), (T)0.5 * s * (a.im + am.im));
}

// Folds a block's largest value into absmax (the bit pattern of a
// non-negative double; NaN counts as +inf): waves by shuffles, then LDS, and
// at most one atomic per block, skipped when the accumulator already holds
// more (thousands of blocks contend for one address otherwise).  Every
// thread of the block calls it.
template <int THREADS>
__device__ inline void fold_block_max(double m, unsigned long long* __restrict__ absmax) {
  if (m != m) m = __builtin_huge_val();
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(m, o);
    m = w > m ? w : m;
  }
  constexpr int NW = (THREADS + 63) / 64;
  __shared__ double wmax[NW];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int i = 0; i < NW; ++i) b = wmax[i] > b ? wmax[i] : b;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(b);
    if (b > 0.0 && bits > __hip_atomic_load(absmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(absmax, bits);
  }
}

// Adds the block's sum of one value per thread to *acc (one atomic per
// block).  Every thread of the block calls it.
template <int THREADS>
__device__ inline void fold_block_sum(double v, double* __restrict__ acc) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  constexpr int NW = (THREADS + 63) / 64;
  __shared__ double wsum[NW];
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int i = 0; i < NW; ++i) b += wsum[i];
    atomicAdd(acc, b);
  }
}

// ===========================================================================
// Wave feedback q_w  (compute_qw_complex!, nonlinear.jl:859-944)
// ===========================================================================

// Y-pass: Bm = mask*B (scaled 1/N) -> T1 = IFFTy(Bm), T3 = IFFTy(i*ky*Bm).
template <typename R, int NY, int V>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_ya_qw(Geo g, const cplx<R>* __restrict__ B, cplx<R>* __restrict__ T1,
        cplx<R>* __restrict__ T3, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, +1, TL::PL>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const Blk blk = xcd_block();
  const int x = fw_x(g, TL::S, blk.x, s), z = blk.y;
  const bool ok = x < g.nx && g.col_kept(x);
  cplx<R>* sl = lds + s * TL::LD;
  const double* __restrict__ ky = stage_k<NY>(g.ky);
  cplx<R> v[F::IT][F::R0], w[F::IT][F::R0];
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      const bool in = ok && F::valid(lane, it);
      const size_t a = ((size_t)z * NY + y) * g.nx + x;
      const bool keep = in && g.keep(x, y);
      const cplx<double> b = scale(cvt<double>(keep_if(ldu(B, a, in && g.in_disk(x, y)), keep)), g.inv_n);
      v[it][t] = cvt<R>(b);
      w[it][t] = cvt<R>(ik_mul(ky[y], b));
    }
  F::run(v, sl, lane, twr);
  if (T3) F::run(w, sl, lane, twr);  // uniform: T3 == nullptr computes T1 only
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      if (ok && F::valid(lane, it)) {
        size_t a = ((size_t)z * NY + y) * g.nx + x;
        T1[a] = v[it][t];
        if (T3) T3[a] = w[it][t];
      }
    }
}

// X-pass: B = IFFTx(T1), Bx = IFFTx(i*kx*T1), By = IFFTx(T3);
// qwr = Re(i/2 (conj(Bx)By - conj(By)Bx)) = Bx.im*By.re - Bx.re*By.im,
// |B|^2; forward real FFTs along x -> Qa (qwr), Qb (|B|^2), half width.
template <typename R, int NX, int V>
__global__ void __launch_bounds__((Tile<R, NX, V>::THREADS))
k_xa_qw(Geo g, const cplx<R>* __restrict__ T1, const cplx<R>* __restrict__ T3,
        cplx<R>* __restrict__ Qa, cplx<R>* __restrict__ Qb, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NX, V>;
  using FI = Fft<R, NX, +1, TL::PL, TL::XWAVE>;
  using FF = Fft<R, NX, -1, TL::PL, TL::XWAVE>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int r = threadIdx.x / TL::T, lane = threadIdx.x % TL::T;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const int row = blockIdx.x * TL::S + r;
  const bool ok = row < g.nz * g.ny;
  cplx<R>* sl = lds + r * TL::LD;
  const size_t base = (size_t)row * NX;
  const double* __restrict__ kx = stage_k<NX>(g.kx);
  cplx<R> v[FI::IT][FI::R0], w[FI::IT][FI::R0], y3[FI::IT][FI::R0];
  R mag[FI::IT][FI::R0];
  // B and Bx; the By row is loaded with them
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      int x = FI::elem(lane, it, t);
      // columns a pruned pass does not carry are zero
      const bool in = ok && FI::valid(lane, it), kc = g.col_kept(x);
      cplx<R> a = keep_if(ldu(T1, base + x, in && kc), kc);
      y3[it][t] = keep_if(ldu(T3, base + x, in && kc), kc);
      v[it][t] = a;
      w[it][t] = ik_mul((R)kx[x], a);
    }
  FI::run(v, sl, lane, twr);
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) mag[it][t] = v[it][t].re * v[it][t].re + v[it][t].im * v[it][t].im;
  FI::run(w, sl, lane, twr);  // w = Bx
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) v[it][t] = y3[it][t];
  FI::run(v, sl, lane, twr);  // v = By
  const size_t hb = (size_t)row * g.Wp;
  if constexpr (TL::XWAVE) {
    // qwr and |B|^2 (both real) in one packed forward transform, unpacked
    // through LDS.  |B|^2 is scaled by the power of two that brings its row
    // maximum to qwr's: qwr carries an x and a y derivative, so the two
    // differ by many orders of magnitude, and unscaled the packed transform
    // would leave round-off of |B|^2's size in qwr's coefficients.  A power
    // of two scales and unscales exactly.  A row's T lanes lie in one wave.
    R mq = 0, mb = 0;
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        const R q = w[it][t].im * v[it][t].re - w[it][t].re * v[it][t].im;
        v[it][t].re = q;
        if (FI::valid(lane, it)) {
          mq = fmax(mq, fabs(q));
          mb = fmax(mb, mag[it][t]);
        }
      }
#pragma unroll
    for (int o = 1; o < TL::T; o <<= 1) {
      mq = fmax(mq, __shfl_xor(mq, o, TL::T));
      mb = fmax(mb, __shfl_xor(mb, o, TL::T));
    }
    int eq = 0, eb = 0;
    (void)frexp(mq, &eq);
    (void)frexp(mb, &eb);
    const int e = (mq > 0 && mb > 0) ? min(max(eq - eb, -100), 100) : 0;
    const R s = ldexp((R)1, e), is = ldexp((R)0.5, -e);
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) v[it][t].im = s * mag[it][t];
    FF::run(v, sl, lane, twr);
    FI::sync();  // the last stage's LDS reads are done
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t)
        if (FI::valid(lane, it)) sl[FI::pad(FI::elem(lane, it, t))] = v[it][t];
    FI::sync();
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        int x = FI::elem(lane, it, t);
        if (ok && FI::valid(lane, it) && x < g.Kx) {
          const cplx<R> a = v[it][t], m = conj(sl[FI::pad(x == 0 ? 0 : NX - x)]);
          Qa[hb + x] = mk<R>((R)0.5 * (a.re + m.re), (R)0.5 * (a.im + m.im));
          Qb[hb + x] = mk<R>(is * (a.im - m.im), -is * (a.re - m.re));
        }
      }
    return;
  }
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      R q = w[it][t].im * v[it][t].re - w[it][t].re * v[it][t].im;
      v[it][t] = mk<R>(q, (R)0);
      w[it][t] = mk<R>(mag[it][t], (R)0);
    }
  FF::run(v, sl, lane, twr);
  FF::run(w, sl, lane, twr);
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      int x = FI::elem(lane, it, t);
      if (ok && FI::valid(lane, it) && x < g.Kx) {
        Qa[hb + x] = v[it][t];
        Qb[hb + x] = w[it][t];
      }
    }
}

// Y-pass (forward, half width): q_w = mask*(F(qwr) - kh2/4 F(|B|^2))/f and
// q_eff = mask ? q - q_w : 0  (replace_q_with_wave_feedback_rhs!, timestep.jl:122-140).
template <typename R, int NY, int V>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_ya2_qeff(Geo g, const cplx<R>* __restrict__ Qa, const cplx<R>* __restrict__ Qb,
           const cplx<R>* __restrict__ q, cplx<R>* __restrict__ qeff,
           cplx<R>* __restrict__ qw_out, double inv_f, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, -1, TL::PL>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const Blk blk = xcd_block();
  const int x = blk.x * TL::S + s, z = blk.y;
  const bool ok = x < g.Kx;
  cplx<R>* sl = lds + s * TL::LD;
  const double* __restrict__ ky = stage_k<NY>(g.ky);
  const double kxv = ok ? g.kx[x] : 0.0;
  cplx<R> v[F::IT][F::R0], w[F::IT][F::R0];
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      bool in = ok && F::valid(lane, it);
      size_t a = ((size_t)z * NY + y) * g.Wp + x;
      // guarded by `in` (no branch forms here, and the clamped form of the
      // other passes costs this one a wave per SIMD in registers)
      v[it][t] = in ? Qa[a] : czero<R>();
      w[it][t] = in ? Qb[a] : czero<R>();
    }
  F::run(v, sl, lane, twr);
  F::run(w, sl, lane, twr);
  const int mx = mirror(x, g.nx);
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      if (!(ok && F::valid(lane, it))) continue;
      const bool keep = g.keep(x, y);
      if (g.pruned() && !keep) continue;  // zero there already
      const double kh2 = kxv * kxv + ky[y] * ky[y];
      cplx<double> fq = cvt<double>(v[it][t]), fm = cvt<double>(w[it][t]);
      cplx<double> qw = keep ? scale(fq - scale(fm, 0.25 * kh2), inv_f) : czero<double>();
      size_t a = ((size_t)z * NY + y) * g.nx + x;
      if (qeff) qeff[a] = cvt<R>(keep ? cvt<double>(q[a]) - qw : czero<double>());
      if (qw_out) qw_out[a] = cvt<R>(qw);
      if (mx != x) {
        const int my = mirror(y, NY);
        size_t am = ((size_t)z * NY + my) * g.nx + mx;
        cplx<double> qwm = conj(qw);
        if (qeff) qeff[am] = cvt<R>(keep ? cvt<double>(q[am]) - qwm : czero<double>());
        if (qw_out) qw_out[am] = cvt<R>(qwm);
      }
    }
}

// ===========================================================================
// Advection and refraction  (nonlinear.jl:414-479, 621-676; operators.jl:195-274)
// ===========================================================================

// Y-pass (inverse, half width) producing the y-transformed Hermitian parts of
// u = Re ifft(-i ky psi), v = Re ifft(i kx psi) (unmasked psi), of
// Re ifft(mask q) and of zeta = Re ifft(mask (-kh2 psi)).
// OUTS: the output set fixed at compile time (0: `outs` at run time).  Each
// fixed set keeps only its own inputs live, which roughly halves the
// registers of the two-output launches.
template <typename R, int NY, int V, int OUTS>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_yb_adv(Geo g, int outs_rt, const cplx<R>* __restrict__ psi, const cplx<R>* __restrict__ q,
         cplx<R>* __restrict__ Pu, cplx<R>* __restrict__ Pv, cplx<R>* __restrict__ Pq,
         cplx<R>* __restrict__ Pz, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, +1, TL::PL>;
  const int outs = OUTS ? OUTS : outs_rt;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const Blk blk = xcd_block();
  const int x = blk.x * TL::S + s, z = blk.y;
  const bool ok = x < g.Kx;
  const int mx = mirror(x, g.nx);
  cplx<R>* sl = lds + s * TL::LD;
  const double kxv = ok ? g.kx[x] : 0.0, kxm = ok ? g.kx[mx] : 0.0;
  const double* __restrict__ ky = stage_k<NY>(g.ky);
  // psi and its mirror are loaded once, in storage precision; every output's
  // transform input is formed from them at the transform precision R
  cplx<R> p[F::IT][F::R0], pm[F::IT][F::R0];
  cplx<R> v[F::IT][F::R0];
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t), my = mirror(y, NY);
      // (x, y) and its mirror are in the disk together
      bool in = ok && F::valid(lane, it) && g.in_disk(x, y);
      // guarded by `in`: the clamped form costs this pass a wave per SIMD
      p[it][t] = in ? psi[((size_t)z * NY + y) * g.nx + x] : czero<R>();
      pm[it][t] = in ? psi[((size_t)z * NY + my) * g.nx + mx] : czero<R>();
    }
  for (int o = 0; o < 4; ++o) {
    const int bit = 1 << o;
    if (!(outs & bit)) continue;  // uniform across the block
    cplx<R>* dst = o == 0 ? Pu : o == 1 ? Pv : o == 2 ? Pq : Pz;
#pragma unroll
    for (int it = 0; it < F::IT; ++it)
#pragma unroll
      for (int t = 0; t < F::R0; ++t) {
        int y = F::elem(lane, it, t), my = mirror(y, NY);
        cplx<R> a, am;
        // u and v as k_yf_omega forms them (the same values whichever pass
        // makes them): the multiplier times the Hermitian part of psi, or
        // its anti-Hermitian part where the multiplier's wavenumber is its
        // own mirror (ky at the row ny/2, kx at the column nx/2)
        if (o == 0) {
          const cplx<R> f = 2 * y == NY ? aherm(p[it][t], pm[it][t], (R)g.inv_n)
                                        : herm(p[it][t], pm[it][t], (R)g.inv_n);
          v[it][t] = ik_mul((R)-ky[y], f);
          continue;
        } else if (o == 1) {
          const cplx<R> f = (mx == x && x != 0) ? aherm(p[it][t], pm[it][t], (R)g.inv_n)
                                                : herm(p[it][t], pm[it][t], (R)g.inv_n);
          v[it][t] = ik_mul((R)kxv, f);
          continue;
        } else if (o == 2) {
          // loaded here, under the mask, rather than with psi: as unconditional
          // loads they would be hoisted to the kernel's start and cost a wave
          // per SIMD in registers
          bool in = ok && F::valid(lane, it);
          bool keep = in && g.keep(x, y);
          a = keep ? q[((size_t)z * NY + y) * g.nx + x] : czero<R>();
          am = keep ? q[((size_t)z * NY + my) * g.nx + mx] : czero<R>();
        } else {
          bool in = ok && F::valid(lane, it);
          bool keep = in && g.keep(x, y);
          double kh2 = kxv * kxv + ky[y] * ky[y];
          double kh2m = kxm * kxm + ky[my] * ky[my];
          a = keep ? scale(p[it][t], (R)-kh2) : czero<R>();
          am = keep ? scale(pm[it][t], (R)-kh2m) : czero<R>();
        }
        v[it][t] = herm(a, am, (R)g.inv_n);
      }
    F::run(v, sl, lane, twr);
#pragma unroll
    for (int it = 0; it < F::IT; ++it)
#pragma unroll
      for (int t = 0; t < F::R0; ++t) {
        int y = F::elem(lane, it, t);
        if (ok && F::valid(lane, it)) dst[((size_t)z * NY + y) * g.Wp + x] = v[it][t];
      }
  }
}

// Load one half-width row as a full Hermitian row (element x >= W is
// conj(element nx-x)); columns from Kx on (those a pruned pass does not carry)
// are zero.
template <typename R, class F>
__device__ inline void load_herm_row(cplx<R> (&v)[F::IT][F::R0], const cplx<R>* __restrict__ src,
                                     size_t hb, int nx, int Kx, int lane, bool ok) {
  const int W = nx / 2 + 1;
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      const int x = F::elem(lane, it, t);
      const bool lo = x < W;
      const int c = lo ? x : nx - x;
      const bool k = c < Kx, in = ok && F::valid(lane, it) && k;
      const cplx<R> a = keep_if(src[in ? hb + c : 0], k);
      v[it][t] = lo ? a : conj(a);
    }
}

// Two half-width rows of real fields' spectra packed as one complex row,
// Z = A + i B over the full row (x >= W from the conjugate mirrors; the
// self-conjugate points x = 0 and nx/2 take their real parts, which is what
// the real part of a single inverse transform keeps).  One inverse transform
// then gives a in .re and b in .im.  The pairs packed this way are fields of
// one kind and scale (u and v; the two components of a gradient).
template <typename R, class F>
__device__ inline void load_herm_pair(cplx<R> (&v)[F::IT][F::R0], const cplx<R>* __restrict__ A,
                                      const cplx<R>* __restrict__ B, size_t hb, int nx, int Kx, int lane,
                                      bool ok) {
  const int W = nx / 2 + 1;
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      const int x = F::elem(lane, it, t);
      const bool lo = x < W, self = x == 0 || 2 * x == nx;
      const int c = lo ? x : nx - x;
      const bool k = c < Kx, in = ok && F::valid(lane, it) && k;
      const size_t i = in ? hb + c : 0;
      cplx<R> a = keep_if(A[i], k), b = keep_if(B[i], k);
      // mirrored points take the conjugate; the self-conjugate ones their real part
      a.im = lo ? (self ? (R)0 : a.im) : -a.im;
      b.im = lo ? (self ? (R)0 : b.im) : -b.im;
      v[it][t] = mk<R>(a.re - b.im, a.im + b.re);
    }
}

// X-pass: real rows from half-width intermediates (C2R), written to `outs`.
// Used for u, v (velocities) and w.
template <typename R, int NX, int V>
__global__ void __launch_bounds__((Tile<R, NX, V>::THREADS))
k_x_c2r(Geo g, int nfields, const cplx<R>* __restrict__ P0, const cplx<R>* __restrict__ P1,
        R* __restrict__ o0, R* __restrict__ o1, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NX, V>;
  using FI = Fft<R, NX, +1, TL::PL, TL::XWAVE>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int r = threadIdx.x / TL::T, lane = threadIdx.x % TL::T;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const int row = blockIdx.x * TL::S + r;
  const bool ok = row < g.nz * g.ny;
  cplx<R>* sl = lds + r * TL::LD;
  cplx<R> v[FI::IT][FI::R0];
  if (nfields == 2) {  // u and v: one packed transform
    load_herm_pair<R, FI>(v, P0, P1, (size_t)row * g.Wp, NX, g.Kx, lane, ok);
  } else {
    load_herm_row<R, FI>(v, P0, (size_t)row * g.Wp, NX, g.Kx, lane, ok);
  }
  FI::run(v, sl, lane, twr);
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      int x = FI::elem(lane, it, t);
      if (ok && FI::valid(lane, it)) {
        o0[(size_t)row * NX + x] = v[it][t].re;
        if (nfields == 2) o1[(size_t)row * NX + x] = v[it][t].im;
      }
    }
}

// X-pass for q advection: q = C2R(Pq); u*q, v*q -> forward real FFTs -> G1, G2.
// UVP = 0: u and v are read as physical rows.  UVP = 1: they are made here
// from their y-transformed half rows Pu, Pv (one packed inverse transform, as
// k_x_c2r does) and also written out to u and v for the B pass that follows.
template <typename R, int NX, int V, int UVP>
__global__ void __launch_bounds__((Tile<R, NX, V>::THREADS))
k_xb_q(Geo g, const cplx<R>* __restrict__ Pq, const cplx<R>* __restrict__ Pu, const cplx<R>* __restrict__ Pv,
       R* __restrict__ u, R* __restrict__ vv, cplx<R>* __restrict__ G1, cplx<R>* __restrict__ G2,
       const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NX, V>;
  using FI = Fft<R, NX, +1, TL::PL, TL::XWAVE>;
  using FF = Fft<R, NX, -1, TL::PL, TL::XWAVE>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int r = threadIdx.x / TL::T, lane = threadIdx.x % TL::T;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const int row = blockIdx.x * TL::S + r;
  const bool ok = row < g.nz * g.ny;
  cplx<R>* sl = lds + r * TL::LD;
  cplx<R> v[FI::IT][FI::R0];
  R qv[FI::IT][FI::R0];
  const size_t hb = (size_t)row * g.Wp;
  if constexpr (UVP) {
    // both inputs are loaded before the first transform
    cplx<R> w[FI::IT][FI::R0];
    load_herm_pair<R, FI>(w, Pu, Pv, hb, NX, g.Kx, lane, ok);
    load_herm_row<R, FI>(v, Pq, hb, NX, g.Kx, lane, ok);
    FI::run(w, sl, lane, twr);
    FI::run(v, sl, lane, twr);
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        int x = FI::elem(lane, it, t);
        if (ok && FI::valid(lane, it)) {
          u[(size_t)row * NX + x] = w[it][t].re;
          vv[(size_t)row * NX + x] = w[it][t].im;
        }
        qv[it][t] = v[it][t].re;
        v[it][t] = mk<R>(w[it][t].re * qv[it][t], w[it][t].im * qv[it][t]);
      }
  } else {
    load_herm_row<R, FI>(v, Pq, hb, NX, g.Kx, lane, ok);
    FI::run(v, sl, lane, twr);
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) qv[it][t] = v[it][t].re;
    // u*q and v*q (one kind and scale) packed as Z = uq + i vq: one forward
    // transform, then F(uq)[k] = (Z[k] + conj Z[-k])/2 and
    // F(vq)[k] = (Z[k] - conj Z[-k])/(2i), with Z[-k] read back through LDS.
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        int x = FI::elem(lane, it, t);
        const bool in = ok && FI::valid(lane, it);
        const R cu = ldu(u, (size_t)row * NX + x, in);
        const R cv = ldu(vv, (size_t)row * NX + x, in);
        v[it][t] = mk<R>(cu * qv[it][t], cv * qv[it][t]);
      }
  }
  FF::run(v, sl, lane, twr);
  FI::sync();  // the last stage's LDS reads are done
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t)
      if (FI::valid(lane, it)) sl[FI::pad(FI::elem(lane, it, t))] = v[it][t];
  FI::sync();
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      int x = FI::elem(lane, it, t);
      if (ok && FI::valid(lane, it) && x < g.Kx) {
        const cplx<R> a = v[it][t], m = conj(sl[FI::pad(x == 0 ? 0 : NX - x)]);
        G1[hb + x] = mk<R>((R)0.5 * (a.re + m.re), (R)0.5 * (a.im + m.im));
        G2[hb + x] = mk<R>((R)0.5 * (a.im - m.im), (R)-0.5 * (a.re - m.re));
      }
    }
}

// X-pass for B: B = IFFTx(T1), zeta = C2R(Pz); u*B, v*B, zeta*B -> forward
// complex FFTs along x.  The x-derivative of the flux and the refraction
// factor are per-kx multipliers, so they are applied here and the two terms
// leave as one field: G3 = -i kx F(uB) - (i/2) F(zeta B), G4 = F(vB) (full
// width).  k_yb2_B adds -i ky F_y(G4) after the y transform.
template <typename R, int NX, int V>
__global__ void __launch_bounds__((Tile<R, NX, V>::THREADS))
k_xb_B(Geo g, int flags, const cplx<R>* __restrict__ T1, const cplx<R>* __restrict__ Pz,
       const R* __restrict__ u, const R* __restrict__ vv, cplx<R>* __restrict__ G3,
       cplx<R>* __restrict__ G4, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NX, V>;
  using FI = Fft<R, NX, +1, TL::PL, TL::XWAVE>;
  using FF = Fft<R, NX, -1, TL::PL, TL::XWAVE>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int r = threadIdx.x / TL::T, lane = threadIdx.x % TL::T;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const int row = blockIdx.x * TL::S + r;
  const bool ok = row < g.nz * g.ny;
  cplx<R>* sl = lds + r * TL::LD;
  const size_t base = (size_t)row * NX;
  const double* __restrict__ kx = stage_k<NX>(g.kx);
  cplx<R> b[FI::IT][FI::R0], v[FI::IT][FI::R0], h[FI::IT][FI::R0];
  R zt[FI::IT][FI::R0], ur[FI::IT][FI::R0], vr[FI::IT][FI::R0];
  const bool adv = (flags & XB_ADV) != 0;  // uniform
  // every input row is loaded before the first transform (u and v included)
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      int x = FI::elem(lane, it, t);
      const bool in = ok && FI::valid(lane, it), kc = g.col_kept(x);
      b[it][t] = keep_if(ldu(T1, base + x, in && kc), kc);  // columns not carried are zero
      ur[it][t] = ldu(u, base + x, in && adv);
      vr[it][t] = ldu(vv, base + x, in && adv);
      h[it][t] = czero<R>();
    }
  if (flags & XB_REFR) load_herm_row<R, FI>(v, Pz, (size_t)row * g.Wp, NX, g.Kx, lane, ok);
  FI::run(b, sl, lane, twr);
  if (flags & XB_REFR) {
    FI::run(v, sl, lane, twr);
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) zt[it][t] = v[it][t].re;
  }
  // f = 0: u*B, 1: v*B, 2: zeta*B
  for (int f = 0; f < 3; ++f) {
    if (f < 2 && !(flags & XB_ADV)) continue;
    if (f == 2 && !(flags & XB_REFR)) continue;
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        const R c = f == 2 ? zt[it][t] : (f == 0 ? ur[it][t] : vr[it][t]);
        v[it][t] = mk<R>(c * b[it][t].re, c * b[it][t].im);
      }
    FF::run(v, sl, lane, twr);
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        int x = FI::elem(lane, it, t);
        if (f == 1) {
          if (ok && FI::valid(lane, it) && g.col_kept(x)) G4[base + x] = v[it][t];
        } else {
          const R m = f == 0 ? (R)-kx[x] : (R)-0.5;
          h[it][t] = h[it][t] + ik_mul(m, v[it][t]);
        }
      }
  }
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      int x = FI::elem(lane, it, t);
      if (ok && FI::valid(lane, it) && g.col_kept(x)) G3[base + x] = h[it][t];
    }
}

// The two advection X-passes of one RHS in one launch (q advected, B
// advected and refracted): what k_xb_q and k_xb_B compute, from one read of
// each input row.  UVP = 1: u and v come from their y-transformed half rows
// (one packed inverse transform) and are not written out.
template <typename R, int NX, int V, int UVP>
__global__ void __launch_bounds__((Tile<R, NX, V>::THREADS))
k_xb_all(Geo g, const cplx<R>* __restrict__ Pq, const cplx<R>* __restrict__ Pz, const cplx<R>* __restrict__ T1,
         const cplx<R>* __restrict__ Pu, const cplx<R>* __restrict__ Pv, const R* __restrict__ u,
         const R* __restrict__ vv, cplx<R>* __restrict__ G1, cplx<R>* __restrict__ G2, cplx<R>* __restrict__ G3,
         cplx<R>* __restrict__ G4, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NX, V>;
  using FI = Fft<R, NX, +1, TL::PL, TL::XWAVE>;
  using FF = Fft<R, NX, -1, TL::PL, TL::XWAVE>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int r = threadIdx.x / TL::T, lane = threadIdx.x % TL::T;
  TwTab<R, TL::NN, TL::PL, sizeof(R) == 8> twr;  // FP64: register twiddles (measured faster)
  twr.load(lane, tw);
  const int row = blockIdx.x * TL::S + r;
  const bool ok = row < g.nz * g.ny;
  cplx<R>* sl = lds + r * TL::LD;
  const size_t base = (size_t)row * NX, hb = (size_t)row * g.Wp;
  const double* __restrict__ kx = stage_k<NX>(g.kx);
  cplx<R> b[FI::IT][FI::R0], v[FI::IT][FI::R0];
  R ur[FI::IT][FI::R0], vr[FI::IT][FI::R0], zr[FI::IT][FI::R0];
  // every input row is loaded before the first transform
  if constexpr (UVP) {
    cplx<R> w[FI::IT][FI::R0];
    load_herm_pair<R, FI>(w, Pu, Pv, hb, NX, g.Kx, lane, ok);
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        int x = FI::elem(lane, it, t);
        const bool kc = g.col_kept(x);
        b[it][t] = keep_if(ldu(T1, base + x, ok && FI::valid(lane, it) && kc), kc);
      }
    load_herm_pair<R, FI>(v, Pq, Pz, hb, NX, g.Kx, lane, ok);
    FI::run(w, sl, lane, twr);
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        ur[it][t] = w[it][t].re;
        vr[it][t] = w[it][t].im;
      }
  } else {
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        int x = FI::elem(lane, it, t);
        const bool in = ok && FI::valid(lane, it), kc = g.col_kept(x);
        b[it][t] = keep_if(ldu(T1, base + x, in && kc), kc);
        ur[it][t] = ldu(u, base + x, in);
        vr[it][t] = ldu(vv, base + x, in);
      }
    load_herm_pair<R, FI>(v, Pq, Pz, hb, NX, g.Kx, lane, ok);
  }
  // q and zeta (both real, of one kind: q = zeta plus the stretching term) in
  // one packed inverse transform: q in .re, zeta in .im.  Then u*q + i v*q,
  // one forward transform, unpacked through LDS (as k_xb_q).
  FI::run(v, sl, lane, twr);
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      const R qv = v[it][t].re;
      zr[it][t] = v[it][t].im;
      v[it][t] = mk<R>(ur[it][t] * qv, vr[it][t] * qv);
    }
  FF::run(v, sl, lane, twr);
  FI::sync();  // the last stage's LDS reads are done
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t)
      if (FI::valid(lane, it)) sl[FI::pad(FI::elem(lane, it, t))] = v[it][t];
  FI::sync();
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      int x = FI::elem(lane, it, t);
      if (ok && FI::valid(lane, it) && x < g.Kx) {
        const cplx<R> a = v[it][t], m = conj(sl[FI::pad(x == 0 ? 0 : NX - x)]);
        G1[hb + x] = mk<R>((R)0.5 * (a.re + m.re), (R)0.5 * (a.im + m.im));
        G2[hb + x] = mk<R>((R)0.5 * (a.im - m.im), (R)-0.5 * (a.re - m.re));
      }
    }
  // B: u*B, v*B, zeta*B (as k_xb_B); FI::run starts with a barrier, so the
  // LDS reads above are done before it writes
  FI::run(b, sl, lane, twr);
  cplx<R> h[FI::IT][FI::R0];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        const R c = f == 2 ? zr[it][t] : (f == 0 ? ur[it][t] : vr[it][t]);
        v[it][t] = mk<R>(c * b[it][t].re, c * b[it][t].im);
      }
    FF::run(v, sl, lane, twr);
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        int x = FI::elem(lane, it, t);
        if (f == 1) {
          if (ok && FI::valid(lane, it) && g.col_kept(x)) G4[base + x] = v[it][t];
        } else if (f == 0) {
          h[it][t] = ik_mul((R)-kx[x], v[it][t]);
        } else {
          h[it][t] = h[it][t] + ik_mul((R)-0.5, v[it][t]);
        }
      }
  }
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      int x = FI::elem(lane, it, t);
      if (ok && FI::valid(lane, it) && g.col_kept(x)) G3[base + x] = h[it][t];
    }
}

// ETD-RK2 coefficients of one point, computed where they are used from its
// kh2 = kx^2 + ky^2 and the model's parameter block (qg_flags.hpp: dt and the
// closure terms, in device memory, rewritten in place when dt changes):
// lambda*dt = dt * sum_i c_i kh2^(o_i/2) (int_factor, src/nonlinear.jl:91-97,
// 1160-1183; zero when inviscid, no terms), then E = exp(-x), h phi_1 and
// h phi_2 with their series below |x| = 1e-6 (_etd_coefficients,
// src/timestep.jl:177-190).  `stage`: E and h1 are wanted too (the final
// update needs h2 only).  Computing them costs a few FP64 operations per
// point; loading them from per-point tables was a dependent round trip each.
struct Etd {
  const EtdParams* __restrict__ p;
  __device__ inline void coefs(int which, double kh2, bool stage, double& E, double& h1, double& h2) const {
    const double h = p->dt;
    const int n = p->n[which];
    double total = 0.0;
    for (int i = 0; i < n; ++i) {
      double pw = 1.0;
      for (int j = 0; j < p->half[which][i]; ++j) pw *= kh2;
      total = total + p->coef[which][i] * pw;
    }
    const double x = h * total;
    E = 1.0;
    h1 = h;
    if (stage) E = exp(-x);
    // Both branches run for every point (the branch is if-converted), so the
    // series scales its terms by the reciprocals of 6, 24, 120 and 720 as
    // constants: a division, even by a constant, is a full FP64 division
    // sequence here, and these were seven of the nine per point.  The terms
    // they scale are below 1e-11 of the leading one for |x| < 1e-6.
    constexpr double r6 = 1.0 / 6, r24 = 1.0 / 24, r120 = 1.0 / 120, r720 = 1.0 / 720;
    if (fabs(x) < 1e-6) {
      const double x2 = x * x;
      if (stage) h1 = h * (1 - x / 2 + x2 * r6 - x2 * x * r24 + x2 * x2 * r120);
      h2 = h * (1.0 / 2 - x * r6 + x2 * r24 - x2 * x * r120 + x2 * x2 * r720);
    } else {
      const double em1 = expm1(-x);
      if (stage) h1 = h * (-em1) / x;
      h2 = h * (em1 + x) / (x * x);
    }
  }
};

// Vertical diffusion nu_z * d2q/dz2 with Neumann ends (dissipation_q_nv!, nonlinear.jl:1010-1044).
// z is the stacked level; the column of its member has nzc levels.
template <typename S>
__device__ inline cplx<double> vdiff(const cplx<S>* __restrict__ q, int z, int nzc, size_t col,
                                     size_t plane, double nu, double inv_dz2) {
  if (nzc <= 1) return czero<double>();
  const int zl = z % nzc;
  cplx<double> c = cvt<double>(q[(size_t)z * plane + col]);
  cplx<double> r;
  if (zl == 0) {
    cplx<double> up = cvt<double>(q[(size_t)(z + 1) * plane + col]);
    r = up - c;
  } else if (zl == nzc - 1) {
    cplx<double> dn = cvt<double>(q[(size_t)(z - 1) * plane + col]);
    r = dn - c;
  } else {
    cplx<double> up = cvt<double>(q[(size_t)(z + 1) * plane + col]);
    cplx<double> dn = cvt<double>(q[(size_t)(z - 1) * plane + col]);
    r = mk<double>(up.re - 2 * c.re + dn.re, up.im - 2 * c.im + dn.im);
  }
  return scale(scale(r, nu), inv_dz2);
}

// Y-pass (forward, half width): nq = mask*(i kx F(uq) + i ky F(vq)); rhs_q and
// the ETD stage/final update of q (timestep.jl:286-295, 362-404).
// FL >= 0 fixes the flag word at compile time (the launcher passes the same
// value), which removes the per-point branches on it; FL = -1 reads `flags`.
// The pass's work for tile blk (blk.x: tile of S columns, blk.y: level);
// lds: the block's TL::S * TL::LD values.
template <typename R, int NY, int V, int FL = -1>
__device__ inline void yb2_q_block(const Geo& g, const Blk blk, cplx<R>* lds, int mode, int flags_rt,
                                   const cplx<R>* __restrict__ G1, const cplx<R>* __restrict__ G2,
                                   const cplx<R>* __restrict__ qd, const cplx<R>* __restrict__ qn,
                                   cplx<R>* __restrict__ rhs0, cplx<R>* __restrict__ qout, const Etd& etd,
                                   double nu, double inv_dz2, int* __restrict__ flag,
                                   const cplx<R>* __restrict__ tw) {
  const int flags = FL >= 0 ? FL : flags_rt;
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, -1, TL::PL>;
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const int x = blk.x * TL::S + s, z = blk.y;
  const bool ok = x < g.Kx;
  cplx<R>* sl = lds + s * TL::LD;
  const double* __restrict__ ky = stage_k<NY>(g.ky);
  const double kxv = ok ? g.kx[x] : 0.0;
  cplx<R> a[F::IT][F::R0], b[F::IT][F::R0];
  if (flags & YQ_ADV) {  // uniform
#pragma unroll
    for (int it = 0; it < F::IT; ++it)
#pragma unroll
      for (int t = 0; t < F::R0; ++t) {
        int y = F::elem(lane, it, t);
        bool in = ok && F::valid(lane, it);
        size_t ad = ((size_t)z * NY + y) * g.Wp + x;
        a[it][t] = ldu(G1, ad, in);
        b[it][t] = ldu(G2, ad, in);
      }
    F::run(a, sl, lane, twr);
    F::run(b, sl, lane, twr);
  }
  const int mx = mirror(x, g.nx);
  const size_t plane = (size_t)g.nx * NY;
  bool fin = true;
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      if (!(ok && F::valid(lane, it))) continue;
      const bool keep = g.keep(x, y);
      if (g.pruned() && !keep) continue;  // q, the bracket: zero there already
      cplx<double> nq = czero<double>();
      if ((flags & YQ_ADV) && keep)
        nq = ik_mul(kxv, cvt<double>(a[it][t])) + ik_mul(ky[y], cvt<double>(b[it][t]));
      const int my = mirror(y, NY);
      // The ETD coefficients depend on (kx, ky) through kh2 only, so the
      // mirror column (mx, my) has those of (x, y).
      const bool upd = keep && !(flags & YQ_FIXED);
      double E = 0.0, h1 = 0.0, h2 = 0.0;
      if (upd) etd.coefs(0, kxv * kxv + ky[y] * ky[y], mode == MODE_STAGE, E, h1, h2);
      for (int side = 0; side < 2; ++side) {
        if (side == 1 && mx == x) break;
        const int cx = side ? mx : x, cy = side ? my : y;
        const size_t col = (size_t)cy * g.nx + cx;
        const size_t ad = (size_t)z * plane + col;
        cplx<double> n = side ? conj(nq) : nq;
        cplx<double> rhs = czero<double>();
        if (upd) {
          cplx<double> dq = (flags & YQ_DIFF) ? vdiff(qd, z, g.nzc, col, plane, nu, inv_dz2) : czero<double>();
          rhs = mk<double>(-n.re + dq.re, -n.im + dq.im);
        }
        // ETD-RK2 (src/timestep.jl:362-404): a = E q0 + h1 r0 and
        // q1 = a + h2 (r1 - r0) = [E q0 + (h1 - h2) r0] + h2 r1; the stage
        // keeps the bracket for the final step, which then needs neither
        // q0 nor r0.
        cplx<double> out = czero<double>();
        if (keep) {
          if (flags & YQ_FIXED) {
            out = cvt<double>(qn[ad]);
            if (mode == MODE_STAGE) rhs0[ad] = czero<R>();
          } else if (mode == MODE_STAGE) {
            const cplx<double> q0 = cvt<double>(qn[ad]);
            out = scale(q0, E) + scale(rhs, h1);
            rhs0[ad] = cvt<R>(scale(q0, E) + scale(rhs, h1 - h2));
          } else {
            out = cvt<double>(rhs0[ad]) + scale(rhs, h2);
          }
        } else if (mode == MODE_STAGE) {
          rhs0[ad] = czero<R>();
        }
        const cplx<R> qs = cvt<R>(out);
        qout[ad] = qs;
        fin = fin && finite_c(qs);
      }
    }
  // flag: the termination check's non-finite test of the new q (every point
  // is written above, the mirror (-x, -y) included)
  if (flag && !fin) atomicOr(flag, 1);
}

template <typename R, int NY, int V, int FL = -1>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_yb2_q(Geo g, int mode, int flags_rt, const cplx<R>* __restrict__ G1, const cplx<R>* __restrict__ G2,
        const cplx<R>* __restrict__ qd, const cplx<R>* __restrict__ qn,
        cplx<R>* __restrict__ rhs0,
        cplx<R>* __restrict__ qout, Etd etd, double nu, double inv_dz2, int* __restrict__ flag,
        const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  yb2_q_block<R, NY, V, FL>(g, xcd_block(), lds, mode, flags_rt, G1, G2, qd, qn, rhs0, qout, etd, nu, inv_dz2,
                            flag, tw);
}

// Y-pass (forward, full width) of G3 = -i kx F_x(uB) - (i/2) F_x(zeta B) and
// G4 = F_x(vB): rhs_B = mask*(F_y(G3) - i ky F_y(G4)) + i (f/2) kh2 A, i.e.
// -nB - (i/2) rB + i (f/2) kh2 A with nB = i kx F(uB) + i ky F(vB) and
// rB = F(zeta B); then the ETD update of B and, with T1, the inverse y
// transforms of the new B that the next pass over B needs.
// The pass's work for tile blk (blk.x: full-width tile, blk.y: level); lds:
// the block's TL::S * TL::LD values.
template <typename R, int NY, int V>
__device__ inline void yb2_B_block(const Geo& g, const Blk blk, cplx<R>* lds, int mode, int flags,
                                   const cplx<R>* __restrict__ G3, const cplx<R>* __restrict__ G4,
                                   const cplx<R>* __restrict__ A, const cplx<R>* __restrict__ Bn,
                                   cplx<R>* __restrict__ rhs0, cplx<R>* __restrict__ Bout, cplx<R>* __restrict__ T1,
                                   cplx<R>* __restrict__ T3, const Etd& etd, double alpha, int* __restrict__ flag,
                                   const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, -1, TL::PL>;
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const int x = fw_x(g, TL::S, blk.x, s), z = blk.y;
  const bool ok = x < g.nx && g.col_kept(x);
  cplx<R>* sl = lds + s * TL::LD;
  const double* __restrict__ ky = stage_k<NY>(g.ky);
  cplx<R> v[F::IT][F::R0], w[F::IT][F::R0];
  const bool has3 = (flags & (YB_ADV | YB_REFR)) != 0, has4 = (flags & YB_ADV) != 0;  // uniform
  // both inputs are loaded before either transform: twice the bytes in flight
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      bool in = ok && F::valid(lane, it);
      const size_t ad = ((size_t)z * NY + y) * g.nx + x;
      v[it][t] = ldu(G3, ad, in && has3);
      w[it][t] = ldu(G4, ad, in && has4);
    }
  // F(G3) = -i kx F(uB) - (i/2) F(zeta B) and F(G4) = F(vB) stay at the
  // transform precision; their FP64 sum is formed per point below
  if (has3) F::run(v, sl, lane, twr);
  if (has4) F::run(w, sl, lane, twr);
  const size_t plane = (size_t)g.nx * NY;
  const double kxv = ok ? g.kx[x] : 0.0;
  // FP32 up to 256-point columns: the update's field inputs (A, and B_n or
  // the kept bracket) are loaded together, from clamped indices, before the
  // update; loaded under each point's mask they would be one round trip per
  // point.  FP64 and longer columns keep the late loads: grouped, they cost
  // those a wave per SIMD in registers.
  constexpr bool GROUP = sizeof(R) == 4 && NY <= 256;
  const cplx<R>* __restrict__ base = mode == MODE_STAGE ? Bn : rhs0;
  cplx<R> av[F::IT][F::R0], bv[F::IT][F::R0];
  if constexpr (GROUP) {
    const bool disp = (flags & YB_DISP) != 0;
#pragma unroll
    for (int it = 0; it < F::IT; ++it)
#pragma unroll
      for (int t = 0; t < F::R0; ++t) {
        const int y = F::elem(lane, it, t);
        const bool in = ok && F::valid(lane, it) && g.in_disk(x, y);
        const size_t ad = ((size_t)z * NY + y) * g.nx + x;
        av[it][t] = ldu(A, ad, in && disp);
        bv[it][t] = ldu(base, ad, in);
      }
  }
  bool fin = true;
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      const cplx<R> g3 = v[it][t];
      v[it][t] = czero<R>();
      if (!(ok && F::valid(lane, it))) continue;
      const bool keep = g.keep(x, y);
      if (g.pruned() && !keep) continue;  // B, the bracket: zero there already
      const size_t col = (size_t)y * g.nx + x;
      const size_t ad = (size_t)z * plane + col;
      cplx<double> rhs = czero<double>();
      if (keep) {
        cplx<double> acc = has3 ? cvt<double>(g3) : czero<double>();
        if (has4) acc = acc + ik_mul(-ky[y], cvt<double>(w[it][t]));  // - i ky F(vB)
        cplx<double> t1 = czero<double>();
        if (flags & YB_DISP) {
          const double kh2 = kxv * kxv + ky[y] * ky[y];
          const double c = alpha * kh2;
          cplx<double> Av = cvt<double>(GROUP ? av[it][t] : A[ad]);
          t1 = mk<double>(-(c * Av.im), c * Av.re);
        }
        rhs = acc + t1;
      }
      // ETD-RK2 as for q: the stage keeps E B0 + (h1 - h2) r0, the final
      // step adds h2 r1 to it
      cplx<double> out = czero<double>();
      if (keep) {
        double E, h1, h2;
        etd.coefs(1, kxv * kxv + ky[y] * ky[y], mode == MODE_STAGE, E, h1, h2);
        if (mode == MODE_STAGE) {
          const cplx<double> b0 = cvt<double>(GROUP ? bv[it][t] : base[ad]);
          out = scale(b0, E) + scale(rhs, h1);
          rhs0[ad] = cvt<R>(scale(b0, E) + scale(rhs, h1 - h2));
        } else {
          out = cvt<double>(GROUP ? bv[it][t] : base[ad]) + scale(rhs, h2);
        }
      } else if (mode == MODE_STAGE) {
        rhs0[ad] = czero<R>();
      }
      v[it][t] = cvt<R>(out);
      Bout[ad] = v[it][t];
      fin = fin && finite_c(v[it][t]);
    }
  // flag: the termination check's non-finite test of the new B
  if (flag && !fin) atomicOr(flag, 1);
  // The next pass over these columns would be the inverse y transform of the
  // new B (ya_qw: T1 = IFFTy(mask B / (nx ny)), T3 = IFFTy(i ky mask B /
  // (nx ny))); it is made here from the stored values, which the column
  // already holds.
  if (T1) {  // uniform
    using FI = Fft<R, NY, +1, TL::PL>;
#pragma unroll
    for (int it = 0; it < F::IT; ++it)
#pragma unroll
      for (int t = 0; t < F::R0; ++t) {
        int y = F::elem(lane, it, t);
        const bool in = ok && F::valid(lane, it);
        const cplx<double> bm =
            (in && g.keep(x, y)) ? scale(cvt<double>(v[it][t]), g.inv_n) : czero<double>();
        v[it][t] = cvt<R>(bm);
        w[it][t] = cvt<R>(ik_mul(ky[y], bm));
      }
    FI::run(v, sl, lane, twr);
    if (T3) FI::run(w, sl, lane, twr);  // uniform
#pragma unroll
    for (int it = 0; it < F::IT; ++it)
#pragma unroll
      for (int t = 0; t < F::R0; ++t) {
        int y = F::elem(lane, it, t);
        if (ok && F::valid(lane, it)) {
          const size_t ad = ((size_t)z * NY + y) * g.nx + x;
          T1[ad] = v[it][t];
          if (T3) T3[ad] = w[it][t];
        }
      }
  }
}

template <typename R, int NY, int V>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_yb2_B(Geo g, int mode, int flags, const cplx<R>* __restrict__ G3, const cplx<R>* __restrict__ G4,
        const cplx<R>* __restrict__ A, const cplx<R>* __restrict__ Bn, cplx<R>* __restrict__ rhs0,
        cplx<R>* __restrict__ Bout, cplx<R>* __restrict__ T1, cplx<R>* __restrict__ T3, Etd etd, double alpha,
        int* __restrict__ flag, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  yb2_B_block<R, NY, V>(g, xcd_block(), lds, mode, flags, G3, G4, A, Bn, rhs0, Bout, T1, T3, etd, alpha, flag, tw);
}

// The q update (k_yb2_q) and the B update (k_yb2_B) of one ETD stage in one
// launch: logical blocks x < nq of each level run the q pass's tile x, the
// others the B pass's tile x - nq.  The two read and write disjoint arrays,
// so their blocks run side by side, as two launches on two streams would,
// without the streams' synchronization.  Both share the tile (variant V) and
// the staged twiddle and wavenumber tables.
template <typename R, int NY, int V, int FL = -1>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_yb2_qB(Geo g, int nq, int mode, int qflags, const cplx<R>* __restrict__ G1, const cplx<R>* __restrict__ G2,
         const cplx<R>* __restrict__ qd, const cplx<R>* __restrict__ qn, cplx<R>* __restrict__ rq0,
         cplx<R>* __restrict__ qout, double nu, double inv_dz2, int bflags, const cplx<R>* __restrict__ G3,
         const cplx<R>* __restrict__ G4, const cplx<R>* __restrict__ A, const cplx<R>* __restrict__ Bn,
         cplx<R>* __restrict__ rB0, cplx<R>* __restrict__ Bout, cplx<R>* __restrict__ T1,
         cplx<R>* __restrict__ T3, double alpha, Etd etd, int* __restrict__ flag,
         const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  Blk b = xcd_block();
  if (b.x < nq) {
    yb2_q_block<R, NY, V, FL>(g, b, lds, mode, qflags, G1, G2, qd, qn, rq0, qout, etd, nu, inv_dz2, flag, tw);
  } else {
    b.x -= nq;
    yb2_B_block<R, NY, V>(g, b, lds, mode, bflags, G3, G4, A, Bn, rB0, Bout, T1, T3, etd, alpha, flag, tw);
  }
}

// ===========================================================================
// Finalize: velocities and the QG omega equation (operators.jl:195-274, 365-493;
// diagnostics.jl:149-229)
// ===========================================================================

// Y-pass (inverse, half width) of the four omega-RHS factors
// (_omega_eqn_rhs_direct!, src/diagnostics.jl:149-229): psi_z (forward
// difference to the level above, zero at the top) times i*kx and i*ky, and
// the two-level average psi-bar (psi itself at the top) times -i*kx*kh2 and
// -i*ky*kh2.
// UV = 1: the same pass also makes the y-transformed half rows of u and v
// (what k_yb_adv writes for OUT_U | OUT_V) from the psi it holds.
// CHK = 1: also the y-transformed half rows of psi itself, Ppsi (the
// termination check's max |ifft2(psi)|, folded in by the X-pass), and psi is
// tested for non-finite values at every point (flag).
// CHK = 2: psi's non-finite test as for CHK = 1, and instead of Ppsi a bound
// of the check's maximum: lsum[z] += the block's share of
// sum over the level's whole spectrum of |h.re| + |h.im|, h = psi's Hermitian
// part / (nx ny), which bounds |Re ifft2(psi)| at every point of the level
// (the X-pass folds the levels' sums into absmax).  The half-width columns
// stand for their mirrors too (x = 0 and x = nx/2 are their own).
//
// Every output is the Hermitian part of a multiplier times psi, psi_z or
// psi-bar at (x, y) and its mirror (-x, -y).  Each multiplier is odd in the
// one wavenumber it carries (kx or ky; kh2 is even), so the output is the
// multiplier times the Hermitian part of the field, except on the column or
// row that is its own mirror (kx or ky at the Nyquist index), where it is the
// multiplier times the anti-Hermitian part.  The pass keeps only the two
// parts' sources, psi(z) and psi(z+1) (the difference and the average follow
// by linearity), as Hermitian parts, plus their anti-Hermitian parts at the
// row ny/2 (one register slot of one lane, for even ny), and forms, transforms
// and stores one output at a time.  The kx outputs of the Nyquist column
// (x = nx/2) come from one extra tile column of blocks (blk.x == the number
// of tiles), whose sources are the anti-Hermitian parts of that column.
template <typename R, int NY, int V, int UV, int CHK>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS), (sizeof(R) == 4 && V != 3 ? 4 : 1))
k_yf_omega(Geo g, const cplx<R>* __restrict__ psi, cplx<R>* __restrict__ Pbx, cplx<R>* __restrict__ Pby,
           cplx<R>* __restrict__ Pxx, cplx<R>* __restrict__ Pxy, cplx<R>* __restrict__ Pu,
           cplx<R>* __restrict__ Pv, cplx<R>* __restrict__ Ppsi, int* __restrict__ flag,
           double* __restrict__ lsum, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, +1, TL::PL>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const Blk blk = xcd_block();
  // (pruned passes carry no Nyquist column and launch no Nyquist blocks)
  const int ntile = (g.Kx + TL::S - 1) / TL::S;
  const bool nyq = blk.x == ntile;  // the Nyquist column's kx outputs (uniform)
  const int x = nyq ? g.nx / 2 : blk.x * TL::S + s, z = blk.y;
  const bool ok = nyq ? s == 0 : x < g.Kx;
  const int mx = mirror(x, g.nx);
  const bool xnyq = mx == x && x != 0;  // the Nyquist column: its own mirror, kx != 0
  cplx<R>* sl = lds + s * TL::LD;
  const double kxv = ok ? g.kx[x] : 0.0;
  const bool top = z % g.nzc == g.nzc - 1;  // top level of this member's column
  const int zu = top ? z : z + 1;
  const R rdz = (R)g.dz, rinv = (R)g.inv_n;
  // psi(z), psi(zu) / (nx ny): Hermitian parts (anti-Hermitian in the
  // Nyquist blocks), and the anti-Hermitian parts at the row ny/2
  cplx<R> h0[F::IT][F::R0], h1[F::IT][F::R0];
  cplx<R> a0 = czero<R>(), a1 = czero<R>();
  bool fin = true;
  double bsum = 0.0;  // CHK = 2: this thread's share of the level's bound
  const double wx = (x == 0 || xnyq) ? 1.0 : 2.0;
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      const int y = F::elem(lane, it, t), my = mirror(y, NY);
      const bool in = ok && F::valid(lane, it);
      // outside the disk (pruned) psi is zero: not loaded
      const bool d = g.in_disk(x, y), ld = in && d;
      const cplx<R> p = keep_if(ldu(psi, ((size_t)z * NY + y) * g.nx + x, ld), d);
      const cplx<R> pm = keep_if(ldu(psi, ((size_t)z * NY + my) * g.nx + mx, ld), d);
      const cplx<R> up = keep_if(ldu(psi, ((size_t)zu * NY + y) * g.nx + x, ld), d);
      const cplx<R> upm = keep_if(ldu(psi, ((size_t)zu * NY + my) * g.nx + mx, ld), d);
      fin = fin && (!in || (finite_c(p) && finite_c(pm)));
      h0[it][t] = nyq ? aherm(p, pm, rinv) : herm(p, pm, rinv);
      h1[it][t] = nyq ? aherm(up, upm, rinv) : herm(up, upm, rinv);
      if (CHK == 2 && in) bsum += wx * ((double)fabs(h0[it][t].re) + (double)fabs(h0[it][t].im));
      if (2 * y == NY) {  // the row ny/2 (even ny), held by one lane
        a0 = aherm(p, pm, rinv);
        a1 = aherm(up, upm, rinv);
      }
    }
  if (CHK && !fin) atomicOr(flag, 1);
  // the regular blocks cover every column once (the Nyquist blocks repeat x = nx/2)
  if (CHK == 2 && !nyq) fold_block_sum<TL::THREADS>(bsum, lsum + z);  // uniform
  // psi_z = (psi(zu) - psi(z)) / dz (exactly 0 at the top, where zu = z) and
  // psi-bar = (psi(zu) + psi(z)) / 2 (psi itself at the top); the division
  // is a multiplication by 1/dz (a division is a full sequence per value)
  const R idz = (R)1 / rdz;
  auto dzf = [&](cplx<R> u1, cplx<R> u0) { return mk<R>((u1.re - u0.re) * idz, (u1.im - u0.im) * idz); };
  auto avg = [&](cplx<R> u1, cplx<R> u0) { return scale(u1 + u0, (R)0.5); };
  const double* __restrict__ gky = stage_k<NY>(g.ky);
  // outputs: 0 i kx psi_z, 1 i ky psi_z, 2 -i kx kh2 psi-bar, 3 -i ky kh2 psi-bar,
  // 4 -i ky psi (u), 5 i kx psi (v), 6 psi
  cplx<R> v[F::IT][F::R0];
#pragma unroll
  for (int o = 0; o < 7; ++o) {
    const bool kxo = o == 0 || o == 2 || o == 5;  // multiplier odd in kx
    if ((o == 4 || o == 5) && !UV) continue;
    if (o == 6 && CHK != 1) continue;
    if (nyq && !kxo) continue;  // uniform
    // ky and kh2 loaded again for every output (kept, they would hold 32
    // registers through every transform)
    lds_cdouble* __restrict__ ky_o = opaque_lds(gky);
#pragma unroll
    for (int it = 0; it < F::IT; ++it)
#pragma unroll
      for (int t = 0; t < F::R0; ++t) {
        const int y = F::elem(lane, it, t);
        const double ky = ky_o[y];
        const double kh2 = kxv * kxv + ky * ky;
        cplx<R> f1 = h1[it][t], f0 = h0[it][t];
        if (!kxo && o != 6 && 2 * y == NY) {  // ky at the row ny/2
          f1 = a1;
          f0 = a0;
        }
        cplx<R> r;
        switch (o) {
          case 0: r = ik_mul((R)kxv, dzf(f1, f0)); break;
          case 1: r = ik_mul((R)ky, dzf(f1, f0)); break;
          case 2: r = ik_mul((R)(-(kxv * kh2)), avg(f1, f0)); break;
          case 3: r = ik_mul((R)(-(ky * kh2)), avg(f1, f0)); break;
          case 4: r = ik_mul((R)-ky, f0); break;
          case 5: r = ik_mul((R)kxv, f0); break;
          default: r = f0; break;
        }
        v[it][t] = r;
      }
    cplx<R>* dst = o == 0 ? Pbx : o == 1 ? Pby : o == 2 ? Pxx : o == 3 ? Pxy : o == 4 ? Pu : o == 5 ? Pv : Ppsi;
    F::run(v, sl, lane, twr);
    // the Nyquist column's kx outputs are stored by the Nyquist blocks
    const bool st = ok && !(kxo && xnyq && !nyq);
#pragma unroll
    for (int it = 0; it < F::IT; ++it)
#pragma unroll
      for (int t = 0; t < F::R0; ++t) {
        const int y = F::elem(lane, it, t);
        if (st && F::valid(lane, it)) dst[((size_t)z * NY + y) * g.Wp + x] = v[it][t];
      }
  }
}

// The termination check's max |ifft2(psi)| over one row: the inverse x
// transform of the row's half spectrum (psi's Hermitian part, y-transformed
// by k_yf_omega or k_y_check: ifft2 of a real field's spectrum), folded into
// absmax.  Shared by k_xf and k_x_psi_max so that the fused and the stand-alone
// check compute the same values.  Every thread of the block calls it.
template <typename R, int NX, int THREADS, class FI, class TW>
__device__ inline void psi_row_max(cplx<R> (&v)[FI::IT][FI::R0], const cplx<R>* __restrict__ Ppsi, size_t hb,
                                   int Kx, bool ok, int lane, cplx<R>* sl, const TW& twr,
                                   unsigned long long* __restrict__ absmax) {
  load_herm_row<R, FI>(v, Ppsi, hb, NX, Kx, lane, ok);
  FI::run(v, sl, lane, twr);
  double m = 0.0;
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      if (ok && FI::valid(lane, it)) {
        const double a = fabs((double)v[it][t].re);
        m = (a > m || a != a) ? a : m;
      }
    }
  fold_block_max<THREADS>(m, absmax);
}

// X-pass: u, v out; omega RHS 2(bx*xy - by*xx) -> forward real FFT -> Rw (half).
// The inverse transforms go in pairs of real fields (u, v), (bx, by), (xx, xy),
// each pair packed into one complex transform.
// CHK = 1 and Ppsi set: also the inverse x transform of psi's half rows (from
// k_yf_omega), whose largest |value| is folded into absmax (psi_row_max).
// lsum set (k_yf_omega's CHK = 2 sums over nlev levels): the first block
// folds their largest into absmax and clears them for the next step.
// FP64 runs the plain pass as the CHK = 0 instantiation (the check's code
// path cost it a fifth of its speed); FP32 runs CHK = 1 either way (the
// form that measured fastest).
template <typename R, int NX, int V, int CHK>
__global__ void __launch_bounds__((Tile<R, NX, V>::THREADS))
k_xf(Geo g, int want_omega, const cplx<R>* __restrict__ Pu, const cplx<R>* __restrict__ Pv,
     const cplx<R>* __restrict__ Pbx, const cplx<R>* __restrict__ Pby, const cplx<R>* __restrict__ Pxx,
     const cplx<R>* __restrict__ Pxy, R* __restrict__ u, R* __restrict__ vv, cplx<R>* __restrict__ Rw,
     const cplx<R>* __restrict__ Ppsi, unsigned long long* __restrict__ absmax,
     double* __restrict__ lsum, int nlev, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NX, V>;
  using FI = Fft<R, NX, +1, TL::PL, TL::XWAVE>;
  using FF = Fft<R, NX, -1, TL::PL, TL::XWAVE>;
  if (lsum && blockIdx.x == 0) {  // uniform
    double m = 0.0;
    for (int i = threadIdx.x; i < nlev; i += TL::THREADS) {
      const double a = lsum[i];
      m = (a > m || a != a) ? a : m;
      lsum[i] = 0.0;
    }
    fold_block_max<TL::THREADS>(m, absmax);
    __syncthreads();  // its block array is free again (psi_row_max uses it too)
  }
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int r = threadIdx.x / TL::T, lane = threadIdx.x % TL::T;
  TwTab<R, TL::NN, TL::PL, sizeof(R) == 8> twr;  // FP64: register twiddles (measured faster)
  twr.load(lane, tw);
  const int row = blockIdx.x * TL::S + r;
  const bool ok = row < g.nz * g.ny;
  cplx<R>* sl = lds + r * TL::LD;
  const size_t hb = (size_t)row * g.Wp;
  cplx<R> v[FI::IT][FI::R0], gb[FI::IT][FI::R0], gx[FI::IT][FI::R0];
  // all three packed rows are loaded before the first transform
  load_herm_pair<R, FI>(v, Pu, Pv, hb, NX, g.Kx, lane, ok);
  if (want_omega) {
    load_herm_pair<R, FI>(gb, Pbx, Pby, hb, NX, g.Kx, lane, ok);
    load_herm_pair<R, FI>(gx, Pxx, Pxy, hb, NX, g.Kx, lane, ok);
  }
  // u and v: one packed transform
  FI::run(v, sl, lane, twr);
#pragma unroll
  for (int it = 0; it < FI::IT; ++it)
#pragma unroll
    for (int t = 0; t < FI::R0; ++t) {
      int x = FI::elem(lane, it, t);
      if (ok && FI::valid(lane, it)) {
        u[(size_t)row * NX + x] = v[it][t].re;
        vv[(size_t)row * NX + x] = v[it][t].im;
      }
    }
  if (want_omega) {  // uniform
    // (bx, by) and (xx, xy): one packed transform each
    FI::run(gb, sl, lane, twr);  // .re = bx, .im = by
    FI::run(gx, sl, lane, twr);  // .re = xx, .im = xy
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t)
        v[it][t] = mk<R>((R)2 * (gb[it][t].re * gx[it][t].im - gb[it][t].im * gx[it][t].re), (R)0);
    FF::run(v, sl, lane, twr);
#pragma unroll
    for (int it = 0; it < FI::IT; ++it)
#pragma unroll
      for (int t = 0; t < FI::R0; ++t) {
        int x = FI::elem(lane, it, t);
        if (ok && FI::valid(lane, it) && x < g.Kx) Rw[hb + x] = v[it][t];
      }
  }
  if constexpr (CHK) {
    if (Ppsi) psi_row_max<R, NX, TL::THREADS, FI>(v, Ppsi, hb, g.Kx, ok, lane, sl, twr, absmax);  // uniform
  }
}

// Y-pass (forward, half width) of a real field, masked, expanded to a full
// Hermitian spectrum in double precision.
template <typename R, int NY, int V>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_y_r2c_mask(Geo g, const cplx<R>* __restrict__ Rw, cplx<R>* __restrict__ out,
             const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, -1, TL::PL>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const Blk blk = xcd_block();
  const int x = blk.x * TL::S + s, z = blk.y;
  const bool ok = x < g.Kx;
  cplx<R>* sl = lds + s * TL::LD;
  cplx<R> v[F::IT][F::R0];
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      v[it][t] = ldu(Rw, ((size_t)z * NY + y) * g.Wp + x, ok && F::valid(lane, it));
    }
  F::run(v, sl, lane, twr);
  const int mx = mirror(x, g.nx);
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      if (!(ok && F::valid(lane, it))) continue;
      const bool keep = g.keep(x, y);
      if (g.pruned() && !keep) continue;  // pruned readers do not look there
      cplx<double> c = keep ? cvt<double>(v[it][t]) : czero<double>();
      out[((size_t)z * NY + y) * g.nx + x] = cvt<R>(c);
      if (mx != x) out[((size_t)z * NY + mirror(y, NY)) * g.nx + mx] = cvt<R>(conj(c));
    }
}

// Y-pass (inverse, half width): Hermitian part of a double spectrum.
template <typename R, int NY, int V>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_y_herm(Geo g, const cplx<R>* __restrict__ in, cplx<R>* __restrict__ P,
         const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, +1, TL::PL>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const Blk blk = xcd_block();
  const int x = blk.x * TL::S + s, z = blk.y;
  const bool ok = x < g.Kx;
  const int mx = mirror(x, g.nx);
  cplx<R>* sl = lds + s * TL::LD;
  cplx<R> v[F::IT][F::R0];
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      cplx<R> a = czero<R>(), am = czero<R>();
      if (ok && F::valid(lane, it) && g.in_disk(x, y)) {
        a = in[((size_t)z * NY + y) * g.nx + x];
        am = in[((size_t)z * NY + mirror(y, NY)) * g.nx + mx];
      }
      v[it][t] = herm(a, am, (R)g.inv_n);
    }
  F::run(v, sl, lane, twr);
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      if (ok && F::valid(lane, it)) P[((size_t)z * NY + y) * g.Wp + x] = v[it][t];
    }
}

// ===========================================================================
// Plain complex transforms (set!/diagnostics/check): double in and out.
// ===========================================================================

// Y-pass C2C of full double fields; inverse applies 1/(nx*ny).
template <typename R, int NY, int DIR, int V>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_y_c2c(Geo g, const cplx<double>* __restrict__ in, cplx<double>* __restrict__ out,
        const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, DIR, TL::PL>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const Blk blk = xcd_block();
  const int x = blk.x * TL::S + s, z = blk.y;
  const bool ok = x < g.nx;
  cplx<R>* sl = lds + s * TL::LD;
  cplx<R> v[F::IT][F::R0];
  const double sc = DIR > 0 ? g.inv_n : 1.0;
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      v[it][t] = (ok && F::valid(lane, it)) ? cvt<R>(scale(in[((size_t)z * NY + y) * g.nx + x], sc))
                                            : czero<R>();
    }
  F::run(v, sl, lane, twr);
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      if (ok && F::valid(lane, it)) out[((size_t)z * NY + y) * g.nx + x] = cvt<double>(v[it][t]);
    }
}

// X-pass C2C of full double fields.
template <typename R, int NX, int DIR, int V>
__global__ void __launch_bounds__((Tile<R, NX, V>::THREADS))
k_x_c2c(Geo g, const cplx<double>* __restrict__ in, cplx<double>* __restrict__ out,
        const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NX, V>;
  using F = Fft<R, NX, DIR, TL::PL, TL::XWAVE>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int r = threadIdx.x / TL::T, lane = threadIdx.x % TL::T;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const int row = blockIdx.x * TL::S + r;
  const bool ok = row < g.nz * g.ny;
  cplx<R>* sl = lds + r * TL::LD;
  cplx<R> v[F::IT][F::R0];
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int x = F::elem(lane, it, t);
      v[it][t] = (ok && F::valid(lane, it)) ? cvt<R>(in[(size_t)row * NX + x]) : czero<R>();
    }
  F::run(v, sl, lane, twr);
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int x = F::elem(lane, it, t);
      if (ok && F::valid(lane, it)) out[(size_t)row * NX + x] = cvt<double>(v[it][t]);
    }
}

// ===========================================================================
// Termination check (_check_termination_conditions!, src/simulation.jl:25-44)
// at the transform precision: any non-finite value of q, B or psi, and
// max |ifft2(psi)| (complex inverse, 1/(nx*ny)).  Both fold into accumulators
// (flag |= 1, absmax = max) that the caller clears, so a run of steps can
// check every step without a host round trip.
// ===========================================================================

// Termination check (_check_termination_conditions!, src/simulation.jl:25-44)
// on its own, Y-pass (half width): P = inverse y transform of psi's Hermitian
// part / (nx*ny), formed as k_yf_omega forms its psi output; q, B and psi are
// tested for non-finite values at the points it loads, (x, y) and the mirror
// (-x, -y), which together are every point of the levels it covers.
template <typename R, int NY, int V>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_y_check(Geo g, const cplx<R>* __restrict__ psi, const cplx<R>* __restrict__ q, const cplx<R>* __restrict__ B,
          cplx<R>* __restrict__ P, int* __restrict__ flag, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, +1, TL::PL>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const Blk blk = xcd_block();
  const int x = blk.x * TL::S + s, z = blk.y;
  const bool ok = x < g.W;
  const int mx = mirror(x, g.nx);
  cplx<R>* sl = lds + s * TL::LD;
  const R rinv = (R)g.inv_n;
  cplx<R> v[F::IT][F::R0];
  bool fin = true;
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      const int y = F::elem(lane, it, t), my = mirror(y, NY);
      const bool in = ok && F::valid(lane, it);
      const size_t ad = ((size_t)z * NY + y) * g.nx + x, adm = ((size_t)z * NY + my) * g.nx + mx;
      const cplx<R> p = ldu(psi, ad, in), pm = ldu(psi, adm, in);
      fin = fin && (!in || (finite_c(p) && finite_c(pm) && finite_c(q[ad]) && finite_c(q[adm]) &&
                            finite_c(B[ad]) && finite_c(B[adm])));
      v[it][t] = herm(p, pm, rinv);
    }
  if (!fin) atomicOr(flag, 1);
  F::run(v, sl, lane, twr);
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      const int y = F::elem(lane, it, t);
      if (ok && F::valid(lane, it)) P[((size_t)z * NY + y) * g.Wp + x] = v[it][t];
    }
}

// X-pass of the stand-alone check: max |ifft2(psi)| from k_y_check's half
// rows into absmax (nothing is stored).
template <typename R, int NX, int V>
__global__ void __launch_bounds__((Tile<R, NX, V>::THREADS))
k_x_psi_max(Geo g, const cplx<R>* __restrict__ P, unsigned long long* __restrict__ absmax,
            const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NX, V>;
  using FI = Fft<R, NX, +1, TL::PL, TL::XWAVE>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int r = threadIdx.x / TL::T, lane = threadIdx.x % TL::T;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const int row = blockIdx.x * TL::S + r;
  const bool ok = row < g.nz * g.ny;
  cplx<R>* sl = lds + r * TL::LD;
  cplx<R> v[FI::IT][FI::R0];
  psi_row_max<R, NX, TL::THREADS, FI>(v, P, (size_t)row * g.Wp, g.Kx, ok, lane, sl, twr, absmax);
}

// Progress maxima (_progress_maxima, src/simulation.jl:46-81), Y-pass of the
// wave velocity LA: the spectrum B + kh2 A / 4 (YBJ+; B itself otherwise),
// formed in FP64 from the stored fields, / (nx*ny), inverse y transform into
// `out`.  k_x_check then folds max |LA| over the grid.
template <typename R, int NY, int V>
__global__ void __launch_bounds__((Tile<R, NY, V>::THREADS))
k_y_la(Geo g, const cplx<R>* __restrict__ B, const cplx<R>* __restrict__ A, int ybj_plus,
       cplx<R>* __restrict__ out, const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NY, V>;
  using F = Fft<R, NY, +1, TL::PL>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int s = threadIdx.x % TL::S, lane = threadIdx.x / TL::S;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const Blk blk = xcd_block();
  const int x = blk.x * TL::S + s, z = blk.y;
  const bool ok = x < g.nx;
  cplx<R>* sl = lds + s * TL::LD;
  cplx<R> v[F::IT][F::R0];
  const double kxv = ok ? g.kx[x] : 0.0;
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      cplx<double> la = czero<double>();
      if (ok && F::valid(lane, it)) {
        const size_t ad = ((size_t)z * NY + y) * g.nx + x;
        la = cvt<double>(B[ad]);
        if (ybj_plus) {
          const double kh2 = kxv * kxv + g.ky[y] * g.ky[y];
          la = la + scale(cvt<double>(A[ad]), 0.25 * kh2);
        }
      }
      v[it][t] = cvt<R>(scale(la, g.inv_n));
    }
  F::run(v, sl, lane, twr);
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int y = F::elem(lane, it, t);
      if (ok && F::valid(lane, it)) out[((size_t)z * NY + y) * g.nx + x] = v[it][t];
    }
}

// X-pass: inverse x transform of the Y-pass output; folds max |value| into
// absmax (non-negative doubles compare like their bit patterns; NaN counts
// as +inf).  Nothing is stored.
template <typename R, int NX, int V>
__global__ void __launch_bounds__((Tile<R, NX, V>::THREADS))
k_x_check(Geo g, const cplx<R>* __restrict__ in, unsigned long long* __restrict__ absmax,
          const cplx<R>* __restrict__ tw) {
  using TL = Tile<R, NX, V>;
  using F = Fft<R, NX, +1, TL::PL, TL::XWAVE>;
  __shared__ cplx<R> lds[TL::S * TL::LD];
  const int r = threadIdx.x / TL::T, lane = threadIdx.x % TL::T;
  TwTab<R, TL::NN, TL::PL> twr;
  twr.load(lane, tw);
  const int row = blockIdx.x * TL::S + r;
  const bool ok = row < g.nz * g.ny;
  cplx<R>* sl = lds + r * TL::LD;
  cplx<R> v[F::IT][F::R0];
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      int x = F::elem(lane, it, t);
      v[it][t] = ldu(in, (size_t)row * NX + x, ok && F::valid(lane, it));
    }
  F::run(v, sl, lane, twr);
  double m = 0.0;
#pragma unroll
  for (int it = 0; it < F::IT; ++it)
#pragma unroll
    for (int t = 0; t < F::R0; ++t) {
      if (ok && F::valid(lane, it)) {
        const cplx<double> c = cvt<double>(v[it][t]);
        const double a = hypot(c.re, c.im);
        m = (a > m || a != a) ? a : m;
      }
    }
  fold_block_max<TL::THREADS>(m, absmax);
}

}  // namespace qg
