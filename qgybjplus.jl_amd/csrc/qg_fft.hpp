// LDS-staged mixed-radix Stockham FFT for one sequence held by a team of
// threads.  Written for CDNA4 wave64: a sequence of length N is processed by
// T lanes; the first stage reads from registers and the last stage leaves its
// result in registers with the SAME element distribution, so a kernel can run
// an inverse transform, do pointwise physics in registers, and feed a forward
// transform without another LDS round trip.
//
// Register slot v[it][t] of lane `lane` holds element  (lane + it*T) + t*(N/R0).
//
// Stage algorithm (decimation in time, self-sorting): with span P (product of
// the previous radices) and radix r, butterfly i in [0, N/r) reads
// x[i + t*N/r], multiplies by w_{rP}^{t*(i mod P)}, applies an r-point DFT and
// writes y[(i - i mod P)*r + i mod P + t*P].
//
// The stage twiddles are staged once per block in LDS (TwTab) and read by
// every forward and inverse transform of the kernel.
#pragma once

#include "qg_common.hpp"

namespace qg {

// cos(2*pi*m/240) for m in [0, 60], correctly rounded; the radices 2, 3, 4,
// 5, 6, 8, 10, 12, 15, 16 and 30 divide 240.
constexpr double kCos240[61] = {
    1.0, 0.9996573249755573, 0.9986295347545738, 0.996917333733128,
    0.9945218953682733, 0.9914448613738104, 0.9876883405951378, 0.9832549075639546,
    0.9781476007338057, 0.9723699203976766, 0.9659258262890683, 0.958819734868193,
    0.9510565162951535, 0.9426414910921784, 0.9335804264972017, 0.9238795325112867,
    0.9135454576426009, 0.9025852843498606, 0.8910065241883679, 0.8788171126619654,
    0.8660254037844386, 0.8526401643540922, 0.838670567945424, 0.8241261886220157,
    0.8090169943749475, 0.7933533402912352, 0.7771459614569709, 0.7604059656000309,
    0.7431448254773942, 0.7253743710122876, 0.7071067811865476, 0.688354575693754,
    0.6691306063588582, 0.6494480483301837, 0.6293203910498375, 0.6087614290087207,
    0.5877852522924731, 0.5664062369248328, 0.5446390350150271, 0.5224985647159489,
    0.5, 0.4771587602596084, 0.4539904997395468, 0.43051109680829514,
    0.4067366430758002, 0.3826834323650898, 0.35836794954530027, 0.3338068592337709,
    0.30901699437494745, 0.2840153447039226, 0.25881904510252074, 0.23344536385590542,
    0.20791169081775934, 0.18223552549214744, 0.15643446504023087, 0.1305261922200516,
    0.10452846326765347, 0.07845909572784494, 0.052335956242943835, 0.026176948307873153,
    0.0};

__host__ __device__ constexpr double cos240(int m) {
  m = ((m % 240) + 240) % 240;
  if (m <= 60) return kCos240[m];
  if (m <= 120) return -kCos240[120 - m];
  if (m <= 180) return -kCos240[m - 120];
  return kCos240[240 - m];
}
__host__ __device__ constexpr double sin240(int m) { return cos240(m - 60); }

// cos(2*pi*m/900) for m in [0, 225], correctly rounded (evaluated to 50
// digits); the roots of the radices 9 and 25, which do not divide 240.
constexpr double kCos900[226] = {
    1.0, 0.9999756307053947, 0.9999025240093042, 0.9997806834748455,
    0.9996101150403544, 0.9993908270190958, 0.9991228300988584, 0.9988061373414341,
    0.998440764181981, 0.9980267284282716, 0.9975640502598242, 0.9970527522269202,
    0.9964928592495044, 0.9958843986159703, 0.9952273999818312, 0.9945218953682733,
    0.9937679191605964, 0.9929655081065369, 0.9921147013144779, 0.9912155402515417,
    0.9902680687415704, 0.9892723329629883, 0.9882283814465528, 0.9871362650729879,
    0.985996037070505, 0.984807753012208, 0.9835714708133859, 0.9822872507286887,
    0.9809551553491915, 0.9795752495993441, 0.9781476007338057, 0.9766722783341679,
    0.9751493543055633, 0.9735789028731603, 0.9719610005785463, 0.9702957262759965,
    0.9685831611286311, 0.9668233886044594, 0.9650164944723114, 0.9631625667976581,
    0.9612616959383189, 0.9593139745400575, 0.9573194975320672, 0.9552783621223436,
    0.953190667792947, 0.9510565162951535, 0.9488760116444965, 0.9466492601156964,
    0.9443763702374811, 0.9420574527872967, 0.9396926207859084, 0.9372819894918915,
    0.9348256763960144, 0.9323238012155122, 0.9297764858882515, 0.9271838545667874,
    0.9245460336123131, 0.9218631515885005, 0.9191353392552344, 0.9163627295622396,
    0.9135454576426009, 0.9106836608061771, 0.9077774785329087, 0.9048270524660196,
    0.9018325264051138, 0.898794046299167, 0.8957117602394129, 0.8925858184521255,
    0.8894163732912975, 0.8862035792312147, 0.882947592858927, 0.8796485728666165,
    0.8763066800438636, 0.8729220772698096, 0.869494929505219, 0.8660254037844386,
    0.8625136692072575, 0.8589598969306644, 0.8553642601605066, 0.8517269341430477,
    0.848048096156426, 0.8443279255020151, 0.8405666034956842, 0.8367643134589616,
    0.8329212407100994, 0.8290375725550417, 0.8251134982782952, 0.821149209133704,
    0.8171448983351285, 0.8131007610470277, 0.8090169943749475, 0.8048937973559142,
    0.8007313709487335, 0.7965299180241963, 0.7922896433551907, 0.7880107536067219,
    0.7836934573258398, 0.7793379649314741, 0.7749444887041796, 0.7705132427757893,
    0.766044443118978, 0.7615383075367368, 0.7569950556517564, 0.7524149088957244,
    0.7477980904985319, 0.7431448254773942, 0.7384553406258838, 0.7337298645028764,
    0.7289686274214116, 0.7241718614374675, 0.7193398003386512, 0.7144726796328034,
    0.7095707365365209, 0.7046342099635947, 0.6996633405133654, 0.6946583704589973,
    0.6896195437356698, 0.6845471059286887, 0.6794413042615165, 0.6743023875837234,
    0.6691306063588582, 0.6639262126522416, 0.6586894601186805, 0.6534206039901055,
    0.6481199010631309, 0.6427876096865394, 0.6374239897486897, 0.6320293026648509,
    0.6266038113644605, 0.6211477802783104, 0.6156614753256583, 0.6101451639012676,
    0.6045991148623748, 0.5990235985155858, 0.5934188866037015, 0.5877852522924731,
    0.5821229701572894, 0.5764323161697933, 0.5707135676844317, 0.564967003424938,
    0.5591929034707468, 0.5533915492433441, 0.5475632234925503, 0.5417082102827399,
    0.5358267949789967, 0.5299192642332049, 0.5239859059700791, 0.5180270093731302,
    0.5120428648705715, 0.5060337641211639, 0.5, 0.49394186658423095,
    0.4878596591387327, 0.48175367410171527, 0.47562420907027525, 0.46947156278589075,
    0.4632960351198617, 0.4570979270586942, 0.45087754068943076, 0.4446351791849275,
    0.4383711467890774, 0.4320857488019823, 0.42577929156507266, 0.4194520824461771,
    0.41310442982454176, 0.4067366430758002, 0.40034903255689497, 0.3939419095909511,
    0.387515586452103, 0.3810703763502741, 0.374606593415912, 0.368124552684678,
    0.3616245700820923, 0.35510696240813705, 0.34857204732181524, 0.3420201433256687,
    0.33545156975025503, 0.32886664673858323, 0.3222656952305111, 0.31564903694710245,
    0.30901699437494745, 0.3023698907504445, 0.2957080500440467, 0.2890317969444716,
    0.2823414568428765, 0.27563735581699916, 0.2689198206152657, 0.2621891786408647,
    0.25544575793579055, 0.2486898871648548, 0.24192189559966773, 0.23514211310259003,
    0.22835087011065575, 0.2215484976194673, 0.2147353271670632, 0.20791169081775934,
    0.20107792114596468, 0.19423435121997198, 0.18738131458572463, 0.18051914525055998,
    0.17364817766693036, 0.16676874671610228, 0.15988118769183488, 0.15298583628403806,
    0.14608302856241162, 0.13917310096006544, 0.13225639025712244, 0.12533323356430426,
    0.11840396830650095, 0.11146893220632548, 0.10452846326765347, 0.09758289975914947,
    0.09063258019778016, 0.0836778433323155, 0.07671902812681865, 0.0697564737441253,
    0.06279051952931337, 0.0558215049931638, 0.048849769795613264, 0.04187565372919963,
    0.03489949670250097, 0.02792163872356888, 0.02094241988335696, 0.013962180339145272,
    0.0069812602979615525, 0.0,
};

__host__ __device__ constexpr double cos900(int m) {
  m = ((m % 900) + 900) % 900;
  if (m <= 225) return kCos900[m];
  if (m <= 450) return -kCos900[450 - m];
  if (m <= 675) return -kCos900[m - 450];
  return kCos900[900 - m];
}
__host__ __device__ constexpr double sin900(int m) { return cos900(m - 225); }

// exp(DIR * 2*pi*i * m / N) for N dividing 240 or 900.
template <typename R, int N, int DIR>
__device__ inline cplx<R> unit_root(int m) {
  if constexpr (240 % N == 0) {
    constexpr int f = 240 / N;
    return mk<R>((R)cos240(m * f), (R)(DIR * sin240(m * f)));
  } else {
    static_assert(900 % N == 0, "radix must divide 240 or 900");
    constexpr int f = 900 / N;
    return mk<R>((R)cos900(m * f), (R)(DIR * sin900(m * f)));
  }
}

// ---------------------------------------------------------------------------
// In-register DFTs.  DIR = -1 forward (exp(-i..)), +1 inverse (unnormalised).
// ---------------------------------------------------------------------------
template <int N, int DIR>
struct Dft;

template <int DIR>
struct Dft<1, DIR> {
  template <typename R>
  __device__ static inline void run(cplx<R>*) {}
};

template <int DIR>
struct Dft<2, DIR> {
  template <typename R>
  __device__ static inline void run(cplx<R>* v) {
    cplx<R> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <int DIR>
struct Dft<3, DIR> {
  template <typename R>
  __device__ static inline void run(cplx<R>* v) {
    const R h = (R)0.8660254037844386;  // sin(2*pi/3)
    cplx<R> a = v[0], b = v[1], c = v[2];
    cplx<R> s = b + c;
    cplx<R> d = b - c;
    cplx<R> m = mk<R>(a.re - (R)0.5 * s.re, a.im - (R)0.5 * s.im);
    cplx<R> r = mk<R>(-(R)DIR * h * d.im, (R)DIR * h * d.re);  // DIR*i*h*d
    v[0] = a + s;
    v[1] = m + r;
    v[2] = m - r;
  }
};

// 5-point DFT: with s1 = a1 + a4, d1 = a1 - a4, s2 = a2 + a3, d2 = a2 - a3,
// X1,4 = a0 + c1 s1 + c2 s2 +- DIR i (n1 d1 + n2 d2) and
// X2,3 = a0 + c2 s1 + c1 s2 +- DIR i (n2 d1 - n1 d2), where c1 = cos(2 pi/5),
// c2 = cos(4 pi/5), n1 = sin(2 pi/5), n2 = sin(4 pi/5).
template <int DIR>
struct Dft<5, DIR> {
  template <typename R>
  __device__ static inline void run(cplx<R>* v) {
    const R c1 = (R)0.30901699437494745, c2 = (R)-0.8090169943749475;
    const R n1 = (R)0.9510565162951535, n2 = (R)0.5877852522924731;
    const cplx<R> a0 = v[0];
    const cplx<R> s1 = v[1] + v[4], d1 = v[1] - v[4], s2 = v[2] + v[3], d2 = v[2] - v[3];
    const cplx<R> m1 = mk<R>(a0.re + c1 * s1.re + c2 * s2.re, a0.im + c1 * s1.im + c2 * s2.im);
    const cplx<R> m2 = mk<R>(a0.re + c2 * s1.re + c1 * s2.re, a0.im + c2 * s1.im + c1 * s2.im);
    const cplx<R> e1 = mk<R>(n1 * d1.re + n2 * d2.re, n1 * d1.im + n2 * d2.im);
    const cplx<R> e2 = mk<R>(n2 * d1.re - n1 * d2.re, n2 * d1.im - n1 * d2.im);
    const cplx<R> r1 = mk<R>(-(R)DIR * e1.im, (R)DIR * e1.re);  // DIR*i*e1
    const cplx<R> r2 = mk<R>(-(R)DIR * e2.im, (R)DIR * e2.re);  // DIR*i*e2
    v[0] = a0 + s1 + s2;
    v[1] = m1 + r1;
    v[4] = m1 - r1;
    v[2] = m2 + r2;
    v[3] = m2 - r2;
  }
};

template <int DIR>
struct Dft<4, DIR> {
  template <typename R>
  __device__ static inline void run(cplx<R>* v) {
    cplx<R> a = v[0], b = v[1], c = v[2], d = v[3];
    cplx<R> s0 = a + c, d0 = a - c, s1 = b + d, d1 = b - d;
    cplx<R> r = mk<R>(-(R)DIR * d1.im, (R)DIR * d1.re);  // DIR*i*d1
    v[0] = s0 + s1;
    v[2] = s0 - s1;
    v[1] = d0 + r;
    v[3] = d0 - r;
  }
};

// Composite DFT of size N1*N2 (Cooley-Tukey, decimation in time):
// input n = N2*n1 + n2, output k = k1 + N1*k2.
template <int N1, int N2, int DIR>
struct DftCT {
  template <typename R>
  __device__ static inline void run(cplx<R>* v) {
    constexpr int N = N1 * N2;
    cplx<R> y[N2][N1];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) y[n2][n1] = v[N2 * n1 + n2];
      Dft<N1, DIR>::run(y[n2]);
#pragma unroll
      for (int k1 = 1; k1 < N1; ++k1) {
        if (n2 != 0) y[n2][k1] = cmul(y[n2][k1], unit_root<R, N, DIR>(n2 * k1));
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      cplx<R> z[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) z[n2] = y[n2][k1];
      Dft<N2, DIR>::run(z);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) v[k1 + N1 * k2] = z[k2];
    }
  }
};

template <int DIR> struct Dft<6, DIR> : DftCT<2, 3, DIR> {};
template <int DIR> struct Dft<10, DIR> : DftCT<2, 5, DIR> {};
template <int DIR> struct Dft<8, DIR> : DftCT<2, 4, DIR> {};
template <int DIR> struct Dft<12, DIR> : DftCT<4, 3, DIR> {};
template <int DIR> struct Dft<16, DIR> : DftCT<4, 4, DIR> {};
template <int DIR> struct Dft<15, DIR> : DftCT<3, 5, DIR> {};
template <int DIR> struct Dft<30, DIR> : DftCT<2, 15, DIR> {};
template <int DIR> struct Dft<9, DIR> : DftCT<3, 3, DIR> {};
template <int DIR> struct Dft<25, DIR> : DftCT<5, 5, DIR> {};

// ---------------------------------------------------------------------------
// Plans: radices per stage (first == last so the register distribution of the
// input and of the output coincide) and lanes per sequence T.
// ---------------------------------------------------------------------------
template <int N, int PL = 0> struct FftPlan;
template <> struct FftPlan<1>    { static constexpr int S = 1, r0 = 1, r1 = 1, T = 1; };
template <> struct FftPlan<2>    { static constexpr int S = 1, r0 = 2, r1 = 1, T = 1; };
template <> struct FftPlan<3>    { static constexpr int S = 1, r0 = 3, r1 = 1, T = 1; };
template <> struct FftPlan<4>    { static constexpr int S = 1, r0 = 4, r1 = 1, T = 1; };
template <> struct FftPlan<5>    { static constexpr int S = 1, r0 = 5, r1 = 1, T = 1; };
template <> struct FftPlan<6>    { static constexpr int S = 1, r0 = 6, r1 = 1, T = 1; };
template <> struct FftPlan<8>    { static constexpr int S = 1, r0 = 8, r1 = 1, T = 1; };
template <> struct FftPlan<9>    { static constexpr int S = 2, r0 = 3, r1 = 1, T = 3; };
template <> struct FftPlan<18>   { static constexpr int S = 3, r0 = 3, r1 = 2, T = 6; };
template <> struct FftPlan<25>   { static constexpr int S = 2, r0 = 5, r1 = 1, T = 5; };
template <> struct FftPlan<27>   { static constexpr int S = 3, r0 = 3, r1 = 3, T = 9; };
template <> struct FftPlan<36>   { static constexpr int S = 2, r0 = 6, r1 = 1, T = 6; };
template <> struct FftPlan<45>   { static constexpr int S = 3, r0 = 3, r1 = 5, T = 15; };
template <> struct FftPlan<50>   { static constexpr int S = 3, r0 = 5, r1 = 2, T = 10; };
template <> struct FftPlan<54>   { static constexpr int S = 3, r0 = 3, r1 = 6, T = 18; };
template <> struct FftPlan<72>   { static constexpr int S = 3, r0 = 6, r1 = 2, T = 12; };
template <> struct FftPlan<75>   { static constexpr int S = 3, r0 = 5, r1 = 3, T = 15; };
template <> struct FftPlan<90>   { static constexpr int S = 3, r0 = 3, r1 = 10, T = 30; };
template <> struct FftPlan<100>  { static constexpr int S = 2, r0 = 10, r1 = 1, T = 10; };
template <> struct FftPlan<108>  { static constexpr int S = 3, r0 = 6, r1 = 3, T = 18; };
template <> struct FftPlan<125>  { static constexpr int S = 3, r0 = 5, r1 = 5, T = 25; };
template <> struct FftPlan<144>  { static constexpr int S = 2, r0 = 12, r1 = 1, T = 12; };
template <> struct FftPlan<150>  { static constexpr int S = 3, r0 = 5, r1 = 6, T = 30; };
template <> struct FftPlan<180>  { static constexpr int S = 3, r0 = 6, r1 = 5, T = 30; };
template <> struct FftPlan<200>  { static constexpr int S = 3, r0 = 10, r1 = 2, T = 20; };
template <> struct FftPlan<216>  { static constexpr int S = 3, r0 = 6, r1 = 6, T = 36; };
template <> struct FftPlan<250>  { static constexpr int S = 3, r0 = 5, r1 = 10, T = 50; };
template <> struct FftPlan<288>  { static constexpr int S = 3, r0 = 12, r1 = 2, T = 24; };
template <> struct FftPlan<300>  { static constexpr int S = 3, r0 = 10, r1 = 3, T = 30; };
template <> struct FftPlan<360>  { static constexpr int S = 3, r0 = 6, r1 = 10, T = 60; };
template <> struct FftPlan<400>  { static constexpr int S = 3, r0 = 10, r1 = 4, T = 40; };
template <> struct FftPlan<432>  { static constexpr int S = 3, r0 = 12, r1 = 3, T = 36; };
template <> struct FftPlan<500>  { static constexpr int S = 3, r0 = 10, r1 = 5, T = 50; };
template <> struct FftPlan<576>  { static constexpr int S = 3, r0 = 12, r1 = 4, T = 48; };
template <> struct FftPlan<600>  { static constexpr int S = 3, r0 = 10, r1 = 6, T = 60; };
template <> struct FftPlan<720>  { static constexpr int S = 3, r0 = 12, r1 = 5, T = 60; };
template <> struct FftPlan<800>  { static constexpr int S = 3, r0 = 10, r1 = 8, T = 80; };
template <> struct FftPlan<864>  { static constexpr int S = 3, r0 = 12, r1 = 6, T = 72; };
template <> struct FftPlan<1000> { static constexpr int S = 3, r0 = 10, r1 = 10, T = 100; };
// lengths whose plan needs a radix of 15 or 30 (a composite in-register DFT)
template <> struct FftPlan<15>   { static constexpr int S = 1, r0 = 15, r1 = 1, T = 1; };
template <> struct FftPlan<30>   { static constexpr int S = 1, r0 = 30, r1 = 1, T = 1; };
template <> struct FftPlan<60>   { static constexpr int S = 3, r0 = 2, r1 = 15, T = 30; };
template <> struct FftPlan<120>  { static constexpr int S = 3, r0 = 2, r1 = 30, T = 60; };
template <> struct FftPlan<135>  { static constexpr int S = 3, r0 = 3, r1 = 15, T = 45; };
template <> struct FftPlan<225>  { static constexpr int S = 2, r0 = 15, r1 = 1, T = 15; };
template <> struct FftPlan<240>  { static constexpr int S = 3, r0 = 4, r1 = 15, T = 60; };
template <> struct FftPlan<375>  { static constexpr int S = 3, r0 = 5, r1 = 15, T = 75; };
template <> struct FftPlan<480>  { static constexpr int S = 3, r0 = 4, r1 = 30, T = 60; };
template <> struct FftPlan<540>  { static constexpr int S = 3, r0 = 6, r1 = 15, T = 90; };
template <> struct FftPlan<750>  { static constexpr int S = 3, r0 = 5, r1 = 30, T = 75; };
template <> struct FftPlan<960>  { static constexpr int S = 3, r0 = 8, r1 = 15, T = 120; };
// with a radix-9 or radix-25 stage
template <> struct FftPlan<81>   { static constexpr int S = 2, r0 = 9, r1 = 1, T = 9; };
template <> struct FftPlan<162>  { static constexpr int S = 3, r0 = 9, r1 = 2, T = 18; };
template <> struct FftPlan<243>  { static constexpr int S = 3, r0 = 9, r1 = 3, T = 27; };
template <> struct FftPlan<270>  { static constexpr int S = 3, r0 = 3, r1 = 30, T = 45; };
template <> struct FftPlan<324>  { static constexpr int S = 3, r0 = 9, r1 = 4, T = 36; };
template <> struct FftPlan<405>  { static constexpr int S = 3, r0 = 9, r1 = 5, T = 45; };
template <> struct FftPlan<450>  { static constexpr int S = 3, r0 = 15, r1 = 2, T = 30; };
template <> struct FftPlan<486>  { static constexpr int S = 3, r0 = 9, r1 = 6, T = 54; };
template <> struct FftPlan<625>  { static constexpr int S = 3, r0 = 5, r1 = 25, T = 125; };
template <> struct FftPlan<900>  { static constexpr int S = 3, r0 = 10, r1 = 9, T = 90; };
template <> struct FftPlan<648>  { static constexpr int S = 3, r0 = 9, r1 = 8, T = 72; };
template <> struct FftPlan<675>  { static constexpr int S = 3, r0 = 15, r1 = 3, T = 45; };
template <> struct FftPlan<729>  { static constexpr int S = 3, r0 = 9, r1 = 9, T = 81; };
template <> struct FftPlan<810>  { static constexpr int S = 3, r0 = 9, r1 = 10, T = 90; };
template <> struct FftPlan<972>  { static constexpr int S = 3, r0 = 9, r1 = 12, T = 108; };
template <> struct FftPlan<10>   { static constexpr int S = 1, r0 = 10, r1 = 1, T = 1; };
template <> struct FftPlan<12>   { static constexpr int S = 1, r0 = 12, r1 = 1, T = 1; };
template <> struct FftPlan<16>   { static constexpr int S = 2, r0 = 4, r1 = 1, T = 4; };
template <> struct FftPlan<20>   { static constexpr int S = 3, r0 = 2, r1 = 5, T = 5; };
template <> struct FftPlan<24>   { static constexpr int S = 3, r0 = 2, r1 = 6, T = 3; };
template <> struct FftPlan<32>   { static constexpr int S = 3, r0 = 4, r1 = 2, T = 8; };
template <> struct FftPlan<40>   { static constexpr int S = 3, r0 = 2, r1 = 10, T = 5; };
template <> struct FftPlan<48>   { static constexpr int S = 3, r0 = 4, r1 = 3, T = 6; };
template <> struct FftPlan<64>   { static constexpr int S = 2, r0 = 8, r1 = 1, T = 8; };
template <> struct FftPlan<80>   { static constexpr int S = 3, r0 = 4, r1 = 5, T = 10; };
template <> struct FftPlan<96>   { static constexpr int S = 3, r0 = 4, r1 = 6, T = 12; };
template <> struct FftPlan<128>  { static constexpr int S = 3, r0 = 4, r1 = 8, T = 16; };
template <> struct FftPlan<160>  { static constexpr int S = 3, r0 = 4, r1 = 10, T = 20; };
template <> struct FftPlan<192>  { static constexpr int S = 3, r0 = 8, r1 = 3, T = 24; };
template <> struct FftPlan<256>  { static constexpr int S = 3, r0 = 8, r1 = 4, T = 32; };
template <> struct FftPlan<320>  { static constexpr int S = 3, r0 = 8, r1 = 5, T = 40; };
template <> struct FftPlan<384>  { static constexpr int S = 3, r0 = 8, r1 = 6, T = 48; };
template <> struct FftPlan<512>  { static constexpr int S = 3, r0 = 8, r1 = 8, T = 64; };
template <> struct FftPlan<640>  { static constexpr int S = 3, r0 = 8, r1 = 10, T = 80; };
template <> struct FftPlan<768>  { static constexpr int S = 3, r0 = 8, r1 = 12, T = 96; };
template <> struct FftPlan<1024> { static constexpr int S = 3, r0 = 16, r1 = 4, T = 64; };
// Alternative plans (PL = 1): fewer lanes per sequence and more elements per
// lane, so a 256-thread block spans twice the columns (128-byte row segments
// of FP32 complex values in the Y-passes).  Where none is defined the
// alternative is the base plan.
template <int N> struct FftPlan<N, 1> : FftPlan<N, 0> {};
template <> struct FftPlan<256, 1> { static constexpr int S = 2, r0 = 16, r1 = 1, T = 16; };

// Stage twiddles, forward sign (the inverse uses the conjugates).  By
// default they are staged once per block in LDS from the length-N table, in
// the layout the stages read: the middle stage's w^(t k step) as [t-1][k],
// k < R0, and the last stage's w^(t i) as [t-1][i], i < N/R0, so the lanes of
// a stage read consecutive words (no bank conflicts) and no lane keeps
// twiddles in registers.  Every thread of the block calls load() at the start
// of the kernel; the first barrier of the first transform orders the fill
// before any read.  REG = true keeps each lane's own stage twiddles in
// registers instead (the passes that measure faster so).
template <typename R, int N, int PL = 0, bool REG = false>
struct TwTab {
  using P = FftPlan<N, PL>;
  static constexpr int R0 = P::r0, R1 = P::r1, T = P::T;
  static constexpr int NB = N / R0;
  static constexpr int MID = (P::S == 3) ? (R1 - 1) * R0 : 0;
  static constexpr int LAST = (P::S >= 2) ? (R0 - 1) * NB : 0;
  static constexpr int SIZE = MID + LAST;
  static constexpr int STEP = (P::S == 3) ? N / (R1 * R0) : 1;
  // register form: per lane, [it][t-1] of the stages it runs
  static constexpr int IT = (NB + T - 1) / T;
  static constexpr int NB1 = N / R1;
  static constexpr int IT1 = (NB1 + T - 1) / T;
  static constexpr int NMID = (REG && P::S == 3) ? IT1 * (R1 - 1) : 1;
  static constexpr int NLAST = (REG && P::S >= 2) ? IT * (R0 - 1) : 1;
  const cplx<R>* tab;
  cplx<R> rmid[NMID];
  cplx<R> rlast[NLAST];

  // tw: length-N table, tw[m] = exp(-2*pi*i*m/N)
  __device__ inline void load(int lane, const cplx<R>* __restrict__ tw) {
    if constexpr (REG) {
      if constexpr (P::S == 3) {
#pragma unroll
        for (int it = 0; it < IT1; ++it) {
          const int i = lane + it * T;
#pragma unroll
          for (int t = 1; t < R1; ++t)
            rmid[it * (R1 - 1) + t - 1] = (NB1 % T == 0 || i < NB1) ? tw[t * (i % R0) * STEP] : mk<R>((R)1, (R)0);
        }
      }
      if constexpr (P::S >= 2) {
#pragma unroll
        for (int it = 0; it < IT; ++it) {
          const int i = lane + it * T;
#pragma unroll
          for (int t = 1; t < R0; ++t)
            rlast[it * (R0 - 1) + t - 1] = (NB % T == 0 || i < NB) ? tw[t * i] : mk<R>((R)1, (R)0);
        }
      }
      tab = nullptr;
    } else {
      __shared__ cplx<R> s[SIZE > 0 ? SIZE : 1];
      for (int e = threadIdx.x; e < SIZE; e += blockDim.x) {
        int m;
        if (e < MID) {
          m = (e / R0 + 1) * (e % R0) * STEP;
        } else {
          const int f = e - MID;
          m = (f / NB + 1) * (f % NB);
        }
        s[e] = tw[m];
      }
      __syncthreads();  // the table is shared by every sequence of the block
      tab = s;
    }
  }
  // middle stage: w^(t k step), 1 <= t < R1, k = i mod R0 of the lane's
  // butterfly i = lane + it T
  __device__ inline cplx<R> mid(int it, int t, int k) const {
    if constexpr (REG) return rmid[it * (R1 - 1) + t - 1];
    else return tab[(t - 1) * R0 + k];
  }
  // last stage: w^(t i), 1 <= t < R0, i = lane + it T < N/R0
  __device__ inline cplx<R> last(int it, int t, int i) const {
    if constexpr (REG) return rlast[it * (R0 - 1) + t - 1];
    else return tab[MID + (t - 1) * NB + i];
  }
  // LDS bytes of the staged table
  static constexpr int bytes() { return (SIZE > 0 ? SIZE : 1) * (int)sizeof(cplx<R>); }
};

template <typename R, int DIR>
__device__ inline cplx<R> twmul(cplx<R> v, cplx<R> w) {
  if (DIR > 0) w.im = -w.im;
  return cmul(v, w);
}

// WAVE: the T lanes of every sequence lie in one wave (a team of T
// consecutive threads, T dividing 64), so the stages' LDS exchanges need only
// the wave's own order (a fence and a wave barrier) instead of a block
// barrier, and the block's waves run through their transforms independently.
template <typename R, int N, int DIR, int PL = 0, bool WAVE = false>
struct Fft {
  using P = FftPlan<N, PL>;
  static constexpr int R0 = P::r0;
  static constexpr int R1 = P::r1;
  static constexpr int T = P::T;
  static constexpr int NB = N / R0;
  static constexpr int IT = (NB + T - 1) / T;
  static constexpr int NE = IT * R0;  // register elements per lane
  // LDS index of element `idx`: one pad slot every R0 elements keeps the
  // stride-R0 first-stage writes of neighbouring lanes in distinct banks.
  static constexpr int LDSZ = N + N / R0;
  __device__ static inline int pad(int idx) { return idx + idx / R0; }

  // Element index held in register slot [it][t].
  __device__ static inline int elem(int lane, int it, int t) { return lane + it * T + t * NB; }
  __device__ static inline bool valid(int lane, int it) { return (NB % T == 0) || (lane + it * T < NB); }

  static_assert(!WAVE || 64 % P::T == 0, "a sequence of a wave-synchronized transform must lie in one wave");
  // Orders this sequence's LDS accesses before and after it: a block barrier,
  // or with WAVE the wave's fence and barrier.
  __device__ static inline void sync() {
    if constexpr (WAVE) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      __syncthreads();
    }
  }

  // s: this sequence's LDS (LDSZ elements).  Contains barriers: every thread
  // of the block (with WAVE: of the wave) must call it the same number of
  // times.
  template <class TW>
  __device__ static inline void run(cplx<R> (&v)[IT][R0], cplx<R>* s, int lane, const TW& w) {
    if constexpr (P::S == 1) {
#pragma unroll
      for (int it = 0; it < IT; ++it)
        if (valid(lane, it)) Dft<R0, DIR>::run(v[it]);
    } else {
      // first stage: registers -> LDS
      sync();
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        if (valid(lane, it)) {
          int i = lane + it * T;
          Dft<R0, DIR>::run(v[it]);
#pragma unroll
          for (int t = 0; t < R0; ++t) s[pad(i * R0 + t)] = v[it][t];
        }
      }
      sync();
      if constexpr (P::S == 3) {
        constexpr int NB1 = N / R1;
        constexpr int IT1 = (NB1 + T - 1) / T;
        cplx<R> u[IT1][R1];
        cplx<R> wm[IT1][R1];  // the stage's twiddles, read with the data (before the barrier)
#pragma unroll
        for (int it = 0; it < IT1; ++it) {
          int i = lane + it * T;
          if (NB1 % T == 0 || i < NB1) {
#pragma unroll
            for (int t = 0; t < R1; ++t) u[it][t] = s[pad(i + t * NB1)];
#pragma unroll
            for (int t = 1; t < R1; ++t) wm[it][t] = w.mid(it, t, i % R0);
          }
        }
        sync();
#pragma unroll
        for (int it = 0; it < IT1; ++it) {
          int i = lane + it * T;
          if (NB1 % T == 0 || i < NB1) {
            int k = i % R0;
#pragma unroll
            for (int t = 1; t < R1; ++t) u[it][t] = twmul<R, DIR>(u[it][t], wm[it][t]);
            Dft<R1, DIR>::run(u[it]);
            int j = (i - k) * R1 + k;
#pragma unroll
            for (int t = 0; t < R1; ++t) s[pad(j + t * R0)] = u[it][t];
          }
        }
        sync();
      }
      // last stage: LDS -> registers, span NB = N/R0, k = i
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        if (valid(lane, it)) {
          int i = lane + it * T;
#pragma unroll
          for (int t = 0; t < R0; ++t) v[it][t] = s[pad(i + t * NB)];
#pragma unroll
          for (int t = 1; t < R0; ++t) v[it][t] = twmul<R, DIR>(v[it][t], w.last(it, t, i));
          Dft<R0, DIR>::run(v[it]);
        }
      }
    }
  }
};

}  // namespace qg
