// em_fft.hpp -- in-register / LDS Stockham FFT engine for gfx950.
//
// One FFT line of length N is owned by T threads; each thread holds E = N/T
// complex values in registers in the "strided" distribution: thread tau holds
// element tau + T*e in v[e].  A line is transformed by a sequence of Stockham
// radix-R stages (R divides E, product of the radices = N).  Stage inputs are
// read in the strided distribution and stage outputs are scattered to their
// Stockham positions; the first stage reads straight from registers and the
// last stage leaves its outputs in registers in the strided distribution
// again, so a line makes (stages - 1) round trips through LDS and none through
// HBM.  Because input and output share one distribution, an inverse FFT, a
// pointwise multiply and a forward FFT chain in registers with no reordering.
//
// Directions: DIR = -1 forward exp(-2 pi i nk/N), DIR = +1 inverse, both
// unnormalised, matching the C2C convention the reference relies on
// (fft_dev.cuh:159-190).
#pragma once
#include <hip/hip_runtime.h>

namespace em {

// --------------------------------------------------------------------------
// complex helpers
// --------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float cnorm(float2 a) { return a.x * a.x + a.y * a.y; }
// multiply by DIR*i
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}
// multiply by exp(DIR * 2 pi i m / n) for a twiddle from the forward table
template <int DIR>
__device__ __forceinline__ float2 twmul(float2 a, float2 w_fwd) {
  return DIR < 0 ? cmul(a, w_fwd) : cmulc(a, w_fwd);
}

// --------------------------------------------------------------------------
// compile-time twiddle constants exp(DIR 2 pi i M / N)
// --------------------------------------------------------------------------
constexpr double kPiD = 3.141592653589793238462643383279502884;
constexpr double cx_sin(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 24; ++i) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 24; ++i) {
    term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
template <int N, int M, int DIR>
struct CTw {
  static constexpr int m = ((M % N) + N) % N;
  // reduce the angle to [-pi, pi]
  static constexpr double ang = DIR * 2.0 * kPiD * (2 * m <= N ? (double)m / N : (double)(m - N) / N);
  static constexpr float re = (float)cx_cos(ang);
  static constexpr float im = (float)cx_sin(ang);
};

template <int N, int M, int DIR>
__device__ __forceinline__ float2 ctw(float2 a) {
  constexpr int m = ((M % N) + N) % N;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (4 * m == N) {  // exp(DIR i pi/2) = DIR*i
    return mul_i<DIR>(a);
  } else if constexpr (2 * m == N) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (4 * m == 3 * N) {
    return mul_i<-DIR>(a);
  } else {
    constexpr float c = CTw<N, M, DIR>::re, s = CTw<N, M, DIR>::im;
    return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// --------------------------------------------------------------------------
// small DFTs on compile-time register slots v[B + n*S], natural order in/out
// --------------------------------------------------------------------------
template <int R, int DIR, int B, int S, int E>
struct Dft;

template <int DIR, int B, int S, int E>
struct Dft<2, DIR, B, S, E> {
  __device__ __forceinline__ static void run(float2 (&v)[E]) {
    const float2 a = v[B], b = v[B + S];
    v[B] = cadd(a, b);
    v[B + S] = csub(a, b);
  }
};

template <int DIR, int B, int S, int E>
struct Dft<3, DIR, B, S, E> {
  __device__ __forceinline__ static void run(float2 (&v)[E]) {
    constexpr float h = 0.86602540378443864676f;  // sqrt(3)/2
    const float2 a0 = v[B], a1 = v[B + S], a2 = v[B + 2 * S];
    const float2 s = cadd(a1, a2), d = csub(a1, a2);
    const float2 m = make_float2(a0.x - 0.5f * s.x, a0.y - 0.5f * s.y);
    const float2 r = cscale(mul_i<DIR>(d), h);
    v[B] = cadd(a0, s);
    v[B + S] = cadd(m, r);
    v[B + 2 * S] = csub(m, r);
  }
};

template <int DIR, int B, int S, int E>
struct Dft<4, DIR, B, S, E> {
  __device__ __forceinline__ static void run(float2 (&v)[E]) {
    const float2 a0 = v[B], a1 = v[B + S], a2 = v[B + 2 * S], a3 = v[B + 3 * S];
    const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    const float2 t2 = cadd(a1, a3), t3 = mul_i<DIR>(csub(a1, a3));
    v[B] = cadd(t0, t2);
    v[B + 2 * S] = csub(t0, t2);
    v[B + S] = cadd(t1, t3);
    v[B + 3 * S] = csub(t1, t3);
  }
};

// A compile-time twiddle w = exp(DIR 2 pi i M / N) in the factored form
// w = g (1 + i tau) (g = cos, tau = tan of its angle), so that w x = g u with
// u = x + i tau x: two FMAs (two adds when tau = +-1), and the real g can
// ride on the FMAs of the butterfly that consumes w x.  Multiples of pi/2
// keep g = 1 and make u a swap / negation.
template <int N, int M, int DIR>
struct TanTw {
  static constexpr int m = ((M % N) + N) % N;
  // 0: w = 1, 1: w = DIR i, 2: w = -1, 3: w = -DIR i, 4: general
  static constexpr int kind = m == 0 ? 0 : (4 * m == N ? 1 : (2 * m == N ? 2 : (4 * m == 3 * N ? 3 : 4)));
  static constexpr double ang = 2.0 * kPiD * (double)m / N;
  static constexpr double c = cx_cos(ang);
  static constexpr double s = DIR * cx_sin(ang);
  static constexpr double g = kind == 4 ? c : 1.0;
  // odd multiples of pi/4: tau = +-1 exactly
  static constexpr bool unit = kind == 4 && (8 * m) % N == 0 && ((8 * m) / N) % 2 == 1;
  static constexpr float tau = kind == 4 ? (float)(s / c) : 0.f;
  __device__ __forceinline__ static float2 u(float2 x) {
    if constexpr (kind == 0) {
      return x;
    } else if constexpr (kind == 1) {
      return mul_i<DIR>(x);
    } else if constexpr (kind == 2) {
      return make_float2(-x.x, -x.y);
    } else if constexpr (kind == 3) {
      return mul_i<-DIR>(x);
    } else if constexpr (unit) {
      return tau > 0.f ? make_float2(x.x - x.y, x.y + x.x) : make_float2(x.x + x.y, x.y - x.x);
    } else {
      return make_float2(__builtin_fmaf(-tau, x.y, x.x), __builtin_fmaf(tau, x.x, x.y));
    }
  }
};

// a + k b and a - k b for a compile-time real k (adds when k = +-1)
template <int K_IS>  // 1: k = 1, -1: k = -1, 0: general
__device__ __forceinline__ void fma_pm(float2 a, float2 b, float k, float2& p, float2& q) {
  if constexpr (K_IS == 1) {
    p = cadd(a, b);
    q = csub(a, b);
  } else if constexpr (K_IS == -1) {
    p = csub(a, b);
    q = cadd(a, b);
  } else {
    p = make_float2(__builtin_fmaf(k, b.x, a.x), __builtin_fmaf(k, b.y, a.y));
    q = make_float2(__builtin_fmaf(-k, b.x, a.x), __builtin_fmaf(-k, b.y, a.y));
  }
}
constexpr int unit_sign(double k) { return k == 1.0 ? 1 : (k == -1.0 ? -1 : 0); }

// Radix-4 butterfly over x_n = v[B + n S] (n = 0..3) after the constant
// twiddles w_n = exp(DIR 2 pi i n K / N), outputs in natural order in place.
// With w_n x_n = g_n u_n:  t0,1 = u0 +- g2 u2;  q+- = u1 +- (g3/g1) u3;
// y0,2 = t0 +- g1 q+;  y1,3 = t1 +- g1 (DIR i q-).  Every twiddle other than
// a multiple of pi/2 costs 2 instructions (u) instead of a complex multiply's
// 4; the g's are FMA operands of adds the butterfly makes anyway.
template <int N, int K, int DIR, int B, int S, int E>
struct Radix4Tw {
  __device__ __forceinline__ static void run(float2 (&v)[E]) {
    using W1 = TanTw<N, K, DIR>;
    using W2 = TanTw<N, 2 * K, DIR>;
    using W3 = TanTw<N, 3 * K, DIR>;
    const float2 u0 = v[B], u1 = W1::u(v[B + S]), u2 = W2::u(v[B + 2 * S]), u3 = W3::u(v[B + 3 * S]);
    float2 t0, t1, qp, qm;
    fma_pm<unit_sign(W2::g)>(u0, u2, (float)W2::g, t0, t1);
    constexpr double r = W3::g / W1::g;
    fma_pm<unit_sign(r)>(u1, u3, (float)r, qp, qm);
    const float2 qmi = mul_i<DIR>(qm);
    float2 y0, y1, y2, y3;
    fma_pm<unit_sign(W1::g)>(t0, qp, (float)W1::g, y0, y2);
    fma_pm<unit_sign(W1::g)>(t1, qmi, (float)W1::g, y1, y3);
    v[B] = y0;
    v[B + S] = y1;
    v[B + 2 * S] = y2;
    v[B + 3 * S] = y3;
  }
};

// --------------------------------------------------------------------------
// Per-lane twiddles in the same factored form.  A twiddled radix-4 butterfly
// with w_r = g_r (1 + i tau_r), tau_r = DIR tan(theta_r) for the forward
// angle theta_r (T_r = tan(theta_r) is stored): u_r = x_r + i tau_r x_r, then
//   t0,1 = x0 +- g2 u2;  q+- = u1 +- (g3/g1) u3;
//   y0,2 = t0 +- g1 q+;  y1,3 = t1 +- g1 (DIR i q-):
// 22 instructions against 28 (three complex multiplies and 16 adds).  The
// tables are built in FP64, where cos(pi/2) = 6.1e-17 is not zero: the
// twiddle -i then reads g = 6.1e-17, T = 1.6e16, whose product is 1 to FP32
// rounding, and u carries 1.6e16 |x| (no overflow for |x| < 1e22).
// --------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ float2 tan_u(float2 x, float T) {
  const float t = DIR > 0 ? T : -T;
  return make_float2(__builtin_fmaf(-t, x.y, x.x), __builtin_fmaf(t, x.x, x.y));
}

// Radix-4 of (x0, g1 u1, g2 u2, g3 u3) in natural order into x0..x3, given
// u1..u3 and rho = g3 / g1.
template <int DIR>
__device__ __forceinline__ void radix4_scaled(float2& x0, float2& x1, float2& x2, float2& x3, float2 u1, float2 u2,
                                              float2 u3, float g1, float g2, float rho) {
  float2 t0, t1, qp, qm;
  fma_pm<0>(x0, u2, g2, t0, t1);
  fma_pm<0>(u1, u3, rho, qp, qm);
  const float2 qmi = mul_i<DIR>(qm);
  fma_pm<0>(t0, qp, g1, x0, x2);
  fma_pm<0>(t1, qmi, g1, x1, x3);
}

// A twiddled radix-4 butterfly from one table entry: c0 = (T1, T2, T3, g1),
// c1 = (g2, g3 / g1).
template <int DIR>
__device__ __forceinline__ void radix4_lane(float2& x0, float2& x1, float2& x2, float2& x3, float4 c0, float2 c1) {
  const float2 u1 = tan_u<DIR>(x1, c0.x), u2 = tan_u<DIR>(x2, c0.y), u3 = tan_u<DIR>(x3, c0.z);
  radix4_scaled<DIR>(x0, x1, x2, x3, u1, u2, u3, c0.w, c1.x, c1.y);
}

// Composite R = R1 * 4 by the four-step split n = 4 n1 + n2, k = k1 + R1 k2:
// R1-point DFTs over n1 for each n2, then per k1 the radix-4 over n2 with the
// twiddles W_R^(n2 k1) folded into its FMAs (Radix4Tw).
template <int R1, int DIR, int B, int S, int E>
struct DftComposite {
  static constexpr int R2 = 4;
  static constexpr int R = R1 * R2;
  template <int N2>
  __device__ __forceinline__ static void step1(float2 (&v)[E]) {
    if constexpr (N2 < R2) {
      Dft<R1, DIR, B + N2 * S, R2 * S, E>::run(v);
      step1<N2 + 1>(v);
    }
  }
  template <int K1>
  __device__ __forceinline__ static void step23(float2 (&v)[E]) {
    if constexpr (K1 < R1) {
      Radix4Tw<R, K1, DIR, B + K1 * R2 * S, S, E>::run(v);
      step23<K1 + 1>(v);
    }
  }
  __device__ __forceinline__ static void run(float2 (&v)[E]) {
    step1<0>(v);
    step23<0>(v);
    // X[k1 + R1 k2] sits at slot R2 k1 + k2: reorder to natural order
    float2 tmp[R];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1)
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) tmp[k1 + R1 * k2] = v[B + (R2 * k1 + k2) * S];
#pragma unroll
    for (int k = 0; k < R; ++k) v[B + k * S] = tmp[k];
  }
};

template <int DIR, int B, int S, int E>
struct Dft<8, DIR, B, S, E> {
  __device__ __forceinline__ static void run(float2 (&v)[E]) { DftComposite<2, DIR, B, S, E>::run(v); }
};
template <int DIR, int B, int S, int E>
struct Dft<12, DIR, B, S, E> {
  __device__ __forceinline__ static void run(float2 (&v)[E]) { DftComposite<3, DIR, B, S, E>::run(v); }
};
template <int DIR, int B, int S, int E>
struct Dft<16, DIR, B, S, E> {
  __device__ __forceinline__ static void run(float2 (&v)[E]) { DftComposite<4, DIR, B, S, E>::run(v); }
};

// --------------------------------------------------------------------------
// Twiddled radix-16 butterflies on slots v[B + r S] (r = 0..15), input r
// carrying the lane's twiddle w_r = exp(DIR i theta r) of its stage, computed
// as the 4 x 4 composite (step 1 over r = n2 + 4 m for each n2, then the
// radix-4s over n2 of the constant twiddles W16^(n2 k1)).
// --------------------------------------------------------------------------
// X[k1 + 4 k2] sits at slot 4 k1 + k2: to natural order
template <int B, int S, int E>
__device__ __forceinline__ void reorder16(float2 (&v)[E]) {
  float2 tmp[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) tmp[k1 + 4 * k2] = v[B + (4 * k1 + k2) * S];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[B + k * S] = tmp[k];
}

// Step 2+3 of radix16_lane: the step-1 outputs X'(n2) are the true outputs
// divided by g_(n2); with W16^(n2 k1) = G (1 + i tau) the radix-4 over n2
// sees scales s_(n2) = g_(n2) G, s_0 = 1: per k1 the lane's S2 = s_2,
// sig = s_3 / s_1 and S1 = s_1.
template <int DIR, int B, int S, int E, int K1>
__device__ __forceinline__ void radix16_lane_step23(float2 (&v)[E], const float (&S2)[4], const float (&sig)[4],
                                                    const float (&S1)[4]) {
  if constexpr (K1 < 4) {
    constexpr int b = B + 4 * K1 * S;
    const float2 u1 = TanTw<16, K1, DIR>::u(v[b + S]);
    const float2 u2 = TanTw<16, 2 * K1, DIR>::u(v[b + 2 * S]);
    const float2 u3 = TanTw<16, 3 * K1, DIR>::u(v[b + 3 * S]);
    radix4_scaled<DIR>(v[b], v[b + S], v[b + 2 * S], v[b + 3 * S], u1, u2, u3, S1[K1], S2[K1], sig[K1]);
    radix16_lane_step23<DIR, B, S, E, K1 + 1>(v, S2, sig, S1);
  }
}

// Every twiddle per lane, from a table entry of 10 float4 (40 floats):
// T_1..T_15 and a pad; per n2: r02 = g_(n2+8)/g_(n2), r13 = g_(n2+12)/g_(n2+4),
// kap = g_(n2+4)/g_(n2); per k1: S2, sig, S1 (above).  Step 1 for sub-
// butterfly n2 is radix4_scaled on u_(n2+4m) with (g1, g2, rho) = (kap, r02,
// r13); its outputs are short of the factor g_(n2) (g_0 = 1), which step 3
// carries.  174 instructions against 204 with the twiddles multiplied in.
template <int DIR, int B, int S, int E>
__device__ __forceinline__ void radix16_lane(float2 (&v)[E], const float4* __restrict__ c) {
  const float4 Ta = c[0], Tb = c[1], Tc = c[2], Td = c[3];
  const float T[16] = {0.f, Ta.x, Ta.y, Ta.z, Ta.w, Tb.x, Tb.y, Tb.z, Tb.w, Tc.x, Tc.y, Tc.z, Tc.w, Td.x, Td.y, Td.z};
  const float4 f02 = c[4], f13 = c[5], fk = c[6], fS2 = c[7], fsg = c[8], fS1 = c[9];
  const float r02[4] = {f02.x, f02.y, f02.z, f02.w}, r13[4] = {f13.x, f13.y, f13.z, f13.w};
  const float kap[4] = {fk.x, fk.y, fk.z, fk.w};
  const float S2[4] = {fS2.x, fS2.y, fS2.z, fS2.w}, sig[4] = {fsg.x, fsg.y, fsg.z, fsg.w};
  const float S1[4] = {fS1.x, fS1.y, fS1.z, fS1.w};
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    const float2 u1 = tan_u<DIR>(v[B + (n2 + 4) * S], T[n2 + 4]);
    const float2 u2 = tan_u<DIR>(v[B + (n2 + 8) * S], T[n2 + 8]);
    const float2 u3 = tan_u<DIR>(v[B + (n2 + 12) * S], T[n2 + 12]);
    if (n2 > 0) v[B + n2 * S] = tan_u<DIR>(v[B + n2 * S], T[n2]);
    radix4_scaled<DIR>(v[B + n2 * S], v[B + (n2 + 4) * S], v[B + (n2 + 8) * S], v[B + (n2 + 12) * S], u1, u2, u3,
                       kap[n2], r02[n2], r13[n2]);
  }
  radix16_lane_step23<DIR, B, S, E, 0>(v, S2, sig, S1);
  reorder16<B, S, E>(v);
}

// Register form for the column passes: the stage's twiddle of input
// r = n2 + 4 m factors as lo_(n2) hi_m with lo = W^(n2 k) (complex, multiplied
// in) and hi = W^(4 m k), whose radix-4 over m takes it in the factored form
// (h0 = (T1, T2, T3, g1), h1 = (g2, g3/g1) of W^(4k)).  Step 3 as Dft<16>.
template <int K1, int DIR, int B, int S, int E>
__device__ __forceinline__ void radix16_step23(float2 (&v)[E]) {
  if constexpr (K1 < 4) {
    Radix4Tw<16, K1, DIR, B + 4 * K1 * S, S, E>::run(v);
    radix16_step23<K1 + 1, DIR, B, S, E>(v);
  }
}
template <int DIR, int B, int S, int E>
__device__ __forceinline__ void radix16_lohi(float2 (&v)[E], const float2 (&lo)[3], float4 h0, float2 h1) {
#pragma unroll
  for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
    for (int m = 0; m < 4; ++m) v[B + (n2 + 4 * m) * S] = twmul<DIR>(v[B + (n2 + 4 * m) * S], lo[n2 - 1]);
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2)
    radix4_lane<DIR>(v[B + n2 * S], v[B + (n2 + 4) * S], v[B + (n2 + 8) * S], v[B + (n2 + 12) * S], h0, h1);
  radix16_step23<0, DIR, B, S, E>(v);
  reorder16<B, S, E>(v);
}

// --------------------------------------------------------------------------
// LDS layout of one line: one pad slot every 16 elements keeps the scattered
// Stockham writes of radix-16 stages off a single bank.
// --------------------------------------------------------------------------
template <int N>
struct LineLds {
  static constexpr int size = N + N / 16;
  __device__ __forceinline__ static int idx(int i) { return i + (i >> 4); }
};

// --------------------------------------------------------------------------
// Per-thread twiddles of one Stockham stage (R, NS): W_{NS R}^{r k} with
// k = j mod NS for the butterflies j = tau + T u of this thread.
//  mode 1 (T multiple of NS):  k = tau mod NS for every u;
//  mode 2 (NS multiple of T):  k = tau + T (u mod NS/T), and the u part
//                              W^{r T u} is a compile-time constant;
//  mode 0 (otherwise):         read from the table at every use.
// For modes 1 and 2 the factors are W^{r k} = W^{r1 k} W^{4 r2 k}
// (r = r1 + 4 r2, r1 < 4): 3 + (R-1)/4 table values per thread, read once
// from tw[m] = exp(-2 pi i m / N) (FP64-generated) and kept in registers.
// --------------------------------------------------------------------------
template <int N, int T, int NS, int R>
struct StageTw {
  static constexpr int NR = NS * R;
  static constexpr int step = N / NR;
  static constexpr int mode = (NS <= 1) ? 3 : ((T % NS == 0) ? 1 : ((NS % T == 0) ? 2 : 0));
  static constexpr int NLO = (R - 1 < 3) ? (R - 1) : 3;
  static constexpr int NHI = (R - 1) / 4;
  float2 lo[NLO > 0 ? NLO : 1];
  float2 hi[NHI > 0 ? NHI : 1];
  int k0;

  __device__ __forceinline__ void init(int tau, const float2* __restrict__ tw) {
    if constexpr (mode == 1 || mode == 2) {
      const int k = (mode == 1) ? (tau % NS) : tau;
#pragma unroll
      for (int m = 1; m <= NLO; ++m) lo[m - 1] = tw[(m * k * step) % N];
#pragma unroll
      for (int m = 1; m <= NHI; ++m) hi[m - 1] = tw[(4 * m * k * step) % N];
    }
    k0 = tau;
  }

  // v *= W^{r k} for input r of butterfly u (compile-time r, u)
  template <int DIR, int U, int RR>
  __device__ __forceinline__ float2 apply(float2 v, const float2* __restrict__ tw) const {
    if constexpr (mode == 0) {
      const int k = (k0 + T * U) % NS;
      return twmul<DIR>(v, tw[(RR * k * step) % N]);
    } else {
      constexpr int r1 = RR % 4, r2 = RR / 4;
      if constexpr (r1 > 0) v = twmul<DIR>(v, lo[r1 - 1]);
      if constexpr (r2 > 0) v = twmul<DIR>(v, hi[r2 - 1]);
      if constexpr (mode == 2) v = ctw<NR, RR * T * (U % (NS / T)), DIR>(v);
      return v;
    }
  }
};

template <int N, int T, int NS, int... RS>
struct TwChain {
  static constexpr bool in_registers = true;
  __device__ __forceinline__ void init(int, const float2*) {}
};
template <int N, int T, int NS, int R, int... RS>
struct TwChain<N, T, NS, R, RS...> {
  // false when some stage reads the global table at every use (mode 0)
  static constexpr bool in_registers =
      StageTw<N, T, NS, R>::mode != 0 && TwChain<N, T, NS * R, RS...>::in_registers;
  StageTw<N, T, NS, R> s;
  TwChain<N, T, NS * R, RS...> next;
  __device__ __forceinline__ void init(int tau, const float2* __restrict__ tw) {
    s.init(tau, tw);
    next.init(tau, tw);
  }
};

// Twiddles of one stage read from a per-workgroup LDS table instead of being
// formed in registers: entry (r - 1) * NS + k holds W_N^{r k step}, where k is
// the butterfly index mod NS.  Lane tau's k is tau mod NS (T a multiple of
// NS) or tau + T (U mod NS/T) (NS a multiple of T), so every twiddle is one
// LDS read at a compile-time offset from the lane's base and one complex
// multiply; lanes read consecutive (or broadcast) entries.  Stages whose k
// follows neither pattern (mode 0, e.g. NS = 12 or 48 with T = 64) keep the
// lane's k of each of its E/R butterflies, computed once at init.
template <int N, int T, int NS, int R>
struct StageTwTab {
  static constexpr int step = N / (NS * R);
  static constexpr int mode = (NS <= 1) ? 3 : ((T % NS == 0) ? 1 : ((NS % T == 0) ? 2 : 0));
  static constexpr int entries = (NS <= 1) ? 0 : (R - 1) * NS;
  static constexpr int NB = N / T / R;  // butterflies per lane
  const float2* base;
  int kb[mode == 0 ? NB : 1];

  __device__ __forceinline__ void init(int tau, const float2* tab) {
    if constexpr (mode == 0) {
      base = tab;
#pragma unroll
      for (int u = 0; u < NB; ++u) kb[u] = (tau + T * u) % NS;
    } else {
      base = tab + ((mode == 1) ? (tau % NS) : tau);
    }
  }
  template <int DIR, int U, int RR>
  __device__ __forceinline__ float2 apply(float2 v, const float2* __restrict__) const {
    if constexpr (mode == 0) {
      return twmul<DIR>(v, base[(RR - 1) * NS + kb[U]]);
    } else {
      constexpr int off = (RR - 1) * NS + ((mode == 2) ? T * (U % (NS / T)) : 0);
      return twmul<DIR>(v, base[off]);
    }
  }
  __device__ __forceinline__ static void fill(float2* tab, const float2* __restrict__ tw, int tid, int nt) {
    for (int i = tid; i < entries; i += nt) {
      const int r = i / NS + 1, k = i % NS;
      tab[i] = tw[(r * k * step) % N];
    }
  }
};

template <int N, int T, int NS, int... RS>
struct TwChainTab {
  static constexpr int entries = 0;
  static constexpr bool ok = true;
  __device__ __forceinline__ void init(int, const float2*) {}
  __device__ __forceinline__ static void fill(float2*, const float2*, int, int) {}
};
template <int N, int T, int NS, int R, int... RS>
struct TwChainTab<N, T, NS, R, RS...> {
  using St = StageTwTab<N, T, NS, R>;
  using Next = TwChainTab<N, T, NS * R, RS...>;
  static constexpr int entries = St::entries + Next::entries;
  static constexpr bool ok = true;
  St s;
  Next next;
  __device__ __forceinline__ void init(int tau, const float2* tab) {
    s.init(tau, tab);
    next.init(tau, tab + St::entries);
  }
  __device__ __forceinline__ static void fill(float2* tab, const float2* __restrict__ tw, int tid, int nt) {
    St::fill(tab, tw, tid, nt);
    Next::fill(tab + St::entries, tw, tid, nt);
  }
};

// Synchronisation between the LDS write and read phases of an exchange.
// SYNC_BLOCK: lines span several wavefronts (workgroup barrier).
// SYNC_WAVE:  every line lives inside one wavefront; LDS requests of one wave
//             complete in order, so only compiler reordering must be fenced.
enum { SYNC_BLOCK = 0, SYNC_WAVE = 1 };
template <int SYNC>
__device__ __forceinline__ void lds_sync() {
  if constexpr (SYNC == SYNC_BLOCK) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// --------------------------------------------------------------------------
// Stockham stages.  Stage (R, NS): butterfly j in [0, N/R), thread tau owns
// j = tau + T*u, u in [0, E/R); its inputs x[j + r N/R] sit in v[u + r E/R];
// twiddle W_{NS R}^{r (j mod NS)}; outputs go to (j/NS) NS R + (j mod NS) + r NS.
// --------------------------------------------------------------------------
template <int N, int T, int DIR, int SYNC>
struct FftLine {
  static constexpr int E = N / T;

  template <int R, int U>
  __device__ __forceinline__ static void butterflies(float2 (&v)[E]) {
    if constexpr (U < E / R) {
      Dft<R, DIR, U, E / R, E>::run(v);
      butterflies<R, U + 1>(v);
    }
  }

  template <int R, int NS, int U, int RR, class St>
  __device__ __forceinline__ static void twiddle_one(float2 (&v)[E], const St& st,
                                                     const float2* __restrict__ tw) {
    if constexpr (U < E / R) {
      if constexpr (RR < R) {
        v[U + RR * (E / R)] = st.template apply<DIR, U, RR>(v[U + RR * (E / R)], tw);
        twiddle_one<R, NS, U, RR + 1>(v, st, tw);
      } else {
        twiddle_one<R, NS, U + 1, 1>(v, st, tw);
      }
    }
  }

  template <int R, int NS>
  __device__ __forceinline__ static void store_stage(const float2 (&v)[E], float2* lds, int tau) {
#pragma unroll
    for (int u = 0; u < E / R; ++u) {
      const int j = tau + T * u;
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[LineLds<N>::idx(base + r * NS)] = v[u + r * (E / R)];
    }
  }

  __device__ __forceinline__ static void load_strided(float2 (&v)[E], const float2* lds, int tau) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = lds[LineLds<N>::idx(tau + T * e)];
  }

  // Chain: TwChain (twiddles formed in registers) or TwChainTab (LDS table)
  template <int NS, int R, int... RS, class Chain>
  __device__ __forceinline__ static void stages(float2 (&v)[E], float2* lds, int tau, const Chain& tws,
                                                const float2* __restrict__ tw) {
    static_assert(E % R == 0, "radix must divide the elements per thread");
    if constexpr (NS > 1) twiddle_one<R, NS, 0, 1>(v, tws.s, tw);
    butterflies<R, 0>(v);
    if constexpr (sizeof...(RS) > 0) {
      store_stage<R, NS>(v, lds, tau);
      lds_sync<SYNC>();
      load_strided(v, lds, tau);
      lds_sync<SYNC>();
      stages<NS * R, RS...>(v, lds, tau, tws.next, tw);
    } else {
      static_assert(NS * R == N, "radices must multiply to N");
    }
  }

  template <int... RS>
  __device__ __forceinline__ static void run(float2 (&v)[E], float2* lds, int tau,
                                             const TwChain<N, T, 1, RS...>& tws,
                                             const float2* __restrict__ tw) {
    stages<1, RS...>(v, lds, tau, tws, tw);
  }
  template <int... RS>
  __device__ __forceinline__ static void run(float2 (&v)[E], float2* lds, int tau,
                                             const TwChainTab<N, T, 1, RS...>& tws,
                                             const float2* __restrict__ tw) {
    stages<1, RS...>(v, lds, tau, tws, tw);
  }

  // The same transform split after its first stage, so the threads that hold
  // a line may change between the two halves: first() runs stage 1 (no
  // twiddles) and scatters its outputs to LDS; rest() reads them in the
  // strided distribution of the (possibly different) owning threads and runs
  // the remaining stages.  The caller synchronises in between.
  template <int R1, int... RS>
  __device__ __forceinline__ static void first(float2 (&v)[E], float2* lds, int tau) {
    butterflies<R1, 0>(v);
    store_stage<R1, 1>(v, lds, tau);
  }
  template <int R1, int... RS>
  __device__ __forceinline__ static void rest(float2 (&v)[E], float2* lds, int tau,
                                              const TwChain<N, T, 1, R1, RS...>& tws,
                                              const float2* __restrict__ tw) {
    load_strided(v, lds, tau);
    lds_sync<SYNC>();
    stages<R1, RS...>(v, lds, tau, tws.next, tw);
  }
};

}  // namespace em
