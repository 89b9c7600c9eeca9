// Deterministic transcendentals for the sensor path, in fp32 and fp64, host and device.
//
// libm, the HIP device library (ocml) and the reference's CPU build round exp/log differently, so
// none of them can be called where results must be bitwise portable between this library's GPU
// kernels, its CPU restatement and the reference's CPU path. Every function here uses only +, -,
// *, /, sqrt and exact operations in a fixed order (see hd.h for the floating-point policy), and
// follows a textbook method named at the function. The methods, their polynomial degrees and the
// order of operations are those the reference fixes for its own transcendentals
// (src/core/pmath/explog.h:1-20, special.h:1-14, trig.h:1-12), because parity with the reference's
// fp32 CPU path requires the same roundings; the constants are derived from their definitions
// (gen_constants.py, and the unit tests re-derive the scalar ones).
//
// Underflow policy: exp(x) is exactly 0 below kMinArg (just above ln of the smallest normal), so no
// subnormal is produced.
#pragma once

#include "hd.h"
#include "pmath_tables.h"

namespace bsx::pm {

template <class T>
struct C;

template <>
struct C<float> {
  static constexpr float kLog2e = 0x1.715476p+0f;    // 1/ln 2, rounded
  static constexpr float kLn2Hi = 0x1.62e400p-1f;    // ln 2 truncated to 16 bits: k * kLn2Hi is exact
  static constexpr float kLn2Lo = 0x1.7f7d1cp-20f;   // ln 2 - kLn2Hi, rounded
  static constexpr float kMinArg = -87.33f;
  static constexpr float kMaxArg = 88.72f;
  static constexpr float kSqrtHalf = 0x1.6a09e6p-1f;
  static constexpr float kTwoOverPi = 0x1.45f306p-1f;
  static constexpr float kPio2_1 = 0x1.922p+0f;      // pi/2 = P1 + P2 + P3, P1 and P2 rounded to 12 bits
  static constexpr float kPio2_2 = -0x1.2aep-18f;
  static constexpr float kPio2_3 = -0x1.de973ep-31f;
  static constexpr float kTrigMaxArg = 6000.0f;      // |quadrant| < 2^12 keeps k * P1, k * P2 exact
};

template <>
struct C<double> {
  static constexpr double kLog2e = 0x1.71547652b82fep+0;
  static constexpr double kLn2Hi = 0x1.62e42fee00000p-1;  // ln 2 truncated to 32 bits
  static constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;
  static constexpr double kMinArg = -708.39;
  static constexpr double kMaxArg = 709.78;
  static constexpr double kSqrtHalf = 0x1.6a09e667f3bcdp-1;
  static constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
  static constexpr double kPio2_1 = 0x1.921fb544p+0;       // 31 bits
  static constexpr double kPio2_2 = 0x1.0b4611a6p-34;      // 32 bits
  static constexpr double kPio2_3 = 0x1.3198a2e037073p-69;
  static constexpr double kTrigMaxArg = 1.0e6;
};

// x * 2^e by exponent-field arithmetic (x * 2^e1 * 2^e2 near the ends of the range).
BSX_HD BSX_INLINE float ldexp_exact(float x, int e) {
  const int e1 = e > 127 ? 127 : (e < -126 ? -126 : e);
  const int e2 = e - e1;
  const float p1 = float_of(static_cast<uint32_t>(e1 + 127) << 23);
  if (e2 == 0) return x * p1;
  return (x * p1) * float_of(static_cast<uint32_t>(e2 + 127) << 23);
}
BSX_HD BSX_INLINE double ldexp_exact(double x, int e) {
  const int e1 = e > 1023 ? 1023 : (e < -1022 ? -1022 : e);
  const int e2 = e - e1;
  const double p1 = double_of(static_cast<uint64_t>(e1 + 1023) << 52);
  if (e2 == 0) return x * p1;
  return (x * p1) * double_of(static_cast<uint64_t>(e2 + 1023) << 52);
}

// m in [0.5, 1) and e with x = m 2^e, for finite x > 0 (subnormals scaled first).
BSX_HD BSX_INLINE float frexp_exact(float x, int* e) {
  int bias = 0;
  if (x < 0x1p-126f) {
    x = x * 0x1p25f;
    bias = 25;
  }
  const uint32_t b = bits_of(x);
  *e = static_cast<int>((b >> 23) & 0xffu) - 126 - bias;
  return float_of((b & 0x807fffffu) | (126u << 23));
}
BSX_HD BSX_INLINE double frexp_exact(double x, int* e) {
  int bias = 0;
  if (x < 0x1p-1022) {
    x = x * 0x1p54;
    bias = 54;
  }
  const uint64_t b = bits_of(x);
  *e = static_cast<int>((b >> 52) & 0x7ffu) - 1022 - bias;
  return double_of((b & 0x800fffffffffffffull) | (uint64_t{1022} << 52));
}

namespace k {

// e^r on |r| <= ln2/2: Taylor polynomial 1 + r + r^2/2! + ... (degree 7 in fp32, 13 in fp64), Horner.
BSX_HD BSX_INLINE float exp_poly(float r) {
  float p = 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  return p * r + 1.0f;
}
BSX_HD BSX_INLINE double exp_poly(double r) {
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  return p * r + 1.0;
}

// ln(1 + f) = 2 atanh(s), s = f / (2 + f): R(z) = 2z/3 + 2z^2/5 + ... with z = s^2 (4 terms fp32,
// 9 terms fp64), so that ln(1 + f) = f - s (f - R).
BSX_HD BSX_INLINE float atanh_tail(float z) {
  float p = 2.0f / 9.0f;
  p = p * z + 2.0f / 7.0f;
  p = p * z + 2.0f / 5.0f;
  p = p * z + 2.0f / 3.0f;
  return p * z;
}
BSX_HD BSX_INLINE double atanh_tail(double z) {
  double p = 2.0 / 19.0;
  p = p * z + 2.0 / 17.0;
  p = p * z + 2.0 / 15.0;
  p = p * z + 2.0 / 13.0;
  p = p * z + 2.0 / 11.0;
  p = p * z + 2.0 / 9.0;
  p = p * z + 2.0 / 7.0;
  p = p * z + 2.0 / 5.0;
  p = p * z + 2.0 / 3.0;
  return p * z;
}

// e^x - 1 on |x| < 0.35: Taylor, x + x^2/2! + ..., degree 8 (fp32) / 14 (fp64).
BSX_HD BSX_INLINE float expm1_poly(float x) {
  float p = 1.0f / 40320.0f;
  p = p * x + 1.0f / 5040.0f;
  p = p * x + 1.0f / 720.0f;
  p = p * x + 1.0f / 120.0f;
  p = p * x + 1.0f / 24.0f;
  p = p * x + 1.0f / 6.0f;
  p = p * x + 0.5f;
  return (p * x) * x + x;
}
BSX_HD BSX_INLINE double expm1_poly(double x) {
  double p = 1.0 / 87178291200.0;
  p = p * x + 1.0 / 6227020800.0;
  p = p * x + 1.0 / 479001600.0;
  p = p * x + 1.0 / 39916800.0;
  p = p * x + 1.0 / 3628800.0;
  p = p * x + 1.0 / 362880.0;
  p = p * x + 1.0 / 40320.0;
  p = p * x + 1.0 / 5040.0;
  p = p * x + 1.0 / 720.0;
  p = p * x + 1.0 / 120.0;
  p = p * x + 1.0 / 24.0;
  p = p * x + 1.0 / 6.0;
  p = p * x + 0.5;
  return (p * x) * x + x;
}

// sin and cos on |r| <= pi/4, Taylor in z = r^2. fp32: the sine's 1/5! coefficient is 1/120 written
// to 8 significant digits (0x1.111110p-7, one ulp under the correctly rounded value), the others
// are 1/n! rounded; fp64: all 1/n! rounded.
BSX_HD BSX_INLINE void sincos_poly(float r, float* s, float* c) {
  const float z = r * r;
  float ps = 1.0f / 362880.0f;
  ps = ps * z - 1.0f / 5040.0f;
  ps = ps * z + 0x1.111110p-7f;
  ps = ps * z - 1.0f / 6.0f;
  *s = r + (r * z) * ps;
  float pc = -1.0f / 3628800.0f;
  pc = pc * z + 1.0f / 40320.0f;
  pc = pc * z - 1.0f / 720.0f;
  pc = pc * z + 1.0f / 24.0f;
  *c = (1.0f - 0.5f * z) + (z * z) * pc;
}
BSX_HD BSX_INLINE void sincos_poly(double r, double* s, double* c) {
  const double z = r * r;
  double ps = 1.0 / 355687428096000.0;
  ps = ps * z - 1.0 / 1307674368000.0;
  ps = ps * z + 1.0 / 6227020800.0;
  ps = ps * z - 1.0 / 39916800.0;
  ps = ps * z + 1.0 / 362880.0;
  ps = ps * z - 1.0 / 5040.0;
  ps = ps * z + 1.0 / 120.0;
  ps = ps * z - 1.0 / 6.0;
  *s = r + (r * z) * ps;
  double pc = -1.0 / 6402373705728000.0;
  pc = pc * z + 1.0 / 20922789888000.0;
  pc = pc * z - 1.0 / 87178291200.0;
  pc = pc * z + 1.0 / 479001600.0;
  pc = pc * z - 1.0 / 3628800.0;
  pc = pc * z + 1.0 / 40320.0;
  pc = pc * z - 1.0 / 720.0;
  pc = pc * z + 1.0 / 24.0;
  *c = (1.0 - 0.5 * z) + (z * z) * pc;
}

}  // namespace k

// e^x: Cody-Waite reduction x = k ln2 + r, then the Taylor kernel and an exact scaling by 2^k.
template <class T>
BSX_HD BSX_INLINE T exp(T x) {
  if (!(x >= C<T>::kMinArg)) return x != x ? x : T(0);
  if (x > C<T>::kMaxArg) return inf_<T>();
  const T kk = rint_(x * C<T>::kLog2e);
  const T r = (x - kk * C<T>::kLn2Hi) - kk * C<T>::kLn2Lo;
  return ldexp_exact(k::exp_poly(r), static_cast<int>(kk));
}

// ln x: x = m 2^e with m in [sqrt(1/2), sqrt 2), ln m = f - s (f - R(s^2)), f = m - 1 (exact).
template <class T>
BSX_HD BSX_INLINE T log(T x) {
  if (!(x > T(0))) return x == T(0) ? -inf_<T>() : nan_<T>();
  if (x == inf_<T>()) return x;
  int e = 0;
  T m = frexp_exact(x, &e);
  if (m < C<T>::kSqrtHalf) {
    m = m + m;
    e -= 1;
  }
  const T f = m - T(1);
  const T s = f / (T(2) + f);
  const T lnm = f - s * (f - k::atanh_tail(s * s));
  const T ef = static_cast<T>(e);
  return ef * C<T>::kLn2Hi + (lnm + ef * C<T>::kLn2Lo);
}

template <class T>
BSX_HD BSX_INLINE T expm1(T x) {
  if (fabs_(x) < T(0.35)) return k::expm1_poly(x);
  return exp(x) - T(1);
}

// ln(1 + x) with Goldberg's correction.
template <class T>
BSX_HD BSX_INLINE T log1p(T x) {
  const T u = T(1) + x;
  if (u == T(1)) return x;
  return log(u) * (x / (u - T(1)));
}

// sin and cos: Cody-Waite reduction by pi/2 in three parts, quadrant by rint(x 2/pi); NaN for
// |x| > kTrigMaxArg.
template <class T>
BSX_HD BSX_INLINE void sincos(T x, T* s, T* c) {
  if (!(fabs_(x) <= C<T>::kTrigMaxArg)) {
    *s = *c = nan_<T>();
    return;
  }
  if (x == T(0)) {
    *s = x;
    *c = T(1);
    return;
  }
  const T q = rint_(x * C<T>::kTwoOverPi);
  const T r = q == T(0) ? x : ((x - q * C<T>::kPio2_1) - q * C<T>::kPio2_2) - q * C<T>::kPio2_3;
  T sr, cr;
  k::sincos_poly(r, &sr, &cr);
  switch (static_cast<int>(q) & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

// Residual exponentials R_j(y) = e^{-y} - sum_{i<j} (-y)^i / i! for y >= 0, j = 0..4: e^{-y} minus
// its Taylor polynomial, by that direct form for y > 1 and by the nested series
// ((-y)^j / j!) (1 - y/(j+1) (1 - y/(j+2) (...))) (10 terms fp32, 17 fp64) for y <= 1.
template <class T>
BSX_HD BSX_INLINE T resid_exp(int j, T y) {
  if (j <= 0) return exp(-y);
  T yj = T(1), jf = T(1);
  for (int i = 1; i <= j; ++i) {
    yj = yj * (-y);
    jf = jf * static_cast<T>(i);
  }
  if (y > T(1)) {
    T t = T(1), s = T(1);
    for (int i = 1; i < j; ++i) {
      t = t * (-y) / static_cast<T>(i);
      s = s + t;
    }
    return exp(-y) - s;
  }
  constexpr int kTerms = sizeof(T) == 8 ? 17 : 10;
  T q = T(1);
  for (int i = kTerms; i >= 1; --i) q = T(1) - y * q / static_cast<T>(j + i);
  return yj * (q / jf);
}

// erfcx(x) = e^{x^2} erfc(x), x >= 0, by Weideman's rational approximation of the Faddeeva
// function on the imaginary axis: 2 p(Z) / (L + x)^2 + (1/sqrt pi) / (L + x), Z = (L - x)/(L + x),
// N = 24 terms in fp32 and 40 in fp64. erfcx(+inf) = 0, negative arguments give 1, NaN propagates.
template <class T>
BSX_HD BSX_INLINE T erfcx(T x) {
  if (!(x > T(0) && x < inf_<T>())) return x > T(0) ? T(0) : (x == x ? T(1) : x);
  constexpr bool kD = sizeof(T) == 8;
  const T lam = static_cast<T>(kD ? tables::kWeidemanL40 : tables::kWeidemanL24);
  const T lx = lam + x;
  const T z = (lam - x) / lx;
  T p;
  if constexpr (kD) {
    p = static_cast<T>(tables::kWeidemanP40[0]);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int i = 1; i < 40; ++i) p = p * z + static_cast<T>(tables::kWeidemanP40[i]);
  } else {
    p = static_cast<T>(tables::kWeidemanP24[0]);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int i = 1; i < 24; ++i) p = p * z + static_cast<T>(tables::kWeidemanP24[i]);
  }
  constexpr double kInvSqrtPi = 0x1.20dd750429b6dp-1;
  return T(2) * p / (lx * lx) + static_cast<T>(kInvSqrtPi) / lx;
}

// Dawson's integral F(x) = e^{-x^2} int_0^x e^{t^2} dt. Below the switch (0.2 fp64, 1 fp32) its
// Taylor series x sum (-2x^2)^n / (2n+1)!! (12 terms, nested); above, Rybicki's sampling sum with
// h = 1/4 and 13 terms; beyond 1e8, 1/(2x).
template <class T>
BSX_HD BSX_INLINE T dawson(T x) {
  const T ax = fabs_(x);
  constexpr bool kD = sizeof(T) == 8;
  const T switch_at = kD ? T(0.2) : T(1);
  if (ax < switch_at) {
    const T two_x2 = T(2) * x * x;
    T q = T(1);
    for (int n = 12; n >= 1; --n) q = T(1) - two_x2 * q / static_cast<T>(2 * n + 1);
    return x * q;
  }
  if (!(ax <= T(1e8))) return ax == ax ? copysign_(T(0.5) / ax, x) : x;
  const T h = T(0.25);
  const int n0 = 2 * static_cast<int>(T(0.5) * ax / h + T(0.5));
  const T xp = ax - static_cast<T>(n0) * h;
  T e1 = exp(T(2) * xp * h);
  const T e2 = e1 * e1;
  T d1 = static_cast<T>(n0 + 1), d2 = static_cast<T>(n0 - 1);
  T sum = T(0);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int i = 0; i < 13; ++i) {
    sum = sum + static_cast<T>(tables::kRybicki[i]) * (e1 / d1 + T(1) / (d2 * e1));
    d1 = d1 + T(2);
    d2 = d2 - T(2);
    e1 = e1 * e2;
  }
  constexpr double kInvSqrtPi = 0x1.20dd750429b6dp-1;
  return copysign_(static_cast<T>(kInvSqrtPi) * exp(-xp * xp) * sum, x);
}

}  // namespace bsx::pm
