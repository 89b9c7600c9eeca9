// Host/device qualifiers for the functions this library compiles both for the gfx950 device and for
// the host (map drawing, the host-side init, unit tests of the maths).
//
// Floating-point policy (applies to every translation unit that includes this header): compiled
// with -ffp-contract=off and without fast-math, so that a*b+c is two correctly rounded operations
// on the host and on the device alike; fp32 division and sqrt are IEEE correctly rounded (HIP's
// default -fhip-fp32-correctly-rounded-divide-sqrt). Under that policy every function below that
// uses only +, -, *, /, sqrt and exact operations (rint, floor, bit casts) rounds identically on
// the CPU and on the GPU, which is what makes the sensor's fp32 results bitwise portable.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BSX_HD __host__ __device__
#define BSX_DEV __device__
#define BSX_INLINE __forceinline__
#define BSX_NOINLINE __noinline__
#else
#define BSX_HD
#define BSX_DEV
#define BSX_INLINE inline
#define BSX_NOINLINE
#endif

// Pointers into device memory that the kernels read from descriptors (not kernel arguments) carry
// the global address space in the device pass, so that their accesses compile to global (not flat)
// memory instructions; the host pass sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define BSX_GLOBAL __attribute__((address_space(1)))
#else
#define BSX_GLOBAL
#endif

namespace bsx {

template <class T>
struct FloatEps;
template <>
struct FloatEps<float> {
  static constexpr float value = 0x1p-23f;  // ulp(1)
};
template <>
struct FloatEps<double> {
  static constexpr double value = 0x1p-52;
};

template <class T>
BSX_HD BSX_INLINE T vmin(T a, T b) {
  return b < a ? b : a;
}
template <class T>
BSX_HD BSX_INLINE T vmax(T a, T b) {
  return a < b ? b : a;
}

BSX_HD BSX_INLINE float fabs_(float x) { return __builtin_fabsf(x); }
BSX_HD BSX_INLINE double fabs_(double x) { return __builtin_fabs(x); }
BSX_HD BSX_INLINE float sqrt_(float x) { return __builtin_sqrtf(x); }
BSX_HD BSX_INLINE double sqrt_(double x) { return __builtin_sqrt(x); }
BSX_HD BSX_INLINE float rint_(float x) { return __builtin_rintf(x); }
BSX_HD BSX_INLINE double rint_(double x) { return __builtin_rint(x); }
BSX_HD BSX_INLINE float floor_(float x) { return __builtin_floorf(x); }
BSX_HD BSX_INLINE double floor_(double x) { return __builtin_floor(x); }
BSX_HD BSX_INLINE float copysign_(float x, float s) { return __builtin_copysignf(x, s); }
BSX_HD BSX_INLINE double copysign_(double x, double s) { return __builtin_copysign(x, s); }

template <class T>
BSX_HD BSX_INLINE T inf_() {
  return static_cast<T>(__builtin_inf());
}
template <class T>
BSX_HD BSX_INLINE T nan_() {
  return static_cast<T>(__builtin_nan(""));
}

BSX_HD BSX_INLINE uint32_t bits_of(float f) { return __builtin_bit_cast(uint32_t, f); }
BSX_HD BSX_INLINE float float_of(uint32_t u) { return __builtin_bit_cast(float, u); }
BSX_HD BSX_INLINE uint64_t bits_of(double f) { return __builtin_bit_cast(uint64_t, f); }
BSX_HD BSX_INLINE double double_of(uint64_t u) { return __builtin_bit_cast(double, u); }

// int64 ns to seconds the way the kernel forms reference-relative times: through fp64 (exact for
// |dt| < 2^53 ns), then rounded once to T.
template <class T>
BSX_HD BSX_INLINE T ns_to_s64(int64_t dt) {
  return static_cast<T>(static_cast<double>(dt) * 1e-9);
}

}  // namespace bsx
