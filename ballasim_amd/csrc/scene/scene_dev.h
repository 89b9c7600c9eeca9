// The analytic procedural scene and the per-pixel sample-time controller, as device code.
//
// One lane per camera pixel. A plan step finds the shapes whose swept bounding circle can cover the
// pixel during the window (a bit mask, at most kMaxShapes) and their largest image speed v_hat; a
// pixel no shape can reach is static and takes no samples. The controller then walks the window's
// dyadic lattice, flags intervals by signature flips and the eta_L midpoint test, and refines the
// flagged ones depth-first in time order, so the run comes out sorted without a sort.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/bsx/bsx_scene.h"

namespace bsx {

inline constexpr int kMaxShapes = 256;
inline constexpr int kMaskWords = kMaxShapes / 64;
inline constexpr int kWindowLog2 = 23;
inline constexpr int64_t kWindowNs = int64_t{1} << kWindowLog2;

struct DevShape {
  double cx, cy, vx, vy, angle, omega;
  float hx, hy, radius;  // radius: the bounding circle
  float lux, contrast, inv_cell, speed;  // speed: |v| + |omega| radius (scene px/s)
  uint32_t seed;
  int32_t kind, octaves;
};

struct SceneDev {
  const DevShape* shapes;
  int32_t n;
  float background;
};

struct ViewDev {
  uint32_t width, height;
  double ox, oy, scale;
};

__device__ __forceinline__ uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}

// Lattice value in [0, 1) at integer (ix, iy) of octave o.
__device__ __forceinline__ float lattice_value(int32_t ix, int32_t iy, uint32_t seed, int o) {
  const uint32_t h = mix32(static_cast<uint32_t>(ix) * 0x8da6b343u ^ static_cast<uint32_t>(iy) * 0xd8163841u ^
                           (seed + static_cast<uint32_t>(o) * 0x9e3779b9u));
  return static_cast<float>(h >> 8) * (1.0f / 16777216.0f);
}

// Smoothstep-interpolated value noise, octaves summed with halving amplitudes, normalized to [0, 1].
__device__ __forceinline__ float value_noise(float u, float v, uint32_t seed, int octaves, float inv_cell) {
  float sum = 0.0f, norm = 0.0f, amp = 1.0f, f = inv_cell;
  for (int o = 0; o < octaves; ++o) {
    const float X = u * f, Y = v * f;
    const float fx = floorf(X), fy = floorf(Y);
    const int32_t ix = static_cast<int32_t>(fx), iy = static_cast<int32_t>(fy);
    const float tx = X - fx, ty = Y - fy;
    const float sx = tx * tx * (3.0f - 2.0f * tx), sy = ty * ty * (3.0f - 2.0f * ty);
    const float a = lattice_value(ix, iy, seed, o), b = lattice_value(ix + 1, iy, seed, o);
    const float c = lattice_value(ix, iy + 1, seed, o), d = lattice_value(ix + 1, iy + 1, seed, o);
    const float top = a + (b - a) * sx, bot = c + (d - c) * sx;
    sum += amp * (top + (bot - top) * sy);
    norm += amp;
    amp *= 0.5f;
    f *= 2.0f;
  }
  return norm > 0.0f ? sum / norm : 0.5f;
}

struct Sample {
  float e;
  uint32_t sig;
};

// The scene at scene point (x, y) and time t (s), over the candidate shapes of the mask: the topmost
// shape that holds the point, found first, and then its texture, so that the lanes of a wave that find
// their shapes at different candidates evaluate the texture together.
__device__ __forceinline__ Sample eval_scene(const SceneDev& sc, const uint64_t* mask, double x, double y,
                                             double t) {
  int hit = -1;
  float hu = 0.0f, hv = 0.0f;
  for (int w = kMaskWords - 1; w >= 0 && hit < 0; --w) {
    uint64_t m = mask[w];
    while (m != 0) {
      const int b = 63 - __clzll(static_cast<long long>(m));
      m &= ~(1ull << b);
      const int i = w * 64 + b;
      const DevShape& s = sc.shapes[i];
      const float dx = static_cast<float>(x - (s.cx + s.vx * t));
      const float dy = static_cast<float>(y - (s.cy + s.vy * t));
      if (dx * dx + dy * dy > s.radius * s.radius) continue;
      double th = s.angle + s.omega * t;
      th -= 6.283185307179586 * floor(th * 0.15915494309189535);
      float sn, cs;
      __sincosf(static_cast<float>(th), &sn, &cs);
      const float u = cs * dx + sn * dy, v = -sn * dx + cs * dy;
      const bool inside = s.kind == BSX_SHAPE_DISK ? (u * u + v * v <= s.hx * s.hx)
                                                   : (fabsf(u) <= s.hx && fabsf(v) <= s.hy);
      if (!inside) continue;
      hit = i;
      hu = u;
      hv = v;
      break;
    }
  }
  if (hit < 0) return Sample{sc.background, 0u};
  const DevShape& s = sc.shapes[hit];
  const float n = value_noise(hu + s.hx, hv + s.hy, s.seed, s.octaves, s.inv_cell);
  return Sample{s.lux * (1.0f + s.contrast * (2.0f * n - 1.0f)), static_cast<uint32_t>(hit + 1)};
}

// The shapes whose bounding circle, swept over [t0, t1] (s), passes within a pixel of (x, y), and
// their largest speed.
__device__ __forceinline__ float plan_pixel(const SceneDev& sc, double x, double y, double t0, double t1,
                                            uint64_t* mask) {
  float v_hat = 0.0f;
  for (int w = 0; w < kMaskWords; ++w) mask[w] = 0;
  for (int i = 0; i < sc.n; ++i) {
    const DevShape& s = sc.shapes[i];
    const double ax = s.cx + s.vx * t0, ay = s.cy + s.vy * t0;
    const double bx = s.cx + s.vx * t1, by = s.cy + s.vy * t1;
    const double ex = bx - ax, ey = by - ay, px = x - ax, py = y - ay;
    const double ee = ex * ex + ey * ey;
    double f = ee > 0.0 ? (px * ex + py * ey) / ee : 0.0;
    f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
    const double qx = px - f * ex, qy = py - f * ey;
    const double r = static_cast<double>(s.radius) + 1.5;
    if (qx * qx + qy * qy <= r * r) {
      mask[i >> 6] |= 1ull << (i & 63);
      v_hat = fmaxf(v_hat, s.speed);
    }
  }
  return v_hat;
}

// eta_L of a triple (reference: midpoint_eta, src/engine/sampling.h:46-58) for one constant light
// group: |(a + b) (-1/2) + m| / (dark + min(a, m, b)), in fp32.
__device__ __forceinline__ float midpoint_eta(float dark, float a, float m, float b) {
  const float dev = fabsf((a + b) * -0.5f + m);
  const float low = fminf(fminf(a, m), b);
  const float num = 0.0f + 1.0f * dev;
  const float den = dark + 1.0f * low;
  return num / den;
}

}  // namespace bsx
