// Device building blocks of the proposer's kernels: bf16 rounding, the 32x32 MFMA tile over operands
// in LDS, and the forward pass of one wave's 64 rows (used by the training kernel and the learned
// sampler).
//
// MFMA: v_mfma_f32_32x32x8bf16_1k multiplies a 32x8 bf16 tile A by an 8x32 bf16 tile B into a 32x32
// fp32 accumulator. Lane l (of 64) supplies A[l % 32][4 (l / 32) + i] and B[4 (l / 32) + i][l % 32]
// for i = 0..3, and holds accumulator element j (of 16) at row 8 (j / 4) + 4 (l / 32) + j % 4,
// column l % 32. (launch_tile_check tests this layout against a host product.)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.h"

namespace bsx {
namespace mlp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

inline constexpr int kS16 = 72;  // row stride (elements) of bf16 LDS matrices: 36 words, no bank aliasing
inline constexpr int kS32 = 65;  // row stride of fp32 LDS matrices

// fp32 -> bf16, round to nearest even (finite inputs).
__host__ __device__ __forceinline__ uint16_t to_bf16(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(uint16_t h) {
  return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16);
}

// Orders this wave's LDS writes before its lanes read each other's values (the waves of a block run
// their own loops, so no block barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
// d silu / dz = s (1 + z (1 - s)), s = 1 / (1 + exp(-z)).
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// acc += A . B over k = 0 .. K - 1 (K a multiple of 8) for the 32x32 tile whose A rows start at a and
// B columns at b: A(m, k) = a[m * am + k * ak], B(k, n) = b[k * bk + n * bn] (bf16 in LDS). Called by
// all 64 lanes of the wave.
__device__ __forceinline__ void mma_tile(f32x16& acc, const uint16_t* a, int am, int ak, const uint16_t* b, int bk,
                                         int bn, int K) {
  const int l = threadIdx.x & 63, r = l & 31, kq = (l >> 5) * 4;
  for (int k0 = 0; k0 < K; k0 += 8) {
    bf16x4 av, bv;
    for (int i = 0; i < 4; ++i) {
      av[i] = static_cast<short>(a[r * am + (k0 + kq + i) * ak]);
      bv[i] = static_cast<short>(b[(k0 + kq + i) * bk + r * bn]);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
  }
}

// Row and column of accumulator element j of this lane.
__device__ __forceinline__ int acc_row(int j) { return 8 * (j >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (j & 3); }
__device__ __forceinline__ int acc_col() { return threadIdx.x & 31; }

// The weights as the kernels hold them in LDS: bf16 matrices (W1 padded to 16 input rows) and fp32
// biases and output weights (w3 as its bf16 value).
struct WeightsLds {
  uint16_t w1[kInPad * kS16];
  uint16_t w2[kHid * kS16];
  float b1[kHid], b2[kHid], w3[kHid];
  float b3;
};

// Loads the weights into LDS (all threads of the block; a barrier follows).
__device__ __forceinline__ void load_weights(WeightsLds& s, const uint16_t* wb, const float* w) {
  for (int i = threadIdx.x; i < kInPad * kHid; i += blockDim.x) {
    const int f = i / kHid, h = i % kHid;
    s.w1[f * kS16 + h] = f < kIn ? wb[kOffW1 + f * kHid + h] : uint16_t{0};
  }
  for (int i = threadIdx.x; i < kHid * kHid; i += blockDim.x) s.w2[(i / kHid) * kS16 + i % kHid] = wb[kOffW2 + i];
  for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
    s.b1[i] = w[kOffB1 + i];
    s.b2[i] = w[kOffB2 + i];
    s.w3[i] = from_bf16(wb[kOffW3 + i]);
  }
  if (threadIdx.x == 0) s.b3 = w[kOffB3];
  __syncthreads();
}

// One wave's activations in LDS: xs = the 64 input rows (bf16, columns 12 .. 31 zero), z1 and z2 the
// pre-activations (fp32), h1 = bf16(silu(z1)).
struct WaveLds {
  uint16_t xs[kRowsPerWave * kS16];
  float z1[kRowsPerWave * kS32];
  float z2[kRowsPerWave * kS32];
  uint16_t h1[kRowsPerWave * kS16];
};

// Writes this lane's features as row (lane) of xs, then syncs the wave.
__device__ __forceinline__ void put_features(WaveLds& s, const float* f) {
  const int l = threadIdx.x & 63;
  for (int c = 0; c < kS16; ++c) s.xs[l * kS16 + c] = c < kIn ? to_bf16(f[c]) : uint16_t{0};
  wave_sync();
}

// The forward pass of the wave's 64 rows (xs filled): z1, h1, z2 in LDS; returns this lane's row's y.
__device__ __forceinline__ float wave_forward(const WeightsLds& w, WaveLds& s) {
  for (int mt = 0; mt < 2; ++mt)
    for (int nt = 0; nt < 2; ++nt) {
      f32x16 acc = {};
      mma_tile(acc, s.xs + mt * 32 * kS16, kS16, 1, w.w1 + nt * 32, kS16, 1, kInPad);
      for (int j = 0; j < 16; ++j) {
        const int r = mt * 32 + acc_row(j), c = nt * 32 + acc_col();
        const float z = acc[j] + w.b1[c];
        s.z1[r * kS32 + c] = z;
        s.h1[r * kS16 + c] = to_bf16(silu(z));
      }
    }
  wave_sync();
  for (int mt = 0; mt < 2; ++mt)
    for (int nt = 0; nt < 2; ++nt) {
      f32x16 acc = {};
      mma_tile(acc, s.h1 + mt * 32 * kS16, kS16, 1, w.w2 + nt * 32, kS16, 1, kHid);
      for (int j = 0; j < 16; ++j) {
        const int r = mt * 32 + acc_row(j), c = nt * 32 + acc_col();
        s.z2[r * kS32 + c] = acc[j] + w.b2[c];
      }
    }
  wave_sync();
  const int l = threadIdx.x & 63;
  float y = 0.0f;
  for (int h = 0; h < kHid; ++h) y += from_bf16(to_bf16(silu(s.z2[l * kS32 + h]))) * w.w3[h];
  return y + w.b3;
}

}  // namespace mlp
}  // namespace bsx
