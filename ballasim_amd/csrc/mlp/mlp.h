// The learned sample-interval proposer: a 12-64-64-1 multilayer perceptron with SiLU hidden layers,
// BF16 weights on the MFMA units with FP32 accumulation, FP32 master weights and Adam state.
//
// Parameter vector layout (5,057 fp32 values, the order of the gradients and of Adam's moments):
//   W1 [12 x 64] (input-major), b1 [64], W2 [64 x 64] (row = layer-1 unit), b2 [64], w3 [64], b3 [1].
// Forward, for a row x of 12 features:
//   z1 = bf16(x) . bf16(W1) + b1,  h1 = bf16(silu(z1)),
//   z2 = h1 . bf16(W2) + b2,       h2 = bf16(silu(z2)),
//   y  = h2 . bf16(w3) + b3,
// products of bf16 values summed in fp32 (MFMA), biases and activations in fp32, silu(z) =
// z / (1 + exp(-z)). The training loss is the mean over the batch of (y - target)^2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bsx {
namespace mlp {

inline constexpr int kIn = 12;      // features
inline constexpr int kInPad = 16;   // features padded to two MFMA k-steps
inline constexpr int kHid = 64;     // units per hidden layer
inline constexpr int kOffW1 = 0;
inline constexpr int kOffB1 = kOffW1 + kIn * kHid;
inline constexpr int kOffW2 = kOffB1 + kHid;
inline constexpr int kOffB2 = kOffW2 + kHid * kHid;
inline constexpr int kOffW3 = kOffB2 + kHid;
inline constexpr int kOffB3 = kOffW3 + kHid;
inline constexpr int kParams = kOffB3 + 1;  // 5057
inline constexpr int kRowsPerWave = 64;     // a wave's rows (two 32-row MFMA tiles)

struct Adam {
  float lr, beta1, beta2, eps;
  int32_t step;  // 1-based step of this update (bias correction)
};

// The bf16 copies the kernels read (round to nearest even of the fp32 masters), in the layout above.
hipError_t launch_to_bf16(const float* w, uint16_t* wb, int n, hipStream_t stream);
// Forward of n rows (features fp32, row-major [n x 12]) into y [n].
hipError_t launch_forward(const uint16_t* wb, const float* b_fp32, const float* x, int64_t n, float* y,
                          hipStream_t stream);
// One training batch of n rows: per-group partial gradients (ceil(n / 64) groups of 64 rows, one row
// of kParams values each in `partial`) of the mean squared error over `batch` rows (batch >= n: the
// loss of a larger batch split over launches), and per-group partial sums of the squared errors in
// loss_partial.
hipError_t launch_gradients(const uint16_t* wb, const float* b_fp32, const float* x, const float* target, int64_t n,
                            int64_t batch, float* partial, float* loss_partial, hipStream_t stream);
// Sums the partial gradients of `waves` groups (in group order) into grad, then one Adam step of the
// fp32 masters w (moments m, v) and their bf16 copies wb.
hipError_t launch_adam(const float* partial, int waves, float* grad, float* w, float* m, float* v, uint16_t* wb,
                       const Adam& a, hipStream_t stream);
// MFMA check: C [32 x 32] fp32 = A [32 x 8] . B [8 x 32], A and B bf16 row-major, through the kernels'
// tile routine (a self-check of the operand layouts).
hipError_t launch_tile_check(const uint16_t* a, const uint16_t* b, float* c, hipStream_t stream);

}  // namespace mlp
}  // namespace bsx
