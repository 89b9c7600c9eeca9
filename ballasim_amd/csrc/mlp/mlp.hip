// Kernels of the learned proposer: bf16 copies, batched forward, per-group gradients (forward and
// backward on the MFMA units, four waves per group of 64 rows), and the reduction of the partial
// gradients with an Adam step.
#include <hip/hip_runtime.h>

#include "mlp_dev.h"

namespace bsx {
namespace mlp {

namespace {

__global__ void to_bf16_kernel(const float* w, uint16_t* wb, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = to_bf16(w[i]);
}

__device__ __forceinline__ void load_row(const float* x, int64_t row, bool valid, float* f) {
  for (int c = 0; c < kIn; ++c) f[c] = valid ? x[row * kIn + c] : 0.0f;
}

// Rows in groups of 64, one group per wave per iteration (grid-stride).
__global__ __launch_bounds__(64) void forward_kernel(const uint16_t* wb, const float* w, const float* x, int64_t n,
                                                     float* y) {
  __shared__ WeightsLds W;
  __shared__ WaveLds S;
  load_weights(W, wb, w);
  const int l = threadIdx.x;
  const int64_t groups = (n + 63) / 64;
  for (int64_t g = blockIdx.x; g < groups; g += gridDim.x) {
    const int64_t row = g * 64 + l;
    const bool valid = row < n;
    float f[kIn];
    load_row(x, row, valid, f);
    put_features(S, f);
    const float out = wave_forward(W, S);
    if (valid) y[row] = out;
    wave_sync();
  }
}

// Gradients of (1 / batch) sum (y - t)^2 over the rows of group g = blockIdx.x (64 rows), written as
// one partial gradient vector per group. Four waves share the group: each takes one 32x32 tile of
// every 64x64 product (wave w: row tile w / 2, column tile w % 2), the element-wise steps are split
// over the 256 threads, and every sum runs over the rows in order, as a single wave would run it.
__global__ __launch_bounds__(256) void gradients_kernel(const uint16_t* wb, const float* w, const float* x,
                                                        const float* target, int64_t n, float two_over_batch,
                                                        float* partial, float* loss_partial) {
  __shared__ WeightsLds W;
  __shared__ WaveLds S;
  __shared__ uint16_t dz2[kRowsPerWave * kS16];
  __shared__ float dys[kRowsPerWave];
  // dz1 lives in z2's storage, which no step reads once dw3 is formed (a barrier separates them): 74 KB
  // of LDS per group, so that two groups fit in a CU's 160 KB.
  static_assert(sizeof(uint16_t) * kRowsPerWave * kS16 <= sizeof(S.z2), "dz1 fits in z2's storage");
  uint16_t* const dz1 = reinterpret_cast<uint16_t*>(S.z2);
  load_weights(W, wb, w);
  const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
  const int mt = wave >> 1, nt = wave & 1;
  const int64_t g = blockIdx.x;
  const int64_t row = g * 64 + l;
  const bool valid = row < n;
  if (wave == 0) {
    float f[kIn];
    load_row(x, row, valid, f);
    for (int c = 0; c < kS16; ++c) S.xs[l * kS16 + c] = c < kIn ? to_bf16(f[c]) : uint16_t{0};
  }
  __syncthreads();
  // layer 1, then layer 2: this wave's tile
  {
    f32x16 acc = {};
    mma_tile(acc, S.xs + mt * 32 * kS16, kS16, 1, W.w1 + nt * 32, kS16, 1, kInPad);
    for (int j = 0; j < 16; ++j) {
      const int r = mt * 32 + acc_row(j), c = nt * 32 + acc_col();
      const float z = acc[j] + W.b1[c];
      S.z1[r * kS32 + c] = z;
      S.h1[r * kS16 + c] = to_bf16(silu(z));
    }
  }
  __syncthreads();
  {
    f32x16 acc = {};
    mma_tile(acc, S.h1 + mt * 32 * kS16, kS16, 1, W.w2 + nt * 32, kS16, 1, kHid);
    for (int j = 0; j < 16; ++j) {
      const int r = mt * 32 + acc_row(j), c = nt * 32 + acc_col();
      S.z2[r * kS32 + c] = acc[j] + W.b2[c];
    }
  }
  __syncthreads();
  float* P = partial + g * kParams;
  if (wave == 0) {  // the output of row l, its error, and the group's squared error
    float y = 0.0f;
    for (int h = 0; h < kHid; ++h) y += from_bf16(to_bf16(silu(S.z2[l * kS32 + h]))) * W.w3[h];
    y += W.b3;
    const float e = valid ? y - target[row] : 0.0f;
    dys[l] = e * two_over_batch;
    float se = e * e;
    for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off);
    if (l == 0) loss_partial[g] = se;
  }
  __syncthreads();
  // dz2 = dy w3 silu'(z2): column l, rows 16 wave .. 16 wave + 15
  for (int r = wave * 16; r < wave * 16 + 16; ++r)
    dz2[r * kS16 + l] = to_bf16(dys[r] * W.w3[l] * silu_grad(S.z2[r * kS32 + l]));
  __syncthreads();
  // dw3[h] = sum_r dy_r h2[r][h] (wave 0), db2[h] = sum_r dz2[r][h] (wave 1), db3 = sum_r dy_r (wave 2)
  if (wave == 0) {
    float s3 = 0.0f;
    for (int r = 0; r < kRowsPerWave; ++r) s3 += dys[r] * from_bf16(to_bf16(silu(S.z2[r * kS32 + l])));
    P[kOffW3 + l] = s3;
  } else if (wave == 1) {
    float s2 = 0.0f;
    for (int r = 0; r < kRowsPerWave; ++r) s2 += from_bf16(dz2[r * kS16 + l]);
    P[kOffB2 + l] = s2;
  } else if (wave == 2 && l == 0) {
    float sb = 0.0f;
    for (int r = 0; r < kRowsPerWave; ++r) sb += dys[r];
    P[kOffB3] = sb;
  }
  // dW2 = h1^T dz2 (rows of W2: layer-1 units), this wave's tile
  {
    f32x16 acc = {};
    mma_tile(acc, S.h1 + mt * 32, 1, kS16, dz2 + nt * 32, kS16, 1, kRowsPerWave);
    for (int j = 0; j < 16; ++j) P[kOffW2 + (mt * 32 + acc_row(j)) * kHid + nt * 32 + acc_col()] = acc[j];
  }
  __syncthreads();  // z2 is read (dw3) before its storage takes dz1
  // dh1 = dz2 W2^T, dz1 = dh1 silu'(z1), this wave's tile
  {
    f32x16 acc = {};
    mma_tile(acc, dz2 + mt * 32 * kS16, kS16, 1, W.w2 + nt * 32 * kS16, 1, kS16, kHid);
    for (int j = 0; j < 16; ++j) {
      const int r = mt * 32 + acc_row(j), c = nt * 32 + acc_col();
      dz1[r * kS16 + c] = to_bf16(acc[j] * silu_grad(S.z1[r * kS32 + c]));
    }
  }
  __syncthreads();
  if (wave == 0) {  // db1[h] = sum_r dz1[r][h]
    float s1 = 0.0f;
    for (int r = 0; r < kRowsPerWave; ++r) s1 += from_bf16(dz1[r * kS16 + l]);
    P[kOffB1 + l] = s1;
  } else if (wave >= 2) {  // dW1 = x^T dz1 (features padded to a 32-row tile; rows 12 .. 31 are zero, not stored)
    const int ct = wave - 2;
    f32x16 acc = {};
    mma_tile(acc, S.xs, 1, kS16, dz1 + ct * 32, kS16, 1, kRowsPerWave);
    for (int j = 0; j < 16; ++j) {
      const int fr = acc_row(j);
      if (fr < kIn) P[kOffW1 + fr * kHid + ct * 32 + acc_col()] = acc[j];
    }
  }
}

__global__ void adam_kernel(const float* partial, int waves, float* grad, float* w, float* m, float* v, uint16_t* wb,
                            float lr, float beta1, float beta2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  // In wave order, as before; eight loads are issued ahead of their additions.
  float g = 0.0f;
  int k = 0;
  for (; k + 8 <= waves; k += 8) {
    float v[8];
    for (int u = 0; u < 8; ++u) v[u] = partial[static_cast<int64_t>(k + u) * kParams + i];
    for (int u = 0; u < 8; ++u) g += v[u];
  }
  for (; k < waves; ++k) g += partial[static_cast<int64_t>(k) * kParams + i];
  grad[i] = g;
  if (lr == 0.0f) return;  // gradients only
  const float mi = beta1 * m[i] + (1.0f - beta1) * g;
  const float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wi;
  wb[i] = to_bf16(wi);
}

__global__ __launch_bounds__(64) void tile_check_kernel(const uint16_t* a, const uint16_t* b, float* c) {
  __shared__ uint16_t as[32 * 8], bs[8 * 32];
  for (int i = threadIdx.x; i < 256; i += 64) {
    as[i] = a[i];
    bs[i] = b[i];
  }
  __syncthreads();
  f32x16 acc = {};
  mma_tile(acc, as, 8, 1, bs, 32, 1, 8);
  for (int j = 0; j < 16; ++j) c[acc_row(j) * 32 + acc_col()] = acc[j];
}

}  // namespace

hipError_t launch_to_bf16(const float* w, uint16_t* wb, int n, hipStream_t stream) {
  to_bf16_kernel<<<(n + 255) / 256, 256, 0, stream>>>(w, wb, n);
  return hipGetLastError();
}

hipError_t launch_forward(const uint16_t* wb, const float* b_fp32, const float* x, int64_t n, float* y,
                          hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int64_t groups = (n + 63) / 64;
  const unsigned grid = static_cast<unsigned>(groups < 8192 ? groups : 8192);
  forward_kernel<<<grid, 64, 0, stream>>>(wb, b_fp32, x, n, y);
  return hipGetLastError();
}

hipError_t launch_gradients(const uint16_t* wb, const float* b_fp32, const float* x, const float* target, int64_t n,
                            int64_t batch, float* partial, float* loss_partial, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int64_t groups = (n + 63) / 64;
  gradients_kernel<<<static_cast<unsigned>(groups), 256, 0, stream>>>(wb, b_fp32, x, target, n,
                                                                     2.0f / static_cast<float>(batch), partial,
                                                                     loss_partial);
  return hipGetLastError();
}

hipError_t launch_adam(const float* partial, int waves, float* grad, float* w, float* m, float* v, uint16_t* wb,
                       const Adam& a, hipStream_t stream) {
  const double bc1 = 1.0 - __builtin_pow(static_cast<double>(a.beta1), a.step);
  const double bc2 = 1.0 - __builtin_pow(static_cast<double>(a.beta2), a.step);
  adam_kernel<<<(kParams + 255) / 256, 256, 0, stream>>>(partial, waves, grad, w, m, v, wb, a.lr, a.beta1, a.beta2,
                                                         a.eps, static_cast<float>(bc1), static_cast<float>(bc2));
  return hipGetLastError();
}

hipError_t launch_tile_check(const uint16_t* a, const uint16_t* b, float* c, hipStream_t stream) {
  tile_check_kernel<<<1, 64, 0, stream>>>(a, b, c);
  return hipGetLastError();
}

}  // namespace mlp
}  // namespace bsx
