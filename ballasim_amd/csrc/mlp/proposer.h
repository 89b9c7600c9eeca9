// Host side of the learned proposer: the fp32 master weights, Adam's moments and the bf16 copies in
// device memory, and the training loop over examples resident in device memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../sensor/model.h"
#include "mlp.h"

namespace bsx {

class Proposer {
 public:
  // Weights drawn from a seeded uniform (Glorot bounds), biases zero.
  explicit Proposer(uint64_t seed);
  void set_params(const float* w);  // host, kParams values
  void get_params(float* w) const;
  // The gradient of the mean squared error over n rows (device x [n x 12], t [n]) into grad (host,
  // kParams values); returns the mean squared error.
  double gradients(const float* x, const float* t, int64_t n, float* grad, hipStream_t stream);
  // `epochs` passes over the n rows in order, in batches of `batch` rows (a multiple of 64), one Adam
  // step per batch; returns the mean squared error of the last pass (measured during it).
  double train(const float* x, const float* t, int64_t n, int batch, int epochs, float lr, hipStream_t stream);
  void predict(const float* x, int64_t n, float* y, hipStream_t stream);
  const uint16_t* weights_bf16() const { return wb_.as<const uint16_t>(); }
  const float* weights() const { return w_.as<const float>(); }
  int64_t steps() const { return step_; }

 private:
  DeviceBuffer w_, m_, v_, wb_, grad_, partial_, loss_;
  int64_t step_ = 0;
};

}  // namespace bsx
