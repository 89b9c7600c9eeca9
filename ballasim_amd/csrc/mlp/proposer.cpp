// Host side of the learned proposer.
#include "proposer.h"

#include <cmath>
#include <stdexcept>
#include <vector>

namespace bsx {

namespace {

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace

Proposer::Proposer(uint64_t seed) {
  using namespace mlp;
  std::vector<float> w(kParams, 0.0f);
  uint64_t s = seed;
  const auto uni = [&](double bound) {
    return static_cast<float>(bound * (2.0 * (static_cast<double>(splitmix(s) >> 11) * 0x1p-53) - 1.0));
  };
  for (int i = 0; i < kIn * kHid; ++i) w[kOffW1 + i] = uni(std::sqrt(6.0 / (kIn + kHid)));
  for (int i = 0; i < kHid * kHid; ++i) w[kOffW2 + i] = uni(std::sqrt(6.0 / (2 * kHid)));
  for (int i = 0; i < kHid; ++i) w[kOffW3 + i] = uni(std::sqrt(6.0 / (kHid + 1)));
  w_.ensure(kParams * 4);
  m_.ensure(kParams * 4);
  v_.ensure(kParams * 4);
  wb_.ensure(kParams * 2);
  grad_.ensure(kParams * 4);
  set_params(w.data());
}

void Proposer::set_params(const float* w) {
  using namespace mlp;
  hip_check(hipMemcpy(w_.get(), w, kParams * 4, hipMemcpyHostToDevice), "upload weights");
  hip_check(hipMemset(m_.get(), 0, kParams * 4), "zero moments");
  hip_check(hipMemset(v_.get(), 0, kParams * 4), "zero moments");
  hip_check(launch_to_bf16(w_.as<const float>(), wb_.as<uint16_t>(), kParams, nullptr), "bf16 weights");
  hip_check(hipDeviceSynchronize(), "set params");
  step_ = 0;
}

void Proposer::get_params(float* w) const {
  hip_check(hipMemcpy(w, w_.get(), mlp::kParams * 4, hipMemcpyDeviceToHost), "download weights");
}

double Proposer::gradients(const float* x, const float* t, int64_t n, float* grad, hipStream_t stream) {
  using namespace mlp;
  if (n <= 0) throw std::invalid_argument("Proposer.gradients: need at least one row");
  const int64_t groups = (n + 63) / 64;
  partial_.ensure(static_cast<size_t>(groups) * kParams * 4);
  loss_.ensure(static_cast<size_t>(groups) * 4);
  hip_check(launch_gradients(wb_.as<const uint16_t>(), w_.as<const float>(), x, t, n, n, partial_.as<float>(),
                             loss_.as<float>(), stream),
            "gradients");
  const Adam none{0.0f, 0.9f, 0.999f, 1e-8f, 1};
  hip_check(launch_adam(partial_.as<const float>(), static_cast<int>(groups), grad_.as<float>(), w_.as<float>(),
                        m_.as<float>(), v_.as<float>(), wb_.as<uint16_t>(), none, stream),
            "reduce gradients");
  std::vector<float> lp(groups);
  hip_check(hipMemcpyAsync(grad, grad_.get(), kParams * 4, hipMemcpyDeviceToHost, stream), "download gradient");
  hip_check(hipMemcpyAsync(lp.data(), loss_.get(), groups * 4, hipMemcpyDeviceToHost, stream), "download loss");
  hip_check(hipStreamSynchronize(stream), "gradients");
  double s = 0.0;
  for (const float v : lp) s += v;
  return s / static_cast<double>(n);
}

double Proposer::train(const float* x, const float* t, int64_t n, int batch, int epochs, float lr, hipStream_t stream) {
  using namespace mlp;
  if (n <= 0 || batch <= 0 || batch % 64 != 0 || epochs < 1 || !(lr > 0.0f)) {
    throw std::invalid_argument("Proposer.train: need rows, a batch that is a positive multiple of 64, epochs >= 1 "
                                "and lr > 0");
  }
  const int64_t groups = (n + 63) / 64;
  partial_.ensure(static_cast<size_t>(batch / 64) * kParams * 4);
  loss_.ensure(static_cast<size_t>(groups) * 4);
  for (int e = 0; e < epochs; ++e) {
    for (int64_t s = 0; s < n; s += batch) {
      const int64_t nb = std::min<int64_t>(batch, n - s);
      hip_check(launch_gradients(wb_.as<const uint16_t>(), w_.as<const float>(), x + s * kIn, t + s, nb, nb,
                                 partial_.as<float>(), loss_.as<float>() + s / 64, stream),
                "gradients");
      const Adam a{lr, 0.9f, 0.999f, 1e-8f, static_cast<int32_t>(++step_)};
      hip_check(launch_adam(partial_.as<const float>(), static_cast<int>((nb + 63) / 64), grad_.as<float>(),
                            w_.as<float>(), m_.as<float>(), v_.as<float>(), wb_.as<uint16_t>(), a, stream),
                "adam");
    }
  }
  std::vector<float> lp(groups);
  hip_check(hipMemcpyAsync(lp.data(), loss_.get(), groups * 4, hipMemcpyDeviceToHost, stream), "download loss");
  hip_check(hipStreamSynchronize(stream), "train");
  double s = 0.0;
  for (const float v : lp) s += v;
  return s / static_cast<double>(n);
}

void Proposer::predict(const float* x, int64_t n, float* y, hipStream_t stream) {
  hip_check(mlp::launch_forward(wb_.as<const uint16_t>(), w_.as<const float>(), x, n, y, stream), "forward");
}

}  // namespace bsx
