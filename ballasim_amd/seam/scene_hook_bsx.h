// The reference engine's RenderHook (src/engine/window.h:66-84) on this library's procedural scene:
// WindowEngine's request batches rendered by bsx_scene_render_requests. A request's time is the
// window's t_w plus its t_off_ns, as the engine's own hook adapter forms it
// (src/engine/bindings/engine_module.cpp:236). The scene has one light: group 0 receives the
// irradiance, further groups 0; the class word marks textured shapes (render/types.h kClassTextured).
// The engine's P&V outputs (render_pv, probe) are not provided and refuse as the base class does.
//
// BsxSceneHook<Space::kDevice> takes the engine's device arrays as they are. BsxSceneHook<Space::kHost>
// serves the host engine (src/engine/window.cpp): each batch goes to the device and its rows come back
// into the engine's out_index rows.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "bsx/bsx_scene.h"
#include "engine/window.h"

namespace balla::engine {

static_assert(sizeof(SampleRequest) == sizeof(bsx_sample_request), "the engine's 24-byte SampleRequest");

template <Space S>
class BsxSceneHook final : public RenderHook {
 public:
  BsxSceneHook(bsx_scene* scene, const bsx_view& view, int32_t groups, void* stream)
      : scene_(scene), view_(view), groups_(groups), stream_(stream) {
    if (scene == nullptr || groups < 1 || groups > 8) throw std::invalid_argument("BsxSceneHook: a scene and 1-8 groups");
  }
  ~BsxSceneHook() override {
    for (void* p : {d_req_, d_irr_, d_sig_, d_cls_}) (void)hipFree(p);
  }
  BsxSceneHook(const BsxSceneHook&) = delete;
  BsxSceneHook& operator=(const BsxSceneHook&) = delete;

  // The base of the requests' t_off_ns: the window the engine samples.
  void set_window(int64_t t_w) { t_w_ = t_w; }

  void render(SampleRequest* requests, size_t n, const uint32_t* out_index, float* irr, uint64_t* sig,
              uint32_t* klass, uint32_t /*subsamples: the scene is evaluated at the pixel centre*/) override {
    if (n == 0) return;
    if constexpr (S == Space::kDevice) {
      check(bsx_scene_render_requests(scene_, &view_, t_w_, reinterpret_cast<const bsx_sample_request*>(requests), n,
                                      out_index, groups_, irr, sig, klass, stream_));
    } else {
      stage(n);
      const auto st = static_cast<hipStream_t>(stream_);
      hip(hipMemcpyAsync(d_req_, requests, n * sizeof(bsx_sample_request), hipMemcpyHostToDevice, st));
      check(bsx_scene_render_requests(scene_, &view_, t_w_, static_cast<const bsx_sample_request*>(d_req_), n,
                                      nullptr, groups_, static_cast<float*>(d_irr_), static_cast<uint64_t*>(d_sig_),
                                      klass != nullptr ? static_cast<uint32_t*>(d_cls_) : nullptr, stream_));
      h_irr_.resize(n * static_cast<size_t>(groups_));
      h_sig_.resize(n);
      h_cls_.resize(n);
      hip(hipMemcpyAsync(h_irr_.data(), d_irr_, h_irr_.size() * sizeof(float), hipMemcpyDeviceToHost, st));
      hip(hipMemcpyAsync(h_sig_.data(), d_sig_, n * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
      if (klass != nullptr) hip(hipMemcpyAsync(h_cls_.data(), d_cls_, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      hip(hipStreamSynchronize(st));
      for (size_t i = 0; i < n; ++i) {
        const size_t row = out_index != nullptr ? out_index[i] : i;
        std::memcpy(irr + row * groups_, h_irr_.data() + i * groups_, static_cast<size_t>(groups_) * sizeof(float));
        sig[row] = h_sig_[i];
        if (klass != nullptr) klass[row] = h_cls_[i];
      }
    }
    ++calls;
    rendered += n;
  }

  size_t calls = 0, rendered = 0;  // render calls and requests served

 private:
  static void check(int status) {
    if (status == BSX_OK) return;
    char msg[512];
    bsx_last_error(msg, sizeof msg);
    throw std::runtime_error(std::string("BsxSceneHook: ") + msg);
  }
  static void hip(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(std::string("BsxSceneHook: ") + hipGetErrorString(e));
  }
  // Device staging for n requests (host engine only), grown as needed.
  void stage(size_t n) {
    if (n <= cap_) return;
    for (void** p : {&d_req_, &d_irr_, &d_sig_, &d_cls_}) {
      (void)hipFree(*p);
      *p = nullptr;
    }
    cap_ = n + n / 2;
    hip(hipMalloc(&d_req_, cap_ * sizeof(bsx_sample_request)));
    hip(hipMalloc(&d_irr_, cap_ * static_cast<size_t>(groups_) * sizeof(float)));
    hip(hipMalloc(&d_sig_, cap_ * sizeof(uint64_t)));
    hip(hipMalloc(&d_cls_, cap_ * sizeof(uint32_t)));
  }

  bsx_scene* scene_;
  bsx_view view_;
  int32_t groups_;
  void* stream_;
  int64_t t_w_ = 0;
  size_t cap_ = 0;
  void *d_req_ = nullptr, *d_irr_ = nullptr, *d_sig_ = nullptr, *d_cls_ = nullptr;
  std::vector<float> h_irr_;
  std::vector<uint64_t> h_sig_;
  std::vector<uint32_t> h_cls_;
};

}  // namespace balla::engine
