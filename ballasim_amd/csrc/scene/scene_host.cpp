// The scene object and its C ABI (include/bsx/bsx_scene.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/bsx/bsx_scene.h"
#include "../../../include/bsx/bsx_mlp.h"
#include "../handles.h"
#include "../mlp/proposer.h"
#include "../sensor/model.h"
#include "learned_kernels.h"
#include "scene_kernels.h"

namespace bsx {

namespace {

struct WindowSlot {
  DeviceBuffer counts, row_start, t, irr;
};

}  // namespace

class Scene {
 public:
  Scene(const bsx_shape* shapes, int32_t n, float background) : n_(n), background_(background) {
    if (n < 0 || n > kMaxShapes) throw std::invalid_argument("bsx_scene_create: 0 to 256 shapes");
    if (!(background > 0.0f) || !std::isfinite(background)) {
      throw std::invalid_argument("bsx_scene_create: background irradiance must be finite and > 0");
    }
    std::vector<DevShape> d(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
      const bsx_shape& s = shapes[i];
      if (!(s.hx > 0) || (s.kind == BSX_SHAPE_RECT && !(s.hy > 0)) || !(s.lux > 0.0f) || !(s.contrast >= 0.0f) ||
          s.contrast >= 1.0f || s.octaves < 0 || s.octaves > 16 || !(s.tex_cell_px > 0.0f) ||
          (s.kind != BSX_SHAPE_RECT && s.kind != BSX_SHAPE_DISK)) {
        throw std::invalid_argument("bsx_scene_create: shape " + std::to_string(i) + " is invalid");
      }
      DevShape& o = d[i];
      o.cx = s.cx;
      o.cy = s.cy;
      o.vx = s.vx;
      o.vy = s.vy;
      o.angle = s.angle;
      o.omega = s.omega;
      o.hx = static_cast<float>(s.hx);
      o.hy = static_cast<float>(s.kind == BSX_SHAPE_DISK ? s.hx : s.hy);
      o.radius = static_cast<float>(s.kind == BSX_SHAPE_DISK ? s.hx : std::sqrt(s.hx * s.hx + s.hy * s.hy));
      o.lux = s.lux;
      o.contrast = s.contrast;
      o.inv_cell = 1.0f / s.tex_cell_px;
      o.speed = static_cast<float>(std::sqrt(s.vx * s.vx + s.vy * s.vy) + std::fabs(s.omega) * o.radius);
      o.seed = s.tex_seed;
      o.kind = s.kind;
      o.octaves = s.octaves;
    }
    shapes_.ensure(std::max<size_t>(1, d.size()) * sizeof(DevShape));
    if (n > 0) hip_check(hipMemcpy(shapes_.get(), d.data(), d.size() * sizeof(DevShape), hipMemcpyHostToDevice), "shapes");
  }

  SceneDev dev() const { return SceneDev{shapes_.as<const DevShape>(), n_, background_}; }

  static ViewDev view_of(const bsx_view& v) {
    if (v.width == 0 || v.height == 0 || !(v.px_scale > 0)) throw std::invalid_argument("bsx_view: bad size or scale");
    return ViewDev{v.width, v.height, v.origin_x, v.origin_y, v.px_scale};
  }

  void render(const bsx_view& view, int64_t t, float* out) {
    const ViewDev v = view_of(view);
    const size_t P = static_cast<size_t>(v.width) * v.height;
    frame_.ensure(P * 4);
    hip_check(launch_render(dev(), v, t, frame_.as<float>(), nullptr), "render");
    hip_check(hipMemcpy(out, frame_.get(), P * 4, hipMemcpyDeviceToHost), "download frame");
  }

  WindowSlot& slot_at(int32_t slot) {
    if (slot < 0 || slot > 100000) throw std::invalid_argument("sample_window: slot out of range");
    if (static_cast<size_t>(slot) >= slots_.size()) slots_.resize(static_cast<size_t>(slot) + 1);
    if (!slots_[slot]) slots_[slot] = std::make_unique<WindowSlot>();
    return *slots_[slot];
  }

  // The pixel ids 0 .. P - 1 of a view's runs. Runs handed out earlier keep pointing at the buffer they
  // were given, so a larger view gets a new buffer and the old ones stay alive with the scene.
  void ensure_ids(size_t P) {
    if (ids_.empty() || ids_.back()->bytes() < P * 4) {
      std::vector<uint32_t> iota(P);
      for (size_t i = 0; i < P; ++i) iota[i] = static_cast<uint32_t>(i);
      ids_.push_back(std::make_unique<DeviceBuffer>());
      ids_.back()->ensure(P * 4);
      hip_check(hipMemcpy(ids_.back()->get(), iota.data(), P * 4, hipMemcpyHostToDevice), "ids");
    }
  }

  bsx_device_runs runs_of(WindowSlot& w, size_t P, int64_t total) const {
    bsx_device_runs out{};
    out.pixel = ids_.back()->as<const uint32_t>();
    out.row_start = w.row_start.as<const int64_t>();
    out.row_count = w.counts.as<const int32_t>();
    out.n = P;
    out.t = w.t.as<const int64_t>();
    out.irr = w.irr.as<const float>();
    out.n_samples = static_cast<size_t>(total);
    return out;
  }

  int64_t* scan_scratch(size_t P) {
    scan_.ensure(scan_scratch_count(P) * 8);
    return scan_.as<int64_t>();
  }

  bsx_device_runs sample_counts(const bsx_view& view, const int32_t* counts, int64_t t_w, int32_t slot,
                                hipStream_t stream, uint64_t* total_out) {
    const ViewDev v = view_of(view);
    if ((t_w & (kWindowNs - 1)) != 0) throw std::invalid_argument("sample_counts: t_w must be a multiple of 2^23 ns");
    const size_t P = static_cast<size_t>(v.width) * v.height;
    int64_t total = 0;
    for (size_t p = 0; p < P; ++p) {
      if (counts[p] < 0 || counts[p] > kWindowNs) {
        throw std::invalid_argument("sample_counts: a pixel's count must lie in [0, 2^23]");
      }
      total += counts[p];
    }
    WindowSlot& w = slot_at(slot);
    w.counts.ensure(P * 4);
    w.row_start.ensure(P * 8);
    scan_.ensure(scan_scratch_count(P) * 8);
    ensure_ids(P);
    w.t.ensure(std::max<int64_t>(total, 1) * 8);
    w.irr.ensure(std::max<int64_t>(total, 1) * 4);
    hip_check(hipMemcpyAsync(w.counts.get(), counts, P * 4, hipMemcpyHostToDevice, stream), "counts");
    hip_check(launch_exclusive_scan(w.counts.as<int32_t>(), P, w.row_start.as<int64_t>(), scan_.as<int64_t>(), stream),
              "scan");
    hip_check(launch_counts_write(dev(), v, t_w, w.counts.as<const int32_t>(), w.row_start.as<const int64_t>(), total,
                                  w.t.as<int64_t>(), w.irr.as<float>(), stream),
              "write samples");
    hip_check(hipStreamSynchronize(stream), "sample counts");
    if (total_out != nullptr) *total_out = static_cast<uint64_t>(total);
    return runs_of(w, P, total);
  }

  bsx_device_runs sample(const bsx_view& view, const bsx_sampler_rules& r, int64_t t_w, int32_t slot,
                         hipStream_t stream, uint64_t* total_out) {
    const ViewDev v = view_of(view);
    if ((t_w & (kWindowNs - 1)) != 0) throw std::invalid_argument("sample_window: t_w must be a multiple of 2^23 ns");
    if (r.fixed_period_ns < 0 || (r.fixed_period_ns == 0 && !(r.d_px > 0)) || r.l_min < 0 || r.l_max > 13 ||
        r.l_min > r.l_max || r.max_refine_levels < 0 || r.max_refine_levels > 7) {
      throw std::invalid_argument("sample_window: rules need d_px > 0, 0 <= l_min <= l_max <= 13, "
                                  "0 <= max_refine_levels <= 7");
    }
    if (r.fixed_period_ns > 0 && r.fixed_period_ns < 64) {
      throw std::invalid_argument("sample_window: fixed_period_ns must be >= 64");
    }
    WindowSlot& w = slot_at(slot);
    const size_t P = static_cast<size_t>(v.width) * v.height;
    w.counts.ensure(P * 4);
    w.row_start.ensure(P * 8);
    scan_.ensure(scan_scratch_count(P) * 8);
    ensure_ids(P);
    const uint32_t cap = capture_slots(P);
    if (cap > 0 && reserve_capture(P, cap)) return sample_once(v, r, t_w, w, P, cap, stream, total_out);
    hip_check(launch_count(dev(), v, r, t_w, w.counts.as<int32_t>(), stream), "count samples");
    hip_check(launch_exclusive_scan(w.counts.as<int32_t>(), P, w.row_start.as<int64_t>(), scan_.as<int64_t>(), stream),
              "scan");
    int64_t total = 0;
    const size_t blocks = scan_scratch_count(P) - 1;
    hip_check(hipMemcpyAsync(&total, scan_.as<int64_t>() + blocks, 8, hipMemcpyDeviceToHost, stream), "total");
    hip_check(hipStreamSynchronize(stream), "sample window");
    w.t.ensure(std::max<int64_t>(total, 1) * 8);
    w.irr.ensure(std::max<int64_t>(total, 1) * 4);
    hip_check(launch_write(dev(), v, r, t_w, w.row_start.as<int64_t>(), w.t.as<int64_t>(), w.irr.as<float>(), stream),
              "write samples");
    hip_check(hipStreamSynchronize(stream), "sample window");
    if (total_out != nullptr) *total_out = static_cast<uint64_t>(total);
    return runs_of(w, P, total);
  }

  // Device memory the one-pass sampler may hold for its capture slots (0: the two-pass sampler).
  void set_capture_budget(uint64_t bytes) {
    capture_budget_ = bytes;
    if (capture_.bytes() > bytes) {
      DeviceBuffer none;
      capture_.swap(none);
    }
  }
  uint64_t capture_budget() const { return capture_budget_; }
  // Capture slots per pixel for a view of P pixels: as many as the budget holds, at most kMaxCaptureSlots.
  uint32_t capture_slots(size_t P) const {
    const uint64_t k = capture_budget_ / (static_cast<uint64_t>(P) * sizeof(CapturedSample));
    return static_cast<uint32_t>(std::min<uint64_t>(k, kMaxCaptureSlots));
  }

 private:
  static constexpr uint32_t kMaxCaptureSlots = 1024;

  // One pass over the scene: every pixel's samples counted and its first `cap` captured; the counts
  // scanned; the captured runs copied into place and the pixels with more samples sampled again there.
  // The capture slots and the largest-count word; when the device cannot hold them, the budget drops to
  // 0 and this scene samples in two passes from then on.
  bool reserve_capture(size_t P, uint32_t cap) {
    try {
      capture_.ensure(P * cap * sizeof(CapturedSample) + 256);
      return true;
    } catch (const std::runtime_error&) {
      (void)hipGetLastError();
      set_capture_budget(0);
      return false;
    }
  }

  bsx_device_runs sample_once(const ViewDev& v, const bsx_sampler_rules& r, int64_t t_w, WindowSlot& w, size_t P,
                              uint32_t cap, hipStream_t stream, uint64_t* total_out) {
    auto* scratch = capture_.as<CapturedSample>();
    auto* max_count = reinterpret_cast<unsigned int*>(capture_.as<uint8_t>() + P * cap * sizeof(CapturedSample));
    hip_check(launch_capture(dev(), v, r, t_w, cap, w.counts.as<int32_t>(), scratch, max_count, stream),
              "capture samples");
    hip_check(launch_exclusive_scan(w.counts.as<int32_t>(), P, w.row_start.as<int64_t>(), scan_.as<int64_t>(), stream),
              "scan");
    int64_t total = 0;
    uint32_t most = 0;
    const size_t blocks = scan_scratch_count(P) - 1;
    hip_check(hipMemcpyAsync(&total, scan_.as<int64_t>() + blocks, 8, hipMemcpyDeviceToHost, stream), "total");
    hip_check(hipMemcpyAsync(&most, max_count, 4, hipMemcpyDeviceToHost, stream), "largest count");
    hip_check(hipStreamSynchronize(stream), "sample window");
    w.t.ensure(std::max<int64_t>(total, 1) * 8);
    w.irr.ensure(std::max<int64_t>(total, 1) * 4);
    hip_check(launch_compact(v, t_w, cap, w.counts.as<const int32_t>(), w.row_start.as<const int64_t>(), scratch,
                             w.t.as<int64_t>(), w.irr.as<float>(), stream),
              "place samples");
    if (most > cap) {
      hip_check(launch_overflow_write(dev(), v, r, t_w, cap, w.counts.as<const int32_t>(),
                                      w.row_start.as<const int64_t>(), w.t.as<int64_t>(), w.irr.as<float>(), stream),
                "write samples");
      ++overflow_windows_;
    }
    hip_check(hipStreamSynchronize(stream), "sample window");
    if (total_out != nullptr) *total_out = static_cast<uint64_t>(total);
    return runs_of(w, P, total);
  }

 public:
  uint64_t overflow_windows() const { return overflow_windows_; }

 private:
  int32_t n_;
  float background_;
  uint64_t capture_budget_ = uint64_t{16} << 30;
  uint64_t overflow_windows_ = 0;
  DeviceBuffer capture_;
  DeviceBuffer shapes_, frame_, scan_;
  std::vector<std::unique_ptr<DeviceBuffer>> ids_;
  std::vector<std::unique_ptr<WindowSlot>> slots_;
};

// Training examples of the proposer from dense traces of the scene (see learned.hip).
void learned_examples(Scene& sc, const bsx_view& view, const bsx_learned_examples_spec& sp, float* x, float* target,
                      hipStream_t stream, uint64_t* n_active_out) {
  const ViewDev v = Scene::view_of(view);
  const size_t P = static_cast<size_t>(v.width) * v.height;
  if (sp.n_dense < sp.max_steps + 3 * 256 + 2 || sp.max_steps < 1 || sp.delta_ns < 1 || sp.n < 0 ||
      !(sp.theta_ref > 0.0f) || !(sp.theta_band > 0.0f) || !(sp.tol_log > 0.0f)) {
    throw std::invalid_argument("learned_examples: need n_dense >= max_steps + 770, max_steps >= 1, delta >= 1, "
                                "n >= 0 and positive thresholds and tolerance");
  }
  DeviceBuffer trace, ref, active, pixels;
  trace.ensure(P * sp.n_dense * 4);
  ref.ensure(P * sp.n_dense * 4);
  active.ensure(P);
  hip_check(launch_dense_trace(sc.dev(), v, sp.t0_ns, sp.delta_ns, sp.n_dense, sp.dark, sp.theta_ref,
                               trace.as<float>(), ref.as<float>(), active.as<uint8_t>(), stream),
            "dense trace");
  std::vector<uint8_t> act(P);
  hip_check(hipMemcpyAsync(act.data(), active.get(), P, hipMemcpyDeviceToHost, stream), "active pixels");
  hip_check(hipStreamSynchronize(stream), "dense trace");
  // Candidates: every varying pixel, and constant ones in a tenth of their number (spread evenly).
  std::vector<uint32_t> cand, quiet;
  for (size_t p = 0; p < P; ++p) (act[p] ? cand : quiet).push_back(static_cast<uint32_t>(p));
  if (n_active_out != nullptr) *n_active_out = cand.size();
  const size_t extra = cand.empty() ? quiet.size() : std::min(quiet.size(), cand.size() / 10);
  for (size_t k = 0; k < extra; ++k) cand.push_back(quiet[k * quiet.size() / std::max<size_t>(extra, 1)]);
  if (cand.empty()) throw std::invalid_argument("learned_examples: the view has no pixels");
  pixels.ensure(cand.size() * 4);
  hip_check(hipMemcpyAsync(pixels.get(), cand.data(), cand.size() * 4, hipMemcpyHostToDevice, stream), "pixels");
  hip_check(launch_examples(trace.as<const float>(), ref.as<const float>(), static_cast<uint32_t>(P), sp.n_dense,
                            pixels.as<const uint32_t>(), static_cast<uint32_t>(cand.size()), sp.delta_ns,
                            sp.max_steps, sp.theta_band, sp.tol_log, sp.seed, sp.n, x, target, stream),
            "examples");
  hip_check(hipStreamSynchronize(stream), "examples");
}

// The learned sampler of one view: the per-pixel controller state and the window passes.
class LearnedSampler {
 public:
  LearnedSampler(Scene& sc, const bsx_view& view, const bsx_learned_rules& r, Proposer& p)
      : sc_(sc), view_(view), p_(p) {
    const ViewDev v = Scene::view_of(view);
    const bsx_sampler_rules& q = r.plain;
    if (!(q.refine_tol_log > 0.0f) || q.max_refine_levels < 0 || q.max_refine_levels > 7 || !(r.tol_log > 0.0f) ||
        !(r.theta_band > 0.0f) || !(r.theta_ref > 0.0f) || !(r.margin > 0.0f) || r.min_dt_ns < 2) {
      throw std::invalid_argument("learned sampler: need positive tolerances, thresholds and margin, "
                                  "0 <= max_refine_levels <= 7 and min_dt_ns >= 2");
    }
    if (!(r.near >= 0.0f) || !(q.d_px > 0) || q.l_min < 0 || q.l_max > 13 || q.l_min > q.l_max) {
      throw std::invalid_argument("learned sampler: need near >= 0, d_px > 0 and 0 <= l_min <= l_max <= 13");
    }
    r_ = LearnedRules{q.dark,      q.refine_tol_log, r.tol_log, r.theta_band, r.theta_ref, r.margin,
                      q.max_refine_levels, q.refine_signature, static_cast<int64_t>(q.min_split_ns),
                      r.min_dt_ns, r.near, q.l_min, q.l_max, q.d_px};
    P_ = static_cast<size_t>(v.width) * v.height;
    t_.ensure(P_ * 4 * 8);
    L_.ensure(P_ * 4 * 4);
    n_.ensure(P_ * 4);
    ref_.ensure(P_ * 4);
    e0_.ensure(P_ * 4);
    sig_.ensure(P_ * 4);
    work_.ensure(16);
  }

  LearnedStateDev state() const {
    return LearnedStateDev{t_.as<int64_t>(), L_.as<float>(), n_.as<int32_t>(), ref_.as<float>(), e0_.as<float>(),
                           sig_.as<uint32_t>()};
  }

  void init(int64_t t0, hipStream_t stream) {
    hip_check(launch_learned_init(sc_.dev(), Scene::view_of(view_), t0, r_.dark, state(), stream), "learned init");
    hip_check(hipStreamSynchronize(stream), "learned init");
    t_init_ = t0;
    initialized_ = true;
  }

  bsx_device_runs window(int64_t t_w, int32_t slot, hipStream_t stream, uint64_t* total_out, uint64_t* proposals,
                         uint64_t* fallbacks, const SensorModel* sensor) {
    if (!initialized_) throw std::logic_error("learned sampler: call init first");
    SensorLink link{};
    if (sensor != nullptr) {
      const bsx_sensor_info_t info = sensor->info();
      if (static_cast<size_t>(info.width) * info.height != P_) {
        throw std::invalid_argument("learned sampler: the sensor's size differs from the view's");
      }
      if (sensor->params().photoreceptor != BSX_PHOTORECEPTOR_INSTANTANEOUS) {
        throw std::invalid_argument("learned sampler: only an instantaneous-response sensor can be linked");
      }
      const CameraDev cd = sensor->current_view();
      link = SensorLink{cd.mp.theta_on, cd.mp.theta_off, cd.mp.scale,   cd.in.i_scale,       cd.in.s3_ref,
                        cd.in.n3_ref,   cd.in.dr_ref,    cd.in.blind_until, cd.kc.eps_ci, cd.kc.ideal_reset,
                        cd.kc.refractory_ns};
    }
    if ((t_w & (kWindowNs - 1)) != 0) throw std::invalid_argument("learned sampler: t_w must be a multiple of 2^23 ns");
    const ViewDev v = Scene::view_of(view_);
    WindowSlot& w = sc_.slot_at(slot);
    w.counts.ensure(P_ * 4);
    w.row_start.ensure(P_ * 8);
    sc_.ensure_ids(P_);
    int64_t* scan = sc_.scan_scratch(P_);
    hip_check(hipMemsetAsync(work_.get(), 0, 16, stream), "zero work");
    hip_check(launch_learned_window(sc_.dev(), v, r_, p_.weights_bf16(), p_.weights(), t_w, state(), link,
                                    w.counts.as<int32_t>(), nullptr, nullptr, nullptr,
                                    work_.as<unsigned long long>(), false, stream),
              "learned count");
    hip_check(launch_exclusive_scan(w.counts.as<int32_t>(), P_, w.row_start.as<int64_t>(), scan, stream), "scan");
    int64_t total = 0;
    unsigned long long wk[2] = {0, 0};
    hip_check(hipMemcpyAsync(&total, scan + scan_scratch_count(P_) - 1, 8, hipMemcpyDeviceToHost, stream), "total");
    hip_check(hipMemcpyAsync(wk, work_.get(), 16, hipMemcpyDeviceToHost, stream), "work");
    hip_check(hipStreamSynchronize(stream), "learned count");
    w.t.ensure(std::max<int64_t>(total, 1) * 8);
    w.irr.ensure(std::max<int64_t>(total, 1) * 4);
    hip_check(launch_learned_window(sc_.dev(), v, r_, p_.weights_bf16(), p_.weights(), t_w, state(), link, nullptr,
                                    w.row_start.as<const int64_t>(), w.t.as<int64_t>(), w.irr.as<float>(), nullptr,
                                    true, stream),
              "learned write");
    hip_check(hipStreamSynchronize(stream), "learned write");
    if (total_out != nullptr) *total_out = static_cast<uint64_t>(total);
    if (proposals != nullptr) *proposals = wk[0];
    if (fallbacks != nullptr) *fallbacks = wk[1];
    return sc_.runs_of(w, P_, total);
  }

 private:
  Scene& sc_;
  bsx_view view_;
  Proposer& p_;
  LearnedRules r_{};
  size_t P_ = 0;
  DeviceBuffer t_, L_, n_, ref_, e0_, sig_, work_;
  int64_t t_init_ = 0;
  bool initialized_ = false;
};

}  // namespace bsx


// Status mapping shared with capi.cpp's conventions.
extern "C" int bsx_internal_set_error(int code, const char* msg);

template <class F>
static int scene_guarded(F&& f) {
  try {
    f();
    return BSX_OK;
  } catch (const std::invalid_argument& e) {
    return bsx_internal_set_error(BSX_E_INVALID, e.what());
  } catch (const std::logic_error& e) {
    return bsx_internal_set_error(BSX_E_LOGIC, e.what());
  } catch (const std::exception& e) {
    return bsx_internal_set_error(BSX_E_RUNTIME, e.what());
  }
}

extern "C" {

int bsx_scene_create(const bsx_shape* shapes, int32_t n_shapes, float background, bsx_scene** out) {
  return scene_guarded([&] {
    if (out == nullptr || (n_shapes > 0 && shapes == nullptr)) throw std::invalid_argument("bsx_scene_create: null");
    *out = nullptr;
    auto* s = new bsx::Scene(shapes, n_shapes, background);
    *out = new bsx_scene{s};
  });
}

int bsx_scene_destroy(bsx_scene* s) {
  return scene_guarded([&] {
    if (s == nullptr) return;
    delete s->s;
    delete s;
  });
}

int bsx_copy_to_host(void* dst, const void* src, size_t bytes) {
  return scene_guarded([&] {
    if (bytes == 0) return;
    if (dst == nullptr || src == nullptr) throw std::invalid_argument("bsx_copy_to_host: null");
    bsx::hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "copy to host");
  });
}

int bsx_scene_render(bsx_scene* s, const bsx_view* view, int64_t t_ns, float* out) {
  return scene_guarded([&] {
    if (s == nullptr || view == nullptr || out == nullptr) throw std::invalid_argument("bsx_scene_render: null");
    s->s->render(*view, t_ns, out);
  });
}

int bsx_scene_sample_window(bsx_scene* s, const bsx_view* view, const bsx_sampler_rules* rules, int64_t t_w,
                            int32_t slot, void* stream, bsx_device_runs* out, uint64_t* total_samples) {
  return scene_guarded([&] {
    if (s == nullptr || view == nullptr || rules == nullptr || out == nullptr) {
      throw std::invalid_argument("bsx_scene_sample_window: null");
    }
    *out = s->s->sample(*view, *rules, t_w, slot, static_cast<hipStream_t>(stream), total_samples);
  });
}

int bsx_scene_set_capture_budget(bsx_scene* s, uint64_t bytes) {
  return scene_guarded([&] {
    if (s == nullptr) throw std::invalid_argument("bsx_scene_set_capture_budget: null");
    s->s->set_capture_budget(bytes);
  });
}

int bsx_scene_capture_budget(bsx_scene* s, uint64_t* bytes, uint64_t* overflow_windows) {
  return scene_guarded([&] {
    if (s == nullptr) throw std::invalid_argument("bsx_scene_capture_budget: null");
    if (bytes != nullptr) *bytes = s->s->capture_budget();
    if (overflow_windows != nullptr) *overflow_windows = s->s->overflow_windows();
  });
}

int bsx_scene_sample_counts(bsx_scene* s, const bsx_view* view, const int32_t* counts, int64_t t_w, int32_t slot,
                            void* stream, bsx_device_runs* out, uint64_t* total_samples) {
  return scene_guarded([&] {
    if (s == nullptr || view == nullptr || counts == nullptr || out == nullptr) {
      throw std::invalid_argument("bsx_scene_sample_counts: null");
    }
    *out = s->s->sample_counts(*view, counts, t_w, slot, static_cast<hipStream_t>(stream), total_samples);
  });
}

int bsx_scene_learned_examples(bsx_scene* s, const bsx_view* view, const bsx_learned_examples_spec* spec, float* x,
                                float* target, void* stream, uint64_t* n_active_pixels) {
  return scene_guarded([&] {
    if (s == nullptr || view == nullptr || spec == nullptr || (spec->n > 0 && (x == nullptr || target == nullptr))) {
      throw std::invalid_argument("bsx_scene_learned_examples: null");
    }
    bsx::learned_examples(*s->s, *view, *spec, x, target, static_cast<hipStream_t>(stream), n_active_pixels);
  });
}

int bsx_learned_sampler_create(bsx_scene* s, const bsx_view* view, const bsx_learned_rules* rules, bsx_mlp* proposer,
                               bsx_learned_sampler** out) {
  return scene_guarded([&] {
    if (s == nullptr || view == nullptr || rules == nullptr || proposer == nullptr || proposer->p == nullptr ||
        out == nullptr) {
      throw std::invalid_argument("bsx_learned_sampler_create: null");
    }
    *out = nullptr;
    auto* ls = new bsx::LearnedSampler(*s->s, *view, *rules, *proposer->p);
    *out = new bsx_learned_sampler{ls};
  });
}

int bsx_learned_sampler_destroy(bsx_learned_sampler* ls) {
  return scene_guarded([&] {
    if (ls == nullptr) return;
    delete ls->ls;
    delete ls;
  });
}

int bsx_learned_sampler_init(bsx_learned_sampler* ls, int64_t t0_ns, void* stream) {
  return scene_guarded([&] {
    if (ls == nullptr || ls->ls == nullptr) throw std::invalid_argument("bsx_learned_sampler_init: null");
    ls->ls->init(t0_ns, static_cast<hipStream_t>(stream));
  });
}

int bsx_learned_sampler_window(bsx_learned_sampler* ls, int64_t t_w, int32_t slot, bsx_sensor* sensor, void* stream,
                               bsx_device_runs* out, uint64_t* total_samples, uint64_t* proposals, uint64_t* fallbacks) {
  return scene_guarded([&] {
    if (ls == nullptr || ls->ls == nullptr || out == nullptr) throw std::invalid_argument("bsx_learned_sampler_window: null");
    if (sensor != nullptr && sensor->model == nullptr) throw std::invalid_argument("bsx_learned_sampler_window: null sensor");
    *out = ls->ls->window(t_w, slot, static_cast<hipStream_t>(stream), total_samples, proposals, fallbacks,
                          sensor != nullptr ? sensor->model : nullptr);
  });
}

int bsx_scene_render_requests(bsx_scene* s, const bsx_view* view, int64_t t_base_ns, const bsx_sample_request* requests,
                              size_t n, const uint32_t* out_index, int32_t groups, float* irr, uint64_t* sig,
                              uint32_t* klass, void* stream) {
  return scene_guarded([&] {
    if (s == nullptr || view == nullptr) throw std::invalid_argument("bsx_scene_render_requests: null");
    if (groups < 1 || groups > BSX_MAX_LIGHT_GROUPS) {
      throw std::invalid_argument("bsx_scene_render_requests: groups must lie in [1, 8]");
    }
    if (n == 0) return;
    if (requests == nullptr || irr == nullptr || sig == nullptr) {
      throw std::invalid_argument("bsx_scene_render_requests: null requests or outputs");
    }
    const bsx::ViewDev v = bsx::Scene::view_of(*view);
    bsx::hip_check(bsx::launch_requests(s->s->dev(), v, t_base_ns, reinterpret_cast<const bsx::SampleRequestDev*>(requests),
                                        n, out_index, groups, irr, sig, klass, static_cast<hipStream_t>(stream)),
                   "render requests");
  });
}

}  // extern "C"
