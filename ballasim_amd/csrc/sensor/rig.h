// A rig: several sensor models (cameras of any resolution) advanced by one call, whose chain of
// kernels covers every camera at once: one phase-A launch over all cameras' segments, one heavy-first
// order of the joint pixel space and one walk launch over all pixels, captured into a hipGraph and
// replayed. Each camera's events, state and statistics are those of advancing it alone.
//
// The reference engine advances its cameras one after the other on one stream
// (src/engine/engine.h:1-8, src/engine/engine_impl.h:1343-1348); a rig replaces that loop of
// per-camera advance_device calls by one launch, so that the ragged work of all cameras (a few heavy
// pixels and many light ones) is balanced over the whole device.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "chain.h"
#include "model.h"

namespace bsx {

struct RigTrial {
  WalkTuning tuning;
  double walk_ms;
};

class Rig {
 public:
  // The cameras are borrowed: each must outlive the rig. All share one photoreceptor kind.
  explicit Rig(std::vector<SensorModel*> cameras);

  // Advances every camera from its own model time to t_end on its runs (runs[c] for camera c; empty
  // runs hold the light). The events of camera c are left in its model's device buffer (out[c]). The
  // checks of every camera pass before any camera changes; a failure after them leaves every camera
  // uninitialized.
  void advance_device(const bsx_device_runs* runs, int64_t t_end_ns, hipStream_t stream, bsx_device_events* out);

  // Times the walk of the given call under candidate tunings (coordinate descent over block size,
  // walk budget, search batch and queue chunk from the current tuning, `reps` timed walks per
  // candidate, minimum taken), keeps the fastest and returns the trials. No camera changes.
  std::vector<RigTrial> autotune(const bsx_device_runs* runs, int64_t t_end_ns, hipStream_t stream, int reps);

  LaunchChain& chain() { return chain_; }
  size_t size() const { return cams_.size(); }
  double kernel_ms() const { return kernel_ms_; }

 private:
  // Checks every camera's runs (one synchronization), then stages their descriptors.
  std::vector<CameraDev> check_and_stage(const bsx_device_runs* runs, int64_t t_end_ns, hipStream_t stream);
  // After a launch (counters_ read back): a camera whose runs hold a value that is not finite refuses
  // the whole call, naming the camera; no camera has advanced.
  void refuse_non_finite();
  RunsDev device_runs(size_t c, const bsx_device_runs& r) const;

  std::vector<SensorModel*> cams_;
  LaunchChain chain_;
  std::vector<unsigned long long> counters_;  // (1 + kStCount) per camera
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  double kernel_ms_ = 0.0;

 public:
  ~Rig();
  Rig(const Rig&) = delete;
  Rig& operator=(const Rig&) = delete;
};

}  // namespace bsx
