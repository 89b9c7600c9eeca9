// SensorModel on one MI355X: the host side of the sensor path.
//
// It owns the frozen maps (drawn on the host in fp64, kept on the device in the kernel's fp32), the
// per-pixel working state (two structure-of-arrays copies in HBM: a call reads one and writes the
// other, so an event-buffer overflow re-runs from the untouched input state without a backup copy),
// the event buffer and the launch loop. The member functions mirror the reference's SensorModel
// (src/sensor/sensor_model.h:98-320) for the device kCuda / precision kFp32 configuration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "kconst.h"
#include "maps.h"
#include "params.h"

namespace bsx {

class LaunchChain;
class Rig;

// Grow-only device allocation.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer();
  void* ensure(size_t bytes);  // keeps the old contents only when no growth is needed
  void* get() const { return p_; }
  size_t bytes() const { return bytes_; }
  void swap(DeviceBuffer& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(bytes_, o.bytes_);
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p_);
  }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

void hip_check(hipError_t e, const char* what);

struct HostEvents {
  std::vector<uint16_t> x, y;
  std::vector<uint8_t> p, label;
  std::vector<int64_t> t_cross, t_reported;
  std::vector<uint16_t> flags;
  std::vector<float> ds, dn, dl, l_first;
  void resize(size_t n);
  size_t size() const { return t_cross.size(); }
};

// The reference's saved per-pixel state (src/sensor/pixel.h:105-121), 200 bytes: the format of
// save_state / load_state.
struct SavedPixelState {
  double z, x1, s2, s3;
  double i_scale;
  double tv;
  double nk[3], nk1[3];
  double ln_a1;
  int64_t cell;
  int64_t t_ref;
  double s3_ref, n3_ref, dr_ref;
  int64_t blind_until;
  int64_t t_search;
  float held[8];
  uint16_t flags;
  uint8_t have_k1;
  uint8_t reserved;
  float rho_m1;
};
static_assert(sizeof(SavedPixelState) == 200, "the saved-state layout is the reference's");

class SensorModel {
 public:
  SensorModel(const SensorParams& params, uint32_t width, uint32_t height, uint64_t seed,
              std::vector<bsx_waveform> groups, uint32_t camera);
  // The reference's backend seam (CudaSensorKernel(kc, maps, tables), src/sensor/sensor_cuda.h:22-23):
  // a model of n_pixels pixels from the reference's fp32 kernel constants, its per-pixel maps and its
  // fp32 noise tables. Its states come from upload_states; set_time gives each call's model time.
  SensorModel(const bsx_kernel_constants& kc, const bsx_pixel_map* maps, size_t n_pixels, const float* tables,
              size_t n_table_values);
  ~SensorModel();
  SensorModel(const SensorModel&) = delete;
  SensorModel& operator=(const SensorModel&) = delete;

  void set_photometry(double illuminance_scale, const float* gain, size_t n_gain);
  void init(int64_t t0_ns, const float* irradiance0);
  void advance(const uint32_t* ids, size_t n, const int64_t* t, const float* irr, size_t m, const int32_t* counts,
               int64_t t_end_ns);
  void advance_runs(const uint32_t* ids, size_t n, const int64_t* t, const float* irr, size_t n_samples,
                    const int64_t* row_start, const int32_t* row_count, int64_t t_end_ns);
  bsx_device_events advance_device(const bsx_device_runs& runs, int64_t t_end_ns, hipStream_t stream);
  // The input each pixel would hold after a call on these runs (runs null or empty: the held inputs),
  // pixels x groups into out (host), after the work queued on stream; the model is not changed.
  void held_frame(const bsx_device_runs* runs, hipStream_t stream, float* out);
  // advance_device for callers that cannot allocate: the call's scratch (run index, checks,
  // counters, descriptors) lives in the library's static device scratch, the raw events go to the
  // caller's device buffer of cap events, and nothing is allocated. The walk runs without phase A and
  // without the heavy-first order (same results). More events than cap: std::invalid_argument, the
  // state unchanged. Returns the event count.
  uint64_t advance_device_noalloc(const bsx_device_runs& runs, int64_t t_end_ns, hipStream_t stream,
                                  RawEventDev* events, uint64_t cap);
  void download_events(const bsx_device_events& ev);
  const HostEvents& events() const { return events_; }

  size_t state_bytes() const;
  void save_state(uint8_t* out, size_t n) const;
  void load_state(const uint8_t* data, size_t n);
  // The reference's per-pixel states (200 B each) to and from the device's working state.
  void upload_states(const SavedPixelState* states, size_t n);
  void download_states(SavedPixelState* out, size_t n) const;
  // The seam's calls: the model time of the next call (after upload_states), and new photometry.
  void set_time(int64_t t_state_ns);
  void set_kernel_photometry(float scale, float irr_max, const bsx_pixel_map* maps, size_t n);
  // The last host call's raw events, sorted by (t_cross, pixel).
  const std::vector<RawEventDev>& raw_events() const { return raw_sorted_; }

  const bsx_stats& stats() const { return stats_; }
  void reset_stats();
  void set_policy(const bsx_launch_policy& p);
  LaunchChain& chain() { return *chain_; }
  bsx_sensor_info_t info() const;
  const SensorMaps& maps() const { return maps_; }
  const Readout& readout() const { return readout_; }
  // The call constants, maps and current state as a descriptor (for a sampler that reads the sensor's
  // reference state between calls).
  CameraDev current_view() const;
  uint32_t width() const { return width_; }
  uint32_t height() const { return height_; }
  const SensorParams& params() const { return params_; }

 private:
  friend class Rig;
  void upload_maps();
  StateDev state_view(int which) const;
  RunsDev runs_view() const;
  // The checks of device runs' ids and bounds and the scan of the run lengths, queued on stream; their
  // results are read by finish_device_checks after the stream is synchronized.
  void enqueue_device_checks(const bsx_device_runs& runs, int64_t t_end_ns, hipStream_t stream);
  // Throws std::invalid_argument for the first failed check (before any state change).
  void finish_device_checks();
  // The samples' values are checked by the chain's phase A, which counts the values that are not
  // finite into c[1 + kStNonFinite] (c: this camera's counters after the launch); the walk then
  // advances no pixel. A nonzero count refuses the call (std::invalid_argument), the model left at
  // its state before the call.
  void refuse_non_finite(const unsigned long long* c);
  // This camera's descriptor for a launch of the chain over the runs (row_of built, pos scanned:
  // total segments), its buffers sized for them.
  CameraDev stage(const RunsDev& runs, const uint32_t* pixel, size_t n_runs, int64_t total);
  // Queues the scan of the run lengths into d_pos_ and the total's read into scan_total_.
  void enqueue_scan(const RunsDev& runs, size_t n_runs, hipStream_t stream);
  // Grows the event buffer after an overflow to hold `need` events (the descriptor's ev and cap).
  void grow_events(unsigned long long need, CameraDev& cd);
  // The end of a call whose chain completed: the written state becomes current, the counters are
  // added to the statistics.
  void commit(const unsigned long long* counters, int64_t t_end);
  // One call on device runs (row_of built, checks passed), with overflow re-runs.
  void run_advance(const RunsDev& runs, const uint32_t* pixel, size_t n_runs, int64_t t_end, hipStream_t stream,
                   bool scanned);
  void finish_host_events(const RawEventDev* raw, size_t n);
  KConst kconst() const;
  KConst constants_of_params() const;
  static void require_device();
  void upload_noise_tables(const float* data, size_t n, int levels, int n_nodes, int stride, float ln_min,
                           float inv_dln, float margin_z);
  void allocate();
  void set_kernel_maps(const bsx_pixel_map* maps, size_t n);
  bool instant() const { return kc_.instant != 0; }

  SensorParams params_;
  uint32_t width_, height_;
  uint64_t seed_;
  uint32_t camera_;
  std::vector<bsx_waveform> groups_;
  int G_;
  size_t P_;
  SensorMaps maps_;
  Readout readout_{};
  double gain_max_ = 1.0;
  float irr_max_ = 0.0f;
  float scale32_ = 1.0f;       // illuminance_scale in fp32 (the maps' scale is scale32_ x gain)
  KConst kc_{};                // the call constants (irr_max and the noise view added by kconst())
  bool noise_ = false;         // the cascade's in-path noise
  bool from_constants_ = false;  // built from the reference's kernel constants (the backend seam)
  bool states_set_ = false;

  DeviceBuffer d_maps_;       // 6 fp32 arrays of P
  DeviceBuffer d_noise_tab_;  // the in-path noise's fp32 tables (noise on)
  NoiseDev noise_dev_{};      // their view (tab null: noise off)
  DeviceBuffer d_state_[2];   // two state copies
  int cur_ = 0;
  DeviceBuffer d_events_;
  unsigned long long cap_ = 0;
  DeviceBuffer d_counters_;   // 1 + kStCount uint64
  DeviceBuffer d_flags_;      // scatter flags + check counters
  DeviceBuffer d_row_of_;
  DeviceBuffer d_t_, d_irr_, d_row_start_, d_row_count_, d_frame_, d_ids_;
  DeviceBuffer d_pos_, d_scan_;  // the runs' segment positions (exclusive scan of the run lengths)
  DeviceBuffer d_work_;  // per run: estimated walk cost (the heavy-first order's key)
  DeviceBuffer d_loud_;  // per run: the first segment that is not quiet under the input state
  DeviceBuffer d_sblocks_;  // per kScreenBlock segment positions: phase A's photocurrent range
  std::unique_ptr<LaunchChain> chain_;
  // Host copies of a device call's check words and scan total (written by queued copies).
  unsigned int check_flags_[4] = {0, 0, 0, 0};
  unsigned long long check_samples_[2] = {0, 0};
  int64_t scan_total_ = 0;
  unsigned long long counters_host_[1 + kStCount] = {};
  CameraDev noalloc_cam_{};  // the no-allocation call's descriptor and header, copied into the scratch
  CallDev noalloc_call_{};
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;

  int64_t t_state_ = 0;
  bool initialized_ = false;
  bsx_stats stats_{};
  HostEvents events_;
  std::vector<RawEventDev> raw_host_;
  std::vector<RawEventDev> raw_sorted_;
};

}  // namespace bsx
