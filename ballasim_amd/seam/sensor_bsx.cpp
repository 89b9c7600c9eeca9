// The reference's GPU backend seam implemented on this library: the members of
// balla::sensor::CudaSensorKernel (src/sensor/sensor_cuda.h:20-68) and sensor_cuda_device_count(),
// over the C ABI of include/bsx/bsx_sensor.h. A maintainer builds the reference with
// BALLASIM_SENSOR_CUDA=1 and this file in place of src/sensor/sensor_cuda.cu, links libbsx.so, and the
// reference's own SensorModel(device = kCuda) runs on an MI355X unchanged.
//
// Features of the reference that this backend does not run (modulated light, transient noise
// diffusion, the high-pass filter, veiling glare, regions of interest, threshold-type hot pixels)
// throw std::invalid_argument where the reference would first use them.
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "bsx/bsx_sensor.h"
#include "sensor/sensor_cuda.h"

namespace balla::sensor {

namespace {

void check(int status, const char* what) {
  if (status == BSX_OK) return;
  char msg[1024];
  bsx_last_error(msg, sizeof msg);
  const std::string text = std::string("bsx ") + what + ": " + msg;
  if (status == BSX_E_INVALID) throw std::invalid_argument(text);
  if (status == BSX_E_LOGIC) throw std::logic_error(text);
  throw std::runtime_error(text);
}

static_assert(sizeof(PixelMap) == sizeof(bsx_pixel_map), "PixelMap and bsx_pixel_map share the 48-byte layout");
static_assert(sizeof(PixelState) == 200, "the 200-byte PixelState");
static_assert(sizeof(RawEvent) == sizeof(bsx_raw_event), "RawEvent and bsx_raw_event share the 32-byte layout");

bsx_kernel_constants constants_of(const KernelConstants<float>& kc) {
  bsx_kernel_constants c{};
  c.k = kc.cc.k;
  c.a2 = kc.cc.a2;
  c.a3 = kc.cc.a3;
  c.tau_o = kc.cc.tau_o;
  c.tau_sf = kc.cc.tau_sf;
  c.eps_accept = kc.cc.eps_accept;
  c.eps_drop = kc.cc.eps_drop;
  c.eps_pl_base = kc.eps_pl_base;
  c.scale = kc.scale;
  c.irr_max = kc.irr_max;
  c.groups = kc.groups;
  c.instant = kc.instant ? 1 : 0;
  c.noise = kc.noise ? 1 : 0;
  c.levels = kc.levels;
  c.margin_z = kc.margin_z;
  c.n_nodes = kc.tables.n_nodes;
  c.stride = kc.tables.stride;
  c.ln_a1_min = kc.tables.ln_a1_min;
  c.inv_dln = kc.tables.inv_dln;
  c.ideal_reset = kc.ideal_reset ? 1 : 0;
  c.refractory_ns = kc.refractory_ns;
  c.eps_ci = kc.eps_ci;
  c.label_cut = kc.label_cut;
  c.seed = kc.seed;
  c.camera = kc.camera;
  c.modulated = kc.modulated ? 1 : 0;
  c.breakpoints = kc.breakpoints ? 1 : 0;
  c.transient = kc.transient ? 1 : 0;
  c.hpf = kc.hpf ? 1 : 0;
  return c;
}

const bsx_pixel_map* maps_of(const std::vector<PixelMap>& maps) {
  return reinterpret_cast<const bsx_pixel_map*>(maps.data());
}

bsx_device_runs runs_of(const DeviceRuns& r) {
  return bsx_device_runs{r.pixel, r.row_start, r.row_count, r.n, r.t, r.irr, r.n_samples};
}

struct Counters {
  bsx_stats s{};
};

// Adds the difference of the model's statistics over a call to the reference's counters.
void add_delta(const bsx_stats& a, const bsx_stats& b, KernelStats& stats, GpuLaunchStats& launch) {
  stats.substeps += b.substeps - a.substeps;
  stats.substep_attempts += b.substep_attempts - a.substep_attempts;
  stats.settled_substeps += b.settled_substeps - a.settled_substeps;
  stats.cells += b.cells - a.cells;
  stats.refinements += b.refinements - a.refinements;
  stats.leaf_intervals += b.leaf_intervals - a.leaf_intervals;
  stats.fallbacks += b.fallbacks - a.fallbacks;
  stats.input_clamps += b.input_clamps - a.input_clamps;
  stats.quarantines += b.quarantines - a.quarantines;
  stats.pieces += b.pieces - a.pieces;
  const double ms = b.kernel_ms - a.kernel_ms;
  launch.launches += b.launches - a.launches;
  launch.reruns += b.overflow_reruns - a.overflow_reruns;
  launch.total_ms += ms;
  if (ms > launch.max_ms) launch.max_ms = ms;
}

}  // namespace

int sensor_cuda_device_count() {
  int32_t n = 0;
  if (bsx_device_count(&n) != BSX_OK) return 0;
  return n;
}

struct CudaSensorKernel::Impl {
  bsx_sensor* s = nullptr;
  size_t n_pixels = 0;
  std::vector<RawEvent> host;

  bsx_stats stats() const {
    bsx_stats st{};
    check(bsx_sensor_stats(s, &st), "stats");
    return st;
  }
};

CudaSensorKernel::CudaSensorKernel(const KernelConstants<float>& kc, const std::vector<PixelMap>& maps,
                                   const std::vector<float>& tables)
    : impl_(std::make_unique<Impl>()) {
  const bsx_kernel_constants c = constants_of(kc);
  check(bsx_sensor_create_from_constants(&c, maps_of(maps), maps.size(), tables.empty() ? nullptr : tables.data(),
                                         tables.size(), &impl_->s),
        "CudaSensorKernel");
  impl_->n_pixels = maps.size();
}

CudaSensorKernel::~CudaSensorKernel() {
  if (impl_ && impl_->s != nullptr) bsx_sensor_destroy(impl_->s);
}

void CudaSensorKernel::upload_states(const std::vector<PixelState>& states, const std::vector<HpfState>& hpf) {
  if (!hpf.empty()) throw std::invalid_argument("bsx: the high-pass filter is not supported by the bsx backend");
  check(bsx_sensor_upload_states(impl_->s, states.data(), states.size()), "upload_states");
}

void CudaSensorKernel::download_states(std::vector<PixelState>& states) const {
  states.resize(impl_->n_pixels);
  check(bsx_sensor_download_states(impl_->s, states.data(), states.size()), "download_states");
}

void CudaSensorKernel::download_hpf(std::vector<HpfState>& hpf) const { hpf.clear(); }

void CudaSensorKernel::enable_hpf(const KernelConstants<float>&, const std::vector<float>&) {
  throw std::invalid_argument("bsx: the high-pass filter is not supported by the bsx backend");
}

void CudaSensorKernel::set_pixel_ids(const std::vector<uint32_t>&) {
  throw std::invalid_argument("bsx: region-of-interest models are not supported by the bsx backend");
}

void CudaSensorKernel::set_policy(const GpuLaunchPolicy& policy) {
  // Only the initial event capacity has a meaning here: this backend's launches are not sliced.
  bsx_launch_policy p{};
  p.event_capacity = policy.event_capacity;
  check(bsx_sensor_set_launch_policy(impl_->s, &p), "set_policy");
}

void CudaSensorKernel::set_photometry(const KernelConstants<float>& kc, const std::vector<PixelMap>& maps) {
  check(bsx_sensor_set_kernel_photometry(impl_->s, kc.scale, kc.irr_max, maps_of(maps), maps.size()),
        "set_photometry");
}

void CudaSensorKernel::set_glare(const GlareRamp<float>& glare) {
  bool zero = true;
  for (int g = 0; g < kMaxLightGroups; ++g) zero = zero && glare.a[g] == 0.0f && glare.b[g] == 0.0f;
  if (glare.on && !zero) throw std::invalid_argument("bsx: veiling glare is not supported by the bsx backend");
}

std::vector<float> CudaSensorKernel::end_frame(const DeviceRuns& runs, void* stream) {
  bsx_sensor_info_t info{};
  check(bsx_sensor_info(impl_->s, &info), "end_frame");
  std::vector<float> out(static_cast<size_t>(info.width) * info.height * static_cast<size_t>(info.groups));
  if (out.empty()) return out;
  const bsx_device_runs r = runs_of(runs);
  check(bsx_sensor_held_frame(impl_->s, runs.n > 0 ? &r : nullptr, stream, out.data()), "end_frame");
  return out;
}

DeviceEvents CudaSensorKernel::advance_device(const DeviceRuns& runs, int64_t t_state, int64_t t_end, void* stream,
                                              bool check_inputs, KernelStats& stats, GpuLaunchStats& launch_stats) {
  (void)check_inputs;  // the runs are always checked (ids, bounds, finite samples) before any state changes
  const bsx_stats before = impl_->stats();
  check(bsx_sensor_set_time(impl_->s, t_state), "advance_device");
  const bsx_device_runs r = runs_of(runs);
  bsx_device_events ev{};
  check(bsx_sensor_advance_device(impl_->s, &r, t_end, stream, &ev), "advance_device");
  add_delta(before, impl_->stats(), stats, launch_stats);
  return DeviceEvents{reinterpret_cast<const RawEvent*>(ev.data), ev.count};
}

std::vector<RawEvent> CudaSensorKernel::download(const DeviceEvents& events) const {
  const bsx_device_events ev{reinterpret_cast<const bsx_raw_event*>(events.data), events.count};
  check(bsx_sensor_download_events(impl_->s, &ev), "download");
  const bsx_raw_event* raw = nullptr;
  size_t n = 0;
  check(bsx_sensor_raw_events(impl_->s, &raw, &n), "download");
  std::vector<RawEvent> out(n);
  if (n > 0) std::memcpy(static_cast<void*>(out.data()), raw, n * sizeof(RawEvent));
  return out;
}

std::vector<RawEvent> CudaSensorKernel::advance(const uint32_t* pixel_ids, size_t n, const int64_t* t,
                                                const float* irr, size_t n_samples, const int64_t* row_start,
                                                const int32_t* row_count, int64_t t_state, int64_t t_end,
                                                KernelStats& stats, GpuLaunchStats& launch_stats) {
  const bsx_stats before = impl_->stats();
  check(bsx_sensor_set_time(impl_->s, t_state), "advance");
  check(bsx_sensor_advance_runs(impl_->s, pixel_ids, n, t, irr, n_samples, row_start, row_count, t_end), "advance");
  bsx_stats after = impl_->stats();
  // SensorModel::advance_raw counted the clamped samples of these host runs itself
  after.input_clamps = before.input_clamps;
  add_delta(before, after, stats, launch_stats);
  const bsx_raw_event* raw = nullptr;
  size_t count = 0;
  check(bsx_sensor_raw_events(impl_->s, &raw, &count), "advance");
  std::vector<RawEvent> out(count);
  if (count > 0) std::memcpy(static_cast<void*>(out.data()), raw, count * sizeof(RawEvent));
  return out;
}

}  // namespace balla::sensor
