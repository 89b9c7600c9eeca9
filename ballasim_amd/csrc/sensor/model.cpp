#include "model.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "chain.h"
#include "kernels.h"
#include "noise_tables.h"
#include "scratch.h"

namespace bsx {

namespace {

constexpr double kMaxSamplePhotocurrent_fA = 1e9;  // the reference's input clamp (src/sensor/pixel.h:97)
constexpr uint32_t kStateMagic = 0x534e5342;      // "BSNS"
constexpr uint32_t kStateVersion = 6;

struct StateHeader {
  uint32_t magic, version;
  uint32_t width, height;
  uint32_t groups, camera;
  uint64_t seed;
  uint64_t fingerprint;
  int64_t t_state;
  uint32_t initialized;
  uint32_t state_size;
};
static_assert(sizeof(StateHeader) == 56, "state header layout");

size_t align256(size_t b) { return (b + 255) & ~size_t{255}; }

// FNV-1a 64 over the parameters, the waveforms and a gain map (when not all ones).
uint64_t fingerprint(const SensorParams& p, const std::vector<bsx_waveform>& groups, const std::vector<float>& gain) {
  uint64_t h = 0xcbf29ce484222325ull;
  const auto mix = [&](const void* d, size_t n) {
    const auto* b = static_cast<const unsigned char*>(d);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
  };
  mix(&p, sizeof(p));
  for (const bsx_waveform& w : groups) mix(&w, sizeof(w));
  if (!std::all_of(gain.begin(), gain.end(), [](float g) { return g == 1.0f; })) mix(gain.data(), gain.size() * 4);
  return h;
}

}  // namespace

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("bsx: ") + what + ": " + hipGetErrorString(e));
}

DeviceBuffer::~DeviceBuffer() {
  if (p_ != nullptr) (void)hipFree(p_);
}

void* DeviceBuffer::ensure(size_t bytes) {
  if (bytes <= bytes_ && p_ != nullptr) return p_;
  if (p_ != nullptr) hip_check(hipFree(p_), "hipFree");
  p_ = nullptr;
  bytes_ = 0;
  const size_t want = std::max<size_t>(bytes, 256);
  hip_check(hipMalloc(&p_, want), "hipMalloc");
  bytes_ = want;
  return p_;
}

void HostEvents::resize(size_t n) {
  x.resize(n);
  y.resize(n);
  p.resize(n);
  label.resize(n);
  t_cross.resize(n);
  t_reported.resize(n);
  flags.resize(n);
  ds.resize(n);
  dn.resize(n);
  dl.resize(n);
  l_first.resize(n);
}

SensorModel::SensorModel(const SensorParams& params, uint32_t width, uint32_t height, uint64_t seed,
                         std::vector<bsx_waveform> groups, uint32_t camera)
    : params_(params), width_(width), height_(height), seed_(seed), camera_(camera), groups_(std::move(groups)) {
  validate(params_);
  if (width == 0 || height == 0 || width > 65535 || height > 65535) {
    throw std::invalid_argument("SensorModel: width and height must lie in [1, 65535]");
  }
  if (camera > 255) throw std::invalid_argument("SensorModel: camera index must be < 256 (purpose.h layout)");
  if (groups_.empty() || groups_.size() > static_cast<size_t>(kMaxGroups)) {
    throw std::invalid_argument("SensorModel: need 1 to 8 light groups");
  }
  for (const bsx_waveform& w : groups_) {
    if (w.type != BSX_WAVE_CONSTANT) {
      throw std::invalid_argument("SensorModel: modulated light groups (fourier, pulse) are not supported by the bsx "
                                  "backend");
    }
  }
  require_device();
  G_ = static_cast<int>(groups_.size());
  P_ = static_cast<size_t>(width) * height;
  maps_ = draw_sensor_maps(params_, seed_, camera_, static_cast<uint32_t>(P_));
  // The kernel needs k * 1 ns * I^2 normal in fp32 (src/sensor/sensor_model.cpp:235-259).
  const double floor_fA = std::sqrt(static_cast<double>(FLT_MIN) / (params_.k * 1e-15 * 1e-9));
  double lowest = INFINITY;
  for (const double d : maps_.dark_fA) lowest = std::min(lowest, d);
  if (!(lowest >= floor_fA)) {
    throw std::invalid_argument("SensorModel: the pixel dark current " + std::to_string(lowest) +
                                " fA is below the fp32 kernel's floor of " + std::to_string(floor_fA) + " fA");
  }
  readout_ = readout_of(params_, seed_, camera_);
  irr_max_ = static_cast<float>(kMaxSamplePhotocurrent_fA / (params_.illuminance_scale * gain_max_));
  scale32_ = static_cast<float>(params_.illuminance_scale);
  kc_ = constants_of_params();
  noise_ = noise_on(params_);
  if (noise_) {
    NoiseModelParams nm{};
    nm.kq = params_.k * kElementaryCharge;
    nm.a2 = 1.0 / params_.tau_o;
    nm.a3 = 1.0 / params_.tau_sf;
    nm.sigma_o = params_.sigma_o;
    nm.levels = params_.noise_leaf_level;
    const NoiseTableSet tabs = build_noise_tables(nm);
    upload_noise_tables(tabs.data32.data(), tabs.data32.size(), tabs.levels, tabs.n_nodes, tabs.stride,
                        static_cast<float>(kNoiseLnA1Min), static_cast<float>(tabs.inv_dln),
                        static_cast<float>(params_.noise_margin_z));
  }
  allocate();
}

SensorModel::SensorModel(const bsx_kernel_constants& kc, const bsx_pixel_map* maps, size_t n_pixels,
                         const float* tables, size_t n_table_values)
    : params_(preset("prophesee_gen3")), width_(0), height_(1), seed_(kc.seed), camera_(kc.camera) {
  // The reference's SensorModel built these from its parameters and checked them; here the features
  // this backend does not run are refused by name.
  if (kc.modulated || kc.breakpoints) {
    throw std::invalid_argument("CudaSensorKernel (bsx): modulated light groups are not supported by the bsx backend");
  }
  if (kc.transient) {
    throw std::invalid_argument("CudaSensorKernel (bsx): noise_diffusion transient is not supported by the bsx backend");
  }
  if (kc.hpf) throw std::invalid_argument("CudaSensorKernel (bsx): the high-pass filter is not supported by the bsx backend");
  if (kc.groups < 1 || kc.groups > kMaxGroups) throw std::invalid_argument("CudaSensorKernel (bsx): need 1 to 8 light groups");
  if (kc.camera > 255) throw std::invalid_argument("CudaSensorKernel (bsx): camera index must be < 256");
  if (n_pixels == 0 || n_pixels > 0xffffffffull || (n_pixels > 0 && maps == nullptr)) {
    throw std::invalid_argument("CudaSensorKernel (bsx): need one PixelMap per pixel");
  }
  require_device();
  from_constants_ = true;
  G_ = kc.groups;
  groups_.assign(static_cast<size_t>(G_), bsx_waveform{});
  P_ = n_pixels;
  width_ = static_cast<uint32_t>(n_pixels);
  set_kernel_maps(maps, n_pixels);
  irr_max_ = kc.irr_max;
  scale32_ = kc.scale;
  KConst k{};
  k.groups = kc.groups;
  k.instant = kc.instant ? 1 : 0;
  k.ideal_reset = kc.ideal_reset ? 1 : 0;
  k.n_pixels = static_cast<int32_t>(P_);
  k.refractory_ns = kc.refractory_ns;
  k.eps_ci = kc.eps_ci;
  k.label_cut = kc.label_cut;
  k.k = kc.k;
  k.a2 = kc.a2;
  k.a3 = kc.a3;
  k.tau_o = kc.tau_o;
  k.tau_sf = kc.tau_sf;
  k.eps_accept = kc.eps_accept;
  k.eps_drop = kc.eps_drop;
  k.eps_pl_base = kc.eps_pl_base;
  k.shot_rate_hz = 0.0;
  k.seed = kc.seed;
  k.camera = kc.camera;
  k.width = width_;
  kc_ = k;
  noise_ = kc.noise != 0 && !kc.instant;
  if (noise_) {
    if (kc.levels < 0 || kc.levels > kNoiseMaxLevels || kc.n_nodes < 2 || kc.stride != noise_stride(kc.levels) ||
        tables == nullptr || n_table_values != static_cast<size_t>(kc.n_nodes) * static_cast<size_t>(kc.stride)) {
      throw std::invalid_argument("CudaSensorKernel (bsx): the noise tables do not match their layout");
    }
    upload_noise_tables(tables, n_table_values, kc.levels, kc.n_nodes, kc.stride, kc.ln_a1_min, kc.inv_dln,
                        kc.margin_z);
  }
  allocate();
}

void SensorModel::require_device() {
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev <= 0) {
    throw std::runtime_error("SensorModel: no HIP device available");
  }
}

void SensorModel::upload_noise_tables(const float* data, size_t n, int levels, int n_nodes, int stride, float ln_min,
                                      float inv_dln, float margin_z) {
  d_noise_tab_.ensure(n * sizeof(float));
  hip_check(hipMemcpy(d_noise_tab_.get(), data, n * sizeof(float), hipMemcpyHostToDevice), "upload noise tables");
  noise_dev_.tab = d_noise_tab_.as<const float>();
  noise_dev_.levels = levels;
  noise_dev_.n_nodes = n_nodes;
  noise_dev_.stride = stride;
  noise_dev_.ln_min = ln_min;
  noise_dev_.inv_dln = inv_dln;
  noise_dev_.margin_z = margin_z;
}

void SensorModel::allocate() {
  size_t state_bytes = 0;
  state_bytes += 9 * align256(P_ * 4) + 4 * align256(P_ * 8) + align256(P_ * 4) + align256(P_ * 2);
  state_bytes += align256(P_ * 4 * G_);
  // the in-path noise: nk[3], nk1[3], ln_a1 (fp32), cell (int64), have_k1 (byte)
  if (noise_) state_bytes += 7 * align256(P_ * 4) + align256(P_ * 8) + align256(P_);
  d_state_[0].ensure(state_bytes);
  d_state_[1].ensure(state_bytes);
  upload_maps();
  d_counters_.ensure((1 + kStCount) * sizeof(unsigned long long));
  d_flags_.ensure(64);
  hip_check(hipEventCreate(&ev0_), "hipEventCreate");
  hip_check(hipEventCreate(&ev1_), "hipEventCreate");
  chain_ = std::make_unique<LaunchChain>();
}

void SensorModel::set_kernel_maps(const bsx_pixel_map* maps, size_t n) {
  maps_ = SensorMaps{};
  maps_.theta_on.resize(n);
  maps_.theta_off.resize(n);
  maps_.leak.resize(n);
  maps_.dark_fA.resize(n);
  maps_.phi0.resize(n);
  maps_.hot.resize(n);
  maps_.hot_rate_hz.assign(n, 0.0);
  maps_.gain.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const bsx_pixel_map& m = maps[i];
    if (m.hot > 1) {
      throw std::invalid_argument("CudaSensorKernel (bsx): threshold-type hot pixels are not supported by the bsx "
                                  "backend");
    }
    maps_.theta_on[i] = m.theta_on;
    maps_.theta_off[i] = m.theta_off;
    maps_.leak[i] = m.leak;
    maps_.dark_fA[i] = m.dark_fA;
    maps_.phi0[i] = m.phi0;
    maps_.hot[i] = m.hot;
    maps_.gain[i] = m.gain;
  }
}

void SensorModel::set_kernel_photometry(float scale, float irr_max, const bsx_pixel_map* maps, size_t n) {
  if (!from_constants_) throw std::logic_error("set_kernel_photometry: the model was built from parameters");
  if (n != P_ || maps == nullptr) throw std::invalid_argument("set_kernel_photometry: one PixelMap per pixel");
  set_kernel_maps(maps, n);
  scale32_ = scale;
  irr_max_ = irr_max;
  upload_maps();
}

void SensorModel::set_time(int64_t t_state_ns) {
  if (!states_set_) throw std::logic_error("SensorModel.set_time: upload the states first");
  t_state_ = t_state_ns;
  initialized_ = true;
}

SensorModel::~SensorModel() {
  if (ev0_ != nullptr) (void)hipEventDestroy(ev0_);
  if (ev1_ != nullptr) (void)hipEventDestroy(ev1_);
}

KConst SensorModel::kconst() const {
  KConst kc = kc_;
  kc.irr_max = irr_max_;
  kc.noise = noise_dev_;
  return kc;
}

KConst SensorModel::constants_of_params() const {
  KConst kc{};
  kc.groups = G_;
  kc.instant = params_.photoreceptor == BSX_PHOTORECEPTOR_INSTANTANEOUS ? 1 : 0;
  kc.ideal_reset = params_.reset_mode == BSX_RESET_IDEAL ? 1 : 0;
  kc.n_pixels = static_cast<int32_t>(P_);
  kc.refractory_ns = std::llround(params_.refractory_s * 1e9);
  kc.irr_max = irr_max_;
  kc.eps_ci = static_cast<float>(params_.eps_ci);
  kc.label_cut = static_cast<float>(params_.label_share_cut);
  // Cascade constants (src/sensor/sensor_model.cpp:375-390): eps_h = eps_L / 8, accepted at eps_h / 1.5.
  const double eps_h = params_.eps_L / 8.0;
  kc.k = static_cast<float>(params_.k * 1e-15);
  kc.a2 = static_cast<float>(1.0 / params_.tau_o);
  kc.a3 = static_cast<float>(1.0 / params_.tau_sf);
  kc.tau_o = static_cast<float>(params_.tau_o);
  kc.tau_sf = static_cast<float>(params_.tau_sf);
  kc.eps_accept = static_cast<float>(eps_h / 1.5);
  kc.eps_drop = static_cast<float>(std::max(eps_h / 1024.0, 32.0 * static_cast<double>(FLT_EPSILON)));
  kc.eps_pl_base = static_cast<float>(params_.eps_L / 8.0);
  kc.shot_rate_hz = params_.shot_rate_hz;
  kc.seed = seed_;
  kc.camera = camera_;
  kc.width = width_;
  return kc;
}

void SensorModel::upload_maps() {
  const size_t a = align256(P_ * 4);
  d_maps_.ensure(7 * a);
  std::vector<float> buf(P_);
  const float scale32 = scale32_;
  const auto put = [&](int slot, const auto& fn) {
    for (size_t i = 0; i < P_; ++i) buf[i] = fn(i);
    hip_check(hipMemcpy(static_cast<char*>(d_maps_.get()) + slot * a, buf.data(), P_ * 4, hipMemcpyHostToDevice),
              "upload maps");
  };
  put(0, [&](size_t i) { return static_cast<float>(maps_.theta_on[i]); });
  put(1, [&](size_t i) { return static_cast<float>(maps_.theta_off[i]); });
  put(2, [&](size_t i) { return static_cast<float>(maps_.leak[i]); });
  put(3, [&](size_t i) { return static_cast<float>(maps_.dark_fA[i]); });
  put(4, [&](size_t i) { return scale32 * maps_.gain[i]; });  // kc.scale x gain in fp32
  put(5, [&](size_t i) { return static_cast<float>(maps_.phi0[i]); });
  hip_check(hipMemcpy(static_cast<char*>(d_maps_.get()) + 6 * a, maps_.hot.data(), P_, hipMemcpyHostToDevice),
            "upload maps");
}

StateDev SensorModel::state_view(int which) const {
  char* b = static_cast<char*>(d_state_[which].get());
  const size_t f4 = align256(P_ * 4), f8 = align256(P_ * 8);
  StateDev s{};
  float** fl[9] = {&s.z, &s.x1, &s.s2, &s.s3, &s.i_scale, &s.tv, &s.s3_ref, &s.n3_ref, &s.dr_ref};
  for (float** f : fl) {
    *f = reinterpret_cast<float*>(b);
    b += f4;
  }
  s.t_ref = reinterpret_cast<int64_t*>(b);
  b += f8;
  s.blind_until = reinterpret_cast<int64_t*>(b);
  b += f8;
  s.t_search = reinterpret_cast<int64_t*>(b);
  b += f8;
  s.shot_next = reinterpret_cast<int64_t*>(b);
  b += f8;
  s.shot_index = reinterpret_cast<uint32_t*>(b);
  b += f4;
  s.flags = reinterpret_cast<uint16_t*>(b);
  b += align256(P_ * 2);
  s.held = reinterpret_cast<float*>(b);
  b += align256(P_ * 4 * G_);
  if (noise_) {
    s.nk = reinterpret_cast<float*>(b);
    b += 3 * f4;
    s.nk1 = reinterpret_cast<float*>(b);
    b += 3 * f4;
    s.ln_a1 = reinterpret_cast<float*>(b);
    b += f4;
    s.cell = reinterpret_cast<int64_t*>(b);
    b += f8;
    s.have_k1 = reinterpret_cast<uint8_t*>(b);
  }
  return s;
}

static MapsDev maps_view(const DeviceBuffer& d, size_t P) {
  const size_t a = align256(P * 4);
  const char* b = static_cast<const char*>(d.get());
  MapsDev m{};
  m.theta_on = reinterpret_cast<const float*>(b + 0 * a);
  m.theta_off = reinterpret_cast<const float*>(b + 1 * a);
  m.leak = reinterpret_cast<const float*>(b + 2 * a);
  m.dark = reinterpret_cast<const float*>(b + 3 * a);
  m.scale = reinterpret_cast<const float*>(b + 4 * a);
  m.phi0 = reinterpret_cast<const float*>(b + 5 * a);
  m.hot = reinterpret_cast<const uint8_t*>(b + 6 * a);
  return m;
}

void SensorModel::set_photometry(double illuminance_scale, const float* gain, size_t n_gain) {
  if (from_constants_) throw std::logic_error("SensorModel::set_photometry: the model was built from kernel constants");
  if (gain != nullptr && n_gain != P_) {
    throw std::invalid_argument("SensorModel::set_photometry: one gain per pixel of the frame (or none)");
  }
  SensorParams p = params_;
  p.illuminance_scale = illuminance_scale;
  validate(p);
  double g_max = gain == nullptr ? 1.0 : 0.0;
  for (size_t i = 0; gain != nullptr && i < n_gain; ++i) {
    if (!(gain[i] > 0.0f) || !std::isfinite(gain[i])) {
      throw std::invalid_argument("SensorModel::set_photometry: gains must be > 0");
    }
    g_max = std::max(g_max, static_cast<double>(gain[i]));
  }
  params_ = p;
  gain_max_ = g_max;
  for (size_t i = 0; i < P_; ++i) maps_.gain[i] = gain == nullptr ? 1.0f : gain[i];
  irr_max_ = static_cast<float>(kMaxSamplePhotocurrent_fA / (params_.illuminance_scale * gain_max_));
  scale32_ = static_cast<float>(params_.illuminance_scale);
  upload_maps();
}

void SensorModel::init(int64_t t0_ns, const float* irradiance0) {
  const size_t count = P_ * G_;
  for (size_t i = 0; i < count; ++i) {
    if (!std::isfinite(irradiance0[i])) throw std::invalid_argument("init: non-finite irradiance");
  }
  initialized_ = false;
  d_frame_.ensure(count * sizeof(float));
  hip_check(hipMemcpy(d_frame_.get(), irradiance0, count * sizeof(float), hipMemcpyHostToDevice), "upload frame");
  hip_check(launch_fill_u64(d_counters_.as<unsigned long long>(), 1 + kStCount, 0, nullptr), "fill");
  hip_check(launch_init(kconst(), maps_view(d_maps_, P_), state_view(cur_), d_frame_.as<float>(), t0_ns,
                        d_counters_.as<unsigned long long>(), nullptr),
            "init kernel");
  unsigned long long c[1 + kStCount];
  hip_check(hipMemcpy(c, d_counters_.get(), sizeof(c), hipMemcpyDeviceToHost), "read counters");
  stats_.input_clamps += c[1 + kStClamps];
  stats_.cells += c[1 + kStCells];
  stats_.refinements += c[1 + kStRefinements];
  t_state_ = t0_ns;
  initialized_ = true;
}

void SensorModel::advance(const uint32_t* ids, size_t n, const int64_t* t, const float* irr, size_t m,
                          const int32_t* counts, int64_t t_end_ns) {
  if (m > static_cast<size_t>(INT32_MAX)) throw std::invalid_argument("SensorModel.advance: too many samples per pixel");
  std::vector<int64_t> start(n);
  std::vector<int32_t> count(n);
  for (size_t i = 0; i < n; ++i) {
    const int32_t c = counts != nullptr ? counts[i] : static_cast<int32_t>(m);
    if (c < 0 || static_cast<size_t>(c) > m) throw std::invalid_argument("SensorModel.advance: count out of range");
    start[i] = static_cast<int64_t>(i * m);
    count[i] = c;
  }
  advance_runs(ids, n, t, irr, n * m, start.data(), count.data(), t_end_ns);
}

void SensorModel::advance_runs(const uint32_t* ids, size_t n, const int64_t* t, const float* irr, size_t n_samples,
                               const int64_t* row_start, const int32_t* row_count, int64_t t_end_ns) {
  if (!initialized_) {
    throw std::logic_error(
        "SensorModel.advance: call init() or load_state() first (a call that failed part-way leaves the model "
        "uninitialized)");
  }
  if (t_end_ns < t_state_) throw std::invalid_argument("SensorModel.advance: t_end_ns is before the model time");
  // The reference's input rules (src/sensor/sensor_model.cpp:1075-1103), checked before any state change.
  std::vector<int32_t> row_of(P_, -1);
  uint64_t clamps = 0;
  for (size_t i = 0; i < n; ++i) {
    const uint32_t p = ids[i];
    if (p >= P_) throw std::invalid_argument("SensorModel.advance: pixel id out of range");
    if (row_of[p] >= 0) throw std::invalid_argument("SensorModel.advance: duplicate pixel id");
    row_of[p] = static_cast<int32_t>(i);
    const int64_t s = row_start[i];
    const int32_t c = row_count[i];
    if (c < 0 || s < 0 || static_cast<size_t>(s) + static_cast<size_t>(c) > n_samples) {
      throw std::invalid_argument("SensorModel.advance: sample run out of range");
    }
    const int64_t* ti = t + s;
    for (int32_t j = 0; j < c; ++j) {
      if (j == 0 ? ti[0] < t_state_ : ti[j] <= ti[j - 1]) {
        throw std::invalid_argument(
            "SensorModel.advance: sample times must be strictly increasing and start at or after the model time");
      }
    }
    if (c > 0 && ti[c - 1] > t_end_ns) throw std::invalid_argument("SensorModel.advance: sample after t_end_ns");
    const float* vi = irr + static_cast<size_t>(s) * G_;
    for (size_t j = 0; j < static_cast<size_t>(c) * G_; ++j) {
      if (!std::isfinite(vi[j])) {
        throw std::invalid_argument("SensorModel.advance: non-finite irradiance (pixel " + std::to_string(p) +
                                    ", sample " + std::to_string(j / G_) + ")");
      }
      const float v = vi[j];
      if (!(v >= 0.0f && v <= irr_max_)) ++clamps;
    }
  }
  stats_.input_clamps += clamps;
  initialized_ = false;
  d_t_.ensure(std::max<size_t>(n_samples, 1) * 8);
  d_irr_.ensure(std::max<size_t>(n_samples * G_, 1) * 4);
  d_row_start_.ensure(std::max<size_t>(n, 1) * 8);
  d_row_count_.ensure(std::max<size_t>(n, 1) * 4);
  d_row_of_.ensure(P_ * 4);
  if (n_samples > 0) {
    hip_check(hipMemcpy(d_t_.get(), t, n_samples * 8, hipMemcpyHostToDevice), "upload t");
    hip_check(hipMemcpy(d_irr_.get(), irr, n_samples * G_ * 4, hipMemcpyHostToDevice), "upload irr");
  }
  if (n > 0) {
    hip_check(hipMemcpy(d_row_start_.get(), row_start, n * 8, hipMemcpyHostToDevice), "upload row_start");
    hip_check(hipMemcpy(d_row_count_.get(), row_count, n * 4, hipMemcpyHostToDevice), "upload row_count");
  }
  hip_check(hipMemcpy(d_row_of_.get(), row_of.data(), P_ * 4, hipMemcpyHostToDevice), "upload row_of");
  d_ids_.ensure(std::max<size_t>(n, 1) * 4);
  if (n > 0) hip_check(hipMemcpy(d_ids_.get(), ids, n * 4, hipMemcpyHostToDevice), "upload ids");
  run_advance(runs_view(), d_ids_.as<const uint32_t>(), n, t_end_ns, nullptr, false);
  initialized_ = true;
  const unsigned long long count = counters_host_[0];
  raw_host_.resize(count);
  if (count > 0) {
    hip_check(hipMemcpy(raw_host_.data(), d_events_.get(), count * sizeof(RawEventDev), hipMemcpyDeviceToHost),
              "download events");
  }
  finish_host_events(raw_host_.data(), count);
}

RunsDev SensorModel::runs_view() const {
  RunsDev r{};
  r.row_of = d_row_of_.as<const int32_t>();
  r.row_start = d_row_start_.as<const int64_t>();
  r.row_count = d_row_count_.as<const int32_t>();
  r.t = d_t_.as<const int64_t>();
  r.irr = d_irr_.as<const float>();
  return r;
}

void SensorModel::enqueue_scan(const RunsDev& runs, size_t n_runs, hipStream_t stream) {
  d_pos_.ensure(std::max<size_t>(n_runs, 1) * 8);
  scan_total_ = 0;
  if (n_runs == 0) return;
  d_scan_.ensure(scan_scratch_count(n_runs) * 8);
  hip_check(launch_exclusive_scan(runs.row_count, n_runs, d_pos_.as<int64_t>(), d_scan_.as<int64_t>(), stream),
            "scan runs");
  hip_check(hipMemcpyAsync(&scan_total_, d_scan_.as<int64_t>() + scan_scratch_count(n_runs) - 1, 8,
                           hipMemcpyDeviceToHost, stream),
            "read total");
}

CameraDev SensorModel::stage(const RunsDev& runs, const uint32_t* pixel, size_t n_runs, int64_t total) {
  const bool inst = instant();
  if (cap_ == 0) cap_ = std::max<unsigned long long>(1ull << 20, 2 * P_);
  d_events_.ensure(cap_ * sizeof(RawEventDev));
  d_work_.ensure(std::max<size_t>(n_runs, 1) * sizeof(float));
  d_loud_.ensure(std::max<size_t>(n_runs, 1) * sizeof(int));
  if (inst) d_sblocks_.ensure((static_cast<size_t>(total) / kScreenBlock + 1) * sizeof(ScreenBlock));
  CameraDev cd{};
  cd.kc = kconst();
  cd.mp = maps_view(d_maps_, P_);
  cd.in = state_view(cur_);
  cd.out = state_view(1 - cur_);
  cd.runs = runs;
  cd.pixel = pixel;
  cd.pos = d_pos_.as<const int64_t>();
  cd.first_loud = inst ? d_loud_.as<int>() : nullptr;
  cd.screen_blocks = inst ? d_sblocks_.as<ScreenBlock>() : nullptr;
  cd.work = d_work_.as<float>();
  cd.ev = d_events_.as<RawEventDev>();
  cd.cap = cap_;
  cd.counters = d_counters_.as<unsigned long long>();
  cd.t_state = t_state_;
  cd.n_runs = static_cast<int64_t>(n_runs);
  cd.total = total;
  return cd;
}

CameraDev SensorModel::current_view() const {
  CameraDev cd{};
  cd.kc = kconst();
  cd.mp = maps_view(d_maps_, P_);
  cd.in = state_view(cur_);
  cd.t_state = t_state_;
  return cd;
}

void SensorModel::grow_events(unsigned long long need, CameraDev& cd) {
  cap_ = std::max<unsigned long long>(cap_ * 2, need + need / 4);
  d_events_.ensure(cap_ * sizeof(RawEventDev));
  cd.ev = d_events_.as<RawEventDev>();
  cd.cap = cap_;
  ++stats_.overflow_reruns;
}

void SensorModel::commit(const unsigned long long* c, int64_t t_end) {
  cur_ = 1 - cur_;
  t_state_ = t_end;
  stats_.events += c[0];
  stats_.substeps += c[1 + kStSubsteps];
  stats_.substep_attempts += c[1 + kStAttempts];
  stats_.settled_substeps += c[1 + kStSettled];
  stats_.cells += c[1 + kStCells];
  stats_.refinements += c[1 + kStRefinements];
  stats_.leaf_intervals += c[1 + kStLeaves];
  stats_.fallbacks += c[1 + kStFallbacks];
  stats_.quarantines += c[1 + kStQuarantines];
  stats_.pieces += c[1 + kStPieces];
}

void SensorModel::run_advance(const RunsDev& runs, const uint32_t* pixel, size_t n_runs, int64_t t_end,
                               hipStream_t stream, bool scanned) {
  if (!scanned) {
    enqueue_scan(runs, n_runs, stream);
    hip_check(hipStreamSynchronize(stream), "scan runs");
  }
  std::vector<CameraDev> cams{stage(runs, pixel, n_runs, scan_total_)};
  unsigned long long* c = counters_host_;
  hip_check(hipEventRecord(ev0_, stream), "event record");
  chain_->launch(cams, instant(), t_end, stream);
  while (true) {
    hip_check(hipEventRecord(ev1_, stream), "event record");
    hip_check(hipMemcpyAsync(c, d_counters_.get(), sizeof(counters_host_), hipMemcpyDeviceToHost, stream),
              "read counters");
    hip_check(hipStreamSynchronize(stream), "advance");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "event time");
    stats_.kernel_ms += ms;
    ++stats_.launches;
    refuse_non_finite(c);
    if (c[0] <= cap_) break;
    grow_events(c[0], cams[0]);
    hip_check(hipEventRecord(ev0_, stream), "event record");
    chain_->rerun_walk(cams, stream);
  }
  commit(c, t_end);
}

void SensorModel::enqueue_device_checks(const bsx_device_runs& runs, int64_t t_end_ns, hipStream_t stream) {
  if (!initialized_) {
    throw std::logic_error(
        "SensorModel.advance: call init() or load_state() first (a call that failed part-way leaves the model "
        "uninitialized)");
  }
  if (t_end_ns < t_state_) throw std::invalid_argument("SensorModel.advance: t_end_ns is before the model time");
  d_row_of_.ensure(P_ * 4);
  unsigned int* flags = d_flags_.as<unsigned int>();
  hip_check(launch_scatter_rows(runs.pixel, runs.row_start, runs.row_count, runs.n, runs.n_samples,
                                d_row_of_.as<int32_t>(), static_cast<uint32_t>(P_), flags, stream),
            "scatter rows");
  hip_check(hipMemcpyAsync(check_flags_, d_flags_.get(), 16, hipMemcpyDeviceToHost, stream), "read flags");
  // The scan reads row_count of runs the scatter may have refused; its result is used only if every
  // check passes. The samples' values are checked by the chain's phase A (refuse_non_finite).
  RunsDev r{};
  r.row_count = runs.row_count;
  enqueue_scan(r, runs.n, stream);
}

void SensorModel::finish_device_checks() {
  if (check_flags_[0] != 0) throw std::invalid_argument("SensorModel.advance: pixel id out of range");
  if (check_flags_[1] != 0) throw std::invalid_argument("SensorModel.advance: duplicate pixel id");
  if (check_flags_[2] != 0) throw std::invalid_argument("SensorModel.advance: sample run out of range");
}

void SensorModel::refuse_non_finite(const unsigned long long* c) {
  if (c[1 + kStNonFinite] == 0) return;
  initialized_ = true;  // the walk advanced no pixel and nothing was committed
  throw std::invalid_argument("SensorModel.advance: " + std::to_string(c[1 + kStNonFinite]) +
                              " non-finite irradiance samples");
}

void SensorModel::held_frame(const bsx_device_runs* runs, hipStream_t stream, float* out) {
  if (out == nullptr && P_ > 0) throw std::invalid_argument("SensorModel.held_frame: null output");
  const bool with_runs = runs != nullptr && runs->n > 0;
  if (with_runs) {
    d_row_of_.ensure(P_ * 4);
    d_flags_.ensure(64);
    hip_check(launch_scatter_rows(runs->pixel, runs->row_start, runs->row_count, runs->n, runs->n_samples,
                                  d_row_of_.as<int32_t>(), static_cast<uint32_t>(P_), d_flags_.as<unsigned int>(),
                                  stream),
              "held frame rows");
  }
  d_frame_.ensure(std::max<size_t>(P_ * G_, 1) * sizeof(float));
  const StateDev st = state_view(cur_);
  hip_check(launch_held_frame(st.held, with_runs ? d_row_of_.as<const int32_t>() : nullptr,
                              with_runs ? runs->row_start : nullptr, with_runs ? runs->row_count : nullptr,
                              with_runs ? runs->irr : nullptr, static_cast<uint32_t>(P_), static_cast<int>(G_),
                              d_frame_.as<float>(), stream),
            "held frame");
  if (P_ > 0) {
    hip_check(hipMemcpyAsync(out, d_frame_.get(), P_ * G_ * sizeof(float), hipMemcpyDeviceToHost, stream),
              "held frame download");
  }
  hip_check(hipStreamSynchronize(stream), "held frame");
}

bsx_device_events SensorModel::advance_device(const bsx_device_runs& runs, int64_t t_end_ns, hipStream_t stream) {
  enqueue_device_checks(runs, t_end_ns, stream);
  hip_check(hipStreamSynchronize(stream), "checks");
  finish_device_checks();
  RunsDev r{};
  r.row_of = d_row_of_.as<const int32_t>();
  r.row_start = runs.row_start;
  r.row_count = runs.row_count;
  r.t = runs.t;
  r.irr = runs.irr;
  initialized_ = false;
  run_advance(r, runs.pixel, runs.n, t_end_ns, stream, true);
  stats_.input_clamps += counters_host_[1 + kStClamps];
  initialized_ = true;
  return bsx_device_events{reinterpret_cast<const bsx_raw_event*>(d_events_.get()),
                           static_cast<size_t>(counters_host_[0])};
}

uint64_t SensorModel::advance_device_noalloc(const bsx_device_runs& runs, int64_t t_end_ns, hipStream_t stream,
                                             RawEventDev* events, uint64_t cap) {
  static_assert(kScratchCounters + (1 + kStCount) * sizeof(unsigned long long) <= kScratchCamera, "scratch layout");
  static_assert(kScratchCamera + sizeof(CameraDev) <= kScratchCall, "scratch layout");
  static_assert(kScratchCall + sizeof(CallDev) <= kScratchRowOf, "scratch layout");
  if (!initialized_) {
    throw std::logic_error(
        "SensorModel.advance: call init() or load_state() first (a call that failed part-way leaves the model "
        "uninitialized)");
  }
  if (t_end_ns < t_state_) throw std::invalid_argument("SensorModel.advance: t_end_ns is before the model time");
  if (P_ > kScratchMaxPixels) {
    throw std::invalid_argument("advance_device_noalloc: the camera has " + std::to_string(P_) +
                                " pixels; the static scratch serves at most " + std::to_string(kScratchMaxPixels));
  }
  if (events == nullptr && cap > 0) throw std::invalid_argument("advance_device_noalloc: events is null");
  char* base = static_cast<char*>(scratch_address());
  unsigned int* flags = reinterpret_cast<unsigned int*>(base + kScratchFlags);
  unsigned long long* chk = reinterpret_cast<unsigned long long*>(base + kScratchChecks);
  unsigned long long* counters = reinterpret_cast<unsigned long long*>(base + kScratchCounters);
  CameraDev* d_cam = reinterpret_cast<CameraDev*>(base + kScratchCamera);
  CallDev* d_call = reinterpret_cast<CallDev*>(base + kScratchCall);
  int32_t* row_of = reinterpret_cast<int32_t*>(base + kScratchRowOf);
  // the checks of the runs, before any state change
  hip_check(launch_scatter_rows(runs.pixel, runs.row_start, runs.row_count, runs.n, runs.n_samples, row_of,
                                static_cast<uint32_t>(P_), flags, stream),
            "scatter rows");
  hip_check(launch_fill_u64(chk, 2, 0, stream), "fill");
  hip_check(launch_check_runs(runs.irr, runs.row_start, runs.row_count, runs.n, runs.n_samples, G_, irr_max_, chk,
                              stream),
            "check samples");
  hip_check(hipMemcpyAsync(check_flags_, flags, 16, hipMemcpyDeviceToHost, stream), "read flags");
  hip_check(hipMemcpyAsync(check_samples_, chk, 16, hipMemcpyDeviceToHost, stream), "read checks");
  hip_check(hipStreamSynchronize(stream), "checks");
  finish_device_checks();
  if (check_samples_[0] != 0) {
    throw std::invalid_argument("SensorModel.advance: " + std::to_string(check_samples_[0]) +
                                " non-finite irradiance samples");
  }
  stats_.input_clamps += check_samples_[1];
  // one walk over every pixel in pixel order, exact searches only
  CameraDev& cd = noalloc_cam_;
  cd = CameraDev{};
  cd.kc = kconst();
  cd.mp = maps_view(d_maps_, P_);
  cd.in = state_view(cur_);
  cd.out = state_view(1 - cur_);
  cd.runs.row_of = row_of;
  cd.runs.row_start = runs.row_start;
  cd.runs.row_count = runs.row_count;
  cd.runs.t = runs.t;
  cd.runs.irr = runs.irr;
  cd.pixel = runs.pixel;
  cd.ev = events;
  cd.cap = cap;
  cd.counters = counters;
  cd.t_state = t_state_;
  cd.n_runs = static_cast<int64_t>(runs.n);
  noalloc_call_ = CallDev{t_end_ns, 0, 0};
  hip_check(hipMemcpyAsync(d_cam, &cd, sizeof(CameraDev), hipMemcpyHostToDevice, stream), "upload camera");
  hip_check(hipMemcpyAsync(d_call, &noalloc_call_, sizeof(CallDev), hipMemcpyHostToDevice, stream), "upload call");
  hip_check(launch_fill_u64(counters, 1 + kStCount, 0, stream), "fill");
  ChainArgs a{};
  a.cams = d_cam;
  a.n_cams = 1;
  a.n_total = static_cast<uint32_t>(P_);
  a.groups = G_;
  a.instant = instant();
  a.noise = noise_dev_.tab != nullptr && !instant();
  a.call = d_call;
  a.tu = chain_->tuning();
  a.tu.chunk = 0;
  hip_check(launch_chain_walk(a, stream), "walk");
  hip_check(hipMemcpyAsync(counters_host_, counters, sizeof(counters_host_), hipMemcpyDeviceToHost, stream),
            "read counters");
  hip_check(hipStreamSynchronize(stream), "advance");
  const unsigned long long n_events = counters_host_[0];
  if (n_events > cap) {
    throw std::invalid_argument("advance_device_noalloc: the call makes " + std::to_string(n_events) +
                                " events, more than the buffer's " + std::to_string(cap) +
                                " (the state is unchanged)");
  }
  ++stats_.launches;
  commit(counters_host_, t_end_ns);
  return n_events;
}

void SensorModel::download_events(const bsx_device_events& ev) {
  raw_host_.resize(ev.count);
  if (ev.count > 0) {
    hip_check(hipMemcpy(raw_host_.data(), ev.data, ev.count * sizeof(RawEventDev), hipMemcpyDeviceToHost),
              "download events");
  }
  finish_host_events(raw_host_.data(), ev.count);
}

void SensorModel::finish_host_events(const RawEventDev* raw, size_t n) {
  // (t_cross, pixel) is unique: a pixel never fires twice in one ns.
  std::vector<RawEventDev>& all = raw_sorted_;
  all.assign(raw, raw + n);
  std::sort(all.begin(), all.end(), [](const RawEventDev& a, const RawEventDev& b) {
    return a.t_cross != b.t_cross ? a.t_cross < b.t_cross : a.pixel < b.pixel;
  });
  events_.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const RawEventDev& e = all[i];
    events_.x[i] = static_cast<uint16_t>(e.pixel % width_);
    events_.y[i] = static_cast<uint16_t>(e.pixel / width_);
    events_.p[i] = e.polarity;
    events_.t_cross[i] = e.t_cross;
    events_.t_reported[i] = reported_time(readout_, e.t_cross);
    events_.label[i] = e.label;
    events_.flags[i] = e.flags;
    events_.ds[i] = e.ds;
    events_.dn[i] = e.dn;
    events_.dl[i] = e.dl;
    events_.l_first[i] = e.l_first;
  }
}

size_t SensorModel::state_bytes() const { return sizeof(StateHeader) + P_ * sizeof(SavedPixelState); }

void SensorModel::download_states(SavedPixelState* out, size_t n) const {
  if (n != P_) throw std::invalid_argument("SensorModel.download_states: one state per pixel");
  const StateDev s = state_view(cur_);
  std::vector<float> f(P_);
  std::vector<int64_t> q(P_);
  std::vector<uint32_t> u(P_);
  std::vector<uint16_t> fl(P_);
  SavedPixelState* ps = out;
  std::memset(static_cast<void*>(ps), 0, P_ * sizeof(SavedPixelState));
  const auto getf = [&](const float* src, double SavedPixelState::*dst) {
    hip_check(hipMemcpy(f.data(), src, P_ * 4, hipMemcpyDeviceToHost), "download state");
    for (size_t i = 0; i < P_; ++i) ps[i].*dst = static_cast<double>(f[i]);
  };
  const auto geti = [&](const int64_t* src, int64_t SavedPixelState::*dst) {
    hip_check(hipMemcpy(q.data(), src, P_ * 8, hipMemcpyDeviceToHost), "download state");
    for (size_t i = 0; i < P_; ++i) ps[i].*dst = q[i];
  };
  getf(s.z, &SavedPixelState::z);
  getf(s.x1, &SavedPixelState::x1);
  getf(s.s2, &SavedPixelState::s2);
  getf(s.s3, &SavedPixelState::s3);
  getf(s.i_scale, &SavedPixelState::i_scale);
  getf(s.tv, &SavedPixelState::tv);
  getf(s.s3_ref, &SavedPixelState::s3_ref);
  getf(s.n3_ref, &SavedPixelState::n3_ref);
  getf(s.dr_ref, &SavedPixelState::dr_ref);
  geti(s.t_ref, &SavedPixelState::t_ref);
  geti(s.blind_until, &SavedPixelState::blind_until);
  geti(s.t_search, &SavedPixelState::t_search);
  if (noise_) {
    // the in-path noise: the nodes, the cell's rate, the cell and whether its end node is drawn
    for (int c = 0; c < 3; ++c) {
      hip_check(hipMemcpy(f.data(), s.nk + static_cast<size_t>(c) * P_, P_ * 4, hipMemcpyDeviceToHost),
                "download state");
      for (size_t i = 0; i < P_; ++i) ps[i].nk[c] = f[i];
      hip_check(hipMemcpy(f.data(), s.nk1 + static_cast<size_t>(c) * P_, P_ * 4, hipMemcpyDeviceToHost),
                "download state");
      for (size_t i = 0; i < P_; ++i) ps[i].nk1[c] = f[i];
    }
    getf(s.ln_a1, &SavedPixelState::ln_a1);
    geti(s.cell, &SavedPixelState::cell);
    std::vector<uint8_t> hk(P_);
    hip_check(hipMemcpy(hk.data(), s.have_k1, P_, hipMemcpyDeviceToHost), "download state");
    for (size_t i = 0; i < P_; ++i) ps[i].have_k1 = hk[i];
  } else if (kc_.shot_rate_hz > 0.0) {
    // The shot-noise process's next arrival and index ride in fields the noise-free kernel leaves
    // unused (cell and nk[0]); without the process they stay 0, as the reference leaves them.
    geti(s.shot_next, &SavedPixelState::cell);
    hip_check(hipMemcpy(u.data(), s.shot_index, P_ * 4, hipMemcpyDeviceToHost), "download state");
    for (size_t i = 0; i < P_; ++i) ps[i].nk[0] = static_cast<double>(u[i]);
  }
  hip_check(hipMemcpy(fl.data(), s.flags, P_ * 2, hipMemcpyDeviceToHost), "download state");
  for (size_t i = 0; i < P_; ++i) ps[i].flags = fl[i];
  for (int g = 0; g < G_; ++g) {
    hip_check(hipMemcpy(f.data(), s.held + static_cast<size_t>(g) * P_, P_ * 4, hipMemcpyDeviceToHost),
              "download state");
    for (size_t i = 0; i < P_; ++i) ps[i].held[g] = f[i];
  }
}

void SensorModel::save_state(uint8_t* out, size_t n) const {
  if (from_constants_) throw std::logic_error("SensorModel.save_state: a model built from kernel constants moves its "
                                              "states with download_states");
  if (n < state_bytes()) throw std::invalid_argument("SensorModel.save_state: buffer too small");
  StateHeader h{};
  h.magic = kStateMagic;
  h.version = kStateVersion;
  h.width = width_;
  h.height = height_;
  h.groups = static_cast<uint32_t>(G_);
  h.camera = camera_;
  h.seed = seed_;
  h.fingerprint = fingerprint(params_, groups_, maps_.gain);
  h.t_state = t_state_;
  h.initialized = initialized_ ? 1 : 0;
  h.state_size = sizeof(SavedPixelState);
  std::memcpy(out, &h, sizeof(h));
  std::vector<SavedPixelState> ps(P_);
  download_states(ps.data(), P_);
  std::memcpy(out + sizeof(h), ps.data(), P_ * sizeof(SavedPixelState));
}

void SensorModel::upload_states(const SavedPixelState* ps, size_t n) {
  if (n != P_) throw std::invalid_argument("SensorModel.upload_states: one state per pixel");
  initialized_ = false;
  const StateDev s = state_view(cur_);
  std::vector<float> f(P_);
  std::vector<int64_t> q(P_);
  std::vector<uint32_t> u(P_);
  std::vector<uint16_t> fl(P_);
  const auto putf = [&](float* dst, double SavedPixelState::*src) {
    for (size_t i = 0; i < P_; ++i) f[i] = static_cast<float>(ps[i].*src);
    hip_check(hipMemcpy(dst, f.data(), P_ * 4, hipMemcpyHostToDevice), "upload state");
  };
  const auto puti = [&](int64_t* dst, int64_t SavedPixelState::*src) {
    for (size_t i = 0; i < P_; ++i) q[i] = ps[i].*src;
    hip_check(hipMemcpy(dst, q.data(), P_ * 8, hipMemcpyHostToDevice), "upload state");
  };
  putf(s.z, &SavedPixelState::z);
  putf(s.x1, &SavedPixelState::x1);
  putf(s.s2, &SavedPixelState::s2);
  putf(s.s3, &SavedPixelState::s3);
  putf(s.i_scale, &SavedPixelState::i_scale);
  putf(s.tv, &SavedPixelState::tv);
  putf(s.s3_ref, &SavedPixelState::s3_ref);
  putf(s.n3_ref, &SavedPixelState::n3_ref);
  putf(s.dr_ref, &SavedPixelState::dr_ref);
  puti(s.t_ref, &SavedPixelState::t_ref);
  puti(s.blind_until, &SavedPixelState::blind_until);
  puti(s.t_search, &SavedPixelState::t_search);
  if (noise_) {
    for (int c = 0; c < 3; ++c) {
      for (size_t i = 0; i < P_; ++i) f[i] = static_cast<float>(ps[i].nk[c]);
      hip_check(hipMemcpy(s.nk + static_cast<size_t>(c) * P_, f.data(), P_ * 4, hipMemcpyHostToDevice),
                "upload state");
      for (size_t i = 0; i < P_; ++i) f[i] = static_cast<float>(ps[i].nk1[c]);
      hip_check(hipMemcpy(s.nk1 + static_cast<size_t>(c) * P_, f.data(), P_ * 4, hipMemcpyHostToDevice),
                "upload state");
    }
    putf(s.ln_a1, &SavedPixelState::ln_a1);
    puti(s.cell, &SavedPixelState::cell);
    std::vector<uint8_t> hk(P_);
    for (size_t i = 0; i < P_; ++i) hk[i] = ps[i].have_k1;
    hip_check(hipMemcpy(s.have_k1, hk.data(), P_, hipMemcpyHostToDevice), "upload state");
  }
  if (kc_.shot_rate_hz > 0.0 && !noise_) {
    puti(s.shot_next, &SavedPixelState::cell);
    for (size_t i = 0; i < P_; ++i) u[i] = static_cast<uint32_t>(ps[i].nk[0]);
  } else {
    // no shot process: no arrival
    for (size_t i = 0; i < P_; ++i) q[i] = INT64_MAX;
    hip_check(hipMemcpy(s.shot_next, q.data(), P_ * 8, hipMemcpyHostToDevice), "upload state");
    std::fill(u.begin(), u.end(), 0u);
  }
  hip_check(hipMemcpy(s.shot_index, u.data(), P_ * 4, hipMemcpyHostToDevice), "upload state");
  for (size_t i = 0; i < P_; ++i) fl[i] = ps[i].flags;
  hip_check(hipMemcpy(s.flags, fl.data(), P_ * 2, hipMemcpyHostToDevice), "upload state");
  for (int g = 0; g < G_; ++g) {
    for (size_t i = 0; i < P_; ++i) f[i] = ps[i].held[g];
    hip_check(hipMemcpy(s.held + static_cast<size_t>(g) * P_, f.data(), P_ * 4, hipMemcpyHostToDevice),
              "upload state");
  }
  states_set_ = true;
}

void SensorModel::load_state(const uint8_t* data, size_t n) {
  if (from_constants_) throw std::logic_error("SensorModel.load_state: a model built from kernel constants moves its "
                                              "states with upload_states");
  if (n != state_bytes()) throw std::invalid_argument("SensorModel.load_state: size does not match this model");
  StateHeader h{};
  std::memcpy(&h, data, sizeof(h));
  if (h.magic != kStateMagic || h.version != kStateVersion || h.state_size != sizeof(SavedPixelState)) {
    throw std::invalid_argument("SensorModel.load_state: not a sensor state of this version");
  }
  if (h.width != width_ || h.height != height_ || h.groups != static_cast<uint32_t>(G_) || h.camera != camera_ ||
      h.seed != seed_ || h.fingerprint != fingerprint(params_, groups_, maps_.gain)) {
    throw std::invalid_argument("SensorModel.load_state: state was saved by a model with other parameters");
  }
  initialized_ = false;
  std::vector<SavedPixelState> ps(P_);
  std::memcpy(ps.data(), data + sizeof(h), P_ * sizeof(SavedPixelState));
  upload_states(ps.data(), P_);
  t_state_ = h.t_state;
  initialized_ = h.initialized != 0;
}

void SensorModel::reset_stats() { stats_ = bsx_stats{}; }

void SensorModel::set_policy(const bsx_launch_policy& p) {
  WalkTuning tu = chain_->tuning();
  if (p.block_size != 0 && (p.block_size % 64 != 0 || p.block_size < 64 || p.block_size > 256)) {
    throw std::invalid_argument("set_launch_policy: block_size must be a multiple of 64 in [64, 256]");
  }
  if (p.walk_budget < 0 || p.walk_budget > 1024) throw std::invalid_argument("set_launch_policy: walk_budget in [0, 1024]");
  if (p.search_batch < 0 || p.search_batch > 64) throw std::invalid_argument("set_launch_policy: search_batch in [0, 64]");
  if (p.queue_chunk < -1 || p.queue_chunk > 64) throw std::invalid_argument("set_launch_policy: queue_chunk in [-1, 64]");
  if (p.head_lanes < 0 || p.head_lanes > 64) throw std::invalid_argument("set_launch_policy: head_lanes in [0, 64]");
  if (p.head_per_mille < -1 || p.head_per_mille > 1000) {
    throw std::invalid_argument("set_launch_policy: head_per_mille in [-1, 1000]");
  }
  if (p.block_size != 0) tu.block = p.block_size;
  if (p.walk_budget != 0) tu.walk_budget = p.walk_budget;
  if (p.search_batch != 0) tu.search_batch = p.search_batch;
  if (p.queue_chunk != 0) tu.chunk = p.queue_chunk < 0 ? 0 : p.queue_chunk;
  if (p.head_lanes != 0) tu.head_lanes = p.head_lanes;
  if (p.head_per_mille != 0) tu.head_per_mille = p.head_per_mille < 0 ? 0 : p.head_per_mille;
  chain_->set_tuning(tu);
  if (p.event_capacity != 0) cap_ = p.event_capacity;
}

bsx_sensor_info_t SensorModel::info() const {
  bsx_sensor_info_t i{};
  i.width = width_;
  i.height = height_;
  i.groups = G_;
  i.initialized = initialized_ ? 1 : 0;
  i.time_ns = t_state_;
  i.readout_phase_ns = readout_.phase_ns;
  i.irr_max = irr_max_;
  i.camera = camera_;
  i.seed = seed_;
  return i;
}

}  // namespace bsx
