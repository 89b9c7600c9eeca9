// A C entry layer over the reference's own SensorModel (src/sensor/sensor_model.h), linked with the
// reference's sources where they lie (oracle/Makefile): the compiled reference, used by the tests
// to pin the C restatement (bsx_oracle.c) and, where it is present, as the timed CPU baseline.
// Test infrastructure only; nothing of the product links it.
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/bsx/bsx_sensor.h"
#include "sensor/noise.h"
#include "sensor/noise_tables.h"
#include "sensor/output_order.h"
#include "sensor/params.h"
#include "sensor/readout.h"
#include "sensor/sensor_model.h"

using namespace balla::sensor;

namespace {

struct RefModel {
  std::unique_ptr<SensorModel> m;
  std::vector<RawEvent> ev;
};

void put_err(char* err, size_t n, const char* msg) {
  if (err != nullptr && n > 0) {
    std::strncpy(err, msg, n - 1);
    err[n - 1] = '\0';
  }
}

SensorParams to_ref(const bsx_params& b) {
  SensorParams p = preset(std::string(b.family, strnlen(b.family, sizeof(b.family))));
  p.photoreceptor = b.photoreceptor == BSX_PHOTORECEPTOR_INSTANTANEOUS ? Photoreceptor::kInstantaneous
                                                                       : Photoreceptor::kCascade;
  p.noise = b.noise != 0;
  p.k = b.k;
  p.tau_o = b.tau_o;
  p.tau_sf = b.tau_sf;
  p.sigma_o = b.sigma_o;
  p.illuminance_scale = b.illuminance_scale;
  p.veiling_glare = b.veiling_glare;
  p.dark_current_fA = b.dark_current_fA;
  p.dark_sigma_ln = b.dark_sigma_ln;
  p.dark_ea_eV = b.dark_ea_eV;
  p.temperature_K = b.temperature_K;
  p.reference_temperature_K = b.reference_temperature_K;
  p.theta_on = b.theta_on;
  p.theta_off = b.theta_off;
  p.sigma_theta_on = b.sigma_theta_on;
  p.sigma_theta_off = b.sigma_theta_off;
  p.theta_onoff_corr = b.theta_onoff_corr;
  p.theta_floor = b.theta_floor;
  p.reset_mode = b.reset_mode == BSX_RESET_IDEAL ? ResetMode::kIdeal : ResetMode::kPhysical;
  p.initial_phase = b.initial_phase == BSX_PHASE_SYMMETRIC ? InitialPhase::kSymmetric
                    : b.initial_phase == BSX_PHASE_ZERO    ? InitialPhase::kZero
                                                           : InitialPhase::kUniform;
  p.refractory_s = b.refractory_s;
  p.eps_ci = b.eps_ci;
  p.leak_rate_hz = b.leak_rate_hz;
  p.leak_sigma_log10 = b.leak_sigma_log10;
  p.leak_ea_eV = b.leak_ea_eV;
  p.tau_hp = b.tau_hp;
  p.hot_fraction = b.hot_fraction;
  p.hot_rate_law = b.hot_rate_law == 1 ? HotRateLaw::kLogNormal : HotRateLaw::kLogUniform;
  p.hot_rate_min_hz = b.hot_rate_min_hz;
  p.hot_rate_max_hz = b.hot_rate_max_hz;
  p.hot_rate_median_hz = b.hot_rate_median_hz;
  p.hot_rate_sigma_ln = b.hot_rate_sigma_ln;
  p.hot_leak_fraction = b.hot_leak_fraction;
  p.hot_reference_current_fA = b.hot_reference_current_fA;
  p.readout_quantum_s = b.readout_quantum_s;
  p.readout_delay_s = b.readout_delay_s;
  p.clock_offset_ns = b.clock_offset_ns;
  p.clock_drift_ppm = b.clock_drift_ppm;
  p.eps_L = b.eps_L;
  p.noise_leaf_level = b.noise_leaf_level;
  p.noise_margin_z = b.noise_margin_z;
  p.label_share_cut = b.label_share_cut;
  return p;
}

}  // namespace

extern "C" {

// The reference's preset(name) in bsx_params form (for the presets' parity test).
int ref_preset(const char* name, bsx_params* b) {
  try {
    const SensorParams p = preset(name);
    std::memset(b, 0, sizeof(*b));
    std::strncpy(b->family, p.family.c_str(), sizeof(b->family) - 1);
    b->photoreceptor = p.photoreceptor == Photoreceptor::kInstantaneous ? BSX_PHOTORECEPTOR_INSTANTANEOUS
                                                                        : BSX_PHOTORECEPTOR_CASCADE;
    b->noise = p.noise ? 1 : 0;
    b->k = p.k;
    b->tau_o = p.tau_o;
    b->tau_sf = p.tau_sf;
    b->sigma_o = p.sigma_o;
    b->illuminance_scale = p.illuminance_scale;
    b->veiling_glare = p.veiling_glare;
    b->dark_current_fA = p.dark_current_fA;
    b->dark_sigma_ln = p.dark_sigma_ln;
    b->dark_ea_eV = p.dark_ea_eV;
    b->temperature_K = p.temperature_K;
    b->reference_temperature_K = p.reference_temperature_K;
    b->theta_on = p.theta_on;
    b->theta_off = p.theta_off;
    b->sigma_theta_on = p.sigma_theta_on;
    b->sigma_theta_off = p.sigma_theta_off;
    b->theta_onoff_corr = p.theta_onoff_corr;
    b->theta_floor = p.theta_floor;
    b->reset_mode = p.reset_mode == ResetMode::kIdeal ? BSX_RESET_IDEAL : BSX_RESET_PHYSICAL;
    b->initial_phase = p.initial_phase == InitialPhase::kSymmetric ? BSX_PHASE_SYMMETRIC
                       : p.initial_phase == InitialPhase::kZero    ? BSX_PHASE_ZERO
                                                                   : BSX_PHASE_UNIFORM;
    b->refractory_s = p.refractory_s;
    b->eps_ci = p.eps_ci;
    b->leak_rate_hz = p.leak_rate_hz;
    b->leak_sigma_log10 = p.leak_sigma_log10;
    b->leak_ea_eV = p.leak_ea_eV;
    b->tau_hp = p.tau_hp;
    b->tau_hp_sigma_ln = p.tau_hp_sigma_ln;
    b->hot_fraction = p.hot_fraction;
    b->hot_rate_law = static_cast<int32_t>(p.hot_rate_law);
    b->limiter = static_cast<int32_t>(p.limiter);
    b->hot_rate_min_hz = p.hot_rate_min_hz;
    b->hot_rate_max_hz = p.hot_rate_max_hz;
    b->hot_rate_median_hz = p.hot_rate_median_hz;
    b->hot_rate_sigma_ln = p.hot_rate_sigma_ln;
    b->hot_leak_fraction = p.hot_leak_fraction;
    b->hot_reference_current_fA = p.hot_reference_current_fA;
    b->readout_quantum_s = p.readout_quantum_s;
    b->readout_delay_s = p.readout_delay_s;
    b->clock_offset_ns = p.clock_offset_ns;
    b->clock_drift_ppm = p.clock_drift_ppm;
    b->eps_L = p.eps_L;
    b->noise_leaf_level = p.noise_leaf_level;
    b->noise_diffusion = static_cast<int32_t>(p.noise_diffusion);
    b->noise_margin_z = p.noise_margin_z;
    b->initial_state = static_cast<int32_t>(p.initial_state);
    b->label_share_cut = p.label_share_cut;
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

void* ref_create(const bsx_params* b, uint32_t width, uint32_t height, uint64_t seed, uint32_t camera,
                 int32_t groups, int32_t num_threads, int32_t fp32, char* err, size_t err_n) {
  try {
    auto r = std::make_unique<RefModel>();
    std::vector<Waveform> g(static_cast<size_t>(groups), make_constant_waveform());
    r->m = std::make_unique<SensorModel>(to_ref(*b), width, height, seed, std::move(g), camera, num_threads,
                                         fp32 ? Precision::kFp32 : Precision::kFp64, Device::kCpu);
    return r.release();
  } catch (const std::exception& e) {
    put_err(err, err_n, e.what());
    return nullptr;
  }
}

void ref_destroy(void* h) { delete static_cast<RefModel*>(h); }

int ref_init(void* h, int64_t t0, const float* irr0, char* err, size_t err_n) {
  try {
    static_cast<RefModel*>(h)->m->init(t0, irr0);
    return 0;
  } catch (const std::exception& e) {
    put_err(err, err_n, e.what());
    return 1;
  }
}

int ref_advance_runs(void* h, const uint32_t* ids, size_t n, const int64_t* t, const float* irr, size_t n_samples,
                     const int64_t* row_start, const int32_t* row_count, int64_t t_end, char* err, size_t err_n) {
  RefModel* r = static_cast<RefModel*>(h);
  try {
    r->ev = r->m->advance_runs_raw(ids, n, t, irr, n_samples, row_start, row_count, t_end);
    return 0;
  } catch (const std::invalid_argument& e) {
    put_err(err, err_n, e.what());
    return 1;
  } catch (const std::exception& e) {
    put_err(err, err_n, e.what());
    return 2;
  }
}

size_t ref_event_count(void* h) { return static_cast<RefModel*>(h)->ev.size(); }

// The reference's noise tables for b's noise model (frozen diffusion): n_nodes rows of stride values, fp64 and
// fp32 (either output may be null).
int ref_noise_tables(const bsx_params* b, double* out64, float* out32, int32_t* n_nodes, int32_t* stride) {
  try {
    NoiseModel nm{};
    nm.kq = b->k * 1.602176634e-19;
    nm.a2 = 1.0 / b->tau_o;
    nm.a3 = 1.0 / b->tau_sf;
    nm.sigma_o = b->sigma_o;
    nm.levels = b->noise_leaf_level;
    nm.margin_z = b->noise_margin_z;
    nm.split = false;
    const NoiseTables t(nm);
    *n_nodes = t.n_nodes();
    *stride = t.layout().stride();
    if (out64 != nullptr) std::memcpy(out64, t.data64().data(), t.data64().size() * sizeof(double));
    if (out32 != nullptr) std::memcpy(out32, t.data32().data(), t.data32().size() * sizeof(float));
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

// The reference's keyed normals (noise.h normals3) in fp64 and fp32 (either output may be null).
void ref_normals3(uint64_t seed, uint32_t purpose, uint32_t camera, uint32_t pixel, uint64_t index, double* out64,
                  float* out32) {
  const auto pu = static_cast<balla::rng::Purpose>(purpose);
  if (out64 != nullptr) normals3<double>(seed, pu, camera, pixel, index, out64);
  if (out32 != nullptr) normals3<float>(seed, pu, camera, pixel, index, out32);
}

// The sorted raw events: the reference's RawEvent has the layout of bsx_raw_event.
void ref_events(void* h, bsx_raw_event* out) {
  const RefModel* r = static_cast<RefModel*>(h);
  static_assert(sizeof(RawEvent) == sizeof(bsx_raw_event), "RawEvent layout");
  if (!r->ev.empty()) std::memcpy(out, r->ev.data(), r->ev.size() * sizeof(RawEvent));
}

void ref_reported(void* h, int64_t* out) {
  const RefModel* r = static_cast<RefModel*>(h);
  for (size_t i = 0; i < r->ev.size(); ++i) out[i] = r->m->reported_time(r->ev[i].t_cross);
}

void ref_maps(void* h, double* theta_on, double* theta_off, double* leak, double* dark, double* phi0, uint8_t* hot,
              double* hot_rate) {
  const SensorMaps& m = static_cast<RefModel*>(h)->m->maps();
  for (size_t i = 0; i < m.pixels.size(); ++i) {
    hot[i] = static_cast<uint8_t>(m.pixels[i].hot);
    hot_rate[i] = m.hot_target_hz[i];
    theta_on[i] = m.pixels[i].theta_on;
    theta_off[i] = m.pixels[i].theta_off;
    leak[i] = m.pixels[i].leak;
    dark[i] = m.pixels[i].dark_fA;
    phi0[i] = m.pixels[i].phi0;
  }
}

// The states, 200 bytes per pixel in the reference's PixelState layout.
void ref_states(void* h, uint8_t* out) {
  const std::vector<PixelState> s = static_cast<RefModel*>(h)->m->states();
  std::memcpy(out, s.data(), s.size() * sizeof(PixelState));
}

void ref_stats(void* h, uint64_t* out) {
  const KernelStats& s = static_cast<RefModel*>(h)->m->stats();
  const uint64_t v[10] = {s.substeps, s.substep_attempts, s.settled_substeps, s.cells,      s.refinements,
                          s.leaf_intervals, s.fallbacks, s.input_clamps,     s.quarantines, s.pieces};
  std::memcpy(out, v, sizeof(v));
}

// The reference's output order (output_order.h) of n events: perm[i] = the index of the i-th event.
int ref_output_order(const int64_t* t_cross, const uint16_t* x, const uint16_t* y, const int64_t* t_reported,
                     size_t n, uint32_t height, uint32_t* perm) {
  std::vector<uint32_t> p;
  output_order({t_cross, n}, {x, n}, {y, n}, {t_reported, n}, order_y_bits(height), p);
  std::memcpy(perm, p.data(), n * sizeof(uint32_t));
  return 0;
}

// reported_time of n crossings, and the window cut of t_end, under one readout.
void ref_readout(int64_t quantum, int64_t delay, int64_t phase, int64_t offset, double drift_ppm,
                 const int64_t* t_cross, size_t n, int64_t* out, int64_t t_end, int64_t* cut) {
  const Readout r{quantum, delay, phase, offset, drift_ppm};
  for (size_t i = 0; i < n; ++i) out[i] = reported_time(r, t_cross[i]);
  *cut = output_cut(r, t_end);
}

// The S1 label of each sensor label (label_s1).
void ref_label_s1(const uint8_t* c3, size_t n, uint8_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = label_s1(c3[i]);
}

}  // extern "C"
