#include "params.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

#include "../core/pmath.h"

namespace bsx {

namespace {

[[noreturn]] void refuse(const std::string& what) { throw std::invalid_argument("SensorParams: " + what); }

void need(bool ok, const char* what) {
  if (!ok) refuse(what);
}

void set_family(SensorParams& p, const char* name) {
  std::memset(p.family, 0, sizeof(p.family));
  std::strncpy(p.family, name, sizeof(p.family) - 1);
}

// The prophesee_gen3 family's priors (reference: src/sensor/param_fields.cpp:369-470) and the
// literal rows' defaults (src/sensor/params.h:105-152).
SensorParams prophesee_gen3() {
  SensorParams p;
  std::memset(&p, 0, sizeof(p));
  set_family(p, "prophesee_gen3");
  p.photoreceptor = BSX_PHOTORECEPTOR_CASCADE;
  p.k = 1.0e16;
  p.tau_o = 25e-6;
  p.tau_sf = 100e-6;
  p.sigma_o = 0.06;
  p.noise = 1;
  p.illuminance_scale = 1.0;
  p.veiling_glare = 0.0;
  p.dark_current_fA = 0.21;
  p.dark_sigma_ln = 0.3;
  p.dark_ea_eV = 0.78;
  p.temperature_K = 298.15;
  p.reference_temperature_K = 298.15;
  p.theta_on = 0.25;
  p.theta_off = 0.25;
  p.sigma_theta_on = 0.03;
  p.sigma_theta_off = 0.03;
  p.theta_onoff_corr = 0.0;
  p.theta_floor = 0.01;
  p.reset_mode = BSX_RESET_PHYSICAL;
  p.refractory_s = 0.3e-3;
  p.eps_ci = 0.0;
  p.initial_phase = BSX_PHASE_UNIFORM;
  p.leak_rate_hz = 0.01;
  p.leak_sigma_log10 = 0.1;
  p.leak_ea_eV = 0.78;
  p.tau_hp = 0.0;
  p.tau_hp_sigma_ln = 0.0;
  p.hot_fraction = 0.0;
  p.hot_rate_law = 0;
  p.hot_rate_min_hz = 10.0;
  p.hot_rate_max_hz = 2000.0;
  p.hot_rate_median_hz = 100.0;
  p.hot_rate_sigma_ln = 1.5;
  p.hot_leak_fraction = 0.5;
  p.hot_reference_current_fA = 100.0;
  p.readout_quantum_s = 1e-6;
  p.readout_delay_s = 300e-6;
  p.clock_offset_ns = 0;
  p.clock_drift_ppm = 0.0;
  p.limiter = 0;
  p.eps_L = 0.002;
  p.noise_leaf_level = 5;
  p.noise_margin_z = 6.0;
  p.noise_diffusion = 0;
  p.initial_state = 0;
  p.label_share_cut = 0.5;
  p.shot_rate_hz = 0.0;
  return p;
}

}  // namespace

std::vector<std::string> preset_names() { return {"prophesee_gen3", "samsung_gen3", "ideal"}; }

SensorParams preset(const std::string& name) {
  SensorParams p = prophesee_gen3();
  if (name == "prophesee_gen3") return p;
  if (name == "samsung_gen3") {  // src/sensor/param_fields.cpp:471-498
    set_family(p, "samsung_gen3");
    p.theta_on = 0.20;
    p.theta_off = 0.20;
    p.refractory_s = 196e-6;
    p.readout_delay_s = 100e-6;
    p.leak_rate_hz = 0.01;
    p.dark_current_fA = 0.08;
    p.tau_hp = 0.0;
    return p;
  }
  if (name == "ideal") {  // src/sensor/param_fields.cpp:499-521
    set_family(p, "ideal");
    p.photoreceptor = BSX_PHOTORECEPTOR_INSTANTANEOUS;
    p.k = 1.0e18;
    p.tau_o = 1e-6;
    p.tau_sf = 2e-6;
    p.noise = 0;
    p.dark_current_fA = 0.01;
    p.dark_sigma_ln = 0.0;
    p.theta_on = 0.20;
    p.theta_off = 0.20;
    p.sigma_theta_on = 0.0;
    p.sigma_theta_off = 0.0;
    p.reset_mode = BSX_RESET_IDEAL;
    p.refractory_s = 0.0;
    p.leak_rate_hz = 0.0;
    p.leak_sigma_log10 = 0.0;
    p.tau_hp = 0.0;
    p.hot_fraction = 0.0;
    p.readout_quantum_s = 0.0;
    p.readout_delay_s = 0.0;
    return p;
  }
  throw std::invalid_argument("unknown sensor preset '" + name + "' (known: prophesee_gen3, samsung_gen3, ideal)");
}

void validate_ranges(const SensorParams& p) {
  const std::string fam(p.family, strnlen(p.family, sizeof(p.family)));
  bool known = false;
  for (const std::string& n : preset_names()) known = known || n == fam;
  if (!known) refuse("family = '" + fam + "': one of prophesee_gen3, samsung_gen3, ideal");
  const auto finite = [](double v) { return std::isfinite(v); };
  // The rows' hard ranges (src/sensor/param_fields.cpp rows).
  need(finite(p.k) && p.k > 0, "k must be > 0");
  need(finite(p.tau_o) && p.tau_o > 0, "tau_o must be > 0");
  need(finite(p.tau_sf) && p.tau_sf > 0, "tau_sf must be > 0");
  need(finite(p.sigma_o) && p.sigma_o >= 0, "sigma_o must be >= 0");
  need(finite(p.illuminance_scale) && p.illuminance_scale > 0, "illuminance_scale must be > 0");
  need(finite(p.veiling_glare) && p.veiling_glare >= 0 && p.veiling_glare < 1, "veiling_glare must lie in [0, 1)");
  need(finite(p.dark_current_fA) && p.dark_current_fA > 0, "dark_current_fA must be > 0");
  need(finite(p.dark_sigma_ln) && p.dark_sigma_ln >= 0, "dark_sigma_ln must be >= 0");
  need(finite(p.dark_ea_eV) && p.dark_ea_eV >= 0, "dark_ea_eV must be >= 0");
  need(finite(p.temperature_K) && p.temperature_K > 0, "temperature_K must be > 0");
  need(finite(p.reference_temperature_K) && p.reference_temperature_K > 0, "reference_temperature_K must be > 0");
  need(finite(p.theta_on) && p.theta_on > 0, "theta_on must be > 0");
  need(finite(p.theta_off) && p.theta_off > 0, "theta_off must be > 0");
  need(finite(p.sigma_theta_on) && p.sigma_theta_on >= 0, "sigma_theta_on must be >= 0");
  need(finite(p.sigma_theta_off) && p.sigma_theta_off >= 0, "sigma_theta_off must be >= 0");
  need(finite(p.theta_onoff_corr) && p.theta_onoff_corr >= -1 && p.theta_onoff_corr <= 1,
       "theta_onoff_corr must lie in [-1, 1]");
  need(finite(p.theta_floor) && p.theta_floor > 0, "theta_floor must be > 0");
  need(finite(p.refractory_s) && p.refractory_s >= 0 && p.refractory_s < 10, "refractory_s must lie in [0, 10)");
  need(finite(p.eps_ci), "eps_ci must be finite");
  need(finite(p.leak_rate_hz) && p.leak_rate_hz >= 0, "leak_rate_hz must be >= 0");
  need(finite(p.leak_sigma_log10) && p.leak_sigma_log10 >= 0, "leak_sigma_log10 must be >= 0");
  need(finite(p.leak_ea_eV) && p.leak_ea_eV >= 0, "leak_ea_eV must be >= 0");
  need(finite(p.tau_hp) && p.tau_hp >= 0 && p.tau_hp <= 100, "tau_hp must lie in [0, 100]");
  need(finite(p.hot_fraction) && p.hot_fraction >= 0 && p.hot_fraction <= 1, "hot_fraction must lie in [0, 1]");
  need(finite(p.hot_rate_min_hz) && p.hot_rate_min_hz > 0, "hot_rate_min_hz must be > 0");
  need(finite(p.hot_rate_max_hz) && p.hot_rate_max_hz > 0, "hot_rate_max_hz must be > 0");
  need(p.hot_rate_law == 0 || p.hot_rate_law == 1, "hot_rate_law must be 0 (log-uniform) or 1 (log-normal)");
  need(finite(p.hot_rate_median_hz) && p.hot_rate_median_hz > 0, "hot_rate_median_hz must be > 0");
  need(finite(p.hot_rate_sigma_ln) && p.hot_rate_sigma_ln >= 0, "hot_rate_sigma_ln must be >= 0");
  need(finite(p.hot_leak_fraction) && p.hot_leak_fraction >= 0 && p.hot_leak_fraction <= 1,
       "hot_leak_fraction must lie in [0, 1]");
  need(finite(p.readout_quantum_s) && p.readout_quantum_s >= 0, "readout_quantum_s must be >= 0");
  need(finite(p.readout_delay_s) && p.readout_delay_s >= 0, "readout_delay_s must be >= 0");
  need(finite(p.clock_drift_ppm), "clock_drift_ppm must be finite");
  need(finite(p.eps_L) && p.eps_L > 0, "eps_L must be > 0");
  need(p.noise_leaf_level >= 0 && p.noise_leaf_level <= 10, "noise_leaf_level must lie in [0, 10]");
  need(finite(p.noise_margin_z) && p.noise_margin_z >= 0, "noise_margin_z must be >= 0");
  need(finite(p.label_share_cut) && p.label_share_cut > 0 && p.label_share_cut <= 1,
       "label_share_cut must lie in (0, 1]");
  need(finite(p.shot_rate_hz) && p.shot_rate_hz >= 0, "shot_rate_hz must be >= 0");
  // Cross-key rules (src/sensor/params.cpp:52-92).
  need(std::fmax(p.tau_o, p.tau_sf) >= 1.1 * std::fmin(p.tau_o, p.tau_sf),
       "tau_sf and tau_o must differ by at least 10 % (the larger >= 1.1 x the smaller)");
  need(!(p.noise && p.sigma_o > 0) || p.tau_o >= 1e-6, "sigma_o > 0 requires tau_o >= 1 us");
  need(p.theta_on > p.theta_floor, "theta_on must exceed theta_floor");
  need(p.theta_off > p.theta_floor, "theta_off must exceed theta_floor");
  need(std::fabs(p.eps_ci) < 0.5 * p.theta_floor, "|eps_ci| must be < theta_floor / 2");
  need(p.hot_rate_max_hz >= p.hot_rate_min_hz, "hot rates must satisfy 0 < hot_rate_min_hz <= hot_rate_max_hz");
  need(!(p.hot_fraction != 0) || p.hot_rate_max_hz * p.refractory_s < 1.0,
       "hot_rate_max_hz must be < 1 / refractory_s");
  need(p.readout_quantum_s == 0 || p.readout_delay_s >= p.readout_quantum_s,
       "readout_delay_s must be >= readout_quantum_s when quantization is on");
}

void validate(const SensorParams& p) {
  validate_ranges(p);
  // Features of the reference this backend does not run: refused by name, never approximated.
  need(p.veiling_glare == 0, "veiling_glare > 0 is not supported by the bsx backend");
  need(p.tau_hp == 0, "the high-pass filter (tau_hp > 0) is not supported by the bsx backend");
  need(p.limiter == 0, "the rate limiter is not supported by the bsx backend");
  need(p.initial_state == 0, "initial_state stationary is not supported by the bsx backend");
  need(!noise_on(p) || p.noise_diffusion == 0,
       "noise_diffusion transient (the in-path noise's transient rule) is not supported by the bsx backend");
  need(!noise_on(p) || p.hot_fraction == 0 || p.hot_leak_fraction >= 1.0,
       "threshold-type hot pixels (hot_leak_fraction < 1 with the in-path noise) are not supported by the bsx backend");
  need(p.shot_rate_hz == 0 || p.photoreceptor == BSX_PHOTORECEPTOR_INSTANTANEOUS,
       "shot_rate_hz > 0 (the shot-noise process) needs the instantaneous photoreceptor");
}

double arrhenius(double ea_eV, double temperature_K, double reference_K) {
  constexpr double kBoltzmannEv = 8.617333262e-5;
  return pm::exp(-ea_eV / kBoltzmannEv * (1.0 / temperature_K - 1.0 / reference_K));
}

double median_dark_fA(const SensorParams& p) {
  return p.dark_current_fA * arrhenius(p.dark_ea_eV, p.temperature_K, p.reference_temperature_K);
}

}  // namespace bsx
