#include "maps.h"

#include <cmath>
#include <thread>

#include "../core/philox.h"
#include "../core/pmath.h"

namespace bsx {

namespace {

constexpr double kTwoPi = 0x1.921fb54442d18p+2;
// Philox blocks of kSensorMap (reference: src/sensor/maps.cpp:21-23).
constexpr uint64_t kMapNormals = 0;   // xi_on, xi_off, xi_leak, xi_dark
constexpr uint64_t kMapUniforms = 1;  // u_phase, u_hot, u_type, u_rate
constexpr uint64_t kMapHotRate = 2;   // xi_rate of the log-normal hot law (hot pixels only)

// Four normals of one block of kSensorMap by Box-Muller on two uniform pairs.
void map_normals(uint64_t seed, uint32_t camera, uint32_t pixel, uint64_t block, double* xi) {
  const Word4 r = draw(seed, kPurposeSensorMap, camera, pixel, block);
  const double u0 = open_uniform64(r.v[0]), u1 = open_uniform64(r.v[1]);
  const double u2 = open_uniform64(r.v[2]), u3 = open_uniform64(r.v[3]);
  const double r0 = std::sqrt(-2.0 * pm::log(u0)), r1 = std::sqrt(-2.0 * pm::log(u2));
  double s0, c0, s1, c1;
  pm::sincos(kTwoPi * u1, &s0, &c0);
  pm::sincos(kTwoPi * u3, &s1, &c1);
  xi[0] = r0 * c0;
  xi[1] = r0 * s0;
  xi[2] = r1 * c1;
  xi[3] = r1 * s1;
}

inline double dmax(double a, double b) { return a < b ? b : a; }
inline double dclamp(double v, double lo, double hi) { return v < lo ? lo : hi < v ? hi : v; }

}  // namespace

SensorMaps draw_sensor_maps(const SensorParams& p, uint64_t seed, uint32_t camera, uint32_t n_pixels,
                            int num_threads) {
  SensorMaps m;
  m.theta_on.resize(n_pixels);
  m.theta_off.resize(n_pixels);
  m.leak.resize(n_pixels);
  m.dark_fA.resize(n_pixels);
  m.phi0.resize(n_pixels);
  m.hot.assign(n_pixels, 0);
  m.hot_rate_hz.assign(n_pixels, 0.0);
  m.gain.assign(n_pixels, 1.0f);
  const double leak_temp = arrhenius(p.leak_ea_eV, p.temperature_K, p.reference_temperature_K);
  const double dark_temp = arrhenius(p.dark_ea_eV, p.temperature_K, p.reference_temperature_K);
  const double rho = p.theta_onoff_corr, rho_c = std::sqrt(1.0 - rho * rho);
  const double ln10 = pm::log(10.0);
  const double ln_rmin = pm::log(p.hot_rate_min_hz), ln_rmax = pm::log(p.hot_rate_max_hz);
  const auto one = [&](uint32_t px) {
    double xi[4];
    map_normals(seed, camera, px, kMapNormals, xi);
    const Word4 u = draw(seed, kPurposeSensorMap, camera, px, kMapUniforms);
    const double u_phase = open_uniform64(u.v[0]), u_hot = open_uniform64(u.v[1]);
    const double u_rate = open_uniform64(u.v[3]);
    const double xi_off = rho == 0.0 ? xi[1] : rho * xi[0] + rho_c * xi[1];
    const double th_on = dmax(p.theta_floor, p.theta_on + p.sigma_theta_on * xi[0]);
    const double th_off = dmax(p.theta_floor, p.theta_off + p.sigma_theta_off * xi_off);
    double leak_rate = p.leak_rate_hz * pm::exp(p.leak_sigma_log10 * xi[2] * ln10) * leak_temp;
    if (u_hot < p.hot_fraction) {
      // A hot pixel: its target rate log-uniform on [min, max], or log-normal clamped to it. Without
      // the cascade's noise (the only mode this library runs) every hot pixel is leak-type (the
      // reference's rule: u_type < hot_leak_fraction, or noise off): its leak cycle meets the rate,
      // 1 / (T_r + 1 / R_L) = R at R_L = R / (1 - R T_r) under the physical reset, R_L = R under the
      // ideal one (validate keeps R T_r < 1).
      double rate;
      if (p.hot_rate_law == 0) {
        rate = pm::exp(ln_rmin + u_rate * (ln_rmax - ln_rmin));
      } else {
        double z[4];
        map_normals(seed, camera, px, kMapHotRate, z);
        rate = dclamp(p.hot_rate_median_hz * pm::exp(p.hot_rate_sigma_ln * z[0]), p.hot_rate_min_hz, p.hot_rate_max_hz);
      }
      m.hot_rate_hz[px] = rate;
      m.hot[px] = 1;
      leak_rate = p.reset_mode == BSX_RESET_PHYSICAL ? rate / (1.0 - rate * p.refractory_s) : rate;
    }
    m.theta_on[px] = th_on;
    m.theta_off[px] = th_off;
    m.dark_fA[px] = p.dark_current_fA * pm::exp(p.dark_sigma_ln * xi[3]) * dark_temp;
    m.leak[px] = th_on * leak_rate;  // the ON drift theta_on R_L of the reference
    m.phi0[px] = p.initial_phase == BSX_PHASE_UNIFORM     ? u_phase * th_on
                 : p.initial_phase == BSX_PHASE_SYMMETRIC ? -th_off + u_phase * (th_on + th_off)
                                                          : 0.0;
  };
  int workers = num_threads > 0 ? num_threads : static_cast<int>(std::thread::hardware_concurrency());
  if (workers < 1) workers = 1;
  if (n_pixels < 65536u) workers = 1;
  if (workers == 1) {
    for (uint32_t px = 0; px < n_pixels; ++px) one(px);
    return m;
  }
  std::vector<std::thread> pool;
  for (int w = 0; w < workers; ++w) {
    pool.emplace_back([&, w] {
      const uint32_t b = static_cast<uint32_t>(static_cast<uint64_t>(n_pixels) * w / workers);
      const uint32_t e = static_cast<uint32_t>(static_cast<uint64_t>(n_pixels) * (w + 1) / workers);
      for (uint32_t px = b; px < e; ++px) one(px);
    });
  }
  for (std::thread& t : pool) t.join();
  return m;
}

Readout readout_of(const SensorParams& p, uint64_t seed, uint32_t camera) {
  Readout r{};
  r.quantum_ns = std::llround(p.readout_quantum_s * 1e9);
  r.delay_ns = std::llround(p.readout_delay_s * 1e9);
  r.phase_ns = r.quantum_ns > 0 ? static_cast<int64_t>(draw(seed, kPurposeReadoutPhase, camera, 0, 0).v[0] %
                                                       static_cast<uint64_t>(r.quantum_ns))
                                : 0;
  r.clock_offset_ns = p.clock_offset_ns;
  r.clock_drift_ppm = p.clock_drift_ppm;
  return r;
}

}  // namespace bsx
