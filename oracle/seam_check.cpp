// The reference's own SensorModel on both of its devices (TEST INFRASTRUCTURE): built from the
// reference's sources with BALLASIM_SENSOR_CUDA=1 and ballasim_amd/seam/sensor_bsx.cpp in place of its
// CUDA backend (oracle/Makefile, target seam), so Device::kCuda runs on libbsx.so. The same seeded calls
// go to a Precision::kFp32 model on Device::kCpu (the reference's CPU path) and on Device::kCuda; their
// events (every field, reported times included), saved states, work counters and the frame means of the
// held inputs (the backend's end_frame on kCuda) must be identical.
//
//   seam_check build   constructs the kCuda model: without a device the reference's own runtime_error
//                      must come back (prints "no device: ..."), with one it must construct
//   seam_check run     the parity cases (prints one line per case; exit status 1 on any difference)
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "sensor/params.h"
#include "sensor/sensor_model.h"

using namespace balla::sensor;

namespace {

struct Case {
  const char* name;
  const char* family;
  void (*tweak)(SensorParams&);
};

const Case kCases[] = {
    {"prophesee_gen3", "prophesee_gen3", [](SensorParams&) {}},
    {"prophesee_noise_off", "prophesee_gen3", [](SensorParams& p) { p.noise = false; }},
    {"prophesee_noise_ideal_leak", "prophesee_gen3",
     [](SensorParams& p) {
       p.reset_mode = ResetMode::kIdeal;
       p.refractory_s = 1e-5;
       p.leak_rate_hz = 0.5;
     }},
    {"ideal_physical_leak", "ideal",
     [](SensorParams& p) {
       p.reset_mode = ResetMode::kPhysical;
       p.refractory_s = 1e-4;
       p.leak_rate_hz = 0.3;
     }},
};

struct Runs {
  std::vector<uint32_t> ids;
  std::vector<int64_t> t, start;
  std::vector<int32_t> count;
  std::vector<float> irr;
};

Runs random_runs(std::mt19937_64& rng, uint32_t P, int64_t t0, int64_t t1) {
  Runs r;
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::normal_distribution<double> N(3.0, 1.5);
  for (uint32_t p = 0; p < P; ++p) {
    if (U(rng) > 0.8) continue;
    const int n = static_cast<int>(U(rng) * 12);
    std::vector<int64_t> ts;
    for (int j = 0; j < n; ++j) ts.push_back(t0 + static_cast<int64_t>(U(rng) * static_cast<double>(t1 - t0)));
    std::sort(ts.begin(), ts.end());
    ts.erase(std::unique(ts.begin(), ts.end()), ts.end());
    r.ids.push_back(p);
    r.start.push_back(static_cast<int64_t>(r.t.size()));
    r.count.push_back(static_cast<int32_t>(ts.size()));
    for (const int64_t t : ts) {
      r.t.push_back(t);
      r.irr.push_back(static_cast<float>(std::exp(N(rng))));
    }
  }
  return r;
}

bool same(const Events& a, const Events& b, std::string& why) {
  if (a.size() != b.size()) {
    why = "event count " + std::to_string(a.size()) + " vs " + std::to_string(b.size());
    return false;
  }
  for (size_t i = 0; i < a.size(); ++i) {
    const bool eq = a.x[i] == b.x[i] && a.y[i] == b.y[i] && a.p[i] == b.p[i] && a.t_cross_ns[i] == b.t_cross_ns[i] &&
                    a.t_reported_ns[i] == b.t_reported_ns[i] && a.label[i] == b.label[i] && a.flags[i] == b.flags[i] &&
                    std::memcmp(&a.ds[i], &b.ds[i], 4) == 0 && std::memcmp(&a.dn[i], &b.dn[i], 4) == 0 &&
                    std::memcmp(&a.dl[i], &b.dl[i], 4) == 0 && std::memcmp(&a.l_first[i], &b.l_first[i], 4) == 0;
    if (!eq) {
      why = "event " + std::to_string(i) + " (t " + std::to_string(a.t_cross_ns[i]) + " vs " +
            std::to_string(b.t_cross_ns[i]) + ")";
      return false;
    }
  }
  return true;
}

bool same_stats(const KernelStats& a, const KernelStats& b, bool instant, std::string& why) {
  const uint64_t va[] = {a.substeps, a.substep_attempts, a.settled_substeps, a.cells, a.refinements,
                         instant ? 0 : a.leaf_intervals, a.fallbacks, a.input_clamps, a.quarantines, a.pieces};
  const uint64_t vb[] = {b.substeps, b.substep_attempts, b.settled_substeps, b.cells, b.refinements,
                         instant ? 0 : b.leaf_intervals, b.fallbacks, b.input_clamps, b.quarantines, b.pieces};
  for (int i = 0; i < 10; ++i) {
    if (va[i] != vb[i]) {
      why = "counter " + std::to_string(i) + ": " + std::to_string(va[i]) + " vs " + std::to_string(vb[i]);
      return false;
    }
  }
  return true;
}

template <class T>
T* to_device(const std::vector<T>& v) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc");
  if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
    throw std::runtime_error("hipMemcpy");
  }
  return static_cast<T*>(p);
}

int run_case(const Case& c) {
  SensorParams p = preset(c.family);
  c.tweak(p);
  const uint32_t W = 48, H = 32, P = W * H;
  const uint64_t seed = 1234 + std::strlen(c.name);
  std::vector<Waveform> groups(1);
  SensorModel cpu(p, W, H, seed, groups, 1, 2, Precision::kFp32, Device::kCpu);
  SensorModel gpu(p, W, H, seed, groups, 1, 0, Precision::kFp32, Device::kCuda);
  // The frame means of the held inputs at every call's end: on kCuda they come from the backend's
  // end_frame (the held frame on the device, with the call's device runs or, for host runs, none).
  cpu.set_track_frame_means(true);
  gpu.set_track_frame_means(true);
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> N(3.0, 1.0);
  std::vector<float> irr0(P);
  for (float& v : irr0) v = static_cast<float>(std::exp(N(rng)));
  const int64_t t0 = 2 * (int64_t{1} << 23) - 33'333;
  cpu.init(t0, irr0.data());
  gpu.init(t0, irr0.data());
  int64_t t = t0;
  size_t events = 0;
  std::string why;
  for (int call = 0; call < 4; ++call) {
    const int64_t t1 = t + (int64_t{1} << 21) + 4321 * call;
    const Runs r = random_runs(rng, P, t, t1);
    Events a = cpu.advance_runs(r.ids.data(), r.ids.size(), r.t.data(), r.irr.data(), r.t.size(), r.start.data(),
                                r.count.data(), t1);
    Events b;
    if (call < 3) {
      b = gpu.advance_runs(r.ids.data(), r.ids.size(), r.t.data(), r.irr.data(), r.t.size(), r.start.data(),
                           r.count.data(), t1);
    } else {  // the last call on device-resident runs (advance_device, download_events)
      DeviceRuns d;
      d.pixel = to_device(r.ids);
      d.row_start = to_device(r.start);
      d.row_count = to_device(r.count);
      d.n = r.ids.size();
      d.t = to_device(r.t);
      d.irr = to_device(r.irr);
      d.n_samples = r.t.size();
      const DeviceEvents ev = gpu.advance_device(d, t1, nullptr);
      b = gpu.download_events(ev);
      (void)hipFree(const_cast<uint32_t*>(d.pixel));
      (void)hipFree(const_cast<int64_t*>(d.row_start));
      (void)hipFree(const_cast<int32_t*>(d.row_count));
      (void)hipFree(const_cast<int64_t*>(d.t));
      (void)hipFree(const_cast<float*>(d.irr));
    }
    if (!same(a, b, why)) {
      std::printf("case %s call %d: events differ: %s\n", c.name, call, why.c_str());
      return 1;
    }
    if (cpu.glare_frame_mean() != gpu.glare_frame_mean() || gpu.glare_frame_mean().size() != groups.size()) {
      std::printf("case %s call %d: frame means of the held inputs differ\n", c.name, call);
      return 1;
    }
    events += a.size();
    t = t1;
  }
  if (cpu.save_state() != gpu.save_state()) {
    std::printf("case %s: saved states differ\n", c.name);
    return 1;
  }
  if (!same_stats(cpu.stats(), gpu.stats(), p.photoreceptor == Photoreceptor::kInstantaneous, why)) {
    std::printf("case %s: %s\n", c.name, why.c_str());
    return 1;
  }
  // a state saved on the GPU model continues on a fresh CPU model exactly as on the GPU
  SensorModel cpu2(p, W, H, seed, groups, 1, 2, Precision::kFp32, Device::kCpu);
  const std::vector<uint8_t> blob = gpu.save_state();
  cpu2.load_state(blob.data(), blob.size());
  const int64_t t1 = t + (int64_t{1} << 20);
  const Runs r = random_runs(rng, P, t, t1);
  const Events a = cpu2.advance_runs(r.ids.data(), r.ids.size(), r.t.data(), r.irr.data(), r.t.size(),
                                     r.start.data(), r.count.data(), t1);
  const Events b = gpu.advance_runs(r.ids.data(), r.ids.size(), r.t.data(), r.irr.data(), r.t.size(), r.start.data(),
                                    r.count.data(), t1);
  if (!same(a, b, why) || cpu2.save_state() != gpu.save_state()) {
    std::printf("case %s: resumed run differs: %s\n", c.name, why.c_str());
    return 1;
  }
  events += a.size();
  std::printf("case %s: %zu events, cells %llu, refinements %llu: identical\n", c.name, events,
              static_cast<unsigned long long>(gpu.stats().cells),
              static_cast<unsigned long long>(gpu.stats().refinements));
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "run";
  try {
    if (mode == "build") {
      try {
        SensorModel m(preset("prophesee_gen3"), 8, 8, 1, std::vector<Waveform>(1), 0, 0, Precision::kFp32,
                      Device::kCuda);
        std::printf("constructed on the device\n");
      } catch (const std::runtime_error& e) {
        std::printf("no device: %s\n", e.what());
      }
      return 0;
    }
    int bad = 0;
    for (const Case& c : kCases) bad |= run_case(c);
    return bad;
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 2;
  }
}
