// The C ABI of include/bsx/bsx_sensor.h over bsx::SensorModel: exceptions become status codes with
// a per-thread message (std::invalid_argument -> BSX_E_INVALID, std::logic_error -> BSX_E_LOGIC,
// anything else -> BSX_E_RUNTIME), as the reference's Python layer maps them to ValueError,
// RuntimeError and so on (src/sensor/sensor_module.cpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bsx/bsx_mlp.h"
#include "../../include/bsx/bsx_output.h"
#include "../../include/bsx/bsx_rig.h"
#include "../../include/bsx/bsx_sensor.h"
#include "mlp/proposer.h"
#include "output/output.h"
#include "output/stream.h"
#include "sensor/kernels.h"
#include "sensor/model.h"
#include "sensor/noise_tables.h"
#include "sensor/params.h"
#include "sensor/rig.h"

#include "handles.h"

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return BSX_OK;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return BSX_E_INVALID;
  } catch (const std::logic_error& e) {
    g_last_error = e.what();
    return BSX_E_LOGIC;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return BSX_E_RUNTIME;
  } catch (...) {
    g_last_error = "bsx: unknown error";
    return BSX_E_RUNTIME;
  }
}

bsx::SensorModel& model_of(bsx_sensor* s) {
  if (s == nullptr || s->model == nullptr) throw std::invalid_argument("bsx: null sensor handle");
  return *s->model;
}

bsx::Rig& rig_of(bsx_rig* r) {
  if (r == nullptr || r->rig == nullptr) throw std::invalid_argument("bsx: null rig handle");
  return *r->rig;
}

bsx::Proposer& mlp_of(bsx_mlp* m) {
  if (m == nullptr || m->p == nullptr) throw std::invalid_argument("bsx: null proposer handle");
  return *m->p;
}

bsx_launch_policy policy_of(const bsx::WalkTuning& t) {
  bsx_launch_policy p{};
  p.block_size = t.block;
  p.walk_budget = t.walk_budget;
  p.search_batch = t.search_batch;
  p.queue_chunk = t.chunk;
  p.head_lanes = t.head_lanes;
  p.head_per_mille = t.head_per_mille;
  return p;
}

void need_ptr(const void* p, const char* what) {
  if (p == nullptr) throw std::invalid_argument(std::string("bsx: null ") + what);
}

}  // namespace

extern "C" {

// The other C-ABI translation units (scene_host.cpp) report through the same per-thread message.
int bsx_internal_set_error(int code, const char* msg) {
  g_last_error = msg;
  return code;
}

int bsx_last_error(char* buf, size_t n) {
  if (buf != nullptr && n > 0) {
    const size_t k = std::min(n - 1, g_last_error.size());
    std::memcpy(buf, g_last_error.data(), k);
    buf[k] = '\0';
  }
  return BSX_OK;
}

int bsx_device_count(int32_t* out) {
  return guarded([&] {
    need_ptr(out, "out");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *out = n;
  });
}

int bsx_math_selfcheck(uint64_t* log_violations, uint64_t* log_hash) {
  return guarded([&] {
    need_ptr(log_violations, "log_violations");
    need_ptr(log_hash, "log_hash");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) throw std::runtime_error("bsx_math_selfcheck: no HIP device");
    unsigned long long* d = nullptr;
    bsx::hip_check(hipMalloc(&d, 16), "bsx_math_selfcheck");
    unsigned long long h[2] = {0, 0};
    hipError_t e = bsx::launch_math_selfcheck(d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(h, d, 16, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    bsx::hip_check(e, "bsx_math_selfcheck");
    *log_violations = h[0];
    *log_hash = h[1];
  });
}

int bsx_params_preset(const char* name, bsx_params* out) {
  return guarded([&] {
    need_ptr(name, "name");
    need_ptr(out, "out");
    *out = bsx::preset(name);
  });
}

int bsx_preset_names(char* buf, size_t n) {
  return guarded([&] {
    std::string all;
    for (const std::string& s : bsx::preset_names()) all += (all.empty() ? "" : ",") + s;
    if (buf != nullptr && n > 0) {
      const size_t k = std::min(n - 1, all.size());
      std::memcpy(buf, all.data(), k);
      buf[k] = '\0';
    }
  });
}

int bsx_params_validate(const bsx_params* p) {
  return guarded([&] {
    need_ptr(p, "params");
    bsx::validate(*p);
  });
}

int bsx_draw_maps(const bsx_params* p, uint64_t seed, uint32_t camera, uint32_t n_pixels, double* theta_on,
                  double* theta_off, double* leak, double* dark_fA, double* phi0) {
  return guarded([&] {
    need_ptr(p, "params");
    bsx::validate_ranges(*p);
    const bsx::SensorMaps m = bsx::draw_sensor_maps(*p, seed, camera, n_pixels);
    const size_t n = m.size();
    if (theta_on != nullptr) std::memcpy(theta_on, m.theta_on.data(), n * 8);
    if (theta_off != nullptr) std::memcpy(theta_off, m.theta_off.data(), n * 8);
    if (leak != nullptr) std::memcpy(leak, m.leak.data(), n * 8);
    if (dark_fA != nullptr) std::memcpy(dark_fA, m.dark_fA.data(), n * 8);
    if (phi0 != nullptr) std::memcpy(phi0, m.phi0.data(), n * 8);
  });
}

int bsx_noise_tables(const bsx_params* p, double* data64, float* data32, size_t cap, int32_t* n_nodes,
                     int32_t* stride) {
  return guarded([&] {
    need_ptr(p, "params");
    need_ptr(n_nodes, "n_nodes");
    need_ptr(stride, "stride");
    bsx::validate_ranges(*p);
    bsx::NoiseModelParams nm{};
    nm.kq = p->k * bsx::kElementaryCharge;
    nm.a2 = 1.0 / p->tau_o;
    nm.a3 = 1.0 / p->tau_sf;
    nm.sigma_o = p->sigma_o;
    nm.levels = p->noise_leaf_level;
    const bsx::NoiseTableSet t = bsx::build_noise_tables(nm);
    *n_nodes = t.n_nodes;
    *stride = t.stride;
    const size_t need = t.data64.size();
    if ((data64 != nullptr || data32 != nullptr) && cap < need) {
      throw std::invalid_argument("bsx_noise_tables: cap is smaller than n_nodes * stride");
    }
    if (data64 != nullptr) std::memcpy(data64, t.data64.data(), need * sizeof(double));
    if (data32 != nullptr) std::memcpy(data32, t.data32.data(), need * sizeof(float));
  });
}

int bsx_sensor_create(const bsx_params* params, uint32_t width, uint32_t height, uint64_t seed,
                      const bsx_waveform* light_groups, int32_t n_groups, uint32_t camera, bsx_sensor** out) {
  return guarded([&] {
    need_ptr(params, "params");
    need_ptr(out, "out");
    *out = nullptr;
    if (n_groups < 0 || (n_groups > 0 && light_groups == nullptr)) {
      throw std::invalid_argument("bsx_sensor_create: light_groups must hold n_groups waveforms");
    }
    std::vector<bsx_waveform> groups(light_groups, light_groups + n_groups);
    auto* m = new bsx::SensorModel(*params, width, height, seed, std::move(groups), camera);
    *out = new bsx_sensor{m};
  });
}

int bsx_sensor_create_from_constants(const bsx_kernel_constants* kc, const bsx_pixel_map* maps, size_t n_pixels,
                                     const float* tables, size_t n_table_values, bsx_sensor** out) {
  static_assert(sizeof(bsx_pixel_map) == 48, "bsx_pixel_map is the reference's 48-byte PixelMap");
  return guarded([&] {
    need_ptr(kc, "kc");
    need_ptr(out, "out");
    *out = nullptr;
    auto* m = new bsx::SensorModel(*kc, maps, n_pixels, tables, n_table_values);
    *out = new bsx_sensor{m};
  });
}

int bsx_sensor_upload_states(bsx_sensor* s, const void* states, size_t n_states) {
  static_assert(sizeof(bsx::SavedPixelState) == 200, "the reference's 200-byte PixelState");
  return guarded([&] {
    need_ptr(states, "states");
    model_of(s).upload_states(static_cast<const bsx::SavedPixelState*>(states), n_states);
  });
}

int bsx_sensor_download_states(bsx_sensor* s, void* states, size_t n_states) {
  return guarded([&] {
    need_ptr(states, "states");
    model_of(s).download_states(static_cast<bsx::SavedPixelState*>(states), n_states);
  });
}

int bsx_sensor_set_time(bsx_sensor* s, int64_t t_state_ns) {
  return guarded([&] { model_of(s).set_time(t_state_ns); });
}

int bsx_sensor_set_kernel_photometry(bsx_sensor* s, float scale, float irr_max, const bsx_pixel_map* maps,
                                     size_t n_pixels) {
  return guarded([&] { model_of(s).set_kernel_photometry(scale, irr_max, maps, n_pixels); });
}

int bsx_sensor_raw_events(bsx_sensor* s, const bsx_raw_event** out, size_t* count) {
  static_assert(sizeof(bsx_raw_event) == sizeof(bsx::RawEventDev), "raw event layout");
  return guarded([&] {
    need_ptr(out, "out");
    need_ptr(count, "count");
    const std::vector<bsx::RawEventDev>& r = model_of(s).raw_events();
    *out = reinterpret_cast<const bsx_raw_event*>(r.data());
    *count = r.size();
  });
}

int bsx_sensor_destroy(bsx_sensor* s) {
  return guarded([&] {
    if (s == nullptr) return;
    delete s->model;
    delete s;
  });
}

int bsx_sensor_set_photometry(bsx_sensor* s, double illuminance_scale, const float* gain, size_t n_gain) {
  return guarded([&] { model_of(s).set_photometry(illuminance_scale, gain, n_gain); });
}

int bsx_sensor_init(bsx_sensor* s, int64_t t0_ns, const float* irradiance0) {
  return guarded([&] {
    need_ptr(irradiance0, "irradiance0");
    model_of(s).init(t0_ns, irradiance0);
  });
}

int bsx_sensor_advance(bsx_sensor* s, const uint32_t* pixel_ids, size_t n, const int64_t* t, const float* irr,
                       size_t m, const int32_t* counts, int64_t t_end_ns) {
  return guarded([&] {
    if (n > 0) need_ptr(pixel_ids, "pixel_ids");
    if (n * m > 0) {
      need_ptr(t, "t");
      need_ptr(irr, "irr");
    }
    model_of(s).advance(pixel_ids, n, t, irr, m, counts, t_end_ns);
  });
}

int bsx_sensor_advance_runs(bsx_sensor* s, const uint32_t* pixel_ids, size_t n, const int64_t* t, const float* irr,
                            size_t n_samples, const int64_t* row_start, const int32_t* row_count, int64_t t_end_ns) {
  return guarded([&] {
    if (n > 0) {
      need_ptr(pixel_ids, "pixel_ids");
      need_ptr(row_start, "row_start");
      need_ptr(row_count, "row_count");
    }
    if (n_samples > 0) {
      need_ptr(t, "t");
      need_ptr(irr, "irr");
    }
    model_of(s).advance_runs(pixel_ids, n, t, irr, n_samples, row_start, row_count, t_end_ns);
  });
}

int bsx_sensor_held_frame(bsx_sensor* s, const bsx_device_runs* runs, void* stream, float* out) {
  return guarded([&] { model_of(s).held_frame(runs, static_cast<hipStream_t>(stream), out); });
}

int bsx_sensor_advance_device(bsx_sensor* s, const bsx_device_runs* runs, int64_t t_end_ns, void* stream,
                              bsx_device_events* out) {
  return guarded([&] {
    need_ptr(runs, "runs");
    need_ptr(out, "out");
    *out = model_of(s).advance_device(*runs, t_end_ns, static_cast<hipStream_t>(stream));
  });
}

int bsx_sensor_advance_device_noalloc(bsx_sensor* s, const bsx_device_runs* runs, int64_t t_end_ns, void* stream,
                                      void* events, uint64_t events_cap, uint64_t* n_events) {
  return guarded([&] {
    need_ptr(runs, "runs");
    need_ptr(n_events, "n_events");
    *n_events = model_of(s).advance_device_noalloc(*runs, t_end_ns, static_cast<hipStream_t>(stream),
                                                   static_cast<bsx::RawEventDev*>(events), events_cap);
  });
}

int bsx_sensor_download_events(bsx_sensor* s, const bsx_device_events* events) {
  return guarded([&] {
    need_ptr(events, "events");
    model_of(s).download_events(*events);
  });
}

int bsx_sensor_events(bsx_sensor* s, bsx_events* out) {
  return guarded([&] {
    need_ptr(out, "out");
    const bsx::HostEvents& e = model_of(s).events();
    out->count = e.size();
    out->x = e.x.data();
    out->y = e.y.data();
    out->p = e.p.data();
    out->t_cross_ns = e.t_cross.data();
    out->t_reported_ns = e.t_reported.data();
    out->label = e.label.data();
    out->flags = e.flags.data();
    out->ds = e.ds.data();
    out->dn = e.dn.data();
    out->dl = e.dl.data();
    out->l_first = e.l_first.data();
  });
}

int bsx_sensor_state_bytes(bsx_sensor* s, size_t* out) {
  return guarded([&] {
    need_ptr(out, "out");
    *out = model_of(s).state_bytes();
  });
}

int bsx_sensor_save_state(bsx_sensor* s, uint8_t* buf, size_t n) {
  return guarded([&] {
    need_ptr(buf, "buf");
    model_of(s).save_state(buf, n);
  });
}

int bsx_sensor_load_state(bsx_sensor* s, const uint8_t* buf, size_t n) {
  return guarded([&] {
    need_ptr(buf, "buf");
    model_of(s).load_state(buf, n);
  });
}

int bsx_sensor_stats(bsx_sensor* s, bsx_stats* out) {
  return guarded([&] {
    need_ptr(out, "out");
    *out = model_of(s).stats();
  });
}

int bsx_sensor_reset_stats(bsx_sensor* s) {
  return guarded([&] { model_of(s).reset_stats(); });
}

int bsx_sensor_set_launch_policy(bsx_sensor* s, const bsx_launch_policy* policy) {
  return guarded([&] {
    need_ptr(policy, "policy");
    model_of(s).set_policy(*policy);
  });
}

int bsx_sensor_info(bsx_sensor* s, bsx_sensor_info_t* out) {
  return guarded([&] {
    need_ptr(out, "out");
    *out = model_of(s).info();
  });
}

int bsx_sensor_maps(bsx_sensor* s, double* theta_on, double* theta_off, double* leak, double* dark_fA, double* phi0,
                    uint8_t* hot) {
  return guarded([&] {
    const bsx::SensorMaps& m = model_of(s).maps();
    const size_t n = m.size();
    if (theta_on != nullptr) std::memcpy(theta_on, m.theta_on.data(), n * 8);
    if (theta_off != nullptr) std::memcpy(theta_off, m.theta_off.data(), n * 8);
    if (leak != nullptr) std::memcpy(leak, m.leak.data(), n * 8);
    if (dark_fA != nullptr) std::memcpy(dark_fA, m.dark_fA.data(), n * 8);
    if (phi0 != nullptr) std::memcpy(phi0, m.phi0.data(), n * 8);
    if (hot != nullptr) std::memcpy(hot, m.hot.data(), n);
  });
}

int bsx_rig_create(bsx_sensor* const* cameras, int32_t n_cameras, bsx_rig** out) {
  return guarded([&] {
    need_ptr(out, "out");
    *out = nullptr;
    if (n_cameras < 1 || n_cameras > 64) throw std::invalid_argument("bsx_rig_create: need 1 to 64 cameras");
    need_ptr(cameras, "cameras");
    std::vector<bsx::SensorModel*> ms;
    for (int32_t c = 0; c < n_cameras; ++c) ms.push_back(&model_of(cameras[c]));
    auto* r = new bsx::Rig(std::move(ms));
    *out = new bsx_rig{r};
  });
}

int bsx_rig_destroy(bsx_rig* rig) {
  return guarded([&] {
    if (rig == nullptr) return;
    delete rig->rig;
    delete rig;
  });
}

int bsx_rig_advance_device(bsx_rig* rig, const bsx_device_runs* runs, int64_t t_end_ns, void* stream,
                           bsx_device_events* out) {
  return guarded([&] {
    need_ptr(runs, "runs");
    need_ptr(out, "out");
    rig_of(rig).advance_device(runs, t_end_ns, static_cast<hipStream_t>(stream), out);
  });
}

int bsx_rig_set_tuning(bsx_rig* rig, const bsx_launch_policy* tuning) {
  return guarded([&] {
    need_ptr(tuning, "tuning");
    bsx::LaunchChain& ch = rig_of(rig).chain();
    bsx::WalkTuning t = ch.tuning();
    if (tuning->block_size != 0) t.block = tuning->block_size;
    if (tuning->walk_budget != 0) t.walk_budget = tuning->walk_budget;
    if (tuning->search_batch != 0) t.search_batch = tuning->search_batch;
    if (tuning->queue_chunk != 0) t.chunk = tuning->queue_chunk < 0 ? 0 : tuning->queue_chunk;
    if (tuning->head_lanes != 0) t.head_lanes = tuning->head_lanes;
    if (tuning->head_per_mille != 0) t.head_per_mille = tuning->head_per_mille < 0 ? 0 : tuning->head_per_mille;
    ch.set_tuning(t);
  });
}

int bsx_rig_get_tuning(bsx_rig* rig, bsx_launch_policy* out) {
  return guarded([&] {
    need_ptr(out, "out");
    *out = policy_of(rig_of(rig).chain().tuning());
  });
}

int bsx_rig_set_graph(bsx_rig* rig, int32_t on) {
  return guarded([&] { rig_of(rig).chain().set_graph(on != 0); });
}

int bsx_rig_stats(bsx_rig* rig, bsx_rig_stats_t* out) {
  return guarded([&] {
    need_ptr(out, "out");
    bsx::Rig& r = rig_of(rig);
    const bsx::ChainStats& s = r.chain().stats();
    *out = bsx_rig_stats_t{s.launches, s.graph_launches, s.captures, s.walk_reruns, r.kernel_ms()};
  });
}

int bsx_rig_autotune(bsx_rig* rig, const bsx_device_runs* runs, int64_t t_end_ns, void* stream, int32_t reps,
                     bsx_launch_policy* tunings, double* walk_ms, int32_t max_trials, int32_t* n_trials) {
  return guarded([&] {
    need_ptr(runs, "runs");
    if (max_trials < 0) throw std::invalid_argument("bsx_rig_autotune: max_trials must be >= 0");
    const std::vector<bsx::RigTrial> tr = rig_of(rig).autotune(runs, t_end_ns, static_cast<hipStream_t>(stream), reps);
    const int32_t k = std::min<int32_t>(max_trials, static_cast<int32_t>(tr.size()));
    for (int32_t i = 0; i < k; ++i) {
      if (tunings != nullptr) tunings[i] = policy_of(tr[i].tuning);
      if (walk_ms != nullptr) walk_ms[i] = tr[i].walk_ms;
    }
    if (n_trials != nullptr) *n_trials = static_cast<int32_t>(tr.size());
  });
}

int bsx_mlp_create(uint64_t seed, bsx_mlp** out) {
  return guarded([&] {
    need_ptr(out, "out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw std::runtime_error("bsx_mlp_create: no HIP device");
    auto* p = new bsx::Proposer(seed);
    *out = new bsx_mlp{p};
  });
}

int bsx_mlp_destroy(bsx_mlp* m) {
  return guarded([&] {
    if (m == nullptr) return;
    delete m->p;
    delete m;
  });
}

int bsx_mlp_set_params(bsx_mlp* m, const float* params) {
  return guarded([&] {
    need_ptr(params, "params");
    for (int i = 0; i < BSX_MLP_PARAMS; ++i)
      if (!std::isfinite(params[i])) throw std::invalid_argument("bsx_mlp_set_params: non-finite parameter");
    mlp_of(m).set_params(params);
  });
}

int bsx_mlp_get_params(bsx_mlp* m, float* params) {
  return guarded([&] {
    need_ptr(params, "params");
    mlp_of(m).get_params(params);
  });
}

int bsx_mlp_gradients(bsx_mlp* m, const float* x, const float* target, int64_t n, void* stream, float* grad,
                      double* loss) {
  return guarded([&] {
    need_ptr(x, "x");
    need_ptr(target, "target");
    need_ptr(grad, "grad");
    const double l = mlp_of(m).gradients(x, target, n, grad, static_cast<hipStream_t>(stream));
    if (loss != nullptr) *loss = l;
  });
}

int bsx_mlp_train(bsx_mlp* m, const float* x, const float* target, int64_t n, int32_t batch, int32_t epochs, float lr,
                  void* stream, double* loss) {
  return guarded([&] {
    need_ptr(x, "x");
    need_ptr(target, "target");
    const double l = mlp_of(m).train(x, target, n, batch, epochs, lr, static_cast<hipStream_t>(stream));
    if (loss != nullptr) *loss = l;
  });
}

int bsx_mlp_predict(bsx_mlp* m, const float* x, int64_t n, float* y, void* stream) {
  return guarded([&] {
    if (n < 0) throw std::invalid_argument("bsx_mlp_predict: n must be >= 0");
    if (n == 0) return;
    need_ptr(x, "x");
    need_ptr(y, "y");
    mlp_of(m).predict(x, n, y, static_cast<hipStream_t>(stream));
  });
}

int bsx_mlp_tile_check(const uint16_t* a, const uint16_t* b, float* c) {
  return guarded([&] {
    need_ptr(a, "a");
    need_ptr(b, "b");
    need_ptr(c, "c");
    bsx::DeviceBuffer da, db, dc;
    da.ensure(512);
    db.ensure(512);
    dc.ensure(4096);
    bsx::hip_check(hipMemcpy(da.get(), a, 512, hipMemcpyHostToDevice), "upload");
    bsx::hip_check(hipMemcpy(db.get(), b, 512, hipMemcpyHostToDevice), "upload");
    bsx::hip_check(bsx::mlp::launch_tile_check(da.as<const uint16_t>(), db.as<const uint16_t>(), dc.as<float>(),
                                               nullptr),
                   "tile check");
    bsx::hip_check(hipMemcpy(c, dc.get(), 4096, hipMemcpyDeviceToHost), "download");
  });
}

namespace {

bsx::OutputStage& output_of(bsx_output* o) {
  if (o == nullptr || o->o == nullptr) throw std::invalid_argument("bsx: null output handle");
  return *o->o;
}

static_assert(sizeof(bsx_stream_event) == sizeof(bsx::StreamEventDev), "stream record layout");
static_assert(offsetof(bsx_stream_event, x) == offsetof(bsx::StreamEventDev, x), "stream record layout");
static_assert(offsetof(bsx_stream_event, flags) == offsetof(bsx::StreamEventDev, flags), "stream record layout");

bsx::EventStream& stream_of(bsx_event_stream* s) {
  if (s == nullptr || s->s == nullptr) throw std::invalid_argument("bsx: null event stream handle");
  return *s->s;
}

void fill_output(const bsx::OutputColumns& c, size_t held, bsx_output_events* out) {
  *out = bsx_output_events{c.x.size(),          c.x.data(),         c.y.data(),       c.p.data(),
                           c.t_us.data(),       c.t_ns_rem.data(),  c.t_reported.data(), c.t_cross.data(),
                           c.label.data(),      c.flags.data(),     held};
}

}  // namespace

int bsx_output_create(bsx_sensor* camera, const bsx_crop* crop, int64_t t_offset_us, bsx_output** out) {
  return guarded([&] {
    need_ptr(out, "out");
    *out = nullptr;
    bsx::SensorModel& m = model_of(camera);
    bsx::CropBox box{};
    if (crop != nullptr) box = bsx::CropBox{crop->x0, crop->y0, crop->x1, crop->y1};
    auto* o = new bsx::OutputStage(m.width(), m.height(), m.readout(), crop != nullptr ? &box : nullptr, t_offset_us);
    *out = new bsx_output{o};
  });
}

int bsx_output_destroy(bsx_output* o) {
  return guarded([&] {
    if (o == nullptr) return;
    delete o->o;
    delete o;
  });
}

int bsx_output_window(bsx_output* o, const bsx_device_events* events, int64_t t_w, void* stream,
                      bsx_output_events* out) {
  return guarded([&] {
    need_ptr(events, "events");
    need_ptr(out, "out");
    bsx::OutputStage& s = output_of(o);
    const bsx::OutputColumns& c = s.window(reinterpret_cast<const bsx::RawEventDev*>(events->data), events->count, t_w,
                                           static_cast<hipStream_t>(stream));
    fill_output(c, s.held(), out);
  });
}

int bsx_output_flush(bsx_output* o, void* stream, bsx_output_events* out) {
  return guarded([&] {
    need_ptr(out, "out");
    bsx::OutputStage& s = output_of(o);
    const bsx::OutputColumns& c = s.flush(static_cast<hipStream_t>(stream));
    fill_output(c, s.held(), out);
  });
}

int bsx_event_stream_create(size_t chunk_bytes, bsx_chunk_sink sink, void* user, bsx_event_stream** out) {
  return guarded([&] {
    need_ptr(out, "out");
    *out = nullptr;
    auto* s = new bsx::EventStream(chunk_bytes, reinterpret_cast<bsx::EventStream::Sink>(sink), user);
    *out = new bsx_event_stream{s};
  });
}

int bsx_event_stream_destroy(bsx_event_stream* s) {
  return guarded([&] {
    if (s == nullptr) return;
    delete s->s;
    delete s;
  });
}

int bsx_output_window_to_stream(bsx_output* o, const bsx_device_events* events, int64_t t_w, void* stream,
                                bsx_event_stream* s, size_t* released, size_t* held) {
  return guarded([&] {
    need_ptr(events, "events");
    bsx::OutputStage& st = output_of(o);
    bsx::EventStream& es = stream_of(s);
    const auto hs = static_cast<hipStream_t>(stream);
    const bsx::OutEventDev* ev = nullptr;
    const size_t m = st.window_device(reinterpret_cast<const bsx::RawEventDev*>(events->data), events->count, t_w,
                                      hs, &ev);
    es.push(ev, m, hs);
    if (released != nullptr) *released = m;
    if (held != nullptr) *held = st.held();
  });
}

int bsx_output_flush_to_stream(bsx_output* o, void* stream, bsx_event_stream* s, size_t* released) {
  return guarded([&] {
    bsx::OutputStage& st = output_of(o);
    bsx::EventStream& es = stream_of(s);
    const auto hs = static_cast<hipStream_t>(stream);
    const bsx::OutEventDev* ev = nullptr;
    const size_t m = st.flush_device(hs, &ev);
    es.push(ev, m, hs);
    if (released != nullptr) *released = m;
  });
}

int bsx_event_stream_finish(bsx_event_stream* s, void* stream) {
  return guarded([&] { stream_of(s).finish(static_cast<hipStream_t>(stream)); });
}

int bsx_event_stream_stats(const bsx_event_stream* s, bsx_stream_stats* out) {
  return guarded([&] {
    need_ptr(out, "out");
    const bsx::EventStreamStats st = stream_of(const_cast<bsx_event_stream*>(s)).stats();
    *out = bsx_stream_stats{st.chunks, st.events, st.bytes, st.wait_ms, st.copy_ms};
  });
}

}  // extern "C"
