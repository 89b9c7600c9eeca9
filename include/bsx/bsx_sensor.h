/* bsx: the per-pixel continuous-time event-sensor path on MI355X (gfx950), as a C ABI.
 *
 * The entry points replace the reference's sensor backend seam: the C++ class
 * balla::sensor::SensorModel (src/sensor/sensor_model.h:98-320) with its CUDA pimpl
 * CudaSensorKernel (src/sensor/sensor_cuda.h:20-68), and the Python module _sensor built on it
 * (src/sensor/sensor_module.cpp:446-584). Each declaration names the reference member it stands in
 * for. Plain pointers and sizes only: no C++ or torch types cross this boundary.
 *
 * Conventions.
 *  - Every function returns a status: BSX_OK (0), or BSX_E_INVALID (the reference throws
 *    std::invalid_argument: bad input, refused before any state change), BSX_E_LOGIC
 *    (std::logic_error: misuse such as advancing an uninitialised model), BSX_E_RUNTIME
 *    (std::runtime_error: no device, a HIP error). bsx_last_error() returns the message of the last
 *    failure on the calling thread.
 *  - A call that fails after its input checks leaves the model uninitialised until init or
 *    load_state (reference: src/sensor/sensor_model.h:177-180).
 *  - Times are int64 ns; irradiance is fp32 in renderer units; one model is used by one thread at a
 *    time; device work is queued on the caller's stream (a hipStream_t passed as void*, 0 = the
 *    null stream).
 */
#ifndef BSX_SENSOR_H_
#define BSX_SENSOR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSX_OK 0
#define BSX_E_INVALID 1
#define BSX_E_LOGIC 2
#define BSX_E_RUNTIME 3

#define BSX_MAX_LIGHT_GROUPS 8
#define BSX_MAX_HARMONICS 8

enum { BSX_PHOTORECEPTOR_CASCADE = 0, BSX_PHOTORECEPTOR_INSTANTANEOUS = 1 };
enum { BSX_RESET_PHYSICAL = 0, BSX_RESET_IDEAL = 1 };
enum { BSX_PHASE_UNIFORM = 0, BSX_PHASE_SYMMETRIC = 1, BSX_PHASE_ZERO = 2 };
enum { BSX_WAVE_CONSTANT = 0, BSX_WAVE_FOURIER = 1, BSX_WAVE_PULSE = 2 };

/* Sensor parameters: the fields of the reference's SensorParams (src/sensor/params.h:65-172), SI
 * units, plus shot_rate_hz (this library's rate-specified shot-noise process; 0 = off, as every
 * reference family). Enum-valued fields hold the BSX_* values above. */
typedef struct bsx_params {
  char family[32];
  int32_t photoreceptor;
  int32_t noise;
  double k, tau_o, tau_sf, sigma_o;
  double illuminance_scale, veiling_glare;
  double dark_current_fA, dark_sigma_ln, dark_ea_eV, temperature_K, reference_temperature_K;
  double theta_on, theta_off, sigma_theta_on, sigma_theta_off, theta_onoff_corr, theta_floor;
  int32_t reset_mode;
  int32_t initial_phase;
  double refractory_s, eps_ci;
  double leak_rate_hz, leak_sigma_log10, leak_ea_eV;
  double tau_hp, tau_hp_sigma_ln;
  double hot_fraction;
  int32_t hot_rate_law;
  int32_t limiter;
  double hot_rate_min_hz, hot_rate_max_hz, hot_rate_median_hz, hot_rate_sigma_ln, hot_leak_fraction,
      hot_reference_current_fA;
  double readout_quantum_s, readout_delay_s;
  int64_t clock_offset_ns;
  double clock_drift_ppm;
  double eps_L;
  int32_t noise_leaf_level;
  int32_t noise_diffusion;
  double noise_margin_z;
  int32_t initial_state;
  int32_t reserved0;
  double label_share_cut;
  double shot_rate_hz;
} bsx_params;

/* A light group's waveform m_g(t) (reference: struct Waveform, src/sensor/waveform.h:38-52). */
typedef struct bsx_waveform {
  int32_t type;
  int32_t n_harmonics;
  double f0_hz;
  double a[BSX_MAX_HARMONICS];
  double phi[BSX_MAX_HARMONICS];
  double period_ns, phase_ns, rise_ns, top_ns, floor_level, peak_level;
} bsx_waveform;

/* One event as the kernel emits it: the reference's RawEvent (src/sensor/pixel.h:123-135), 32 B. */
typedef struct bsx_raw_event {
  uint32_t pixel;
  uint8_t polarity; /* 1 = ON */
  uint8_t label;    /* 0 signal, 1 noise, 2 leak, 3 hot */
  uint16_t flags;
  int64_t t_cross;
  float ds, dn, dl;
  float l_first;
} bsx_raw_event;

/* Sample runs in device memory (reference: DeviceRuns, src/sensor/sensor_model.h:81-89): listed
 * pixel pixel[i] takes the row_count[i] samples starting at sample row_start[i] of t and
 * irr (n_samples x groups). */
typedef struct bsx_device_runs {
  const uint32_t* pixel;
  const int64_t* row_start;
  const int32_t* row_count;
  size_t n;
  const int64_t* t;
  const float* irr;
  size_t n_samples;
} bsx_device_runs;

/* Raw events left in device memory by advance_device, unordered, valid until the next advance
 * (reference: DeviceEvents, src/sensor/sensor_model.h:93-96). */
typedef struct bsx_device_events {
  const bsx_raw_event* data;
  size_t count;
} bsx_device_events;

/* One pixel's frozen maps as the reference hands them to its GPU backend: PixelMap
 * (src/sensor/maps.h:17-27), 48 B. hot: 0 none, 1 leak-type, 2 threshold-type. */
typedef struct bsx_pixel_map {
  double theta_on, theta_off, leak, dark_fA, phi0;
  uint8_t hot;
  uint8_t pad[3];
  float gain;
} bsx_pixel_map;

/* The fp32 kernel constants the reference hands to its GPU backend (KernelConstants<float>,
 * src/sensor/pixel.h:216-252), as plain fields: the cascade's constants (CascadeConstants,
 * src/sensor/substep.h:26-34), the photometry, the noise tables' view and the reset. The last four
 * flags name features of the reference this backend refuses (BSX_E_INVALID when set). */
typedef struct bsx_kernel_constants {
  float k, a2, a3, tau_o, tau_sf, eps_accept, eps_drop;
  float eps_pl_base, scale, irr_max;
  int32_t groups;
  int32_t instant, noise, levels;
  float margin_z;
  int32_t n_nodes, stride; /* the noise tables' shape (noise on) */
  float ln_a1_min, inv_dln;
  int32_t ideal_reset;
  int64_t refractory_ns;
  float eps_ci, label_cut;
  uint64_t seed;
  uint32_t camera;
  int32_t modulated, breakpoints, transient, hpf;
} bsx_kernel_constants;

/* Host view of the events of the last advance/download, sorted by (t_cross, pixel), as the
 * reference's Events (src/sensor/sensor_model.h:66-75). Pointers into model-owned memory, valid
 * until the next call on the model. */
typedef struct bsx_events {
  size_t count;
  const uint16_t* x;
  const uint16_t* y;
  const uint8_t* p;
  const int64_t* t_cross_ns;
  const int64_t* t_reported_ns;
  const uint8_t* label;
  const uint16_t* flags;
  const float* ds;
  const float* dn;
  const float* dl;
  const float* l_first;
} bsx_events;

/* Work counters (reference: KernelStats, src/sensor/pixel.h:137-169), plus launch timing. */
typedef struct bsx_stats {
  uint64_t substeps, substep_attempts, settled_substeps, cells, refinements, leaf_intervals, fallbacks,
      input_clamps, quarantines, pieces;
  uint64_t launches, overflow_reruns, events;
  double kernel_ms; /* summed device time of the advance kernels (HIP events) */
} bsx_stats;

/* Launch policy (reference: GpuLaunchPolicy, src/sensor/launch.h:16-27): the initial event
 * capacity (0 = automatic), which the model grows on overflow, and the walk's launch shape (0 in a
 * field = keep the current value; the autotuner of a rig sets them). */
typedef struct bsx_launch_policy {
  uint64_t event_capacity;
  int32_t block_size;   /* threads per workgroup, a multiple of 64 in [64, 256] (default 256) */
  int32_t walk_budget;  /* segments a walking lane does before its wave re-plans, [1, 1024] (default 8) */
  int32_t search_batch; /* waiting lanes that start a joint exact-search pass, [1, 64] (default 8) */
  int32_t queue_chunk;  /* 64-pixel groups per grab of a persistent work queue, [0, 64]; 0 = no queue;
                           -1 sets 0 (default 0) */
  int32_t head_lanes;     /* pixels a walk group takes at the head of the heavy-first pixel order, [1, 64]
                             (default 64) */
  int32_t head_per_mille; /* the head's share of the order, in thousandths of the pixels, [0, 1000]; -1 sets
                             0 (default 0: every group takes 64 pixels) */
} bsx_launch_policy;

typedef struct bsx_sensor_info_t {
  uint32_t width, height;
  int32_t groups;
  int32_t initialized;
  int64_t time_ns;
  int64_t readout_phase_ns;
  float irr_max;
  uint32_t camera;
  uint64_t seed;
} bsx_sensor_info_t;

typedef struct bsx_sensor bsx_sensor;

/* The message of the last failure on this thread (truncated to n - 1 bytes); BSX_OK. */
int bsx_last_error(char* buf, size_t n);

/* Visible HIP devices. Reference: sensor_cuda_available(), src/sensor/sensor_model.h:52. */
int bsx_device_count(int32_t* out);

/* Self-check of the device math on the current device (no reference counterpart). Over every
 * positive finite fp32 x (bit patterns 1 .. 0x7f7fffff): the number of x where the device log of
 * the next float is below the log of x (the exact searches of the instantaneous response rely on
 * none), and an order-independent hash of all log bit patterns, sum of splitmix64((u << 32) | bits)
 * mod 2^64, to compare with a host evaluation. */
int bsx_math_selfcheck(uint64_t* log_violations, uint64_t* log_hash);

/* Reference: preset(name), src/sensor/params.h:176; preset_names(), :177 (names joined by ','). */
int bsx_params_preset(const char* name, bsx_params* out);
int bsx_preset_names(char* buf, size_t n);
/* Reference: SensorParams::validate(), src/sensor/params.cpp:44-96. */
int bsx_params_validate(const bsx_params* p);
/* The frozen per-pixel maps of a camera, drawn on the host (no device needed). Reference:
 * draw_sensor_maps(p, seed, camera, n_pixels), src/sensor/maps.h:57 / maps.cpp:78-156. Any output
 * pointer may be null; each non-null one receives n_pixels values. */
int bsx_draw_maps(const bsx_params* p, uint64_t seed, uint32_t camera, uint32_t n_pixels, double* theta_on,
                  double* theta_off, double* leak, double* dark_fA, double* phi0);

/* The in-path noise's tables of a cascade parameter set, built on the host in long double (no
 * device needed). Reference: NoiseTables(NoiseModel{k q, 1/tau_o, 1/tau_sf, sigma_o,
 * noise_leaf_level}), src/sensor/noise_tables.h:86-108 / noise_tables.cpp:341-411 (frozen
 * diffusion). *n_nodes and *stride receive the shape; data64 / data32 (either may be null) receive
 * n_nodes * stride values when cap holds that many, else BSX_E_INVALID. */
int bsx_noise_tables(const bsx_params* p, double* data64, float* data32, size_t cap, int32_t* n_nodes,
                     int32_t* stride);

/* The reference's GPU backend seam (CudaSensorKernel, src/sensor/sensor_cuda.h:20-68), for its own
 * SensorModel to run on this backend: a model of n_pixels pixels built from the reference's fp32
 * kernel constants, its PixelMaps and its fp32 noise tables (n_table_values = n_nodes * stride, or 0
 * with the noise off). Reference: CudaSensorKernel(kc, maps, tables), sensor_cuda.h:22-23. */
int bsx_sensor_create_from_constants(const bsx_kernel_constants* kc, const bsx_pixel_map* maps, size_t n_pixels,
                                     const float* tables, size_t n_table_values, bsx_sensor** out);
/* The reference's 200-B PixelStates (src/sensor/pixel.h:105-121) in and out of the device's working
 * state. Reference: upload_states / download_states, sensor_cuda.h:30-31. */
int bsx_sensor_upload_states(bsx_sensor* s, const void* states, size_t n_states);
int bsx_sensor_download_states(bsx_sensor* s, void* states, size_t n_states);
/* The model time of the next call (the reference passes t_state to each advance:
 * sensor_cuda.h:50-51, 59-61), after upload_states. */
int bsx_sensor_set_time(bsx_sensor* s, int64_t t_state_ns);
/* New photometry of a model built from constants: kc.scale, kc.irr_max and the maps (their gains).
 * Reference: CudaSensorKernel::set_photometry(kc, maps), sensor_cuda.h:40. */
int bsx_sensor_set_kernel_photometry(bsx_sensor* s, float scale, float irr_max, const bsx_pixel_map* maps,
                                     size_t n_pixels);
/* The raw events of the last host advance or download, sorted by (t_cross, pixel): *out points into
 * model-owned memory, valid until the next call on the model. Reference: the RawEvents that
 * CudaSensorKernel::advance and download return (sensor_cuda.h:54, 59-61). */
int bsx_sensor_raw_events(bsx_sensor* s, const bsx_raw_event** out, size_t* count);

/* Reference: SensorModel::SensorModel(params, width, height, seed, light_groups, camera, ...,
 * Precision::kFp32, Device::kCuda), src/sensor/sensor_model.h:107-110 (num_threads, precision and
 * roi have no meaning here: one GPU, fp32, full frame). */
int bsx_sensor_create(const bsx_params* params, uint32_t width, uint32_t height, uint64_t seed,
                      const bsx_waveform* light_groups, int32_t n_groups, uint32_t camera, bsx_sensor** out);
/* Reference: SensorModel::~SensorModel. */
int bsx_sensor_destroy(bsx_sensor* s);

/* Reference: SensorModel::set_photometry(scale, gain), src/sensor/sensor_model.h:116 (gain: one
 * per pixel, or null for 1 everywhere). */
int bsx_sensor_set_photometry(bsx_sensor* s, double illuminance_scale, const float* gain, size_t n_gain);

/* Reference: SensorModel::init(t0_ns, irradiance0), src/sensor/sensor_model.h:121. irradiance0 is
 * a host array of (width * height) x groups. */
int bsx_sensor_init(bsx_sensor* s, int64_t t0_ns, const float* irradiance0);

/* Reference: SensorModel::advance(ids, n, t, irr, m, counts, t_end), src/sensor/sensor_model.h:150.
 * Host arrays; the events are then in bsx_sensor_events. */
int bsx_sensor_advance(bsx_sensor* s, const uint32_t* pixel_ids, size_t n, const int64_t* t, const float* irr,
                       size_t m, const int32_t* counts, int64_t t_end_ns);
/* Reference: SensorModel::advance_runs(...), src/sensor/sensor_model.h:157-158. Host arrays. */
int bsx_sensor_advance_runs(bsx_sensor* s, const uint32_t* pixel_ids, size_t n, const int64_t* t, const float* irr,
                            size_t n_samples, const int64_t* row_start, const int32_t* row_count, int64_t t_end_ns);
/* Reference: CudaSensorKernel::end_frame(runs, stream), src/sensor/sensor_cuda.h:47-49 (SensorModel's
 * held frame, sensor_model.cpp:748-760): the input each pixel would hold after a call on these device
 * runs, its run's last sample (raw) or else its held input; runs null or empty: the held inputs.
 * out: pixels x groups floats on the host, written after the work queued on `stream`. The model is not
 * changed. */
int bsx_sensor_held_frame(bsx_sensor* s, const bsx_device_runs* runs, void* stream, float* out);
/* Reference: SensorModel::advance_device(runs, t_end, stream), src/sensor/sensor_model.h:166. The
 * runs are in device memory; the events stay there (out). */
int bsx_sensor_advance_device(bsx_sensor* s, const bsx_device_runs* runs, int64_t t_end_ns, void* stream,
                              bsx_device_events* out);
/* Reference: SensorModel::download_events(events), src/sensor/sensor_model.h:167. Sorts by
 * (t_cross, pixel) and fills the host view of bsx_sensor_events. */
int bsx_sensor_download_events(bsx_sensor* s, const bsx_device_events* events);
/* advance_device that allocates nothing. It has no counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The call's scratch (run index, check words, counters, launch descriptors)
 * lives in the library's static device scratch of 16 MiB + 4 KiB per GPU (held until the process
 * ends); cameras of more than 2^22 pixels are refused (BSX_E_INVALID). The raw events (unordered, the
 * 32-byte bsx_raw_event layout) go to the caller's device buffer `events` of `events_cap` entries;
 * *n_events receives their count. A call that would make more than events_cap events fails with
 * BSX_E_INVALID and leaves the model's state unchanged. Results equal bsx_sensor_advance_device's.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU (from several
 * threads or streams at once) race. */
int bsx_sensor_advance_device_noalloc(bsx_sensor* s, const bsx_device_runs* runs, int64_t t_end_ns, void* stream,
                                      void* events, uint64_t events_cap, uint64_t* n_events);
/* The host events of the last advance / advance_runs / download_events. */
int bsx_sensor_events(bsx_sensor* s, bsx_events* out);

/* Reference: state_bytes / save_state / load_state, src/sensor/sensor_model.h:181-183. The blob is
 * the reference's layout: a 56-byte header ("BSNS") followed by one 200-byte PixelState per pixel. */
int bsx_sensor_state_bytes(bsx_sensor* s, size_t* out);
int bsx_sensor_save_state(bsx_sensor* s, uint8_t* buf, size_t n);
int bsx_sensor_load_state(bsx_sensor* s, const uint8_t* buf, size_t n);

/* Reference: stats() / reset_stats(), src/sensor/sensor_model.h:187-192. */
int bsx_sensor_stats(bsx_sensor* s, bsx_stats* out);
int bsx_sensor_reset_stats(bsx_sensor* s);
/* Reference: set_gpu_launch_policy(policy), src/sensor/sensor_model.h:194. */
int bsx_sensor_set_launch_policy(bsx_sensor* s, const bsx_launch_policy* policy);
/* width, height, groups, initialized, time_ns, readout_phase_ns, irr_max (reference accessors,
 * src/sensor/sensor_model.h:195-219). */
int bsx_sensor_info(bsx_sensor* s, bsx_sensor_info_t* out);
/* The frozen maps (reference: SensorModel::maps(), src/sensor/sensor_model.h:186), one value per
 * pixel in pixel order; any pointer may be null. */
int bsx_sensor_maps(bsx_sensor* s, double* theta_on, double* theta_off, double* leak, double* dark_fA, double* phi0,
                    uint8_t* hot);

#ifdef __cplusplus
}
#endif

#endif /* BSX_SENSOR_H_ */
