/* bsx scene and sample-time controller: the caller side of the sensor path on the GPU.
 *
 * On the GPU the reference's renderer is replaced by an analytic procedural scene evaluated in the
 * kernels (textured rectangles and disks with affine motion over a uniform background), and its
 * window engine's plain error-tolerance controller is re-implemented on the device:
 *   - the per-window dyadic lattice of level l = clamp(ceil(log2(v_hat W / d_px)), l_min, l_max)
 *     (reference: src/engine/motion.cpp:1330-1340, plan_schedule src/engine/motion.h:301-309);
 *   - the round-0 flags (signature flip, eta_L midpoint test) and up to max_refine_levels rounds of
 *     midpoints and half flags, intervals shorter than min_split never split (reference:
 *     src/engine/sampling.h:46-58 midpoint_eta, :471-567 Round0Flags, :591-607 Midpoints,
 *     :612-733 HalfFlags);
 *   - the ragged run assembly, one run per pixel in time order (reference: sampling.h:764-839).
 * The runs are left in device memory in the bsx_device_runs form bsx_sensor_advance_device takes.
 * Conventions as include/bsx/bsx_sensor.h (status codes, bsx_last_error).
 */
#ifndef BSX_SCENE_H_
#define BSX_SCENE_H_

#include <stddef.h>
#include <stdint.h>

#include "bsx_sensor.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { BSX_SHAPE_RECT = 0, BSX_SHAPE_DISK = 1 };

/* One textured shape with affine motion: centre c(t) = (cx, cy) + (vx, vy) t, rotation
 * angle + omega t (t in seconds from 0); irradiance lux (1 + contrast (2 n - 1)) with n in [0, 1]
 * the seeded value noise of `octaves` octaves whose coarsest cell is tex_cell_px pixels. Shapes
 * later in the list are drawn over earlier ones. */
typedef struct bsx_shape {
  int32_t kind;
  uint32_t tex_seed;
  double cx, cy, vx, vy;
  double angle, omega;
  double hx, hy; /* rectangle half sizes; a disk's radius is hx */
  float lux, contrast;
  int32_t octaves;
  float tex_cell_px;
} bsx_shape;

/* A camera's view of the scene: camera pixel (x, y) samples scene point
 * (origin_x + (x + 0.5) px_scale, origin_y + (y + 0.5) px_scale). */
typedef struct bsx_view {
  uint32_t width, height;
  double origin_x, origin_y, px_scale;
} bsx_view;

/* The controller's rules (reference: SampleRules, src/engine/sampling.h:90-122, and the engine's
 * defaults, src/engine/engine.h:41-61). fixed_period_ns > 0 selects the fixed-rate grid (every
 * pixel at t_w + k period, no flags), the dense reference of the error measures. */
typedef struct bsx_sampler_rules {
  double d_px;             /* 0.25 */
  float refine_tol_log;    /* eta_L tolerance; <= 0: no midpoint test */
  float dark;              /* the dark term of eta_L in renderer units (dark current / scale) */
  int32_t l_min, l_max;    /* 1 and 13 (min_dt 1024 ns) */
  int32_t max_refine_levels; /* 4 */
  uint32_t min_split_ns;   /* 2048 = 2 min_dt */
  int32_t refine_signature; /* 1 */
  int32_t reserved;
  int64_t fixed_period_ns; /* 0: the adaptive controller */
} bsx_sampler_rules;

typedef struct bsx_scene bsx_scene;

int bsx_scene_create(const bsx_shape* shapes, int32_t n_shapes, float background, bsx_scene** out);
int bsx_scene_destroy(bsx_scene* s);
/* The scene's irradiance at t_ns for every pixel of the view, into a host array of width*height
 * floats (the frame init takes). */
int bsx_scene_render(bsx_scene* s, const bsx_view* view, int64_t t_ns, float* out);
/* The samples of window [t_w, t_w + 2^23 ns] for every pixel of the view (t_w a multiple of 2^23;
 * the sample at t_w itself is carried, not emitted): runs in pixel order (run i is pixel i, possibly
 * empty), left in device buffer set `slot` of the scene (valid until that slot is sampled again or
 * the scene is destroyed). Work is queued on `stream`; the call returns when the runs are complete.
 * *total_samples receives the window's sample count. */
int bsx_scene_sample_window(bsx_scene* s, const bsx_view* view, const bsx_sampler_rules* rules, int64_t t_w,
                            int32_t slot, void* stream, bsx_device_runs* out, uint64_t* total_samples);

/* The device memory bsx_scene_sample_window may hold for its one-pass form (default 16 GiB). With
 * k = min(1024, bytes / (8 width height)) >= 1 slots per pixel it renders each sample once: every pixel's
 * first k samples are captured as 8-byte (offset, value) slots while it is counted, then copied into place
 * once the counts are scanned; a pixel with more samples is sampled again in place. With bytes too small
 * for one slot per pixel it counts in one pass and writes in a second. The runs are the same either way.
 * No reference counterpart. When the device cannot hold the slots, the budget drops to 0. */
int bsx_scene_set_capture_budget(bsx_scene* s, uint64_t bytes);
/* The capture budget now in force, and the windows so far whose largest pixel count exceeded the slots. */
int bsx_scene_capture_budget(bsx_scene* s, uint64_t* bytes, uint64_t* overflow_windows);

/* Samples of window (t_w, t_w + 2^23 ns] at imposed per-pixel counts (a host array of width*height
 * values in [0, 2^23]): pixel p's n samples at t_w + floor((k + 1) 2^23 / n), k = 0 .. n - 1, with the
 * scene's irradiance at each; runs in pixel order in buffer set `slot`, as bsx_scene_sample_window.
 * No reference counterpart: the workload generator of ragged batches whose sample counts follow a
 * given distribution rather than the controller. */
int bsx_scene_sample_counts(bsx_scene* s, const bsx_view* view, const int32_t* counts, int64_t t_w, int32_t slot,
                            void* stream, bsx_device_runs* out, uint64_t* total_samples);

/* The engine's sample request (reference: balla::backend::SampleRequest, src/backend/api.h:25-33), the same
 * 24-byte layout, so the engine's request arrays are passed as they are. */
typedef struct bsx_sample_request {
  uint16_t cam;
  uint16_t flags;
  uint32_t pixel;     /* y * width + x */
  uint32_t k_begin;   /* sub-sample range (the analytic scene is evaluated at the pixel centre for all of them) */
  uint32_t k_count;
  int32_t t_off_ns;   /* time = t_base_ns + t_off_ns */
  uint32_t pose_slot;
} bsx_sample_request;

/* The engine's RenderHook::render on the device scene (reference: src/engine/window.h:66-80; the time base as
 * the engine's hook adapter forms it, src/engine/bindings/engine_module.cpp:236): request i writes the scene's
 * irradiance at its pixel centre and time into row out_index[i] of irr (groups floats per row: the scene's
 * one light in group 0, 0 in the others), its signature (1 + shape index, 0 = background) into
 * sig[out_index[i]] and, when klass is not null, the class word (1 << 13, textured, on a shape; 0 on the
 * background) into klass[out_index[i]]. out_index null: row i. All arrays in device memory; work queued on
 * `stream`. */
int bsx_scene_render_requests(bsx_scene* s, const bsx_view* view, int64_t t_base_ns, const bsx_sample_request* requests,
                              size_t n, const uint32_t* out_index, int32_t groups, float* irr, uint64_t* sig,
                              uint32_t* klass, void* stream);

/* Copies `bytes` from device memory to host memory (synchronous): a caller without its own HIP
 * bindings reads runs or raw events with it. */
int bsx_copy_to_host(void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* BSX_SCENE_H_ */
