/* bsx learned sample-interval proposer: a 12-64-64-1 SiLU perceptron trained and run on the GPU by
 * this library's own kernels (BF16 weights on the MFMA units, FP32 accumulation, FP32 master weights
 * and Adam state).
 *
 * No reference counterpart: the reference's sampler has only the plain error-tolerance controller
 * (src/engine/sampling.h:90-122). The proposer proposes each pixel's next sample interval from its
 * recent samples (12 features, see bsx_scene_sample_learned in bsx_scene.h); the plain controller is
 * its fallback and the measure of its error.
 *
 * Parameter vector (BSX_MLP_PARAMS fp32 values): W1 [12 x 64] (input-major), b1 [64], W2 [64 x 64]
 * (row = layer-1 unit), b2 [64], w3 [64], b3. Forward: z1 = bf16(x) bf16(W1) + b1, h1 = bf16(silu(z1)),
 * z2 = h1 bf16(W2) + b2, y = bf16(silu(z2)) bf16(w3) + b3. Conventions as bsx_sensor.h.
 */
#ifndef BSX_MLP_H_
#define BSX_MLP_H_

#include "bsx_scene.h"
#include "bsx_sensor.h"

#ifdef __cplusplus
extern "C" {
#endif

#define BSX_MLP_INPUTS 12
#define BSX_MLP_PARAMS 5057

typedef struct bsx_mlp bsx_mlp;

/* Weights from a seeded uniform (Glorot bounds), biases zero. */
int bsx_mlp_create(uint64_t seed, bsx_mlp** out);
int bsx_mlp_destroy(bsx_mlp* m);
/* Host arrays of BSX_MLP_PARAMS values; set_params also clears Adam's moments and step count. */
int bsx_mlp_set_params(bsx_mlp* m, const float* params);
int bsx_mlp_get_params(bsx_mlp* m, float* params);
/* Gradient of the mean squared error over n rows (device x [n x 12] row-major, device target [n]) into
 * grad (host, BSX_MLP_PARAMS values); *loss = the mean squared error. */
int bsx_mlp_gradients(bsx_mlp* m, const float* x, const float* target, int64_t n, void* stream, float* grad,
                      double* loss);
/* `epochs` passes over the n device rows in order, batches of `batch` rows (a positive multiple of 64),
 * one Adam step (beta 0.9, 0.999, eps 1e-8) per batch at learning rate lr; *loss = the mean squared
 * error of the last pass, measured during it. */
int bsx_mlp_train(bsx_mlp* m, const float* x, const float* target, int64_t n, int32_t batch, int32_t epochs, float lr,
                  void* stream, double* loss);
/* y [n] (device) = the network's output for the n device rows x [n x 12]. */
int bsx_mlp_predict(bsx_mlp* m, const float* x, int64_t n, float* y, void* stream);
/* Self-check of the MFMA operand layout: c [32 x 32] = a [32 x 8] . b [8 x 32] (host arrays, a and b
 * bf16 bit patterns, row-major) through the kernels' tile routine on the device. */
int bsx_mlp_tile_check(const uint16_t* a, const uint16_t* b, float* c);

/* ---- Training examples and the learned sampler -------------------------------------------------------
 *
 * The controller's signal is L = ln(E + dark) of a pixel's irradiance E (renderer units). At each
 * decision it holds the pixel's four latest samples and a reference level ref that moves by theta_ref
 * per crossing (an ideal event sensor's), and feeds the network 12 features: 10 (L0 - ref); 10 (L_k - L0)
 * for k = 1..3; asinh of the three slopes (log units per ms); (ln h_k - ln 1e5) / 2 of the three gaps h_k
 * (ns); 10 (max - min of the four L); 10 (theta_band - |L0 - ref|). The output y proposes the interval
 * exp(2 y + ln 1e5) ns. */
typedef struct bsx_learned_examples_spec {
  float dark;          /* renderer units added to E (the controller's dark term) */
  float theta_ref;     /* the tracked reference's step */
  float theta_band;    /* half-width of the band about ref in which no event starts */
  float tol_log;       /* log-error tolerance of a chord */
  int64_t t0_ns;       /* start of the dense traces */
  int64_t delta_ns;    /* dense sampling interval (10 us) */
  int32_t n_dense;     /* dense samples per pixel trace (at least max_steps + 770) */
  int32_t max_steps;   /* the longest interval, in dense steps */
  uint64_t seed;
  int64_t n;           /* examples to write */
} bsx_learned_examples_spec;

/* Dense traces of every pixel of the view (n_dense samples delta_ns apart from t0_ns), then n training
 * examples: a pixel (one whose trace varies; constant ones in a tenth of their number), a decision
 * point and a history at log-uniform gaps of 1 to 256 dense steps before it, the 12 features into
 * x [n x 12] and into target [n] the scaled log of the longest interval, in dense steps up to
 * max_steps, whose chord (linear in E) keeps the log error of every dense point inside it within
 * tol_log, or within the distance to the band edge less tol_log while the trace stays inside the band
 * about ref (x and target in device memory). *n_active_pixels = the varying pixels. */
int bsx_scene_learned_examples(bsx_scene* s, const bsx_view* view, const bsx_learned_examples_spec* spec, float* x,
                               float* target, void* stream, uint64_t* n_active_pixels);

typedef struct bsx_learned_rules {
  bsx_sampler_rules plain; /* the fallback's rules: refine_tol_log (eta_L), dark, max_refine_levels, min_split_ns,
                              refine_signature (the lattice fields are not used) */
  float tol_log;           /* the a-posteriori check's log tolerance */
  float theta_band, theta_ref;
  float margin;            /* factor on the proposed interval */
  float near;              /* band room below which the controller samples as the plain one (0: never) */
  int64_t min_dt_ns;       /* lower bound of a proposal */
} bsx_learned_rules;

typedef struct bsx_learned_sampler bsx_learned_sampler;

/* The learned controller for one view of the scene, proposing with the given network (borrowed: it
 * must outlive the sampler). Per window and pixel (pixels no shape reaches take no samples, as with
 * the plain controller): from the newest sample, the network proposes an interval (times margin, at
 * least min_dt_ns, ending at the window's end at the latest); the scene is rendered at its end and at
 * its midpoint; if the chord's log error at the midpoint is within the allowed error (tol_log, or the
 * band room less tol_log while the three values stay inside the band) both become samples, otherwise
 * the plain controller refines the two halves (eta_L and signature flips, max_refine_levels,
 * min_split_ns) and every rendered point becomes a sample. Every render is a sample. */
int bsx_learned_sampler_create(bsx_scene* s, const bsx_view* view, const bsx_learned_rules* rules, bsx_mlp* proposer,
                               bsx_learned_sampler** out);
int bsx_learned_sampler_destroy(bsx_learned_sampler* ls);
/* The controller state at t0 (each pixel's sample at t0, the reference there). */
int bsx_learned_sampler_init(bsx_learned_sampler* ls, int64_t t0_ns, void* stream);
/* Window (t_w, t_w + 2^23] of every pixel: runs in pixel order in the scene's buffer set `slot` (as
 * bsx_scene_sample_window), the sample count, the network's proposals and the fallbacks taken.
 * sensor: the instantaneous-response model these runs feed (same pixels), or null. When given, the
 * controller reads the sensor's state at the window start: each pixel's band is the sensor's own
 * reference and thresholds, a pending release leaves it unknown (near mode) until it passes, and
 * crossings within the window move it as the sensor's reset rule does (a crossing time estimated
 * between samples). Near a band edge (room below `near`) the interval is two of the plain
 * controller's lattice steps for the pixel, judged by eta_L and signature flips at the midpoint. */
int bsx_learned_sampler_window(bsx_learned_sampler* ls, int64_t t_w, int32_t slot, bsx_sensor* sensor, void* stream,
                               bsx_device_runs* out, uint64_t* total_samples, uint64_t* proposals, uint64_t* fallbacks);

#ifdef __cplusplus
}
#endif

#endif /* BSX_MLP_H_ */
