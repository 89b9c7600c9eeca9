// Host-side launch wrappers of learned.hip: dense traces and training examples for the proposer, and
// the learned sampler's window passes.
#pragma once

#include <hip/hip_runtime.h>

#include "scene_dev.h"

namespace bsx {

// ln(1e5): the centre of the proposer's log-interval scale (features and output are (ln dt - this) / 2).
inline constexpr double kLnDtCenter = 11.512925464970229;

// The learned sampler's rules: the plain controller's (eta tolerance, dark, depth, min split,
// signature test) for its fallback, the log tolerance of the a-posteriori check, the band of no event
// start (theta_band about the tracked reference, which moves by theta_ref per crossing), the margin on
// the proposed interval and its lower bound.
struct LearnedRules {
  float dark;
  float tol_eta;
  float tol_log;
  float theta_band;
  float theta_ref;
  float margin;
  int32_t max_levels;
  int32_t signature;
  int64_t min_split;
  int64_t min_dt;
  // Near a threshold edge (room below `near`) the controller samples as the plain one: intervals of two
  // lattice steps of the plain level rule (d_px, l_min, l_max), judged by eta_L at their midpoint.
  float near;
  int32_t l_min, l_max;
  double d_px;
};

// The sensor a learned sampler feeds, read at each window start (null pointers: not coupled, the
// controller tracks an ideal reference of step theta_ref instead): per pixel its thresholds, scale and
// reference state (the instantaneous response's Delta = ln(I / i_scale) - s3_ref - n3_ref + dr_ref),
// a pending release (blind_until) and the reset rule.
struct SensorLink {
  const float* theta_on;
  const float* theta_off;
  const float* scale;
  const float* i_scale;
  const float* s3_ref;
  const float* n3_ref;
  const float* dr_ref;
  const int64_t* blind_until;
  float eps_ci;
  int32_t ideal_reset;
  int64_t refractory_ns;
};

// Per-pixel controller state (structure of arrays, P pixels): four sample times and log values
// ([k * P + p], k = 0 the newest), the count of valid ones, the tracked reference, the newest sample's
// irradiance and signature.
struct LearnedStateDev {
  int64_t* t;
  float* L;
  int32_t* n;
  float* ref;
  float* e0;
  uint32_t* sig;
};

hipError_t launch_dense_trace(const SceneDev& sc, const ViewDev& view, int64_t t0, int64_t delta, int J, float dark,
                              float theta_ref, float* trace, float* ref, uint8_t* active, hipStream_t stream);
hipError_t launch_examples(const float* trace, const float* ref, uint32_t P, int J, const uint32_t* pixels,
                           uint32_t n_pixels, int64_t delta, int max_steps, float theta_band, float tol, uint64_t seed,
                           int64_t n, float* x, float* target, hipStream_t stream);
hipError_t launch_learned_init(const SceneDev& sc, const ViewDev& view, int64_t t0, float dark,
                               const LearnedStateDev& s, hipStream_t stream);
hipError_t launch_learned_window(const SceneDev& sc, const ViewDev& view, const LearnedRules& r, const uint16_t* wb,
                                 const float* w, int64_t t_w, const LearnedStateDev& s, const SensorLink& link,
                                 int32_t* counts, const int64_t* row_start, int64_t* t_out, float* irr_out,
                                 unsigned long long* work, bool write, hipStream_t stream);

}  // namespace bsx
