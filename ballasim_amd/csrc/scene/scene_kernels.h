// Host-side launch wrappers of scene.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "scene_dev.h"

namespace bsx {

// The engine's sample request (reference: balla::backend::SampleRequest, src/backend/api.h:25-33), 24 bytes.
struct SampleRequestDev {
  uint16_t cam;
  uint16_t flags;
  uint32_t pixel;
  uint32_t k_begin;
  uint32_t k_count;
  int32_t t_off_ns;
  uint32_t pose_slot;
};
static_assert(sizeof(SampleRequestDev) == 24, "the engine's 24-byte SampleRequest");
// The class word of a sample on a textured shape (reference: kClassTextured, src/render/types.h:47).
inline constexpr uint32_t kClassTexturedWord = 1u << 13;

hipError_t launch_requests(const SceneDev& sc, const ViewDev& view, int64_t t_base, const SampleRequestDev* req,
                           size_t n, const uint32_t* out_index, int32_t groups, float* irr, uint64_t* sig,
                           uint32_t* klass, hipStream_t stream);

// One captured sample of the one-pass sampler: its offset from t_w and its value.
struct CapturedSample {
  uint32_t dt;
  float e;
};
static_assert(sizeof(CapturedSample) == 8, "8-byte capture slots");

hipError_t launch_render(const SceneDev& sc, const ViewDev& view, int64_t t_ns, float* out, hipStream_t stream);
hipError_t launch_count(const SceneDev& sc, const ViewDev& view, const bsx_sampler_rules& r, int64_t t_w,
                        int32_t* counts, hipStream_t stream);
hipError_t launch_write(const SceneDev& sc, const ViewDev& view, const bsx_sampler_rules& r, int64_t t_w,
                        const int64_t* row_start, int64_t* t, float* irr, hipStream_t stream);
// The one-pass sampler: counts, the first `cap` samples of every pixel in scratch slots [p cap, (p + 1) cap)
// and the largest count in *max_count (zeroed first by a fill kernel); then, once the counts are scanned,
// the captured runs copied into place and the pixels with more than `cap` samples sampled again in place.
hipError_t launch_capture(const SceneDev& sc, const ViewDev& view, const bsx_sampler_rules& r, int64_t t_w,
                          uint32_t cap, int32_t* counts, CapturedSample* scratch, unsigned int* max_count,
                          hipStream_t stream);
hipError_t launch_compact(const ViewDev& view, int64_t t_w, uint32_t cap, const int32_t* counts,
                          const int64_t* row_start, const CapturedSample* scratch, int64_t* t, float* irr,
                          hipStream_t stream);
hipError_t launch_overflow_write(const SceneDev& sc, const ViewDev& view, const bsx_sampler_rules& r, int64_t t_w,
                                 uint32_t cap, const int32_t* counts, const int64_t* row_start, int64_t* t, float* irr,
                                 hipStream_t stream);
// Samples at imposed per-pixel counts (scanned into row_start, total of them), evenly spread over
// (t_w, t_w + 2^23]; the scene's irradiance at each.
hipError_t launch_counts_write(const SceneDev& sc, const ViewDev& view, int64_t t_w, const int32_t* counts,
                               const int64_t* row_start, int64_t total, int64_t* t, float* irr, hipStream_t stream);
// Scratch of scan_scratch_count(n) int64 values; out[i] = sum of counts[0..i), scratch[blocks] = the total.
size_t scan_scratch_count(size_t n);
hipError_t launch_exclusive_scan(const int32_t* counts, size_t n, int64_t* out, int64_t* scratch, hipStream_t stream);

}  // namespace bsx
