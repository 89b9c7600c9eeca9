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

hipError_t launch_render(const SceneDev& sc, const ViewDev& view, int64_t t_ns, float* out, hipStream_t stream);
hipError_t launch_count(const SceneDev& sc, const ViewDev& view, const bsx_sampler_rules& r, int64_t t_w,
                        int32_t* counts, hipStream_t stream);
hipError_t launch_write(const SceneDev& sc, const ViewDev& view, const bsx_sampler_rules& r, int64_t t_w,
                        const int64_t* row_start, int64_t* t, float* irr, hipStream_t stream);
// Samples at imposed per-pixel counts (scanned into row_start, total of them), evenly spread over
// (t_w, t_w + 2^23]; the scene's irradiance at each.
hipError_t launch_counts_write(const SceneDev& sc, const ViewDev& view, int64_t t_w, const int32_t* counts,
                               const int64_t* row_start, int64_t total, int64_t* t, float* irr, hipStream_t stream);
// Scratch of scan_scratch_count(n) int64 values; out[i] = sum of counts[0..i), scratch[blocks] = the total.
size_t scan_scratch_count(size_t n);
hipError_t launch_exclusive_scan(const int32_t* counts, size_t n, int64_t* out, int64_t* scratch, hipStream_t stream);

}  // namespace bsx
