// Host-side launch wrappers of the sensor kernels (kernels.hip). counters[0] is the event counter,
// counters[1 + StatIdx] the work counters.
#pragma once

#include <hip/hip_runtime.h>

#include "kconst.h"

namespace bsx {

hipError_t launch_init(const KConst& kc, const MapsDev& mp, const StateDev& st, const float* irr0, int64_t t0,
                       unsigned long long* counters, hipStream_t stream);
// One launch of the advance over one camera or a rig of cameras: descriptors and the call header in
// device memory (so that a captured graph replays any call whose extents fit its grids), the joint
// pixel order's scratch (key: n_total bytes, hist: 128 words), the order (perm: n_total words) and
// the persistent queue's counter.
struct ChainArgs {
  const CameraDev* cams;
  int n_cams;
  uint32_t n_total;      // pixels of all cameras
  int groups;            // the largest light-group count of the cameras (the kernels' register tier)
  bool instant;          // instantaneous response (else the cascade)
  bool noise;            // the cascade with its in-path noise
  const CallDev* call;
  int64_t seg_capacity;  // grid extents, at least the call's n_segments and n_runs
  int64_t run_capacity;
  uint8_t* key;
  uint32_t* hist;
  uint32_t* perm;
  unsigned int* queue;
  WalkTuning tu;
};
// Zeroes the counters and per-run arrays, runs phase A (instantaneous: every segment's prepared values,
// each run's estimated cost and first loud segment; cascade: the run costs) and the heavy-first pixel
// order of the joint pixel space.
hipError_t launch_chain_prepare(const ChainArgs& a, hipStream_t stream);
// The walk over every camera's pixels (persistent when tu.chunk > 0: chain_walk_blocks blocks, as many
// as are resident at once).
hipError_t launch_chain_walk(const ChainArgs& a, hipStream_t stream);
int chain_walk_blocks(const ChainArgs& a);
// Zeroes the counters and the queue only (before a re-run of the walk after an event-buffer overflow).
hipError_t launch_chain_reset(const ChainArgs& a, hipStream_t stream);
hipError_t launch_fill_f32(float* dst, size_t n, float v, hipStream_t stream);
// out[0] = pm::log monotonicity violations, out[1] = hash of its bits over every positive float.
hipError_t launch_math_selfcheck(unsigned long long* out, hipStream_t stream);
// Exclusive scan of int32 counts into int64 offsets (scratch: scan_scratch_count(n) values; the
// total lands in scratch[scan_scratch_count(n) - 1]). Defined with the scene kernels.
size_t scan_scratch_count(size_t n);
hipError_t launch_exclusive_scan(const int32_t* counts, size_t n, int64_t* out, int64_t* scratch, hipStream_t stream);
hipError_t launch_scatter_rows(const uint32_t* pixel, const int64_t* row_start, const int32_t* row_count, size_t n,
                               size_t n_samples, int32_t* row_of, uint32_t n_pixels, unsigned int* flags,
                               hipStream_t stream);
hipError_t launch_fill_u64(unsigned long long* dst, size_t n, unsigned long long v, hipStream_t stream);
// out[p * groups + g]: pixel p's run's last sample (row_of[p] >= 0 and a nonempty run; row_of null: no
// runs), else held[g * n_pixels + p].
hipError_t launch_held_frame(const float* held, const int32_t* row_of, const int64_t* row_start,
                             const int32_t* row_count, const float* irr, uint32_t n_pixels, int groups, float* out,
                             hipStream_t stream);
// Non-finite (out[0]) and clamped (out[1]) values of the samples the runs reference, one wave per run.
hipError_t launch_check_runs(const float* irr, const int64_t* row_start, const int32_t* row_count, size_t n_runs,
                             size_t n_samples, int groups, float irr_max, unsigned long long* out,
                             hipStream_t stream);

}  // namespace bsx
