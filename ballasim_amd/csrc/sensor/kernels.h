// Host-side launch wrappers of the sensor kernels (kernels.hip). counters[0] is the event counter,
// counters[1 + StatIdx] the work counters.
#pragma once

#include <hip/hip_runtime.h>

#include "kconst.h"

namespace bsx {

hipError_t launch_init(const KConst& kc, const MapsDev& mp, const StateDev& st, const float* irr0, int64_t t0,
                       unsigned long long* counters, hipStream_t stream);
// Phase A: the prepared values of every segment of every run (pos: the exclusive scan of the run
// lengths, total = pos[n_runs]), each run's estimated walk cost (work, added to) and its first segment
// that is not quiet under the input state (first_loud);
// phase B: the per-pixel walk (prep null: exact searches only; first_loud null: from the start).
hipError_t launch_instant_prep(const KConst& kc, const MapsDev& mp, const StateDev& in, const RunsDev& runs,
                               const uint32_t* pixel, const int64_t* pos, int64_t n_runs, int64_t total,
                               int64_t t_state, InstantPrep* out, float* work, int* first_loud, hipStream_t stream);
// The instantaneous walk over the pixels of n_cams cameras (descriptors in device memory), lane i
// taking pixel order[i] of the joint pixel space (order null: index order); queue (one word) is the
// persistent queue's counter when tu.chunk > 0.
hipError_t launch_instant_walk(const CameraDev* cams, int n_cams, uint32_t n_total, int groups, const uint32_t* order,
                               unsigned int* queue, const WalkTuning& tu, int64_t t_state, int64_t t_end,
                               hipStream_t stream);
// The cascade without noise (the run costs for its order: launch_run_cost, zero-filling work first).
hipError_t launch_run_cost(const KConst& kc, const MapsDev& mp, const StateDev& in, const RunsDev& runs,
                           const uint32_t* pixel, const int64_t* pos, int64_t n_runs, int64_t total, float* work,
                           hipStream_t stream);
hipError_t launch_cascade(const KConst& kc, const MapsDev& mp, const StateDev& in, const StateDev& out,
                          const RunsDev& runs, int64_t t_state, int64_t t_end, RawEventDev* ev, unsigned long long cap,
                          unsigned long long* counters, const uint32_t* order, int block, hipStream_t stream);
// Heavy-first order of the joint pixel space of n_cams cameras from their runs' estimated costs
// (CameraDev::work, filled by phase A or launch_run_cost): key (n_total bytes) and hist (128 words)
// are scratch, perm receives the order.
hipError_t launch_cost_order(const CameraDev* cams, int n_cams, uint32_t n_total, uint8_t* key, uint32_t* hist,
                             uint32_t* perm, hipStream_t stream);
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
hipError_t launch_check_samples(const float* irr, size_t count, float irr_max, unsigned long long* out,
                                hipStream_t stream);

}  // namespace bsx
