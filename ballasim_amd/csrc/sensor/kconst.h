// Device-side data of one sensor model: the call constants, the per-pixel maps and the per-pixel
// working state, all structure-of-arrays so that a wave of 64 lanes (64 consecutive pixels) reads
// and writes each field with one coalesced 256-byte (fp32) or 512-byte (int64) access.
#pragma once

#include "../core/hd.h"

namespace bsx {

inline constexpr int kMaxGroups = 8;
inline constexpr int64_t kNotBlind = INT64_MIN;
inline constexpr int kPlainSegmentLog2 = 23;  // pieces never cross a multiple of 2^23 ns
inline constexpr int64_t kPlainSegmentNs = int64_t{1} << kPlainSegmentLog2;
inline constexpr int64_t kNoBreakpoint = INT64_MAX;

// Event labels and flag bits (reference: src/sensor/pixel.h:70-78).
inline constexpr uint8_t kLabelSignal = 0, kLabelNoise = 1, kLabelLeak = 2, kLabelHot = 3;
inline constexpr uint16_t kFlagHotPixel = 1u << 5;
inline constexpr uint16_t kFlagQuarantined = 1u << 9;
inline constexpr uint16_t kFlagSolverFallback = 1u << 10;
inline constexpr uint16_t kFlagFirstAfterInit = 1u << 12;
// This library's flag for events of the rate-specified shot-noise process.
inline constexpr uint16_t kFlagShot = 1u << 13;

// The cascade's in-path noise (noise on): the fp32 tables over ln a1 and their grid.
struct NoiseDev {
  const BSX_GLOBAL float* tab;  // n_nodes x stride (null: noise off)
  int32_t levels;               // leaf level L of the bridge tree (leaves of 2^(16 - L) ns)
  int32_t n_nodes, stride;
  float ln_min, inv_dln;        // the grid: node i at ln a1 = ln_min + i / inv_dln
  float margin_z;               // the exclusion margin's z
};

// Constants of one call of the fp32 kernel (kernel units: fA, s, log-e).
struct KConst {
  int32_t groups;
  int32_t instant;      // instantaneous log response
  int32_t ideal_reset;
  int32_t n_pixels;
  int64_t refractory_ns;
  float irr_max;
  float eps_ci;
  float label_cut;
  // Cascade (reference: CascadeConstants, src/sensor/substep.h:26-34).
  float k, a2, a3, tau_o, tau_sf, eps_accept, eps_drop;
  float eps_pl_base;
  // Shot noise (this library's extension): rate per pixel, Hz; 0 = off.
  double shot_rate_hz;
  uint64_t seed;
  uint32_t camera;
  uint32_t width;
  NoiseDev noise;
};

// Per-pixel constants in the kernel's precision (the reference casts its fp64 maps to T once per
// call: src/sensor/pixel.h:771-775).
struct MapsDev {
  const BSX_GLOBAL float* theta_on;
  const BSX_GLOBAL float* theta_off;
  const BSX_GLOBAL float* leak;
  const BSX_GLOBAL float* dark;
  const BSX_GLOBAL float* scale;  // illuminance_scale x gain, formed in fp32 as the reference's scale_
  const BSX_GLOBAL float* phi0;
  const BSX_GLOBAL uint8_t* hot;  // 0: none, 1: hot (its events are flagged and labelled HOT)
};

// Working state (the fp32 kernel's view of the reference's PixelState, src/sensor/pixel.h:105-121;
// the fp32 values round-trip exactly through the reference's fp64 storage, so the saved-state format
// is produced on the host by widening).
struct StateDev {
  BSX_GLOBAL float* z;
  BSX_GLOBAL float* x1;
  BSX_GLOBAL float* s2;
  BSX_GLOBAL float* s3;
  BSX_GLOBAL float* i_scale;
  BSX_GLOBAL float* tv;
  BSX_GLOBAL float* s3_ref;
  BSX_GLOBAL float* n3_ref;
  BSX_GLOBAL float* dr_ref;
  BSX_GLOBAL int64_t* t_ref;
  BSX_GLOBAL int64_t* blind_until;
  BSX_GLOBAL int64_t* t_search;
  BSX_GLOBAL int64_t* shot_next;  // next shot-noise arrival (ns), this library's extension
  BSX_GLOBAL uint32_t* shot_index;
  BSX_GLOBAL uint16_t* flags;
  BSX_GLOBAL float* held;  // [g * n_pixels + p]
  // The in-path noise (null when off): the noise (n1, n2, n3) at the current cell's start and end
  // nodes ([i * n_pixels + p]), the cell's frozen rate ln a1, the cell index and whether the end
  // node is drawn.
  BSX_GLOBAL float* nk;
  BSX_GLOBAL float* nk1;
  BSX_GLOBAL float* ln_a1;
  BSX_GLOBAL int64_t* cell;
  BSX_GLOBAL uint8_t* have_k1;
};

// Sample runs of a call as the kernels read them: row_of[p] is pixel p's run or -1.
struct RunsDev {
  const BSX_GLOBAL int32_t* row_of;
  const BSX_GLOBAL int64_t* row_start;
  const BSX_GLOBAL int32_t* row_count;
  const BSX_GLOBAL int64_t* t;
  const BSX_GLOBAL float* irr;
};

// Phase A: consecutive positions per thread, and the positions of a 256-thread block, to which every
// camera's segment base is aligned.
inline constexpr int kPrepPerThread = 8;  // consecutive segment positions per phase-A thread
inline constexpr int64_t kSegmentAlign = 256 * kPrepPerThread;

struct RawEventDev;

// One camera's view for a launch. A launch covers one camera or a rig of several: their pixels form
// one index space (camera c holds [base, base + kc.n_pixels)), their runs one run space ([run_base,
// run_base + n_runs)) and their segments one segment space ([seg_base, seg_base + total)); run_base
// is a multiple of 256 and seg_base of kSegmentAlign, so that no block of a run- or segment-indexed
// kernel spans two cameras.
// The photocurrent range of kScreenBlock consecutive segments of one run (aligned positions of the
// camera's segment space), written by phase A: the least lower and the greatest upper end of the
// segments' photocurrent ranges, as InstantLane::segment_range forms them, and whether every one of
// them has equal end values in every group (steady: the photocurrent is then lo = hi at every time of
// the block). lo = 0, hi = +inf: the block holds a position that is not a segment of one run.
inline constexpr int kScreenBlockLog2 = 4;
inline constexpr int kScreenBlock = 1 << kScreenBlockLog2;
struct ScreenBlock {
  float lo, hi;
  uint32_t steady;
};

struct CameraDev {
  KConst kc;
  MapsDev mp;
  StateDev in, out;
  RunsDev runs;
  const BSX_GLOBAL uint32_t* pixel;       // the runs' pixel ids
  const BSX_GLOBAL int64_t* pos;          // the runs' segment positions (exclusive scan of the run lengths)
  BSX_GLOBAL int* first_loud;             // per run (null: every walk from its start)
  BSX_GLOBAL ScreenBlock* screen_blocks;  // per kScreenBlock positions (null: none)
  BSX_GLOBAL float* work;                 // per run: estimated walk cost (null: no order information)
  BSX_GLOBAL RawEventDev* ev;
  unsigned long long cap;
  unsigned long long* counters;  // [0] events, [1 + StatIdx] work counters
  int64_t t_state;             // the camera's model time at the start of the call
  uint32_t base;
  int64_t n_runs, run_base;
  int64_t total, seg_base;
};

// What a launch's kernels read besides the cameras: the call's end time and the extents of the joint
// run and segment spaces (padding included). Kept in device memory, so that a captured graph of the
// launch replays any call whose extents fit the grids it was captured with.
struct CallDev {
  int64_t t_end;
  int64_t n_segments;
  int64_t n_runs;
};

// The walk's launch shape, chosen by the autotuner: threads per block, segments a walking lane does
// before the wave looks again, waiting lanes that start a joint search, groups a wave takes per grab
// of the persistent queue (0: one group per wave, no queue), and the head of the heavy-first order
// walked in groups of fewer than 64 pixels.
struct WalkTuning {
  int32_t block;
  int32_t walk_budget;
  int32_t search_batch;
  int32_t chunk;
  int32_t head_lanes = 64;     // pixels a group takes at the head of the heavy-first order, [1, 64]
  int32_t head_per_mille = 0;  // the head's share of the order's pixels, in thousandths
  int32_t head_groups = 0;     // the head's groups, set per launch from the above (head_groups * head_lanes
                               // <= the walk's pixels); 0 where the walk has no heavy-first order
};

struct RawEventDev {
  uint32_t pixel;
  uint8_t polarity;
  uint8_t label;
  uint16_t flags;
  int64_t t_cross;
  float ds, dn, dl;
  float l_first;
};
static_assert(sizeof(RawEventDev) == 32, "RawEventDev must keep the reference's 32-byte RawEvent layout");

// Per-call work counters, summed over a wave before one atomic add each. kStClamps and kStNonFinite
// count the input samples of the call's runs (the clamp changes a value; a value is not finite) and
// are counted by the chain's phase A, before the walk.
enum StatIdx : int {
  kStSubsteps = 0,
  kStAttempts,
  kStSettled,
  kStCells,
  kStRefinements,
  kStLeaves,
  kStFallbacks,
  kStClamps,
  kStQuarantines,
  kStPieces,
  kStNonFinite,
  kStCount
};

}  // namespace bsx
