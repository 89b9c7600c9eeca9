// gfx950 kernels of the sensor path: init, the instantaneous-response advance, the input checks and
// the run scatter. One lane per pixel; a 64-lane wave covers 64 consecutive pixels, so every state,
// map and row-index access is one coalesced vector access; events leave through a wave-aggregated
// sink (one atomic per wave and push site, consecutive slots for the active lanes).
#include <hip/hip_runtime.h>

#include "cascade.h"
#include "kernels.h"
#include "instant.h"
#include "noise.h"

namespace bsx {

namespace {

__device__ __forceinline__ float clamp_irr(float v, float vmax_) { return v >= 0.0f ? (v <= vmax_ ? v : vmax_) : 0.0f; }

// Sum of a per-lane counter over the wave, added once to the global counter.
__device__ __forceinline__ void wave_add(unsigned long long* dst, uint32_t v) {
  unsigned long long s = v;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (__lane_id() == 0 && s != 0) atomicAdd(dst, s);
}

// The input check of one sample value, as advance_runs makes it on the host: a value that is not
// finite, or one the clamp changes.
struct InputTally {
  uint32_t bad = 0, clamped = 0;
  __device__ __forceinline__ void add(float x, float vmax_) {
    const bool finite = x - x == 0.0f;
    bad += finite ? 0u : 1u;
    clamped += finite && clamp_irr(x, vmax_) != x ? 1u : 0u;
  }
  __device__ __forceinline__ void flush(unsigned long long* counters) {
    wave_add(counters + 1 + kStNonFinite, bad);
    wave_add(counters + 1 + kStClamps, clamped);
  }
};

template <int kMaxG>
__global__ __launch_bounds__(256) void init_kernel(KConst kc, MapsDev mp, StateDev st, const float* irr0, int64_t t0,
                                                   unsigned long long* counters) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t clamps = 0, cells = 0, refinements = 0;
  if (p < kc.n_pixels) {
    NoiseLane<kMaxG> L;
    L.G = kc.groups;
    L.dark = mp.dark[p];
    L.scale = mp.scale[p];
    L.irr_max = kc.irr_max;
    L.n = 0;
    L.skip = 0;
    L.t_state = t0;
    float mean_I = L.dark;
    for (int g = 0; g < kMaxG; ++g) {
      if (g >= kc.groups) continue;
      const float raw = irr0[static_cast<int64_t>(p) * kc.groups + g];
      const float v = clamp_irr(raw, kc.irr_max);
      if (!(v == raw)) ++clamps;
      L.held[g] = v;
      L.va[g] = v;
      mean_I = mean_I + L.scale * v;
      st.held[static_cast<int64_t>(g) * kc.n_pixels + p] = v;
    }
    L.ta = t0;
    L.tb = kNoBreakpoint;  // the held segment: I = dark + s x held
    L.i_scale = mean_I;
    st.i_scale[p] = mean_I;
    st.z[p] = 1.0f / mean_I;
    st.x1[p] = 0.0f;
    st.s2[p] = 0.0f;
    st.s3[p] = 0.0f;
    st.tv[p] = 0.0f;
    const float s3_ref = kc.instant ? L.log_ratio(L.photocurrent(t0)) : 0.0f;
    float n3_0 = 0.0f;
    if (!kc.instant && kc.noise.tab != nullptr) {
      // the cascade's in-path noise: the stationary noise at the cell's start node, extended to t0
      L.pixel = static_cast<uint32_t>(p);
      L.seed = kc.seed;
      L.camera = kc.camera;
      L.nd = kc.noise;
      L.margin_z = kc.noise.margin_z;
      L.cc = cascade::Consts{kc.k, kc.a2, kc.a3, kc.tau_o, kc.tau_sf, kc.eps_accept, kc.eps_drop};
      L.z = 1.0f / mean_I;
      L.x1 = 0.0f;
      L.s2 = 0.0f;
      L.s3 = 0.0f;
      L.n_cells = L.n_refinements = 0;
      L.mid = nullptr;
      L.mid_stride = 0;
      n3_0 = L.init(t0, mean_I);
      for (int i = 0; i < 3; ++i) {
        st.nk[static_cast<int64_t>(i) * kc.n_pixels + p] = L.nk[i];
        st.nk1[static_cast<int64_t>(i) * kc.n_pixels + p] = L.nk1[i];
      }
      st.ln_a1[p] = L.ln_a1;
      st.cell[p] = L.cell;
      st.have_k1[p] = L.have_k1 ? 1 : 0;
      cells = L.n_cells;
      refinements = L.n_refinements;
    }
    st.t_ref[p] = t0;
    st.s3_ref[p] = s3_ref;
    st.n3_ref[p] = n3_0;
    st.dr_ref[p] = mp.phi0[p];
    st.blind_until[p] = kNotBlind;
    st.t_search[p] = t0 + 1;
    st.flags[p] = kFlagFirstAfterInit;
    st.shot_next[p] = kc.shot_rate_hz > 0.0 ? t0 + InstantLane<kMaxG>::shot_gap(kc.shot_rate_hz, kc.seed, kc.camera,
                                                                                   static_cast<uint32_t>(p), 0, nullptr)
                                            : INT64_MAX;
    st.shot_index[p] = 0;
  }
  wave_add(counters + 1 + kStClamps, clamps);
  wave_add(counters + 1 + kStCells, cells);
  wave_add(counters + 1 + kStRefinements, refinements);
}

// Sum of v over the lanes of the wave that share `key` (keys non-decreasing along the lanes), added
// once per key to dst[key] by the key's last lane (key < 0: no contribution).
__device__ __forceinline__ void wave_segmented_add(float* dst, int key, float v) {
  const int lane = __lane_id();
  for (int off = 1; off < 64; off <<= 1) {
    const float vu = __shfl_up(v, off);
    const int ku = __shfl_up(key, off);
    if (lane >= off && ku == key) v += vu;
  }
  const int kn = __shfl_down(key, 1);
  if (key >= 0 && (lane == 63 || kn != key)) atomicAdd(dst + key, v);
}

// Phase A of the instantaneous response: one thread per segment of every run (position q of the
// runs laid end to end, pos the exclusive scan of the run lengths), all runs at once, so the
// transcendental work of the searches is spread evenly over the device whatever the runs' lengths.
// It also estimates each run's walk cost for the heavy-first pixel order: one unit per segment plus
// kSearchWeight per threshold's worth of log-intensity change (where the searches and events are).
constexpr float kSearchWeight = 32.0f;

// wave_segmented_add for counts.
__device__ __forceinline__ void wave_segmented_count(uint32_t* dst, int key, uint32_t v) {
  const int lane = __lane_id();
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t vu = __shfl_up(v, off);
    const int ku = __shfl_up(key, off);
    if (lane >= off && ku == key) v += vu;
  }
  const int kn = __shfl_down(key, 1);
  if (key >= 0 && (lane == 63 || kn != key) && v != 0) atomicAdd(dst + key, v);
}

// The run i of position q (the last i with pos[i] <= q; pos non-decreasing), by bisection.
__device__ __forceinline__ int64_t run_of_position(const int64_t* pos, int64_t n_runs, int64_t q) {
  int64_t lo = 0, hi = n_runs;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (pos[mid] <= q) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The last run at or after run i (pos[i] <= q) whose position is <= q: a few steps forward, then a
// bisection, so that a long stretch of empty runs (pixels without samples share one position) costs
// a logarithm, not a step per run.
__device__ __forceinline__ int64_t forward_run(const int64_t* pos, int64_t n_runs, int64_t i, int64_t q) {
  for (int s = 0; s < 4; ++s) {
    if (!(i + 1 < n_runs && pos[i + 1] <= q)) return i;
    ++i;
  }
  if (!(i + 1 < n_runs && pos[i + 1] <= q)) return i;
  int64_t a = i + 1, b = n_runs;  // pos[a] <= q; the answer lies in [a, b)
  while (b - a > 1) {
    const int64_t mid = a + ((b - a) >> 1);
    if (pos[mid] <= q) a = mid;
    else b = mid;
  }
  return a;
}

// The same for the block's threads at once (q = the block's first position + threadIdx.x): one
// bisection for the block's first position, then forward_run per thread. Called by every thread of
// the block.
__device__ __forceinline__ int64_t block_run_of_position(const int64_t* pos, int64_t n_runs, int64_t q0,
                                                         int64_t q, int64_t total) {
  __shared__ int64_t first;
  if (threadIdx.x == 0) first = q0 < total ? run_of_position(pos, n_runs, q0) : 0;
  __syncthreads();
  return q < total ? forward_run(pos, n_runs, first, q) : first;
}

// The camera of a launch that holds offset q of the joint run (kRuns) or segment space (cameras in
// offset order, a handful at most; an empty camera shares its offset with the next and is skipped).
enum class Space { kRuns, kSegments };
template <Space kSpace>
__device__ __forceinline__ int camera_at(const CameraDev* cams, int n_cams, int64_t q) {
  int c = 0;
  while (c + 1 < n_cams && q >= (kSpace == Space::kRuns ? cams[c + 1].run_base : cams[c + 1].seg_base)) ++c;
  return c;
}

// The runs of a 256-thread block of consecutive positions q0 + threadIdx.x: the runs [first, first +
// 256) staged in LDS (first = the run of q0), so that each thread finds its run by bisection in LDS
// and reads its pixel and row start there. Positions beyond the staged runs (a block that spans more
// than 256 runs: short or empty ones) continue with a bisection in global memory. Called by every
// thread of the block.
struct BlockRuns {
  int64_t pos[256];
  int64_t row_start[256];
  uint32_t pixel[256];
  int64_t first;
  int count;
};

struct RunRef {
  int64_t i, pos, row_start;
  uint32_t pixel;
};

__device__ __forceinline__ RunRef stage_block_runs(BlockRuns& br, const int64_t* pos, const RunsDev& runs,
                                                   const uint32_t* pixel, int64_t n_runs, int64_t q0, int64_t q,
                                                   int64_t total) {
  if (threadIdx.x == 0) br.first = q0 < total ? run_of_position(pos, n_runs, q0) : 0;
  __syncthreads();
  const int64_t first = br.first;
  const int64_t left = n_runs - first;
  const int m = static_cast<int>(left < 256 ? left : 256);
  if (threadIdx.x == 0) br.count = m;
  if (static_cast<int>(threadIdx.x) < m) {
    const int64_t r = first + threadIdx.x;
    br.pos[threadIdx.x] = pos[r];
    br.row_start[threadIdx.x] = runs.row_start[r];
    br.pixel[threadIdx.x] = pixel[r];
  }
  __syncthreads();
  RunRef ref{first, 0, 0, 0};
  if (!(q < total) || m <= 0) return ref;
  int lo = 0, hi = m;  // br.pos[lo] <= q (br.pos[0] <= q0 <= q); the answer lies in [lo, hi)
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (br.pos[mid] <= q) lo = mid;
    else hi = mid;
  }
  int64_t i = first + lo;
  if (lo == m - 1 && i + 1 < n_runs && pos[i + 1] <= q) {  // past the staged runs: bisect the rest
    int64_t a = i + 1, b = n_runs;                          // pos[a] <= q; the answer lies in [a, b)
    while (b - a > 1) {
      const int64_t mid = a + ((b - a) >> 1);
      if (pos[mid] <= q) a = mid;
      else b = mid;
    }
    i = a;
  }
  if (i == first + lo) {
    ref.i = i;
    ref.pos = br.pos[lo];
    ref.row_start = br.row_start[lo];
    ref.pixel = br.pixel[lo];
  } else {
    ref.i = i;
    ref.pos = pos[i];
    ref.row_start = runs.row_start[i];
    ref.pixel = pixel[i];
  }
  return ref;
}

// Moves ref forward to the run of position q >= ref's position (staged runs from LDS, the others from
// global memory).
__device__ __forceinline__ void advance_run(RunRef& ref, const BlockRuns& br, const int64_t* pos, const RunsDev& runs,
                                            const uint32_t* pixel, int64_t n_runs, int64_t q) {
  for (int s = 0; s < 4 && ref.i + 1 < n_runs; ++s) {
    const int64_t k = ref.i + 1 - br.first;
    const bool staged = k < br.count;
    const int64_t next = staged ? br.pos[k] : pos[ref.i + 1];
    if (next > q) return;
    ref.i += 1;
    ref.pos = next;
    ref.row_start = staged ? br.row_start[k] : runs.row_start[ref.i];
    ref.pixel = staged ? br.pixel[k] : pixel[ref.i];
  }
  const int64_t i = forward_run(pos, n_runs, ref.i, q);  // a long stretch of empty runs
  if (i != ref.i) {
    ref.i = i;
    ref.pos = pos[i];
    ref.row_start = runs.row_start[i];
    ref.pixel = pixel[i];
  }
}

// Sum of cost and minimum of loud over the lanes of the wave that share `key` (keys non-decreasing
// along the lanes), folded once per key into work[key] and first_loud[key] by the key's last lane
// (key < 0: no contribution; a null destination is skipped).
__device__ __forceinline__ void wave_segmented_cost_loud(float* work, int* first_loud, int key, float cost, int loud) {
  const int lane = __lane_id();
  for (int off = 1; off < 64; off <<= 1) {
    const float cu = __shfl_up(cost, off);
    const int lu = __shfl_up(loud, off);
    const int ku = __shfl_up(key, off);
    if (lane >= off && ku == key) {
      cost += cu;
      loud = lu < loud ? lu : loud;
    }
  }
  const int kn = __shfl_down(key, 1);
  if (key >= 0 && (lane == 63 || kn != key)) {
    if (work != nullptr) atomicAdd(work + key, cost);
    if (first_loud != nullptr && loud != INT32_MAX) atomicMin(first_loud + key, loud);
  }
}

// Phase A of the instantaneous response as a screen of every segment of every camera's runs, without
// logarithms: one block of kSegmentAlign consecutive positions of one camera, each thread
// kPrepPerThread consecutive positions (a pixel's constants and photocurrent bounds formed once for its
// consecutive segments, a segment's end sample carried into the next). A segment is quiet under the
// call's input state when no release or shot arrival is pending before its end, it holds no 2^23-ns
// cut, and its photocurrent stays inside the pixel's screen bounds over the whole call
// (InstantLane::screen_bounds from t_state to t_end - 1, InstantLane::screened): no integer of it fires
// under that state. Each run's first segment that is not quiet (the walk starts there) and its walk
// cost (one unit per loud segment plus kSearchWeight per threshold's worth of log-intensity change, fast
// logs: a schedule only) are folded per run. Nothing is written per segment. The kernel also makes the
// call's input check: every value of every run's samples is read here exactly once (a segment's end
// sample; a run's first sample when it lies at the model time and ends no segment), and the
// camera's kStNonFinite and kStClamps counters take the values that are not finite or that the clamp
// changes.
template <int kMaxG>
__global__ __launch_bounds__(256) void chain_screen_kernel(const CameraDev* cams, int n_cams, const CallDev* call) {
  const int64_t g0 = static_cast<int64_t>(blockIdx.x) * kSegmentAlign;
  if (g0 >= call->n_segments) return;  // beyond this call's segments (a grid sized for a larger one)
  const CameraDev& cd = cams[camera_at<Space::kSegments>(cams, n_cams, g0)];
  const KConst kc = cd.kc;
  const MapsDev mp = cd.mp;
  const StateDev in = cd.in;
  const RunsDev runs = cd.runs;
  const uint32_t* const pixel = cd.pixel;
  const int64_t* const pos = cd.pos;
  float* const work = cd.work;
  int* const first_loud = cd.first_loud;
  const int64_t t_state = cd.t_state, total = cd.total, n_runs = cd.n_runs, t_end = call->t_end;
  const int64_t q0 = g0 - cd.seg_base, qt = q0 + static_cast<int64_t>(threadIdx.x) * kPrepPerThread;
  __shared__ BlockRuns br;
  RunRef rr = stage_block_runs(br, pos, runs, pixel, n_runs, q0, qt, total);
  int acc_run = -1;
  float cost_acc = 0.0f;
  int loud_acc = INT32_MAX;
  InputTally tally;
  InstantLane<kMaxG> L;
  bool usable = false;    // the pixel's screen applies (leak >= 0, n3_ref = 0, not blind)
  int64_t loaded = -1;    // the run whose pixel constants L holds
  int64_t prev_end = -1;  // the position whose segment's end sample L.tb, L.vb hold (-1: none)
  // this thread's half of a screen block: its positions all segments of run blk_run, their range
  bool blk_ok = true, blk_steady = true;
  int blk_run = -1;
  float blk_lo = INFINITY, blk_hi = -INFINITY;
  // The thread's positions, a run at a time: the run's lookup, bookkeeping and constants once per run,
  // then its positions in a plain loop.
  const int64_t q_end = vmin(qt + kPrepPerThread, total);
  if (qt + kPrepPerThread > total) blk_ok = false;
  for (int64_t q = qt; q < q_end;) {
    advance_run(rr, br, pos, runs, pixel, n_runs, q);
    const int run = static_cast<int>(rr.i);
    const int64_t k1 = rr.i + 1 - br.first;
    const int64_t next_pos = rr.i + 1 < n_runs ? (k1 < br.count ? br.pos[k1] : pos[rr.i + 1]) : total;
    const int64_t q_stop = vmin(q_end, next_pos);
    blk_ok &= blk_run < 0 || run == blk_run;
    blk_run = run;
    if (run != acc_run) {
      if (acc_run >= 0) {
        if (work != nullptr) atomicAdd(work + acc_run, cost_acc);
        if (first_loud != nullptr && loud_acc != INT32_MAX) atomicMin(first_loud + acc_run, loud_acc);
      }
      acc_run = run;
      cost_acc = 0.0f;
      loud_acc = INT32_MAX;
    }
    const uint32_t p = rr.pixel;
    if (rr.i != loaded) {
      loaded = rr.i;
      L.G = kc.groups;
      L.irr_max = kc.irr_max;
      L.dark = mp.dark[p];
      L.scale = mp.scale[p];
      L.leak = mp.leak[p];
      L.theta_on = mp.theta_on[p];
      L.theta_off = mp.theta_off[p];
      L.i_scale = in.i_scale[p];
      L.s3_ref = in.s3_ref[p];
      L.n3_ref = in.n3_ref[p];
      L.dr_ref = in.dr_ref[p];
      L.t_ref = in.t_ref[p];
      L.blind_until = in.blind_until[p];
      L.shot_next = in.shot_next[p];
      usable = L.blind_until == kNotBlind && L.leak >= 0.0f && L.n3_ref == 0.0f;
      if (usable) L.screen_bounds(t_state, t_end - 1);
    }
    const int64_t* t = runs.t + rr.row_start;
    const float* v = runs.irr + rr.row_start * kc.groups;
    for (; q < q_stop; ++q) {
      const int64_t j = q - rr.pos;
      if (j == 0) {
        if (t[0] == t_state) {  // the first sample is the model time's: no segment ends at it
          for (int g = 0; g < kMaxG; ++g)
            if (g < kc.groups) tally.add(v[g], kc.irr_max);
          prev_end = -1;
          blk_ok = false;
          continue;
        }
        L.ta = t_state;
        for (int g = 0; g < kMaxG; ++g)
          if (g < kc.groups) L.va[g] = in.held[static_cast<int64_t>(g) * kc.n_pixels + p];
      } else if (prev_end == q - 1) {  // the previous position's segment ends where this one starts
        L.ta = L.tb;
        for (int g = 0; g < kMaxG; ++g)
          if (g < kc.groups) L.va[g] = L.vb[g];
      } else {
        L.ta = t[j - 1];
        for (int g = 0; g < kMaxG; ++g)
          if (g < kc.groups) L.va[g] = L.clampv(v[(j - 1) * kc.groups + g]);
      }
      L.tb = t[j];
      for (int g = 0; g < kMaxG; ++g) {
        if (g < kc.groups) {
          const float x = v[j * kc.groups + g];
          tally.add(x, kc.irr_max);
          L.vb[g] = L.clampv(x);
        }
      }
      prev_end = q;
      float seg_lo, seg_hi;
      L.segment_range(seg_lo, seg_hi);
      blk_lo = fminf(blk_lo, seg_lo);
      blk_hi = fmaxf(blk_hi, seg_hi);
      blk_steady &= L.steady_segment();
      const bool quiet = usable & (L.shot_next >= L.tb) &
                         (((L.ta >> kPlainSegmentLog2) + 1) * kPlainSegmentNs >= L.tb) & L.in_screen(seg_lo, seg_hi);
      if (!quiet) {
        loud_acc = vmin(loud_acc, static_cast<int>(j));
        float Ia = L.dark, Ib = L.dark;
        for (int g = 0; g < kMaxG; ++g) {
          if (g < kc.groups) {
            Ia += L.scale * L.va[g];
            Ib += L.scale * L.vb[g];
          }
        }
        // (a schedule only: hardware logarithm and reciprocal)
        cost_acc += 1.0f + kSearchWeight * fabs_(__logf(Ib) - __logf(Ia)) *
                               __builtin_amdgcn_rcpf(vmin(L.theta_on, L.theta_off));
      }
    }
  }
  wave_segmented_cost_loud(work, first_loud, acc_run, cost_acc, loud_acc);
  tally.flush(cd.counters);
  // Screen blocks: the two threads of an even-odd lane pair cover one aligned block of kScreenBlock
  // positions (kPrepPerThread each); the even one writes its summary.
  static_assert(2 * kPrepPerThread == kScreenBlock, "a screen block is two threads' positions");
  const int p_run = __shfl_xor(blk_run, 1);
  const bool p_ok = __shfl_xor(static_cast<int>(blk_ok), 1) != 0;
  const bool p_steady = __shfl_xor(static_cast<int>(blk_steady), 1) != 0;
  const float p_lo = __shfl_xor(blk_lo, 1), p_hi = __shfl_xor(blk_hi, 1);
  if (cd.screen_blocks != nullptr && (threadIdx.x & 1) == 0 && qt < total) {
    const bool ok = blk_ok & p_ok & (blk_run == p_run);
    cd.screen_blocks[qt >> kScreenBlockLog2] =
        ok ? ScreenBlock{fminf(blk_lo, p_lo), fmaxf(blk_hi, p_hi), (blk_steady & p_steady) ? 1u : 0u}
           : ScreenBlock{0.0f, INFINITY, 0u};
  }
}

// Before phase A: zero each camera's counters, the order histogram and the queue counter (block 0),
// and set every run's cost to 0 and first loud segment to none (one thread per run, cameras' run
// bases multiples of 256). walk_only: the walk's counters only, phase A's input counts kept.
__device__ __forceinline__ void reset_counters(const CameraDev* cams, int n_cams, unsigned int* queue,
                                               bool walk_only) {
  for (int k = threadIdx.x; k < n_cams * (1 + kStCount); k += blockDim.x) {
    const int idx = k % (1 + kStCount);
    if (walk_only && (idx == 1 + kStClamps || idx == 1 + kStNonFinite)) continue;
    cams[k / (1 + kStCount)].counters[idx] = 0ull;
  }
  if (threadIdx.x == 0 && queue != nullptr) *queue = 0u;
}

__global__ __launch_bounds__(256) void chain_init_kernel(const CameraDev* cams, int n_cams, const CallDev* call,
                                                         uint32_t* hist, unsigned int* queue) {
  if (blockIdx.x == 0) {
    reset_counters(cams, n_cams, queue, false);
    if (threadIdx.x < 128) hist[threadIdx.x] = 0u;
  }
  const int64_t g0 = static_cast<int64_t>(blockIdx.x) * blockDim.x;
  if (g0 >= call->n_runs) return;
  const CameraDev& c = cams[camera_at<Space::kRuns>(cams, n_cams, g0)];
  const int64_t r = g0 - c.run_base + threadIdx.x;
  if (r < c.n_runs) {
    if (c.work != nullptr) c.work[r] = 0.0f;
    if (c.first_loud != nullptr) c.first_loud[r] = INT32_MAX;
  }
}

// Before a re-run of the walk: the walk's counters and the queue counter only.
__global__ void chain_reset_kernel(const CameraDev* cams, int n_cams, unsigned int* queue) {
  reset_counters(cams, n_cams, queue, true);
}

// The pixel order of a walk: pixels by estimated cost, heaviest first (64 logarithmic buckets, bucket
// 0 = no run), so that the long walks start first and waves hold pixels of like cost. A schedule
// only: every pixel's result is independent of where and when it runs.
__device__ __forceinline__ int cost_bucket(float w) {
  if (!(w > 0.0f)) return 0;
  const int b = 1 + static_cast<int>(4.0f * __log2f(1.0f + w));
  return b < 1 ? 1 : (b > 63 ? 63 : b);
}

// The camera of index g of a launch's pixel space (cameras in base order, a handful at most).
__device__ __forceinline__ int camera_of(const CameraDev* cams, int n_cams, uint32_t g) {
  int c = 0;
  while (c + 1 < n_cams && g >= cams[c + 1].base) ++c;
  return c;
}

// Pixels per thread of the order kernels (fewer blocks: fewer global atomics on the 64 counters).
constexpr int kOrderPerThread = 16;

// For the lanes of a wave with a bucket b >= 0: one LDS atomic per distinct bucket adds the lanes of
// that bucket to cnt[b]; returns this lane's slot among them (the old count plus its rank among the
// wave's lanes of the same bucket).
__device__ __forceinline__ uint32_t wave_bucket_slot(uint32_t* cnt, int b) {
  const int lane = __lane_id();
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  unsigned long long todo = __ballot(b >= 0);
  uint32_t slot = 0;
  while (todo != 0) {
    const int leader = __ffsll(static_cast<long long>(todo)) - 1;
    const int bl = __shfl(b, leader);
    const unsigned long long same = __ballot(b == bl) & todo;
    uint32_t w = 0;
    if (lane == leader) w = atomicAdd(&cnt[bl], static_cast<uint32_t>(__popcll(same)));
    w = __shfl(w, leader);
    if (b == bl) slot = w + static_cast<uint32_t>(__popcll(same & below));
    todo &= ~same;
  }
  return slot;
}

__global__ __launch_bounds__(256) void order_count_kernel(const CameraDev* cams, int n_cams, uint32_t n, uint8_t* key,
                                                          uint32_t* hist) {
  __shared__ uint32_t cnt[64];
  if (threadIdx.x < 64) cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * (256u * kOrderPerThread);
  for (int k = 0; k < kOrderPerThread; ++k) {
    const uint32_t g = base + static_cast<uint32_t>(k) * 256u + threadIdx.x;
    int b = -1;
    if (g < n) {
      const CameraDev& c = cams[camera_of(cams, n_cams, g)];
      const uint32_t p = g - c.base;
      const int32_t r = c.runs.row_of != nullptr ? c.runs.row_of[p] : -1;
      b = r >= 0 && c.work != nullptr ? cost_bucket(c.work[r]) : 0;
      key[g] = static_cast<uint8_t>(b);
    }
    (void)wave_bucket_slot(cnt, b);
  }
  __syncthreads();
  if (threadIdx.x < 64 && cnt[threadIdx.x] != 0) atomicAdd(hist + threadIdx.x, cnt[threadIdx.x]);
}

// cursor[b] = the number of pixels in buckets above b (one wave).
__global__ void order_offsets_kernel(const uint32_t* hist, uint32_t* cursor) {
  const int b = threadIdx.x;
  uint32_t above = 0;
  for (int k = b + 1; k < 64; ++k) above += hist[k];
  cursor[b] = above;
}

__global__ __launch_bounds__(256) void order_scatter_kernel(const uint8_t* key, uint32_t n, uint32_t* cursor,
                                                            uint32_t* perm) {
  __shared__ uint32_t cnt[64], base[64];
  if (threadIdx.x < 64) cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t first = blockIdx.x * (256u * kOrderPerThread);
  int bk[kOrderPerThread];
  uint32_t slot[kOrderPerThread];
  for (int k = 0; k < kOrderPerThread; ++k) {
    const uint32_t p = first + static_cast<uint32_t>(k) * 256u + threadIdx.x;
    bk[k] = p < n ? key[p] : -1;
    slot[k] = wave_bucket_slot(cnt, bk[k]);
  }
  __syncthreads();
  if (threadIdx.x < 64 && cnt[threadIdx.x] != 0) base[threadIdx.x] = atomicAdd(cursor + threadIdx.x, cnt[threadIdx.x]);
  __syncthreads();
  for (int k = 0; k < kOrderPerThread; ++k) {
    const uint32_t p = first + static_cast<uint32_t>(k) * 256u + threadIdx.x;
    if (bk[k] >= 0) perm[base[bk[k]] + slot[k]] = p;
  }
}

constexpr int kInstantBlock = 256;

// Events of lanes of several cameras: per camera, the lanes that push together reserve consecutive
// slots of that camera's buffer with one atomic add (ballot, popcount rank, shuffle of the base);
// slots past the capacity are counted, not written (the host re-runs from the untouched input).
struct RigSink {
  const CameraDev* cam;
  int cam_index;
  __device__ __forceinline__ void push(const RawEventDev& e) {
    unsigned long long mask = __ballot(1);
    const int lane = __lane_id();
    while (mask != 0) {
      const int leader = __ffsll(static_cast<long long>(mask)) - 1;
      const int c0 = __shfl(cam_index, leader);
      const unsigned long long same = __ballot(cam_index == c0) & mask;
      if (cam_index == c0) {
        const unsigned rank = __popcll(same & ((1ull << lane) - 1ull));
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(cam->counters, static_cast<unsigned long long>(__popcll(same)));
        base = __shfl(base, leader);
        const unsigned long long slot = base + rank;
        if (slot < cam->cap) cam->ev[slot] = e;
      }
      mask &= ~same;
    }
  }
};

// Whether phase A found a value that is not finite in some camera's runs: the call is then refused
// and the walk advances no pixel (read once per block, before any lane work).
__device__ __forceinline__ bool call_refused(const CameraDev* cams, int n_cams) {
  bool refused = false;
  for (int c = 0; c < n_cams; ++c) refused |= cams[c].counters[1 + kStNonFinite] != 0;
  return refused;
}

// Adds each lane's v to counters[idx] of its camera: one atomic per camera present in the wave
// (cam_index < 0: no contribution). Called by every lane of the wave.
__device__ __forceinline__ void camera_add(const CameraDev* cams, int cam_index, int idx, uint32_t v) {
  unsigned long long mask = __ballot(cam_index >= 0 && v != 0);
  while (mask != 0) {
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    const int c0 = __shfl(cam_index, leader);
    unsigned long long s = cam_index == c0 ? v : 0u;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (__lane_id() == leader) atomicAdd(cams[c0].counters + idx, s);
    mask &= ~__ballot(cam_index == c0);
  }
}

// One 64-pixel group of the walk: lane i of the launch's pixel order takes pixel order[i] (of its
// camera), walks it to t_end and writes its state. The wave's lanes walk their pixels; a lane that
// reaches an exact search waits until search_batch lanes wait (or none walks), and then all waiting
// lanes search together, so the searches of different pixels run side by side.
template <int kMaxG>
__device__ __forceinline__ void instant_group(const CameraDev* cams, int n_cams, uint32_t n_total,
                                              const uint32_t* order, uint32_t i, const WalkTuning& tu,
                                              int64_t t_end, uint32_t* stacks) {
  InstantLane<kMaxG> L;
  L.n_substeps = L.n_pieces = L.n_leaves = 0;
  int ci = -1;
  const CameraDev* cam = cams;
  int p = 0;
  if (i < n_total) {
    const uint32_t g = order != nullptr ? order[i] : i;
    ci = camera_of(cams, n_cams, g);
    cam = cams + ci;
    p = static_cast<int>(g - cam->base);
  }
  RigSink sink{cam, ci};
  const KConst& kc = cam->kc;
  if (ci >= 0) {
    const MapsDev& mp = cam->mp;
    const StateDev& in = cam->in;
    const RunsDev& runs = cam->runs;
    L.pixel = static_cast<uint32_t>(p);
    L.theta_on = mp.theta_on[p];
    L.theta_off = mp.theta_off[p];
    L.leak = mp.leak[p];
    L.dark = mp.dark[p];
    L.scale = mp.scale[p];
    L.irr_max = kc.irr_max;
    L.eps_ci = kc.eps_ci;
    L.label_cut = kc.label_cut;
    L.refractory_ns = kc.refractory_ns;
    L.G = kc.groups;
    L.ideal_reset = kc.ideal_reset != 0;
    L.hot = mp.hot[p] != 0;
    const int32_t row = runs.row_of != nullptr ? runs.row_of[p] : -1;
    int loud_start = 0;
    if (row >= 0) {
      const int64_t start = runs.row_start[row];
      L.ts = runs.t + start;
      L.irr = runs.irr + start * kc.groups;
      L.n = runs.row_count[row];
      if (cam->first_loud != nullptr) loud_start = vmin(cam->first_loud[row], L.n);
      if (cam->screen_blocks != nullptr) {
        L.qb = cam->screen_blocks;
        L.q_base = cam->pos[row];
      }
    } else {
      L.ts = nullptr;
      L.irr = nullptr;
      L.n = 0;
    }
    L.t_state = cam->t_state;
    L.skip = L.n > 0 && L.ts[0] == L.t_state ? 1 : 0;
    for (int g = 0; g < kMaxG; ++g)
      if (g < kc.groups) L.held[g] = in.held[static_cast<int64_t>(g) * kc.n_pixels + p];
    L.i_scale = in.i_scale[p];
    L.s3_ref = in.s3_ref[p];
    L.n3_ref = in.n3_ref[p];
    L.dr_ref = in.dr_ref[p];
    L.t_ref = in.t_ref[p];
    L.blind_until = in.blind_until[p];
    L.t_search = in.t_search[p];
    L.flags = in.flags[p];
    L.shot_rate = kc.shot_rate_hz;
    L.seed = kc.seed;
    L.camera = kc.camera;
    L.shot_next = in.shot_next[p];
    L.shot_index = in.shot_index[p];
    L.begin(t_end, loud_start);
  }
  int mode = ci >= 0 ? L.walk(sink, tu.walk_budget) : InstantLane<kMaxG>::kDone;
  while (true) {
    const unsigned long long searching = __ballot(mode == InstantLane<kMaxG>::kSearch);
    const unsigned long long walking = __ballot(mode == InstantLane<kMaxG>::kWalk);
    if ((searching | walking) == 0) break;
    if (searching != 0 && (walking == 0 || __popcll(searching) >= tu.search_batch)) {
      if (mode == InstantLane<kMaxG>::kSearch) {
        L.search_step(sink, stacks, blockDim.x);
        mode = InstantLane<kMaxG>::kWalk;
      }
    } else if (mode == InstantLane<kMaxG>::kWalk) {
      mode = L.walk(sink, tu.walk_budget);
    }
  }
  if (ci >= 0) {
    const StateDev& in = cam->in;
    const StateDev& out = cam->out;
    L.end();
    for (int g = 0; g < kMaxG; ++g)
      if (g < kc.groups) out.held[static_cast<int64_t>(g) * kc.n_pixels + p] = L.held[g];
    out.z[p] = in.z[p];
    out.x1[p] = in.x1[p];
    out.s2[p] = in.s2[p];
    out.s3[p] = in.s3[p];
    out.tv[p] = in.tv[p];
    out.i_scale[p] = L.i_scale;
    out.s3_ref[p] = L.s3_ref;
    out.n3_ref[p] = L.n3_ref;
    out.dr_ref[p] = L.dr_ref;
    out.t_ref[p] = L.t_ref;
    out.blind_until[p] = L.blind_until;
    out.t_search[p] = L.t_search;
    out.flags[p] = static_cast<uint16_t>(L.flags);
    out.shot_next[p] = L.shot_next;
    out.shot_index[p] = L.shot_index;
  }
  camera_add(cams, ci, 1 + kStSubsteps, L.n_substeps);
  camera_add(cams, ci, 1 + kStPieces, L.n_pieces);
  camera_add(cams, ci, 1 + kStLeaves, L.n_leaves);
}

// The walk's groups over n pixels of the heavy-first order: tu.head_groups groups of tu.head_lanes
// pixels at the head (the heaviest pixels, spread over more waves so that fewer lanes share each
// wave's turns), then groups of 64.
BSX_HD inline uint32_t walk_group_count(uint32_t n, const WalkTuning& tu) {
  const uint32_t head = static_cast<uint32_t>(tu.head_groups) * static_cast<uint32_t>(tu.head_lanes);
  return static_cast<uint32_t>(tu.head_groups) + (n - head + 63) / 64;
}

// The order position lane `lane` of group `group` takes (past the pixel space for an idle head lane).
__device__ __forceinline__ uint32_t walk_index(uint32_t group, uint32_t lane, const WalkTuning& tu) {
  const uint32_t ng = static_cast<uint32_t>(tu.head_groups), s = static_cast<uint32_t>(tu.head_lanes);
  if (group < ng) return lane < s ? group * s + lane : 0xffffffffu;
  return ng * s + (group - ng) * 64 + lane;
}

// The walk over the pixels of n_cams cameras, in a launch's pixel order. Without a queue (tu.chunk =
// 0) each wave takes one group; with one, waves grab tu.chunk groups at a time from the queue counter
// until the pixel space is exhausted (a persistent grid: every wave exits once a grab is empty).
template <class Group>
__device__ __forceinline__ void walk_groups(uint32_t n_total, unsigned int* queue, const WalkTuning& tu,
                                            const Group& group_at) {
  const uint32_t lane = __lane_id();
  const uint32_t n_groups = walk_group_count(n_total, tu);
  if (tu.chunk <= 0 || queue == nullptr) {
    const uint32_t group = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    if (group < n_groups) group_at(walk_index(group, lane, tu));
    return;
  }
  while (true) {
    uint32_t g0 = 0;
    if (lane == 0) g0 = atomicAdd(queue, static_cast<unsigned int>(tu.chunk));
    g0 = __shfl(g0, 0);
    if (g0 >= n_groups) break;
    const uint32_t g1 = vmin(g0 + static_cast<uint32_t>(tu.chunk), n_groups);
    for (uint32_t group = g0; group < g1; ++group) group_at(walk_index(group, lane, tu));
  }
}

// Three waves per SIMD: the register allocation is held to 168 (a few bytes of scratch per lane
// beyond it), which measured faster than 173 registers at two waves.
template <int kMaxG>
__global__ __launch_bounds__(kInstantBlock) __attribute__((amdgpu_waves_per_eu(3))) void instant_walk_kernel(const CameraDev* cams, int n_cams, uint32_t n_total,
                                                                   const uint32_t* order, unsigned int* queue,
                                                                   WalkTuning tu, const CallDev* call) {
  if (call_refused(cams, n_cams)) return;
  // The lanes' bisection stacks, one column per lane (conflict-free: a level is a row of the block;
  // blockDim.x <= kInstantBlock by the launch bound).
  __shared__ uint32_t stacks[InstantLane<kMaxG>::kSearchDepth * kInstantBlock];
  const int64_t t_end = call->t_end;
  walk_groups(n_total, queue, tu, [&](uint32_t i) {
    instant_group<kMaxG>(cams, n_cams, n_total, order, i, tu, t_end, stacks + threadIdx.x);
  });
}

// Estimated cost of each run's cascade walk for the heavy-first pixel order, over every segment of
// every camera's runs: one unit per segment plus kSearchWeight per threshold's worth of
// log-intensity change (fast logs: a schedule only). Segment j's end sample is sample j of its run,
// so the kernel also makes the call's input check as chain_screen_kernel does.
__global__ __launch_bounds__(256) void chain_cost_kernel(const CameraDev* cams, int n_cams, const CallDev* call) {
  const int64_t g0 = static_cast<int64_t>(blockIdx.x) * blockDim.x;
  if (g0 >= call->n_segments) return;
  const CameraDev& c = cams[camera_at<Space::kSegments>(cams, n_cams, g0)];
  const int64_t q0 = g0 - c.seg_base, q = q0 + threadIdx.x;
  const int64_t i_q = block_run_of_position(c.pos, c.n_runs, q0, q, c.total);
  const KConst& kc = c.kc;
  const MapsDev& mp = c.mp;
  int run = -1;
  float cost = 0.0f;
  InputTally tally;
  if (q < c.total) {
    const int64_t i = i_q, j = q - c.pos[i];
    run = static_cast<int>(i);
    const uint32_t p = c.pixel[i];
    const float* v = c.runs.irr + c.runs.row_start[i] * kc.groups;
    float Ia = mp.dark[p], Ib = mp.dark[p];
    for (int g = 0; g < kc.groups; ++g) {
      const float x = v[j * kc.groups + g];
      tally.add(x, kc.irr_max);
      Ia += mp.scale[p] * (j == 0 ? c.in.held[static_cast<int64_t>(g) * kc.n_pixels + p] : v[(j - 1) * kc.groups + g]);
      Ib += mp.scale[p] * x;
    }
    const float theta = vmin(mp.theta_on[p], mp.theta_off[p]);
    cost = 1.0f + kSearchWeight * fabs_(__logf(Ib) - __logf(Ia)) / theta;
  }
  if (c.work != nullptr) wave_segmented_add(c.work, run, cost);
  tally.flush(c.counters);
}

// A cascade lane's pixel p of camera cam: its constants, its run and its state, then begin(t_end).
template <int kMaxG>
__device__ __forceinline__ void cascade_setup(CascadeLane<kMaxG>& L, const CameraDev* cam, int p, int64_t t_end) {
  const KConst& kc = cam->kc;
  const MapsDev& mp = cam->mp;
  const StateDev& in = cam->in;
  const RunsDev& runs = cam->runs;
  L.pixel = static_cast<uint32_t>(p);
  L.theta_on = mp.theta_on[p];
  L.theta_off = mp.theta_off[p];
  L.leak = mp.leak[p];
  L.dark = mp.dark[p];
  L.scale = mp.scale[p];
  L.irr_max = kc.irr_max;
  L.eps_ci = kc.eps_ci;
  L.label_cut = kc.label_cut;
  L.refractory_ns = kc.refractory_ns;
  L.G = kc.groups;
  L.ideal_reset = kc.ideal_reset != 0;
  L.hot = mp.hot[p] != 0;
  L.cc = cascade::Consts{kc.k, kc.a2, kc.a3, kc.tau_o, kc.tau_sf, kc.eps_accept, kc.eps_drop};
  const int32_t row = runs.row_of != nullptr ? runs.row_of[p] : -1;
  if (row >= 0) {
    const int64_t start = runs.row_start[row];
    L.ts = runs.t + start;
    L.irr = runs.irr + start * kc.groups;
    L.n = runs.row_count[row];
  } else {
    L.ts = nullptr;
    L.irr = nullptr;
    L.n = 0;
  }
  L.t_state = cam->t_state;
  L.skip = L.n > 0 && L.ts[0] == L.t_state ? 1 : 0;
  for (int g = 0; g < kMaxG; ++g)
    if (g < kc.groups) L.held[g] = in.held[static_cast<int64_t>(g) * kc.n_pixels + p];
  L.i_scale = in.i_scale[p];
  L.z = in.z[p];
  L.x1 = in.x1[p];
  L.s2 = in.s2[p];
  L.s3 = in.s3[p];
  L.tv = in.tv[p];
  L.s3_ref = in.s3_ref[p];
  L.n3_ref = in.n3_ref[p];
  L.dr_ref = in.dr_ref[p];
  L.t_ref = in.t_ref[p];
  L.blind_until = in.blind_until[p];
  L.t_search = in.t_search[p];
  L.flags = in.flags[p];
  L.shot_rate = 0.0;
  L.shot_next = INT64_MAX;
  L.shot_index = 0;
  L.begin(t_end);
}

// A finished cascade lane's end() and its pixel's new state.
template <int kMaxG>
__device__ __forceinline__ void cascade_store(CascadeLane<kMaxG>& L, const CameraDev* cam, int p) {
  const KConst& kc = cam->kc;
  const StateDev& in = cam->in;
  const StateDev& out = cam->out;
  L.end();
  for (int g = 0; g < kMaxG; ++g)
    if (g < kc.groups) out.held[static_cast<int64_t>(g) * kc.n_pixels + p] = L.held[g];
  out.z[p] = L.z;
  out.x1[p] = L.x1;
  out.s2[p] = L.s2;
  out.s3[p] = L.s3;
  out.tv[p] = L.tv;
  out.i_scale[p] = L.i_scale;
  out.s3_ref[p] = L.s3_ref;
  out.n3_ref[p] = L.n3_ref;
  out.dr_ref[p] = L.dr_ref;
  out.t_ref[p] = L.t_ref;
  out.blind_until[p] = L.blind_until;
  out.t_search[p] = L.t_search;
  out.flags[p] = static_cast<uint16_t>(L.flags);
  out.shot_next[p] = in.shot_next[p];
  out.shot_index[p] = in.shot_index[p];
}

// The lanes' work counters into their cameras' (every lane of the wave calls it; ci < 0: none).
template <int kMaxG>
__device__ __forceinline__ void cascade_count(const CameraDev* cams, int ci, CascadeLane<kMaxG>& L) {
  camera_add(cams, ci, 1 + kStSubsteps, L.n_substeps);
  camera_add(cams, ci, 1 + kStAttempts, L.n_attempts);
  camera_add(cams, ci, 1 + kStSettled, L.n_settled);
  camera_add(cams, ci, 1 + kStLeaves, L.n_leaves);
  camera_add(cams, ci, 1 + kStFallbacks, L.n_fallbacks);
  camera_add(cams, ci, 1 + kStQuarantines, L.n_quarantines);
  camera_add(cams, ci, 1 + kStPieces, L.n_pieces);
}

// The cascade without noise: lane i of the launch's pixel order takes pixel order[i] of its camera
// and advances it to t_end. Each lane runs its call as a resumable machine and the wave's lanes share
// one loop, one sub-step try per lane per turn, so that they share each try whatever piece or segment
// each is in (a lane's own steps are those of the loops over its segments, pieces and sub-steps, in
// order). (Lanes that take their pixels one at a time from a queue in heavy-first order measured
// slower: 40.7 against 36.4 ms on the first 1280x720 windows.)
template <int kMaxG>
__device__ __forceinline__ void cascade_lane(const CameraDev* cams, int n_cams, uint32_t n_total, const uint32_t* order,
                                             uint32_t i, int64_t t_end) {
  using Lane = CascadeLane<kMaxG>;
  Lane L;
  L.n_substeps = L.n_pieces = L.n_leaves = 0;
  L.n_attempts = L.n_settled = L.n_fallbacks = L.n_quarantines = 0;
  int state = Lane::kFinished;
  int ci = -1, p = 0;
  const CameraDev* cam = cams;
  if (i < n_total) {
    const uint32_t g = order != nullptr ? order[i] : i;
    ci = camera_of(cams, n_cams, g);
    cam = cams + ci;
    p = static_cast<int>(g - cam->base);
  }
  RigSink sink{cam, ci};
  if (ci >= 0) {
    cascade_setup(L, cam, p, t_end);
    state = L.prepare(sink);
  }
  while (__ballot(state == Lane::kAttempt) != 0) {
    if (state == Lane::kAttempt) {
      L.attempt(sink);
      state = L.prepare(sink);
    }
  }
  if (ci >= 0) cascade_store(L, cam, p);
  cascade_count(cams, ci, L);
}

// The cascade's helpers (sub-steps, the comparator's search) are inlined into the lane program: as
// calls they saved and restored the caller's live registers through scratch on every sub-step attempt.
// Two waves per SIMD (256 registers a lane) keep the occupancy the calls had, with a fifth of their
// scratch.
template <int kMaxG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void cascade_walk_kernel(const CameraDev* cams, int n_cams, uint32_t n_total,
                                                           const uint32_t* order, unsigned int* queue, WalkTuning tu,
                                                           const CallDev* call) {
  if (call_refused(cams, n_cams)) return;
  const int64_t t_end = call->t_end;
  walk_groups(n_total, queue, tu, [&](uint32_t i) { cascade_lane<kMaxG>(cams, n_cams, n_total, order, i, t_end); });
}

// The cascade with its in-path noise: lane i of the launch's pixel order takes pixel order[i] of its
// camera and advances it to t_end; mid is the lane's column of the tree walk's path memory (LDS). (Run
// as machines in one loop, as cascade_lane runs its lanes, the noise walk measured no faster: 103 ms
// against 100 ms on the first 1280x720 windows, with 576 bytes of scratch a lane against 432.)
template <int kMaxG>
__device__ __forceinline__ void noise_lane(const CameraDev* cams, int n_cams, uint32_t n_total, const uint32_t* order,
                                           uint32_t i, int64_t t_end, float* mid, int mid_stride) {
  NoiseLane<kMaxG> L;
  L.n_substeps = L.n_pieces = L.n_leaves = 0;
  L.n_attempts = L.n_settled = L.n_fallbacks = L.n_quarantines = 0;
  L.n_cells = L.n_refinements = 0;
  int ci = -1;
  const CameraDev* cam = cams;
  int p = 0;
  if (i < n_total) {
    const uint32_t g = order != nullptr ? order[i] : i;
    ci = camera_of(cams, n_cams, g);
    cam = cams + ci;
    p = static_cast<int>(g - cam->base);
  }
  RigSink sink{cam, ci};
  if (ci >= 0) {
    const KConst& kc = cam->kc;
    const MapsDev& mp = cam->mp;
    const StateDev& in = cam->in;
    const StateDev& out = cam->out;
    const RunsDev& runs = cam->runs;
    const int64_t np = kc.n_pixels;
    L.pixel = static_cast<uint32_t>(p);
    L.theta_on = mp.theta_on[p];
    L.theta_off = mp.theta_off[p];
    L.leak = mp.leak[p];
    L.dark = mp.dark[p];
    L.scale = mp.scale[p];
    L.irr_max = kc.irr_max;
    L.eps_ci = kc.eps_ci;
    L.label_cut = kc.label_cut;
    L.refractory_ns = kc.refractory_ns;
    L.G = kc.groups;
    L.ideal_reset = kc.ideal_reset != 0;
    L.hot = mp.hot[p] != 0;
    L.seed = kc.seed;
    L.camera = kc.camera;
    L.cc = cascade::Consts{kc.k, kc.a2, kc.a3, kc.tau_o, kc.tau_sf, kc.eps_accept, kc.eps_drop};
    L.nd = kc.noise;
    L.margin_z = kc.noise.margin_z;
    L.mid = mid;
    L.mid_stride = mid_stride;
    const int32_t row = runs.row_of != nullptr ? runs.row_of[p] : -1;
    if (row >= 0) {
      const int64_t start = runs.row_start[row];
      L.ts = runs.t + start;
      L.irr = runs.irr + start * kc.groups;
      L.n = runs.row_count[row];
    } else {
      L.ts = nullptr;
      L.irr = nullptr;
      L.n = 0;
    }
    L.t_state = cam->t_state;
    L.skip = L.n > 0 && L.ts[0] == L.t_state ? 1 : 0;
    for (int g = 0; g < kMaxG; ++g)
      if (g < kc.groups) L.held[g] = in.held[static_cast<int64_t>(g) * np + p];
    L.i_scale = in.i_scale[p];
    L.z = in.z[p];
    L.x1 = in.x1[p];
    L.s2 = in.s2[p];
    L.s3 = in.s3[p];
    L.tv = in.tv[p];
    L.s3_ref = in.s3_ref[p];
    L.n3_ref = in.n3_ref[p];
    L.dr_ref = in.dr_ref[p];
    L.t_ref = in.t_ref[p];
    L.blind_until = in.blind_until[p];
    L.t_search = in.t_search[p];
    L.flags = in.flags[p];
    for (int c = 0; c < 3; ++c) {
      L.nk[c] = in.nk[c * np + p];
      L.nk1[c] = in.nk1[c * np + p];
    }
    L.ln_a1 = in.ln_a1[p];
    L.cell = in.cell[p];
    L.have_k1 = in.have_k1[p] != 0;
    L.shot_rate = 0.0;
    L.shot_next = INT64_MAX;
    L.shot_index = 0;
    L.resume();
    L.advance(t_end, sink);
    for (int g = 0; g < kMaxG; ++g)
      if (g < kc.groups) out.held[static_cast<int64_t>(g) * np + p] = L.held[g];
    out.z[p] = L.z;
    out.x1[p] = L.x1;
    out.s2[p] = L.s2;
    out.s3[p] = L.s3;
    out.tv[p] = L.tv;
    out.i_scale[p] = L.i_scale;
    out.s3_ref[p] = L.s3_ref;
    out.n3_ref[p] = L.n3_ref;
    out.dr_ref[p] = L.dr_ref;
    out.t_ref[p] = L.t_ref;
    out.blind_until[p] = L.blind_until;
    out.t_search[p] = L.t_search;
    out.flags[p] = static_cast<uint16_t>(L.flags);
    out.shot_next[p] = in.shot_next[p];
    out.shot_index[p] = in.shot_index[p];
    for (int c = 0; c < 3; ++c) {
      out.nk[c * np + p] = L.nk[c];
      out.nk1[c * np + p] = L.nk1[c];
    }
    out.ln_a1[p] = L.ln_a1;
    out.cell[p] = L.cell;
    out.have_k1[p] = L.have_k1 ? 1 : 0;
  }
  camera_add(cams, ci, 1 + kStSubsteps, L.n_substeps);
  camera_add(cams, ci, 1 + kStAttempts, L.n_attempts);
  camera_add(cams, ci, 1 + kStSettled, L.n_settled);
  camera_add(cams, ci, 1 + kStLeaves, L.n_leaves);
  camera_add(cams, ci, 1 + kStFallbacks, L.n_fallbacks);
  camera_add(cams, ci, 1 + kStQuarantines, L.n_quarantines);
  camera_add(cams, ci, 1 + kStPieces, L.n_pieces);
  camera_add(cams, ci, 1 + kStCells, L.n_cells);
  camera_add(cams, ci, 1 + kStRefinements, L.n_refinements);
}

template <int kMaxG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void noise_walk_kernel(const CameraDev* cams, int n_cams, uint32_t n_total,
                                                         const uint32_t* order, unsigned int* queue, WalkTuning tu,
                                                         const CallDev* call) {
  if (call_refused(cams, n_cams)) return;
  // The lanes' path memory of the tree walk, one column per lane (blockDim.x <= 256 by the launch
  // bound).
  __shared__ float mid[noise::kMidSlots * 256];
  const int64_t t_end = call->t_end;
  walk_groups(n_total, queue, tu, [&](uint32_t i) {
    noise_lane<kMaxG>(cams, n_cams, n_total, order, i, t_end, mid + threadIdx.x, blockDim.x);
  });
}

// row_of[pixel[i]] = i, flagging out-of-range ids (flags[0]), duplicates (flags[1]) and runs outside
// the sample array (flags[2]), so that no kernel reads past the runs it was given.
__global__ void scatter_rows_kernel(const uint32_t* pixel, const int64_t* row_start, const int32_t* row_count,
                                    size_t n, size_t n_samples, int32_t* row_of, uint32_t n_pixels,
                                    unsigned int* flags) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t s = row_start[i];
  const int32_t c = row_count[i];
  if (s < 0 || c < 0 || static_cast<uint64_t>(s) + static_cast<uint64_t>(c) > n_samples) {
    atomicOr(flags + 2, 1u);
    return;
  }
  const uint32_t p = pixel[i];
  if (p >= n_pixels) {
    atomicOr(flags + 0, 1u);
    return;
  }
  const int32_t prev = atomicExch(row_of + p, static_cast<int32_t>(i));
  if (prev >= 0) atomicOr(flags + 1, 1u);
}

// The input each pixel holds after a call on the runs (row_of: its run, or -1): its run's last
// sample, raw, else its held input (reference: end_frame_kernel, sensor_cuda.cu:390-404); out is
// pixels x groups.
__global__ void held_frame_kernel(const float* held, const int32_t* row_of, const int64_t* row_start,
                                  const int32_t* row_count, const float* irr, uint32_t n_pixels, int groups,
                                  float* out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pixels) return;
  const int32_t row = row_of != nullptr ? row_of[p] : -1;
  const bool from_run = row >= 0 && row_count[row] > 0;
  const float* v = from_run ? irr + static_cast<int64_t>(row_start[row] + row_count[row] - 1) * groups : nullptr;
  for (int g = 0; g < groups; ++g)
    out[static_cast<int64_t>(p) * groups + g] = from_run ? v[g] : held[static_cast<int64_t>(g) * n_pixels + p];
}

__global__ void fill_i32_kernel(int32_t* dst, size_t n, int32_t v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

__global__ void fill_u64_kernel(unsigned long long* dst, size_t n, unsigned long long v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

// Input check of device runs, as the reference's check_inputs_kernel (sensor_cuda.cu:415-440): one
// wave per run over the c*G values that run references, counting the non-finite values (out[0]) and
// the values the clamp changes (out[1]). Samples no run references are not read, and a run outside the
// sample array is skipped (scatter_rows_kernel flags it).
__global__ __launch_bounds__(256) void check_runs_kernel(const float* irr, const int64_t* row_start,
                                                         const int32_t* row_count, size_t n_runs, size_t n_samples,
                                                         int groups, float vmax_, unsigned long long* out) {
  const size_t run = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  if (run >= n_runs) return;  // whole waves: the run is uniform across the wave
  const int64_t s = row_start[run];
  const int32_t c = row_count[run];
  if (s < 0 || c < 0 || static_cast<uint64_t>(s) + static_cast<uint64_t>(c) > n_samples) return;
  const float* v = irr + static_cast<size_t>(s) * static_cast<size_t>(groups);
  const size_t n = static_cast<size_t>(c) * static_cast<size_t>(groups);
  uint32_t bad = 0, clamped = 0;
  for (size_t j = __lane_id(); j < n; j += 64) {
    const float x = v[j];
    if (!(x - x == 0.0f)) ++bad;
    else if (clamp_irr(x, vmax_) != x) ++clamped;
  }
  wave_add(out + 0, bad);
  wave_add(out + 1, clamped);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Every positive finite float: monotonicity violations of pm::log and the hash of its bits
// (grid-stride over the bit patterns u = 1 .. 0x7f7fffff).
__global__ __launch_bounds__(256) void log_check_kernel(unsigned long long* out) {
  uint64_t bad = 0, hash = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t u = 1 + blockIdx.x * blockDim.x + threadIdx.x; u < 0x7f800000u; u += stride) {
    const float y = pm::log(__uint_as_float(u));
    hash += splitmix64((static_cast<uint64_t>(u) << 32) | __float_as_uint(y));
    if (u + 1 < 0x7f800000u && pm::log(__uint_as_float(u + 1)) < y) ++bad;
  }
  for (int off = 32; off > 0; off >>= 1) {
    bad += __shfl_xor(bad, off);
    hash += __shfl_xor(hash, off);
  }
  if (__lane_id() == 0) {
    atomicAdd(out + 0, static_cast<unsigned long long>(bad));
    atomicAdd(out + 1, static_cast<unsigned long long>(hash));
  }
}

inline dim3 grid_for(size_t n, int block) { return dim3(static_cast<unsigned>((n + block - 1) / block)); }

}  // namespace

hipError_t launch_init(const KConst& kc, const MapsDev& mp, const StateDev& st, const float* irr0, int64_t t0,
                       unsigned long long* counters, hipStream_t stream) {
  const int block = 256;
  const dim3 grid = grid_for(kc.n_pixels, block);
  if (kc.groups <= 1) init_kernel<1><<<grid, block, 0, stream>>>(kc, mp, st, irr0, t0, counters);
  else if (kc.groups <= 2) init_kernel<2><<<grid, block, 0, stream>>>(kc, mp, st, irr0, t0, counters);
  else if (kc.groups <= 4) init_kernel<4><<<grid, block, 0, stream>>>(kc, mp, st, irr0, t0, counters);
  else init_kernel<8><<<grid, block, 0, stream>>>(kc, mp, st, irr0, t0, counters);
  return hipGetLastError();
}

namespace {

// Blocks of the walk kernel resident on the device at once (a persistent grid's size), per kernel,
// light-group tier and block size.
int resident_blocks(int kind, int tier, int block) {
  static int cache[3][4][4] = {};
  int& r = cache[kind][tier][block / 64 - 1];
  if (r == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    static const void* const kernels[3][4] = {
        {reinterpret_cast<const void*>(cascade_walk_kernel<1>), reinterpret_cast<const void*>(cascade_walk_kernel<2>),
         reinterpret_cast<const void*>(cascade_walk_kernel<4>), reinterpret_cast<const void*>(cascade_walk_kernel<8>)},
        {reinterpret_cast<const void*>(instant_walk_kernel<1>), reinterpret_cast<const void*>(instant_walk_kernel<2>),
         reinterpret_cast<const void*>(instant_walk_kernel<4>), reinterpret_cast<const void*>(instant_walk_kernel<8>)},
        {reinterpret_cast<const void*>(noise_walk_kernel<1>), reinterpret_cast<const void*>(noise_walk_kernel<2>),
         reinterpret_cast<const void*>(noise_walk_kernel<4>), reinterpret_cast<const void*>(noise_walk_kernel<8>)}};
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernels[kind][tier], block, 0) != hipSuccess)
      per_cu = 1;
    r = vmax(1, cus) * vmax(1, per_cu);
  }
  return r;
}

int group_tier(int groups) { return groups <= 1 ? 0 : groups <= 2 ? 1 : groups <= 4 ? 2 : 3; }

hipError_t chain_check(const ChainArgs& a) {
  if (a.n_cams < 1 || a.cams == nullptr || a.call == nullptr) return hipErrorInvalidValue;
  if (a.tu.block < 64 || a.tu.block > kInstantBlock || a.tu.block % 64 != 0) return hipErrorInvalidValue;
  if (a.seg_capacity < 0 || a.run_capacity < 0 || a.groups < 1 || a.groups > kMaxGroups) return hipErrorInvalidValue;
  return hipSuccess;
}

}  // namespace

hipError_t launch_chain_prepare(const ChainArgs& a, hipStream_t stream) {
  const hipError_t e = chain_check(a);
  if (e != hipSuccess) return e;
  const dim3 run_grid = grid_for(static_cast<size_t>(vmax<int64_t>(a.run_capacity, 1)), 256);
  chain_init_kernel<<<run_grid, 256, 0, stream>>>(a.cams, a.n_cams, a.call, a.hist, a.queue);
  if (a.seg_capacity > 0) {
    const dim3 grid = grid_for(static_cast<size_t>(a.seg_capacity), 256);
    const dim3 prep_grid = grid_for(static_cast<size_t>(a.seg_capacity), kSegmentAlign);
    if (!a.instant) chain_cost_kernel<<<grid, 256, 0, stream>>>(a.cams, a.n_cams, a.call);
    else if (a.groups <= 1) chain_screen_kernel<1><<<prep_grid, 256, 0, stream>>>(a.cams, a.n_cams, a.call);
    else if (a.groups <= 2) chain_screen_kernel<2><<<prep_grid, 256, 0, stream>>>(a.cams, a.n_cams, a.call);
    else if (a.groups <= 4) chain_screen_kernel<4><<<prep_grid, 256, 0, stream>>>(a.cams, a.n_cams, a.call);
    else chain_screen_kernel<8><<<prep_grid, 256, 0, stream>>>(a.cams, a.n_cams, a.call);
  }
  if (a.n_total > 0) {
    const dim3 order_grid = grid_for(a.n_total, 256 * kOrderPerThread);
    order_count_kernel<<<order_grid, 256, 0, stream>>>(a.cams, a.n_cams, a.n_total, a.key, a.hist);
    order_offsets_kernel<<<1, 64, 0, stream>>>(a.hist, a.hist + 64);
    order_scatter_kernel<<<order_grid, 256, 0, stream>>>(a.key, a.n_total, a.hist + 64, a.perm);
  }
  return hipGetLastError();
}

int chain_walk_blocks(const ChainArgs& a) {
  const int block = a.tu.block;
  const unsigned full = static_cast<unsigned>((uint64_t{walk_group_count(a.n_total, a.tu)} * 64 + block - 1) / block);
  if (a.tu.chunk <= 0 || a.queue == nullptr) return static_cast<int>(full);
  const int kind = a.instant ? 1 : a.noise ? 2 : 0;
  return static_cast<int>(vmin(full, static_cast<unsigned>(resident_blocks(kind, group_tier(a.groups), block))));
}

hipError_t launch_chain_walk(const ChainArgs& a, hipStream_t stream) {
  const hipError_t e = chain_check(a);
  if (e != hipSuccess) return e;
  if (a.n_total == 0) return hipSuccess;
  const int block = a.tu.block, tier = group_tier(a.groups);
  const dim3 grid(static_cast<unsigned>(chain_walk_blocks(a)));
#define BSX_WALK(K, G) K<G><<<grid, block, 0, stream>>>(a.cams, a.n_cams, a.n_total, a.perm, a.queue, a.tu, a.call)
  if (a.instant) {
    if (tier == 0) BSX_WALK(instant_walk_kernel, 1);
    else if (tier == 1) BSX_WALK(instant_walk_kernel, 2);
    else if (tier == 2) BSX_WALK(instant_walk_kernel, 4);
    else BSX_WALK(instant_walk_kernel, 8);
  } else if (a.noise) {
    if (tier == 0) BSX_WALK(noise_walk_kernel, 1);
    else if (tier == 1) BSX_WALK(noise_walk_kernel, 2);
    else if (tier == 2) BSX_WALK(noise_walk_kernel, 4);
    else BSX_WALK(noise_walk_kernel, 8);
  } else {
    if (tier == 0) BSX_WALK(cascade_walk_kernel, 1);
    else if (tier == 1) BSX_WALK(cascade_walk_kernel, 2);
    else if (tier == 2) BSX_WALK(cascade_walk_kernel, 4);
    else BSX_WALK(cascade_walk_kernel, 8);
  }
#undef BSX_WALK
  return hipGetLastError();
}

hipError_t launch_chain_reset(const ChainArgs& a, hipStream_t stream) {
  const hipError_t e = chain_check(a);
  if (e != hipSuccess) return e;
  chain_reset_kernel<<<1, 256, 0, stream>>>(a.cams, a.n_cams, a.queue);
  return hipGetLastError();
}

hipError_t launch_math_selfcheck(unsigned long long* out, hipStream_t stream) {
  fill_u64_kernel<<<1, 64, 0, stream>>>(out, 2, 0);
  log_check_kernel<<<2048, 256, 0, stream>>>(out);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* dst, size_t n, float v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  fill_i32_kernel<<<grid_for(n, 256), 256, 0, stream>>>(reinterpret_cast<int32_t*>(dst), n, __builtin_bit_cast(int32_t, v));
  return hipGetLastError();
}

hipError_t launch_scatter_rows(const uint32_t* pixel, const int64_t* row_start, const int32_t* row_count, size_t n,
                               size_t n_samples, int32_t* row_of, uint32_t n_pixels, unsigned int* flags,
                               hipStream_t stream) {
  const int block = 256;
  fill_i32_kernel<<<grid_for(n_pixels, block), block, 0, stream>>>(row_of, n_pixels, -1);
  fill_i32_kernel<<<1, 64, 0, stream>>>(reinterpret_cast<int32_t*>(flags), 4, 0);
  if (n > 0)
    scatter_rows_kernel<<<grid_for(n, block), block, 0, stream>>>(pixel, row_start, row_count, n, n_samples, row_of,
                                                                  n_pixels, flags);
  return hipGetLastError();
}

hipError_t launch_held_frame(const float* held, const int32_t* row_of, const int64_t* row_start,
                             const int32_t* row_count, const float* irr, uint32_t n_pixels, int groups, float* out,
                             hipStream_t stream) {
  if (n_pixels == 0) return hipSuccess;
  held_frame_kernel<<<grid_for(n_pixels, 256), 256, 0, stream>>>(held, row_of, row_start, row_count, irr, n_pixels,
                                                                 groups, out);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* dst, size_t n, unsigned long long v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  fill_u64_kernel<<<grid_for(n, 256), 256, 0, stream>>>(dst, n, v);
  return hipGetLastError();
}

hipError_t launch_check_runs(const float* irr, const int64_t* row_start, const int32_t* row_count, size_t n_runs,
                             size_t n_samples, int groups, float irr_max, unsigned long long* out,
                             hipStream_t stream) {
  if (n_runs == 0) return hipSuccess;
  check_runs_kernel<<<grid_for(n_runs * 64, 256), 256, 0, stream>>>(irr, row_start, row_count, n_runs, n_samples,
                                                                    groups, irr_max, out);
  return hipGetLastError();
}

}  // namespace bsx
