// gfx950 kernels of the sensor path: init, the instantaneous-response advance, the input checks and
// the run scatter. One lane per pixel; a 64-lane wave covers 64 consecutive pixels, so every state,
// map and row-index access is one coalesced vector access; events leave through a wave-aggregated
// sink (one atomic per wave and push site, consecutive slots for the active lanes).
#include <hip/hip_runtime.h>

#include "cascade.h"
#include "kernels.h"
#include "instant.h"

namespace bsx {

namespace {

// Wave-aggregated event sink: the lanes that reach a push together reserve consecutive slots with
// one atomic add (ballot over the active lanes, rank by a popcount of the lower lanes), then each
// writes its 32-byte event; slots past the capacity are counted, not written (the host re-runs the
// call from the untouched input state with a larger buffer).
struct WaveSink {
  RawEventDev* ev;
  unsigned long long cap;
  unsigned long long* count;
  __device__ __forceinline__ void push(const RawEventDev& e) {
    const unsigned long long mask = __ballot(1);
    const int lane = __lane_id();
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    const unsigned rank = __popcll(mask & ((1ull << lane) - 1ull));
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(count, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, leader);
    const unsigned long long slot = base + rank;
    if (slot < cap) ev[slot] = e;
  }
};

__device__ __forceinline__ float clamp_irr(float v, float vmax_) { return v >= 0.0f ? (v <= vmax_ ? v : vmax_) : 0.0f; }

// Sum of a per-lane counter over the wave, added once to the global counter.
__device__ __forceinline__ void wave_add(unsigned long long* dst, uint32_t v) {
  unsigned long long s = v;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (__lane_id() == 0 && s != 0) atomicAdd(dst, s);
}

template <int kMaxG>
__global__ __launch_bounds__(256) void init_kernel(KConst kc, MapsDev mp, StateDev st, const float* irr0, int64_t t0,
                                                   unsigned long long* counters) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t clamps = 0;
  if (p < kc.n_pixels) {
    InstantLane<kMaxG> L;
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
    st.t_ref[p] = t0;
    st.s3_ref[p] = s3_ref;
    st.n3_ref[p] = 0.0f;
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

// Minimum of v over the lanes of the wave that share `key` (keys non-decreasing along the lanes),
// folded once per key into dst[key] by the key's last lane (key < 0: no contribution).
__device__ __forceinline__ void wave_segmented_min(int* dst, int key, int v) {
  const int lane = __lane_id();
  for (int off = 1; off < 64; off <<= 1) {
    const int vu = __shfl_up(v, off);
    const int ku = __shfl_up(key, off);
    if (lane >= off && ku == key) v = vu < v ? vu : v;
  }
  const int kn = __shfl_down(key, 1);
  if (key >= 0 && (lane == 63 || kn != key) && v != INT32_MAX) atomicMin(dst + key, v);
}

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

// The same for the block's threads at once (q = the block's first position + threadIdx.x): one
// bisection for the block's first position, then a short forward walk per thread (a block's 256
// consecutive positions span few runs). Called by every thread of the block.
__device__ __forceinline__ int64_t block_run_of_position(const int64_t* pos, int64_t n_runs, int64_t q0,
                                                         int64_t q, int64_t total) {
  __shared__ int64_t first;
  if (threadIdx.x == 0) first = q0 < total ? run_of_position(pos, n_runs, q0) : 0;
  __syncthreads();
  int64_t i = first;
  if (q < total)
    while (i + 1 < n_runs && pos[i + 1] <= q) ++i;
  return i;
}

template <int kMaxG>
__global__ __launch_bounds__(256) void instant_prep_kernel(KConst kc, MapsDev mp, StateDev in, RunsDev runs,
                                                           const uint32_t* pixel, const int64_t* pos, int64_t n_runs,
                                                           int64_t total, int64_t t_state, InstantPrep* out,
                                                           float* work, int* first_loud) {
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * blockDim.x, q = q0 + threadIdx.x;
  const int64_t i_q = block_run_of_position(pos, n_runs, q0, q, total);
  int run = -1;
  float cost = 0.0f;
  int loud = INT32_MAX;
  if (q < total) {
    const int64_t i = i_q, j = q - pos[i];
    run = static_cast<int>(i);
    const uint32_t p = pixel[i];
    const int64_t s = runs.row_start[i];
    InstantLane<kMaxG> L;
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
    L.t_search = in.t_search[p];
    L.shot_next = in.shot_next[p];
    const int64_t* t = runs.t + s;
    const float* v = runs.irr + s * kc.groups;
    bool empty = false;
    if (j == 0) {
      if (t[0] == t_state) {  // the first sample is the model time's: no segment ends at it
        out[q] = InstantPrep{0.0f, 0.0f, 0.0f, 0.0f};
        empty = true;
      } else {
        L.ta = t_state;
        for (int g = 0; g < kMaxG; ++g)
          if (g < kc.groups) L.va[g] = in.held[static_cast<int64_t>(g) * kc.n_pixels + p];
      }
    } else {
      L.ta = t[j - 1];
      for (int g = 0; g < kMaxG; ++g)
        if (g < kc.groups) L.va[g] = L.clampv(v[(j - 1) * kc.groups + g]);
    }
    if (!empty) {
      L.tb = t[j];
      for (int g = 0; g < kMaxG; ++g)
        if (g < kc.groups) L.vb[g] = L.clampv(v[j * kc.groups + g]);
      const InstantPrep r = L.prep_segment();
      out[q] = r;
      const float theta = vmin(L.theta_on, L.theta_off);
      const bool quiet = L.quiet_segment(r);
      loud = quiet ? INT32_MAX : static_cast<int>(j);
      cost = quiet ? 0.0f : 1.0f + kSearchWeight * fabs_(r.x_c - r.x_a) / theta;
    }
  }
  if (work != nullptr) wave_segmented_add(work, run, cost);
  if (first_loud != nullptr) wave_segmented_min(first_loud, run, loud);
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

__global__ __launch_bounds__(256) void order_count_kernel(const CameraDev* cams, int n_cams, uint32_t n, uint8_t* key,
                                                          uint32_t* hist) {
  __shared__ uint32_t cnt[64];
  if (threadIdx.x < 64) cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) {
    const CameraDev& c = cams[camera_of(cams, n_cams, g)];
    const uint32_t p = g - c.base;
    const int32_t r = c.runs.row_of != nullptr ? c.runs.row_of[p] : -1;
    const int b = r >= 0 && c.work != nullptr ? cost_bucket(c.work[r]) : 0;
    key[g] = static_cast<uint8_t>(b);
    atomicAdd(&cnt[b], 1u);
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
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = p < n ? key[p] : 0;
  const uint32_t local = p < n ? atomicAdd(&cnt[b], 1u) : 0u;
  __syncthreads();
  if (threadIdx.x < 64 && cnt[threadIdx.x] != 0) base[threadIdx.x] = atomicAdd(cursor + threadIdx.x, cnt[threadIdx.x]);
  __syncthreads();
  if (p < n) perm[base[b] + local] = p;
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
                                              int64_t t_state, int64_t t_end, uint32_t* stacks) {
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
    const int32_t row = runs.row_of != nullptr ? runs.row_of[p] : -1;
    L.prep = nullptr;
    int loud_start = 0;
    if (row >= 0) {
      const int64_t start = runs.row_start[row];
      L.ts = runs.t + start;
      L.irr = runs.irr + start * kc.groups;
      L.n = runs.row_count[row];
      if (cam->prep != nullptr) L.prep = cam->prep + cam->pos[row];
      if (cam->first_loud != nullptr) loud_start = vmin(cam->first_loud[row], L.n);
    } else {
      L.ts = nullptr;
      L.irr = nullptr;
      L.n = 0;
    }
    L.t_state = t_state;
    L.skip = L.n > 0 && L.ts[0] == t_state ? 1 : 0;
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

// The instantaneous walk over the pixels of n_cams cameras. Without a queue (tu.chunk = 0) each wave
// takes one 64-pixel group; with one, waves grab tu.chunk groups at a time from the queue counter
// until the pixel space is exhausted (a persistent grid: every wave exits once a grab is empty).
template <int kMaxG>
__global__ __launch_bounds__(kInstantBlock) void instant_walk_kernel(const CameraDev* cams, int n_cams, uint32_t n_total,
                                                                   const uint32_t* order, unsigned int* queue,
                                                                   WalkTuning tu, int64_t t_state, int64_t t_end) {
  // The lanes' bisection stacks, one column per lane (conflict-free: a level is a row of the block;
  // blockDim.x <= kInstantBlock by the launch bound).
  __shared__ uint32_t stacks[InstantLane<kMaxG>::kSearchDepth * kInstantBlock];
  const uint32_t lane = __lane_id();
  const uint32_t n_groups = (n_total + 63) / 64;
  if (tu.chunk <= 0 || queue == nullptr) {
    const uint32_t group = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    if (group < n_groups)
      instant_group<kMaxG>(cams, n_cams, n_total, order, group * 64 + lane, tu, t_state, t_end, stacks + threadIdx.x);
    return;
  }
  while (true) {
    uint32_t g0 = 0;
    if (lane == 0) g0 = atomicAdd(queue, static_cast<unsigned int>(tu.chunk));
    g0 = __shfl(g0, 0);
    if (g0 >= n_groups) break;
    const uint32_t g1 = vmin(g0 + static_cast<uint32_t>(tu.chunk), n_groups);
    for (uint32_t group = g0; group < g1; ++group)
      instant_group<kMaxG>(cams, n_cams, n_total, order, group * 64 + lane, tu, t_state, t_end, stacks + threadIdx.x);
  }
}

// Estimated cost of each run's cascade walk for the heavy-first pixel order: one unit per segment
// plus kSearchWeight per threshold's worth of log-intensity change (fast logs: a schedule only).
__global__ __launch_bounds__(256) void run_cost_kernel(KConst kc, MapsDev mp, StateDev in, RunsDev runs,
                                                       const uint32_t* pixel, const int64_t* pos, int64_t n_runs,
                                                       int64_t total, float* work) {
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * blockDim.x, q = q0 + threadIdx.x;
  const int64_t i_q = block_run_of_position(pos, n_runs, q0, q, total);
  int run = -1;
  float cost = 0.0f;
  if (q < total) {
    const int64_t i = i_q, j = q - pos[i];
    run = static_cast<int>(i);
    const uint32_t p = pixel[i];
    const float* v = runs.irr + runs.row_start[i] * kc.groups;
    float Ia = mp.dark[p], Ib = mp.dark[p];
    for (int g = 0; g < kc.groups; ++g) {
      Ia += mp.scale[p] * (j == 0 ? in.held[static_cast<int64_t>(g) * kc.n_pixels + p] : v[(j - 1) * kc.groups + g]);
      Ib += mp.scale[p] * v[j * kc.groups + g];
    }
    const float theta = vmin(mp.theta_on[p], mp.theta_off[p]);
    cost = 1.0f + kSearchWeight * fabs_(__logf(Ib) - __logf(Ia)) / theta;
  }
  wave_segmented_add(work, run, cost);
}

// The cascade without noise: one lane per pixel, in the heavy-first order (null: pixel order).
template <int kMaxG>
__global__ __launch_bounds__(256) void cascade_kernel(KConst kc, MapsDev mp, StateDev in, StateDev out, RunsDev runs,
                                                      int64_t t_state, int64_t t_end, RawEventDev* ev,
                                                      unsigned long long cap, unsigned long long* counters,
                                                      const uint32_t* order) {
  const int lane_index = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = order != nullptr && lane_index < kc.n_pixels ? static_cast<int>(order[lane_index]) : lane_index;
  CascadeLane<kMaxG> L;
  L.n_substeps = L.n_pieces = L.n_leaves = 0;
  L.n_attempts = L.n_settled = L.n_fallbacks = L.n_quarantines = 0;
  WaveSink sink{ev, cap, counters};
  if (p < kc.n_pixels) {
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
    L.cc = cascade::Consts{kc.k, kc.a2, kc.a3, kc.tau_o, kc.tau_sf, kc.eps_accept, kc.eps_drop};
    const int32_t row = runs.row_of != nullptr ? runs.row_of[p] : -1;
    L.prep = nullptr;
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
    L.t_state = t_state;
    L.skip = L.n > 0 && L.ts[0] == t_state ? 1 : 0;
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
    L.advance(t_end, sink);
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
  wave_add(counters + 1 + kStSubsteps, L.n_substeps);
  wave_add(counters + 1 + kStAttempts, L.n_attempts);
  wave_add(counters + 1 + kStSettled, L.n_settled);
  wave_add(counters + 1 + kStLeaves, L.n_leaves);
  wave_add(counters + 1 + kStFallbacks, L.n_fallbacks);
  wave_add(counters + 1 + kStQuarantines, L.n_quarantines);
  wave_add(counters + 1 + kStPieces, L.n_pieces);
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

__global__ void fill_i32_kernel(int32_t* dst, size_t n, int32_t v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

__global__ void fill_u64_kernel(unsigned long long* dst, size_t n, unsigned long long v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

// Input check of device runs: non-finite samples (out[0]) and clamped samples (out[1]); one lane
// per sample, coalesced, wave-reduced.
__global__ void check_samples_kernel(const float* irr, size_t count, float vmax_, unsigned long long* out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t bad = 0, clamped = 0;
  if (i < count) {
    const float v = irr[i];
    if (!(v - v == 0.0f)) bad = 1;
    else if (clamp_irr(v, vmax_) != v) clamped = 1;
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

hipError_t launch_instant_prep(const KConst& kc, const MapsDev& mp, const StateDev& in, const RunsDev& runs,
                               const uint32_t* pixel, const int64_t* pos, int64_t n_runs, int64_t total,
                               int64_t t_state, InstantPrep* out, float* work, int* first_loud, hipStream_t stream) {
  if (n_runs <= 0) return hipSuccess;
  fill_i32_kernel<<<grid_for(static_cast<size_t>(n_runs), 256), 256, 0, stream>>>(first_loud, n_runs, INT32_MAX);
  if (total <= 0) return hipGetLastError();
  const dim3 grid = grid_for(static_cast<size_t>(total), 256);
#define BSX_PREP(G)                                                                                         \
  instant_prep_kernel<G><<<grid, 256, 0, stream>>>(kc, mp, in, runs, pixel, pos, n_runs, total, t_state, out, work, \
                                                   first_loud)
  if (kc.groups <= 1) BSX_PREP(1);
  else if (kc.groups <= 2) BSX_PREP(2);
  else if (kc.groups <= 4) BSX_PREP(4);
  else BSX_PREP(8);
#undef BSX_PREP
  return hipGetLastError();
}

hipError_t launch_instant_walk(const CameraDev* cams, int n_cams, uint32_t n_total, int groups, const uint32_t* order,
                               unsigned int* queue, const WalkTuning& tu, int64_t t_state, int64_t t_end,
                               hipStream_t stream) {
  if (n_total == 0) return hipSuccess;
  const int block = tu.block;
  dim3 grid = grid_for(n_total, block);
  if (tu.chunk > 0 && queue != nullptr) {
    // persistent: as many blocks as are resident at once (at most one per 64-pixel group set)
    static int resident[4] = {0, 0, 0, 0};
    const int gi = groups <= 1 ? 0 : groups <= 2 ? 1 : groups <= 4 ? 2 : 3;
    if (resident[gi] == 0) {
      int dev = 0, cus = 0, per_cu = 0;
      (void)hipGetDevice(&dev);
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      hipError_t e = hipSuccess;
      if (gi == 0) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, instant_walk_kernel<1>, block, 0);
      else if (gi == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, instant_walk_kernel<2>, block, 0);
      else if (gi == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, instant_walk_kernel<4>, block, 0);
      else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, instant_walk_kernel<8>, block, 0);
      if (e != hipSuccess) return e;
      resident[gi] = vmax(1, cus * vmax(1, per_cu) * 256 / block);
    }
    grid = dim3(vmin(grid.x, static_cast<unsigned>(resident[gi])));
    fill_i32_kernel<<<1, 64, 0, stream>>>(reinterpret_cast<int32_t*>(queue), 1, 0);
  }
#define BSX_WALK(G) \
  instant_walk_kernel<G><<<grid, block, 0, stream>>>(cams, n_cams, n_total, order, queue, tu, t_state, t_end)
  if (groups <= 1) BSX_WALK(1);
  else if (groups <= 2) BSX_WALK(2);
  else if (groups <= 4) BSX_WALK(4);
  else BSX_WALK(8);
#undef BSX_WALK
  return hipGetLastError();
}

hipError_t launch_run_cost(const KConst& kc, const MapsDev& mp, const StateDev& in, const RunsDev& runs,
                           const uint32_t* pixel, const int64_t* pos, int64_t n_runs, int64_t total, float* work,
                           hipStream_t stream) {
  if (n_runs <= 0) return hipSuccess;
  fill_i32_kernel<<<grid_for(static_cast<size_t>(n_runs), 256), 256, 0, stream>>>(reinterpret_cast<int32_t*>(work),
                                                                                  n_runs, 0);
  if (total > 0)
    run_cost_kernel<<<grid_for(static_cast<size_t>(total), 256), 256, 0, stream>>>(kc, mp, in, runs, pixel, pos,
                                                                                   n_runs, total, work);
  return hipGetLastError();
}

hipError_t launch_cascade(const KConst& kc, const MapsDev& mp, const StateDev& in, const StateDev& out,
                          const RunsDev& runs, int64_t t_state, int64_t t_end, RawEventDev* ev, unsigned long long cap,
                          unsigned long long* counters, const uint32_t* order, int block, hipStream_t stream) {
  const dim3 grid = grid_for(kc.n_pixels, block);
#define BSX_CAS(G) \
  cascade_kernel<G><<<grid, block, 0, stream>>>(kc, mp, in, out, runs, t_state, t_end, ev, cap, counters, order)
  if (kc.groups <= 1) BSX_CAS(1);
  else if (kc.groups <= 2) BSX_CAS(2);
  else if (kc.groups <= 4) BSX_CAS(4);
  else BSX_CAS(8);
#undef BSX_CAS
  return hipGetLastError();
}

hipError_t launch_cost_order(const CameraDev* cams, int n_cams, uint32_t n_total, uint8_t* key, uint32_t* hist,
                             uint32_t* perm, hipStream_t stream) {
  if (n_total == 0) return hipSuccess;
  const int block = 256;
  fill_i32_kernel<<<1, 128, 0, stream>>>(reinterpret_cast<int32_t*>(hist), 128, 0);
  order_count_kernel<<<grid_for(n_total, block), block, 0, stream>>>(cams, n_cams, n_total, key, hist);
  order_offsets_kernel<<<1, 64, 0, stream>>>(hist, hist + 64);
  order_scatter_kernel<<<grid_for(n_total, block), block, 0, stream>>>(key, n_total, hist + 64, perm);
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

hipError_t launch_fill_u64(unsigned long long* dst, size_t n, unsigned long long v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  fill_u64_kernel<<<grid_for(n, 256), 256, 0, stream>>>(dst, n, v);
  return hipGetLastError();
}

hipError_t launch_check_samples(const float* irr, size_t count, float irr_max, unsigned long long* out,
                                hipStream_t stream) {
  if (count == 0) return hipSuccess;
  check_samples_kernel<<<grid_for(count, 256), 256, 0, stream>>>(irr, count, irr_max, out);
  return hipGetLastError();
}

}  // namespace bsx
