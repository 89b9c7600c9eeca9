// Kernels of the procedural scene and the sample-time controller (scene_dev.h): a frame render, the
// two passes of a window's sampling (count, then write at the scanned offsets) and an exclusive scan.
#include <hip/hip_runtime.h>

#include "scene_dev.h"
#include "scene_kernels.h"

namespace bsx {

namespace {

struct RulesDev {
  double d_px;
  float tol, dark;
  int32_t l_min, l_max, max_levels;
  uint32_t min_split;
  int32_t signature;
  int64_t fixed_period;
};

__global__ void render_kernel(SceneDev sc, ViewDev view, double t, float* out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= view.width * view.height) return;
  const double x = view.ox + (static_cast<double>(p % view.width) + 0.5) * view.scale;
  const double y = view.oy + (static_cast<double>(p / view.width) + 0.5) * view.scale;
  uint64_t mask[kMaskWords];
  for (int w = 0; w < kMaskWords; ++w) mask[w] = 0;
  for (int i = 0; i < sc.n; ++i) mask[i >> 6] |= 1ull << (i & 63);
  out[p] = eval_scene(sc, mask, x, y, t).e;
}

// What one pass of sample_pixel does with each sample: count it, write it into the pixel's run (t_out,
// irr_out), or capture the first `cap` of them as (dt, value) pairs into the pixel's scratch slots.
enum class Emit { kCount, kWrite, kCapture };

// One pixel's samples over (t_w, t_w + W] under mode M; returns the pixel's sample count.
template <Emit M>
__device__ uint32_t sample_pixel(const SceneDev& sc, const ViewDev& view, const RulesDev& r, int64_t t_w, uint32_t p,
                                 int64_t* t_out, float* irr_out, CapturedSample* cap_out = nullptr, uint32_t cap = 0) {
  const double x = view.ox + (static_cast<double>(p % view.width) + 0.5) * view.scale;
  const double y = view.oy + (static_cast<double>(p / view.width) + 0.5) * view.scale;
  const double tw_s = static_cast<double>(t_w) * 1e-9;
  uint64_t mask[kMaskWords];
  const float v_hat = plan_pixel(sc, x, y, tw_s, tw_s + kWindowNs * 1e-9, mask);
  uint32_t count = 0;
  const auto emit = [&](uint32_t dt, float e) {
    if constexpr (M == Emit::kWrite) {
      t_out[count] = t_w + dt;
      irr_out[count] = e;
    } else if constexpr (M == Emit::kCapture) {
      if (count < cap) cap_out[count] = CapturedSample{dt, e};
    }
    ++count;
  };
  // A sample's time in seconds as the engine's render hook forms it: (t_w + dt) 1e-9.
  const auto at = [&](uint32_t dt) { return eval_scene(sc, mask, x, y, static_cast<double>(t_w + dt) * 1e-9); };
  if (r.fixed_period > 0) {
    for (int64_t d = r.fixed_period; d <= kWindowNs; d += r.fixed_period) emit(static_cast<uint32_t>(d), at(d).e);
    return count;
  }
  if (v_hat <= 0.0f) return 0;  // static: no shape reaches the pixel this window
  // Level: clamp(ceil(log2(v_hat W / d_px)), l_min, l_max), v_hat in camera px/s.
  const double lv = ceil(log2(static_cast<double>(v_hat) / view.scale * (kWindowNs * 1e-9) / r.d_px));
  const int level = lv != lv ? r.l_min : static_cast<int>(fmin(fmax(lv, static_cast<double>(r.l_min)),
                                                               static_cast<double>(r.l_max)));
  const uint32_t n = 1u << level, step = 1u << (kWindowLog2 - level);
  const bool mid = r.tol > 0.0f;
  // Sliding lattice window: slots j-2, j-1, j, j+1.
  Sample sm2{0.0f, 0u}, sm1 = at(0), s0 = at(step), sp1{0.0f, 0u};
  if (n >= 2) sp1 = at(2 * step);
  struct Node {
    uint32_t lo, hi;
    float e_lo, e_hi;
    uint32_t s_lo, s_hi;
    int32_t depth;  // 0: emit hi's... unused; > 0: refine at this depth; -1: emit (lo, e_lo)
  };
  for (uint32_t j = 1; j <= n; ++j) {
    const uint32_t lo = (j - 1) * step, hi = j * step;
    const bool s = r.signature && sm1.sig != s0.sig;
    bool m = false;
    if (mid) {
      m = (j + 1 <= n && midpoint_eta(r.dark, sm1.e, s0.e, sp1.e) > r.tol) ||
          (j >= 2 && midpoint_eta(r.dark, sm2.e, sm1.e, s0.e) > r.tol);
    }
    if ((s || m) && step >= r.min_split && r.max_levels > 0) {
      Node stack[16];
      int sp = 0;
      stack[sp++] = Node{lo, hi, sm1.e, s0.e, sm1.sig, s0.sig, 1};
      while (sp > 0) {
        const Node nd = stack[--sp];
        if (nd.depth < 0) {
          emit(nd.lo, nd.e_lo);
          continue;
        }
        const uint32_t g = (nd.lo + nd.hi) >> 1;
        const Sample sg = at(g);
        const bool s1 = r.signature && nd.s_lo != sg.sig, s2 = r.signature && sg.sig != nd.s_hi;
        const bool bad = mid && midpoint_eta(r.dark, nd.e_lo, sg.e, nd.e_hi) > r.tol;
        const bool p1 = (s1 || bad) && g - nd.lo >= r.min_split;
        const bool p2 = (s2 || bad) && nd.hi - g >= r.min_split;
        const bool deeper = nd.depth < r.max_levels;
        if (p2 && deeper) stack[sp++] = Node{g, nd.hi, sg.e, nd.e_hi, sg.sig, nd.s_hi, nd.depth + 1};
        stack[sp++] = Node{g, 0u, sg.e, 0.0f, 0u, 0u, -1};
        if (p1 && deeper) stack[sp++] = Node{nd.lo, g, nd.e_lo, sg.e, nd.s_lo, sg.sig, nd.depth + 1};
      }
    }
    emit(hi, s0.e);
    sm2 = sm1;
    sm1 = s0;
    s0 = sp1;
    if (j + 2 <= n) sp1 = at((j + 2) * step);
  }
  return count;
}

__global__ __launch_bounds__(256) void count_kernel(SceneDev sc, ViewDev view, RulesDev r, int64_t t_w,
                                                    int32_t* counts) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= view.width * view.height) return;
  counts[p] = static_cast<int32_t>(sample_pixel<Emit::kCount>(sc, view, r, t_w, p, nullptr, nullptr));
}

// sample_pixel's adaptive controller as a resumable machine, so that a wave's lanes share every
// scene evaluation whatever lattice slot or refinement node each is at: a lane states the offset it
// needs evaluated next (want), the wave evaluates all of them at once, and feed(value) runs the lane's
// controller on until it needs the next one. A lane makes the same evaluations and emits the same
// samples in the same order as sample_pixel.
struct PixelSampler {
  struct Node {
    uint32_t lo, hi;
    float e_lo, e_hi;
    uint32_t s_lo, s_hi;
    int32_t depth;  // > 0: refine at this depth; -1: emit (lo, e_lo)
  };
  enum : int { kCarried, kFirst, kSecond, kNode, kSlide, kDone };
  RulesDev r;
  uint32_t n, step, j;
  int stage;
  bool mid;
  Sample sm2, sm1, s0, sp1;
  Node* stack;  // the lane's refinement stack (16 nodes), apart from the rest so that the rest stays in registers
  int sp;
  Node cur;          // the node whose midpoint is being evaluated
  uint32_t want;     // the offset to evaluate next (stage != kDone)
  CapturedSample* out;
  uint32_t cap, count;

  __device__ __forceinline__ void emit(uint32_t dt, float e) {
    if (count < cap) out[count] = CapturedSample{dt, e};
    ++count;
  }

  // The flags of lattice interval j, then its refinement (or, with none, its end sample).
  __device__ __forceinline__ void interval() {
    const uint32_t lo = (j - 1) * step, hi = j * step;
    const bool s = r.signature && sm1.sig != s0.sig;
    bool m = false;
    if (mid) {
      m = (j + 1 <= n && midpoint_eta(r.dark, sm1.e, s0.e, sp1.e) > r.tol) ||
          (j >= 2 && midpoint_eta(r.dark, sm2.e, sm1.e, s0.e) > r.tol);
    }
    sp = 0;
    if ((s || m) && step >= r.min_split && r.max_levels > 0) stack[sp++] = Node{lo, hi, sm1.e, s0.e, sm1.sig, s0.sig, 1};
    pop();
  }

  // Pops the refinement stack until a midpoint needs evaluating; with the stack empty, emits the
  // interval's end sample and slides the lattice window.
  __device__ __forceinline__ void pop() {
    while (sp > 0) {
      const Node nd = stack[--sp];
      if (nd.depth < 0) {
        emit(nd.lo, nd.e_lo);
        continue;
      }
      cur = nd;
      want = (nd.lo + nd.hi) >> 1;
      stage = kNode;
      return;
    }
    emit(j * step, s0.e);
    sm2 = sm1;
    sm1 = s0;
    s0 = sp1;
    if (j + 2 <= n) {
      want = (j + 2) * step;
      stage = kSlide;
      return;
    }
    next();
  }

  __device__ __forceinline__ void next() {
    if (++j > n) {
      stage = kDone;
      return;
    }
    interval();
  }

  __device__ __forceinline__ void feed(const Sample& v) {
    switch (stage) {
      case kCarried:
        sm1 = v;
        want = step;
        stage = kFirst;
        return;
      case kFirst:
        s0 = v;
        if (n >= 2) {
          want = 2 * step;
          stage = kSecond;
          return;
        }
        j = 1;
        interval();
        return;
      case kSecond:
        sp1 = v;
        j = 1;
        interval();
        return;
      case kNode: {
        const Node nd = cur;
        const uint32_t g = want;
        const bool s1 = r.signature && nd.s_lo != v.sig, s2 = r.signature && v.sig != nd.s_hi;
        const bool bad = mid && midpoint_eta(r.dark, nd.e_lo, v.e, nd.e_hi) > r.tol;
        const bool p1 = (s1 || bad) && g - nd.lo >= r.min_split;
        const bool p2 = (s2 || bad) && nd.hi - g >= r.min_split;
        const bool deeper = nd.depth < r.max_levels;
        if (p2 && deeper) stack[sp++] = Node{g, nd.hi, v.e, nd.e_hi, v.sig, nd.s_hi, nd.depth + 1};
        stack[sp++] = Node{g, 0u, v.e, 0.0f, 0u, 0u, -1};
        if (p1 && deeper) stack[sp++] = Node{nd.lo, g, nd.e_lo, v.e, nd.s_lo, v.sig, nd.depth + 1};
        pop();
        return;
      }
      case kSlide:
        sp1 = v;
        next();
        return;
      default:
        return;
    }
  }
};

// The one-pass form: every pixel's count, its first `cap` samples captured into scratch slots
// [p cap, (p + 1) cap), and the window's largest count (atomicMax of the waves' maxima) in *max_count.
// The adaptive controller runs as PixelSampler machines, one scene evaluation per lane per turn.
__global__ __launch_bounds__(256) void capture_kernel(SceneDev sc, ViewDev view, RulesDev r, int64_t t_w, uint32_t cap,
                                                      int32_t* counts, CapturedSample* scratch,
                                                      unsigned int* max_count) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t c = 0;
  if (r.fixed_period > 0) {
    if (p < view.width * view.height) {
      c = sample_pixel<Emit::kCapture>(sc, view, r, t_w, p, nullptr, nullptr,
                                       scratch + static_cast<size_t>(p) * cap, cap);
      counts[p] = static_cast<int32_t>(c);
    }
  } else {
    const bool live = p < view.width * view.height;
    const double x = live ? view.ox + (static_cast<double>(p % view.width) + 0.5) * view.scale : 0.0;
    const double y = live ? view.oy + (static_cast<double>(p / view.width) + 0.5) * view.scale : 0.0;
    const double tw_s = static_cast<double>(t_w) * 1e-9;
    uint64_t mask[kMaskWords];
    for (int w = 0; w < kMaskWords; ++w) mask[w] = 0;
    PixelSampler::Node stack[16];
    PixelSampler m;
    m.stack = stack;
    m.r = r;
    m.mid = r.tol > 0.0f;
    m.out = scratch + static_cast<size_t>(live ? p : 0) * cap;
    m.cap = cap;
    m.count = 0;
    m.stage = PixelSampler::kDone;
    if (live) {
      const float v_hat = plan_pixel(sc, x, y, tw_s, tw_s + kWindowNs * 1e-9, mask);
      if (v_hat > 0.0f) {  // a pixel no shape reaches is static: no samples
        const double lv = ceil(log2(static_cast<double>(v_hat) / view.scale * (kWindowNs * 1e-9) / r.d_px));
        const int level = lv != lv ? r.l_min : static_cast<int>(fmin(fmax(lv, static_cast<double>(r.l_min)),
                                                                     static_cast<double>(r.l_max)));
        m.n = 1u << level;
        m.step = 1u << (kWindowLog2 - level);
        m.sm2 = Sample{0.0f, 0u};
        m.sp1 = Sample{0.0f, 0u};
        m.want = 0;
        m.stage = PixelSampler::kCarried;
      }
    }
    while (__ballot(m.stage != PixelSampler::kDone) != 0) {
      if (m.stage != PixelSampler::kDone)
        m.feed(eval_scene(sc, mask, x, y, static_cast<double>(t_w + m.want) * 1e-9));
    }
    if (live) {
      c = m.count;
      counts[p] = static_cast<int32_t>(c);
    }
  }
  for (int off = 32; off > 0; off >>= 1) c = max(c, static_cast<uint32_t>(__shfl_xor(static_cast<int>(c), off)));
  if ((threadIdx.x & 63) == 0 && c > 0) atomicMax(max_count, c);
}

// The captured runs moved to their place in the window's runs: one wave per pixel, its samples copied
// 64 at a time. A pixel with more samples than slots is left to overflow_write_kernel.
__global__ __launch_bounds__(256) void compact_kernel(uint32_t n_pixels, int64_t t_w, uint32_t cap,
                                                      const int32_t* counts, const int64_t* row_start,
                                                      const CapturedSample* scratch, int64_t* t, float* irr) {
  const uint32_t p = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= n_pixels) return;
  const uint32_t c = static_cast<uint32_t>(counts[p]);
  if (c > cap) return;
  const CapturedSample* src = scratch + static_cast<size_t>(p) * cap;
  const int64_t base = row_start[p];
  for (uint32_t k = threadIdx.x & 63; k < c; k += 64) {
    const CapturedSample v = src[k];
    t[base + k] = t_w + v.dt;
    irr[base + k] = v.e;
  }
}

__global__ __launch_bounds__(256) void write_kernel(SceneDev sc, ViewDev view, RulesDev r, int64_t t_w,
                                                    const int64_t* row_start, int64_t* t, float* irr) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= view.width * view.height) return;
  const int64_t s = row_start[p];
  sample_pixel<Emit::kWrite>(sc, view, r, t_w, p, t + s, irr + s);
}

// write_kernel for the pixels whose count exceeds the capture slots (the others return at once).
__global__ __launch_bounds__(256) void overflow_write_kernel(SceneDev sc, ViewDev view, RulesDev r, int64_t t_w,
                                                             uint32_t cap, const int32_t* counts,
                                                             const int64_t* row_start, int64_t* t, float* irr) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= view.width * view.height || static_cast<uint32_t>(counts[p]) <= cap) return;
  const int64_t s = row_start[p];
  sample_pixel<Emit::kWrite>(sc, view, r, t_w, p, t + s, irr + s);
}

// Samples at imposed per-pixel counts: pixel p's n = counts[p] samples at t_w + floor((k + 1) W / n),
// k = 0 .. n - 1 (evenly spread, the last at the window's end), one thread per sample (its pixel by
// bisection of the scanned offsets), every shape of the scene evaluated.
__global__ __launch_bounds__(256) void counts_write_kernel(SceneDev sc, ViewDev view, int64_t t_w, const int32_t* counts,
                                                           const int64_t* row_start, uint32_t n_pixels, int64_t total,
                                                           int64_t* t, float* irr) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= total) return;
  uint32_t lo = 0, hi = n_pixels;  // the last pixel whose run starts at or before s
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (row_start[mid] <= s) lo = mid;
    else hi = mid;
  }
  const uint32_t p = lo;
  const int64_t k = s - row_start[p], n = counts[p];
  const int64_t dt = (k + 1) * kWindowNs / n;
  const double x = view.ox + (static_cast<double>(p % view.width) + 0.5) * view.scale;
  const double y = view.oy + (static_cast<double>(p / view.width) + 0.5) * view.scale;
  uint64_t mask[kMaskWords];
  for (int w = 0; w < kMaskWords; ++w) mask[w] = 0;
  for (int i = 0; i < sc.n; ++i) mask[i >> 6] |= 1ull << (i & 63);
  t[s] = t_w + dt;
  irr[s] = eval_scene(sc, mask, x, y, static_cast<double>(t_w + dt) * 1e-9).e;
}

// The requests of the engine's RenderHook: request i (pixel, time t_base + t_off_ns) writes the scene's
// irradiance at the pixel centre into row out_index[i] of irr (group 0; further groups 0), its
// signature into sig[out_index[i]] and, when klass is not null, the class word (textured shape or
// background) into klass[out_index[i]].
__global__ __launch_bounds__(256) void requests_kernel(SceneDev sc, ViewDev view, int64_t t_base,
                                                       const SampleRequestDev* req, size_t n, const uint32_t* out_index,
                                                       int32_t groups, float* irr, uint64_t* sig, uint32_t* klass) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SampleRequestDev r = req[i];
  const uint32_t p = r.pixel;
  const double x = view.ox + (static_cast<double>(p % view.width) + 0.5) * view.scale;
  const double y = view.oy + (static_cast<double>(p / view.width) + 0.5) * view.scale;
  uint64_t mask[kMaskWords];
  for (int w = 0; w < kMaskWords; ++w) mask[w] = 0;
  for (int k = 0; k < sc.n; ++k) mask[k >> 6] |= 1ull << (k & 63);
  const int64_t t = t_base + r.t_off_ns;
  const Sample s = eval_scene(sc, mask, x, y, static_cast<double>(t) * 1e-9);
  const uint32_t row = out_index != nullptr ? out_index[i] : static_cast<uint32_t>(i);
  for (int g = 0; g < groups; ++g) irr[static_cast<int64_t>(row) * groups + g] = g == 0 ? s.e : 0.0f;
  sig[row] = s.sig;
  if (klass != nullptr) klass[row] = s.sig != 0 ? kClassTexturedWord : 0u;
}

// Exclusive scan of int32 counts into int64 offsets: per-block scans (wave shuffles, then the
// block's wave totals through LDS), the block totals scanned by one block, then added back.
constexpr int kScanBlock = 1024;

__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* total) {
  __shared__ int64_t wave_sums[kScanBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t inc = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t o = __shfl_up(inc, off);
    if (lane >= off) inc += o;
  }
  if (lane == 63) wave_sums[wave] = inc;
  __syncthreads();
  if (wave == 0) {
    int64_t w = lane < kScanBlock / 64 ? wave_sums[lane] : 0;
    for (int off = 1; off < 64; off <<= 1) {
      const int64_t o = __shfl_up(w, off);
      if (lane >= off) w += o;
    }
    if (lane < kScanBlock / 64) wave_sums[lane] = w;
  }
  __syncthreads();
  const int64_t before = wave == 0 ? 0 : wave_sums[wave - 1];
  *total = wave_sums[kScanBlock / 64 - 1];
  __syncthreads();
  return before + inc - v;
}

__global__ __launch_bounds__(kScanBlock) void scan_blocks_kernel(const int32_t* counts, size_t n, int64_t* out,
                                                                 int64_t* block_sums) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kScanBlock + threadIdx.x;
  const int64_t v = i < n ? counts[i] : 0;
  int64_t total;
  const int64_t ex = block_exclusive_scan(v, &total);
  if (i < n) out[i] = ex;
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void scan_sums_kernel(int64_t* block_sums, size_t n_blocks, int64_t* total) {
  // One block scans the block totals in chunks of kScanBlock, carrying the running sum.
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (size_t base = 0; base < n_blocks; base += kScanBlock) {
    const size_t i = base + threadIdx.x;
    const int64_t v = i < n_blocks ? block_sums[i] : 0;
    int64_t chunk;
    const int64_t ex = block_exclusive_scan(v, &chunk);
    const int64_t c = carry;
    if (i < n_blocks) block_sums[i] = c + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry = c + chunk;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void scan_add_kernel(int64_t* out, size_t n, const int64_t* block_sums) {
  const size_t i = static_cast<size_t>(blockIdx.x) * kScanBlock + threadIdx.x;
  if (i < n) out[i] += block_sums[blockIdx.x];
}

__global__ void fill_u32_kernel(unsigned int* dst, size_t n, unsigned int v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

RulesDev rules_dev(const bsx_sampler_rules& r) {
  RulesDev d{};
  d.d_px = r.d_px;
  d.tol = r.refine_tol_log;
  d.dark = r.dark;
  d.l_min = r.l_min;
  d.l_max = r.l_max;
  d.max_levels = r.max_refine_levels;
  d.min_split = r.min_split_ns;
  d.signature = r.refine_signature;
  d.fixed_period = r.fixed_period_ns;
  return d;
}

}  // namespace

hipError_t launch_render(const SceneDev& sc, const ViewDev& view, int64_t t_ns, float* out, hipStream_t stream) {
  const size_t P = static_cast<size_t>(view.width) * view.height;
  render_kernel<<<static_cast<unsigned>((P + 255) / 256), 256, 0, stream>>>(sc, view, static_cast<double>(t_ns) * 1e-9,
                                                                           out);
  return hipGetLastError();
}

hipError_t launch_count(const SceneDev& sc, const ViewDev& view, const bsx_sampler_rules& r, int64_t t_w,
                        int32_t* counts, hipStream_t stream) {
  const size_t P = static_cast<size_t>(view.width) * view.height;
  count_kernel<<<static_cast<unsigned>((P + 255) / 256), 256, 0, stream>>>(sc, view, rules_dev(r), t_w, counts);
  return hipGetLastError();
}

hipError_t launch_write(const SceneDev& sc, const ViewDev& view, const bsx_sampler_rules& r, int64_t t_w,
                        const int64_t* row_start, int64_t* t, float* irr, hipStream_t stream) {
  const size_t P = static_cast<size_t>(view.width) * view.height;
  write_kernel<<<static_cast<unsigned>((P + 255) / 256), 256, 0, stream>>>(sc, view, rules_dev(r), t_w, row_start, t,
                                                                          irr);
  return hipGetLastError();
}

hipError_t launch_capture(const SceneDev& sc, const ViewDev& view, const bsx_sampler_rules& r, int64_t t_w,
                          uint32_t cap, int32_t* counts, CapturedSample* scratch, unsigned int* max_count,
                          hipStream_t stream) {
  const size_t P = static_cast<size_t>(view.width) * view.height;
  fill_u32_kernel<<<1, 64, 0, stream>>>(max_count, 1, 0u);
  capture_kernel<<<static_cast<unsigned>((P + 255) / 256), 256, 0, stream>>>(sc, view, rules_dev(r), t_w, cap, counts,
                                                                            scratch, max_count);
  return hipGetLastError();
}

hipError_t launch_compact(const ViewDev& view, int64_t t_w, uint32_t cap, const int32_t* counts,
                          const int64_t* row_start, const CapturedSample* scratch, int64_t* t, float* irr,
                          hipStream_t stream) {
  const size_t P = static_cast<size_t>(view.width) * view.height;
  compact_kernel<<<static_cast<unsigned>((P + 3) / 4), 256, 0, stream>>>(static_cast<uint32_t>(P), t_w, cap, counts,
                                                                        row_start, scratch, t, irr);
  return hipGetLastError();
}

hipError_t launch_overflow_write(const SceneDev& sc, const ViewDev& view, const bsx_sampler_rules& r, int64_t t_w,
                                 uint32_t cap, const int32_t* counts, const int64_t* row_start, int64_t* t, float* irr,
                                 hipStream_t stream) {
  const size_t P = static_cast<size_t>(view.width) * view.height;
  overflow_write_kernel<<<static_cast<unsigned>((P + 255) / 256), 256, 0, stream>>>(sc, view, rules_dev(r), t_w, cap,
                                                                                   counts, row_start, t, irr);
  return hipGetLastError();
}

hipError_t launch_counts_write(const SceneDev& sc, const ViewDev& view, int64_t t_w, const int32_t* counts,
                               const int64_t* row_start, int64_t total, int64_t* t, float* irr, hipStream_t stream) {
  if (total <= 0) return hipSuccess;
  counts_write_kernel<<<static_cast<unsigned>((total + 255) / 256), 256, 0, stream>>>(
      sc, view, t_w, counts, row_start, view.width * view.height, total, t, irr);
  return hipGetLastError();
}

hipError_t launch_requests(const SceneDev& sc, const ViewDev& view, int64_t t_base, const SampleRequestDev* req,
                           size_t n, const uint32_t* out_index, int32_t groups, float* irr, uint64_t* sig,
                           uint32_t* klass, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  requests_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(sc, view, t_base, req, n, out_index,
                                                                             groups, irr, sig, klass);
  return hipGetLastError();
}

size_t scan_scratch_count(size_t n) { return (n + kScanBlock - 1) / kScanBlock + 1; }

hipError_t launch_exclusive_scan(const int32_t* counts, size_t n, int64_t* out, int64_t* scratch, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + kScanBlock - 1) / kScanBlock;
  int64_t* total = scratch + blocks;
  scan_blocks_kernel<<<static_cast<unsigned>(blocks), kScanBlock, 0, stream>>>(counts, n, out, scratch);
  scan_sums_kernel<<<1, kScanBlock, 0, stream>>>(scratch, blocks, total);
  scan_add_kernel<<<static_cast<unsigned>(blocks), kScanBlock, 0, stream>>>(out, n, scratch);
  return hipGetLastError();
}

}  // namespace bsx
