// The learned sample-time controller on the GPU: the proposer's features, the training examples built
// from dense traces of the scene, and the learned sampler (one lane per pixel, the wave's proposals
// batched through the network on the MFMA units, the plain controller's refinement as the fallback).
#include <hip/hip_runtime.h>

#include "../mlp/mlp_dev.h"
#include "learned_kernels.h"

namespace bsx {

namespace {

using mlp::kIn;

constexpr int64_t kNotBlindLink = INT64_MIN;  // the sensor state's "not blind" (no pending release)

// The controller's view of a pixel: up to four latest samples (index 0 the newest) of the log
// irradiance L = ln(E + dark), their times, and the reference level it tracks as an ideal event
// sensor of threshold theta_ref would (ref moves by theta_ref per crossing).
struct History {
  int64_t t[4];
  float L[4];
  int n;
  float ref;
  float e0;      // the newest sample's irradiance
  uint32_t sig;  // and its signature (the shape seen, 0: background)
  // The band about ref in which no event starts: [ref - th_off, ref + th_on]; a pending release of a
  // physical reset (kNoRelease: none) leaves the reference unknown until it passes.
  float th_on, th_off;
  bool coupled, ideal;
  float eps_ci;
  int64_t refractory;
  int64_t release;
  static constexpr int64_t kNoRelease = INT64_MIN;

  // Room to the nearer band edge at level L (0 while a release is pending).
  __device__ __forceinline__ float room(float Lv) const {
    if (release != kNoRelease) return 0.0f;
    return fminf(ref + th_on - Lv, Lv - (ref - th_off));
  }

  __device__ __forceinline__ void push(int64_t tn, float en, uint32_t sn, float Ln, float dark) {
    const int64_t tp = t[0];
    const float ep = e0;
    for (int k = 3; k > 0; --k) {
      t[k] = t[k - 1];
      L[k] = L[k - 1];
    }
    t[0] = tn;
    L[0] = Ln;
    e0 = en;
    sig = sn;
    n = n < 4 ? n + 1 : 4;
    // The sensor interpolates linearly in irradiance between samples: the log level at time tx.
    const auto at = [&](int64_t tx) {
      const float w = tn > tp ? static_cast<float>(static_cast<double>(tx - tp) / static_cast<double>(tn - tp)) : 1.0f;
      return logf(ep + (en - ep) * w + dark);
    };
    if (!coupled) {  // an ideal sensor of step th_on
      while (Ln - ref >= th_on) ref += th_on;
      while (ref - Ln >= th_on) ref -= th_on;
      return;
    }
    if (release != kNoRelease && release <= tn) {
      ref = at(release) + eps_ci;
      release = kNoRelease;
    }
    if (release != kNoRelease) return;
    const float d = Ln - ref;
    if (!(d >= th_on || d <= -th_off)) return;
    if (ideal) {
      while (Ln - ref >= th_on) ref += th_on;
      while (ref - Ln >= th_off) ref -= th_off;
      return;
    }
    // physical reset: the crossing between the two samples (log-linear estimate), then blind for the
    // refractory period and a new reference at the release
    const float Lp = at(tp);
    const float edge = d >= th_on ? ref + th_on : ref - th_off;
    const float f = Ln != Lp ? fminf(fmaxf((edge - Lp) / (Ln - Lp), 0.0f), 1.0f) : 1.0f;
    release = tp + static_cast<int64_t>(f * static_cast<float>(tn - tp)) + refractory;
    if (release <= tn) {
      ref = at(release) + eps_ci;
      release = kNoRelease;
    }
  }
};

// The 12 features of a decision at the newest sample (missing history: the oldest sample repeated
// 100 us earlier):
//   0: 10 (L0 - ref);  1-3: 10 (L_k - L0);  4-6: asinh of the slopes (L_{k-1} - L_k) / h_k per ms;
//   7-9: (ln h_k - ln 1e5) / 2 for the gaps h_k = t_{k-1} - t_k (ns);  10: 10 (max L - min L);
//   11: 10 (theta_band - |L0 - ref|), the headroom to the band where no event can start.
__device__ __forceinline__ void features(const History& h, float theta_band, float* f) {
  float L[4];
  double t[4];
  for (int k = 0; k < 4; ++k) {
    const int src = k < h.n ? k : h.n - 1;
    L[k] = h.L[src];
    t[k] = static_cast<double>(h.t[src]) - (k < h.n ? 0.0 : 1e5 * (k - h.n + 1));
  }
  f[0] = 10.0f * (L[0] - h.ref);
  float lo = L[0], hi = L[0];
  for (int k = 1; k < 4; ++k) {
    f[k] = 10.0f * (L[k] - L[0]);
    const float hk = static_cast<float>(t[k - 1] - t[k]);
    f[3 + k] = asinhf((L[k - 1] - L[k]) / (hk * 1e-6f));
    f[6 + k] = (logf(hk) - kLnDtCenter) * 0.5f;
    lo = fminf(lo, L[k]);
    hi = fmaxf(hi, L[k]);
  }
  f[10] = 10.0f * (hi - lo);
  f[11] = h.coupled ? 10.0f * h.room(L[0]) : 10.0f * (theta_band - fabsf(L[0] - h.ref));
}

__device__ __forceinline__ float log_irr(float e, float dark) { return logf(e + dark); }

// ---- Dense traces and training examples -----------------------------------------------------------

// trace[j * P + p] = L of pixel p at t0 + j delta (j < J), ref[j * P + p] = the ideal sensor's
// reference after that sample (started at the first), active[p] = 1 if the trace varies.
__global__ __launch_bounds__(256) void dense_trace_kernel(SceneDev sc, ViewDev view, int64_t t0, int64_t delta, int J,
                                                          float dark, float theta_ref, float* trace, float* ref,
                                                          uint8_t* active) {
  const uint32_t P = view.width * view.height;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double x = view.ox + (static_cast<double>(p % view.width) + 0.5) * view.scale;
  const double y = view.oy + (static_cast<double>(p / view.width) + 0.5) * view.scale;
  uint64_t mask[kMaskWords];
  const double t0s = static_cast<double>(t0) * 1e-9;
  (void)plan_pixel(sc, x, y, t0s, t0s + static_cast<double>(delta) * J * 1e-9, mask);
  float r = 0.0f, lo = 0.0f, hi = 0.0f;
  for (int j = 0; j < J; ++j) {
    const double ts = t0s + static_cast<double>(delta) * j * 1e-9;
    const float L = log_irr(eval_scene(sc, mask, x, y, ts).e, dark);
    if (j == 0) {
      r = L;
      lo = hi = L;
    }
    while (L - r >= theta_ref) r += theta_ref;
    while (r - L >= theta_ref) r -= theta_ref;
    lo = fminf(lo, L);
    hi = fmaxf(hi, L);
    trace[static_cast<int64_t>(j) * P + p] = L;
    ref[static_cast<int64_t>(j) * P + p] = r;
  }
  active[p] = hi - lo > 1e-6f ? 1 : 0;
}

__device__ __forceinline__ uint64_t hash64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unit(uint64_t h) { return static_cast<double>(h >> 11) * 0x1p-53; }

// Does the chord from dense point j0 to j0 + m (linear in E + dark, the sensor's interpolation) stay
// within the allowed error of the trace at every dense point between? Allowed: tol, or, while the
// trace stays inside the band |L - ref| < theta_band over [j0, j0 + m], the distance to the band edge
// less tol (no event can start there, so a larger error changes no event).
__device__ __forceinline__ bool chord_ok(const float* tr, uint32_t P, uint32_t p, int j0, int m, float ref,
                                         float theta_band, float tol) {
  const float La = tr[static_cast<int64_t>(j0) * P + p], Lb = tr[static_cast<int64_t>(j0 + m) * P + p];
  const float Ea = expf(La), Eb = expf(Lb);
  bool inside = fabsf(La - ref) < theta_band && fabsf(Lb - ref) < theta_band;
  float worst_tight = 0.0f, worst_band = -1.0f;
  for (int j = 1; j < m; ++j) {
    const float L = tr[static_cast<int64_t>(j0 + j) * P + p];
    const float w = static_cast<float>(j) / static_cast<float>(m);
    const float err = fabsf(logf(Ea + (Eb - Ea) * w) - L);
    const float room = theta_band - fabsf(L - ref);
    inside = inside && room > 0.0f;
    worst_tight = fmaxf(worst_tight, err - tol);
    worst_band = fmaxf(worst_band, err - fmaxf(tol, room - tol));
  }
  return inside ? worst_band <= 0.0f : worst_tight <= 0.0f;
}

// Example e: a pixel of the candidate list, a decision point j0 and a history at gaps of 1 .. 256
// dense steps (log-uniform) before it; target = (ln(m* delta) - ln 1e5) / 2 for the largest step m*
// (dense steps, at most max_steps) whose chord passes chord_ok (doubling, then bisection).
__global__ __launch_bounds__(256) void examples_kernel(const float* trace, const float* ref, uint32_t P, int J,
                                                       const uint32_t* pixels, uint32_t n_pixels, int64_t delta,
                                                       int max_steps, float theta_band, float tol, uint64_t seed,
                                                       int64_t n, float* x, float* target) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  uint64_t s = hash64(seed ^ hash64(static_cast<uint64_t>(e)));
  const uint32_t p = pixels[hash64(s++) % n_pixels];
  int gaps[3];
  for (int k = 0; k < 3; ++k) gaps[k] = static_cast<int>(exp(unit(hash64(s++)) * log(256.0)) + 0.5);
  const int hist = gaps[0] + gaps[1] + gaps[2];
  const int span = J - 1 - max_steps - hist;
  const int j0 = hist + static_cast<int>(unit(hash64(s++)) * span);
  History h{};
  h.n = 4;
  h.ref = ref[static_cast<int64_t>(j0) * P + p];
  h.coupled = false;
  h.release = History::kNoRelease;
  int j = j0;
  for (int k = 0; k < 4; ++k) {
    h.t[k] = static_cast<int64_t>(j) * delta;
    h.L[k] = trace[static_cast<int64_t>(j) * P + p];
    if (k < 3) j -= gaps[k];
  }
  float f[kIn];
  features(h, theta_band, f);
  for (int c = 0; c < kIn; ++c) x[e * kIn + c] = f[c];
  int good = 0, bad = max_steps + 1;
  for (int m = 1; m <= max_steps; m = m * 2 > max_steps && m < max_steps ? max_steps : m * 2) {
    if (chord_ok(trace, P, p, j0, m, h.ref, theta_band, tol)) {
      good = m;
      if (m == max_steps) break;
    } else {
      bad = m;
      break;
    }
  }
  while (bad - good > 1 && good >= 1) {
    const int mid = (good + bad) / 2;
    if (chord_ok(trace, P, p, j0, mid, h.ref, theta_band, tol)) good = mid;
    else bad = mid;
  }
  const double steps = good >= 1 ? good : 0.5;  // not even one dense step: half of one
  target[e] = static_cast<float>((log(steps * static_cast<double>(delta)) - kLnDtCenter) * 0.5);
}

// ---- The learned sampler -------------------------------------------------------------------------

struct LearnedState {
  int64_t* t;  // [4 * P]
  float* L;    // [4 * P]
  int32_t* n;
  float* ref;
  float* e0;
  uint32_t* sig;
};

__global__ __launch_bounds__(256) void learned_init_kernel(SceneDev sc, ViewDev view, int64_t t0, float dark,
                                                           LearnedState st) {
  const uint32_t P = view.width * view.height;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double x = view.ox + (static_cast<double>(p % view.width) + 0.5) * view.scale;
  const double y = view.oy + (static_cast<double>(p / view.width) + 0.5) * view.scale;
  uint64_t mask[kMaskWords];
  for (int w = 0; w < kMaskWords; ++w) mask[w] = 0;
  for (int i = 0; i < sc.n; ++i) mask[i >> 6] |= 1ull << (i & 63);
  const Sample s0 = eval_scene(sc, mask, x, y, static_cast<double>(t0) * 1e-9);
  const float L = log_irr(s0.e, dark);
  for (int k = 0; k < 4; ++k) {
    st.t[static_cast<int64_t>(k) * P + p] = t0;
    st.L[static_cast<int64_t>(k) * P + p] = L;
  }
  st.n[p] = 1;
  st.ref[p] = L;
  st.e0[p] = s0.e;
  st.sig[p] = s0.sig;
}

// One window (t_w, t_w + W] of every pixel. kWrite false: counts[p] = the pixel's samples and the
// work counters; true: the samples at row_start[p] and the new controller state.
template <bool kWrite>
__global__ __launch_bounds__(64) void learned_window_kernel(SceneDev sc, ViewDev view, LearnedRules r, const uint16_t* wb,
                                                            const float* w, int64_t t_w, LearnedState st,
                                                            SensorLink link, int32_t* counts,
                                                            const int64_t* row_start, int64_t* t_out, float* irr_out,
                                                            unsigned long long* work) {
  __shared__ mlp::WeightsLds W;
  __shared__ mlp::WaveLds S;
  mlp::load_weights(W, wb, w);
  const uint32_t P = view.width * view.height;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = p < P;
  const double x = view.ox + (static_cast<double>(valid ? p % view.width : 0) + 0.5) * view.scale;
  const double y = view.oy + (static_cast<double>(valid ? p / view.width : 0) + 0.5) * view.scale;
  const double tw_s = static_cast<double>(t_w) * 1e-9;
  uint64_t mask[kMaskWords];
  const float v_hat = plan_pixel(sc, x, y, tw_s, tw_s + kWindowNs * 1e-9, mask);
  History h{};
  if (valid) {
    for (int k = 0; k < 4; ++k) {
      h.t[k] = st.t[static_cast<int64_t>(k) * P + p];
      h.L[k] = st.L[static_cast<int64_t>(k) * P + p];
    }
    h.n = st.n[p];
    h.ref = st.ref[p];
    h.e0 = st.e0[p];
    h.sig = st.sig[p];
  }
  h.release = History::kNoRelease;
  h.coupled = valid && link.theta_on != nullptr;
  h.th_on = h.th_off = r.theta_ref;
  if (h.coupled) {
    // The sensor's band at the window start: Delta = L + ln(scale / i_scale) - s3_ref - n3_ref + dr_ref (the
    // controller's L = ln(E + dark) stands for ln(dark_p / scale + E)).
    h.ref = link.s3_ref[p] + link.n3_ref[p] - link.dr_ref[p] - logf(link.scale[p] / link.i_scale[p]);
    h.th_on = link.theta_on[p];
    h.th_off = link.theta_off[p];
    h.ideal = link.ideal_reset != 0;
    h.eps_ci = link.eps_ci;
    h.refractory = link.refractory_ns;
    const int64_t bu = link.blind_until[p];
    if (bu != kNotBlindLink && bu >= h.t[0]) h.release = bu;
  }
  // The plain controller's lattice step of this pixel (its level rule), the interval of the near mode.
  int64_t lattice = kWindowNs;
  if (v_hat > 0.0f) {
    const double lv = ceil(log2(static_cast<double>(v_hat) / view.scale * (kWindowNs * 1e-9) / r.d_px));
    const int level = lv != lv ? r.l_min : static_cast<int>(fmin(fmax(lv, static_cast<double>(r.l_min)),
                                                                 static_cast<double>(r.l_max)));
    lattice = kWindowNs >> level;
  }
  // A pixel that took no samples in earlier windows was static there (no shape reached it): its value at
  // t_w is its newest sample's. The history moves to t_w with that value (no render, no sample), so that
  // proposals start at the window and not at the stale sample.
  if (valid && h.t[0] < t_w) h.push(t_w, h.e0, h.sig, h.L[0], r.dark);
  const int64_t t_end = t_w + kWindowNs;
  bool active = valid && v_hat > 0.0f && h.t[0] < t_end;
  uint32_t count = 0, proposals = 0, fallbacks = 0;
  const int64_t base = kWrite && valid ? row_start[p] : 0;
  const auto emit = [&](int64_t ts, float e, uint32_t sg) {
    if constexpr (kWrite) {
      t_out[base + count] = ts;
      irr_out[base + count] = e;
    }
    ++count;
    h.push(ts, e, sg, log_irr(e, r.dark), r.dark);
  };
  const auto at = [&](int64_t ts) { return eval_scene(sc, mask, x, y, static_cast<double>(ts) * 1e-9); };
  while (__ballot(active) != 0) {
    float f[kIn];
    if (active) features(h, r.theta_band, f);
    else
      for (int c = 0; c < kIn; ++c) f[c] = 0.0f;
    mlp::put_features(S, f);
    const float out = mlp::wave_forward(W, S);
    if (active) {
      ++proposals;
      const double dt_pred = exp(2.0 * static_cast<double>(out) + kLnDtCenter) * r.margin;
      // Near a band edge (or with the reference unknown) the plain controller's density: two lattice steps,
      // judged by eta_L at the midpoint as the plain controller judges its lattice triples.
      const float room0 = h.room(h.L[0]);
      const bool near = room0 < r.near;
      double dt_max = near ? static_cast<double>(2 * lattice) : 0x1p23;
      if (!near && h.n >= 2) {
        // far: no longer than the recent slope needs to reach the near zone
        const double slope = fabs(static_cast<double>(h.L[0] - h.L[1])) / static_cast<double>(h.t[0] - h.t[1]);
        if (slope > 0.0) dt_max = fmin(dt_max, static_cast<double>(room0 - r.near) / slope);
      }
      const int64_t dt = static_cast<int64_t>(fmin(fmax(fmin(dt_pred, dt_max), static_cast<double>(r.min_dt)), 0x1p23));
      const int64_t ta = h.t[0];
      const int64_t tb = ta + dt < t_end ? ta + dt : t_end;
      const float Ea = h.e0;
      const uint32_t sig_a = h.sig;
      const Sample sb = at(tb);
      if (tb - ta < 2) {
        emit(tb, sb.e, sb.sig);
      } else {
        // The a-posteriori check at the interval's midpoint (rendered, a sample): far from the edges the
        // chord's log error against the allowed error (tol_log, or the band room less tol_log while the
        // three values stay inside the band); near them eta_L, as the plain rules. A signature flip
        // (another shape or the background seen) fails it in both.
        const int64_t tm = ta + (tb - ta) / 2;
        const Sample sm = at(tm);
        const float La = h.L[0], Lb = log_irr(sb.e, r.dark), Lm = log_irr(sm.e, r.dark);
        bool ok;
        if (near) {
          ok = midpoint_eta(r.dark, Ea, sm.e, sb.e) <= r.tol_eta &&
               !(r.signature && (sig_a != sm.sig || sm.sig != sb.sig));
        } else {
          const float err = fabsf(logf(0.5f * (Ea + sb.e) + r.dark) - Lm);
          const float room = h.coupled ? fminf(fminf(h.room(La), h.room(Lb)), h.room(Lm))
                                       : r.theta_band - fmaxf(fmaxf(fabsf(La - h.ref), fabsf(Lb - h.ref)),
                                                              fabsf(Lm - h.ref));
          const float allowed = room > 0.0f ? fmaxf(r.tol_log, room - r.tol_log) : r.tol_log;
          // an interval that reaches the near zone is refined as the plain controller would
          ok = err <= allowed && !(r.signature && (sig_a != sm.sig || sm.sig != sb.sig)) &&
               !(h.coupled && room < r.near);
        }
        if (ok) {
          emit(tm, sm.e, sm.sig);
          emit(tb, sb.e, sb.sig);
        } else {
          // Fallback: the plain controller's refinement of the two halves (eta_L and signature flips,
          // down to max_levels, no split below min_split), depth-first in time order: interior
          // points as they are decided, then tb.
          ++fallbacks;
          struct Node {
            int64_t lo, hi;
            float e_lo, e_hi;
            uint32_t s_lo, s_hi;
            int32_t depth;  // > 0: refine at this depth; -1: emit (lo, e_lo, s_lo)
          };
          Node stack[24];
          int sp = 0;
          const bool deep0 = r.max_levels >= 1;
          if (deep0 && tb - tm >= r.min_split) stack[sp++] = Node{tm, tb, sm.e, sb.e, sm.sig, sb.sig, 1};
          stack[sp++] = Node{tm, 0, sm.e, 0.0f, sm.sig, 0u, -1};
          if (deep0 && tm - ta >= r.min_split) stack[sp++] = Node{ta, tm, Ea, sm.e, sig_a, sm.sig, 1};
          while (sp > 0) {
            const Node nd = stack[--sp];
            if (nd.depth < 0) {
              emit(nd.lo, nd.e_lo, nd.s_lo);
              continue;
            }
            const int64_t g = nd.lo + (nd.hi - nd.lo) / 2;
            const Sample sg = at(g);
            const bool s1 = r.signature && nd.s_lo != sg.sig, s2 = r.signature && sg.sig != nd.s_hi;
            const bool bad = midpoint_eta(r.dark, nd.e_lo, sg.e, nd.e_hi) > r.tol_eta;
            const bool p1 = (s1 || bad) && g - nd.lo >= r.min_split;
            const bool p2 = (s2 || bad) && nd.hi - g >= r.min_split;
            const bool deeper = nd.depth < r.max_levels;
            if (p2 && deeper) stack[sp++] = Node{g, nd.hi, sg.e, nd.e_hi, sg.sig, nd.s_hi, nd.depth + 1};
            stack[sp++] = Node{g, 0, sg.e, 0.0f, sg.sig, 0u, -1};
            if (p1 && deeper) stack[sp++] = Node{nd.lo, g, nd.e_lo, sg.e, nd.s_lo, sg.sig, nd.depth + 1};
          }
          emit(tb, sb.e, sb.sig);
        }
      }
      active = h.t[0] < t_end;
    }
  }
  if (valid) {
    if constexpr (kWrite) {
      for (int k = 0; k < 4; ++k) {
        st.t[static_cast<int64_t>(k) * P + p] = h.t[k];
        st.L[static_cast<int64_t>(k) * P + p] = h.L[k];
      }
      st.n[p] = h.n;
      st.ref[p] = h.ref;
      st.e0[p] = h.e0;
      st.sig[p] = h.sig;
    } else {
      counts[p] = static_cast<int32_t>(count);
    }
  }
  if constexpr (!kWrite) {
    unsigned long long a = proposals, b = fallbacks;
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off);
      b += __shfl_xor(b, off);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(work + 0, a);
      atomicAdd(work + 1, b);
    }
  }
}

}  // namespace

hipError_t launch_dense_trace(const SceneDev& sc, const ViewDev& view, int64_t t0, int64_t delta, int J, float dark,
                              float theta_ref, float* trace, float* ref, uint8_t* active, hipStream_t stream) {
  const uint32_t P = view.width * view.height;
  dense_trace_kernel<<<(P + 255) / 256, 256, 0, stream>>>(sc, view, t0, delta, J, dark, theta_ref, trace, ref, active);
  return hipGetLastError();
}

hipError_t launch_examples(const float* trace, const float* ref, uint32_t P, int J, const uint32_t* pixels,
                           uint32_t n_pixels, int64_t delta, int max_steps, float theta_band, float tol, uint64_t seed,
                           int64_t n, float* x, float* target, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  examples_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(
      trace, ref, P, J, pixels, n_pixels, delta, max_steps, theta_band, tol, seed, n, x, target);
  return hipGetLastError();
}

hipError_t launch_learned_init(const SceneDev& sc, const ViewDev& view, int64_t t0, float dark,
                               const LearnedStateDev& s, hipStream_t stream) {
  const uint32_t P = view.width * view.height;
  learned_init_kernel<<<(P + 255) / 256, 256, 0, stream>>>(sc, view, t0, dark,
                                                           LearnedState{s.t, s.L, s.n, s.ref, s.e0, s.sig});
  return hipGetLastError();
}

hipError_t launch_learned_window(const SceneDev& sc, const ViewDev& view, const LearnedRules& r, const uint16_t* wb,
                                 const float* w, int64_t t_w, const LearnedStateDev& s, const SensorLink& link,
                                 int32_t* counts, const int64_t* row_start, int64_t* t_out, float* irr_out,
                                 unsigned long long* work, bool write, hipStream_t stream) {
  const uint32_t P = view.width * view.height;
  const LearnedState st{s.t, s.L, s.n, s.ref, s.e0, s.sig};
  if (write)
    learned_window_kernel<true><<<(P + 63) / 64, 64, 0, stream>>>(sc, view, r, wb, w, t_w, st, link, counts,
                                                                  row_start, t_out, irr_out, work);
  else
    learned_window_kernel<false><<<(P + 63) / 64, 64, 0, stream>>>(sc, view, r, wb, w, t_w, st, link, counts,
                                                                   row_start, t_out, irr_out, work);
  return hipGetLastError();
}

}  // namespace bsx
