// The instantaneous log response of one pixel over one call, as a lane program: one GPU lane runs
// one pixel from the model time to t_end (kernels.hip).
//
// Model (reference: PixelKernel::advance_instant, src/sensor/pixel.h:870-907; instant_search
// :1884-1916; instant_may_fire :1843-1876; fire :1531-1588). The photocurrent is
// I(t) = I_dk + sum_g s I_g(t), each I_g linear between the pixel's samples and held after the last.
// Delta(t) = ln(I(t) / i_scale) - s3_ref - n3_ref + leak (t - t_ref) + dr_ref, and an event fires at
// the first integer ns where Delta >= theta_on or <= -theta_off, found per sample segment by a
// time-ordered bisection that drops a node when the bounds of I keep Delta inside the levels. The
// reset is physical (blind on (t_c, t_c + T_r], re-reference at release) or ideal (r += p theta).
//
// Every arithmetic step is the reference's, in its order, because the fp32 results must equal the
// reference's fp32 CPU path bit for bit (events, crossing times, increments and state).
#pragma once

#include "../core/pmath.h"
#include "kconst.h"

namespace bsx {

template <int kMaxG>
struct InstantLane {
  // Per-pixel constants.
  uint32_t pixel;
  float theta_on, theta_off, leak, dark, scale;
  float irr_max, eps_ci, label_cut;
  int64_t refractory_ns;
  int G;
  bool ideal_reset;
  // The call's samples, augmented by the held input at the model time (sample 0).
  const int64_t* ts;
  const float* irr;
  int n, skip;
  int64_t t_state;
  // State.
  float held[kMaxG];
  float i_scale, s3_ref, n3_ref, dr_ref;
  int64_t t_ref, blind_until, t_search;
  uint32_t flags;
  uint32_t live;
  // Counters.
  uint32_t n_substeps, n_pieces, n_leaves;

  BSX_HD BSX_INLINE int n_aug() const { return 1 + n - skip; }
  BSX_HD BSX_INLINE int64_t sample_time(int i) const { return i == 0 ? t_state : ts[i - 1 + skip]; }
  BSX_HD BSX_INLINE float clampv(float v) const { return v >= 0.0f ? (v <= irr_max ? v : irr_max) : 0.0f; }
  BSX_HD BSX_INLINE float sample_value(int i, int g) const {
    if (i == 0 && skip == 0) return held[g];
    return clampv(irr[static_cast<int64_t>(i - 1 + skip) * G + g]);
  }

  // Live groups: a non-zero value somewhere in the augmented samples (a dead group adds exactly 0).
  BSX_HD BSX_INLINE void scan_input() {
    const uint32_t all = (1u << G) - 1u;
    uint32_t m = 0;
    if (skip == 0) {
      for (int g = 0; g < kMaxG; ++g)
        if (g < G && held[g] != 0.0f) m |= 1u << g;
    }
    for (int i = 0; i < n && m != all; ++i) {
      for (int g = 0; g < kMaxG; ++g)
        if (g < G && clampv(irr[static_cast<int64_t>(i) * G + g]) != 0.0f) m |= 1u << g;
    }
    live = m;
  }

  BSX_HD BSX_INLINE int segment_of(int seg, int64_t t) const {
    const int last = n_aug() - 1;
    while (seg < last && sample_time(seg + 1) <= t) ++seg;
    return seg;
  }
  BSX_HD BSX_INLINE int64_t segment_end(int seg) const {
    return seg + 1 < n_aug() ? sample_time(seg + 1) : kNoBreakpoint;
  }

  // I(t) in segment seg: dark + sum over live groups of s times the linear interpolation.
  BSX_HD BSX_INLINE float photocurrent(int seg, int64_t t) const {
    float I = dark;
    const bool held_seg = seg + 1 >= n_aug();
    int64_t t0 = 0, t1 = 0;
    float f = 0.0f;
    if (!held_seg) {
      t0 = sample_time(seg);
      t1 = sample_time(seg + 1);
      f = static_cast<float>(t - t0) / static_cast<float>(t1 - t0);
    }
    for (int g = 0; g < kMaxG; ++g) {
      if (g >= G || !((live >> g) & 1u)) continue;
      float v;
      if (held_seg) {
        v = sample_value(seg, g);
      } else {
        const float v0 = sample_value(seg, g), v1 = sample_value(seg + 1, g);
        v = v0 + (v1 - v0) * f;
      }
      I = I + scale * v;
    }
    return I;
  }

  BSX_HD BSX_INLINE float log_ratio(float I) const { return pm::log(I / i_scale); }

  BSX_HD BSX_INLINE float delta_at(int64_t t, float x) const {
    const float ds = x - s3_ref;
    const float dn = 0.0f - n3_ref;
    const float dl = leak * ns_to_s64<float>(t - t_ref) + dr_ref;
    return ds + dn + dl;
  }

  // Can Delta reach a level at an integer of (a, c] of a segment whose photocurrent is linear in t
  // (constant waveforms: curvature 0, no range cap)?
  BSX_HD BSX_INLINE bool may_fire(int64_t a, int64_t c, float Ia, float Ic) const {
    constexpr float eps = FloatEps<float>::value;
    float hi = vmax(Ia, Ic), lo = vmin(Ia, Ic);
    const float h = static_cast<float>(c - a) * 1e-9f;
    const float hh = h * h * 0.0f;
    const float margin = 64.0f * eps * (hi + 2.0f * 0.0f);
    if (leak == 0.0f && hi - lo >= 0.5f * hh) {
      const float D = delta_at(c, log_ratio(Ic));
      return D >= theta_on || D <= -theta_off;
    }
    if (Ic >= Ia) return delta_at(c, log_ratio(Ic)) >= theta_on;
    hi = vmin(hi + 0.125f * hh, inf_<float>()) + margin;
    lo = vmax(lo - 0.125f * hh, dark) - margin;
    const float dl_a = leak * ns_to_s64<float>(a - t_ref) + dr_ref;
    const float dl_c = leak * ns_to_s64<float>(c - t_ref) + dr_ref;
    const float base = 0.0f - n3_ref - s3_ref;
    const float x_hi = log_ratio(hi);
    if (x_hi + base + dl_c + 16.0f * eps * (1.0f + fabs_(x_hi) + fabs_(s3_ref) + fabs_(dl_c)) >= theta_on)
      return true;
    if (!(lo > 0.0f)) return true;
    const float x_lo = log_ratio(lo);
    return x_lo + base + dl_a - 16.0f * eps * (1.0f + fabs_(x_lo) + fabs_(s3_ref) + fabs_(dl_a)) <= -theta_off;
  }

  struct Hit {
    bool found, on;
    int64_t t;
    float x;
  };

  BSX_HD BSX_INLINE bool fires(int64_t t, Hit& h) const {
    const float D = delta_at(t, h.x);
    h.found = D >= theta_on || D <= -theta_off;
    h.on = D >= theta_on;
    h.t = t;
    return h.found;
  }

  // First integer t in [lo, hi) of segment seg where Delta fires: time-ordered depth-first bisection
  // of (lo, hi - 1]. The pending right siblings are kept as their right ends only: in time order a
  // popped node always starts where the last finished one ended.
  BSX_HD BSX_INLINE Hit search(int seg, int64_t lo, int64_t hi) {
    const float I_lo = photocurrent(seg, lo);
    Hit out{false, false, lo, log_ratio(I_lo)};
    if (fires(lo, out)) return out;
    if (hi - lo == 1) return out;
    int64_t pend_c[32];
    float pend_Ic[32];
    int sp = 0;
    int64_t a = lo, c = hi - 1;
    float Ia = I_lo, Ic = photocurrent(seg, hi - 1);
    while (true) {
      bool descend = false;
      if (c - a == 1) {
        out.t = c;
        out.x = log_ratio(Ic);
        if (fires(c, out)) return out;
      } else {
        ++n_leaves;
        if (may_fire(a, c, Ia, Ic)) descend = true;
      }
      if (descend) {
        const int64_t m = a + (c - a) / 2;
        const float Im = photocurrent(seg, m);
        pend_c[sp] = c;
        pend_Ic[sp] = Ic;
        ++sp;
        c = m;
        Ic = Im;
        continue;
      }
      if (sp == 0) break;
      --sp;
      a = c;
      Ia = Ic;
      c = pend_c[sp];
      Ic = pend_Ic[sp];
    }
    out.found = false;
    return out;
  }

  BSX_HD BSX_INLINE void set_reference(int64_t t, float s3, float n3, float offset) {
    t_ref = t;
    s3_ref = s3;
    n3_ref = n3;
    dr_ref = offset;
  }

  BSX_HD BSX_INLINE static uint8_t split_label(float ds, float dl, float D, float cut) {
    const float a_s = ds / D, a_l = dl / D;
    if (a_s >= cut && a_s >= a_l) return kLabelSignal;
    if (a_l >= cut && a_l > a_s) return kLabelLeak;
    return kLabelNoise;
  }

  template <class Sink>
  BSX_HD BSX_INLINE void fire(int64_t t, bool on, float s3, float n3, Sink& sink) {
    const float ds = s3 - s3_ref;
    const float dn = n3 - n3_ref;
    const float l = leak * ns_to_s64<float>(t - t_ref);
    const float dl = l + dr_ref;
    const float D = ds + dn + dl;
    const bool first = (flags & kFlagFirstAfterInit) != 0;
    uint8_t label;
    if (first) {
      const float Dl = ds + dn + l;
      label = (on ? Dl > 0.0f : Dl < 0.0f) ? split_label(ds, l, Dl, label_cut) : kLabelNoise;
    } else {
      label = split_label(ds, dl, D, label_cut);
    }
    RawEventDev ev;
    ev.pixel = pixel;
    ev.polarity = on ? 1 : 0;
    ev.label = label;
    ev.flags = static_cast<uint16_t>(flags);
    ev.t_cross = t;
    ev.ds = ds;
    ev.dn = dn;
    ev.dl = dl;
    ev.l_first = first ? l : 0.0f;
    sink.push(ev);
    flags = 0;
    if (ideal_reset) {
      set_reference(t, s3, n3, D - (on ? theta_on : -theta_off));
      t_search = t + 1 + refractory_ns;
    } else if (refractory_ns == 0) {
      set_reference(t, s3, n3, -eps_ci);
      t_search = t + 1;
    } else {
      blind_until = t + refractory_ns;
      t_search = blind_until + 1;
    }
  }

  // The call: pieces of one sample segment, cut at multiples of 2^23 ns.
  template <class Sink>
  BSX_HD BSX_INLINE void advance(int64_t t_end, Sink& sink) {
    int seg = 0;  // the model time is sample 0's
    int64_t t = t_state;
    while (t < t_end) {
      seg = segment_of(seg, t);
      const int64_t e =
          vmin(vmin(t_end, segment_end(seg)), ((t >> kPlainSegmentLog2) + 1) * kPlainSegmentNs);
      ++n_substeps;
      ++n_pieces;
      while (true) {
        if (blind_until != kNotBlind) {
          if (blind_until >= e) break;
          set_reference(blind_until, log_ratio(photocurrent(seg, blind_until)), 0.0f, -eps_ci);
          blind_until = kNotBlind;
        }
        const int64_t lo = vmax(t_search, t);
        if (lo >= e) break;
        const Hit hit = search(seg, lo, e);
        if (!hit.found) {
          t_search = vmax(t_search, e);
          break;
        }
        fire(hit.t, hit.on, hit.x, 0.0f, sink);
      }
      t = e;
    }
    if (n > 0) {
      for (int g = 0; g < kMaxG; ++g)
        if (g < G) held[g] = sample_value(n_aug() - 1, g);
    }
  }
};

}  // namespace bsx
