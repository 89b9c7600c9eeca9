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
// A hot pixel's events are flagged kFlagHotPixel and labelled HOT unless SIGNAL (fire :1563-1569).
//
// Every arithmetic step is the reference's, in its order, because the fp32 results must equal the
// reference's fp32 CPU path bit for bit (events, crossing times, increments and state). What differs
// is the data movement, which leaves results unchanged:
//   - segments are walked once, in order, with both end samples in registers (the reference
//     re-reads them per evaluation), and the next sample is loaded one segment ahead so its memory
//     latency hides behind the current segment's work;
//   - the live-group scan is dropped: a dead group adds scale x (+-0) to a sum that starts at the
//     dark current > 0, which leaves it bit for bit unchanged.
#pragma once

#include "../core/philox.h"
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
  bool hot = false;  // a hot pixel: its events carry kFlagHotPixel, and HOT for any label but SIGNAL
  // The call's input run.
  const BSX_GLOBAL int64_t* ts;
  const BSX_GLOBAL float* irr;
  int n, skip;
  int64_t t_state;
  // State.
  float held[kMaxG];
  float i_scale, s3_ref, n3_ref, dr_ref;
  int64_t t_ref, blind_until, t_search;
  uint32_t flags;
  // The current segment [ta, tb) with its end values (tb = kNoBreakpoint: the held segment).
  int64_t ta, tb;
  float va[kMaxG], vb[kMaxG];
  // The rate-specified shot-noise process (this library's extension; rate 0 = off).
  double shot_rate;
  uint64_t seed;
  uint32_t camera;
  int64_t shot_next;
  uint32_t shot_index;
  // Counters.
  uint32_t n_substeps, n_pieces, n_leaves;

  BSX_HD BSX_INLINE float clampv(float v) const { return v >= 0.0f ? (v <= irr_max ? v : irr_max) : 0.0f; }

  // I(t) in the current segment: dark + sum over groups of s times the linear interpolation (or the
  // held value in the last segment), as the reference's photocurrent/group_value.
  BSX_HD BSX_INLINE float photocurrent(int64_t t) const {
    float I = dark;
    if (tb == kNoBreakpoint) {
      for (int g = 0; g < kMaxG; ++g)
        if (g < G) I = I + scale * va[g];
      return I;
    }
    const float f = static_cast<float>(t - ta) / static_cast<float>(tb - ta);
    for (int g = 0; g < kMaxG; ++g)
      if (g < G) I = I + scale * (va[g] + (vb[g] - va[g]) * f);
    return I;
  }

  // photocurrent(ta) of a segment with an end sample, without its interpolation: there f = +0, so
  // (vb - va) f is a zero and va + that zero is va (a zero's sign can differ, which the sum that
  // starts at dark then absorbs: the same bits).
  BSX_HD BSX_INLINE float photocurrent_start() const {
    float I = dark;
    for (int g = 0; g < kMaxG; ++g)
      if (g < G) I = I + scale * va[g];
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

  // ---- The first crossing by monotone transition searches ---------------------------------------
  // The bisection below (the reference's instant_search) returns the first integer t of [lo, hi)
  // where Delta fires, whenever its exclusion test is conservative: a node (a, c] holding a firing
  // integer is never dropped, so the time-ordered walk reaches the first one. It is conservative
  // here because every step of Delta(t) is monotone in t on a segment whose groups all rise or all
  // fall: the interpolation fraction, the products and sums of I(t) and its division by i_scale are
  // correctly rounded operations of monotone arguments; pm::log is monotone on every positive float
  // (checked exhaustively); the leak term leak * (float)((double)(t - t_ref) * 1e-9) rises with t
  // (leak >= 0); and with n3_ref = 0 the bound of the falling-segment test is formed by the same
  // operations as Delta(t) with larger (smaller) arguments. So the first crossing may be found by any
  // means, and these searches find it from a few evaluations of the same fp32 Delta(t):
  //   - Delta non-decreasing (rising or flat I) or non-increasing (falling I without leak): only
  //     one level can be crossed after lo, and "crossed at t" is monotone in t: the transition is
  //     bracketed by interpolation and bisection;
  //   - falling I with leak: Delta(t) lies between fl(S(t) + dl(lo)) and fl(S(t) + dl(c)), both
  //     non-increasing (S = the leak-free part), so the first crossing of the lower level lies in
  //     [tL, tU] between their transitions, which are found as above. Inside, Delta = fl(S + dl)
  //     with S non-increasing and dl(t) non-decreasing, so fl(S(c') + dl(a')) bounds Delta from
  //     below on a node [a', c'], and where dl(a') = dl(c') Delta is non-increasing on the node: a
  //     time-ordered bisection of [tL, tU] drops the nodes whose bound stays above the level and
  //     brackets the crossing in the first node of constant dl that reaches it. (The generic
  //     bisection's bounds of I cannot drop nodes where Delta runs level with -theta_off, where the
  //     leak balances a slow fall, and visits most of their integers.) A segment where the upper
  //     envelope allows an ON crossing takes the same bisection with both levels (two_sided_search).
  // Returns false when the generic bisection must decide (mixed group slopes, n3_ref != 0).

  BSX_HD BSX_INLINE float s_at(int64_t t, float& x) const {
    x = log_ratio(photocurrent(t));
    return (x - s3_ref) + (0.0f - n3_ref);
  }
  BSX_HD BSX_INLINE float dl_at(int64_t t) const { return leak * ns_to_s64<float>(t - t_ref) + dr_ref; }

  // The first t of (a, b] where pred holds, pred monotone (false at a, true at b); v_a, v_b the
  // values of the bracket ends, interpolated towards `level` (a guess only; the predicate decides).
  template <class F>
  BSX_HD BSX_INLINE int64_t first_true(int64_t a, float v_a, int64_t b, float v_b, float level, F&& pred) {
    bool interp = true;
    while (b - a > 1) {
      int64_t m = a + (b - a) / 2;
      if (interp) {
        const float frac = (level - v_a) / (v_b - v_a);
        if (frac > 0.0f && frac < 1.0f) {
          m = a + static_cast<int64_t>(frac * static_cast<float>(b - a));
          m = vmin(vmax(m, a + 1), b - 1);
        }
      }
      const int64_t width = b - a;
      float v_m;
      ++n_leaves;
      if (pred(m, v_m)) {
        b = m;
        v_b = v_m;
      } else {
        a = m;
        v_a = v_m;
      }
      interp = (b - a) * 2 <= width;  // bisect next when interpolation stalled
    }
    return b;
  }

  BSX_HD BSX_INLINE bool transition_search(int64_t lo, int64_t hi, Hit& out, uint32_t* stack, int stride) {
    if (n3_ref != 0.0f || !(leak >= 0.0f)) return false;
    bool up = true, down = true;
    if (tb != kNoBreakpoint) {
      for (int g = 0; g < kMaxG; ++g) {
        if (g < G) {
          if (vb[g] < va[g]) up = false;
          if (vb[g] > va[g]) down = false;
        }
      }
    }
    if (!up && !down) return false;
    float x_lo;
    const float S_lo = s_at(lo, x_lo);
    const float dl_lo = dl_at(lo);
    const float D_lo = S_lo + dl_lo;
    ++n_leaves;
    out.t = lo;
    out.x = x_lo;
    out.found = D_lo >= theta_on || D_lo <= -theta_off;
    out.on = D_lo >= theta_on;
    if (out.found) return true;
    if (hi - lo == 1) return true;
    const int64_t c = hi - 1;
    float x_c;
    const float S_c = s_at(c, x_c);
    ++n_leaves;
    if (up || leak == 0.0f) {
      const float D_c = S_c + dl_at(c);
      const bool on = up;
      const float level = on ? theta_on : -theta_off;
      if (on ? !(D_c >= theta_on) : !(D_c <= -theta_off)) return true;  // no crossing
      const int64_t t = first_true(lo, D_lo, c, D_c, level, [&](int64_t m, float& v) {
        float xm;
        v = s_at(m, xm) + dl_at(m);
        return on ? v >= theta_on : v <= -theta_off;
      });
      out.found = true;
      out.on = on;
      out.t = t;
      out.x = log_ratio(photocurrent(t));
      return true;
    }
    // Falling I under leak: OFF only, unless the upper envelope reaches theta_on somewhere.
    const float dl_c = dl_at(c);
    if (S_lo + dl_c >= theta_on) return two_sided_search(lo, c, out, stack, stride);
    const float L_c = S_c + dl_lo;
    if (!(L_c <= -theta_off)) return true;  // the lower envelope stays above: no crossing
    const int64_t tL = first_true(lo, D_lo, c, L_c, -theta_off, [&](int64_t m, float& v) {
      float xm;
      v = s_at(m, xm) + dl_lo;
      return v <= -theta_off;
    });
    int64_t tU = c + 1;
    const float U_c = S_c + dl_c;
    if (U_c <= -theta_off) {
      float x_l;
      const float U_l = s_at(tL, x_l) + dl_c;
      ++n_leaves;
      tU = U_l <= -theta_off ? tL : first_true(tL, U_l, c, U_c, -theta_off, [&](int64_t m, float& v) {
        float xm;
        v = s_at(m, xm) + dl_c;
        return v <= -theta_off;
      });
    }
    // Nodes [tL + a, tL + c'] (offsets: a piece spans less than 2^23 ns, kSearchDepth levels); the
    // pending right siblings are kept as their right ends, a popped node starting after the last one.
    const int64_t last = vmin(tU, c);
    uint32_t a = 0, e = static_cast<uint32_t>(last - tL);
    int sp = 0;
    while (true) {
      float x_e;
      const float S_e = s_at(tL + e, x_e);
      const float dl_a = dl_at(tL + a), dl_e = dl_at(tL + e);
      ++n_leaves;
      if (S_e + dl_a <= -theta_off) {
        if (dl_a == dl_e) {  // Delta non-increasing on the node, and Delta(e) = S_e + dl_e fires
          int64_t t = tL + a;
          if (a != e) {
            float x_a;
            const float D_a = s_at(t, x_a) + dl_a;
            ++n_leaves;
            if (!(D_a <= -theta_off))
              t = first_true(t, D_a, tL + e, S_e + dl_e, -theta_off, [&](int64_t m, float& v) {
                float xm;
                v = s_at(m, xm) + dl_at(m);
                return v <= -theta_off;
              });
          }
          out.found = true;
          out.on = false;
          out.t = t;
          out.x = log_ratio(photocurrent(t));
          return true;
        }
        const uint32_t m = a + (e - a) / 2;  // e > a: dl differs at the ends
        stack[sp * stride] = e;
        ++sp;
        e = m;
        continue;
      }
      if (sp == 0) break;
      --sp;
      a = e + 1;
      e = stack[sp * stride];
    }
    return true;  // tU = c + 1: no crossing (a finite tU always holds one)
  }

  // Falling I under leak with both levels within reach: the first t of [lo, c] where Delta fires, by a
  // time-ordered bisection on the exact bounds fl(S(b) + dl(a)) <= Delta <= fl(S(a) + dl(b)) of a
  // node [a, b] (S does not increase, dl does not decrease). A node whose bounds stay inside both
  // levels is dropped. On a node of constant dl, Delta does not increase: it fires at a, or first
  // crosses -theta_off where the monotone search finds it, or not at all; on a node of constant S it
  // does not decrease, and the same holds with theta_on. Other nodes are split. Delta(lo) is known
  // not to fire.
  BSX_HD BSX_INLINE bool two_sided_search(int64_t lo, int64_t c, Hit& out, uint32_t* stack, int stride) {
    uint32_t a = 0, e = static_cast<uint32_t>(c - lo);
    int sp = 0;
    out.found = false;
    while (true) {
      float x_a, x_e;
      const float S_a = s_at(lo + a, x_a), S_e = s_at(lo + e, x_e);
      const float dl_a = dl_at(lo + a), dl_e = dl_at(lo + e);
      n_leaves += 2;
      const bool may = (S_a + dl_e >= theta_on) | (S_e + dl_a <= -theta_off);
      if (may) {
        const bool dl_flat = dl_a == dl_e, s_flat = S_a == S_e;
        if (dl_flat || s_flat) {
          const float D_a = S_a + dl_a, D_e = S_e + dl_e;
          int64_t t = -1;
          bool on = false;
          if (D_a >= theta_on || D_a <= -theta_off) {
            t = lo + a;
            on = D_a >= theta_on;
          } else if (dl_flat ? D_e <= -theta_off : D_e >= theta_on) {
            on = !dl_flat;
            const float level = on ? theta_on : -theta_off;
            t = first_true(lo + a, D_a, lo + e, D_e, level, [&](int64_t m, float& v) {
              float xm;
              v = s_at(m, xm) + dl_at(m);
              return on ? v >= theta_on : v <= -theta_off;
            });
          }
          if (t >= 0) {
            out.found = true;
            out.on = on;
            out.t = t;
            out.x = log_ratio(photocurrent(t));
            return true;
          }
        } else {
          const uint32_t m = a + (e - a) / 2;  // e > a: dl differs at the ends
          stack[sp * stride] = e;
          ++sp;
          e = m;
          continue;
        }
      }
      if (sp == 0) break;
      --sp;
      a = e + 1;
      e = stack[sp * stride];
    }
    return true;
  }

  // First integer t in [lo, hi) of the current segment where Delta fires: time-ordered depth-first
  // bisection of (lo, hi - 1]. The pending right siblings are kept as their right ends only (offsets
  // from lo: a piece spans less than 2^23 ns), in the lane's column of a stack with `stride` entries
  // per level (kSearchDepth levels): in time order a popped node always starts where the last
  // finished one ended, and its I at the right end is recomputed, as photocurrent is a pure function.
  static constexpr int kSearchDepth = 24;

  BSX_HD BSX_INLINE Hit search(int64_t lo, int64_t hi, uint32_t* stack, int stride) {
    const float I_lo = photocurrent(lo);
    Hit out{false, false, lo, log_ratio(I_lo)};
    if (fires(lo, out)) return out;
    if (hi - lo == 1) return out;
    uint32_t a = 0, c = static_cast<uint32_t>(hi - 1 - lo);
    float Ia = I_lo, Ic = photocurrent(hi - 1);
    int sp = 0;
    while (true) {
      if (c - a == 1) {
        out.t = lo + c;
        out.x = log_ratio(Ic);
        if (fires(lo + c, out)) return out;
      } else {
        ++n_leaves;
        if (may_fire(lo + a, lo + c, Ia, Ic)) {
          const uint32_t m = a + (c - a) / 2;
          stack[sp * stride] = c;
          ++sp;
          c = m;
          Ic = photocurrent(lo + m);
          continue;
        }
      }
      if (sp == 0) break;
      --sp;
      a = c;
      Ia = Ic;
      c = stack[sp * stride];
      Ic = photocurrent(lo + c);
    }
    out.found = false;
    return out;
  }

  BSX_HD BSX_INLINE void set_reference(int64_t t, float s3, float n3, float offset) {
    t_ref = t;
    s3_ref = s3;
    n3_ref = n3;
    dr_ref = offset;
    scr_valid = false;
  }

  // ---- The photocurrent screen ------------------------------------------------------------------
  // Between two reference changes, Delta(t) = fl(fl(fl(x(t) - s3_ref) + (0 - n3_ref)) + dl(t)) with
  // x(t) = pm::log(fl(I(t) / i_scale)) is non-decreasing in the photocurrent I(t) (every step is a
  // correctly rounded monotone operation, and pm::log is monotone on every positive float) and, with
  // leak >= 0, in t through dl. Its fp32 value lies within 2^-13 of the real-valued
  // ln(I / i_scale) - s3_ref + dl(t) (|x|, |s3_ref| and |dl| stay below 64: photocurrents lie in
  // [dark, 1e9 fA] and a reference is an x value; a handful of roundings of 2^-18 at most, and the log's
  // own error). So on [t_from, t_end) no integer fires while I(t) stays above
  //   scr_lo = i_scale exp(-theta_off + s3_ref - dl(t_from)) (1 + 2^-9)
  // and below
  //   scr_hi = i_scale exp(theta_on + s3_ref - dl(t_end - 1)) (1 - 2^-9),
  // a margin of 2^-9 in log space against an error below 2^-13 (the exp's own error is far smaller).
  // A segment's computed I(t) lies within a relative 2^-20 of the range of its end values, so the
  // test widens that range by 2^-12. It needs no logarithm: a segment that passes it is decided
  // without the exact search, whose result (no crossing) it is.
  float scr_lo = 0.0f, scr_hi = 0.0f;
  bool scr_valid = false;
  bool screen = true;  // the screen on (off: every piece takes the exact search, for the tests)

  BSX_HD BSX_INLINE void screen_bounds(int64_t t_from, int64_t t_last) {
    const float dl_lo = dl_at(t_from), dl_hi = dl_at(t_last);
    scr_hi = i_scale * pm::exp(theta_on + s3_ref - dl_hi) * (1.0f - 0x1p-9f);
    scr_lo = i_scale * pm::exp(-theta_off + s3_ref - dl_lo) * (1.0f + 0x1p-9f);
    scr_valid = true;
  }

  // The range of the current segment's end photocurrents, lo <= I(t) <= hi up to the rounding the
  // screen allows for.
  BSX_HD BSX_INLINE void segment_range(float& lo, float& hi) const {
    lo = dark;
    hi = dark;
    const bool held = tb == kNoBreakpoint;
    for (int g = 0; g < kMaxG; ++g) {
      if (g < G) {
        const float b = held ? va[g] : vb[g];
        lo = lo + scale * vmin(va[g], b);
        hi = hi + scale * vmax(va[g], b);
      }
    }
  }

  // True when the photocurrent range [lo, hi], widened by 2^-12, lies strictly inside (scr_lo, scr_hi).
  BSX_HD BSX_INLINE bool in_screen(float lo, float hi) const {
    return (hi * (1.0f + 0x1p-12f) < scr_hi) & (lo * (1.0f - 0x1p-12f) > scr_lo);
  }

  // True when I(t) over the current segment stays strictly inside (scr_lo, scr_hi).
  BSX_HD BSX_INLINE bool screened() const {
    float lo, hi;
    segment_range(lo, hi);
    return in_screen(lo, hi);
  }

  // ---- Steady light -------------------------------------------------------------------------------
  // When every group's segment ends are equal (or the segment is the held one), photocurrent(t) is the
  // one value I = photocurrent_start() at every t ((vb - va) f is +0), so x(t) = log_ratio(I) is one
  // value and Delta(t) = fl(fl(x - s3_ref) + (0 - n3_ref)) + dl(t)) is non-decreasing in t (dl is,
  // with leak >= 0). Then no integer of [lo, e) fires if and only if Delta(lo) > -theta_off and
  // Delta(e - 1) < theta_on: the exact answer of the search, from one logarithm, which is kept while
  // the light and i_scale stay the same. It decides the pixels whose light is constant and whose
  // Delta sits within the screen's margin of a level (a static scene seen near a threshold).
  float st_I = -1.0f, st_scale = 0.0f, st_x = 0.0f;

  BSX_HD BSX_INLINE bool steady_segment() const {
    if (tb == kNoBreakpoint) return true;
    bool eq = true;
    for (int g = 0; g < kMaxG; ++g)
      if (g < G) eq &= va[g] == vb[g];
    return eq;
  }

  BSX_HD BSX_INLINE float steady_x(float I) {
    if (!(I == st_I && i_scale == st_scale)) {
      st_x = log_ratio(I);
      st_I = I;
      st_scale = i_scale;
    }
    return st_x;
  }

  // Under steady light I over [lo, last] (lo <= last): true when no integer of it fires.
  BSX_HD BSX_INLINE bool steady_quiet(float I, int64_t lo, int64_t last) {
    const float x = steady_x(I);
    return (delta_at(last, x) < theta_on) & (delta_at(lo, x) > -theta_off);
  }

  BSX_HD BSX_INLINE static uint8_t split_label(float ds, float dl, float D, float cut) {
    const float a_s = ds / D, a_l = dl / D;
    if (a_s >= cut && a_s >= a_l) return kLabelSignal;
    if (a_l >= cut && a_l > a_s) return kLabelLeak;
    return kLabelNoise;
  }

  // shot: an event of the shot-noise process (labelled NOISE, flagged kFlagShot, reset as any event).
  template <class Sink>
  BSX_HD BSX_INLINE void fire(int64_t t, bool on, float s3, float n3, Sink& sink, bool shot = false) {
    const float ds = s3 - s3_ref;
    const float dn = n3 - n3_ref;
    const float l = leak * ns_to_s64<float>(t - t_ref);
    const float dl = l + dr_ref;
    const float D = ds + dn + dl;
    const bool first = (flags & kFlagFirstAfterInit) != 0;
    uint8_t label;
    if (shot) {
      label = kLabelNoise;
    } else if (first) {
      const float Dl = ds + dn + l;
      label = (on ? Dl > 0.0f : Dl < 0.0f) ? split_label(ds, l, Dl, label_cut) : kLabelNoise;
    } else {
      label = split_label(ds, dl, D, label_cut);
    }
    if (hot && label != kLabelSignal) label = kLabelHot;
    RawEventDev ev;
    ev.pixel = pixel;
    ev.polarity = on ? 1 : 0;
    ev.label = label;
    ev.flags = static_cast<uint16_t>(flags | (shot ? kFlagShot : 0u) | (hot ? kFlagHotPixel : 0u));
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

  // The fast part of one piece [t, e) of the current segment (the body of advance_instant's piece
  // loop up to its exact search): releases, dead time, and the photocurrent screen. Returns true when
  // the piece is finished, false when the exact search of [s_lo, e) must run next (search_step).
  template <class Sink>
  BSX_HD BSX_INLINE bool piece_fast(int64_t t, int64_t e, Sink& sink) {
    while (true) {
      if (blind_until != kNotBlind) {
        if (blind_until >= e) return true;
        set_reference(blind_until, log_ratio(photocurrent(blind_until)), 0.0f, -eps_ci);
        blind_until = kNotBlind;
      }
      const int64_t lo = vmax(t_search, t);
      if (lo >= e) return true;
      if (screen && n3_ref == 0.0f && leak >= 0.0f) {
        if (!scr_valid) screen_bounds(lo, t_end_ - 1);
        if (screened() || (steady_segment() && steady_quiet(photocurrent_start(), lo, e - 1))) {
          t_search = vmax(t_search, e);
          return true;
        }
      }
      s_lo = lo;
      (void)sink;
      return false;
    }
  }

  BSX_HD BSX_INLINE void load_sample(int i, int64_t& t, float* v) const {
    t = ts[i];
    for (int g = 0; g < kMaxG; ++g)
      if (g < G) v[g] = clampv(irr[static_cast<int64_t>(i) * G + g]);
  }

  // ---- Shot noise ---------------------------------------------------------------------------------
  // Per pixel a Poisson process of rate r: arrival k follows arrival k - 1 (or the init time) after
  // ceil(-ln(U) / r * 1e9) ns (at least 1), U the open uniform of word 0 of
  // Philox(seed, kPurposeShot, camera, pixel, k), whose word 1 gives the polarity (low bit). The
  // arithmetic is fp64 with correctly rounded operations and the library's log, so host and device
  // draw the same schedule.
  BSX_HD BSX_INLINE static int64_t shot_gap(double rate, uint64_t seed, uint32_t camera, uint32_t pixel,
                                            uint64_t index, bool* on) {
    const Word4 w = draw(seed, kPurposeShot, camera, pixel, index);
    const double e = -pm::log(open_uniform64(w.v[0]));
    const double ns = __builtin_ceil(e / rate * 1e9);
    if (on != nullptr) *on = (w.v[1] & 1u) != 0;
    if (!(ns >= 1.0)) return 1;
    if (ns >= 0x1p62) return int64_t{1} << 62;
    return static_cast<int64_t>(ns);
  }

  // The arrival at T = shot_next (T inside the current segment): an event unless T falls in the
  // dead time (T < t_search, which covers the physical blind interval), then the next arrival.
  template <class Sink>
  BSX_HD BSX_INLINE void shot_arrival(int64_t T, Sink& sink) {
    bool on = false;
    (void)shot_gap(shot_rate, seed, camera, pixel, shot_index, &on);
    if (T >= t_search) fire(T, on, log_ratio(photocurrent(T)), 0.0f, sink, true);
    shot_index += 1;
    shot_next = T + shot_gap(shot_rate, seed, camera, pixel, shot_index, nullptr);
  }

  // ---- The call as a resumable lane program --------------------------------------------------
  // advance_instant walks the segments of the augmented samples in order, cutting each into pieces at
  // multiples of 2^23 ns. Here that walk is split at its exact searches: walk() runs until the next
  // exact search is due (kSearch), the call is finished (kDone) or `budget` segments are done
  // (kWalk); search_step() runs the due search and its event. Run in that alternation, each lane
  // performs exactly advance_instant's operations in its order, so a scheduler may interleave the
  // lanes of a wave freely (the kernel gathers the searches of many lanes into one pass).
  static constexpr int kWalk = 0, kSearch = 1, kDone = 2;

  // The segment ends are loaded kAhead samples ahead of the walk (raw: clamped when a segment opens),
  // so that a lane's walk through quiet segments does not wait on one load at a time.
  static constexpr int kAhead = 4;
  int next_i;              // the input sample that ends the next segment to open
  int64_t ft[kAhead];      // ts[next_i + k] (kNoBreakpoint past the run)
  float fv[kAhead][kMaxG];
  bool seg_open, in_piece;
  int64_t t_cur, stop, t_end_, s_lo, e_cur;
  // Phase A's photocurrent ranges of aligned blocks of kScreenBlock segments (null: none), and the
  // camera position of the run's first sample.
  const BSX_GLOBAL ScreenBlock* qb = nullptr;
  int64_t q_base = 0;
  uint32_t n_block_skips = 0;  // blocks passed at once (a diagnostic; no reference counterpart)

  BSX_HD BSX_INLINE void fetch(int k, int i) {
    ft[k] = kNoBreakpoint;
    if (i < n) {
      ft[k] = ts[i];
      for (int g = 0; g < kMaxG; ++g)
        if (g < G) fv[k][g] = irr[static_cast<int64_t>(i) * G + g];
    }
  }

  // The next segment [ta, tb): its end from the head of the prefetch ring (kNoBreakpoint: the held
  // segment after the last sample), and the ring moved one sample on.
  BSX_HD BSX_INLINE void open_next() {
    tb = ft[0];
    for (int g = 0; g < kMaxG; ++g)
      if (g < G) vb[g] = clampv(fv[0][g]);
    for (int k = 0; k + 1 < kAhead; ++k) {
      ft[k] = ft[k + 1];
      for (int g = 0; g < kMaxG; ++g)
        if (g < G) fv[k][g] = fv[k + 1][g];
    }
    ++next_i;
    fetch(kAhead - 1, next_i + kAhead - 1);
  }

  // At a segment boundary, kScreenBlock segments at once: when the next segment starts an aligned block
  // of the run whose photocurrent range (phase A's summary, the extremes of the segments' own ranges)
  // passes the current screen, with no release pending, no shot arrival up to the block's last sample
  // and no 2^23-ns cut inside it, each of its segments would be screened in turn (or lie in dead time)
  // and leave exactly this: one sub-step and one piece each, t_search at least the block's end, and
  // the block's last sample as the next segment's start. The bounds, when formed here, are formed from
  // max(t_search, t_cur), as the block's first piece would form them; bounds formed earlier are valid
  // for later times too (dl does not decrease).
  // Blocks are tried kSkipAhead at a time: their summaries, last sample times and last values are
  // loaded together, and then passed in order while each passes, so that a lane crossing quiet blocks
  // waits on one round of loads per kSkipAhead blocks rather than on two dependent loads per block.
  // Each block is decided as alone: the checks of block j + 1 see the state block j left. (Four at a
  // time spilled 64 bytes a lane against 32 and ran no faster.)
  static constexpr int kSkipAhead = 2;

  // Returns the number of blocks passed (0: none).
  BSX_HD BSX_INLINE int try_block_skip() {
    if (qb == nullptr || !screen || blind_until != kNotBlind || n3_ref != 0.0f || !(leak >= 0.0f)) return 0;
    const int64_t q = q_base + next_i;
    if ((q & (kScreenBlock - 1)) != 0 || next_i + kScreenBlock > n) return 0;
    int64_t t_lasts[kSkipAhead];
    ScreenBlock blocks[kSkipAhead];
    float v_lasts[kSkipAhead][kMaxG];
    int m = 0;  // blocks inside the run
    for (int j = 0; j < kSkipAhead; ++j) {
      if (next_i + (j + 1) * kScreenBlock <= n) {
        const int last = next_i + (j + 1) * kScreenBlock - 1;
        t_lasts[j] = ts[last];
        blocks[j] = qb[(q >> kScreenBlockLog2) + j];
        for (int g = 0; g < kMaxG; ++g)
          if (g < G) v_lasts[j][g] = irr[static_cast<int64_t>(last) * G + g];
        m = j + 1;
      }
    }
    int passed = 0;
    for (int j = 0; j < kSkipAhead; ++j) {
      if (j >= m || (j > 0 && !(t_cur < t_end_))) break;
      const int64_t t_last = t_lasts[j];
      if (!(shot_next > t_last) || ((t_cur >> kPlainSegmentLog2) + 1) * kPlainSegmentNs < t_last) break;
      const ScreenBlock b = blocks[j];
      const int64_t from = vmax(t_search, t_cur);
      if (!scr_valid) screen_bounds(from, t_end_ - 1);
      // a steady block (the light b.lo at every time of it, the current start value's) is decided
      // exactly by its two ends; a block wholly in dead time fires nothing
      if (!in_screen(b.lo, b.hi) && !(b.steady != 0u && (from >= t_last || steady_quiet(b.lo, from, t_last - 1))))
        break;
      n_substeps += kScreenBlock;
      n_pieces += kScreenBlock;
      ++n_block_skips;
      t_search = vmax(t_search, t_last);
      ta = t_last;
      for (int g = 0; g < kMaxG; ++g)
        if (g < G) va[g] = clampv(v_lasts[j][g]);
      t_cur = t_last;
      next_i += kScreenBlock;
      ++passed;
    }
    if (passed > 0)
      for (int k = 0; k < kAhead; ++k) fetch(k, next_i + k);
    return passed;
  }

  // first_loud: the index of the first input sample ending a segment that phase A did not find quiet
  // under the input state (n: none). The walk starts at that segment, with the quiet ones before it
  // left as the walk would leave them (t_search, the piece counts; phase A's screens are not counted
  // as leaf intervals).
  BSX_HD BSX_INLINE void begin(int64_t t_end, int first_loud = 0) {
    t_end_ = t_end;
    next_i = skip;
    ta = t_state;
    for (int g = 0; g < kMaxG; ++g)
      if (g < G) va[g] = held[g];
    if (skip) load_sample(0, ta, va);
    if (first_loud > skip) {
      const uint32_t quiet = static_cast<uint32_t>(first_loud - skip);
      n_substeps += quiet;
      n_pieces += quiet;
      next_i = first_loud;
      load_sample(first_loud - 1, ta, va);
      t_search = vmax(t_search, ta);
    }
    for (int k = 0; k < kAhead; ++k) fetch(k, next_i + k);
    t_cur = ta;
    seg_open = false;
    in_piece = false;
    tb = kNoBreakpoint;
  }

  template <class Sink>
  BSX_HD BSX_INLINE int walk(Sink& sink, int budget) {
    while (true) {
      if (!seg_open) {
        if (!(t_cur < t_end_)) return kDone;
        if (const int passed = try_block_skip()) {
          budget -= passed;
          if (budget <= 0) return kWalk;
          continue;
        }
        open_next();
        stop = vmin(t_end_, tb);
        seg_open = true;
      }
      if (t_cur < stop) {
        if (!in_piece && shot_next == t_cur) {
          shot_arrival(t_cur, sink);
          continue;
        }
        e_cur = vmin(vmin(stop, ((t_cur >> kPlainSegmentLog2) + 1) * kPlainSegmentNs), shot_next);
        if (!in_piece) {
          ++n_substeps;
          ++n_pieces;
          in_piece = true;
        }
        if (!piece_fast(t_cur, e_cur, sink)) return kSearch;
        t_cur = e_cur;
        in_piece = false;
        continue;
      }
      if (tb == kNoBreakpoint) return kDone;
      ta = tb;
      for (int g = 0; g < kMaxG; ++g)
        if (g < G) va[g] = vb[g];
      seg_open = false;
      if (--budget <= 0) return kWalk;
    }
  }

  // The exact search of [s_lo, e_cur) that walk() stopped at; its event (if any) keeps the piece
  // open, as the piece loop continues after a crossing.
  template <class Sink>
  BSX_HD BSX_INLINE void search_step(Sink& sink, uint32_t* stack, int stride) {
    Hit hit{false, false, s_lo, 0.0f};
    if (!transition_search(s_lo, e_cur, hit, stack, stride)) hit = search(s_lo, e_cur, stack, stride);
    if (!hit.found) {
      t_search = vmax(t_search, e_cur);
      t_cur = e_cur;
      in_piece = false;
      return;
    }
    fire(hit.t, hit.on, hit.x, 0.0f, sink);
  }

  // After the call: the last input sample is held until the next call.
  BSX_HD BSX_INLINE void end() {
    if (n > 0) {
      for (int g = 0; g < kMaxG; ++g)
        if (g < G) held[g] = clampv(irr[static_cast<int64_t>(n - 1) * G + g]);
    }
  }

  // The whole call on one lane, in order.
  template <class Sink>
  BSX_HD BSX_INLINE void advance(int64_t t_end, Sink& sink) {
    uint32_t stack[kSearchDepth];
    begin(t_end);
    for (int m = walk(sink, 1 << 30); m != kDone; m = walk(sink, 1 << 30))
      if (m == kSearch) search_step(sink, stack, 1);
    end();
  }
};

}  // namespace bsx
