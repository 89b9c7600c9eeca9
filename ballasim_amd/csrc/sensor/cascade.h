// The photoreceptor cascade without noise, as a lane program: one GPU lane runs one pixel's call.
//
// Model (reference: PixelKernel::advance_slice / run_segment / make_substep / comparator_plain,
// src/sensor/pixel.h:909-945, 1455-1515, 1590-1628; stage 1, src/sensor/stage1.h:26-92; sub-steps,
// src/sensor/substep.h:43-241; the comparator input and its first firing, src/sensor/expcubic.h:33-282,
// pixel.h:322-387). Stage 1 dz/dt = k (1 - I z) is advanced in closed form on each linear piece of
// photocurrent; x1 = -ln(z i_scale) is replaced per sub-step by its cubic Hermite, accepted by a
// midpoint residual and a boundary-layer bound (dyadic halving to a 1 ns floor); stages 2 and 3
// (first-order lags tau_o, tau_sf) are solved exactly for that cubic, giving the comparator input
// s3(u) = cubic + C0 R_j(a2 u) + C1 R_j(a3 u), whose first integer-ns threshold crossing is found by
// a hull test, Polya-Rolle root isolation of the monotone pieces and an interpolation search.
//
// Every arithmetic step is the reference's, in its order: the fp32 results equal the reference's
// fp32 CPU path bit for bit. The sample cursor, photocurrent, fire and reset are InstantLane's (they
// are the same functions of the reference). Constant light only (the waveform groups are constant).
#pragma once

#include "instant.h"

namespace bsx {
namespace cascade {

// g(u) = c0 + c1 u + c2 u^2 + c3 u^3 + C0 R_j(a0 u) + C1 R_j(a1 u), u in seconds (j = order).
struct ExpCubic {
  float c[4];
  float C[2];
  float a[2];
  int order;

  BSX_HD BSX_INLINE float poly(float u) const { return c[0] + u * (c[1] + u * (c[2] + u * c[3])); }
  BSX_HD BSX_INLINE float term(int i, float u) const {
    return C[i] == 0.0f ? 0.0f : C[i] * pm::resid_exp(order, a[i] * u);
  }
  BSX_HD BSX_INLINE float operator()(float u) const { return poly(u) + term(0, u) + term(1, u); }
  BSX_HD BSX_INLINE float operator()(float u, float r0, float r1) const {
    return poly(u) + (C[0] == 0.0f ? 0.0f : C[0] * r0) + (C[1] == 0.0f ? 0.0f : C[1] * r1);
  }
};

BSX_HD BSX_INLINE ExpCubic derivative(const ExpCubic& g, int n) {
  ExpCubic d = g;
  for (int k = 0; k < n; ++k) {
    d.c[0] = d.c[1];
    d.c[1] = 2.0f * d.c[2];
    d.c[2] = 3.0f * d.c[3];
    d.c[3] = 0.0f;
    d.C[0] = -d.a[0] * d.C[0];
    d.C[1] = -d.a[1] * d.C[1];
  }
  d.order = g.order > n ? g.order - n : 0;
  return d;
}

// g(d + u) in u (residual terms re-split so that large amplitudes meet small residuals).
BSX_HD BSX_INLINE ExpCubic shifted(const ExpCubic& g, float d) {
  ExpCubic s = g;
  s.c[0] = g.c[0] + d * (g.c[1] + d * (g.c[2] + d * g.c[3]));
  s.c[1] = g.c[1] + d * (2.0f * g.c[2] + d * (3.0f * g.c[3]));
  s.c[2] = g.c[2] + d * (3.0f * g.c[3]);
  s.c[3] = g.c[3];
  for (int i = 0; i < 2; ++i) {
    if (g.C[i] == 0.0f) continue;
    const float ad = g.a[i] * d;
    s.C[i] = g.C[i] * pm::exp(-ad);
    float coef = g.C[i];
    for (int m = 0; m < g.order; ++m) {
      s.c[m] = s.c[m] + coef * pm::resid_exp(g.order - m, ad);
      coef = coef * (-g.a[i]) / static_cast<float>(m + 1);
    }
  }
  return s;
}

struct Range {
  float lo, hi;
};

// The cubic's range on [u0, u1] from its Bernstein coefficients.
BSX_HD BSX_INLINE Range cubic_range(const float* c, float u0, float u1) {
  const float w = u1 - u0;
  const float q0 = c[0] + u0 * (c[1] + u0 * (c[2] + u0 * c[3]));
  const float q1 = (c[1] + u0 * (2.0f * c[2] + u0 * (3.0f * c[3]))) * w;
  const float q2 = (c[2] + u0 * (3.0f * c[3])) * (w * w);
  const float q3 = c[3] * (w * w * w);
  const float b0 = q0;
  const float b1 = q0 + q1 / 3.0f;
  const float b2 = q0 + (2.0f * q1 + q2) / 3.0f;
  const float b3 = q0 + q1 + q2 + q3;
  return {vmin(vmin(b0, b1), vmin(b2, b3)), vmax(vmax(b0, b1), vmax(b2, b3))};
}

// Conservative range of g: the cubic's hull plus each monotone exponential term's end values.
BSX_HD BSX_INLINE Range range(const ExpCubic& g, float u0, float u1) {
  Range r = cubic_range(g.c, u0, u1);
  for (int i = 0; i < 2; ++i) {
    if (g.C[i] == 0.0f) continue;
    const float e0 = g.term(i, u0), e1 = g.term(i, u1);
    r.lo = r.lo + vmin(e0, e1);
    r.hi = r.hi + vmax(e0, e1);
  }
  return r;
}

struct RootList {
  float v[6];
  int n;
};

BSX_HD BSX_INLINE float root_tolerance(float hi) { return vmax(1e-11f, 16.0f * 0x1p-23f * fabs_(hi)); }

// ITP root of f on [a, b] with f(a) f(b) < 0 (Oliveira and Takahashi 2020; kappa1 = 0.2 / (b - a),
// kappa2 = 2, n0 = 1).
BSX_HD BSX_INLINE float itp_root(const ExpCubic& f, float a, float b, float fa, float fb, float tol) {
  const float width0 = b - a;
  const float k1 = 0.2f / width0;
  int n_half = 0;
  for (float w = width0; w > 2.0f * tol && n_half < 200; w = w * 0.5f) ++n_half;
  const int n_max = n_half + 1;
  float scale = tol;
  for (int j = 0; j < n_max; ++j) scale = scale * 2.0f;
  for (int j = 0; j < 400 && b - a > 2.0f * tol; ++j) {
    const float xh = 0.5f * (a + b);
    const float r = scale - 0.5f * (b - a);
    const float xf = (fb * a - fa * b) / (fb - fa);
    const float sg = xh - xf >= 0.0f ? 1.0f : -1.0f;
    const float delta = k1 * (b - a) * (b - a);
    const float xt = delta <= fabs_(xh - xf) ? xf + sg * delta : xh;
    const float xi = fabs_(xt - xh) <= r ? xt : xh - sg * r;
    if (!(xi > a && xi < b)) break;
    const float yi = f(xi);
    if (yi == 0.0f) return xi;
    if ((yi > 0.0f) == (fa > 0.0f)) {
      a = xi;
      fa = yi;
    } else {
      b = xi;
      fb = yi;
    }
    scale = scale * 0.5f;
  }
  return 0.5f * (a + b);
}

// The zero of the fourth derivative (a sum of two exponentials) in (lo, hi), in log space.
BSX_HD BSX_INLINE void fourth_derivative_root(const ExpCubic& g, float lo, float hi, RootList& out) {
  if (g.C[0] == 0.0f || g.C[1] == 0.0f) return;
  if ((g.C[0] > 0.0f) == (g.C[1] > 0.0f)) return;
  if (g.a[0] == g.a[1]) return;
  const float ln_ratio = pm::log(fabs_(g.C[1]) / fabs_(g.C[0])) + 4.0f * pm::log(g.a[1] / g.a[0]);
  const float u = ln_ratio / (g.a[1] - g.a[0]);
  if (u > lo && u < hi) out.v[out.n++] = u;
}

// Sign-change zeros of g^(N) in (lo, hi), sorted: the Polya-Rolle recursion.
template <int N>
BSX_HD BSX_INLINE void level_roots(const ExpCubic& g, float lo, float hi, float tol, RootList& out) {
  if constexpr (N >= 4) {
    fourth_derivative_root(g, lo, hi, out);
  } else {
    const ExpCubic f = derivative(g, N);
    const ExpCubic fd = derivative(g, N + 1);
    RootList pts;
    pts.n = 0;
    pts.v[pts.n++] = lo;
    const Range rd = range(fd, lo, hi);
    if (rd.lo < 0.0f && rd.hi > 0.0f) {
      RootList crit;
      crit.n = 0;
      level_roots<N + 1>(g, lo, hi, tol, crit);
      for (int i = 0; i < crit.n; ++i) pts.v[pts.n++] = crit.v[i];
    }
    pts.v[pts.n++] = hi;
    float fa = f(pts.v[0]);
    for (int i = 0; i + 1 < pts.n; ++i) {
      const float a = pts.v[i], b = pts.v[i + 1];
      const float fb = f(b);
      if ((fa < 0.0f && fb > 0.0f) || (fa > 0.0f && fb < 0.0f)) out.v[out.n++] = itp_root(f, a, b, fa, fb, tol);
      fa = fb;
    }
  }
}

BSX_HD BSX_INLINE RootList critical_points(const ExpCubic& g, float lo, float hi) {
  RootList crit;
  crit.n = 0;
  level_roots<1>(g, lo, hi, root_tolerance(hi), crit);
  return crit;
}

// First integer k in [k_lo, k_hi] with G(k 1e-9) >= 0, or -1; G monotone between the breakpoints.
BSX_HD BSX_INLINE int64_t first_firing(const ExpCubic& G, const RootList& crit, int64_t k_lo, int64_t k_hi) {
  if (k_lo > k_hi) return -1;
  const float kNs = 1e-9f;
  int64_t piece_lo = k_lo;
  for (int i = 0; i <= crit.n && piece_lo <= k_hi; ++i) {
    const int64_t piece_hi =
        i < crit.n ? vmin(static_cast<int64_t>(floor_(crit.v[i] / kNs)), k_hi) : k_hi;
    if (piece_hi < piece_lo) continue;
    int64_t a = piece_lo, b = piece_hi;
    piece_lo = piece_hi + 1;
    float ga = G(static_cast<float>(a) * kNs);
    if (ga >= 0.0f) return a;
    if (b == a) continue;
    float gb = G(static_cast<float>(b) * kNs);
    if (gb >= 0.0f) {
      bool interpolate = true;
      while (b - a > 1) {
        int64_t m;
        if (interpolate) {
          const float frac = (-ga) / (gb - ga);
          m = a + static_cast<int64_t>(frac * static_cast<float>(b - a));
          m = vmin(vmax(m, a + 1), b - 1);
        } else {
          m = a + (b - a) / 2;
        }
        const int64_t width = b - a;
        const float gm = G(static_cast<float>(m) * kNs);
        if (gm >= 0.0f) {
          b = m;
          gb = gm;
        } else {
          a = m;
          ga = gm;
        }
        interpolate = (b - a) * 2 <= width;
      }
      return b;
    }
  }
  return -1;
}

struct Hit {
  bool found, on;
  int64_t t;
};

// Can Delta, on the integers [max(lo, e0), e1), reach a level? (the search's hull test)
BSX_HD BSX_INLINE bool may_cross(const ExpCubic& delta, int64_t e0, int64_t e1, int64_t lo, float theta_on,
                                 float theta_off) {
  const int64_t k_lo = vmax(lo, e0) - e0, k_hi = e1 - 1 - e0;
  if (k_lo > k_hi) return false;
  const Range r = range(delta, static_cast<float>(k_lo) * 1e-9f, static_cast<float>(k_hi) * 1e-9f);
  return r.hi >= theta_on || r.lo <= -theta_off;
}

// The first crossing of Delta on [max(lo, e0), e1) (the reference's search after its hull test).
BSX_HD BSX_INLINE Hit search(const ExpCubic& delta, int64_t e0, int64_t e1, int64_t lo, float theta_on,
                               float theta_off) {
  const Hit none{false, false, 0};
  const int64_t k_lo = vmax(lo, e0) - e0, k_hi = e1 - 1 - e0;
  if (k_lo > k_hi) return none;
  const float u_lo = static_cast<float>(k_lo) * 1e-9f, u_hi = static_cast<float>(k_hi) * 1e-9f;
  const Range r = range(delta, u_lo, u_hi);
  const bool on = r.hi >= theta_on, off = r.lo <= -theta_off;
  if (!(on || off)) return none;
  ExpCubic gon = delta, goff = delta;
  gon.c[0] = gon.c[0] - theta_on;
  for (int m = 0; m < 4; ++m) goff.c[m] = -goff.c[m];
  goff.C[0] = -goff.C[0];
  goff.C[1] = -goff.C[1];
  goff.c[0] = goff.c[0] - theta_off;
  RootList crit;
  crit.n = 0;
  if (u_hi > u_lo) crit = critical_points(delta, u_lo, u_hi);
  const int64_t k_on = on ? first_firing(gon, crit, k_lo, k_hi) : -1;
  const int64_t k_off = off ? first_firing(goff, crit, k_lo, k_hi) : -1;
  if (k_on < 0 && k_off < 0) return none;
  if (k_off < 0 || (k_on >= 0 && k_on <= k_off)) return {true, true, e0 + k_on};
  return {true, false, e0 + k_off};
}

// ---- Stage 1 -------------------------------------------------------------------------------------

BSX_HD BSX_INLINE float integrated_pole(float I0, float I1, float h, float k) { return k * h * (0.5f * (I0 + I1)); }

BSX_HD BSX_INLINE float gl8_pair(float kh, float I0, float dI, double xi, double w) {
  const float u_lo = static_cast<float>(0.5 - 0.5 * xi), u_hi = static_cast<float>(0.5 + 0.5 * xi);
  const float g_lo = kh * (1.0f - u_lo) * (I0 + 0.5f * dI * (1.0f + u_lo));
  const float g_hi = kh * (1.0f - u_hi) * (I0 + 0.5f * dI * (1.0f + u_hi));
  return static_cast<float>(w) * (pm::exp(-g_lo) + pm::exp(-g_hi));
}

// z after h from z0 for I linear I0 -> I1: GL8 (P <= 1/2, with the fp32 small-slope expansion),
// zero-order hold (slope below one ulp), Dawson (rising), erfcx (falling).
BSX_HD BSX_INLINE float z_step(float z0, float I0, float I1, float h, float k) {
  const float kh = k * h;
  const float Im = 0.5f * (I0 + I1);
  const float P = kh * Im;
  const float eP = pm::exp(-P);
  const float dI = I1 - I0;
  if (P <= 0.5f) {
    if (P > 0.0f && P * fabs_(dI) <= 2.4e-4f * Im) {
      const float x = kh * I1;
      const float b = 0.5f * k * dI / h;
      const float one_m_E = -pm::expm1(-x);
      const float g = 1.0f / 3.0f + x * (-0.25f + x * (0.1f + x * (-1.0f / 36.0f + x * (1.0f / 168.0f))));
      const float J = one_m_E / x * h + b * h * h * h * g;
      return z0 * eP + k * J;
    }
    float J = gl8_pair(kh, I0, dI, tables::kGl8Xi[0], tables::kGl8W[0]);
    J = J + gl8_pair(kh, I0, dI, tables::kGl8Xi[1], tables::kGl8W[1]);
    J = J + gl8_pair(kh, I0, dI, tables::kGl8Xi[2], tables::kGl8W[2]);
    J = J + gl8_pair(kh, I0, dI, tables::kGl8Xi[3], tables::kGl8W[3]);
    return z0 * eP + kh * J;
  }
  if (fabs_(dI) <= 0x1p-23f * Im) return 1.0f / Im + (z0 - 1.0f / Im) * eP;
  const float s = fabs_(dI) / h;
  const float r = sqrt_(k / (2.0f * s));
  const float k_over_beta = sqrt_(2.0f * k / s);
  if (dI > 0.0f) return z0 * eP + k_over_beta * (pm::dawson(r * I1) - pm::dawson(r * I0) * eP);
  constexpr double kHalfSqrtPi = 0x1.c5bf891b4ef6ap-1;  // sqrt(pi) / 2
  return z0 * eP + k_over_beta * static_cast<float>(kHalfSqrtPi) * (pm::erfcx(r * I1) - pm::erfcx(r * I0) * eP);
}

// ---- Sub-steps -----------------------------------------------------------------------------------

struct Consts {
  float k, a2, a3, tau_o, tau_sf, eps_accept, eps_drop;
};

struct State {
  float z, s2, s3;
};

struct SubStep {
  int64_t t_a, t_b;
  ExpCubic s3;
  float x1_a, x1_b;
  float int_ku;
  bool fallback;
};

struct Piece {
  int64_t t_a, t_b;
  float I_a, I_b;
  BSX_HD BSX_INLINE float at(int64_t t) const {
    if (t == t_b) return I_b;
    return I_a + (I_b - I_a) * (static_cast<float>(t - t_a) / static_cast<float>(t_b - t_a));
  }
};

BSX_HD BSX_INLINE void particular(const float* c, float tau, float* b) {
  b[3] = c[3];
  b[2] = c[2] - 3.0f * tau * b[3];
  b[1] = c[1] - 2.0f * tau * b[2];
  b[0] = c[0] - tau * b[1];
}

BSX_HD BSX_INLINE float boundary_layer_factor(float P) {
  const float p2 = P * P;
  return vmin(p2 * p2 / 384.0f, 0.15f * P + 1.0f);
}

// Attempts the sub-step [0, h] with linear input I_a -> I_b; on success fills sub (but t_a, t_b) and
// next. at_floor forces acceptance (1 ns).
BSX_HD BSX_INLINE bool try_substep(const Consts& cc, const State& st, float i_scale, float I_a, float I_b, float h,
                                     bool at_floor, SubStep& sub, State& next) {
  const float k = cc.k;
  const float z_a = st.z;
  const float z_b = z_step(z_a, I_a, I_b, h, k);
  const float I_m = 0.5f * (I_a + I_b);
  const float z_m = z_step(z_a, I_a, I_m, 0.5f * h, k);
  const float x1a = -pm::log(z_a * i_scale), x1b = -pm::log(z_b * i_scale), x1m = -pm::log(z_m * i_scale);
  const float da = k * (I_a - 1.0f / z_a), db = k * (I_b - 1.0f / z_b);
  float c[4];
  const float delta = (x1b - x1a) / h;
  c[0] = x1a;
  c[1] = da;
  c[2] = (3.0f * delta - 2.0f * da - db) / h;
  c[3] = (da + db - 2.0f * delta) / (h * h);
  const float hh = 0.5f * h;
  const float mid_residual = (c[0] + hh * (c[1] + hh * (c[2] + hh * c[3]))) - x1m;
  const float P = integrated_pole(I_a, I_b, h, k);
  const float int_ku = pm::log(z_b / z_a) + P;
  const float xi0 = -pm::log(z_a * I_a) + (I_b - I_a) / (h * k * I_a * I_a);
  const float P_star = vmax(P, int_ku);
  const bool accurate = fabs_(mid_residual) <= cc.eps_accept &&
                        fabs_(xi0) * boundary_layer_factor(P_star) <= cc.eps_accept;
  const bool linear = fabs_(c[2]) * h * h + fabs_(c[3]) * h * h * h <= cc.eps_drop;
  const bool residual_form = vmax(cc.a2, cc.a3) * h <= 1.0f;
  const bool conditioned = residual_form || linear || 2.0f * vmin(cc.a2, cc.a3) * h >= 1.0f;
  if (!at_floor && !(accurate && conditioned)) return false;
  float p2[4], p3[4];
  particular(c, cc.tau_o, p2);
  particular(p2, cc.tau_sf, p3);
  const float A2 = st.s2 - p2[0];
  const float A3 = st.s3 - p3[0];
  const float C_o = A2 * cc.a3 / (cc.a3 - cc.a2);
  const float C_s = A3 - C_o;
  float r2, r3 = 0.0f;
  ExpCubic& g = sub.s3;
  g.a[0] = cc.a2;
  g.a[1] = cc.a3;
  g.C[0] = C_o;
  g.C[1] = C_s;
  if (residual_form) {
    const float d1[4] = {c[0], c[1], 2.0f * c[2], 6.0f * c[3]};
    float d2[4], d3[4];
    d2[0] = st.s2;
    d3[0] = st.s3;
    for (int j = 0; j < 3; ++j) {
      d2[j + 1] = cc.a2 * (d1[j] - d2[j]);
      d3[j + 1] = cc.a3 * (d2[j] - d3[j]);
    }
    g.c[0] = d3[0];
    g.c[1] = d3[1];
    g.c[2] = d3[2] / 2.0f;
    g.c[3] = d3[3] / 6.0f;
    g.order = 4;
    r2 = pm::resid_exp(4, cc.a2 * h);
    r3 = pm::resid_exp(4, cc.a3 * h);
    next.s2 = d2[0] + h * (d2[1] + h * (d2[2] / 2.0f + h * (d2[3] / 6.0f))) + A2 * r2;
  } else {
    for (int m = 0; m < 4; ++m) g.c[m] = p3[m];
    g.order = 0;
    r2 = pm::exp(-cc.a2 * h);
    if (C_s != 0.0f) r3 = pm::exp(-cc.a3 * h);
    next.s2 = p2[0] + h * (p2[1] + h * (p2[2] + h * p2[3])) + A2 * r2;
  }
  next.s3 = g(h, r2, r3);
  next.z = z_b;
  if (fabs_(g.C[0]) * (residual_form ? r2 : 1.0f) <= cc.eps_drop &&
      fabs_(g.C[1]) * (residual_form ? r3 : 1.0f) <= cc.eps_drop) {
    g.C[0] = 0.0f;
    g.C[1] = 0.0f;
  }
  sub.x1_a = x1a;
  sub.x1_b = x1b;
  sub.int_ku = int_ku;
  sub.fallback = at_floor && !(accurate && conditioned);
  return true;
}

BSX_HD BSX_INLINE bool settled(const Consts& cc, const State& st, float x1, float I) {
  return fabs_(st.z * I - 1.0f) <= cc.eps_drop && fabs_(st.s2 - x1) <= cc.eps_drop && fabs_(st.s3 - x1) <= cc.eps_drop;
}

}  // namespace cascade

// One pixel's cascade call: InstantLane's cursor, photocurrent and fire, plus the cascade state.
template <int kMaxG>
struct CascadeLane : InstantLane<kMaxG> {
  using Base = InstantLane<kMaxG>;
  using Base::blind_until;
  using Base::dark;
  using Base::eps_ci;
  using Base::G;
  using Base::i_scale;
  using Base::leak;
  using Base::n;
  using Base::n_leaves;
  using Base::n_pieces;
  using Base::n_substeps;
  using Base::scale;
  using Base::s3_ref;
  using Base::dr_ref;
  using Base::t_ref;
  using Base::t_search;
  using Base::ta;
  using Base::tb;
  using Base::theta_off;
  using Base::theta_on;
  using Base::va;
  using Base::vb;
  using Base::flags;

  cascade::Consts cc;
  float z, x1, s2, s3, tv;
  uint32_t n_attempts, n_settled, n_fallbacks, n_quarantines;

  // The comparator input Delta on [e0, ...) of the sub-step, in u = (t - e0) s (noise off).
  BSX_HD BSX_INLINE cascade::ExpCubic delta_fn(const cascade::SubStep& sub, int64_t e0) const {
    cascade::ExpCubic g = e0 == sub.t_a ? sub.s3 : cascade::shifted(sub.s3, static_cast<float>(e0 - sub.t_a) * 1e-9f);
    const float c0 = g.c[0] - s3_ref;
    const float since_ref = ns_to_s64<float>(e0 - t_ref);
    g.c[0] = c0 + (leak * since_ref + dr_ref);
    g.c[1] = g.c[1] + leak;
    return g;
  }

  BSX_HD BSX_INLINE float s3_at(const cascade::SubStep& sub, int64_t t) const {
    return sub.s3(static_cast<float>(t - sub.t_a) * 1e-9f);
  }

  // The plain comparator on one sub-step: releases, then first crossings until none is left.
  template <class Sink>
  BSX_HD BSX_INLINE void compare(const cascade::SubStep& sub, Sink& sink) {
    while (true) {
      if (blind_until != kNotBlind) {
        if (blind_until >= sub.t_b) return;
        const int64_t t = blind_until;
        this->set_reference(t, s3_at(sub, t), 0.0f, -eps_ci);
        blind_until = kNotBlind;
      }
      const int64_t lo = vmax(t_search, sub.t_a);
      if (lo >= sub.t_b) return;
      ++n_leaves;
      const cascade::ExpCubic d = delta_fn(sub, lo);
      if (!cascade::may_cross(d, lo, sub.t_b, lo, theta_on, theta_off)) {
        t_search = vmax(t_search, sub.t_b);
        return;
      }
      const cascade::Hit hit = cascade::search(d, lo, sub.t_b, lo, theta_on, theta_off);
      if (!hit.found) {
        t_search = vmax(t_search, sub.t_b);
        return;
      }
      this->fire(hit.t, hit.on, s3_at(sub, hit.t), 0.0f, sink);
    }
  }

  // The sub-steps of one linear piece as a resumable machine, so that a wave's lanes can share each
  // sub-step attempt whatever piece or segment they are in: piece_begin(pc); then piece_prepare(cmp)
  // runs the settled sub-steps (each followed by cmp) and opens the next sub-step's first try, and
  // returns false once the piece is done; piece_attempt(cmp) makes one try_substep and, on success,
  // commits the sub-step and runs the comparator on it (cmp(sub): the plain one, or the noise lane's
  // tree comparator). Each lane makes the same steps in the same order as a loop over the piece.
  cascade::Piece pc_;
  int64_t sub_ta_, sub_tb_, h_hint_;
  float sub_Ia_;
  bool attempting_;

  BSX_HD BSX_INLINE void piece_begin(const cascade::Piece& pc) {
    pc_ = pc;
    sub_ta_ = pc.t_a;
    h_hint_ = 0;
    attempting_ = false;
  }

  template <class Cmp>
  BSX_HD BSX_INLINE bool piece_prepare(Cmp&& cmp) {
    while (!attempting_) {
      if (!(sub_ta_ < pc_.t_b)) return false;
      const cascade::State st{z, s2, s3};
      sub_Ia_ = pc_.at(sub_ta_);
      ++n_substeps;
      if (pc_.I_a == pc_.I_b && cascade::settled(cc, st, x1, sub_Ia_)) {
        // settled under constant input: the rest of the piece in one step, no transcendental
        cascade::SubStep sub;
        sub.t_a = sub_ta_;
        const float h = static_cast<float>(pc_.t_b - sub_ta_) * 1e-9f;
        sub.s3 = cascade::ExpCubic{{s3, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f}, {cc.a2, cc.a3}, 0};
        sub.x1_a = sub.x1_b = x1;
        sub.int_ku = cc.k * h / z;
        sub.fallback = false;
        ++n_settled;
        if (tv != 0.0f) tv = tv * pm::exp(-sub.int_ku);
        sub.t_b = pc_.t_b;
        sub_ta_ = pc_.t_b;
        cmp(sub);
      } else {
        sub_tb_ = h_hint_ > 0 ? vmin(pc_.t_b, sub_ta_ + 2 * h_hint_) : pc_.t_b;
        attempting_ = true;
      }
    }
    return true;
  }

  template <class Cmp>
  BSX_HD BSX_INLINE void piece_attempt(Cmp&& cmp) {
    const int64_t t_a = sub_ta_, t_b = sub_tb_, h_ns = t_b - t_a;
    const cascade::State st{z, s2, s3};
    cascade::State next{0.0f, 0.0f, 0.0f};
    cascade::SubStep sub;
    ++n_attempts;
    if (!cascade::try_substep(cc, st, i_scale, sub_Ia_, pc_.at(t_b), static_cast<float>(h_ns) * 1e-9f, h_ns <= 1, sub,
                              next)) {
      sub_tb_ = t_a + h_ns / 2;
      return;
    }
    attempting_ = false;
    if (t_b < pc_.t_b || h_hint_ > 0) h_hint_ = t_b - t_a;
    if (sub.fallback) {
      flags |= kFlagSolverFallback;
      ++n_fallbacks;
    }
    if (!(next.z - next.z == 0.0f && next.z > 0.0f && next.s2 - next.s2 == 0.0f && next.s3 - next.s3 == 0.0f)) {
      // quarantine: a non-finite cascade restarts at the steady state of the current input
      const float I = pc_.at(t_b);
      z = 1.0f / I;
      x1 = -pm::log(z * i_scale);
      s2 = s3 = x1;
      tv = 0.0f;
      sub.s3 = cascade::ExpCubic{{s3, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f}, {cc.a2, cc.a3}, 0};
      sub.x1_a = sub.x1_b = x1;
      flags |= kFlagQuarantined;
      ++n_quarantines;
    } else {
      z = next.z;
      s2 = next.s2;
      s3 = next.s3;
      x1 = sub.x1_b;
      tv = tv * pm::exp(-sub.int_ku) + fabs_(sub.x1_b - sub.x1_a);
    }
    sub.t_a = t_a;
    sub.t_b = t_b;
    sub_ta_ = t_b;
    cmp(sub);
  }

  // The sub-steps of one linear piece in nested loops, each followed by the comparator (cmp(sub): the
  // noise lane's tree comparator). The same steps in the same order as piece_begin, piece_prepare and
  // piece_attempt; the noise lane keeps its loops, which hold its state in fewer registers.
  template <class Cmp>
  BSX_HD BSX_INLINE void run_piece(const cascade::Piece& pc, Cmp&& cmp) {
    int64_t t_a = pc.t_a, h_hint = 0;
    while (t_a < pc.t_b) {
      cascade::SubStep sub;
      cascade::State st{z, s2, s3}, next{0.0f, 0.0f, 0.0f};
      const float I_a = pc.at(t_a);
      ++n_substeps;
      sub.t_a = t_a;
      int64_t t_b;
      if (pc.I_a == pc.I_b && cascade::settled(cc, st, x1, I_a)) {
        // settled under constant input: the rest of the piece in one step, no transcendental
        t_b = pc.t_b;
        const float h = static_cast<float>(pc.t_b - t_a) * 1e-9f;
        sub.s3 = cascade::ExpCubic{{s3, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f}, {cc.a2, cc.a3}, 0};
        sub.x1_a = sub.x1_b = x1;
        sub.int_ku = cc.k * h / z;
        sub.fallback = false;
        ++n_settled;
        if (tv != 0.0f) tv = tv * pm::exp(-sub.int_ku);
      } else {
        t_b = h_hint > 0 ? vmin(pc.t_b, t_a + 2 * h_hint) : pc.t_b;
        while (true) {
          const int64_t h_ns = t_b - t_a;
          ++n_attempts;
          if (cascade::try_substep(cc, st, i_scale, I_a, pc.at(t_b), static_cast<float>(h_ns) * 1e-9f, h_ns <= 1, sub,
                                   next))
            break;
          t_b = t_a + h_ns / 2;
        }
        if (t_b < pc.t_b || h_hint > 0) h_hint = t_b - t_a;
        if (sub.fallback) {
          flags |= kFlagSolverFallback;
          ++n_fallbacks;
        }
        if (!(next.z - next.z == 0.0f && next.z > 0.0f && next.s2 - next.s2 == 0.0f && next.s3 - next.s3 == 0.0f)) {
          // quarantine: a non-finite cascade restarts at the steady state of the current input
          const float I = pc.at(t_b);
          z = 1.0f / I;
          x1 = -pm::log(z * i_scale);
          s2 = s3 = x1;
          tv = 0.0f;
          sub.s3 = cascade::ExpCubic{{s3, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f}, {cc.a2, cc.a3}, 0};
          sub.x1_a = sub.x1_b = x1;
          flags |= kFlagQuarantined;
          ++n_quarantines;
        } else {
          z = next.z;
          s2 = next.s2;
          s3 = next.s3;
          x1 = sub.x1_b;
          tv = tv * pm::exp(-sub.int_ku) + fabs_(sub.x1_b - sub.x1_a);
        }
      }
      sub.t_b = t_b;
      t_a = t_b;
      cmp(sub);
    }
  }


  // The call as a resumable machine over the segments and their pieces: begin(t_end); then
  // prepare(sink) returns kAttempt when a sub-step try is due (attempt(sink) makes it) and kFinished
  // once the call is done; then end().
  static constexpr int kAttempt = 1, kFinished = 2;

  template <class Sink>
  BSX_HD BSX_INLINE int prepare(Sink& sink) {
    const auto cmp = [&](const cascade::SubStep& sub) { compare(sub, sink); };
    while (true) {
      if (this->in_piece) {
        if (piece_prepare(cmp)) return kAttempt;
        this->in_piece = false;
        this->t_cur = pc_.t_b;
      }
      if (!this->seg_open) {
        if (!(this->t_cur < this->t_end_)) return kFinished;
        this->open_next();
        this->stop = vmin(this->t_end_, tb);
        this->seg_open = true;
      }
      if (this->t_cur < this->stop) {
        const int64_t t = this->t_cur;
        const int64_t e = vmin(this->stop, ((t >> kPlainSegmentLog2) + 1) * kPlainSegmentNs);
        ++n_pieces;
        piece_begin(cascade::Piece{t, e, this->photocurrent(t), this->photocurrent(e)});
        this->in_piece = true;
        continue;
      }
      // the segment is done: the next one starts at its end sample
      if (tb == kNoBreakpoint) return kFinished;
      this->seg_open = false;
      ta = tb;
      for (int g = 0; g < kMaxG; ++g)
        if (g < G) va[g] = vb[g];
    }
  }

  template <class Sink>
  BSX_HD BSX_INLINE void attempt(Sink& sink) {
    piece_attempt([&](const cascade::SubStep& sub) { compare(sub, sink); });
  }

  // The call: the segments of the augmented samples in order, cut at multiples of 2^23 ns, each a
  // linear piece of photocurrent (constant waveforms); the segment ends are fetched kAhead samples ahead.
  template <class Sink>
  BSX_HD BSX_INLINE void advance(int64_t t_end, Sink& sink) {
    this->begin(t_end);
    while (prepare(sink) == kAttempt) attempt(sink);
    this->end();
  }
};

}  // namespace bsx
