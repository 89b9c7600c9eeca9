// The cascade with its in-path noise, as a lane program: one GPU lane runs one pixel's call.
//
// Model (reference: the noise of M3 §D in src/sensor/noise.h:22-191, the tables' interface
// src/sensor/noise_tables.h:118-139, and PixelKernel's noise path src/sensor/pixel.h:322-357, 389,
// 423-435, 566-594, 645-741, 821-845, 909-940, 1270-1453, 1599-1611, 1630-1652). Time is cut into
// cells of 2^16 ns. In each cell the noise (n1, n2, n3) is a linear Gaussian process with a frozen
// rate a1, taken from a look-ahead of stage 1 over the cell. Its values at the cell's nodes come from
// the lattice step x_{k+1} = Phi x_k + chol(Q) xi, and inside the cell from a binary bridge tree:
// the midpoint of a node is A_m x_L + K x_R + chol xi, down to leaves of 2^(16 - L) ns, where n3 is
// the cubic Hermite with slopes a3 (n2 - n3). Every xi is three normals of one Philox draw keyed by
// the node's time, so the noise is a pure function of (seed, camera, pixel, time), whatever the
// calls' boundaries. The comparator adds n3 to Delta: a cell's (or node's) Hermite plus an exclusion
// margin z sigma_br + g |W e| bounds where the tree may cross, and the tree is refined, in time order,
// only where a crossing is possible.
//
// Every arithmetic step is the reference's, in its order, so the fp32 results equal the reference's
// fp32 CPU path bit for bit. What differs is the data movement: the depth-first walk of a cell's
// tree keeps, instead of a stack of nodes, the midpoint made at each level on the current path
// (node values follow from the path: a node's left end is the midpoint of the level where the path
// last went right, its right end that of the level where it last went left), in a per-lane column of
// caller-provided memory (LDS on the device).
#pragma once

#include "cascade.h"

namespace bsx {
namespace noise {

inline constexpr int kCellLog2 = 16;
inline constexpr int64_t kCellNs = int64_t{1} << kCellLog2;
inline constexpr int kMaxLevels = 10;
inline constexpr int kMidSlots = 3 * kMaxLevels;  // per-lane floats of the tree walk's path memory

// Three standard normals of one keyed draw: Box-Muller of the draw's four 23-bit open uniforms, with
// the library's log and sincos (reference: normals3, src/sensor/noise.h:32-45).
BSX_HD BSX_INLINE void normals3(uint64_t seed, uint32_t purpose, uint32_t camera, uint32_t pixel, uint64_t index,
                                float* out) {
  const Word4 r = draw(seed, purpose, camera, pixel, index);
  const float u0 = open_uniform32(r.v[0]), u1 = open_uniform32(r.v[1]);
  const float u2 = open_uniform32(r.v[2]), u3 = open_uniform32(r.v[3]);
  constexpr double kTwoPi = 0x1.921fb54442d18p+2;
  const float r0 = sqrt_(-2.0f * pm::log(u0)), r1 = sqrt_(-2.0f * pm::log(u2));
  float s0, c0, s1, c1;
  pm::sincos(static_cast<float>(kTwoPi) * u1, &s0, &c0);
  pm::sincos(static_cast<float>(kTwoPi) * u3, &s1, &c1);
  out[0] = r0 * c0;
  out[1] = r0 * s0;
  out[2] = r1 * c1;
}

// y = L x for a packed lower-triangular 3x3 L (00, 10, 11, 20, 21, 22).
BSX_HD BSX_INLINE void lower3_mul(const float* L, const float* x, float* y) {
  y[0] = L[0] * x[0];
  y[1] = L[1] * x[0] + L[2] * x[1];
  y[2] = L[3] * x[0] + L[4] * x[1] + L[5] * x[2];
}

// The cubic Hermite of n3 on a node of length h (s): values n3 at both ends, slopes a3 (n2 - n3).
BSX_HD BSX_INLINE void hermite(const float* xl, const float* xr, float a3, float h, float* c) {
  const float dl = a3 * (xl[1] - xl[2]), dr = a3 * (xr[1] - xr[2]);
  const float delta = (xr[2] - xl[2]) / h;
  c[0] = xl[2];
  c[1] = dl;
  c[2] = (3.0f * delta - 2.0f * dl - dr) / h;
  c[3] = (dl + dr - 2.0f * delta) / (h * h);
}

BSX_HD BSX_INLINE float hermite_at(const float* c, float u) { return c[0] + u * (c[1] + u * (c[2] + u * c[3])); }

// The exclusion margin z sigma_br + g |W e| of a node with ends e = (x_L, x_R) (m: sigma_br, g and the
// 21 packed entries of W).
BSX_HD BSX_INLINE float margin(const float* m, float z, const float* xl, const float* xr) {
  const float e[6] = {xl[0], xl[1], xl[2], xr[0], xr[1], xr[2]};
  const float* W = m + 2;
  float norm2 = 0.0f;
  int idx = 0;
  for (int i = 0; i < 6; ++i) {
    float s = 0.0f;
    for (int j = 0; j <= i; ++j) s = s + W[idx++] * e[j];
    norm2 = norm2 + s * s;
  }
  return z * m[0] + m[1] * sqrt_(norm2);
}

// The table position of a rate (clamped to the grid; NaN lands on the first node).
struct Pos {
  int i;
  float f;
};

BSX_HD BSX_INLINE Pos position(const NoiseDev& nd, float ln_a1) {
  float x = (ln_a1 - nd.ln_min) * nd.inv_dln;
  const float x_max = static_cast<float>(nd.n_nodes - 1);
  x = x > 0.0f ? (x < x_max ? x : x_max) : 0.0f;
  int i = static_cast<int>(x);
  if (i > nd.n_nodes - 2) i = nd.n_nodes - 2;
  return {i, x - static_cast<float>(i)};
}

// count entries from offset, interpolated between the position's two nodes.
BSX_HD BSX_INLINE void lookup(const NoiseDev& nd, Pos p, int offset, int count, float* out) {
  const BSX_GLOBAL float* a = nd.tab + static_cast<int64_t>(p.i) * nd.stride + offset;
  const BSX_GLOBAL float* b = a + nd.stride;
  for (int j = 0; j < count; ++j) out[j] = a[j] + p.f * (b[j] - a[j]);
}

BSX_HD BSX_INLINE int bridge_offset(int l) { return 18 + 24 * (l - 1); }
BSX_HD BSX_INLINE int margin_offset(int levels, int l) { return 18 + 24 * levels + 23 * l; }
BSX_HD BSX_INLINE int64_t level_length(int level) { return kCellNs >> level; }

}  // namespace noise

// One pixel's cascade call with noise: CascadeLane's sub-steps and fire, plus the noise state, the
// cells and the tree comparator. The sample cursor is the segment [ta, tb) with its end values and
// jb, the index of the input sample at tb (tb = kNoBreakpoint: held after the last sample).
template <int kMaxG>
struct NoiseLane : CascadeLane<kMaxG> {
  using Base = CascadeLane<kMaxG>;
  using Base::blind_until;
  using Base::cc;
  using Base::eps_ci;
  using Base::G;
  using Base::leak;
  using Base::n;
  using Base::n_leaves;
  using Base::n_pieces;
  using Base::n_settled;
  using Base::n_substeps;
  using Base::s2;
  using Base::s3;
  using Base::s3_ref;
  using Base::n3_ref;
  using Base::dr_ref;
  using Base::t_ref;
  using Base::t_search;
  using Base::ta;
  using Base::tb;
  using Base::theta_off;
  using Base::theta_on;
  using Base::tv;
  using Base::va;
  using Base::vb;
  using Base::x1;
  using Base::z;

  NoiseDev nd;
  float margin_z;
  float nk[3], nk1[3], ln_a1, hc0[4], margin0;
  int64_t cell;
  bool have_k1;
  int jb;
  uint32_t n_cells, n_refinements;
  // The tree walk's path memory: mid[(3 (l - 1) + c) * mid_stride], l = 1..L, c = 0..2.
  float* mid;
  int mid_stride;

  // ---- The sample cursor ------------------------------------------------------------------------
  BSX_HD BSX_INLINE void cursor_begin() {
    ta = this->t_state;
    for (int g = 0; g < kMaxG; ++g)
      if (g < G) va[g] = this->held[g];
    jb = this->skip;
    if (this->skip) this->load_sample(0, ta, va);
    tb = kNoBreakpoint;
    if (jb < n) this->load_sample(jb, tb, vb);
  }

  // Moves the cursor to the segment holding t (t not before the segment's start).
  BSX_HD BSX_INLINE void seek(int64_t t) {
    while (tb != kNoBreakpoint && tb <= t) {
      ta = tb;
      for (int g = 0; g < kMaxG; ++g)
        if (g < G) va[g] = vb[g];
      ++jb;
      tb = kNoBreakpoint;
      if (jb < n) this->load_sample(jb, tb, vb);
    }
  }

  // The linear piece of photocurrent from t to the segment's end or limit, whichever is first
  // (reference: next_piece for constant light, pixel.h:1190-1230).
  BSX_HD BSX_INLINE cascade::Piece piece(int64_t t, int64_t limit) {
    seek(t);
    const int64_t e = vmin(limit, tb);
    return cascade::Piece{t, e, this->photocurrent(t), this->photocurrent(e)};
  }

  // ---- Cells and the lattice ----------------------------------------------------------------------
  // The level-0 Hermite of n3 and exclusion margin of the current cell.
  BSX_HD BSX_INLINE void level0() {
    noise::hermite(nk, nk1, cc.a3, 65536.0f * 1e-9f, hc0);
    if (nd.levels > 0) {
      float m0[23];
      noise::lookup(nd, noise::position(nd, ln_a1), noise::margin_offset(nd.levels, 0), 23, m0);
      margin0 = noise::margin(m0, margin_z, nk, nk1);
    }
  }

  // The end node of the current cell: x_{k+1} = Phi x_k + chol(Q) xi, xi keyed by the node's time.
  BSX_HD BSX_INLINE void draw_lattice(int64_t tc1) {
    float xi[3], step[12], a[3], b[3];
    noise::normals3(this->seed, kPurposeNoise, this->camera, this->pixel, static_cast<uint64_t>(tc1), xi);
    noise::lookup(nd, noise::position(nd, ln_a1), 0, 12, step);
    noise::lower3_mul(step, nk, a);
    noise::lower3_mul(step + 6, xi, b);
    for (int i = 0; i < 3; ++i) nk1[i] = a[i] + b[i];
    level0();
    have_k1 = true;
    ++n_cells;
  }

  // The frozen rate of a cell in which stage 1 is settled under constant light.
  BSX_HD BSX_INLINE void settled_rate(int64_t t, float pre, int64_t tc1) {
    const float h_cell = 65536.0f * 1e-9f;
    const float int_ku = pre + cc.k * static_cast<float>(tc1 - t) * 1e-9f / z;
    ln_a1 = pm::log(int_ku / h_cell);
  }

  // The cell holding t: its frozen rate from a stage-1 look-ahead to the cell's end (z only, over
  // the pieces of the cell), then its lattice draw. pre: the rate integral already accrued in the cell
  // before t (init inside a cell).
  BSX_HD BSX_INLINE void enter_cell(int64_t t, float pre) {
    cell = t >> noise::kCellLog2;
    const int64_t tc1 = (cell + 1) * noise::kCellNs;
    cascade::Piece pc = piece(t, tc1);
    const cascade::State st{z, s2, s3};
    if (pc.t_b == tc1 && pc.I_a == pc.I_b && cascade::settled(cc, st, x1, pc.I_a)) {
      settled_rate(t, pre, tc1);
    } else {
      // the look-ahead moves the cursor; it is put back afterwards
      const int64_t ta0 = ta, tb0 = tb;
      const int jb0 = jb;
      float va0[kMaxG], vb0[kMaxG];
      for (int g = 0; g < kMaxG; ++g) {
        va0[g] = va[g];
        vb0[g] = vb[g];
      }
      const float h_cell = 65536.0f * 1e-9f;
      float zz = z, int_kI = 0.0f;
      while (true) {
        const float h = static_cast<float>(pc.t_b - pc.t_a) * 1e-9f;
        if (!(pc.I_a == pc.I_b && fabs_(zz * pc.I_a - 1.0f) <= cc.eps_drop)) zz = cascade::z_step(zz, pc.I_a, pc.I_b, h, cc.k);
        int_kI = int_kI + cascade::integrated_pole(pc.I_a, pc.I_b, h, cc.k);
        if (pc.t_b >= tc1) break;
        pc = piece(pc.t_b, tc1);
      }
      const float int_ku = pre + pm::log(zz / z) + int_kI;
      const float r = int_ku / h_cell;
      ln_a1 = pm::log(r < 1e-30f ? 1e-30f : r);
      ta = ta0;
      tb = tb0;
      jb = jb0;
      for (int g = 0; g < kMaxG; ++g) {
        va[g] = va0[g];
        vb[g] = vb0[g];
      }
    }
    draw_lattice(tc1);
  }

  BSX_HD BSX_INLINE void next_cell() {
    for (int i = 0; i < 3; ++i) nk[i] = nk1[i];
    have_k1 = false;
    ++cell;
  }

  // ---- The comparator with noise ------------------------------------------------------------------
  // Delta on [e0, ...) of a sub-step, with the noise Hermite hc of the node starting at node_t.
  BSX_HD BSX_INLINE cascade::ExpCubic delta_fn(const cascade::SubStep& sub, int64_t e0, const float* hc,
                                               int64_t node_t) const {
    cascade::ExpCubic g = e0 == sub.t_a ? sub.s3 : cascade::shifted(sub.s3, static_cast<float>(e0 - sub.t_a) * 1e-9f);
    float c0 = g.c[0] - s3_ref;
    const float d = static_cast<float>(e0 - node_t) * 1e-9f;
    c0 = c0 + ((hc[0] + d * (hc[1] + d * (hc[2] + d * hc[3]))) - n3_ref);
    g.c[1] = g.c[1] + (hc[1] + d * (2.0f * hc[2] + d * (3.0f * hc[3])));
    g.c[2] = g.c[2] + (hc[2] + d * (3.0f * hc[3]));
    g.c[3] = g.c[3] + hc[3];
    const float since_ref = ns_to_s64<float>(e0 - t_ref);
    g.c[0] = c0 + (leak * since_ref + dr_ref);
    g.c[1] = g.c[1] + leak;
    return g;
  }

  // Can Delta, widened by the margin, reach a level on the integers [max(lo, e0), e1)?
  BSX_HD BSX_INLINE bool may_cross(const cascade::ExpCubic& d, int64_t e0, int64_t e1, int64_t lo, float m) const {
    const int64_t k_lo = vmax(lo, e0) - e0, k_hi = e1 - 1 - e0;
    if (k_lo > k_hi) return false;
    const cascade::Range r = cascade::range(d, static_cast<float>(k_lo) * 1e-9f, static_cast<float>(k_hi) * 1e-9f);
    return r.hi + m >= theta_on || r.lo - m <= -theta_off;
  }

  // The midpoint at t_mid of a level-(level - 1) node with ends xl, xr.
  BSX_HD BSX_INLINE void bridge_child(noise::Pos pos, int level, int64_t t_mid, const float* xl, const float* xr,
                                      float* xm) const {
    float xi[3], br[24], c[3];
    noise::normals3(this->seed, kPurposeNoise, this->camera, this->pixel, static_cast<uint64_t>(t_mid), xi);
    noise::lookup(nd, pos, noise::bridge_offset(level), 24, br);
    noise::lower3_mul(br + 18, xi, c);
    for (int i = 0; i < 3; ++i) {
      xm[i] = (br[3 * i] * xl[0] + br[3 * i + 1] * xl[1] + br[3 * i + 2] * xl[2]) +
              (br[9 + 3 * i] * xr[0] + br[9 + 3 * i + 1] * xr[1] + br[9 + 3 * i + 2] * xr[2]) + c[i];
    }
  }

  // n3 at t in the current cell: the tree descended to the leaf holding t.
  BSX_HD BSX_INLINE float n3_at(int64_t t) {
    const int64_t tc0 = cell * noise::kCellNs;
    if (t == tc0) return nk[2];
    if (t == tc0 + noise::kCellNs) return nk1[2];
    const noise::Pos pos = noise::position(nd, ln_a1);
    float xl[3], xr[3], xm[3];
    for (int i = 0; i < 3; ++i) {
      xl[i] = nk[i];
      xr[i] = nk1[i];
    }
    int64_t node_t = tc0;
    const int L = nd.levels;
    for (int l = 1; l <= L; ++l) {
      const int64_t m = node_t + noise::level_length(l);
      bridge_child(pos, l, m, xl, xr, xm);
      ++n_refinements;
      if (t == m) return xm[2];
      if (t < m) {
        for (int i = 0; i < 3; ++i) xr[i] = xm[i];
      } else {
        for (int i = 0; i < 3; ++i) xl[i] = xm[i];
        node_t = m;
      }
    }
    float hc[4];
    noise::hermite(xl, xr, cc.a3, static_cast<float>(noise::level_length(L)) * 1e-9f, hc);
    return noise::hermite_at(hc, static_cast<float>(t - node_t) * 1e-9f);
  }

  struct Found {
    bool found, on;
    int64_t t;
    float s3, n3;
  };

  // The node's end values from the path memory: level l, index i in the level.
  BSX_HD BSX_INLINE void node_ends(int l, uint32_t i, float* xl, float* xr) const {
    if (i == 0) {
      for (int c = 0; c < 3; ++c) xl[c] = nk[c];
    } else {
      const int lv = l - __builtin_ctz(i);
      for (int c = 0; c < 3; ++c) xl[c] = mid[(3 * (lv - 1) + c) * mid_stride];
    }
    if (i + 1 == (1u << l)) {
      for (int c = 0; c < 3; ++c) xr[c] = nk1[c];
    } else {
      const int lv = l - __builtin_ctz(i + 1);
      for (int c = 0; c < 3; ++c) xr[c] = mid[(3 * (lv - 1) + c) * mid_stride];
    }
  }

  // The first crossing of Delta with noise on [lo, sub.t_b) in the cell starting at tc0: the cell's
  // tree walked depth first in time order, a node refined only when its Hermite widened by its margin
  // may reach a level, a leaf searched exactly (reference: walk_impl, pixel.h:645-730). The root is
  // refined unconditionally (its test is the caller's).
  BSX_HD BSX_INLINE Found walk_tree(int64_t tc0, const cascade::SubStep& sub, int64_t lo) {
    Found out{false, false, 0, 0.0f, 0.0f};
    const int L = nd.levels;
    const noise::Pos pos = noise::position(nd, ln_a1);
    int l = 0;
    uint32_t i = 0;
    bool top = L > 0;
    while (true) {
      const int64_t length = noise::level_length(l);
      const int64_t node_t = tc0 + static_cast<int64_t>(i) * length;
      const int64_t span_lo = vmax(node_t, lo), span_hi = vmin(node_t + length, sub.t_b);
      bool descend = false;
      if (span_lo < span_hi) {
        float xl[3], xr[3];
        node_ends(l, i, xl, xr);
        const int64_t e0 = vmax(sub.t_a, span_lo), e1 = vmin(sub.t_b, span_hi);
        if (l < L) {
          descend = true;
          if (!top) {
            float hc[4], mb[23];
            noise::hermite(xl, xr, cc.a3, static_cast<float>(length) * 1e-9f, hc);
            noise::lookup(nd, pos, noise::margin_offset(L, l), 23, mb);
            const float m = noise::margin(mb, margin_z, xl, xr);
            descend = e0 < e1 && may_cross(delta_fn(sub, e0, hc, node_t), e0, e1, lo, m);
          }
          top = false;
          if (descend) {
            float xm[3];
            bridge_child(pos, l + 1, node_t + noise::level_length(l + 1), xl, xr, xm);
            ++n_refinements;
            for (int c = 0; c < 3; ++c) mid[(3 * l + c) * mid_stride] = xm[c];
          }
        } else if (e0 < e1) {
          float hc[4];
          noise::hermite(xl, xr, cc.a3, static_cast<float>(length) * 1e-9f, hc);
          ++n_leaves;
          const cascade::Hit hit = cascade::search(delta_fn(sub, e0, hc, node_t), e0, e1, lo, theta_on, theta_off);
          if (hit.found) {
            out.found = true;
            out.on = hit.on;
            out.t = hit.t;
            out.s3 = sub.s3(static_cast<float>(hit.t - sub.t_a) * 1e-9f);
            out.n3 = noise::hermite_at(hc, static_cast<float>(hit.t - node_t) * 1e-9f);
            return out;
          }
        }
      }
      if (descend) {
        ++l;
        i = 2 * i;
        continue;
      }
      // the next node in time order: up while a right child, then the right sibling
      while (l > 0 && (i & 1u)) {
        i >>= 1;
        --l;
      }
      if (l == 0) return out;
      ++i;
    }
  }

  // The physical release inside the sub-step (the reference at s3 + n3 of the release, less eps_ci);
  // false while blind past the sub-step's end.
  BSX_HD BSX_INLINE bool release(const cascade::SubStep& sub) {
    if (blind_until == kNotBlind) return true;
    if (blind_until >= sub.t_b) return false;
    const int64_t t = blind_until;
    const float n3 = n3_at(t);
    this->set_reference(t, sub.s3(static_cast<float>(t - sub.t_a) * 1e-9f), n3, -eps_ci);
    blind_until = kNotBlind;
    return true;
  }

  // The comparator with noise on one sub-step (reference: comparator_tree, pixel.h:1630-1652).
  template <class Sink>
  BSX_HD BSX_INLINE void compare_tree(const cascade::SubStep& sub, Sink& sink) {
    const int64_t tc0 = cell * noise::kCellNs;
    while (true) {
      if (!release(sub)) return;
      const int64_t lo = vmax(t_search, sub.t_a);
      if (lo >= sub.t_b) return;
      if (nd.levels > 0 && !may_cross(delta_fn(sub, lo, hc0, tc0), lo, sub.t_b, lo, margin0)) {
        t_search = vmax(t_search, sub.t_b);
        return;
      }
      const Found f = walk_tree(tc0, sub, lo);
      if (!f.found) {
        t_search = vmax(t_search, sub.t_b);
        return;
      }
      this->fire(f.t, f.on, f.s3, f.n3, sink);
    }
  }

  // A whole cell under one constant piece with stage 1 settled, advanced in one step when the
  // comparator finds no crossing in it (reference: quiet_cell, pixel.h:1411-1453); false when the
  // cell needs the general path (its lattice draw may then be done).
  template <class Sink>
  BSX_HD BSX_INLINE bool quiet_cell(int64_t t, Sink&) {
    if (blind_until != kNotBlind || nd.levels == 0) return false;
    const int64_t tc1 = t + noise::kCellNs;
    const cascade::Piece pc = piece(t, tc1);
    const cascade::State st{z, s2, s3};
    if (pc.t_b != tc1 || pc.I_a != pc.I_b || !cascade::settled(cc, st, x1, pc.I_a)) return false;
    cell = t >> noise::kCellLog2;
    settled_rate(t, 0.0f, tc1);
    draw_lattice(tc1);
    cascade::SubStep sub;
    sub.t_a = t;
    sub.t_b = tc1;
    sub.s3 = cascade::ExpCubic{{s3, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f}, {cc.a2, cc.a3}, 0};
    sub.int_ku = cc.k * (static_cast<float>(tc1 - t) * 1e-9f) / z;
    const int64_t lo = vmax(t_search, t);
    if (lo < tc1 && may_cross(delta_fn(sub, lo, hc0, t), lo, tc1, lo, margin0)) {
      if (walk_tree(t, sub, lo).found) return false;
    }
    ++n_substeps;
    ++n_settled;
    ++n_pieces;
    if (tv != 0.0f) tv = tv * pm::exp(-sub.int_ku);
    t_search = vmax(t_search, tc1);
    next_cell();
    return true;
  }

  // After loading a state inside a cell: the cell's level-0 hull as its lattice draw left it.
  BSX_HD BSX_INLINE void resume() {
    if (have_k1) level0();
  }

  // The call (reference: advance_slice with noise, pixel.h:909-940): cells in order, each either
  // quiet in one step or cut into the pieces of its samples, each piece's sub-steps followed by the
  // tree comparator.
  template <class Sink>
  BSX_HD BSX_INLINE void advance(int64_t t_end, Sink& sink) {
    cursor_begin();
    int64_t t = this->t_state;
    while (t < t_end) {
      if (!have_k1 && t + noise::kCellNs <= t_end && quiet_cell(t, sink)) {
        t += noise::kCellNs;
        continue;
      }
      if (!have_k1) enter_cell(t, 0.0f);
      const int64_t boundary = (cell + 1) * noise::kCellNs;
      const int64_t stop = vmin(t_end, boundary);
      int64_t tt = t;
      while (tt < stop) {
        const cascade::Piece pc = piece(tt, stop);
        ++n_pieces;
        this->run_piece(pc, [&](const cascade::SubStep& sub) { compare_tree(sub, sink); });
        tt = pc.t_b;
      }
      if (stop == boundary) next_cell();
      t = stop;
    }
    this->end();
  }

  // init's noise part: the stationary state at the cell's start node, extended to t0 inside the
  // cell; returns n3 at t0 (reference: init, pixel.h:821-845). The lane holds the init light as its
  // held segment (no samples).
  BSX_HD BSX_INLINE float init(int64_t t0, float mean_I) {
    cell = t0 >> noise::kCellLog2;
    const int64_t tk = cell * noise::kCellNs;
    float xi[3], chol[6];
    noise::normals3(this->seed, kPurposeNoiseInit, this->camera, this->pixel, static_cast<uint64_t>(tk), xi);
    noise::lookup(nd, noise::position(nd, pm::log(cc.k * mean_I)), 12, 6, chol);
    noise::lower3_mul(chol, xi, nk);
    have_k1 = false;
    if (t0 != tk) {
      cursor_begin();
      enter_cell(t0, cc.k * mean_I * static_cast<float>(t0 - tk) * 1e-9f);
      return n3_at(t0);
    }
    return nk[2];
  }
};

}  // namespace bsx
