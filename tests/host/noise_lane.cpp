// The noise lane program (noise.h) compiled for the host with the device build's floating-point
// policy, as a small C library for the tests: the init kernel's and the walk kernel's per-pixel work
// over plain arrays, in the kernels' structure-of-arrays layout, so that a test can compare it with
// the oracle on the CPU before the same program runs on the GPU.
#include <cstdint>
#include <cstring>
#include <vector>

#include "noise.h"

using bsx::NoiseLane;
using bsx::RawEventDev;

namespace {

constexpr int kG = 2;

struct VecSink {
  std::vector<RawEventDev>* ev;
  void push(const RawEventDev& e) { ev->push_back(e); }
};

}  // namespace

extern "C" {

// Call constants (the KConst fields the lane reads).
struct HostConst {
  int64_t n_pixels, groups, refractory_ns, ideal_reset;
  uint64_t seed;
  int64_t camera;
  float k, a2, a3, tau_o, tau_sf, eps_accept, eps_drop, eps_ci, label_cut, irr_max;
  float ln_min, inv_dln, margin_z;
  int32_t levels, n_nodes, stride, pad;
};

struct HostMaps {
  const float *theta_on, *theta_off, *leak, *dark, *scale, *phi0;
  const uint8_t* hot;
};

struct HostState {
  float *z, *x1, *s2, *s3, *i_scale, *tv, *s3_ref, *n3_ref, *dr_ref;
  int64_t *t_ref, *blind_until, *t_search;
  uint16_t* flags;
  float* held;  // [g * P + p]
  float *nk, *nk1, *ln_a1;
  int64_t* cell;
  uint8_t* have_k1;
};

// counters: cells, refinements, substeps, attempts, settled, leaves, fallbacks, quarantines, pieces.
static void setup(NoiseLane<kG>& L, const HostConst& c, const float* tab, const HostMaps& m, int p) {
  L.pixel = static_cast<uint32_t>(p);
  L.theta_on = m.theta_on[p];
  L.theta_off = m.theta_off[p];
  L.leak = m.leak[p];
  L.dark = m.dark[p];
  L.scale = m.scale[p];
  L.irr_max = c.irr_max;
  L.eps_ci = c.eps_ci;
  L.label_cut = c.label_cut;
  L.refractory_ns = c.refractory_ns;
  L.G = static_cast<int>(c.groups);
  L.ideal_reset = c.ideal_reset != 0;
  L.hot = m.hot[p] != 0;
  L.seed = c.seed;
  L.camera = static_cast<uint32_t>(c.camera);
  L.cc = bsx::cascade::Consts{c.k, c.a2, c.a3, c.tau_o, c.tau_sf, c.eps_accept, c.eps_drop};
  L.nd = bsx::NoiseDev{tab, c.levels, c.n_nodes, c.stride, c.ln_min, c.inv_dln, c.margin_z};
  L.margin_z = c.margin_z;
  L.n_substeps = L.n_pieces = L.n_leaves = 0;
  L.n_attempts = L.n_settled = L.n_fallbacks = L.n_quarantines = 0;
  L.n_cells = L.n_refinements = 0;
  L.shot_rate = 0.0;
  L.shot_next = INT64_MAX;
  L.shot_index = 0;
}

static void count(const NoiseLane<kG>& L, uint64_t* counters) {
  counters[0] += L.n_cells;
  counters[1] += L.n_refinements;
  counters[2] += L.n_substeps;
  counters[3] += L.n_attempts;
  counters[4] += L.n_settled;
  counters[5] += L.n_leaves;
  counters[6] += L.n_fallbacks;
  counters[7] += L.n_quarantines;
  counters[8] += L.n_pieces;
}

// init_kernel's cascade-with-noise work for every pixel.
int noise_lane_init(const HostConst* c, const float* tab, HostMaps m, const float* irr0, int64_t t0, HostState s,
                    uint64_t* counters) {
  const int P = static_cast<int>(c->n_pixels), G = static_cast<int>(c->groups);
  for (int p = 0; p < P; ++p) {
    NoiseLane<kG> L{};
    setup(L, *c, tab, m, p);
    L.n = 0;
    L.skip = 0;
    L.t_state = t0;
    float mean_I = L.dark;
    for (int g = 0; g < G; ++g) {
      const float raw = irr0[static_cast<int64_t>(p) * G + g];
      const float v = L.clampv(raw);
      L.held[g] = v;
      mean_I = mean_I + L.scale * v;
      s.held[static_cast<int64_t>(g) * P + p] = v;
    }
    L.i_scale = mean_I;
    L.z = 1.0f / mean_I;
    L.x1 = L.s2 = L.s3 = 0.0f;
    L.mid = nullptr;
    L.mid_stride = 0;
    const float n3 = L.init(t0, mean_I);
    s.z[p] = L.z;
    s.x1[p] = 0.0f;
    s.s2[p] = 0.0f;
    s.s3[p] = 0.0f;
    s.i_scale[p] = mean_I;
    s.tv[p] = 0.0f;
    s.t_ref[p] = t0;
    s.s3_ref[p] = 0.0f;
    s.n3_ref[p] = n3;
    s.dr_ref[p] = m.phi0[p];
    s.blind_until[p] = bsx::kNotBlind;
    s.t_search[p] = t0 + 1;
    s.flags[p] = bsx::kFlagFirstAfterInit;
    for (int i = 0; i < 3; ++i) {
      s.nk[i * P + p] = L.nk[i];
      s.nk1[i * P + p] = L.nk1[i];
    }
    s.ln_a1[p] = L.ln_a1;
    s.cell[p] = L.cell;
    s.have_k1[p] = L.have_k1 ? 1 : 0;
    count(L, counters);
  }
  return 0;
}

// The walk kernel's work for every pixel: row_of[p] is pixel p's run (-1: none). Events are appended
// to ev (capacity cap); returns the event count (more than cap: not all written).
int64_t noise_lane_advance(const HostConst* c, const float* tab, HostMaps m, const int32_t* row_of,
                           const int64_t* row_start, const int32_t* row_count, const int64_t* t, const float* irr,
                           int64_t t_state, int64_t t_end, HostState s, RawEventDev* ev, int64_t cap,
                           uint64_t* counters) {
  const int P = static_cast<int>(c->n_pixels), G = static_cast<int>(c->groups);
  std::vector<RawEventDev> out;
  std::vector<float> mid(bsx::noise::kMidSlots);
  for (int p = 0; p < P; ++p) {
    NoiseLane<kG> L{};
    setup(L, *c, tab, m, p);
    const int32_t row = row_of[p];
    if (row >= 0) {
      L.ts = t + row_start[row];
      L.irr = irr + row_start[row] * G;
      L.n = row_count[row];
    } else {
      L.ts = nullptr;
      L.irr = nullptr;
      L.n = 0;
    }
    L.t_state = t_state;
    L.skip = L.n > 0 && L.ts[0] == t_state ? 1 : 0;
    for (int g = 0; g < G; ++g) L.held[g] = s.held[static_cast<int64_t>(g) * P + p];
    L.i_scale = s.i_scale[p];
    L.z = s.z[p];
    L.x1 = s.x1[p];
    L.s2 = s.s2[p];
    L.s3 = s.s3[p];
    L.tv = s.tv[p];
    L.s3_ref = s.s3_ref[p];
    L.n3_ref = s.n3_ref[p];
    L.dr_ref = s.dr_ref[p];
    L.t_ref = s.t_ref[p];
    L.blind_until = s.blind_until[p];
    L.t_search = s.t_search[p];
    L.flags = s.flags[p];
    for (int i = 0; i < 3; ++i) {
      L.nk[i] = s.nk[i * P + p];
      L.nk1[i] = s.nk1[i * P + p];
    }
    L.ln_a1 = s.ln_a1[p];
    L.cell = s.cell[p];
    L.have_k1 = s.have_k1[p] != 0;
    L.mid = mid.data();
    L.mid_stride = 1;
    VecSink sink{&out};
    L.resume();
    L.advance(t_end, sink);
    for (int g = 0; g < G; ++g) s.held[static_cast<int64_t>(g) * P + p] = L.held[g];
    s.z[p] = L.z;
    s.x1[p] = L.x1;
    s.s2[p] = L.s2;
    s.s3[p] = L.s3;
    s.tv[p] = L.tv;
    s.s3_ref[p] = L.s3_ref;
    s.n3_ref[p] = L.n3_ref;
    s.dr_ref[p] = L.dr_ref;
    s.t_ref[p] = L.t_ref;
    s.blind_until[p] = L.blind_until;
    s.t_search[p] = L.t_search;
    s.flags[p] = static_cast<uint16_t>(L.flags);
    for (int i = 0; i < 3; ++i) {
      s.nk[i * P + p] = L.nk[i];
      s.nk1[i * P + p] = L.nk1[i];
    }
    s.ln_a1[p] = L.ln_a1;
    s.cell[p] = L.cell;
    s.have_k1[p] = L.have_k1 ? 1 : 0;
    count(L, counters);
  }
  const int64_t n = static_cast<int64_t>(out.size());
  std::memcpy(ev, out.data(), static_cast<size_t>(n < cap ? n : cap) * sizeof(RawEventDev));
  return n;
}

}  // extern "C"
