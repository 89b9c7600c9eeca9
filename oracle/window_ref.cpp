// The reference's own window engine on this library's scene (TEST INFRASTRUCTURE): WindowEngine<kHost>
// compiled from src/engine/window.cpp and par_host.cpp where they lie (oracle/Makefile, target window),
// rendering through ballasim_amd/seam/scene_hook_bsx.h on a bsx scene. The tests hand it a scene created
// through libbsx.so (the same loaded library, so the handle is shared), the active pixels of a window and
// their lattice levels, and compare its runs with bsx_scene_sample_window's: the reference's lattice slots,
// round-0 flags, refinement rounds and run assembly (src/engine/sampling.h, window_impl.h) decide the sample
// set, the hook only renders.
//
//   wr_run     samples one window; returns a handle holding the runs (null on error, message in err)
//   wr_fetch   copies the runs (offsets n + 1, t, irr) and the window's counters out of the handle
//   wr_free    releases the handle
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <vector>

#include "bsx/bsx_scene.h"
#include "engine/window.h"
#include "../ballasim_amd/seam/scene_hook_bsx.h"

using namespace balla::engine;

namespace {

struct Result {
  std::vector<int64_t> offsets, t;
  std::vector<float> irr;
  // render calls, signature flags, midpoint flags, cap hits, then the samples of rounds 0..7
  uint64_t stats[12] = {};
};

}  // namespace

extern "C" {

void* wr_run(bsx_scene* scene, const bsx_view* view, int64_t t_w, float tol, float dark, int32_t max_refine_levels,
             uint32_t min_split, int32_t signature, const uint32_t* pixel, const uint8_t* level, size_t n, char* err,
             size_t err_len) {
  try {
    SampleConfig cfg;
    cfg.rules.fixed = false;
    cfg.rules.refine_signature = signature != 0;
    cfg.rules.midpoint = tol > 0.0f;
    cfg.rules.mid = midpoint_params(tol, dark, {waveform_range(balla::sensor::Waveform{})});
    cfg.rules.max_refine_levels = max_refine_levels;
    cfg.rules.min_split = min_split;
    cfg.groups = 1;
    BsxSceneHook<Space::kHost> hook(scene, *view, 1, nullptr);
    hook.set_window(t_w);
    // Each pixel's carried sample: its value at t_w, rendered by the same hook.
    std::vector<SampleRequest> req(n);
    std::vector<uint32_t> rows(n);
    for (size_t a = 0; a < n; ++a) {
      req[a] = SampleRequest{};
      req[a].pixel = pixel[a];
      req[a].k_count = cfg.subsamples;
      req[a].t_off_ns = 0;
      rows[a] = static_cast<uint32_t>(a);
    }
    std::vector<float> carried_irr(n);
    std::vector<uint64_t> carried_sig(n);
    hook.render(req.data(), n, rows.data(), carried_irr.data(), carried_sig.data(), nullptr, cfg.subsamples);
    WindowEngine<Space::kHost> engine;
    WindowEngine<Space::kHost>::Input in;
    in.t_w = t_w;
    in.pixel = pixel;
    in.level = level;
    in.n = n;
    in.carried_irr = carried_irr.data();
    in.carried_sig = carried_sig.data();
    const WindowStats st = engine.run(in, cfg, hook, size_t{1} << 36, nullptr);
    auto r = std::make_unique<Result>();
    r->offsets.assign(engine.offsets(), engine.offsets() + n + 1);
    r->t.assign(engine.t(), engine.t() + engine.samples());
    r->irr.assign(engine.irr(), engine.irr() + engine.samples());
    r->stats[0] = st.render_calls;
    r->stats[1] = st.flagged_signature;
    r->stats[2] = st.flagged_midpoint;
    r->stats[3] = st.refine_cap_hits;
    for (size_t k = 0; k < st.samples_by_round.size() && k < 8; ++k) r->stats[4 + k] = st.samples_by_round[k];
    return r.release();
  } catch (const std::exception& e) {
    if (err != nullptr && err_len > 0) {
      std::strncpy(err, e.what(), err_len - 1);
      err[err_len - 1] = '\0';
    }
    return nullptr;
  }
}

size_t wr_samples(const void* h) { return static_cast<const Result*>(h)->t.size(); }

void wr_fetch(const void* h, int64_t* offsets, int64_t* t, float* irr, uint64_t* stats) {
  const Result& r = *static_cast<const Result*>(h);
  std::memcpy(offsets, r.offsets.data(), r.offsets.size() * sizeof(int64_t));
  std::memcpy(t, r.t.data(), r.t.size() * sizeof(int64_t));
  std::memcpy(irr, r.irr.data(), r.irr.size() * sizeof(float));
  std::memcpy(stats, r.stats, sizeof r.stats);
}

void wr_free(void* h) { delete static_cast<Result*>(h); }

}  // extern "C"
