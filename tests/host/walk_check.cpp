// The instantaneous lane program on the host, on random runs: the walk as the kernels run it (from
// phase A's first loud segment, found by the log-free photocurrent screen under the input state, each
// piece screened with the current reference before any exact search, in the walk/search alternation
// with small budgets; and again with phase A's screen blocks, aligned 16-segment photocurrent ranges
// at a random offset of the run in the camera's segment space, which the walk passes at once when
// they clear its screen) must give exactly the events and final state of the walk on exact searches
// alone, segment by segment. Prints the counts; exit status 1 on any difference.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "instant.h"

using bsx::InstantLane;
using bsx::RawEventDev;

struct VecSink {
  std::vector<RawEventDev> ev;
  void push(const RawEventDev& e) { ev.push_back(e); }
};

static bool same_event(const RawEventDev& a, const RawEventDev& b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }

int main(int argc, char** argv) {
  const int trials = argc > 1 ? std::atoi(argv[1]) : 2000;
  const unsigned long long seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 7;
  constexpr int kG = 2;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::normal_distribution<double> N01(0.0, 1.0);
  long n_events = 0, n_screened = 0, n_block_skips = 0, bad = 0;
  for (int it = 0; it < trials; ++it) {
    const int G = 1 + it % 2;
    const int n = 20 + static_cast<int>(U(rng) * 300);
    const int64_t t_state = static_cast<int64_t>(U(rng) * 1e9);
    // samples: a log random walk with rare jumps; some runs quiet, some moving
    std::vector<int64_t> ts(n);
    std::vector<float> irr(static_cast<size_t>(n) * G);
    const double vol = it % 3 == 0 ? 0.002 : it % 3 == 1 ? 0.02 : 0.1;
    double lv[kG] = {std::log(0.5 + U(rng)), std::log(0.2 + U(rng))};
    int64_t t = t_state + (it % 7 == 0 ? 0 : 1 + static_cast<int64_t>(U(rng) * 5000));
    const bool steady = it % 3 == 2;  // long stretches of constant light, some near a level
    for (int i = 0; i < n; ++i) {
      ts[i] = t;
      const bool hold = steady && i > 0 && U(rng) < 0.97;
      for (int g = 0; g < G; ++g) {
        if (hold) {
          irr[static_cast<size_t>(i) * G + g] = irr[static_cast<size_t>(i - 1) * G + g];
          continue;
        }
        lv[g] += vol * N01(rng) + (U(rng) < 0.02 ? 0.5 * N01(rng) : 0.0);
        irr[static_cast<size_t>(i) * G + g] = static_cast<float>(std::exp(lv[g]));
      }
      t += 1 + static_cast<int64_t>(U(rng) * (it % 4 == 0 ? 2e6 : 3e4));
    }
    const int64_t t_end = ts[n - 1] + static_cast<int64_t>(U(rng) * 1e5);
    InstantLane<kG> base{};
    base.pixel = static_cast<uint32_t>(it);
    base.G = G;
    base.theta_on = static_cast<float>(0.15 + 0.05 * (U(rng) - 0.5));
    base.theta_off = static_cast<float>(0.15 + 0.05 * (U(rng) - 0.5));
    base.leak = it % 5 == 0 ? 0.0f : static_cast<float>(base.theta_on * (0.05 + 5.0 * U(rng)));
    base.dark = static_cast<float>(1e-3 * U(rng) + 1e-6);
    base.scale = 1.0f;
    base.irr_max = 1e30f;
    base.eps_ci = 1e-4f;
    base.label_cut = 0.5f;
    base.refractory_ns = it % 2 == 0 ? 50000 : 0;
    base.ideal_reset = it % 6 == 5;
    base.ts = ts.data();
    base.irr = irr.data();
    base.n = n;
    base.t_state = t_state;
    base.skip = ts[0] == t_state ? 1 : 0;
    for (int g = 0; g < G; ++g) base.held[g] = static_cast<float>(0.5 + U(rng));
    base.i_scale = static_cast<float>(0.3 + U(rng));
    base.s3_ref = static_cast<float>(std::log(base.held[0] / base.i_scale) + 0.05 * N01(rng));
    if (steady && it % 2 == 0) {  // the held light's Delta within 2^-11 of a level: inside the screen's margin
      float I0 = base.dark;
      for (int g = 0; g < G; ++g) I0 = I0 + base.scale * base.held[g];
      const float x0 = base.log_ratio(I0);
      base.s3_ref = x0 - (it % 4 == 0 ? base.theta_on : -base.theta_off) * (1.0f - 0x1p-11f);
    }
    base.n3_ref = 0.0f;
    base.dr_ref = 0.0f;
    base.t_ref = t_state;
    base.blind_until = bsx::kNotBlind;
    base.t_search = t_state + 1;
    base.flags = 0;
    base.shot_rate = it % 4 == 3 ? 2e3 : 0.0;
    base.seed = 99;
    base.camera = 0;
    base.shot_next = base.shot_rate > 0.0 ? t_state + InstantLane<kG>::shot_gap(base.shot_rate, 99, 0, base.pixel, 0, nullptr)
                                          : INT64_MAX;
    base.shot_index = 0;
    base.n_substeps = base.n_pieces = base.n_leaves = 0;
    VecSink s0, s2;
    InstantLane<kG> A = base, C = base;
    A.screen = false;  // the reference: every piece by the exact search
    A.advance(t_end, s0);
    // the kernel's path: phase A's log-free screen of the segments under the input state (the rule of
    // chain_screen_kernel) gives the first loud segment, and the walk screens each piece from there
    int first_loud = n;
    {
      InstantLane<kG> S = base;
      const bool usable = S.blind_until == bsx::kNotBlind && S.leak >= 0.0f && S.n3_ref == 0.0f;
      if (usable) S.screen_bounds(t_state, t_end - 1);
      for (int j = 0; j < n; ++j) {
        if (j == 0 && ts[0] == t_state) continue;
        if (j == 0) {
          S.ta = t_state;
          for (int g = 0; g < G; ++g) S.va[g] = S.held[g];
        } else {
          S.ta = ts[j - 1];
          for (int g = 0; g < G; ++g) S.va[g] = S.clampv(irr[static_cast<size_t>(j - 1) * G + g]);
        }
        S.tb = ts[j];
        for (int g = 0; g < G; ++g) S.vb[g] = S.clampv(irr[static_cast<size_t>(j) * G + g]);
        const bool quiet = usable && S.shot_next >= S.tb &&
                           ((S.ta >> bsx::kPlainSegmentLog2) + 1) * bsx::kPlainSegmentNs >= S.tb && S.screened();
        if (!quiet) {
          first_loud = j;
          break;
        }
      }
    }
    n_screened += first_loud - C.skip;
    // phase A's screen blocks (chain_screen_kernel's rule): aligned blocks of 16 camera positions, the
    // run's first sample at position q_base; a block is valid when its 16 positions are segments of
    // the run, and holds the extremes of their segment_range
    const int64_t q_base = static_cast<int64_t>(U(rng) * 40);
    std::vector<bsx::ScreenBlock> blocks(static_cast<size_t>((q_base + n) / bsx::kScreenBlock + 1),
                                         bsx::ScreenBlock{0.0f, INFINITY, 0u});
    {
      InstantLane<kG> S = base;
      for (int64_t qb0 = (q_base + bsx::kScreenBlock - 1) / bsx::kScreenBlock * bsx::kScreenBlock;
           qb0 + bsx::kScreenBlock <= q_base + n; qb0 += bsx::kScreenBlock) {
        bool valid = true, steady_blk = true;
        float lo = INFINITY, hi = -INFINITY;
        for (int64_t q = qb0; q < qb0 + bsx::kScreenBlock; ++q) {
          const int j = static_cast<int>(q - q_base);
          if (j == 0 && ts[0] == t_state) {
            valid = false;
            continue;
          }
          S.ta = j == 0 ? t_state : ts[j - 1];
          for (int g = 0; g < G; ++g)
            S.va[g] = j == 0 ? S.held[g] : S.clampv(irr[static_cast<size_t>(j - 1) * G + g]);
          S.tb = ts[j];
          for (int g = 0; g < G; ++g) S.vb[g] = S.clampv(irr[static_cast<size_t>(j) * G + g]);
          float a, b;
          S.segment_range(a, b);
          lo = std::fmin(lo, a);
          hi = std::fmax(hi, b);
          steady_blk = steady_blk && S.steady_segment();
        }
        if (valid) blocks[static_cast<size_t>(qb0 / bsx::kScreenBlock)] = bsx::ScreenBlock{lo, hi, steady_blk ? 1u : 0u};
      }
    }
    InstantLane<kG> E = base;
    E.qb = blocks.data();
    E.q_base = q_base;
    VecSink s3;
    const int budget = 1 + it % 4;
    for (InstantLane<kG>* X : {&C, &E}) {
      VecSink& sx = X == &C ? s2 : s3;
      uint32_t stack2[InstantLane<kG>::kSearchDepth];
      X->begin(t_end, first_loud);
      for (int m = X->walk(sx, budget); m != InstantLane<kG>::kDone; m = X->walk(sx, budget))
        if (m == InstantLane<kG>::kSearch) X->search_step(sx, stack2, 1);
      X->end();
    }
    n_block_skips += E.n_block_skips;
    for (InstantLane<kG>* X : {&C, &E}) {
      const VecSink& sx = X == &C ? s2 : s3;
      bool ok = s0.ev.size() == sx.ev.size();
      for (size_t k = 0; ok && k < s0.ev.size(); ++k) ok = same_event(s0.ev[k], sx.ev[k]);
      ok = ok && A.t_search == X->t_search && A.t_ref == X->t_ref && A.blind_until == X->blind_until &&
           std::memcmp(&A.s3_ref, &X->s3_ref, 4) == 0 && std::memcmp(&A.dr_ref, &X->dr_ref, 4) == 0 &&
           A.flags == X->flags && A.shot_next == X->shot_next && A.shot_index == X->shot_index &&
           A.n_substeps == X->n_substeps && A.n_pieces == X->n_pieces &&
           std::memcmp(A.held, X->held, sizeof(A.held)) == 0;
      if (!ok && ++bad <= 10)
        std::printf("difference at trial %d (%s): events %zu vs %zu, t_search %lld vs %lld, pieces %u vs %u\n", it,
                    X == &C ? "screened walk" : "with screen blocks", s0.ev.size(), sx.ev.size(),
                    static_cast<long long>(A.t_search), static_cast<long long>(X->t_search), A.n_pieces, X->n_pieces);
    }
    n_events += static_cast<long>(s0.ev.size());
  }
  std::printf("walk trials %d events %ld screened %ld block_skips %ld differences %ld\n", trials, n_events, n_screened,
              n_block_skips, bad);
  return bad != 0;
}
