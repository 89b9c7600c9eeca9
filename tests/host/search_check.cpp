// The instantaneous lane's searches on the host: on falling segments under leak, with Delta started
// just above -theta_off or just below theta_on and the photocurrent's fall close to the leak's rise
// (Delta running level with a threshold over long stretches), transition_search must return exactly what the
// generic time-ordered bisection (the reference's instant_search, src/sensor/pixel.h:1884-1916)
// returns: found, polarity, crossing time and log value. Prints the counts and the evaluations of
// both; exit status 1 on any mismatch. With a third argument "screen", checks the photocurrent screen
// instead (screen_check).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "instant.h"

using bsx::InstantLane;

// The photocurrent screen (InstantLane::screen_bounds, screened) on random segments: one to three
// groups rising, falling or mixed, with and without leak, photocurrents from 1e-3 to 1e9 of i_scale,
// Delta at the segment's start placed near either level (inside the screen's margin, just outside it,
// or anywhere between the levels). Whenever the screen passes a segment, the generic bisection over
// [t_a, t_b) must find no crossing.
static int screen_check(int trials, unsigned long long seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<uint32_t> stack(InstantLane<4>::kSearchDepth);
  long quiet = 0, near_quiet = 0, loud = 0, bad = 0;
  for (int it = 0; it < trials; ++it) {
    InstantLane<4> L{};
    L.G = 1 + it % 3;
    L.theta_on = 0.15f + 0.03f * static_cast<float>(U(rng) - 0.5);
    L.theta_off = 0.15f + 0.03f * static_cast<float>(U(rng) - 0.5);
    L.leak = it % 4 == 0 ? 0.0f : static_cast<float>(0.01 + 2.0 * U(rng));
    L.dark = static_cast<float>(1e-3 * U(rng) + 1e-6);
    L.scale = 1.0f;
    L.irr_max = 1e30f;
    L.i_scale = static_cast<float>(std::exp(U(rng) * 12.0 - 6.0));
    const int64_t span = 1 + static_cast<int64_t>(U(rng) * (it % 5 == 0 ? 8e6 : 2e4));
    L.ta = static_cast<int64_t>(U(rng) * 1e9);
    L.tb = L.ta + span;
    L.t_ref = L.ta - static_cast<int64_t>(U(rng) * 5e7);
    L.dr_ref = static_cast<float>(0.2 * (U(rng) - 0.5));
    const double swing = (U(rng) - 0.5) * (it % 3 == 0 ? 1e-3 : 0.3);  // ln I change over the segment
    const double I0 = L.i_scale * std::exp(U(rng) * 27.0 - 7.0);
    for (int g = 0; g < L.G; ++g) {
      const double w = g == 0 ? 1.0 : 0.3 * U(rng);
      const double sg = g == 0 || U(rng) < 0.7 ? swing : -swing;  // some segments mix directions
      L.va[g] = static_cast<float>(I0 * w);
      L.vb[g] = static_cast<float>(I0 * w * std::exp(sg));
    }
    float x0;
    L.s3_ref = 0.0f;
    (void)L.s_at(L.ta, x0);
    // Delta(t_a) near either level: within a few margins of it, or anywhere between the levels
    const double r = U(rng), m = 0x1p-9;
    double target = (U(rng) - 0.5) * 0.2;
    if (r < 0.35) target = L.theta_on - m * (0.5 + 3.0 * U(rng)) - std::fabs(swing);
    else if (r < 0.7) target = -L.theta_off + m * (0.5 + 3.0 * U(rng)) + std::fabs(swing);
    L.s3_ref = static_cast<float>(static_cast<double>(x0) + L.dl_at(L.ta) - target);
    if (it % 11 == 0) L.va[0] = -0.0f;  // a negative zero sample (clampv passes it)
    if (it % 13 == 0) L.dark = 0.0f;
    L.screen_bounds(L.ta, L.tb - 1);
    if (!L.screened()) {
      ++loud;
      continue;
    }
    ++quiet;
    if (r < 0.7) ++near_quiet;
    const InstantLane<4>::Hit b = L.search(L.ta, L.tb, stack.data(), 1);
    if (b.found && ++bad <= 10)
      std::printf("screen mismatch at trial %d: screened, but the bisection fires at %lld (t_a %lld)\n", it,
                  static_cast<long long>(b.t), static_cast<long long>(L.ta));
  }
  std::printf("screen trials %d quiet %ld near_quiet %ld loud %ld mismatches %ld\n", trials, quiet, near_quiet, loud,
              bad);
  return bad != 0;
}

int main(int argc, char** argv) {
  const int trials = argc > 1 ? std::atoi(argv[1]) : 20000;
  const unsigned long long seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 12345;
  if (argc > 3 && std::string(argv[3]) == "screen") return screen_check(trials, seed);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<uint32_t> stack(InstantLane<4>::kSearchDepth);
  long n_found = 0, n_on = 0, n_generic = 0, bad = 0, leaves_transition = 0, leaves_generic = 0;
  for (int it = 0; it < trials; ++it) {
    InstantLane<4> L{};
    L.G = 1 + it % 3;
    L.theta_on = 0.15f + 0.03f * static_cast<float>(U(rng) - 0.5);
    L.theta_off = 0.15f + 0.03f * static_cast<float>(U(rng) - 0.5);
    L.leak = static_cast<float>(it % 7 == 0 ? 5.0 * U(rng) : 0.01 + 0.5 * U(rng));
    L.dark = static_cast<float>(1e-3 * U(rng) + 1e-6);
    L.scale = 1.0f;
    L.irr_max = 1e30f;
    L.i_scale = static_cast<float>(0.5 + U(rng));
    L.n3_ref = 0.0f;
    const int64_t span = 1 + static_cast<int64_t>(U(rng) * (it % 5 == 0 ? 8e6 : 2e5));
    L.ta = static_cast<int64_t>(U(rng) * 1e9);
    L.tb = L.ta + span + 1;
    L.t_ref = L.ta - static_cast<int64_t>(U(rng) * 5e7);
    L.dr_ref = static_cast<float>(0.2 * (U(rng) - 0.5));
    // ln I falls by 0.5 .. 1.5 times the leak's rise over the segment.
    const double fall = L.leak * static_cast<double>(span) * 1e-9 * (0.5 + U(rng));
    const double I0 = 0.1 + U(rng), I1 = I0 * std::exp(-fall);
    for (int g = 0; g < L.G; ++g) {
      const double w = g == 0 ? 1.0 : 0.3 * U(rng);
      L.va[g] = static_cast<float>(I0 * w);
      L.vb[g] = static_cast<float>(I1 * w);
      if (L.vb[g] > L.va[g]) L.vb[g] = L.va[g];
    }
    // s3_ref places Delta(ta) at -theta_off + delta.
    float x0;
    L.s3_ref = 0.0f;
    (void)L.s_at(L.ta, x0);
    // Delta(ta) just above -theta_off, or (odd trials) just below theta_on, where the leak may carry
    // a slowly falling Delta up to the ON level (both levels within reach)
    const double delta = U(rng) * 2e-6 * (it % 3 == 0 ? 100.0 : 1.0);
    const double level = it % 2 == 1 ? static_cast<double>(L.theta_on) - delta : -static_cast<double>(L.theta_off) + delta;
    L.s3_ref = static_cast<float>(static_cast<double>(x0) + L.dl_at(L.ta) - level);
    const int64_t lo = L.ta + static_cast<int64_t>(U(rng) * 0.1 * static_cast<double>(span));
    const int64_t hi = L.tb;
    InstantLane<4>::Hit a{false, false, lo, 0.0f};
    L.n_leaves = 0;
    const bool decided = L.transition_search(lo, hi, a, stack.data(), 1);
    leaves_transition += L.n_leaves;
    L.n_leaves = 0;
    const InstantLane<4>::Hit b = L.search(lo, hi, stack.data(), 1);
    leaves_generic += L.n_leaves;
    if (!decided) {
      ++n_generic;
      continue;
    }
    n_found += b.found;
    n_on += b.found && b.on;
    const bool same = a.found == b.found && (!a.found || (a.on == b.on && a.t == b.t && a.x == b.x));
    if (!same && ++bad <= 10)
      std::printf("mismatch at trial %d: transition found %d t %lld on %d, bisection found %d t %lld on %d\n", it,
                  a.found, static_cast<long long>(a.t), a.on, b.found, static_cast<long long>(b.t), b.on);
  }
  std::printf("trials %d found %ld on %ld generic %ld mismatches %ld evaluations transition %ld bisection %ld\n",
              trials, n_found, n_on, n_generic, bad, leaves_transition, leaves_generic);
  return bad != 0;
}
