// The instantaneous lane's searches on the host: on falling segments under leak, with Delta started
// just above -theta_off and the photocurrent's fall close to the leak's rise (Delta running level
// with the OFF threshold over long stretches), transition_search must return exactly what the
// generic time-ordered bisection (the reference's instant_search, src/sensor/pixel.h:1884-1916)
// returns: found, polarity, crossing time and log value. Prints the counts and the evaluations of
// both; exit status 1 on any mismatch.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "instant.h"

using bsx::InstantLane;

int main(int argc, char** argv) {
  const int trials = argc > 1 ? std::atoi(argv[1]) : 20000;
  const unsigned long long seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 12345;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<uint32_t> stack(InstantLane<4>::kSearchDepth);
  long n_found = 0, n_generic = 0, bad = 0, leaves_transition = 0, leaves_generic = 0;
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
    const double delta = U(rng) * 2e-6 * (it % 3 == 0 ? 100.0 : 1.0);
    L.s3_ref = static_cast<float>(static_cast<double>(x0) + L.dl_at(L.ta) + L.theta_off - delta);
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
    const bool same = a.found == b.found && (!a.found || (a.on == b.on && a.t == b.t && a.x == b.x));
    if (!same && ++bad <= 10)
      std::printf("mismatch at trial %d: transition found %d t %lld on %d, bisection found %d t %lld on %d\n", it,
                  a.found, static_cast<long long>(a.t), a.on, b.found, static_cast<long long>(b.t), b.on);
  }
  std::printf("trials %d found %ld generic %ld mismatches %ld evaluations transition %ld bisection %ld\n", trials,
              n_found, n_generic, bad, leaves_transition, leaves_generic);
  return bad != 0;
}
