// Tables of the cascade's in-path noise, built once per model on the host in long double and read by
// the kernels in fp32 (reference: NoiseTables, src/sensor/noise_tables.h:43-139, construction
// src/sensor/noise_tables.cpp:53-265, 341-411).
//
// The linearised noise n = (n1, n2, n3) of the three-stage photoreceptor obeys dn = A n dt + G dW
// with A = [[-a1, 0, 0], [a2, -a2, 0], [0, a3, -a3]] and G G^T = diag(2 k q a1, 2 sigma_o^2 a2, 0),
// a1 the frozen rate of a 2^16-ns cell. Only a1 varies, so every matrix the kernel needs is a row of
// a table over ln a1 in [ln 1e-2, ln 1e9] at 64 nodes per decade, interpolated linearly:
//   [0, 6)    the lattice step Phi(h0), lower triangle (00, 10, 11, 20, 21, 22)
//   [6, 12)   chol Q(h0), lower triangle
//   [12, 18)  chol of the stationary covariance (the initial state)
//   per level l = 1..L at 18 + 24 (l - 1): the midpoint bridge A_m (9, row-major), K (9), chol of its
//             residual covariance (6): x_m = A_m x_L + K x_R + chol xi
//   per level l = 0..L-1 at 18 + 24 L + 23 l: the exclusion margin's sigma_br, g and W (21, the
//             packed lower triangle of chol(C_ee)^-1)
// Frozen diffusion only: the reference's split tables of transient diffusion are not built.
// The table's bits follow the host's long double and libm's expl, as the reference's do.
#pragma once

#include <stdint.h>

#include <vector>

namespace bsx {

inline constexpr int kNoiseCellLog2 = 16;
inline constexpr int64_t kNoiseCellNs = int64_t{1} << kNoiseCellLog2;
inline constexpr double kNoiseLnA1Min = -4.605170185988091;  // ln 1e-2
inline constexpr double kNoiseLnA1Max = 20.72326583694641;   // ln 1e9
inline constexpr int kNoiseNodesPerDecade = 64;
inline constexpr int kNoiseMaxLevels = 10;
inline constexpr double kElementaryCharge = 1.602176634e-19;  // C (k q is the input node's variance)

inline constexpr int noise_stride(int levels) { return 18 + 47 * levels; }
inline constexpr int noise_bridge_offset(int l) { return 18 + 24 * (l - 1); }
inline constexpr int noise_margin_offset(int levels, int l) { return 18 + 24 * levels + 23 * l; }

struct NoiseModelParams {
  double kq;       // k q: the stationary input-node variance (log-e^2)
  double a2, a3;   // 1 / tau_o, 1 / tau_sf (1/s)
  double sigma_o;  // unfiltered output-node noise (log-e)
  int levels;      // leaf level L of the bridge tree
};

struct NoiseTableSet {
  int n_nodes = 0;
  int stride = 0;
  int levels = 0;
  double inv_dln = 0.0;        // 1 / (ln 10 / 64)
  std::vector<double> data64;  // n_nodes x stride
  std::vector<float> data32;   // its fp32 copy (what the kernels read)
};

// Throws std::invalid_argument for levels outside [0, 10].
NoiseTableSet build_noise_tables(const NoiseModelParams& model);

}  // namespace bsx
