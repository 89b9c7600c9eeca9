// Host construction of the in-path noise tables in long double (reference:
// src/sensor/noise_tables.cpp:53-265 for the matrices, :341-411 for the rows). The operations and
// their order are the reference's, because the fp32 tables must equal its tables bit for bit: the
// kernels' noise events are functions of these bits.
//
// Phi(h) = e^{A h} and Q(h) = int_0^h e^{A s} G G^T e^{A^T s} ds come from scaling and squaring:
// a 16-term Taylor series on h / 2^s with ||A h / 2^s|| <= 1/64, then s doublings
// Q(2h) = Phi(h) Q(h) Phi(h)^T + Q(h), Phi(2h) = Phi(h)^2. A is Metzler, so Phi and Q are
// entrywise non-negative and neither the series nor the doublings cancel.
#include "noise_tables.h"

#include <cmath>
#include <stdexcept>

#include "../core/pmath.h"

namespace bsx {

namespace {

using LD = long double;

struct Mat3 {
  LD m[3][3] = {};

  static Mat3 eye() {
    Mat3 r;
    for (int i = 0; i < 3; ++i) r.m[i][i] = 1;
    return r;
  }
};

Mat3 operator*(const Mat3& a, const Mat3& b) {
  Mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      LD s = 0;
      for (int k = 0; k < 3; ++k) s += a.m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
  return r;
}

Mat3 operator+(const Mat3& a, const Mat3& b) {
  Mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
  return r;
}

Mat3 operator*(const Mat3& a, LD s) {
  Mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] * s;
  return r;
}

Mat3 transposed(const Mat3& a) {
  Mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}

// Drift and diffusion of the noise at frozen rate a1.
struct Drift {
  Mat3 A, GG;
};

Drift drift_of(const NoiseModelParams& p, LD a1) {
  Drift d;
  d.A.m[0][0] = -a1;
  d.A.m[1][0] = p.a2;
  d.A.m[1][1] = -p.a2;
  d.A.m[2][1] = p.a3;
  d.A.m[2][2] = -p.a3;
  d.GG.m[0][0] = 2 * static_cast<LD>(p.kq) * a1;  // keeps Var(n1) = k q
  d.GG.m[1][1] = 2 * static_cast<LD>(p.sigma_o) * p.sigma_o * p.a2;
  return d;
}

void transition(const Drift& d, LD h, Mat3& phi, Mat3& q) {
  LD norm = 0;
  for (int i = 0; i < 3; ++i) {
    LD row = 0;
    for (int j = 0; j < 3; ++j) row += std::fabs(d.A.m[i][j]);
    norm = std::max(norm, row);
  }
  int halvings = 0;
  LD delta = h;
  while (norm * delta > LD(1) / 64 && halvings < 200) {
    delta /= 2;
    ++halvings;
  }
  const Mat3 Ad = d.A * delta;
  Mat3 term = Mat3::eye();
  phi = term;
  for (int n = 1; n <= 16; ++n) {
    term = (term * Ad) * (LD(1) / n);
    phi = phi + term;
  }
  // Q(delta) = sum_n M_n delta^(n+1) / (n+1)!, M_0 = G G^T, M_n = A M_{n-1} + M_{n-1} A^T
  Mat3 Mn = d.GG;
  LD coef = delta;
  q = Mn * coef;
  const Mat3 At = transposed(d.A);
  for (int n = 1; n <= 16; ++n) {
    Mn = d.A * Mn + Mn * At;
    coef = coef * delta / (n + 1);
    q = q + Mn * coef;
  }
  for (int k = 0; k < halvings; ++k) {
    q = (phi * q) * transposed(phi) + q;
    phi = phi * phi;
  }
}

// The stationary covariance (the lower-bidiagonal Lyapunov equation solved row by row).
Mat3 stationary(const NoiseModelParams& p, LD a1) {
  const LD a2 = p.a2, a3 = p.a3;
  Mat3 s;
  s.m[0][0] = p.kq;
  s.m[1][0] = a2 * s.m[0][0] / (a1 + a2);
  s.m[1][1] = s.m[1][0] + static_cast<LD>(p.sigma_o) * p.sigma_o;
  s.m[2][0] = a3 * s.m[1][0] / (a1 + a3);
  s.m[2][1] = (a3 * s.m[1][1] + a2 * s.m[2][0]) / (a2 + a3);
  s.m[2][2] = s.m[2][1];
  s.m[0][1] = s.m[1][0];
  s.m[0][2] = s.m[2][0];
  s.m[1][2] = s.m[2][1];
  return s;
}

// Cholesky factor of a symmetric positive semidefinite N x N matrix, pivots at or below a tiny
// fraction of the trace taken as 0 (a numerically singular matrix still factors).
template <int N>
void cholesky(const LD (&a)[N][N], LD (&l)[N][N]) {
  LD trace = 0;
  for (int i = 0; i < N; ++i) trace += std::fabs(a[i][i]);
  const LD tiny = trace * LD(1e-36) + LD(1e-300);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) l[i][j] = 0;
  for (int j = 0; j < N; ++j) {
    LD d = a[j][j];
    for (int k = 0; k < j; ++k) d -= l[j][k] * l[j][k];
    d = d > tiny ? d : 0;
    l[j][j] = std::sqrt(d);
    for (int i = j + 1; i < N; ++i) {
      LD s = a[i][j];
      for (int k = 0; k < j; ++k) s -= l[i][k] * l[j][k];
      l[i][j] = l[j][j] > 0 ? s / l[j][j] : 0;
    }
  }
}

// (L L^T) x = b by forward and back substitution.
template <int N>
void chol_solve(const LD (&l)[N][N], const LD (&b)[N], LD (&x)[N]) {
  LD y[N];
  for (int i = 0; i < N; ++i) {
    LD s = b[i];
    for (int k = 0; k < i; ++k) s -= l[i][k] * y[k];
    y[i] = l[i][i] > 0 ? s / l[i][i] : 0;
  }
  for (int i = N - 1; i >= 0; --i) {
    LD s = y[i];
    for (int k = i + 1; k < N; ++k) s -= l[k][i] * x[k];
    x[i] = l[i][i] > 0 ? s / l[i][i] : 0;
  }
}

void symmetric_part(const Mat3& a, LD (&out)[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[i][j] = 0.5L * (a.m[i][j] + a.m[j][i]);
}

Mat3 chol3(const Mat3& a) {
  LD in[3][3], l[3][3];
  symmetric_part(a, in);
  cholesky<3>(in, l);
  Mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = l[i][j];
  return r;
}

Mat3 inverse_spd(const Mat3& a) {
  LD in[3][3], l[3][3];
  symmetric_part(a, in);
  cholesky<3>(in, l);
  Mat3 r;
  for (int j = 0; j < 3; ++j) {
    LD e[3] = {0, 0, 0}, x[3];
    e[j] = 1;
    chol_solve<3>(l, e, x);
    for (int i = 0; i < 3; ++i) r.m[i][j] = x[i];
  }
  return r;
}

void put_lower(const Mat3& a, double* out) {
  out[0] = static_cast<double>(a.m[0][0]);
  out[1] = static_cast<double>(a.m[1][0]);
  out[2] = static_cast<double>(a.m[1][1]);
  out[3] = static_cast<double>(a.m[2][0]);
  out[4] = static_cast<double>(a.m[2][1]);
  out[5] = static_cast<double>(a.m[2][2]);
}

// The exclusion margin of a node of length h: the largest bridge residual std of n3 over the
// node's interior points u = i / 20, the largest mismatch g between the optimal predictor and the
// cubic Hermite of n3 (slopes a3 (n2 - n3)) in the endpoint metric, and W = chol(C_ee)^-1 of the
// stationary endpoint pair's covariance C_ee.
void margin_row(const NoiseModelParams& p, const Drift& d, LD a1, LD h, double* out) {
  const Mat3 P = stationary(p, a1);
  constexpr int kPoints = 20;
  Mat3 phi_step, q_unused;
  transition(d, h / kPoints, phi_step, q_unused);
  Mat3 powers[kPoints + 1];
  powers[0] = Mat3::eye();
  for (int i = 1; i <= kPoints; ++i) powers[i] = powers[i - 1] * phi_step;
  const Mat3& phi_h = powers[kPoints];
  LD Cee[6][6];
  const Mat3 PphiT = P * transposed(phi_h), phiP = phi_h * P;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Cee[i][j] = P.m[i][j];
      Cee[i][j + 3] = PphiT.m[i][j];
      Cee[i + 3][j] = phiP.m[i][j];
      Cee[i + 3][j + 3] = P.m[i][j];
    }
  LD L6[6][6];
  cholesky<6>(Cee, L6);
  LD sigma_br = 0, g = 0;
  const LD a3 = p.a3;
  for (int iu = 1; iu < kPoints; ++iu) {
    const LD u = static_cast<LD>(iu) / kPoints;
    const Mat3 cov_l = powers[iu] * P, cov_r = P * transposed(powers[kPoints - iu]);
    LD c[6], w[6];
    for (int j = 0; j < 3; ++j) {
      c[j] = cov_l.m[2][j];
      c[j + 3] = cov_r.m[2][j];
    }
    chol_solve<6>(L6, c, w);
    LD var_br = P.m[2][2];
    for (int j = 0; j < 6; ++j) var_br -= c[j] * w[j];
    const LD H00 = 2 * u * u * u - 3 * u * u + 1, H10 = u * u * u - 2 * u * u + u;
    const LD H01 = -2 * u * u * u + 3 * u * u, H11 = u * u * u - u * u;
    const LD wh[6] = {0, H10 * h * a3, H00 - H10 * h * a3, 0, H11 * h * a3, H01 - H11 * h * a3};
    LD dw[6], g2 = 0;
    for (int j = 0; j < 6; ++j) dw[j] = w[j] - wh[j];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) g2 += dw[i] * Cee[i][j] * dw[j];
    sigma_br = std::max(sigma_br, std::sqrt(std::max(var_br, LD(0))));
    g = std::max(g, std::sqrt(std::max(g2, LD(0))));
  }
  out[0] = static_cast<double>(sigma_br);
  out[1] = static_cast<double>(g);
  LD W[6][6] = {};
  for (int j = 0; j < 6; ++j) {
    W[j][j] = L6[j][j] > 0 ? 1 / L6[j][j] : 0;
    for (int i = j + 1; i < 6; ++i) {
      LD s = 0;
      for (int k = j; k < i; ++k) s -= L6[i][k] * W[k][j];
      W[i][j] = L6[i][i] > 0 ? s / L6[i][i] : 0;
    }
  }
  int idx = 2;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) out[idx++] = static_cast<double>(W[i][j]);
}

}  // namespace

NoiseTableSet build_noise_tables(const NoiseModelParams& model) {
  if (model.levels < 0 || model.levels > kNoiseMaxLevels) {
    throw std::invalid_argument("noise tables: noise_leaf_level must lie in [0, 10]");
  }
  NoiseTableSet t;
  const double dln = pm::log(10.0) / kNoiseNodesPerDecade;
  t.inv_dln = 1.0 / dln;
  t.n_nodes = static_cast<int>(std::lround((kNoiseLnA1Max - kNoiseLnA1Min) * t.inv_dln)) + 1;
  t.levels = model.levels;
  t.stride = noise_stride(model.levels);
  t.data64.assign(static_cast<size_t>(t.n_nodes) * t.stride, 0.0);
  const LD h0 = static_cast<LD>(kNoiseCellNs) * 1e-9L;
  const int L = model.levels;
  for (int node = 0; node < t.n_nodes; ++node) {
    const LD a1 = std::exp(static_cast<LD>(kNoiseLnA1Min) + static_cast<LD>(node) * static_cast<LD>(dln));
    double* row = t.data64.data() + static_cast<size_t>(node) * t.stride;
    const Drift d = drift_of(model, a1);
    Mat3 phis[kNoiseMaxLevels + 1], qs[kNoiseMaxLevels + 1];
    for (int l = 0; l <= L; ++l) transition(d, h0 / static_cast<LD>(int64_t{1} << l), phis[l], qs[l]);
    put_lower(phis[0], row + 0);
    put_lower(chol3(qs[0]), row + 6);
    put_lower(chol3(stationary(model, a1)), row + 12);
    for (int l = 1; l <= L; ++l) {
      // The midpoint of a level-(l-1) node given both its ends: Gaussian conditioning, with the
      // parent's covariance Phi_l Q_l Phi_l^T + Q_l (the doubling step).
      const Mat3& ph = phis[l];
      const Mat3& qh = qs[l];
      const Mat3 parent = (ph * qh) * transposed(ph) + qh;
      const Mat3 K = (qh * transposed(ph)) * inverse_spd(parent);
      const Mat3 Am = ph + ((K * ph) * ph) * LD(-1);
      const Mat3 sig = qh + ((K * ph) * qh) * LD(-1);
      double* b = row + noise_bridge_offset(l);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          b[3 * i + j] = static_cast<double>(Am.m[i][j]);
          b[9 + 3 * i + j] = static_cast<double>(K.m[i][j]);
        }
      put_lower(chol3(sig), b + 18);
    }
    for (int l = 0; l < L; ++l) {
      margin_row(model, d, a1, h0 / static_cast<LD>(int64_t{1} << l), row + noise_margin_offset(L, l));
    }
  }
  t.data32.resize(t.data64.size());
  for (size_t i = 0; i < t.data64.size(); ++i) t.data32[i] = static_cast<float>(t.data64[i]);
  return t;
}

}  // namespace bsx
