// The in-plane swimmer ISF table (host, FP64): the reference's isf2dFast
// (motility-ddm src/motility_fit.js:168-219; legendreP :120-132).  72 rows on a uniform ln Z grid
// over [0.5, 100]; row i holds g(Lambda; Z_i) = P_Z(1/sqrt(1+L^2)) /
// (1+L^2)^((Z+1)/2) at ln Lambda = u0 + k du, k < n_i + 2, with du_i =
// min(0.04, 0.1 / ((Z_i + 1/2)/2)) (one step advances the oscillation phase by
// at most 0.1 rad) and rows ending where the envelope falls below e^-25.
// Lookups: Catmull-Rom along ln Lambda within a row, quadratic Lagrange across
// the three rows around Z.  The rows are stored back to back (row i at off[i])
// so the same flat arrays serve the host evaluation and the device copy the
// per-q LM interpolates.
#pragma once

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace ddm {

struct Isf2dTable {
  static constexpr int nz = 72;
  static constexpr double lser = 1e-4;   // below: the small-Lambda series
  double lz0 = 0, dlz = 0;
  std::vector<double> Z, u0, du, v;
  std::vector<int> n, off;
};

// P_nu(c) by the trapezoid rule on Laplace's first integral (:120-132)
inline double legendre_p_host(double nu, double c) {
  const double s = std::sqrt(std::max(0.0, 1 - c * c));
  const int n = 16 + (int)std::ceil(nu * std::atan2(s, c));
  double acc = 0;
  for (int i = 0; i <= n; i++) {
    const double im = s * std::cos(M_PI * i / n);
    const double val = std::exp(0.5 * nu * std::log(c * c + im * im)) * std::cos(nu * std::atan2(im, c));
    acc += (i == 0 || i == n) ? 0.5 * val : val;
  }
  return acc / n;
}

inline double isf2d_lambda_host(double Lm, double Z) {
  const double u = 1 + Lm * Lm;
  const double ln_pre = -0.5 * (Z + 1) * std::log(u);
  if (ln_pre < -25) return 0;
  return legendre_p_host(Z, 1 / std::sqrt(u)) * std::exp(ln_pre);
}

inline const Isf2dTable& isf2d_table() {
  static Isf2dTable t;
  static std::once_flag once;
  std::call_once(once, [] {
    t.lz0 = std::log(0.5);
    t.dlz = (std::log(100.0) - t.lz0) / (Isf2dTable::nz - 1);
    for (int i = 0; i < Isf2dTable::nz; i++) {
      const double Z = std::exp(t.lz0 + i * t.dlz);
      const double Lc = std::min(1e4, std::sqrt(std::expm1(50 / (Z + 1))));
      const double u0 = std::log(Isf2dTable::lser), u1 = std::log(Lc);
      const double du = std::min(0.04, 0.1 / ((Z + 0.5) / 2));
      const int n = (int)std::ceil((u1 - u0) / du) + 1;
      t.Z.push_back(Z);
      t.u0.push_back(u0);
      t.du.push_back(du);
      t.n.push_back(n);
      t.off.push_back((int)t.v.size());
      for (int k = 0; k < n + 2; k++) t.v.push_back(isf2d_lambda_host(std::exp(u0 + k * du), Z));
    }
  });
  return t;
}

}  // namespace ddm
