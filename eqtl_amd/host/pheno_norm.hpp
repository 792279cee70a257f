// Per-gene phenotype normalisation (--pheno-norm int; scan_common.hpp:124-126,
// util.cpp:132-205): rank-based inverse normal transform of the finite values,
// average ranks for ties, Phi^-1((rank - 0.5) / n_finite); non-finite values
// are kept. Header-only so that a unit check can build it on its own.
#pragma once
#include <algorithm>
#include <cmath>
#include <limits>
#include <utility>
#include <vector>

namespace eqtl {

// Standard normal quantile: the Abramowitz & Stegun 26.2.23 rational start
// (|error| < 4.5e-4) refined by Halley steps on the erfc-based CDF to full
// double precision. p outside (0, 1) is clamped to [1e-15, 1 - 1e-15].
inline double normal_quantile(double p) {
  if (!std::isfinite(p)) return std::numeric_limits<double>::quiet_NaN();
  p = std::min(std::max(p, 1e-15), 1.0 - 1e-15);
  const double pl = p < 0.5 ? p : 1.0 - p;
  const double t = std::sqrt(-2.0 * std::log(pl));
  double x = t - (2.515517 + t * (0.802853 + t * 0.010328)) / (1.0 + t * (1.432788 + t * (0.189269 + t * 0.001308)));
  if (p < 0.5) x = -x;
  const double inv_sqrt2pi = 0.39894228040143267794;
  for (int it = 0; it < 4; ++it) {
    // f = Phi(x) - p, computed in the tail that keeps it accurate
    const double f = x < 0 ? 0.5 * std::erfc(-x / std::sqrt(2.0)) - p : (1.0 - p) - 0.5 * std::erfc(x / std::sqrt(2.0));
    const double d = inv_sqrt2pi * std::exp(-0.5 * x * x);
    if (!(d > 0)) break;
    const double step = f / d;
    x -= step / (1.0 + 0.5 * x * step);  // Halley: f'' = -x f'
  }
  return x;
}

// inverse_normal_transform (util.cpp:177-205) in place on n values.
inline void inverse_normal_transform(double* y, size_t n) {
  std::vector<std::pair<double, size_t>> v;
  for (size_t i = 0; i < n; ++i)
    if (std::isfinite(y[i])) v.emplace_back(y[i], i);
  const size_t nf = v.size();
  if (n <= 1 || nf <= 1) return;
  std::sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  for (size_t i = 0; i < nf;) {
    size_t j = i + 1;
    while (j < nf && v[j].first == v[i].first) ++j;
    const double rank = 0.5 * static_cast<double>(i + 1 + j);  // average of the 1-based ranks i+1..j
    const double q = (rank - 0.5) / static_cast<double>(nf);
    const double z = normal_quantile(q);
    for (size_t k = i; k < j; ++k) y[v[k].second] = z;
    i = j;
  }
}

}  // namespace eqtl
