// Beta-approximated permutation p-value (beta_approx_p, stats_extra.cpp:79-158;
// the QTLtools procedure): fit Beta(a, b) to a gene's permutation minimum
// p-values by moments, refine by maximum likelihood with a Nelder-Mead
// simplex, and report the Beta CDF at the observed minimum p.
//
// The reference drives GSL's `nmsimplex2` minimizer (GSL is not part of the
// reference tree; its version is the system's). The simplex below follows that
// published algorithm: start simplex x0, x0 + step_i e_i with steps a/10, b/10;
// each iteration reflects the worst corner through the centroid of the
// simplex (corner move coefficient -1), tries an expansion (-2) when the
// reflection beats the best, a one-dimensional contraction (0.5) when it does
// not beat the second worst, and otherwise shrinks every corner half-way to the
// best; the size is the root-mean-square distance of the corners from their
// centre, updated incrementally as corners move; the loop stops when the size
// falls below 0.01 or after 1000 iterations. Shape bounds a in [0.1, 10],
// b in [5, 1e6] enter as a 1e300 penalty, and a failed or out-of-bounds fit
// falls back to the moment estimate (QTLtools behaviour).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "numerics.hpp"

namespace eqtl_amd {

struct BetaFitStats {
  double s1, s2, n;  // sum log p, sum log(1 - p), count
};

__device__ inline bool beta_in_bounds(double a, double b) {
  return a >= 0.1 && a <= 10.0 && b >= 5.0 && b <= 1e6;
}

// Stirling remainder of ln Gamma(z) for z >= 20: 1/(12z) - 1/(360z^3) +
// 1/(1260z^5) - 1/(1680z^7) + 1/(1188z^9) (next term < 1e-17).
__device__ inline double lgamma_stirling_rem(double z) {
  const double r = 1.0 / z, r2 = r * r;
  return r * (1.0 / 12 + r2 * (-1.0 / 360 + r2 * (1.0 / 1260 + r2 * (-1.0 / 1680 + r2 * (1.0 / 1188)))));
}

// ln B(a, b) to full precision for the shape range of the fit (the
// reference evaluates gsl_sf_lnbeta, accurate to rounding). For b >= 20,
// ln Gamma(b) - ln Gamma(a+b) is taken from the Stirling series as
// -a ln b - (a+b-1/2) log1p(a/b) + a + rem(b) - rem(a+b), which avoids the
// cancellation of two values near b ln b.
__device__ inline double lnbeta_fit(double a, double b) {
  if (b < 20.0) return lgamma(a) + lgamma(b) - lgamma(a + b);
  const double d = -a * log(b) - (a + b - 0.5) * log1p(a / b) + a + lgamma_stirling_rem(b) -
                   lgamma_stirling_rem(a + b);
  return lgamma(a) + d;
}

__device__ inline double beta_fit_nll(double a, double b, const BetaFitStats& st) {
  if (!beta_in_bounds(a, b)) return 1e300;
  return -((a - 1.0) * st.s1 + (b - 1.0) * st.s2 - st.n * lnbeta_fit(a, b));
}

// Nelder-Mead on (a, b) from (a0, b0); returns false when the size test did
// not succeed or the result leaves the bounds.
__device__ inline bool beta_fit_simplex(const BetaFitStats& st, double& a, double& b) {
  constexpr int P = 3;  // corners
  double x[P][2], y[P];
  x[0][0] = a; x[0][1] = b;
  x[1][0] = a + a / 10.0; x[1][1] = b;
  x[2][0] = a; x[2][1] = b + b / 10.0;
  for (int i = 0; i < P; ++i) y[i] = beta_fit_nll(x[i][0], x[i][1], st);
  double c[2];
  auto center = [&]() {
    for (int j = 0; j < 2; ++j) c[j] = (x[0][j] + x[1][j] + x[2][j]) / P;
  };
  auto size2 = [&]() {
    double s = 0;
    for (int i = 0; i < P; ++i)
      for (int j = 0; j < 2; ++j) s += (x[i][j] - c[j]) * (x[i][j] - c[j]);
    return s / P;
  };
  center();
  double S2 = size2();
  // corner `k` moved by coefficient `coeff` about the centroid of the others
  auto corner_move = [&](double coeff, int k, double* out) {
    const double al = (1.0 - coeff) * P / (P - 1.0), be = (P * coeff - 1.0) / (P - 1.0);
    for (int j = 0; j < 2; ++j) out[j] = al * c[j] + be * x[k][j];
    return beta_fit_nll(out[0], out[1], st);
  };
  auto update = [&](int k, const double* xn, double val) {
    double d2 = 0, xmcd = 0;
    for (int j = 0; j < 2; ++j) {
      const double dl = xn[j] - x[k][j];
      d2 += dl * dl;
      xmcd += (x[k][j] - c[j]) * dl;
    }
    S2 += (2.0 / P) * xmcd + ((P - 1.0) / P) * d2 / P;
    for (int j = 0; j < 2; ++j) {
      c[j] += (xn[j] - x[k][j]) / P;
      x[k][j] = xn[j];
    }
    y[k] = val;
  };
  bool success = false;
  for (int iter = 1; iter <= 1000; ++iter) {
    // highest, second highest and lowest corners
    int hi = 0, lo = 0, s_hi = 1;
    double dhi = y[0], dlo = y[0], ds_hi = y[1];
    for (int i = 1; i < P; ++i) {
      const double v = y[i];
      if (v < dlo) {
        dlo = v;
        lo = i;
      } else if (v > dhi) {
        ds_hi = dhi;
        s_hi = hi;
        dhi = v;
        hi = i;
      } else if (v > ds_hi) {
        ds_hi = v;
        s_hi = i;
      }
    }
    double xc[2], xc2[2];
    const double val = corner_move(-1.0, hi, xc);
    bool bad = false;
    if (isfinite(val) && val < y[lo]) {
      const double val2 = corner_move(-2.0, hi, xc2);
      if (isfinite(val2) && val2 < y[lo]) update(hi, xc2, val2);
      else update(hi, xc, val);
    } else if (!isfinite(val) || val > y[s_hi]) {
      if (isfinite(val) && val <= y[hi]) update(hi, xc, val);
      const double val2 = corner_move(0.5, hi, xc2);
      if (isfinite(val2) && val2 <= y[hi]) {
        update(hi, xc2, val2);
      } else {
        for (int i = 0; i < P; ++i) {
          if (i == lo) continue;
          for (int j = 0; j < 2; ++j) x[i][j] = 0.5 * (x[i][j] + x[lo][j]);
          y[i] = beta_fit_nll(x[i][0], x[i][1], st);
          if (!isfinite(y[i])) bad = true;
        }
        center();
        S2 = size2();
      }
    } else {
      update(hi, xc, val);
    }
    if (bad) break;
    const double size = S2 > 0 ? sqrt(S2) : sqrt(size2());
    if (size < 0.01) {
      success = true;
      break;
    }
  }
  int best = 0;
  for (int i = 1; i < P; ++i)
    if (y[i] < y[best]) best = i;
  a = x[best][0];
  b = x[best][1];
  return success && isfinite(a) && isfinite(b) && beta_in_bounds(a, b);
}

// Regularized incomplete beta I_x(a, b) (beta_cdf, util.cpp:208-257).
__device__ inline double beta_cdf_dev(double x, double a, double b) {
  if (a <= 0 || b <= 0) return __builtin_nan("");
  if (x <= 0) return 0;
  if (x >= 1) return 1;
  const double lbeta = lgamma(a) + lgamma(b) - lgamma(a + b);
  const double front = exp(log(x) * a + log(1 - x) * b - lbeta) / a;
  if (x < (a + 1) / (a + b + 2)) return front * betacf_dev(a, b, x);
  return 1.0 - front * betacf_dev(b, a, 1 - x) * a / b;
}

// One thread per gene: perm_min_p row k (B values), obs[k] -> p_beta, shapes.
__global__ void k_beta_approx(const double* perm_min_p, const double* obs, int G, int B, double* p_beta,
                              double* shape1, double* shape2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  const double* v = perm_min_p + static_cast<size_t>(k) * B;
  const double nan = __builtin_nan("");
  int cnt = 0;
  double sum = 0;
  for (int i = 0; i < B; ++i)
    if (isfinite(v[i]) && v[i] > 0.0 && v[i] < 1.0) {
      ++cnt;
      sum += v[i];
    }
  if (cnt < 10 || !isfinite(obs[k])) {
    p_beta[k] = shape1[k] = shape2[k] = nan;
    return;
  }
  // moment matching (beta_fit_mm)
  const double mean = sum / cnt;
  double var = 0;
  BetaFitStats st{0, 0, 0};
  for (int i = 0; i < B; ++i)
    if (isfinite(v[i]) && v[i] > 0.0 && v[i] < 1.0) {
      const double d = v[i] - mean;
      var += d * d;
      st.s1 += log(v[i]);
      st.s2 += log(1.0 - v[i]);
      st.n += 1;
    }
  var /= static_cast<double>(cnt - 1);
  const double t = (var > 0 && mean > 0 && mean < 1) ? mean * (1.0 - mean) / var - 1.0 : -1.0;
  if (!(t > 0)) {
    p_beta[k] = shape1[k] = shape2[k] = nan;
    return;
  }
  const double mm1 = mean * t, mm2 = (1.0 - mean) * t;
  double a = mm1, b = mm2;
  if (!beta_fit_simplex(st, a, b)) {
    a = mm1;
    b = mm2;
  }
  double o = obs[k];
  if (o < 1e-300) o = 1e-300;
  if (o > 1.0 - 1e-16) o = 1.0 - 1e-16;
  const double pb = beta_cdf_dev(o, a, b);
  p_beta[k] = isfinite(pb) ? pb : nan;
  shape1[k] = a;
  shape2[k] = b;
}

}  // namespace eqtl_amd
