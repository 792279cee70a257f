// eqtl oracle — FP64 CPU restatement of the eqtl association engine's hot path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (the eqtl_amd package, its
// HIP library or its C++ host code) links, loads or calls this file. It is
// imported only by tests/, by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg, where it is the checker (and the timed CPU baseline),
// never the thing measured or shipped.
//
// Every function names the reference file:line it restates. Layouts:
//   dosage  m x n, SNP-major (row j = SNP j's n dosages)
//   Q       n x n column-major (column c = eigenvector c), lambda ascending
//   Y       n x G column-major (gene k contiguous), X n x p column-major
//
// Pinning (what checks each function beyond this restatement) is listed in
// the header comment of oracle/__init__.py.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <random>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

// ---------------------------------------------------------------------------
// Student-t two-sided tail: util.cpp:209-264 (betacf / betai / p_from_t).
// Numerical-Recipes-style Lentz continued fraction, eps 3e-12, <=200 iters.
// ---------------------------------------------------------------------------
double betacf(double a, double b, double x) {
  const int maxit = 200;
  const double eps = 3e-12, fpmin = 1e-30;
  const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0, d = 1.0 - qab * x / qap;
  if (std::fabs(d) < fpmin) d = fpmin;
  d = 1.0 / d;
  double h = d;
  for (int m = 1; m <= maxit; ++m) {
    const int m2 = 2 * m;
    double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
    d = 1.0 + aa * d;
    if (std::fabs(d) < fpmin) d = fpmin;
    c = 1.0 + aa / c;
    if (std::fabs(c) < fpmin) c = fpmin;
    d = 1.0 / d;
    h *= d * c;
    aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
    d = 1.0 + aa * d;
    if (std::fabs(d) < fpmin) d = fpmin;
    c = 1.0 + aa / c;
    if (std::fabs(c) < fpmin) c = fpmin;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) < eps) break;
  }
  return h;
}

double betai(double a, double b, double x) {
  if (x < 0 || x > 1) return kNaN;
  if (x == 0 || x == 1) return x;
  const double lbeta = std::lgamma(a) + std::lgamma(b) - std::lgamma(a + b);
  const double front = std::exp(std::log(x) * a + std::log(1 - x) * b - lbeta) / a;
  if (x < (a + 1) / (a + b + 2)) return front * betacf(a, b, x);
  return 1.0 - front * betacf(b, a, 1 - x) * a / b;
}

double p_from_t(double t, double df) {
  if (!std::isfinite(t) || df <= 0) return 1.0;
  const double x = df / (df + t * t);
  const double p = betai(0.5 * df, 0.5, x);
  return std::min(1.0, std::max(0.0, p));
}

// ---------------------------------------------------------------------------
// ACAT accumulator: stats_extra.cpp:16-31 (x87 long double tan sum).
// ---------------------------------------------------------------------------
struct Acat {
  long double tan_sum = 0.0L;
  int used = 0;
  void add(double p) {
    static const long double kPi = 3.14159265358979323846L;
    if (!std::isfinite(p) || p < 0.0) return;
    if (p >= 1.0 - 1e-12) return;
    if (p <= 0.0) p = 1e-300;
    tan_sum += std::tan((0.5L - static_cast<long double>(p)) * kPi);
    ++used;
  }
  double p() const {
    static const long double kPi = 3.14159265358979323846L;
    if (used == 0) return 1.0;
    const long double T = tan_sum / static_cast<long double>(used);
    double ac = 0.5 - static_cast<double>(std::atan(T) / kPi);
    if (ac < 0) ac = 0;
    if (ac > 1) ac = 1;
    return ac;
  }
};

// ---------------------------------------------------------------------------
// LDLT with diagonal pivoting on the not-yet-updated trailing diagonal, the
// pivot rule of Eigen 3.4's ldlt_inplace<Lower>::unblocked (the dependency the
// reference's model_lmm.cpp:23,141 and model_lm.cpp:25 call).
// A is p x p column-major (lower triangle used); on return L is unit-lower in
// A's strict lower part, D in d[], perm[] the transpositions.
// Returns false on a non-finite result.
// ---------------------------------------------------------------------------
struct Ldlt {
  int p = 0;
  std::vector<double> L, D;
  std::vector<int> tr;
  bool ok = false;
  void compute(const double* A, int pp) {
    p = pp;
    L.assign(A, A + static_cast<size_t>(p) * p);
    D.assign(p, 0.0);
    tr.assign(p, 0);
    std::vector<double> temp(p);
    auto m = [&](int i, int j) -> double& { return L[static_cast<size_t>(j) * p + i]; };
    // symmetrize from the lower triangle
    for (int j = 0; j < p; ++j)
      for (int i = 0; i < j; ++i) m(i, j) = m(j, i);
    ok = true;
    for (int k = 0; k < p; ++k) {
      int big = k;
      double bv = std::fabs(m(k, k));
      for (int i = k + 1; i < p; ++i)
        if (std::fabs(m(i, i)) > bv) { bv = std::fabs(m(i, i)); big = i; }
      tr[k] = big;
      if (big != k) {
        // swap rows/cols k <-> big in the lower-triangular storage
        for (int j = 0; j < k; ++j) std::swap(m(k, j), m(big, j));
        for (int i = big + 1; i < p; ++i) std::swap(m(i, k), m(i, big));
        std::swap(m(k, k), m(big, big));
        for (int i = k + 1; i < big; ++i) {
          const double t = m(i, k);
          m(i, k) = m(big, i);
          m(big, i) = t;
        }
      }
      if (k > 0) {
        double s = 0;
        for (int j = 0; j < k; ++j) { temp[j] = D[j] * m(k, j); s += m(k, j) * temp[j]; }
        m(k, k) -= s;
        for (int i = k + 1; i < p; ++i) {
          double t = 0;
          for (int j = 0; j < k; ++j) t += m(i, j) * temp[j];
          m(i, k) -= t;
        }
      }
      const double akk = m(k, k);
      D[k] = akk;
      if (std::fabs(akk) > 0.0) {
        for (int i = k + 1; i < p; ++i) m(i, k) /= akk;
      } else if (k == 0) {
        // all-zero diagonal: leave as is; solve() will yield zeros
      }
      if (!std::isfinite(akk)) ok = false;
    }
  }
  // x := A^{-1} b
  void solve(const double* b, double* x) const {
    std::vector<double> y(b, b + p);
    for (int k = 0; k < p; ++k) std::swap(y[k], y[tr[k]]);
    auto l = [&](int i, int j) { return L[static_cast<size_t>(j) * p + i]; };
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < i; ++j) y[i] -= l(i, j) * y[j];
    for (int i = 0; i < p; ++i) y[i] = (D[i] != 0.0) ? y[i] / D[i] : 0.0;
    for (int i = p - 1; i >= 0; --i)
      for (int j = i + 1; j < p; ++j) y[i] -= l(j, i) * y[j];
    for (int k = p - 1; k >= 0; --k) std::swap(y[k], y[tr[k]]);
    std::memcpy(x, y.data(), sizeof(double) * p);
  }
};

// ---------------------------------------------------------------------------
// Spectral LMM pieces: model_lmm.cpp:10-86 (reml_negll, optimize_delta),
// :115-153 (prep_lmm), :160-223 (test_lmm) and scan_trans_lmm.cpp:11-70
// (test_lmm_gtil).
// ---------------------------------------------------------------------------
double reml_negll(double delta, const double* yt, const double* Xt, const double* lam, int n,
                  int p, int df) {
  std::vector<double> dinv(n);
  double logdet_d = 0;
  for (int i = 0; i < n; ++i) {
    double v = delta * lam[i] + 1.0;
    if (v < 1e-12) v = 1e-12;
    dinv[i] = 1.0 / v;
    logdet_d += std::log(v);
  }
  std::vector<double> A(static_cast<size_t>(p) * p, 0.0), b(p, 0.0), beta(p);
  for (int a = 0; a < p; ++a) {
    const double* xa = Xt + static_cast<size_t>(a) * n;
    for (int c = 0; c <= a; ++c) {
      const double* xc = Xt + static_cast<size_t>(c) * n;
      double s = 0;
      for (int i = 0; i < n; ++i) s += xa[i] * dinv[i] * xc[i];
      A[static_cast<size_t>(c) * p + a] = s;
      A[static_cast<size_t>(a) * p + c] = s;
    }
    double s = 0;
    for (int i = 0; i < n; ++i) s += xa[i] * (dinv[i] * yt[i]);
    b[a] = s;
  }
  Ldlt f;
  f.compute(A.data(), p);
  if (!f.ok) return 1e300;
  f.solve(b.data(), beta.data());
  double ydy = 0;
  for (int i = 0; i < n; ++i) ydy += yt[i] * (dinv[i] * yt[i]);
  double q = ydy;
  for (int a = 0; a < p; ++a) q -= b[a] * beta[a];
  if (q <= 0) q = 1e-12;
  const double sigma2 = q / df;
  double logdet_x = 0;
  for (int a = 0; a < p; ++a) {
    if (f.D[a] <= 0) return 1e300;
    logdet_x += std::log(f.D[a]);
  }
  return 0.5 * (df * std::log(sigma2) + logdet_d + logdet_x);
}

double optimize_delta(const double* yt, const double* Xt, const double* lam, int n, int p) {
  const int df = n - p;
  if (df <= 0) return 1.0;
  double best_d = 1.0;
  double best_ll = reml_negll(1.0, yt, Xt, lam, n, p, df);
  for (double d = 1e-5; d <= 1e5; d *= 2.0) {
    const double v = reml_negll(d, yt, Xt, lam, n, p, df);
    if (v < best_ll) { best_ll = v; best_d = d; }
  }
  double lo = best_d / 4.0, hi = best_d * 4.0;
  if (lo < 1e-8) lo = 1e-8;
  if (hi > 1e6) hi = 1e6;
  const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  for (int it = 0; it < 40; ++it) {
    const double m1 = hi - (hi - lo) / phi;
    const double m2 = lo + (hi - lo) / phi;
    const double f1 = reml_negll(m1, yt, Xt, lam, n, p, df);
    const double f2 = reml_negll(m2, yt, Xt, lam, n, p, df);
    if (f1 < f2) hi = m2;
    else lo = m1;
  }
  double d = 0.5 * (lo + hi);
  double cur = reml_negll(d, yt, Xt, lam, n, p, df);
  for (int it = 0; it < 8; ++it) {
    const double h = std::max(d * 1e-4, 1e-9);
    const double fm = reml_negll(d - h, yt, Xt, lam, n, p, df);
    const double fp = reml_negll(d + h, yt, Xt, lam, n, p, df);
    const double g1 = (fp - fm) / (2.0 * h);
    const double g2 = (fp - 2.0 * cur + fm) / (h * h);
    if (!(g2 > 1e-12)) break;
    const double step = g1 / g2;
    const double dn = d - step;
    if (!(dn > 1e-8 && dn < 1e6)) break;
    const double vn = reml_negll(dn, yt, Xt, lam, n, p, df);
    if (!(vn < cur)) break;
    d = dn;
    cur = vn;
    if (std::abs(step) < d * 1e-8) break;
  }
  return d;
}

struct PrepLmm {
  int n = 0, p = 0;
  std::vector<double> yt, Xt, dinv, chi0;
  double delta = 1, rss_null = 0, y_dy = 0;
  bool has_a00 = false;
  Ldlt a00;
};

// model_lmm.cpp:124-153 with Q^T y, Q^T X supplied (yt, Xt) and delta either
// searched (delta_in <= 0) or given.
void prep_lmm(const double* yt, const double* Xt, const double* lam, int n, int p,
              double delta_in, PrepLmm& P) {
  P.n = n;
  P.p = p;
  P.yt.assign(yt, yt + n);
  P.Xt.assign(Xt, Xt + static_cast<size_t>(n) * p);
  P.delta = delta_in > 0 ? delta_in : optimize_delta(yt, Xt, lam, n, p);
  P.dinv.resize(n);
  for (int i = 0; i < n; ++i) {
    double v = P.delta * lam[i] + 1.0;
    if (v < 1e-12) v = 1e-12;
    P.dinv[i] = 1.0 / v;
  }
  std::vector<double> A(static_cast<size_t>(p) * p), b(p);
  for (int a = 0; a < p; ++a) {
    const double* xa = Xt + static_cast<size_t>(a) * n;
    for (int c = 0; c <= a; ++c) {
      const double* xc = Xt + static_cast<size_t>(c) * n;
      double s = 0;
      for (int i = 0; i < n; ++i) s += xa[i] * P.dinv[i] * xc[i];
      A[static_cast<size_t>(c) * p + a] = s;
      A[static_cast<size_t>(a) * p + c] = s;
    }
    double s = 0;
    for (int i = 0; i < n; ++i) s += xa[i] * (P.dinv[i] * yt[i]);
    b[a] = s;
  }
  double ydy = 0;
  for (int i = 0; i < n; ++i) ydy += yt[i] * (P.dinv[i] * yt[i]);
  P.y_dy = ydy;
  P.a00.compute(A.data(), p);
  P.chi0.assign(p, 0.0);
  if (P.a00.ok) {
    P.a00.solve(b.data(), P.chi0.data());
    double r = ydy;
    for (int a = 0; a < p; ++a) r -= b[a] * P.chi0[a];
    P.rss_null = r < 0 ? 0 : r;
    P.has_a00 = true;
  } else {
    P.rss_null = ydy;
    P.has_a00 = false;
  }
}

struct Hit {
  double beta = 0, se = 0, stat = 0, p = 1, r2 = 0;
};

// scan_trans_lmm.cpp:11-70 (test_lmm_gtil), gt = Q^T g precomputed.
Hit test_lmm_gtil(const PrepLmm& P, const double* gt) {
  Hit h;
  const int n = P.n, p = P.p;
  const int df = n - p - 1;
  if (P.has_a00) {
    std::vector<double> Dg(n), a(p), u(p);
    double gg = 0;
    for (int i = 0; i < n; ++i) { Dg[i] = P.dinv[i] * gt[i]; gg += Dg[i] * gt[i]; }
    if (gg < 1e-12) { h.p = 1.0; return h; }
    for (int c = 0; c < p; ++c) {
      const double* xc = P.Xt.data() + static_cast<size_t>(c) * n;
      double s = 0;
      for (int i = 0; i < n; ++i) s += xc[i] * Dg[i];
      a[c] = s;
    }
    double yg = 0;
    for (int i = 0; i < n; ++i) yg += Dg[i] * P.yt[i];
    P.a00.solve(a.data(), u.data());
    double aTu = 0, ac = 0;
    for (int c = 0; c < p; ++c) { aTu += a[c] * u[c]; ac += a[c] * P.chi0[c]; }
    const double S = gg - aTu;
    if (S <= 1e-15) { h.p = 1.0; return h; }
    const double bg = (yg - ac) / S;
    double q = P.rss_null - S * bg * bg;
    if (q < 0) q = 0;
    const double sigma2 = (df > 0) ? (q / df) : 1.0;
    h.beta = bg;
    h.se = std::sqrt(std::max(sigma2 / S, 0.0));
    h.stat = (h.se > 0) ? (h.beta / h.se) : 0.0;
    h.p = p_from_t(h.stat, df);
    h.r2 = (P.rss_null > 1e-15) ? std::max(0.0, 1.0 - q / P.rss_null) : 0.0;
    return h;
  }
  // slow path (model_lmm.cpp:197-222): full (p+1) LDLT
  const int p1 = p + 1;
  double gws = 0;
  for (int i = 0; i < n; ++i) gws += gt[i] * P.dinv[i] * gt[i];
  if (gws < 1e-12) { h.p = 1.0; return h; }
  auto col = [&](int c) { return c < p ? P.Xt.data() + static_cast<size_t>(c) * n : gt; };
  std::vector<double> A(static_cast<size_t>(p1) * p1), b(p1), beta(p1), e(p1, 0.0), cc(p1);
  for (int a = 0; a < p1; ++a) {
    for (int c = 0; c <= a; ++c) {
      double s = 0;
      const double *xa = col(a), *xc = col(c);
      for (int i = 0; i < n; ++i) s += xa[i] * P.dinv[i] * xc[i];
      A[static_cast<size_t>(c) * p1 + a] = s;
      A[static_cast<size_t>(a) * p1 + c] = s;
    }
    double s = 0;
    const double* xa = col(a);
    for (int i = 0; i < n; ++i) s += xa[i] * (P.dinv[i] * P.yt[i]);
    b[a] = s;
  }
  Ldlt f;
  f.compute(A.data(), p1);
  if (!f.ok) { h.p = 1.0; return h; }
  f.solve(b.data(), beta.data());
  h.beta = beta[p];
  double q = P.y_dy;
  for (int a = 0; a < p1; ++a) q -= b[a] * beta[a];
  if (q < 0) q = 0;
  const double sigma2 = (df > 0) ? (q / df) : 1.0;
  e[p] = 1.0;
  f.solve(e.data(), cc.data());
  h.se = std::sqrt(std::max(sigma2 * cc[p], 0.0));
  h.stat = (h.se > 0) ? (h.beta / h.se) : 0.0;
  h.p = p_from_t(h.stat, df);
  h.r2 = (P.rss_null > 1e-15) ? std::max(0.0, 1.0 - q / P.rss_null) : 0.0;
  return h;
}

// ---------------------------------------------------------------------------
// LM: model_lm.cpp:19-57, scan_trans_lm.cpp:11-36.
// ---------------------------------------------------------------------------
struct PrepLm {
  int n = 0, p = 0;
  std::vector<double> X, XtXinv, ys;
  double yty = 0;
};

void prep_lm(const double* y, const double* X, int n, int p, PrepLm& P) {
  P.n = n;
  P.p = p;
  P.X.assign(X, X + static_cast<size_t>(n) * p);
  std::vector<double> A(static_cast<size_t>(p) * p), e(p), col(p);
  for (int a = 0; a < p; ++a)
    for (int c = 0; c <= a; ++c) {
      double s = 0;
      for (int i = 0; i < n; ++i) s += X[static_cast<size_t>(a) * n + i] * X[static_cast<size_t>(c) * n + i];
      A[static_cast<size_t>(c) * p + a] = s;
      A[static_cast<size_t>(a) * p + c] = s;
    }
  Ldlt f;
  f.compute(A.data(), p);
  P.XtXinv.assign(static_cast<size_t>(p) * p, 0.0);
  for (int c = 0; c < p; ++c) {
    std::fill(e.begin(), e.end(), 0.0);
    e[c] = 1.0;
    f.solve(e.data(), col.data());
    for (int a = 0; a < p; ++a) P.XtXinv[static_cast<size_t>(c) * p + a] = col[a];
  }
  std::vector<double> Xty(p, 0.0), beta(p, 0.0);
  for (int a = 0; a < p; ++a) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += X[static_cast<size_t>(a) * n + i] * y[i];
    Xty[a] = s;
  }
  for (int a = 0; a < p; ++a) {
    double s = 0;
    for (int c = 0; c < p; ++c) s += P.XtXinv[static_cast<size_t>(c) * p + a] * Xty[c];
    beta[a] = s;
  }
  P.ys.resize(n);
  double yy = 0;
  for (int i = 0; i < n; ++i) {
    double fit = 0;
    for (int a = 0; a < p; ++a) fit += X[static_cast<size_t>(a) * n + i] * beta[a];
    P.ys[i] = y[i] - fit;
    yy += P.ys[i] * P.ys[i];
  }
  P.yty = yy;
}

// residualize_g (scan_trans_lm.cpp:11-19) → g_s, gtg
double residualize_g(const PrepLm& P, const double* g, double* gs) {
  const int n = P.n, p = P.p;
  std::vector<double> Xtg(p), w(p);
  for (int a = 0; a < p; ++a) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += P.X[static_cast<size_t>(a) * n + i] * g[i];
    Xtg[a] = s;
  }
  for (int a = 0; a < p; ++a) {
    double s = 0;
    for (int c = 0; c < p; ++c) s += P.XtXinv[static_cast<size_t>(c) * p + a] * Xtg[c];
    w[a] = s;
  }
  double gtg = 0;
  for (int i = 0; i < n; ++i) {
    double fit = 0;
    for (int a = 0; a < p; ++a) fit += P.X[static_cast<size_t>(a) * n + i] * w[a];
    gs[i] = g[i] - fit;
    gtg += gs[i] * gs[i];
  }
  return gtg;
}

// hit_from_gty (scan_trans_lm.cpp:21-36) == tail of test_lm (model_lm.cpp:38-56)
Hit hit_from_gty(const PrepLm& P, double gtg, double gty) {
  Hit h;
  if (gtg < 1e-12) { h.p = 1.0; return h; }
  h.beta = gty / gtg;
  const double df = P.n - P.p - 1;
  if (df <= 0) { h.p = 1.0; return h; }
  const double rss = P.yty - h.beta * gty;
  const double s2 = std::max(rss / df, 0.0);
  h.se = std::sqrt(s2 / gtg);
  h.stat = (h.se > 0) ? (h.beta / h.se) : 0.0;
  h.p = p_from_t(h.stat, df);
  h.r2 = (P.yty > 0) ? (1.0 - rss / P.yty) : 0.0;
  return h;
}

// subset_maf_or_nan: scan_common.cpp:133-151
double subset_maf_or_nan(const double* g, int n, double* maf_out) {
  if (n <= 0) return kNaN;
  double sum = 0.0, sum2 = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!std::isfinite(g[i])) return kNaN;
    sum += g[i];
    sum2 += g[i] * g[i];
  }
  double af = (sum / static_cast<double>(n)) / 2.0;
  if (af < 0) af = 0;
  if (af > 1) af = 1;
  if (maf_out) *maf_out = af;
  const double maf = af > 0.5 ? 1.0 - af : af;
  if (maf < 1e-12) return kNaN;
  const double var = sum2 / n - (sum / n) * (sum / n);
  if (var < 1e-12) return kNaN;
  return af;
}

// Cyclic Jacobi eigensolver (FP64). Test reference for the eigendecomposition
// (make_lmm_basis, model_lmm.cpp:106-113): ascending eigenvalues clamped at 0.
void jacobi_eigh(const double* K, int n, double* Q, double* lam) {
  std::vector<double> A(K, K + static_cast<size_t>(n) * n);
  auto a = [&](int i, int j) -> double& { return A[static_cast<size_t>(j) * n + i]; };
  std::vector<double> V(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) V[static_cast<size_t>(i) * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double v = a(i, j) * a(i, j);
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int pidx = 0; pidx < n - 1; ++pidx)
      for (int qidx = pidx + 1; qidx < n; ++qidx) {
        const double apq = a(pidx, qidx);
        if (std::fabs(apq) < 1e-300) continue;
        const double app = a(pidx, pidx), aqq = a(qidx, qidx);
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = a(k, pidx), akq = a(k, qidx);
          a(k, pidx) = c * akp - s * akq;
          a(k, qidx) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a(pidx, k), aqk = a(qidx, k);
          a(pidx, k) = c * apk - s * aqk;
          a(qidx, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double& vp = V[static_cast<size_t>(pidx) * n + k];
          double& vq = V[static_cast<size_t>(qidx) * n + k];
          const double x = vp, y = vq;
          vp = c * x - s * y;
          vq = s * x + c * y;
        }
      }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  for (int c = 0; c < n; ++c) {
    lam[c] = std::max(a(ord[c], ord[c]), 0.0);
    std::memcpy(Q + static_cast<size_t>(c) * n, V.data() + static_cast<size_t>(ord[c]) * n,
                sizeof(double) * n);
  }
}

unsigned fnv1a(const char* s) {  // util.hpp:51-58
  unsigned long long h = 14695981039346656037ULL;
  for (const unsigned char* c = reinterpret_cast<const unsigned char*>(s); *c; ++c) {
    h ^= *c;
    h *= 1099511628211ULL;
  }
  return static_cast<unsigned>(h & 0xFFFFFFFFu);
}

// Permutation index stream: scan_cis.hpp:106-111 / scan_perm.cpp:84-88 —
// std::mt19937((seed>=0?seed:1) + 9973*b + fnv1a(gene)), std::shuffle(iota(n)).
void perm_index(int seed, int b, const char* gene, int n, int* idx) {
  std::mt19937 rng(static_cast<unsigned>(seed >= 0 ? seed : 1) + static_cast<unsigned>(b) * 9973u +
                   fnv1a(gene));
  std::vector<int> v(n);
  std::iota(v.begin(), v.end(), 0);
  std::shuffle(v.begin(), v.end(), rng);
  std::memcpy(idx, v.data(), sizeof(int) * n);
}

}  // namespace

// ===========================================================================
// C ABI of the oracle (ctypes from oracle/__init__.py)
// ===========================================================================
extern "C" {

double orc_p_from_t(double t, double df) { return p_from_t(t, df); }
double orc_betai(double a, double b, double x) { return betai(a, b, x); }

void orc_p_from_t_vec(const double* t, int64_t k, double df, double* p) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < k; ++i) p[i] = p_from_t(t[i], df);
}

double orc_acat(const double* p, int64_t k) {
  Acat a;
  for (int64_t i = 0; i < k; ++i) a.add(p[i]);
  return a.p();
}

// p_emp_count: stats_extra.cpp:39-46
double orc_p_emp(double Tobs, const double* Tperm, int32_t B) {
  int ge = 0;
  for (int b = 0; b < B; ++b)
    if (Tperm[b] >= Tobs) ++ge;
  return (1.0 + ge) / (1.0 + B);
}

// beta_approx_p (stats_extra.cpp:79-158): moment fit, Nelder-Mead ML refinement
// as GSL's nmsimplex2 does it (start simplex x0, x0 + x0_i/10 e_i; reflect -1,
// expand -2, 1-D contraction 0.5, shrink to best; RMS corner size < 0.01 or
// 1000 iterations), bounds by a 1e300 penalty, CDF by the continued fraction.
// out: p_beta, shape1, shape2 (NaN when fewer than 10 minima in (0,1) or the
// moment fit fails).
// ln B(a, b) as gsl_sf_lnbeta gives it (to rounding): for b >= 20 the
// difference ln Gamma(b) - ln Gamma(a+b) comes from Stirling's series,
// -a ln b - (a+b-1/2) log1p(a/b) + a + R(b) - R(a+b), with
// R(z) = 1/(12z) - 1/(360z^3) + 1/(1260z^5) - 1/(1680z^7) + 1/(1188z^9).
static double stirling_rem(double z) {
  const double r = 1.0 / z, r2 = r * r;
  return r * (1.0 / 12 + r2 * (-1.0 / 360 + r2 * (1.0 / 1260 + r2 * (-1.0 / 1680 + r2 * (1.0 / 1188)))));
}

double orc_lnbeta(double a, double b) {
  if (b < 20.0) return std::lgamma(a) + std::lgamma(b) - std::lgamma(a + b);
  return std::lgamma(a) - a * std::log(b) - (a + b - 0.5) * std::log1p(a / b) + a + stirling_rem(b) -
         stirling_rem(a + b);
}

static double beta_nll_orc(double a, double b, double s1, double s2, double n) {
  if (!(a >= 0.1 && a <= 10.0 && b >= 5.0 && b <= 1e6)) return 1e300;
  return -((a - 1.0) * s1 + (b - 1.0) * s2 - n * orc_lnbeta(a, b));
}

void orc_beta_approx(const double* pm, int32_t B, double obs, double* out) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  out[0] = out[1] = out[2] = nan;
  std::vector<double> x;
  for (int i = 0; i < B; ++i)
    if (std::isfinite(pm[i]) && pm[i] > 0.0 && pm[i] < 1.0) x.push_back(pm[i]);
  if (x.size() < 10) return;
  double mean = 0;
  for (double v : x) mean += v;
  mean /= x.size();
  double var = 0;
  for (double v : x) var += (v - mean) * (v - mean);
  var /= static_cast<double>(x.size() - 1);
  if (var <= 0 || mean <= 0 || mean >= 1) return;
  const double t = mean * (1.0 - mean) / var - 1.0;
  if (t <= 0) return;
  const double mm1 = mean * t, mm2 = (1.0 - mean) * t;
  double s1 = 0, s2 = 0;
  for (double v : x) {
    s1 += std::log(v);
    s2 += std::log(1.0 - v);
  }
  const double nn = static_cast<double>(x.size());
  auto F = [&](const double* q) { return beta_nll_orc(q[0], q[1], s1, s2, nn); };
  const int P = 3;
  double X[3][2] = {{mm1, mm2}, {mm1 + mm1 / 10.0, mm2}, {mm1, mm2 + mm2 / 10.0}}, Y[3];
  for (int i = 0; i < P; ++i) Y[i] = F(X[i]);
  double C[2];
  auto centre = [&] { for (int j = 0; j < 2; ++j) C[j] = (X[0][j] + X[1][j] + X[2][j]) / P; };
  auto sz2 = [&] {
    double r = 0;
    for (int i = 0; i < P; ++i)
      for (int j = 0; j < 2; ++j) r += (X[i][j] - C[j]) * (X[i][j] - C[j]);
    return r / P;
  };
  centre();
  double S2 = sz2();
  auto move = [&](double cf, int k, double* o) {
    const double al = (1 - cf) * P / (P - 1.0), be = (P * cf - 1.0) / (P - 1.0);
    for (int j = 0; j < 2; ++j) o[j] = al * C[j] + be * X[k][j];
    return F(o);
  };
  auto put = [&](int k, const double* q, double v) {
    double d2 = 0, xd = 0;
    for (int j = 0; j < 2; ++j) {
      const double dl = q[j] - X[k][j];
      d2 += dl * dl;
      xd += (X[k][j] - C[j]) * dl;
    }
    S2 += (2.0 / P) * xd + ((P - 1.0) / P) * d2 / P;
    for (int j = 0; j < 2; ++j) {
      C[j] += (q[j] - X[k][j]) / P;
      X[k][j] = q[j];
    }
    Y[k] = v;
  };
  bool ok = false;
  for (int it = 0; it < 1000; ++it) {
    int hi = 0, lo = 0, shi = 1;
    double dhi = Y[0], dlo = Y[0], dshi = Y[1];
    for (int i = 1; i < P; ++i) {
      if (Y[i] < dlo) { dlo = Y[i]; lo = i; }
      else if (Y[i] > dhi) { dshi = dhi; shi = hi; dhi = Y[i]; hi = i; }
      else if (Y[i] > dshi) { dshi = Y[i]; shi = i; }
    }
    double q1[2], q2[2];
    const double v1 = move(-1.0, hi, q1);
    if (v1 < Y[lo]) {
      const double v2 = move(-2.0, hi, q2);
      if (v2 < Y[lo]) put(hi, q2, v2);
      else put(hi, q1, v1);
    } else if (v1 > Y[shi]) {
      if (v1 <= Y[hi]) put(hi, q1, v1);
      const double v2 = move(0.5, hi, q2);
      if (v2 <= Y[hi]) {
        put(hi, q2, v2);
      } else {
        for (int i = 0; i < P; ++i) {
          if (i == lo) continue;
          for (int j = 0; j < 2; ++j) X[i][j] = 0.5 * (X[i][j] + X[lo][j]);
          Y[i] = F(X[i]);
        }
        centre();
        S2 = sz2();
      }
    } else {
      put(hi, q1, v1);
    }
    if ((S2 > 0 ? std::sqrt(S2) : std::sqrt(sz2())) < 0.01) { ok = true; break; }
  }
  int best = 0;
  for (int i = 1; i < P; ++i)
    if (Y[i] < Y[best]) best = i;
  double a = X[best][0], b = X[best][1];
  if (!(ok && std::isfinite(a) && std::isfinite(b) && a >= 0.1 && a <= 10.0 && b >= 5.0 && b <= 1e6)) {
    a = mm1;
    b = mm2;
  }
  double o = obs;
  if (o < 1e-300) o = 1e-300;
  if (o > 1.0 - 1e-16) o = 1.0 - 1e-16;
  const double pb = betai(a, b, o);
  out[0] = std::isfinite(pb) ? pb : nan;
  out[1] = a;
  out[2] = b;
}

// bh_adjust: scan_common.cpp:201-217
void orc_bh(const double* acat, int32_t G, double* q) {
  if (G <= 0) return;
  std::vector<int> idx(G);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return acat[a] < acat[b]; });
  double prev = 1.0;
  for (int rank = G; rank >= 1; --rank) {
    const int gi = idx[rank - 1];
    double v = acat[gi] * static_cast<double>(G) / static_cast<double>(rank);
    if (v > prev) v = prev;
    if (v > 1.0) v = 1.0;
    q[gi] = v;
    prev = v;
  }
}

double orc_subset_maf_or_nan(const double* g, int32_t n, double* maf_out) {
  return subset_maf_or_nan(g, n, maf_out);
}

// decode_row: plink_bed.cpp:263-320 (LUT {0->2, 1->missing, 2->1, 3->0}),
// applied to m SNP rows of bytes_per_snp bytes. keep[j]=1 if the SNP passes.
// dosage m x n (imputed per policy; NaN kept under no-impute), miss m x n.
void orc_decode_bed(const uint8_t* rows, int64_t m, int32_t bytes_per_snp, int32_t n_file,
                    const int32_t* sample_col, int32_t n, int32_t impute, double max_miss,
                    double maf_min, double* dosage, int8_t* miss, uint8_t* keep, double* af) {
  static const int lut[4] = {2, -1, 1, 0};
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < m; ++j) {
    const uint8_t* row = rows + j * static_cast<int64_t>(bytes_per_snp);
    double* d = dosage + j * static_cast<int64_t>(n);
    int8_t* mm = miss + j * static_cast<int64_t>(n);
    int n_miss = 0, n_ok = 0;
    double sum = 0.0;
    bool bad = false;
    for (int i = 0; i < n; ++i) {
      const int col = sample_col[i];
      if (col < 0 || col >= n_file) { bad = true; break; }
      const int pair = (row[col / 4] >> (2 * (col % 4))) & 3;
      const int a1c = lut[pair];
      if (a1c < 0) {
        d[i] = kNaN;
        mm[i] = 1;
        ++n_miss;
        continue;
      }
      mm[i] = 0;
      d[i] = static_cast<double>(a1c);
      sum += a1c;
      ++n_ok;
    }
    keep[j] = 0;
    af[j] = kNaN;
    if (bad || n_ok == 0) continue;
    const double miss_frac = static_cast<double>(n_miss) / static_cast<double>(n);
    if (miss_frac > max_miss + 1e-15) continue;
    if (!impute && n_miss > 0) continue;
    if (n_miss > 0 && impute) {
      const double mu = sum / n_ok;
      for (int i = 0; i < n; ++i)
        if (!std::isfinite(d[i])) d[i] = mu;
    }
    double a = (sum / static_cast<double>(n_ok)) / 2.0;
    if (a < 0) a = 0;
    if (a > 1) a = 1;
    if (a < 1e-12 || a > 1.0 - 1e-12) continue;
    if (maf_min > 0.0 && (a < maf_min || a > 1.0 - maf_min)) continue;  // pass_maf, util.hpp:61-64
    af[j] = a;
    keep[j] = 1;
  }
}

// ---- VCF record (VcfSession::parse_record, vcf_session.cpp:160-286, and the
// pass_maf gate of for_each_snp, :288-306) over the VCF text of one data line.
// The reference reads the line through htslib (absent from the image); the
// field semantics restated here are those of the VCF 4.x text format as
// htslib exposes them to parse_record: GT alleles separated by '/' or '|',
// '.' a missing allele (bcf_gt_missing), a sample with fewer alleles than
// the record's maximum padded with vector_end; FORMAT/DS (used when the header
// declares it as Float and the record carries it) a float per sample, '.'
// missing; a sample column with fewer ':' subfields than FORMAT keys has the
// trailing ones missing; a record with n_allele != 2 is skipped before
// parsing (multi-allelic, or ALT '.').
static std::vector<std::string> split_on(const char* s, size_t len, char c) {
  std::vector<std::string> out;
  size_t a = 0;
  for (size_t i = 0; i <= len; ++i)
    if (i == len || s[i] == c) {
      out.emplace_back(s + a, i - a);
      a = i + 1;
    }
  return out;
}

static bool haploid_chrom(std::string k) {  // vcf_session.cpp:149-158
  if (k.size() > 3 && (k[0] == 'c' || k[0] == 'C') && (k[1] == 'h' || k[1] == 'H') && (k[2] == 'r' || k[2] == 'R'))
    k = k.substr(3);
  return k == "X" || k == "x" || k == "Y" || k == "y" || k == "M" || k == "m" || k == "MT" || k == "mt" || k == "Mt";
}

// Returns -1 (skipped: n_allele != 2), -2 (sample column count differs from
// n_file: a read error in the reference), 0 (parse_record or pass_maf false)
// or 1 (kept). dosage[n] (analysis order): the SnpRec dosage — NaN for a missing
// call, replaced by the non-missing mean when impute and kept; af_out: out.maf;
// stats[3] (nullable): n_miss, ALT dosage sum, called allele copies; raw[n]
// (nullable): the dosage before imputation.
int32_t orc_vcf_record(const char* line, int64_t len, int32_t ds_float, int32_t n_file, const int32_t* sample_col,
                       int32_t n, int32_t impute, double max_miss, double maf_min, double* dosage, double* af_out,
                       double* stats, double* raw) {
  const auto cols = split_on(line, static_cast<size_t>(len), '\t');
  *af_out = kNaN;
  for (int i = 0; i < n; ++i) dosage[i] = kNaN;
  if (cols.size() < 8) return 0;
  const std::string& alt = cols[4];
  int n_allele = 1;
  if (alt != ".") n_allele = 2 + static_cast<int>(std::count(alt.begin(), alt.end(), ','));
  if (n_allele != 2) return -1;
  if (n <= 0 || n_file <= 0) return 0;
  if (static_cast<int>(cols.size()) != 9 + n_file) return -2;  // htslib: column count differs from the header
  const bool hap = haploid_chrom(cols[0]);
  const auto keys = split_on(cols[8].data(), cols[8].size(), ':');
  int gt_key = -1, ds_key = -1;
  for (size_t k = 0; k < keys.size(); ++k) {
    if (keys[k] == "GT" && gt_key < 0) gt_key = static_cast<int>(k);
    if (keys[k] == "DS" && ds_float && ds_key < 0) ds_key = static_cast<int>(k);
  }
  auto subfield = [&](int col, int key, std::string& out) -> bool {
    const auto sub = split_on(cols[9 + col].data(), cols[9 + col].size(), ':');
    if (key >= static_cast<int>(sub.size())) return false;
    out = sub[key];
    return true;
  };
  int n_miss = 0, n_ok = 0;
  double sum = 0.0, ploidy_sum = 0.0;
  std::string f;
  if (ds_key >= 0) {  // vcf_session.cpp:188-221
    for (int i = 0; i < n; ++i) {
      float v = std::numeric_limits<float>::quiet_NaN();
      if (subfield(sample_col[i], ds_key, f) && !f.empty() &&
          !(f[0] == '.' && !(f.size() > 1 && f[1] >= '0' && f[1] <= '9'))) {
        char* e = nullptr;
        const double d = std::strtod(f.c_str(), &e);
        if (e != f.c_str()) v = static_cast<float>(d);
      }
      if (!std::isfinite(v)) {
        ++n_miss;
        continue;
      }
      dosage[i] = static_cast<double>(v);
      sum += dosage[i];
      ploidy_sum += hap ? 1.0 : 2.0;
      ++n_ok;
    }
  } else {  // vcf_session.cpp:223-260
    if (gt_key < 0) return 0;
    for (int i = 0; i < n; ++i) {
      std::vector<int> al;       // allele indices, -1 = '.'
      if (subfield(sample_col[i], gt_key, f)) {
        size_t q = 0;
        while (q < f.size()) {
          const char c = f[q];
          if (c == '/' || c == '|') { ++q; continue; }
          if (c == '.') { al.push_back(-1); ++q; continue; }
          if (c >= '0' && c <= '9') {
            int v = 0;
            while (q < f.size() && f[q] >= '0' && f[q] <= '9') {
              if (v < 10000000) v = v * 10 + (f[q] - '0');
              ++q;
            }
            al.push_back(v);
            continue;
          }
          ++q;
        }
      }
      const bool miss_gt = al.empty() || al[0] < 0 || (al.size() > 1 && al[1] < 0);
      if (miss_gt) {
        ++n_miss;
        continue;
      }
      int d = al[0], n_alleles = 1;
      if (al.size() > 1) {
        d += al[1];
        ++n_alleles;
      }
      const bool this_hap = hap || n_alleles == 1;
      dosage[i] = static_cast<double>(d);
      sum += d;
      ploidy_sum += this_hap ? 1.0 : 2.0;
      ++n_ok;
    }
  }
  if (raw)
    for (int i = 0; i < n; ++i) raw[i] = dosage[i];
  if (stats) {
    stats[0] = n_miss;
    stats[1] = sum;
    stats[2] = ploidy_sum;
  }
  if (n_ok == 0) return 0;
  const double miss_frac = static_cast<double>(n_miss) / static_cast<double>(n);
  if (miss_frac > max_miss + 1e-15) return 0;
  double af = ploidy_sum > 0 ? sum / ploidy_sum : 0.0;
  if (af < 0) af = 0;
  if (af > 1) af = 1;
  if (af < 1e-12 || af > 1.0 - 1e-12) return 0;
  if (!impute && n_miss > 0) return 0;
  if (n_miss > 0 && impute) {
    const double mu = sum / n_ok;
    for (int i = 0; i < n; ++i)
      if (!std::isfinite(dosage[i])) dosage[i] = mu;
  }
  if (maf_min > 0.0 && (af < maf_min || af > 1.0 - maf_min)) return 0;  // pass_maf, util.hpp:61-64
  *af_out = af;
  return 1;
}

// make_lmm_basis's eigendecomposition (model_lmm.cpp:106-113) with the
// engine's sign convention: column c flipped so that sum_i u_i Q_ic > 0,
// u_i = 1 + frac(0.6180339887498949 (i + 1)) / 2 (LAPACK leaves the signs
// arbitrary; the LMM permutation streams depend on them).
void orc_eigh(const double* K, int32_t n, double* Q, double* lam) {
  jacobi_eigh(K, n, Q, lam);
  for (int c = 0; c < n; ++c) {
    double* q = Q + static_cast<size_t>(c) * n;
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      const double f = 0.6180339887498949 * (i + 1);
      s += q[i] * (1.0 + 0.5 * (f - std::floor(f)));
    }
    if (s < 0)
      for (int i = 0; i < n; ++i) q[i] = -q[i];
  }
}

double orc_reml_negll(double delta, const double* yt, const double* Xt, const double* lam,
                      int32_t n, int32_t p) {
  return reml_negll(delta, yt, Xt, lam, n, p, n - p);
}

// optimize_delta for G genes (Ỹ n x G), OpenMP over genes.
void orc_optimize_delta(const double* Yt, int32_t G, const double* Xt, const double* lam, int32_t n,
                        int32_t p, double* delta) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int k = 0; k < G; ++k)
    delta[k] = optimize_delta(Yt + static_cast<size_t>(k) * n, Xt, lam, n, p);
}

// Q^T A for a n x c column-major A (FP64 reference rotation).
void orc_rotate(const double* Q, int32_t n, const double* A, int64_t c, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < c; ++j)
    for (int r = 0; r < n; ++r) {
      double s = 0;
      const double* q = Q + static_cast<size_t>(r) * n;
      const double* a = A + static_cast<size_t>(j) * n;
      for (int i = 0; i < n; ++i) s += q[i] * a[i];
      out[static_cast<size_t>(j) * n + r] = s;
    }
}

// Per-gene summary written by the trans scans (scan_trans.hpp:53-110).
struct OrcGeneState {
  double delta;
  double acat_p;
  double best_p;
  double best_beta, best_se, best_stat, best_r2, best_af;
  int64_t best_snp;
  int32_t n_tested;
  int32_t n_sig;
};

// Trans LMM scan over decoded SNPs (scan_trans_lmm.cpp:82-251, same_keep path).
//  dosage m x n (imputed), snp_keep[j] decode pass flag, snp_chrom/snp_pos for
//  the cis exclusion; gene_chrom/gene_cstart/gene_cend: cis window
//  [start-w, end+w] (gene_chrom < 0: gene has no location → never excluded).
//  delta_in (nullable): fixed per-gene delta instead of the REML search.
//  Dense outputs (nullable) are m x G, SNP-major, NaN where not tested.
//  Hits (p <= pthr) are written SNP-major, gene order, up to hit_cap.
int64_t orc_trans_lmm_scan(const double* dosage, const uint8_t* snp_keep, const int32_t* snp_chrom,
                           const int64_t* snp_pos, int64_t m, int32_t n, const double* Q,
                           const double* lam, const double* Y, int32_t G, const double* X,
                           int32_t p, const int32_t* gene_chrom, const int64_t* gene_cstart,
                           const int64_t* gene_cend, const double* delta_in, double pthr,
                           OrcGeneState* state, double* d_beta, double* d_se, double* d_stat,
                           double* d_p, int64_t* hit_snp, int32_t* hit_gene, double* hit_vals,
                           int64_t hit_cap, double* loop_seconds) {
  std::vector<double> Yt(static_cast<size_t>(n) * G), Xt(static_cast<size_t>(n) * p);
  orc_rotate(Q, n, Y, G, Yt.data());
  orc_rotate(Q, n, X, p, Xt.data());
  std::vector<PrepLmm> preps(G);
#pragma omp parallel for schedule(dynamic, 4)
  for (int k = 0; k < G; ++k)
    prep_lmm(Yt.data() + static_cast<size_t>(k) * n, Xt.data(), lam, n, p,
             delta_in ? delta_in[k] : -1.0, preps[k]);
  std::vector<Acat> acc(G);
  for (int k = 0; k < G; ++k) {
    state[k] = OrcGeneState{};
    state[k].delta = preps[k].delta;
    state[k].best_p = 2.0;
    state[k].best_snp = -1;
  }
  std::vector<double> gt(n), ps(G);
  std::vector<Hit> hs(G);
  std::vector<char> tested(G);
  int64_t nh = 0;
#ifdef _OPENMP
  const double t_loop0 = omp_get_wtime();
#endif
  for (int64_t j = 0; j < m; ++j) {
    if (!snp_keep[j]) continue;
    const double* g = dosage + j * static_cast<int64_t>(n);
    double maf_sub = 0;
    if (!std::isfinite(subset_maf_or_nan(g, n, &maf_sub))) continue;
    for (int r = 0; r < n; ++r) {
      double s = 0;
      const double* q = Q + static_cast<size_t>(r) * n;
      for (int i = 0; i < n; ++i) s += q[i] * g[i];
      gt[r] = s;
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < G; ++k) {
      tested[k] = 0;
      if (gene_chrom[k] >= 0 && gene_chrom[k] == snp_chrom[j] && snp_pos[j] >= gene_cstart[k] &&
          snp_pos[j] <= gene_cend[k])
        continue;
      hs[k] = test_lmm_gtil(preps[k], gt.data());
      tested[k] = 1;
    }
    for (int k = 0; k < G; ++k) {
      const size_t o = static_cast<size_t>(j) * G + k;
      if (!tested[k]) {
        if (d_p) { d_beta[o] = d_se[o] = d_stat[o] = d_p[o] = kNaN; }
        continue;
      }
      const Hit& h = hs[k];
      if (d_p) { d_beta[o] = h.beta; d_se[o] = h.se; d_stat[o] = h.stat; d_p[o] = h.p; }
      if (!std::isfinite(h.p)) continue;
      OrcGeneState& st = state[k];
      ++st.n_tested;
      acc[k].add(h.p);
      if (h.p <= pthr) {
        ++st.n_sig;
        if (nh < hit_cap) {
          hit_snp[nh] = j;
          hit_gene[nh] = k;
          double* v = hit_vals + nh * 6;
          v[0] = maf_sub; v[1] = h.beta; v[2] = h.se; v[3] = h.stat; v[4] = h.p; v[5] = h.r2;
        }
        ++nh;
      }
      if (h.p < st.best_p) {
        st.best_p = h.p;
        st.best_snp = j;
        st.best_beta = h.beta; st.best_se = h.se; st.best_stat = h.stat; st.best_r2 = h.r2;
        st.best_af = maf_sub;
      }
    }
  }
  for (int k = 0; k < G; ++k) state[k].acat_p = acc[k].p();
#ifdef _OPENMP
  if (loop_seconds) *loop_seconds = omp_get_wtime() - t_loop0;
#endif
  return nh;
}

// Trans LM scan (scan_trans_lm.cpp:100-222, same_keep path).
int64_t orc_trans_lm_scan(const double* dosage, const uint8_t* snp_keep, const int32_t* snp_chrom,
                          const int64_t* snp_pos, int64_t m, int32_t n, const double* Y, int32_t G,
                          const double* X, int32_t p, const int32_t* gene_chrom,
                          const int64_t* gene_cstart, const int64_t* gene_cend, double pthr,
                          OrcGeneState* state, double* d_beta, double* d_se, double* d_stat,
                          double* d_p, int64_t* hit_snp, int32_t* hit_gene, double* hit_vals,
                          int64_t hit_cap) {
  std::vector<PrepLm> preps(G);
#pragma omp parallel for schedule(dynamic, 4)
  for (int k = 0; k < G; ++k) prep_lm(Y + static_cast<size_t>(k) * n, X, n, p, preps[k]);
  std::vector<Acat> acc(G);
  for (int k = 0; k < G; ++k) {
    state[k] = OrcGeneState{};
    state[k].delta = kNaN;
    state[k].best_p = 2.0;
    state[k].best_snp = -1;
  }
  std::vector<double> gs(n);
  std::vector<Hit> hs(G);
  std::vector<char> tested(G);
  int64_t nh = 0;
  for (int64_t j = 0; j < m; ++j) {
    if (!snp_keep[j]) continue;
    const double* g = dosage + j * static_cast<int64_t>(n);
    double maf_sub = 0;
    if (!std::isfinite(subset_maf_or_nan(g, n, &maf_sub))) continue;
    const double gtg = residualize_g(preps[0], g, gs.data());
    const bool degenerate = gtg < 1e-12;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < G; ++k) {
      tested[k] = 0;
      if (gene_chrom[k] >= 0 && gene_chrom[k] == snp_chrom[j] && snp_pos[j] >= gene_cstart[k] &&
          snp_pos[j] <= gene_cend[k])
        continue;
      tested[k] = 1;
      if (degenerate) { hs[k] = Hit{}; hs[k].p = 1.0; continue; }
      double gty = 0;
      const double* ys = preps[k].ys.data();
      for (int i = 0; i < n; ++i) gty += ys[i] * gs[i];
      hs[k] = hit_from_gty(preps[k], gtg, gty);
    }
    for (int k = 0; k < G; ++k) {
      const size_t o = static_cast<size_t>(j) * G + k;
      if (!tested[k]) {
        if (d_p) { d_beta[o] = d_se[o] = d_stat[o] = d_p[o] = kNaN; }
        continue;
      }
      const Hit& h = hs[k];
      if (d_p) { d_beta[o] = h.beta; d_se[o] = h.se; d_stat[o] = h.stat; d_p[o] = h.p; }
      if (!std::isfinite(h.p)) continue;
      OrcGeneState& st = state[k];
      ++st.n_tested;
      acc[k].add(h.p);
      if (h.p <= pthr) {
        ++st.n_sig;
        if (nh < hit_cap) {
          hit_snp[nh] = j;
          hit_gene[nh] = k;
          double* v = hit_vals + nh * 6;
          v[0] = maf_sub; v[1] = h.beta; v[2] = h.se; v[3] = h.stat; v[4] = h.p; v[5] = h.r2;
        }
        ++nh;
      }
      if (h.p < st.best_p) {
        st.best_p = h.p;
        st.best_snp = j;
        st.best_beta = h.beta; st.best_se = h.se; st.best_stat = h.stat; st.best_r2 = h.r2;
        st.best_af = maf_sub;
      }
    }
  }
  for (int k = 0; k < G; ++k) state[k].acat_p = acc[k].p();
  return nh;
}

// ---- NB GLM score test (model_glm.cpp:1-132) ---------------------------------
// Cholesky solve of the p x p SPD system A x = b (A col-major, copied); false
// when A is not positive definite (Eigen's LDLT of the reference reports no
// failure for a PD matrix; both give the same solution to rounding).
static bool spd_solve(std::vector<double> A, int p, const double* b, double* x) {
  for (int j = 0; j < p; ++j) {
    double s = A[j * p + j];
    for (int k = 0; k < j; ++k) s -= A[k * p + j] * A[k * p + j];
    if (!(s > 0)) return false;
    const double l = std::sqrt(s);
    A[j * p + j] = l;
    for (int i = j + 1; i < p; ++i) {
      double t = A[j * p + i];
      for (int k = 0; k < j; ++k) t -= A[k * p + i] * A[k * p + j];
      A[j * p + i] = t / l;
    }
  }
  std::vector<double> z(p);
  for (int i = 0; i < p; ++i) {
    double t = b[i];
    for (int k = 0; k < i; ++k) t -= A[k * p + i] * z[k];
    z[i] = t / A[i * p + i];
  }
  for (int i = p - 1; i >= 0; --i) {
    double t = z[i];
    for (int k = i + 1; k < p; ++k) t -= A[i * p + k] * x[k];
    x[i] = t / A[i * p + i];
  }
  return true;
}

struct PrepGlm {
  int n = 0, p = 0;
  const double* y = nullptr;
  const double* X = nullptr;   // n x p col-major
  std::vector<double> mu, w, XtWX_inv;
  double phi = 1.0;
  bool converged = false;
};

// nb_mom_phi, model_glm.cpp:14-21
static double nb_mom_phi(const double* y, const std::vector<double>& mu, int n, int df) {
  double num = 0;
  for (int i = 0; i < n; ++i) {
    const double m = std::max(mu[i], 1e-8);
    num += ((y[i] - m) * (y[i] - m) - m) / (m * m);
  }
  return std::max(1e-8, num / static_cast<double>(std::max(1, df)));
}

// nb_irls, model_glm.cpp:28-74 (offset 0, phi re-estimated between rounds)
static void nb_irls(const double* y, const double* X, int n, int p, double& phi, std::vector<double>& beta,
                    std::vector<double>& mu, bool& converged) {
  const int df = std::max(1, n - p);
  beta.assign(p, 0.0);
  mu.resize(n);
  for (int i = 0; i < n; ++i) mu[i] = std::max(y[i], 0.1);
  converged = false;
  std::vector<double> A(static_cast<size_t>(p) * p), b(p), bn(p), z(n), w(n);
  for (int outer = 0; outer < 10; ++outer) {
    for (int it = 0; it < 50; ++it) {
      for (int i = 0; i < n; ++i) {
        const double m = std::max(mu[i], 1e-8);
        const double var = m + phi * m * m;
        w[i] = (m * m) / std::max(var, 1e-12);
        z[i] = std::log(m) + (y[i] - m) / m;
      }
      for (int a = 0; a < p; ++a) {
        for (int c = 0; c <= a; ++c) {
          double s = 0;
          for (int i = 0; i < n; ++i) s += X[static_cast<size_t>(a) * n + i] * w[i] * X[static_cast<size_t>(c) * n + i];
          A[static_cast<size_t>(c) * p + a] = A[static_cast<size_t>(a) * p + c] = s;
        }
        double s = 0;
        for (int i = 0; i < n; ++i) s += X[static_cast<size_t>(a) * n + i] * (w[i] * z[i]);
        b[a] = s;
      }
      if (!spd_solve(A, p, b.data(), bn.data())) return;
      double diff = 0;
      for (int a = 0; a < p; ++a) diff = std::max(diff, std::fabs(bn[a] - beta[a]));
      beta = bn;
      for (int i = 0; i < n; ++i) {
        double eta = 0;
        for (int a = 0; a < p; ++a) eta += X[static_cast<size_t>(a) * n + i] * beta[a];
        mu[i] = std::exp(eta);
      }
      if (diff < 1e-6) break;
    }
    const double phi_new = nb_mom_phi(y, mu, n, df);
    if (std::fabs(phi_new - phi) < 1e-6 * std::max(1.0, phi)) {
      phi = phi_new;
      converged = true;
      return;
    }
    phi = phi_new;
  }
  converged = true;
}

// prep_glm_nb, model_glm.cpp:76-98
static void prep_glm(const double* y, const double* X, int n, int p, PrepGlm& P, std::vector<double>* beta_out) {
  P.n = n;
  P.p = p;
  P.y = y;
  P.X = X;
  P.phi = 1.0;
  std::vector<double> beta;
  nb_irls(y, X, n, p, P.phi, beta, P.mu, P.converged);
  if (beta_out) *beta_out = beta;
  P.w.resize(n);
  for (int i = 0; i < n; ++i) {
    const double m = std::max(P.mu[i], 1e-8);
    P.w[i] = m / (1.0 + P.phi * m);
  }
  std::vector<double> A(static_cast<size_t>(p) * p);
  for (int a = 0; a < p; ++a)
    for (int c = 0; c <= a; ++c) {
      double s = 0;
      for (int i = 0; i < n; ++i) s += X[static_cast<size_t>(a) * n + i] * P.w[i] * X[static_cast<size_t>(c) * n + i];
      A[static_cast<size_t>(c) * p + a] = A[static_cast<size_t>(a) * p + c] = s;
    }
  P.XtWX_inv.assign(static_cast<size_t>(p) * p, kNaN);
  std::vector<double> e(p), col(p);
  for (int c = 0; c < p; ++c) {
    std::fill(e.begin(), e.end(), 0.0);
    e[c] = 1.0;
    if (!spd_solve(A, p, e.data(), col.data())) {
      P.converged = false;
      return;
    }
    for (int a = 0; a < p; ++a) P.XtWX_inv[static_cast<size_t>(c) * p + a] = col[a];
  }
}

// test_glm_nb, model_glm.cpp:100-132 (g: the imputed dosage, uncentred)
static Hit test_glm(const PrepGlm& P, const double* g) {
  Hit h;
  h.r2 = 0;
  if (!P.converged) {
    h.p = kNaN;
    return h;
  }
  const int n = P.n, p = P.p;
  double s1 = 0, s2 = 0;
  for (int i = 0; i < n; ++i) {
    s1 += g[i];
    s2 += g[i] * g[i];
  }
  const double gvar = s2 / n - (s1 / n) * (s1 / n);
  if (gvar < 1e-12) {
    h.p = 1.0;
    return h;
  }
  double U = 0, gwg = 0;
  std::vector<double> xwg(p, 0.0);
  for (int i = 0; i < n; ++i) {
    const double r = (P.y[i] - P.mu[i]) / (1.0 + P.phi * P.mu[i]);
    U += g[i] * r;
    const double wg = P.w[i] * g[i];
    gwg += g[i] * wg;
    for (int a = 0; a < p; ++a) xwg[a] += P.X[static_cast<size_t>(a) * n + i] * wg;
  }
  double quad = 0;
  for (int a = 0; a < p; ++a) {
    double t = 0;
    for (int c = 0; c < p; ++c) t += P.XtWX_inv[static_cast<size_t>(c) * p + a] * xwg[c];
    quad += xwg[a] * t;
  }
  const double info = gwg - quad;
  if (info < 1e-12) {
    h.p = 1.0;
    return h;
  }
  h.beta = U / info;
  h.se = 1.0 / std::sqrt(info);
  h.stat = U / std::sqrt(info);
  double pv = std::erfc(std::fabs(h.stat) / std::sqrt(2.0));  // pnorm_two_sided, util.cpp:116-123
  if (pv < 0) pv = 0;
  if (pv > 1) pv = 1;
  h.p = std::isfinite(h.stat) ? pv : 1.0;
  return h;
}

// Null fit of one gene: returns converged; phi, mu (n) and beta (p) out.
int32_t orc_glm_prep(const double* y, const double* X, int32_t n, int32_t p, double* phi, double* mu,
                     double* beta, double* w) {
  PrepGlm P;
  std::vector<double> b;
  prep_glm(y, X, n, p, P, &b);
  *phi = P.phi;
  for (int i = 0; i < n; ++i) {
    mu[i] = P.mu[i];
    if (w) w[i] = P.w[i];
  }
  for (int a = 0; a < p && a < static_cast<int>(b.size()); ++a) beta[a] = b[a];
  return P.converged ? 1 : 0;
}

// Score test of one SNP (imputed dosage g) for one gene: out = beta, se, stat, p.
void orc_glm_test(const double* y, const double* X, int32_t n, int32_t p, const double* g, double* out) {
  PrepGlm P;
  prep_glm(y, X, n, p, P, nullptr);
  const Hit h = test_glm(P, g);
  out[0] = h.beta;
  out[1] = h.se;
  out[2] = h.stat;
  out[3] = h.p;
}

// NB GLM scan over decoded SNPs for every gene (the gene-outer scan of
// scan.cpp:228-300 with scan_gene_snps, scan_cis.hpp:15-66, without
// permutations): a SNP is tested for gene k unless it lies in the gene's
// window (trans) — or, when cis != 0, only if it lies there. Same outputs as
// orc_trans_lm_scan plus phi / converged per gene.
int64_t orc_glm_scan(const double* dosage, const uint8_t* snp_keep, const int32_t* snp_chrom, const int64_t* snp_pos,
                     int64_t m, int32_t n, const double* Y, int32_t G, const double* X, int32_t p,
                     const int32_t* gene_chrom, const int64_t* gene_cstart, const int64_t* gene_cend, int32_t cis,
                     double pthr, OrcGeneState* state, double* phi_out, int32_t* conv_out, double* d_beta,
                     double* d_se, double* d_stat, double* d_p, int64_t* hit_snp, int32_t* hit_gene,
                     double* hit_vals, int64_t hit_cap) {
  std::vector<PrepGlm> preps(G);
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < G; ++k) prep_glm(Y + static_cast<size_t>(k) * n, X, n, p, preps[k], nullptr);
  std::vector<Acat> acc(G);
  for (int k = 0; k < G; ++k) {
    state[k] = OrcGeneState{};
    state[k].delta = kNaN;
    state[k].best_p = 2.0;
    state[k].best_snp = -1;
    phi_out[k] = preps[k].phi;
    conv_out[k] = preps[k].converged ? 1 : 0;
  }
  std::vector<Hit> hs(G);
  std::vector<char> tested(G);
  int64_t nh = 0;
  for (int64_t j = 0; j < m; ++j) {
    const double* g = dosage + j * static_cast<int64_t>(n);
    const bool snp_ok = snp_keep[j] != 0;
    double maf_sub = 0;
    const bool maf_ok = snp_ok && std::isfinite(subset_maf_or_nan(g, n, &maf_sub));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < G; ++k) {
      tested[k] = 0;
      if (!maf_ok) continue;
      const bool inwin = gene_chrom[k] >= 0 && gene_chrom[k] == snp_chrom[j] && snp_pos[j] >= gene_cstart[k] &&
                         snp_pos[j] <= gene_cend[k];
      if (cis ? !inwin : inwin) continue;
      tested[k] = 1;
      hs[k] = test_glm(preps[k], g);
    }
    for (int k = 0; k < G; ++k) {
      const size_t o = static_cast<size_t>(j) * G + k;
      if (!tested[k]) {
        if (d_p) { d_beta[o] = d_se[o] = d_stat[o] = d_p[o] = kNaN; }
        continue;
      }
      const Hit& h = hs[k];
      if (d_p) { d_beta[o] = h.beta; d_se[o] = h.se; d_stat[o] = h.stat; d_p[o] = h.p; }
      if (!std::isfinite(h.p)) continue;
      OrcGeneState& st = state[k];
      ++st.n_tested;
      acc[k].add(h.p);
      if (h.p <= pthr) {
        ++st.n_sig;
        if (nh < hit_cap) {
          hit_snp[nh] = j;
          hit_gene[nh] = k;
          double* v = hit_vals + nh * 6;
          v[0] = maf_sub; v[1] = h.beta; v[2] = h.se; v[3] = h.stat; v[4] = h.p; v[5] = h.r2;
        }
        ++nh;
      }
      if (h.p < st.best_p) {
        st.best_p = h.p;
        st.best_snp = j;
        st.best_beta = h.beta; st.best_se = h.se; st.best_stat = h.stat; st.best_r2 = h.r2;
        st.best_af = maf_sub;
      }
    }
  }
  for (int k = 0; k < G; ++k) state[k].acat_p = acc[k].p();
  return nh;
}

// GRM: grm.cpp:112-224 (standardise, A += Z Z^T, GCTA per-pair divisor).
// dosage m x n imputed; miss m x n mask; snp_keep as decoded. K n x n
// column-major, N_packed n(n+1)/2 row-major lower incl. diagonal.
// Returns the number of SNPs used.
int64_t orc_grm(const double* dosage, const int8_t* miss, const uint8_t* snp_keep, int64_t m,
                int32_t n, double* K, int32_t* N_packed) {
  std::vector<double> A(static_cast<size_t>(n) * n, 0.0);
  std::vector<int> sub_miss(n, 0);
  std::vector<int> joint(static_cast<size_t>(n) * n, 0);
  std::vector<double> Z(n);
  std::vector<int> midx;
  int64_t used = 0;
  for (int64_t j = 0; j < m; ++j) {
    if (!snp_keep[j]) continue;
    const double* g = dosage + j * static_cast<int64_t>(n);
    double mu = 0;
    for (int i = 0; i < n; ++i) mu += g[i];
    mu /= n;
    const double pf = mu / 2.0;
    if (pf < 1e-8 || pf > 1.0 - 1e-8) continue;
    const double sd = std::sqrt(2.0 * pf * (1.0 - pf));
    if (sd < 1e-12) continue;
    for (int i = 0; i < n; ++i) Z[i] = (g[i] - mu) / sd;
    ++used;
    midx.clear();
    const int8_t* mm = miss + j * static_cast<int64_t>(n);
    for (int i = 0; i < n; ++i)
      if (mm[i]) {
        ++sub_miss[i];
        for (int q : midx) joint[static_cast<size_t>(q) * n + i] += 1;  // (i, q) lower
        midx.push_back(i);
      }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; ++c) {
      const double zc = Z[c];
      double* col = A.data() + static_cast<size_t>(c) * n;
      for (int r = c; r < n; ++r) col[r] += Z[r] * zc;
    }
  }
  if (used == 0) return 0;
  size_t k = 0;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < i; ++j) {
      const int sub_N = static_cast<int>(used) - sub_miss[i] - sub_miss[j] +
                        joint[static_cast<size_t>(j) * n + i];
      const double v = sub_N > 0 ? A[static_cast<size_t>(j) * n + i] / sub_N : 0.0;
      K[static_cast<size_t>(j) * n + i] = v;
      K[static_cast<size_t>(i) * n + j] = v;
      N_packed[k++] = sub_N;
    }
    const int sd = static_cast<int>(used) - sub_miss[i];
    K[static_cast<size_t>(i) * n + i] = sd > 0 ? A[static_cast<size_t>(i) * n + i] / sd : 0.0;
    N_packed[k++] = sd;
  }
  return used;
}

uint32_t orc_fnv1a(const char* s) { return fnv1a(s); }

void orc_perm_index(int32_t seed, int32_t b, const char* gene, int32_t n, int32_t* idx) {
  perm_index(seed, b, gene, n, idx);
}

// LMM permutation base (scan_cis.hpp:77-87): b0 = (X̃ᵀD⁻¹X̃)⁻¹ X̃ᵀD⁻¹ỹ by
// LDLT, whitened spectral residuals w = (ỹ − X̃b0)∘√dinv. When the LDLT
// reports failure, base keeps its input (the raw phenotype) and Xb0 stays empty.
static void lmm_perm_base(const PrepLmm& lmm, std::vector<double>& base, std::vector<double>& Xb0) {
  const int n = lmm.n, p = lmm.p;
  std::vector<double> A(static_cast<size_t>(p) * p), bb(p), b0(p);
  for (int a = 0; a < p; ++a) {
    for (int cc = 0; cc <= a; ++cc) {
      double s = 0;
      for (int i = 0; i < n; ++i)
        s += lmm.Xt[static_cast<size_t>(a) * n + i] * lmm.dinv[i] * lmm.Xt[static_cast<size_t>(cc) * n + i];
      A[static_cast<size_t>(cc) * p + a] = s;
      A[static_cast<size_t>(a) * p + cc] = s;
    }
    double s = 0;
    for (int i = 0; i < n; ++i) s += lmm.Xt[static_cast<size_t>(a) * n + i] * (lmm.dinv[i] * lmm.yt[i]);
    bb[a] = s;
  }
  Ldlt f;
  f.compute(A.data(), p);
  if (!f.ok) return;
  f.solve(bb.data(), b0.data());
  Xb0.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (int a = 0; a < p; ++a) s += lmm.Xt[static_cast<size_t>(a) * n + i] * b0[a];
    Xb0[i] = s;
    base[i] = (lmm.yt[i] - s) * std::sqrt(lmm.dinv[i]);
  }
}

// Pseudo-phenotype of one LMM permutation (scan_cis.hpp:135-148): un-whiten
// the permuted residuals, r̃ = w_perm / √max(dinv, 1e-12), and rotate back,
// y = Q (X̃b0 + r̃). Q is column-major n x n (column r = eigenvector r).
static void lmm_perm_phenotype(const double* Q, const PrepLmm& lmm, const std::vector<double>& Xb0,
                               const double* wperm, double* rt, double* yy) {
  const int n = lmm.n;
  for (int i = 0; i < n; ++i) rt[i] = Xb0[i] + wperm[i] / std::sqrt(std::max(lmm.dinv[i], 1e-12));
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (int r = 0; r < n; ++r) s += Q[static_cast<size_t>(r) * n + i] * rt[r];
    yy[i] = s;
  }
}

// Cis gene scan with gene-level permutations (scan_cis.hpp:15-185) for LM
// (model 0) or LMM (model 1) over one gene's window.
//  y (n), X (n x p), window dosages W (c x n, only SNPs that passed decode),
//  lmm: Q, lam. Outputs per window SNP (nullable, c each): beta/se/stat/p.
//  obs: [n_tested, best_p]; perm_min_p (B) and T_perm (B).
// Returns p_emp (NaN when B == 0 or nothing tested).
double orc_cis_gene(int32_t model, const double* y, const double* X, int32_t n, int32_t p,
                    const double* W, int32_t c, const double* Q, const double* lam, int32_t B,
                    int32_t seed, const char* gene, double* w_beta, double* w_se, double* w_stat,
                    double* w_p, double* obs, double* perm_min_p, double* T_perm,
                    double* delta_out) {
  PrepLm lm;
  PrepLmm lmm;
  std::vector<double> Yt, Xt;
  if (model == 0) {
    prep_lm(y, X, n, p, lm);
  } else {
    Yt.resize(n);
    Xt.resize(static_cast<size_t>(n) * p);
    orc_rotate(Q, n, y, 1, Yt.data());
    orc_rotate(Q, n, X, p, Xt.data());
    prep_lmm(Yt.data(), Xt.data(), lam, n, p, -1.0, lmm);
    if (delta_out) *delta_out = lmm.delta;
  }
  std::vector<int> cached;
  int n_tested = 0;
  double best = 2.0;
  // LMM: Q^T g of each window SNP (test_lmm forms it per test, model_lmm.cpp:170;
  // the same deterministic FP64 sum is memoised here, so every permutation
  // reuses bit-identical rotated genotypes instead of recomputing them)
  // LM: g_s and g_s^T g_s of each window SNP (residualize_g depends only on g
  // and X, not on the permuted phenotype; memoised the same way)
  std::vector<double> rot(static_cast<size_t>(c) * n), gtg_of(c);
  std::vector<char> have_rot(c, 0);
  auto test_one = [&](int s, const PrepLm* L, const PrepLmm* M) -> Hit {
    const double* g = W + static_cast<size_t>(s) * n;
    if (L) {
      double* gsr = rot.data() + static_cast<size_t>(s) * n;
      if (!have_rot[s]) {
        gtg_of[s] = residualize_g(*L, g, gsr);
        have_rot[s] = 1;
      }
      const double gtg = gtg_of[s];
      double gty = 0;
      for (int i = 0; i < n; ++i) gty += gsr[i] * L->ys[i];
      if (gtg < 1e-12) { Hit h; h.p = 1.0; return h; }
      return hit_from_gty(*L, gtg, gty);
    }
    double* gr = rot.data() + static_cast<size_t>(s) * n;
    if (!have_rot[s]) {
      for (int r = 0; r < n; ++r) {
        double acc = 0;
        const double* q = Q + static_cast<size_t>(r) * n;
        for (int i = 0; i < n; ++i) acc += q[i] * g[i];
        gr[r] = acc;
      }
      have_rot[s] = 1;
    }
    return test_lmm_gtil(*M, gr);
  };
  for (int s = 0; s < c; ++s) {
    const double* g = W + static_cast<size_t>(s) * n;
    double maf;
    Hit h;
    h.p = kNaN;
    if (std::isfinite(subset_maf_or_nan(g, n, &maf)))
      h = test_one(s, model == 0 ? &lm : nullptr, model == 0 ? nullptr : &lmm);
    if (w_p) { w_beta[s] = h.beta; w_se[s] = h.se; w_stat[s] = h.stat; w_p[s] = h.p; }
    if (!std::isfinite(h.p)) continue;
    ++n_tested;
    if (h.p < best) best = h.p;
    cached.push_back(s);
  }
  obs[0] = n_tested;
  obs[1] = best;
  if (B <= 0 || n_tested == 0) return kNaN;
  // permutation base (scan_cis.hpp:74-87)
  std::vector<double> base(y, y + n), Xb0;
  if (model == 0)
    base = lm.ys;
  else
    lmm_perm_base(lmm, base, Xb0);
  std::vector<int> idx(n);
  std::vector<double> yb(n), rt(n), yy(n);
  for (int b = 0; b < B; ++b) {
    perm_index(seed, b, gene, n, idx.data());
    for (int i = 0; i < n; ++i) yb[i] = base[idx[i]];
    PrepLm lmb;
    PrepLmm lmmb;
    if (model == 0) {
      prep_lm(yb.data(), X, n, p, lmb);
    } else {
      if (!Xb0.empty()) {
        lmm_perm_phenotype(Q, lmm, Xb0, yb.data(), rt.data(), yy.data());
      } else {
        yy = yb;
      }
      std::vector<double> yt2(n);
      orc_rotate(Q, n, yy.data(), 1, yt2.data());
      prep_lmm(yt2.data(), lmm.Xt.data(), lam, n, p, -1.0, lmmb);
    }
    double minp = 1.0;
    for (int s : cached) {
      const Hit h = test_one(s, model == 0 ? &lmb : nullptr, model == 0 ? nullptr : &lmmb);
      if (std::isfinite(h.p) && h.p < minp) minp = h.p;
    }
    perm_min_p[b] = minp;
    T_perm[b] = -std::log10(std::max(minp, 1e-300));
  }
  const double Tobs = -std::log10(std::max(best, 1e-300));
  return orc_p_emp(Tobs, T_perm, B);
}

// LMM permutation base of one phenotype and the pseudo-phenotype of one
// permutation idx (scan_cis.hpp:77-87,135-148), for tests. Q column-major
// n x n, X column-major n x p, delta > 0 given (<= 0: searched). Writes
// base_out (whitened residuals) and y_out (pseudo-phenotype); returns 0 and
// writes nothing when the LDLT of X̃ᵀD⁻¹X̃ reports failure.
int32_t orc_lmm_perm_pheno(const double* Q, const double* lam, const double* y, const double* X,
                           int32_t n, int32_t p, double delta, const int32_t* idx, double* base_out,
                           double* y_out) {
  std::vector<double> yt(n), Xt(static_cast<size_t>(n) * p);
  orc_rotate(Q, n, y, 1, yt.data());
  orc_rotate(Q, n, X, p, Xt.data());
  PrepLmm lmm;
  prep_lmm(yt.data(), Xt.data(), lam, n, p, delta, lmm);
  std::vector<double> base(y, y + n), Xb0, wperm(n), rt(n);
  lmm_perm_base(lmm, base, Xb0);
  if (Xb0.empty()) return 0;
  for (int i = 0; i < n; ++i) wperm[i] = base[idx[i]];
  lmm_perm_phenotype(Q, lmm, Xb0, wperm.data(), rt.data(), y_out);
  std::copy(base.begin(), base.end(), base_out);
  return 1;
}

// Direct LMM Wald at a given (prep) for one rotated genotype — single-test
// entry used by small parity tests.
void orc_lmm_test(const double* yt, const double* Xt, const double* lam, int32_t n, int32_t p,
                  double delta, const double* gt, double* out5, double* delta_out) {
  PrepLmm P;
  prep_lmm(yt, Xt, lam, n, p, delta, P);
  const Hit h = test_lmm_gtil(P, gt);
  out5[0] = h.beta; out5[1] = h.se; out5[2] = h.stat; out5[3] = h.p; out5[4] = h.r2;
  if (delta_out) *delta_out = P.delta;
}

int32_t orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
