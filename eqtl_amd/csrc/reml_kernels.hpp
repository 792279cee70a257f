// Gene-side FP64 kernels (one workgroup of 256 threads per gene):
//   K5 k_reml        — per-gene REML delta: the reference's grid + golden +
//                      numeric-Newton search restated (model_lmm.cpp:39-86), or
//                      a safeguarded Newton on log(delta) driven by the analytic
//                      gradient/Hessian (adjoint of reml_negll, model_lmm.cpp:10-37)
//   k_reml_grad_hess — f, df/dlog(delta), d2f/dlog(delta)^2 for all genes
//   k_lmm_basis      — whitened orthonormal covariate basis W_k, null residual
//                      r_k and the column blocks the fused Wald kernel reads
//                      (prep_lmm, model_lmm.cpp:124-153, in orthonormal form)
//   k_lm_basis / k_lm_gene — prep_lm (model_lm.cpp:19-30) in orthonormal form
//
// reml_negll(delta) = 1/2 [df ln(q/df) + sum ln v_i + ln det(X~^T D X~)],
// v_i = max(delta*lambda_i + 1, 1e-12), D = diag(1/v), q = y~^T D y~ - b^T A^-1 b.
// With theta = ln(delta), d_i = 1/v_i, e_i = 1 - d_i:
//   dD/dtheta = -d e,  d2D/dtheta2 = d e (e - d),  d(sum ln v)/dtheta = sum e,
//   d2(sum ln v)/dtheta2 = sum d e,  q' = r^T D' r,  q'' = r^T D'' r - 2 u^T A^-1 u
// with r = y~ - X~ beta_hat and u = X~^T D' r (envelope theorem), and
//   (ln det A)' = tr(A^-1 A'),  (ln det A)'' = tr(A^-1 A'') - tr(A^-1 A' A^-1 A').
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mfma.hpp"

namespace eqtl_amd {

constexpr int kPMax = 32;   // max covariates (incl. intercept)
constexpr int kChunk = 128; // sample rows staged in LDS per pass
constexpr int kMaxTasks = 8;

struct RemlSmem {
  double Xs[kChunk][kPMax];
  double ys[kChunk];
  double w0[kChunk], w1[kChunk], w2[kChunk];
  double rs[kChunk];
  double A[kPMax * kPMax];   // A (later: LDLT factors)
  double A1[kPMax * kPMax];
  double A2[kPMax * kPMax];
  double M[kPMax * kPMax];   // A^-1 A1 (columns)
  double b[kPMax], beta[kPMax], u[kPMax], wv[kPMax], D[kPMax];
  int tr[kPMax];
  double red[256];
  double scal[16];
  int flag;
};

__device__ inline double block_sum(double v, RemlSmem& sm) {
  sm.red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm.red[threadIdx.x] += sm.red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sm.red[0];
  __syncthreads();
  return r;
}

// LDLT with diagonal pivoting on the not-yet-updated trailing diagonal — the
// pivot rule of Eigen 3.4's LDLT (model_lmm.cpp:23/141). In place on L
// (p x p col-major, full symmetric input). Single thread. Returns false if a
// pivot is not finite.
__device__ inline bool ldlt_factor(double* L, double* D, int* tr, int p) {
  double temp[kPMax];
  auto m = [&](int i, int j) -> double& { return L[j * kPMax + i]; };
  bool ok = true;
  for (int k = 0; k < p; ++k) {
    int big = k;
    double bv = fabs(m(k, k));
    for (int i = k + 1; i < p; ++i)
      if (fabs(m(i, i)) > bv) { bv = fabs(m(i, i)); big = i; }
    tr[k] = big;
    if (big != k) {
      for (int j = 0; j < k; ++j) { const double t = m(k, j); m(k, j) = m(big, j); m(big, j) = t; }
      for (int i = big + 1; i < p; ++i) { const double t = m(i, k); m(i, k) = m(i, big); m(i, big) = t; }
      { const double t = m(k, k); m(k, k) = m(big, big); m(big, big) = t; }
      for (int i = k + 1; i < big; ++i) { const double t = m(i, k); m(i, k) = m(big, i); m(big, i) = t; }
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
    if (fabs(akk) > 0.0)
      for (int i = k + 1; i < p; ++i) m(i, k) /= akk;
    if (!isfinite(akk)) ok = false;
  }
  return ok;
}

// x := A^{-1} b using the factors (any thread; b and x may alias).
__device__ inline void ldlt_solve(const double* L, const double* D, const int* tr, int p, const double* b,
                                  double* x) {
  double y[kPMax];
  for (int k = 0; k < p; ++k) y[k] = b[k];
  for (int k = 0; k < p; ++k) { const double t = y[k]; y[k] = y[tr[k]]; y[tr[k]] = t; }
  for (int i = 0; i < p; ++i)
    for (int j = 0; j < i; ++j) y[i] -= L[j * kPMax + i] * y[j];
  for (int i = 0; i < p; ++i) y[i] = (D[i] != 0.0) ? y[i] / D[i] : 0.0;
  for (int i = p - 1; i >= 0; --i)
    for (int j = i + 1; j < p; ++j) y[i] -= L[i * kPMax + j] * y[j];
  for (int k = p - 1; k >= 0; --k) { const double t = y[k]; y[k] = y[tr[k]]; y[tr[k]] = t; }
  for (int k = 0; k < p; ++k) x[k] = y[k];
}

// Pass over all samples accumulating A (and A1, A2 when DERIV), b and the
// per-row scalars. Collective over the 256 threads.
template <bool DERIV>
__device__ void reml_pass1(const double* __restrict__ Xt, const double* __restrict__ yt,
                           const double* __restrict__ lam, int n, int p, double delta, RemlSmem& sm) {
  const int TA = p * (p + 1) / 2;
  const int T = (DERIV ? 3 * TA : TA) + p;
  double acc[kMaxTasks];
#pragma unroll
  for (int t = 0; t < kMaxTasks; ++t) acc[t] = 0;
  double s_logv = 0, s_ydy = 0, s_e = 0, s_de = 0;
  for (int i0 = 0; i0 < n; i0 += kChunk) {
    const int rows = min(kChunk, n - i0);
    for (int q = threadIdx.x; q < kChunk * p; q += 256) {
      const int a = q / kChunk, r = q % kChunk;
      sm.Xs[r][a] = r < rows ? Xt[static_cast<size_t>(a) * n + i0 + r] : 0.0;
    }
    if (threadIdx.x < kChunk) {
      const int r = threadIdx.x;
      double w0 = 0, w1 = 0, w2 = 0, y = 0;
      if (r < rows) {
        const int i = i0 + r;
        double v = delta * lam[i] + 1.0;
        if (v < 1e-12) v = 1e-12;
        const double d = 1.0 / v;
        y = yt[i];
        w0 = d;
        s_logv += log(v);
        s_ydy += y * (d * y);
        if (DERIV) {
          const double e = 1.0 - d;
          w1 = -d * e;
          w2 = d * e * (e - d);
          s_e += e;
          s_de += d * e;
        }
      }
      sm.w0[r] = w0; sm.w1[r] = w1; sm.w2[r] = w2; sm.ys[r] = y;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kMaxTasks; ++s) {
      const int t = threadIdx.x + 256 * s;
      if (t >= T) break;
      double v = 0;
      if (t < (DERIV ? 3 * TA : TA)) {
        const int which = t / TA, tt = t % TA;
        // triangular index → (a, c), c <= a
        int a = static_cast<int>((sqrt(8.0 * tt + 1.0) - 1.0) * 0.5);
        while ((a + 1) * (a + 2) / 2 <= tt) ++a;
        while (a * (a + 1) / 2 > tt) --a;
        const int c = tt - a * (a + 1) / 2;
        const double* w = which == 0 ? sm.w0 : (which == 1 ? sm.w1 : sm.w2);
        for (int r = 0; r < rows; ++r) v += sm.Xs[r][a] * w[r] * sm.Xs[r][c];
      } else {
        const int a = t - (DERIV ? 3 * TA : TA);
        for (int r = 0; r < rows; ++r) v += sm.Xs[r][a] * (sm.w0[r] * sm.ys[r]);
      }
      acc[s] += v;
    }
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < kMaxTasks; ++s) {
    const int t = threadIdx.x + 256 * s;
    if (t >= T) break;
    if (t < (DERIV ? 3 * TA : TA)) {
      const int which = t / TA, tt = t % TA;
      int a = static_cast<int>((sqrt(8.0 * tt + 1.0) - 1.0) * 0.5);
      while ((a + 1) * (a + 2) / 2 <= tt) ++a;
      while (a * (a + 1) / 2 > tt) --a;
      const int c = tt - a * (a + 1) / 2;
      double* M = which == 0 ? sm.A : (which == 1 ? sm.A1 : sm.A2);
      M[c * kPMax + a] = acc[s];
      M[a * kPMax + c] = acc[s];
    } else {
      sm.b[t - (DERIV ? 3 * TA : TA)] = acc[s];
    }
  }
  const double logv = block_sum(s_logv, sm);
  const double ydy = block_sum(s_ydy, sm);
  if (threadIdx.x == 0) { sm.scal[0] = logv; sm.scal[1] = ydy; }
  if (DERIV) {
    const double se = block_sum(s_e, sm);
    const double sde = block_sum(s_de, sm);
    if (threadIdx.x == 0) { sm.scal[2] = se; sm.scal[3] = sde; }
  }
  __syncthreads();
}

// f(delta) (reml_negll). Collective; returns the same value on every thread.
__device__ double reml_f(const double* Xt, const double* yt, const double* lam, int n, int p, double delta,
                         RemlSmem& sm) {
  reml_pass1<false>(Xt, yt, lam, n, p, delta, sm);
  if (threadIdx.x == 0) {
    const int df = n - p;
    double f;
    if (!ldlt_factor(sm.A, sm.D, sm.tr, p)) {
      f = 1e300;
    } else {
      ldlt_solve(sm.A, sm.D, sm.tr, p, sm.b, sm.beta);
      double q = sm.scal[1];
      for (int a = 0; a < p; ++a) q -= sm.b[a] * sm.beta[a];
      if (q <= 0) q = 1e-12;
      const double sigma2 = q / df;
      double logdet_x = 0;
      bool bad = false;
      for (int a = 0; a < p; ++a) {
        if (sm.D[a] <= 0) { bad = true; break; }
        logdet_x += log(sm.D[a]);
      }
      f = bad ? 1e300 : 0.5 * (df * log(sigma2) + sm.scal[0] + logdet_x);
    }
    sm.scal[8] = f;
  }
  __syncthreads();
  const double f = sm.scal[8];
  __syncthreads();
  return f;
}

// f, df/dtheta, d2f/dtheta2 at delta = exp(theta). Collective.
__device__ void reml_fgh(const double* Xt, const double* yt, const double* lam, int n, int p, double delta,
                         RemlSmem& sm, double* f, double* g, double* h) {
  reml_pass1<true>(Xt, yt, lam, n, p, delta, sm);
  if (threadIdx.x == 0) {
    sm.flag = ldlt_factor(sm.A, sm.D, sm.tr, p) ? 1 : 0;
    for (int a = 0; a < p && sm.flag; ++a)
      if (!(sm.D[a] > 0)) sm.flag = 0;
    if (sm.flag) ldlt_solve(sm.A, sm.D, sm.tr, p, sm.b, sm.beta);
  }
  __syncthreads();
  const bool ok = sm.flag != 0;
  // columns of M = A^-1 A1 and of A^-1 A2 (diagonal only needed)
  if (ok && threadIdx.x < p) {
    const int c = threadIdx.x;
    double col[kPMax], x[kPMax];
    for (int a = 0; a < p; ++a) col[a] = sm.A1[c * kPMax + a];
    ldlt_solve(sm.A, sm.D, sm.tr, p, col, x);
    for (int a = 0; a < p; ++a) sm.M[c * kPMax + a] = x[a];
    for (int a = 0; a < p; ++a) col[a] = sm.A2[c * kPMax + a];
    ldlt_solve(sm.A, sm.D, sm.tr, p, col, x);
    sm.wv[c] = x[c];  // diag of A^-1 A2
  }
  // pass 2: residual r = y~ - X~ beta, q' and q'' pieces, u = X~^T D' r
  double s_q1 = 0, s_q2 = 0;
  double uacc = 0;
  for (int i0 = 0; i0 < n && ok; i0 += kChunk) {
    const int rows = min(kChunk, n - i0);
    for (int q = threadIdx.x; q < kChunk * p; q += 256) {
      const int a = q / kChunk, r = q % kChunk;
      sm.Xs[r][a] = r < rows ? Xt[static_cast<size_t>(a) * n + i0 + r] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < kChunk) {
      const int r = threadIdx.x;
      double w1 = 0, rr = 0;
      if (r < rows) {
        const int i = i0 + r;
        double v = delta * lam[i] + 1.0;
        if (v < 1e-12) v = 1e-12;
        const double d = 1.0 / v, e = 1.0 - d;
        w1 = -d * e;
        const double w2 = d * e * (e - d);
        rr = yt[i];
        for (int a = 0; a < p; ++a) rr -= sm.Xs[r][a] * sm.beta[a];
        s_q1 += w1 * rr * rr;
        s_q2 += w2 * rr * rr;
      }
      sm.rs[r] = rr;
      sm.w1[r] = w1;
    }
    __syncthreads();
    if (threadIdx.x < p) {
      const int a = threadIdx.x;
      for (int r = 0; r < rows; ++r) uacc += sm.Xs[r][a] * sm.w1[r] * sm.rs[r];
    }
    __syncthreads();
  }
  const double q1 = block_sum(s_q1, sm);
  const double q2 = block_sum(s_q2, sm);
  if (threadIdx.x < p) sm.u[threadIdx.x] = uacc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int df = n - p;
    double F = 1e300, Gd = 0, H = 0;
    if (ok) {
      double q = sm.scal[1];
      for (int a = 0; a < p; ++a) q -= sm.b[a] * sm.beta[a];
      if (q <= 0) q = 1e-12;
      double logdet_x = 0, tr1 = 0, tr2 = 0, tr11 = 0;
      for (int a = 0; a < p; ++a) {
        logdet_x += log(sm.D[a]);
        tr1 += sm.M[a * kPMax + a];
        tr2 += sm.wv[a];
        for (int c = 0; c < p; ++c) tr11 += sm.M[c * kPMax + a] * sm.M[a * kPMax + c];
      }
      double w[kPMax];
      ldlt_solve(sm.A, sm.D, sm.tr, p, sm.u, w);
      double uw = 0;
      for (int a = 0; a < p; ++a) uw += sm.u[a] * w[a];
      const double q2f = q2 - 2.0 * uw;
      F = 0.5 * (df * log(q / df) + sm.scal[0] + logdet_x);
      Gd = 0.5 * (df * q1 / q + sm.scal[2] + tr1);
      H = 0.5 * (df * (q2f / q - (q1 / q) * (q1 / q)) + sm.scal[3] + tr2 - tr11);
    }
    sm.scal[8] = F; sm.scal[9] = Gd; sm.scal[10] = H;
  }
  __syncthreads();
  *f = sm.scal[8]; *g = sm.scal[9]; *h = sm.scal[10];
  __syncthreads();
}

// K5: per-gene delta. mode 0 = reference search, 1 = grid + Newton.
__global__ __launch_bounds__(256) void k_reml(const double* Xt, const double* Yt, const double* lam, int n, int p,
                                              int G, int mode, double* delta_out) {
  __shared__ RemlSmem sm;
  const int k = blockIdx.x;
  if (k >= G) return;
  const double* yt = Yt + static_cast<size_t>(k) * n;
  const int df = n - p;
  if (df <= 0) {
    if (threadIdx.x == 0) delta_out[k] = 1.0;
    return;
  }
  double best_d = 1.0;
  double best_ll = reml_f(Xt, yt, lam, n, p, 1.0, sm);
  for (double d = 1e-5; d <= 1e5; d *= 2.0) {
    const double v = reml_f(Xt, yt, lam, n, p, d, sm);
    if (v < best_ll) { best_ll = v; best_d = d; }
  }
  double lo = best_d / 4.0, hi = best_d * 4.0;
  if (lo < 1e-8) lo = 1e-8;
  if (hi > 1e6) hi = 1e6;
  double d;
  if (mode == 0) {
    const double phi = (1.0 + sqrt(5.0)) / 2.0;
    for (int it = 0; it < 40; ++it) {
      const double m1 = hi - (hi - lo) / phi;
      const double m2 = lo + (hi - lo) / phi;
      const double f1 = reml_f(Xt, yt, lam, n, p, m1, sm);
      const double f2 = reml_f(Xt, yt, lam, n, p, m2, sm);
      if (f1 < f2) hi = m2;
      else lo = m1;
    }
    d = 0.5 * (lo + hi);
    double cur = reml_f(Xt, yt, lam, n, p, d, sm);
    for (int it = 0; it < 8; ++it) {
      const double h = fmax(d * 1e-4, 1e-9);
      const double fm = reml_f(Xt, yt, lam, n, p, d - h, sm);
      const double fp = reml_f(Xt, yt, lam, n, p, d + h, sm);
      const double g1 = (fp - fm) / (2.0 * h);
      const double g2 = (fp - 2.0 * cur + fm) / (h * h);
      if (!(g2 > 1e-12)) break;
      const double step = g1 / g2;
      const double dn = d - step;
      if (!(dn > 1e-8 && dn < 1e6)) break;
      const double vn = reml_f(Xt, yt, lam, n, p, dn, sm);
      if (!(vn < cur)) break;
      d = dn;
      cur = vn;
      if (fabs(step) < d * 1e-8) break;
    }
  } else {
    // safeguarded Newton on theta = ln(delta) inside the reference's bracket
    const double tlo = log(lo), thi = log(hi);
    double th = log(best_d);
    double f, g, h;
    reml_fgh(Xt, yt, lam, n, p, exp(th), sm, &f, &g, &h);
    for (int it = 0; it < 60; ++it) {
      double step = (h > 0) ? -g / h : (g > 0 ? -0.5 : 0.5);
      step = fmax(-1.0, fmin(1.0, step));
      double tn = fmin(thi, fmax(tlo, th + step));
      if (fabs(tn - th) < 1e-12) break;
      double fn, gn, hn;
      reml_fgh(Xt, yt, lam, n, p, exp(tn), sm, &fn, &gn, &hn);
      int back = 0;
      while (fn > f + 1e-13 * fabs(f) && back < 30) {
        tn = th + 0.5 * (tn - th);
        reml_fgh(Xt, yt, lam, n, p, exp(tn), sm, &fn, &gn, &hn);
        ++back;
      }
      if (fn > f + 1e-13 * fabs(f)) break;
      const double moved = fabs(tn - th);
      th = tn; f = fn; g = gn; h = hn;
      if (moved < 1e-10) break;
    }
    d = exp(th);
  }
  if (threadIdx.x == 0) delta_out[k] = d;
}

__global__ __launch_bounds__(256) void k_reml_grad_hess(const double* Xt, const double* Yt, const double* lam, int n,
                                                        int p, int G, const double* log_delta, double* f,
                                                        double* g, double* h) {
  __shared__ RemlSmem sm;
  const int k = blockIdx.x;
  if (k >= G) return;
  double F, Gd, H;
  reml_fgh(Xt, Yt + static_cast<size_t>(k) * n, lam, n, p, exp(log_delta[k]), sm, &F, &Gd, &H);
  if (threadIdx.x == 0) { f[k] = F; g[k] = Gd; h[k] = H; }
}

// Plain Cholesky (no pivoting) of sm.A into lower Lc (in sm.A1). Thread 0.
__device__ inline bool chol_lower(RemlSmem& sm, int p) {
  for (int j = 0; j < p; ++j) {
    double s = sm.A[j * kPMax + j];
    for (int k = 0; k < j; ++k) s -= sm.A1[k * kPMax + j] * sm.A1[k * kPMax + j];
    if (!(s > 0)) return false;
    const double ljj = sqrt(s);
    sm.A1[j * kPMax + j] = ljj;
    for (int i = j + 1; i < p; ++i) {
      double t = sm.A[j * kPMax + i];
      for (int k = 0; k < j; ++k) t -= sm.A1[k * kPMax + i] * sm.A1[k * kPMax + j];
      sm.A1[j * kPMax + i] = t / ljj;
    }
  }
  return true;
}

// LMM gene basis at the chosen delta. Column blocks (nq = p + 2):
//   b = 0: d_i ; b = 1..p: s_i W_ic ; b = p+1: s_i r_i   (s = sqrt(d))
// where W = (s X~) Lc^-T is orthonormal (Lc Lc^T = X~^T D X~) and
// r = s y~ - W W^T (s y~) is the whitened null residual; rss_null = ||r||^2.
__global__ __launch_bounds__(256) void k_lmm_basis(const double* Xt, const double* Yt, const double* lam, int n,
                                                   int n_pad, int p, int G, const double* delta, GeneCols gc,
                                                   double* rss_null, int* valid, double* u64) {
  __shared__ RemlSmem sm;
  const int k = blockIdx.x;
  if (k >= G) return;
  const double* yt = Yt + static_cast<size_t>(k) * n;
  const double dk = delta[k];
  reml_pass1<false>(Xt, yt, lam, n, p, dk, sm);
  if (threadIdx.x == 0) sm.flag = chol_lower(sm, p) ? 1 : 0;
  __syncthreads();
  const int nq = p + 2;
  const bool ok = sm.flag != 0;
  // pass A: W rows, u = W^T (s y~)
  double ucol[kPMax];
#pragma unroll
  for (int c = 0; c < kPMax; ++c) ucol[c] = 0;
  for (int i = threadIdx.x; i < n_pad; i += 256) {
    if (i >= n || !ok) {
      for (int b = 0; b < nq; ++b) put_val(gc, k, b, i, 0.0);
      continue;
    }
    double v = dk * lam[i] + 1.0;
    if (v < 1e-12) v = 1e-12;
    const double d = 1.0 / v, s = sqrt(d);
    const double yw = s * yt[i];
    double w[kPMax];
#pragma unroll
    for (int c = 0; c < kPMax; ++c) {
      if (c < p) {
        double t = s * Xt[static_cast<size_t>(c) * n + i];
        for (int q = 0; q < c; ++q) t -= w[q] * sm.A1[q * kPMax + c];
        w[c] = t / sm.A1[c * kPMax + c];
        ucol[c] += w[c] * yw;
      }
    }
    put_val(gc, k, 0, i, d);
#pragma unroll
    for (int c = 0; c < kPMax; ++c)
      if (c < p) put_val(gc, k, 1 + c, i, s * w[c]);
  }
#pragma unroll
  for (int c = 0; c < kPMax; ++c) {
    if (c < p) {
      const double t = block_sum(ucol[c], sm);
      if (threadIdx.x == 0) sm.u[c] = t;
    }
  }
  __syncthreads();
  double rss = 0;
  for (int i = threadIdx.x; i < n && ok; i += 256) {
    double v = dk * lam[i] + 1.0;
    if (v < 1e-12) v = 1e-12;
    const double d = 1.0 / v, s = sqrt(d);
    const double yw = s * yt[i];
    double w[kPMax];
    double r = yw;
#pragma unroll
    for (int c = 0; c < kPMax; ++c) {
      if (c < p) {
        double t = s * Xt[static_cast<size_t>(c) * n + i];
        for (int q = 0; q < c; ++q) t -= w[q] * sm.A1[q * kPMax + c];
        w[c] = t / sm.A1[c * kPMax + c];
        r -= w[c] * sm.u[c];
      }
    }
    rss += r * r;
    put_val(gc, k, p + 1, i, s * r);
    u64[static_cast<size_t>(k) * n + i] = s * r;
  }
  rss = block_sum(rss, sm);
  if (threadIdx.x == 0) {
    rss_null[k] = ok ? rss : 0.0;
    valid[k] = ok ? 1 : 0;
  }
}

// LM: orthonormal basis Wx (n x p, col-major) of X (single workgroup).
__global__ __launch_bounds__(256) void k_lm_basis(const double* X, const double* zeros, int n, int p, double* Wx,
                                                  int* ok_out) {
  __shared__ RemlSmem sm;
  // delta = 0 → weights 1: A = X^T X
  reml_pass1<false>(X, zeros, zeros, n, p, 0.0, sm);
  if (threadIdx.x == 0) sm.flag = chol_lower(sm, p) ? 1 : 0;
  __syncthreads();
  const bool ok = sm.flag != 0;
  for (int i = threadIdx.x; i < n && ok; i += 256) {
    double w[kPMax];
#pragma unroll
    for (int c = 0; c < kPMax; ++c) {
      if (c < p) {
        double t = X[static_cast<size_t>(c) * n + i];
        for (int q = 0; q < c; ++q) t -= w[q] * sm.A1[q * kPMax + c];
        w[c] = t / sm.A1[c * kPMax + c];
        Wx[static_cast<size_t>(c) * n + i] = w[c];
      }
    }
  }
  if (threadIdx.x == 0) *ok_out = ok ? 1 : 0;
}

// LM gene: y_s = y - Wx Wx^T y (written as column block 0), yty = ||y_s||^2.
__global__ __launch_bounds__(256) void k_lm_gene(const double* Y, const double* Wx, int n, int n_pad, int p, int G,
                                                 GeneCols gc, double* yty, int* valid, double* u64) {
  __shared__ RemlSmem sm;
  const int k = blockIdx.x;
  if (k >= G) return;
  const double* y = Y + static_cast<size_t>(k) * n;
  for (int c = 0; c < p; ++c) {
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 256) s += Wx[static_cast<size_t>(c) * n + i] * y[i];
    s = block_sum(s, sm);
    if (threadIdx.x == 0) sm.u[c] = s;
  }
  __syncthreads();
  double ss = 0;
  for (int i = threadIdx.x; i < n_pad; i += 256) {
    double r = 0;
    if (i < n) {
      r = y[i];
      for (int c = 0; c < p; ++c) r -= Wx[static_cast<size_t>(c) * n + i] * sm.u[c];
      ss += r * r;
      u64[static_cast<size_t>(k) * n + i] = r;
    }
    put_val(gc, k, 0, i, r);
  }
  ss = block_sum(ss, sm);
  if (threadIdx.x == 0) {
    yty[k] = ss;
    valid[k] = 1;
  }
}

}  // namespace eqtl_amd
