// Gene side of the negative-binomial score test (prep_glm_nb,
// model_glm.cpp:28-98), one workgroup of 256 threads per gene, FP64:
//   k_glm_null — IRLS for the NB2 log-link null model y ~ X with the
//     dispersion phi held fixed, re-estimated by moments between rounds until
//     it stops moving (nb_irls / nb_mom_phi); then the null working weights
//     w = mu / (1 + phi mu), the Cholesky factor Lc of X^T W X, and the column
//     blocks the fused kernel contracts against the centred dosage:
//       b = 0: w ; b = 1..p: sqrt(w) Z_c, Z = (sqrt(w) X) Lc^-T ;
//       b = p+1: r = (y - mu) / (1 + phi mu)
//     so that gg - |Z^T(sqrt(w) g)|^2 is the score test's information and
//     r^T (g - mean) + mean * sum(r) its score U (gene_scal = sum r).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mfma.hpp"
#include "reml_kernels.hpp"

namespace eqtl_amd {

// A = X^T diag(w) X and b = X^T (w z) over all samples, with (w_i, z_i) from
// row(i, w, z). Collective over the 256 threads; X staged in LDS by chunks.
template <class Row>
__device__ __attribute__((always_inline)) inline void glm_pass(const double* __restrict__ X, int n, int p, Row row, RemlSmem& sm) {
  const int TA = p * (p + 1) / 2;
  const int T = TA + p;
  double acc[kMaxTasks];
#pragma unroll
  for (int t = 0; t < kMaxTasks; ++t) acc[t] = 0;
  for (int i0 = 0; i0 < n; i0 += kChunk) {
    const int rows = min(kChunk, n - i0);
    for (int q = threadIdx.x; q < kChunk * p; q += 256) {
      const int a = q / kChunk, r = q % kChunk;
      sm.Xs[r][a] = r < rows ? X[static_cast<size_t>(a) * n + i0 + r] : 0.0;
    }
    if (threadIdx.x < kChunk) {
      const int r = threadIdx.x;
      double w = 0, z = 0;
      if (r < rows) row(i0 + r, w, z);
      sm.w0[r] = w;
      sm.ys[r] = z;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kMaxTasks; ++s) {
      const int t = threadIdx.x + 256 * s;
      if (t >= T) break;
      double v = 0;
      if (t < TA) {
        int a = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((a + 1) * (a + 2) / 2 <= t) ++a;
        while (a * (a + 1) / 2 > t) --a;
        const int c = t - a * (a + 1) / 2;
        for (int r = 0; r < rows; ++r) v += sm.Xs[r][a] * sm.w0[r] * sm.Xs[r][c];
      } else {
        const int a = t - TA;
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
    if (t < TA) {
      int a = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
      while ((a + 1) * (a + 2) / 2 <= t) ++a;
      while (a * (a + 1) / 2 > t) --a;
      const int c = t - a * (a + 1) / 2;
      sm.A[c * kPMax + a] = acc[s];
      sm.A[a * kPMax + c] = acc[s];
    } else {
      sm.b[t - TA] = acc[s];
    }
  }
  __syncthreads();
}

// mu: G x n scratch (the gene's current mean). Outputs per gene: phi, valid
// (fit converged and X^T W X positive definite), gene_scal = sum r, u64 = r,
// v64 = w (FP64, n each), Lc (p x p, lower, col-major) and the column blocks.
__global__ __launch_bounds__(256) void k_glm_null(const double* X, const double* Y, int n, int n_pad, int p, int G,
                                                  GeneCols gc, double* mu_buf, double* phi_out, int* valid,
                                                  double* gscal, double* u64, double* v64, double* Lc_out) {
  __shared__ RemlSmem sm;
  __shared__ double beta[kPMax], bnew[kPMax];
  __shared__ double s_phi, s_diff;
  __shared__ int s_fail, s_done;
  const int k = blockIdx.x;
  if (k >= G) return;
  const double* y = Y + static_cast<size_t>(k) * n;
  double* mu = mu_buf + static_cast<size_t>(k) * n;
  for (int i = threadIdx.x; i < n; i += 256) mu[i] = fmax(y[i], 0.1);
  if (threadIdx.x < kPMax) beta[threadIdx.x] = 0.0;
  if (threadIdx.x == 0) {
    s_phi = 1.0;
    s_fail = 0;
    s_done = 0;
  }
  __syncthreads();
  const int df = max(1, n - p);
  bool converged = false;
  for (int outer = 0; outer < 10 && !s_fail && !s_done; ++outer) {
    // IRLS with phi fixed (model_glm.cpp:40-63)
    for (int it = 0; it < 50; ++it) {
      const double phi = s_phi;
      glm_pass(X, n, p,
               [&](int i, double& w, double& z) {
                 const double m = fmax(mu[i], 1e-8);
                 const double var = m + phi * m * m;
                 w = (m * m) / fmax(var, 1e-12);
                 z = log(m) + (y[i] - m) / m;
               },
               sm);
      // Lc and u = Lc^-1 b by the block; thread 0 back-substitutes Lc^T x = u
      const bool good = chol_lower_block(sm, p);
      if (threadIdx.x == 0) {
        if (!good) {
          s_fail = 1;
        } else {
          for (int i = p - 1; i >= 0; --i) {
            double t = sm.u[i];
            for (int k2 = i + 1; k2 < p; ++k2) t -= sm.A1[i * kPMax + k2] * bnew[k2];
            bnew[i] = t / sm.A1[i * kPMax + i];
          }
          double diff = 0;
          for (int a = 0; a < p; ++a) diff = fmax(diff, fabs(bnew[a] - beta[a]));
          for (int a = 0; a < p; ++a) beta[a] = bnew[a];
          s_diff = diff;
        }
      }
      __syncthreads();
      if (s_fail) break;
      for (int i = threadIdx.x; i < n; i += 256) {
        double eta = 0;
        for (int a = 0; a < p; ++a) eta += X[static_cast<size_t>(a) * n + i] * beta[a];
        mu[i] = exp(eta);
      }
      __syncthreads();
      if (s_diff < 1e-6) break;
    }
    if (s_fail) break;
    // MoM phi at the current mean (nb_mom_phi, :14-21) and the outer stop (:64-72)
    double num = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
      const double m = fmax(mu[i], 1e-8);
      num += ((y[i] - m) * (y[i] - m) - m) / (m * m);
    }
    num = block_sum(num, sm);
    if (threadIdx.x == 0) {
      const double phi_new = fmax(1e-8, num / static_cast<double>(df));
      if (fabs(phi_new - s_phi) < 1e-6 * fmax(1.0, s_phi)) s_done = 1;
      s_phi = phi_new;
    }
    __syncthreads();
  }
  converged = !s_fail;  // ten rounds without the phi criterion still count as converged (:73)
  const double phi = s_phi;
  // null working weights and X^T W X (prep_glm_nb, :88-97)
  bool ok = converged;
  if (ok) {
    glm_pass(X, n, p,
             [&](int i, double& w, double& z) {
               const double m = fmax(mu[i], 1e-8);
               w = m / (1.0 + phi * m);
               z = 0.0;
             },
             sm);
    ok = chol_lower_block(sm, p);
  }
  for (int idx = threadIdx.x; idx < p * p; idx += 256) {
    const int i = idx % p, jj = idx / p;
    Lc_out[static_cast<size_t>(k) * p * p + idx] = (ok && i >= jj) ? sm.A1[jj * kPMax + i] : 0.0;
  }
  double rsum = 0;
  for (int i = threadIdx.x; i < n_pad; i += 256) {
    if (i >= n || !ok) {
      for (int b = 0; b < p + 2; ++b) put_val(gc, k, b, i, 0.0);
      if (i < n) {
        u64[static_cast<size_t>(k) * n + i] = 0.0;
        v64[static_cast<size_t>(k) * n + i] = 0.0;
      }
      continue;
    }
    const double m = fmax(mu[i], 1e-8);
    const double w = m / (1.0 + phi * m), s = sqrt(w);
    double zr[kPMax];
#pragma unroll
    for (int c = 0; c < kPMax; ++c) {
      if (c < p) {
        double t = s * X[static_cast<size_t>(c) * n + i];
        for (int q = 0; q < c; ++q) t -= zr[q] * sm.A1[q * kPMax + c];
        zr[c] = t / sm.A1[c * kPMax + c];
      }
    }
    put_val(gc, k, 0, i, w);
#pragma unroll
    for (int c = 0; c < kPMax; ++c)
      if (c < p) put_val(gc, k, 1 + c, i, s * zr[c]);
    const double r = (y[i] - mu[i]) / (1.0 + phi * mu[i]);  // test_glm_nb uses the unclamped mean
    put_val(gc, k, p + 1, i, r);
    u64[static_cast<size_t>(k) * n + i] = r;
    v64[static_cast<size_t>(k) * n + i] = w;
    rsum += r;
  }
  rsum = block_sum(rsum, sm);
  if (threadIdx.x == 0) {
    phi_out[k] = phi;
    valid[k] = ok ? 1 : 0;
    gscal[k] = ok ? rsum : 0.0;
  }
}

}  // namespace eqtl_amd
