// Gene-side FP64 kernels (one workgroup of 256 threads per gene):
//   k_lmm_basis      — whitened orthonormal covariate basis W_k, null residual
//                      r_k and the column blocks the fused Wald kernel reads
//                      (prep_lmm, model_lmm.cpp:124-153, in orthonormal form)
//   k_lm_basis / k_lm_gene — prep_lm (model_lm.cpp:19-30) in orthonormal form
// The per-gene REML delta search lives in reml_batch.hpp.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mfma.hpp"

namespace eqtl_amd {

constexpr int kPMax = 32;   // max covariates (incl. intercept)
constexpr int kChunk = 128; // sample rows staged in LDS per pass
constexpr int kMaxTasks = 3;  // per-thread dot products of a pass: p(p+1)/2 + p <= 3 x 256
static_assert(kPMax * (kPMax + 1) / 2 + kPMax <= 256 * kMaxTasks, "pass tasks");

struct RemlSmem {
  double Xs[kChunk][kPMax];  // a chunk of X~ rows
  double ys[kChunk];         // the chunk's y~ (or working response)
  double w0[kChunk];         // the chunk's weights
  double A[kPMax * kPMax];   // X^T W X (full, symmetric)
  double A1[kPMax * kPMax];  // its Cholesky factor Lc (lower, column-major)
  double b[kPMax], u[kPMax]; // X^T W y and u = Lc^-1 b
  double red[4];             // per-wave partial sums of block_sum
  int flag;
};

__device__ inline double block_sum(double v, RemlSmem& sm) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) sm.red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (sm.red[0] + sm.red[1]) + (sm.red[2] + sm.red[3]);
  __syncthreads();
  return r;
}

// Pass over all samples accumulating A = X~^T D X~ (sm.A, both triangles) and
// b = X~^T D y~ (sm.b), d_i = 1 / max(delta lam_i + 1, 1e-12). Collective over
// the 256 threads; X~ is staged in LDS by chunks of kChunk rows.
__device__ __attribute__((always_inline)) inline void reml_pass1(const double* __restrict__ Xt,
                                                                 const double* __restrict__ yt,
                                                                 const double* __restrict__ lam, int n, int p,
                                                                 double delta, RemlSmem& sm) {
  const int TA = p * (p + 1) / 2;
  const int T = TA + p;
  double acc[kMaxTasks];
#pragma unroll
  for (int t = 0; t < kMaxTasks; ++t) acc[t] = 0;
  for (int i0 = 0; i0 < n; i0 += kChunk) {
    const int rows = min(kChunk, n - i0);
    for (int q = threadIdx.x; q < kChunk * p; q += 256) {
      const int a = q / kChunk, r = q % kChunk;
      sm.Xs[r][a] = r < rows ? Xt[static_cast<size_t>(a) * n + i0 + r] : 0.0;
    }
    if (threadIdx.x < kChunk) {
      const int r = threadIdx.x;
      double w0 = 0, y = 0;
      if (r < rows) {
        const int i = i0 + r;
        double v = delta * lam[i] + 1.0;
        if (v < 1e-12) v = 1e-12;
        w0 = 1.0 / v;
        y = yt[i];
      }
      sm.w0[r] = w0;
      sm.ys[r] = y;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kMaxTasks; ++s) {
      const int t = threadIdx.x + 256 * s;
      if (t >= T) break;
      double v = 0;
      if (t < TA) {
        // triangular index -> (a, c), c <= a
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

// Rank-deficient covariates. A column whose remaining (Schur) variance after
// the earlier columns is at most kDependentTol of its own is a linear
// combination of them: its pivot and factor column are set to zero, so the
// basis spans the covariates' column space without it (with an exactly
// dependent column the reference's LDLT meets a zero pivot and its solves
// take the pseudo-inverse, model_lmm.cpp:141-151 / model_lm.cpp:25, which is
// the same projection). A pivot below -kDependentTol of the column's own
// variance means the matrix is not positive semidefinite: failure.
constexpr double kDependentTol = 1e-10;

// 1: pivot sqrt(s); 0: dependent column (pivot 0); -1: not positive semidefinite
__device__ inline int semidef_pivot(double s, double ajj, double* l) {
  if (s > kDependentTol * ajj) {
    *l = sqrt(s);
    return 1;
  }
  *l = 0.0;
  return s >= -kDependentTol * ajj ? 0 : -1;
}

// Cholesky (no pivoting) of sm.A into lower Lc (in sm.A1), dependent columns
// zeroed as above. Thread 0.
__device__ inline bool chol_lower(RemlSmem& sm, int p) {
  for (int j = 0; j < p; ++j) {
    double s = sm.A[j * kPMax + j];
    for (int k = 0; k < j; ++k) s -= sm.A1[k * kPMax + j] * sm.A1[k * kPMax + j];
    double ljj;
    if (semidef_pivot(s, sm.A[j * kPMax + j], &ljj) < 0) return false;
    sm.A1[j * kPMax + j] = ljj;
    for (int i = j + 1; i < p; ++i) {
      double t = sm.A[j * kPMax + i];
      for (int k = 0; k < j; ++k) t -= sm.A1[k * kPMax + i] * sm.A1[k * kPMax + j];
      sm.A1[j * kPMax + i] = ljj > 0 ? t / ljj : 0.0;
    }
  }
  return true;
}

// The same factorisation by the whole block (every thread calls it), column by
// column: thread 0 forms the pivot, threads j < i < p the column below it, each
// element with the serial form's operation order (so the factor is identical).
// Thread p also carries the border row b^T through the same steps, leaving
// u = Lc^-1 b in sm.u (0 for a dependent column). Returns false when the
// matrix is not positive semidefinite.
__device__ inline bool chol_lower_block(RemlSmem& sm, int p) {
  const int t = threadIdx.x;
  if (t == 0) sm.flag = 1;
  __syncthreads();
  for (int j = 0; j < p; ++j) {
    if (t == 0) {
      double s = sm.A[j * kPMax + j];
      for (int k = 0; k < j; ++k) s -= sm.A1[k * kPMax + j] * sm.A1[k * kPMax + j];
      double ljj;
      if (semidef_pivot(s, sm.A[j * kPMax + j], &ljj) < 0) sm.flag = 0;
      sm.A1[j * kPMax + j] = ljj;
    }
    __syncthreads();
    if (!sm.flag) break;
    const double ljj = sm.A1[j * kPMax + j];
    if (t > j && t < p) {
      double x = sm.A[j * kPMax + t];
      for (int k = 0; k < j; ++k) x -= sm.A1[k * kPMax + t] * sm.A1[k * kPMax + j];
      sm.A1[j * kPMax + t] = ljj > 0 ? x / ljj : 0.0;
    } else if (t == p) {
      double x = sm.b[j];
      for (int k = 0; k < j; ++k) x -= sm.A1[k * kPMax + j] * sm.u[k];
      sm.u[j] = ljj > 0 ? x / ljj : 0.0;
    }
    __syncthreads();
  }
  const bool ok = sm.flag != 0;
  __syncthreads();
  return ok;
}

// LMM gene basis at the chosen delta. Column blocks (nq = p + 2):
//   b = 0: d_i ; b = 1..p: s_i W_ic ; b = p+1: s_i r_i   (s = sqrt(d))
// where W = (s X~) Lc^-T is orthonormal (Lc Lc^T = X~^T D X~) and
// r = s y~ - W u, u = W^T (s y~) = Lc^-1 X~^T D y~, is the whitened null
// residual; rss_null = ||r||^2. One pass over the samples: row i's W entries
// come from a forward substitution against Lc, and u from the same
// substitution applied once to X~^T D y~ (no per-column reductions).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_lmm_basis(const double* Xt, const double* Yt, const double* lam, int n,
                                                   int n_pad, int p, int G, const double* delta, GeneCols gc,
                                                   double* rss_null, int* valid, double* u64, double* Lc_out) {
  __shared__ RemlSmem sm;
  const int k = blockIdx.x;
  if (k >= G) return;
  const double* yt = Yt + static_cast<size_t>(k) * n;
  const double dk = delta[k];
  reml_pass1(Xt, yt, lam, n, p, dk, sm);
  const bool ok = chol_lower_block(sm, p);  // Lc, and u = Lc^-1 b
  for (int idx = threadIdx.x; idx < p * p; idx += 256) {
    const int i = idx % p, jj = idx / p;
    Lc_out[static_cast<size_t>(k) * p * p + idx] = (i >= jj && ok) ? sm.A1[jj * kPMax + i] : 0.0;
  }
  const int nq = p + 2;
  double rss = 0;
  for (int i = threadIdx.x; i < n_pad; i += 256) {
    if (i >= n || !ok) {
      for (int b = 0; b < nq; ++b) put_val(gc, k, b, i, 0.0);
      continue;
    }
    // keep the Lc reads inside the loop (hoisting the p(p+1)/2 loop-invariant
    // LDS values into registers spills)
    asm volatile("" ::: "memory");
    double v = dk * lam[i] + 1.0;
    if (v < 1e-12) v = 1e-12;
    const double d = 1.0 / v, s = sqrt(d);
    double r = s * yt[i];
    double w[kPMax];
    put_val(gc, k, 0, i, d);
#pragma unroll
    for (int c = 0; c < kPMax; ++c) {
      if (c < p) {
        double t = s * Xt[static_cast<size_t>(c) * n + i];
        for (int q = 0; q < c; ++q) t -= w[q] * sm.A1[q * kPMax + c];
        const double lcc = sm.A1[c * kPMax + c];
        w[c] = lcc > 0 ? t / lcc : 0.0;  // a dependent covariate adds no direction
        r -= w[c] * sm.u[c];
        put_val(gc, k, 1 + c, i, s * w[c]);
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

// Interpolation nodes of the low-rank trans LMM scan: at node k (REML delta
// node_delta[k]) the whitened covariate projection Z = D X~ Lc^-T (n x p, the
// columns s o W of k_lmm_basis), written to column col0[k] + a * stride[k] of
// Z (column-major, leading dimension n_pad, rows past n zero). One workgroup
// per node; ok_out[k] = 0 when X~^T D X~ is not positive semidefinite.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_lr_node_basis(
    const double* Xt, const double* zeros, const double* lam, int n, int n_pad, int p, const double* node_delta,
    const int32_t* col0, const int32_t* stride, double* Z, int* ok_out) {
  __shared__ RemlSmem sm;
  const int k = blockIdx.x;
  const double dk = node_delta[k];
  reml_pass1(Xt, zeros, lam, n, p, dk, sm);
  const bool ok = chol_lower_block(sm, p);
  const size_t c0 = static_cast<size_t>(col0[k]), cs = static_cast<size_t>(stride[k]);
  for (int i = threadIdx.x; i < n_pad; i += 256) {
    if (i >= n || !ok) {
      for (int c = 0; c < p; ++c) Z[(c0 + c * cs) * n_pad + i] = 0.0;
      continue;
    }
    asm volatile("" ::: "memory");
    double v = dk * lam[i] + 1.0;
    if (v < 1e-12) v = 1e-12;
    const double s = sqrt(1.0 / v);
    double w[kPMax];
#pragma unroll
    for (int c = 0; c < kPMax; ++c) {
      if (c < p) {
        double t = s * Xt[static_cast<size_t>(c) * n + i];
        for (int q = 0; q < c; ++q) t -= w[q] * sm.A1[q * kPMax + c];
        const double lcc = sm.A1[c * kPMax + c];
        w[c] = lcc > 0 ? t / lcc : 0.0;
        Z[(c0 + c * cs) * n_pad + i] = s * w[c];
      }
    }
  }
  if (threadIdx.x == 0) ok_out[k] = ok ? 1 : 0;
}

// LM: orthonormal basis Wx (n x p, col-major) of X (single workgroup).
__global__ __launch_bounds__(256) void k_lm_basis(const double* X, const double* zeros, int n, int p, double* Wx,
                                                  int* ok_out) {
  __shared__ RemlSmem sm;
  // delta = 0 → weights 1: A = X^T X
  reml_pass1(X, zeros, zeros, n, p, 0.0, sm);
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
        const double lcc = sm.A1[c * kPMax + c];
        w[c] = lcc > 0 ? t / lcc : 0.0;  // a dependent covariate adds no direction
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
