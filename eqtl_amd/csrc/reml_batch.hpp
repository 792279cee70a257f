// Batched REML evaluation: reml_negll (model_lmm.cpp:10-37) and its first two
// derivatives in theta = ln(delta) for many (gene, delta) pairs at once.
//
// With x_i the i-th row of X~ = Q^T X and d_i = 1/max(delta*lambda_i + 1, 1e-12):
//   A(delta) = sum_i d_i x_i x_i^T,   b = sum_i d_i y~_i x_i,   c0 = sum_i d_i y~_i^2.
// Every weighted Gram is one GEMM against the per-sample outer products
// P[i, t] = x_ia x_ic (t = a(a+1)/2 + c, c <= a), so an evaluation costs a
// GEMM column instead of a pass over the samples with a p x p accumulation:
//   A = P^T w,  b = X~^T (w o y~)   (w = d, d' or d'' for A, A1, A2 / b, b1, b2).
// The small p x p work (pivoted LDLT with the reference's pivot rule, the
// inverse, traces) runs one wavefront per evaluation.
//
// The grid stage of the reference search evaluates every gene at the same 35
// deltas, where A(delta) is gene-independent: 35 factorizations plus two GEMMs
// (b for all genes: (X~ o d_g)^T Y~; c0: d_g^T (Y~ o Y~)) cover all genes.
//
// Derivatives (e_i = 1 - d_i; D' = -d e, D'' = d e (e - d)):
//   q  = c0 - 2 b^T beta + beta^T A beta, beta = A^-1 b  (= c0 - b^T beta, in
//        the form that is stationary in beta)
//   q' = c1 - 2 beta^T b1 + beta^T A1 beta
//   u  = b1 - A1 beta,  q'' = c2 - 2 beta^T b2 + beta^T A2 beta - 2 u^T A^-1 u
//   f   = 1/2 [df ln(q/df) + sum ln v + ln det A]
//   f'  = 1/2 [df q'/q + sum e + tr(A^-1 A1)]
//   f'' = 1/2 [df (q''/q - (q'/q)^2) + sum d e + tr(A^-1 A2) - tr((A^-1 A1)^2)]
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace eqtl_amd {

constexpr int kRP = 32;        // max covariates
constexpr int kLD = kRP + 1;   // LDS leading dimension of the p x p tiles
constexpr int kNumScal = 8;    // per-evaluation sums: c0 c1 c2 logv se sde

__host__ __device__ inline int tri_index(int a, int c) { return a >= c ? a * (a + 1) / 2 + c : c * (c + 1) / 2 + a; }

// P[t * n + i] = X~[i, a] X~[i, c]
__global__ void k_reml_outer(const double* Xt, int n, int p, double* P) {
  const int t = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int a = 0;
  while ((a + 1) * (a + 2) / 2 <= t) ++a;
  const int c = t - a * (a + 1) / 2;
  P[static_cast<size_t>(t) * n + i] = Xt[static_cast<size_t>(a) * n + i] * Xt[static_cast<size_t>(c) * n + i];
}

__device__ inline double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ inline double block_sum256(double v, double* red) {
  v = wave_sum_d(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  for (int q = 0; q < (static_cast<int>(blockDim.x) >> 6); ++q) s += red[q];
  return s;
}

// Grid weights: Dg[g * n + i] = d_i(delta_g), logv[g] = sum ln v_i. One block per g.
__global__ __launch_bounds__(256) void k_reml_grid_weights(const double* lam, int n, const double* dgrid,
                                                           double* Dg, double* logv) {
  __shared__ double red[4];
  const int g = blockIdx.x;
  const double dl = dgrid[g];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double v = dl * lam[i] + 1.0;
    if (v < 1e-12) v = 1e-12;
    Dg[static_cast<size_t>(g) * n + i] = 1.0 / v;
    s += log(v);
  }
  s = block_sum256(s, red);
  if (threadIdx.x == 0) logv[g] = s;
}

// XD[(g * p + a) * n + i] = X~[i, a] d_i(delta_g)
__global__ void k_reml_grid_xd(const double* Xt, const double* Dg, int n, int p, double* XD) {
  const int ga = blockIdx.y;  // g * p + a
  const int g = ga / p, a = ga % p;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  XD[static_cast<size_t>(ga) * n + i] = Xt[static_cast<size_t>(a) * n + i] * Dg[static_cast<size_t>(g) * n + i];
}

__global__ void k_square(const double* x, size_t N, double* y) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < N) y[i] = x[i] * x[i];
}

// Per-evaluation weights for gene-specific deltas, with the derivatives in
// ln delta. W / Z: n x 3E, column w * E + e holds weight set w (d, d', d'') of
// evaluation e and the same times y~ of its gene.
// S[e * kNumScal + .] = c0 c1 c2 logv se sde. One block per e.
__global__ __launch_bounds__(256) void k_reml_eval_weights_fgh(const double* lam, const double* Yt, int n, int E,
                                                               const int32_t* ek, const double* ed, double* W,
                                                               double* Z, double* S) {
  __shared__ double red[4];
  const int e = blockIdx.x;
  const double dl = ed[e];
  const double* y = Yt + static_cast<size_t>(ek[e]) * n;
  double c0 = 0, c1 = 0, c2 = 0, lv = 0, se = 0, sde = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double v = dl * lam[i] + 1.0;
    if (v < 1e-12) v = 1e-12;
    const double d = 1.0 / v, yi = y[i];
    W[static_cast<size_t>(e) * n + i] = d;
    Z[static_cast<size_t>(e) * n + i] = d * yi;
    c0 += yi * (d * yi);
    lv += log(v);
    const double ee = 1.0 - d;
    const double w1 = -d * ee, w2 = d * ee * (ee - d);
    W[(static_cast<size_t>(E) + e) * n + i] = w1;
    Z[(static_cast<size_t>(E) + e) * n + i] = w1 * yi;
    W[(2 * static_cast<size_t>(E) + e) * n + i] = w2;
    Z[(2 * static_cast<size_t>(E) + e) * n + i] = w2 * yi;
    c1 += w1 * yi * yi;
    c2 += w2 * yi * yi;
    se += ee;
    sde += d * ee;
  }
  c0 = block_sum256(c0, red);
  lv = block_sum256(lv, red);
  c1 = block_sum256(c1, red);
  c2 = block_sum256(c2, red);
  se = block_sum256(se, red);
  sde = block_sum256(sde, red);
  if (threadIdx.x == 0) {
    double* s = S + static_cast<size_t>(e) * kNumScal;
    s[0] = c0; s[1] = c1; s[2] = c2; s[3] = lv; s[4] = se; s[5] = sde;
  }
}

// Inputs of EV f-only evaluations per one-wave block (p <= PB): the weights
// d_i for the X~^T D X~ GEMM (W), and directly b = X~^T (d o y~) (CB, p values),
// c0 = y~^T D y~ and sum ln v_i (S[0], S[3]). X~ (n x p, column-major) is
// shared by all evaluations; each row of it is loaded once per block and used
// for the block's EV evaluations. One wave per block keeps the closing
// reductions to a single wave (EV (p + 2) butterfly sums).
template <int PB, int EV>
__global__ __launch_bounds__(64) void k_reml_eval_weights_f(const double* lam, const double* Yt, const double* Xt,
                                                             int n, int p, int E_arg, const int* E_dev,
                                                             const int32_t* ek, const double* ed, double* W,
                                                             double* CB, double* S) {
  constexpr int NV = PB + 2;  // per evaluation: c0, sum ln v, b[0..p)
  const int E = E_dev ? *E_dev : E_arg;  // evaluations counted on the device (E_arg: the grid's bound)
  const int e0 = blockIdx.x * EV;
  if (e0 >= E) return;
  double dl[EV];
  const double* y[EV];
#pragma unroll
  for (int v = 0; v < EV; ++v) {
    const int e = min(e0 + v, E - 1);  // a spare slot repeats the last evaluation and writes nothing
    dl[v] = ed[e];
    y[v] = Yt + static_cast<size_t>(ek[e]) * n;
  }
  double acc[EV][NV];
#pragma unroll
  for (int v = 0; v < EV; ++v)
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[v][q] = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double x[PB];
#pragma unroll
    for (int a = 0; a < PB; ++a) x[a] = a < p ? Xt[static_cast<size_t>(a) * n + i] : 0.0;
    const double li = lam[i];
#pragma unroll
    for (int v = 0; v < EV; ++v) {
      double vv = dl[v] * li + 1.0;
      if (vv < 1e-12) vv = 1e-12;
      const double d = 1.0 / vv, yi = y[v][i], dy = d * yi;
      if (e0 + v < E) W[static_cast<size_t>(e0 + v) * n + i] = d;
      acc[v][0] += yi * dy;
      acc[v][1] += log(vv);
#pragma unroll
      for (int a = 0; a < PB; ++a) acc[v][2 + a] += x[a] * dy;
    }
  }
  const int lane = threadIdx.x;
#pragma unroll
  for (int v = 0; v < EV; ++v) {
    const int e = e0 + v;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double t = wave_sum_d(acc[v][q]);
      if (lane == q && e < E && q < p + 2) {
        if (q == 0) S[static_cast<size_t>(e) * kNumScal + 0] = t;
        else if (q == 1) S[static_cast<size_t>(e) * kNumScal + 3] = t;
        else CB[static_cast<size_t>(e) * p + (q - 2)] = t;
      }
    }
  }
}

// ---- one wavefront (= one 64-thread block) per p x p problem ----------------

struct WaveLdlt {
  double L[kRP * kLD];    // input A (full), factors on return
  double Ai[kRP * kLD];   // A^-1 (column c at Ai[c * kLD])
  double D[kRP];
  double tmp[kRP];
  int tr[kRP];
  int ok;
};

// Load a packed triangle (TA entries, column of a GEMM result) as a full matrix.
__device__ inline void load_sym(const double* tri, int p, double* M) {
  for (int idx = threadIdx.x; idx < p * p; idx += 64) {
    const int a = idx / p, c = idx % p;
    M[c * kLD + a] = tri[tri_index(a, c)];
  }
}

// LDLT with the pivot taken on the not-yet-updated trailing diagonal (the rule
// of Eigen 3.4's LDLT, model_lmm.cpp:23). Sums keep the sequential order of
// the single-threaded form. m(i, j) = L[j * kLD + i], lower triangle used.
// BORDER: row p holds b^T and m(p, p) = c; the border never pivots, and its
// last step leaves the Schur complement c - b^T A^-1 b in D[p].
// LANES threads (64, or 32 for two problems per wave) share one problem;
// ld: leading dimension of L.
template <bool BORDER, int LANES>
__device__ inline void wave_ldlt_factor(double* L, int ld, double* D, double* tmp, int* tr, int* ok, int p) {
  const int lane = threadIdx.x % LANES;
  const int pe = BORDER ? p + 1 : p;
  auto m = [&](int i, int j) -> double& { return L[j * ld + i]; };
  if (lane == 0) *ok = 1;
  __syncthreads();
  for (int k = 0; k < pe; ++k) {
    if (k < p) {
      // pivot: largest |m(i,i)|, p > i >= k, first index on ties
      double bv = -1.0;
      int big = p;
      if (lane >= k && lane < p) { bv = fabs(m(lane, lane)); big = lane; }
#pragma unroll
      for (int o = LANES / 2; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o);
        const int oi = __shfl_xor(big, o);
        if (ov > bv || (ov == bv && oi < big)) { bv = ov; big = oi; }
      }
      if (!(big < p)) big = k;  // all candidates NaN: keep k (as the sequential scan does)
      if (big != k) {
        if (lane < k) { const double t = m(k, lane); m(k, lane) = m(big, lane); m(big, lane) = t; }
        else if (lane > big && lane < pe) { const double t = m(lane, k); m(lane, k) = m(lane, big); m(lane, big) = t; }
        else if (lane > k && lane < big) { const double t = m(lane, k); m(lane, k) = m(big, lane); m(big, lane) = t; }
        else if (lane == k) { const double t = m(k, k); m(k, k) = m(big, big); m(big, big) = t; }
      }
      if (lane == 0) tr[k] = big;
      __syncthreads();
    }
    if (k > 0) {
      if (lane < k) tmp[lane] = D[lane] * m(k, lane);
      __syncthreads();
      if (lane == k) {
        double s = 0;
        for (int j = 0; j < k; ++j) s += m(k, j) * tmp[j];
        m(k, k) -= s;
      } else if (lane > k && lane < pe) {
        double t = 0;
        for (int j = 0; j < k; ++j) t += m(lane, j) * tmp[j];
        m(lane, k) -= t;
      }
      __syncthreads();
    }
    const double akk = m(k, k);
    if (lane == 0) {
      D[k] = akk;
      if (k < p && !isfinite(akk)) *ok = 0;
    }
    if (lane > k && lane < pe && fabs(akk) > 0.0) m(lane, k) /= akk;
    __syncthreads();
  }
}

// Factor (above), then the explicit inverse.
__device__ inline void wave_ldlt_inverse(WaveLdlt& w, int p) {
  const int lane = threadIdx.x;
  auto m = [&](int i, int j) -> double& { return w.L[j * kLD + i]; };
  wave_ldlt_factor<false, 64>(w.L, kLD, w.D, w.tmp, w.tr, &w.ok, p);
  // inverse: lane c solves A x = e_c in column c of Ai
  if (lane < p) {
    double* y = w.Ai + lane * kLD;
    for (int i = 0; i < p; ++i) y[i] = i == lane ? 1.0 : 0.0;
    for (int k = 0; k < p; ++k) { const int t = w.tr[k]; const double v = y[k]; y[k] = y[t]; y[t] = v; }
    for (int i = 0; i < p; ++i) {
      double s = y[i];
      for (int j = 0; j < i; ++j) s -= m(i, j) * y[j];
      y[i] = s;
    }
    for (int i = 0; i < p; ++i) y[i] = w.D[i] != 0.0 ? y[i] / w.D[i] : 0.0;
    for (int i = p - 1; i >= 0; --i) {
      double s = y[i];
      for (int j = i + 1; j < p; ++j) s -= m(j, i) * y[j];
      y[i] = s;
    }
    for (int k = p - 1; k >= 0; --k) { const int t = w.tr[k]; const double v = y[k]; y[k] = y[t]; y[t] = v; }
  }
  __syncthreads();
}

// log det A from the factors; false when a pivot is not positive (the reference
// returns 1e300 then).
__device__ inline bool ldlt_logdet(const WaveLdlt& w, int p, double* ld) {
  if (!w.ok) return false;
  double s = 0;
  for (int a = 0; a < p; ++a) {
    if (!(w.D[a] > 0)) return false;
    s += log(w.D[a]);
  }
  *ld = s;
  return true;
}

// Factor the grid's A(delta_g): Ainv[g] (p x p, ld p), logdet[g], ok[g].
__global__ __launch_bounds__(64) void k_reml_grid_factor(const double* CA, int TA, int p, double* Ainv,
                                                         double* logdet, int32_t* ok) {
  __shared__ WaveLdlt w;
  const int g = blockIdx.x;
  load_sym(CA + static_cast<size_t>(g) * TA, p, w.L);
  __syncthreads();
  wave_ldlt_inverse(w, p);
  for (int idx = threadIdx.x; idx < p * p; idx += 64) {
    const int a = idx % p, c = idx / p;
    Ainv[static_cast<size_t>(g) * p * p + idx] = w.Ai[c * kLD + a];
  }
  if (threadIdx.x == 0) {
    double ld = 0;
    const bool good = ldlt_logdet(w, p, &ld);
    logdet[g] = ld;
    ok[g] = good ? 1 : 0;
  }
}

// f at every (grid point g, gene k): Bg (ng p x G), C0 (ng x G) -> F[k * ng + g].
__global__ __launch_bounds__(256) void k_reml_grid_f(const double* Ainv, const double* CAg, int TA,
                                                     const double* logdet, const int32_t* okg,
                                                     const double* logv, const double* Bg, const double* C0, int ng,
                                                     int p, int G, int df, double* F) {
  __shared__ double Ai[kRP * kRP], Am[kRP * kRP];
  const int g = blockIdx.y;
  for (int idx = threadIdx.x; idx < p * p; idx += blockDim.x) {
    Ai[idx] = Ainv[static_cast<size_t>(g) * p * p + idx];
    Am[idx] = CAg[static_cast<size_t>(g) * TA + tri_index(idx / p, idx % p)];
  }
  __syncthreads();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  double f = 1e300;
  if (okg[g]) {
    const double* b = Bg + static_cast<size_t>(k) * ng * p + static_cast<size_t>(g) * p;
    double bb[kRP];
#pragma unroll
    for (int a = 0; a < kRP; ++a) bb[a] = a < p ? b[a] : 0.0;
    double beta[kRP];
#pragma unroll
    for (int a = 0; a < kRP; ++a) {
      double t = 0;
      if (a < p)
#pragma unroll
        for (int c = 0; c < kRP; ++c)
          if (c < p) t += Ai[a * p + c] * bb[c];
      beta[a] = t;
    }
    // q = c0 - 2 b^T beta + beta^T A beta (stationary in beta)
    double q = C0[static_cast<size_t>(k) * ng + g];
#pragma unroll
    for (int a = 0; a < kRP; ++a) {
      if (a < p) {
        double t = 0;
#pragma unroll
        for (int c = 0; c < kRP; ++c)
          if (c < p) t += Am[a * p + c] * beta[c];
        q += beta[a] * t - 2.0 * bb[a] * beta[a];
      }
    }
    if (q <= 0) q = 1e-12;
    f = 0.5 * (df * log(q / df) + logv[g] + logdet[g]);
  }
  F[static_cast<size_t>(k) * ng + g] = f;
}

// Reference grid argmin per gene: index 0 is delta = 1, then ascending grid;
// strict improvement keeps the first minimum (model_lmm.cpp:43-48).
__global__ void k_reml_grid_argmin(const double* F, const double* dgrid, int ng, int G, double* best_d,
                                   double* best_f) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  double bf = F[static_cast<size_t>(k) * ng], bd = dgrid[0];
  for (int g = 1; g < ng; ++g) {
    const double v = F[static_cast<size_t>(k) * ng + g];
    if (v < bf) { bf = v; bd = dgrid[g]; }
  }
  best_d[k] = bd;
  best_f[k] = bf;
}

// f alone: the LDLT of the bordered matrix [A b; b^T c0] (pivots within A)
// gives log det A from the first p pivots and q = c0 - b^T A^-1 b as the last,
// with no inverse. LANES = 32 runs two evaluations per wave (p <= 31), 64 one.
// Dynamic LDS per evaluation: reml_eval_f_bytes(p), sized to p so that many
// evaluations are in flight per CU (the factorisation is latency-bound).
__host__ __device__ constexpr int reml_eval_f_doubles(int p) { return (p + 1) * (p + 1) + 2 * (p + 1) + (p + 1); }
__host__ __device__ constexpr size_t reml_eval_f_bytes(int p) { return sizeof(double) * reml_eval_f_doubles(p) + 8; }

template <int LANES>
__global__ __launch_bounds__(64) void k_reml_eval_f(const double* CA, const double* CB, const double* S, int E_arg,
                                                    const int* E_dev, int TA, int p, int df, double* Fo) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_f[];
  const int E = E_dev ? *E_dev : E_arg;  // evaluations counted on the device (E_arg: the grid's bound)
  constexpr int kPer = 64 / LANES;
  if (static_cast<int>(blockIdx.x) * kPer >= E) return;
  const int sub = threadIdx.x / LANES, lane = threadIdx.x % LANES;
  const int e_raw = blockIdx.x * kPer + sub;
  const int e = e_raw < E ? e_raw : E - 1;  // a spare half repeats the last evaluation and writes nothing
  const int ld = p + 1;
  double* L = reinterpret_cast<double*>(smem_f + sub * reml_eval_f_bytes(p));
  double* D = L + ld * (p + 1);
  double* tmp = D + (p + 1);
  int* tr = reinterpret_cast<int*>(tmp + (p + 1));
  int* ok = tr + (p + 1);
  const double* s = S + static_cast<size_t>(e) * kNumScal;
  const double* tri = CA + static_cast<size_t>(e) * TA;
  for (int idx = lane; idx < p * p; idx += LANES) {
    const int a = idx / p, c = idx % p;
    L[c * ld + a] = tri[tri_index(a, c)];
  }
  if (lane < p) L[lane * ld + p] = CB[static_cast<size_t>(e) * p + lane];
  if (lane == 0) L[p * ld + p] = s[0];
  __syncthreads();
  wave_ldlt_factor<true, LANES>(L, ld, D, tmp, tr, ok, p);
  if (lane == 0 && e_raw < E) {
    bool good = *ok != 0;
    double lg = 0;
    for (int a = 0; a < p && good; ++a) {
      if (!(D[a] > 0)) good = false;
      else lg += log(D[a]);
    }
    double q = D[p];
    if (q <= 0) q = 1e-12;
    Fo[e] = good ? 0.5 * (df * log(q / df) + s[3] + lg) : 1e300;
  }
}

// f, f' and f'' (in ln delta), one evaluation per 64-thread block, from the
// GEMM results CA (TA x 3E), CB (p x 3E) and the sums S.
__global__ __launch_bounds__(64) void k_reml_eval_fgh(const double* CA, const double* CB, const double* S, int E, int TA,
                                                  int p, int df, double* Fo, double* Go, double* Ho) {
  __shared__ WaveLdlt w;
  __shared__ double A1[kRP * kLD], A2[kRP * kLD];
  __shared__ double vb[3][kRP], beta[kRP], u[kRP], wv[kRP];
  const int e = blockIdx.x;
  const int lane = threadIdx.x;
  load_sym(CA + static_cast<size_t>(e) * TA, p, w.L);
  load_sym(CA + (static_cast<size_t>(E) + e) * TA, p, A1);
  load_sym(CA + (2 * static_cast<size_t>(E) + e) * TA, p, A2);
  if (lane < p) {
    vb[0][lane] = CB[static_cast<size_t>(e) * p + lane];
    vb[1][lane] = CB[(static_cast<size_t>(E) + e) * p + lane];
    vb[2][lane] = CB[(2 * static_cast<size_t>(E) + e) * p + lane];
  }
  __syncthreads();
  wave_ldlt_inverse(w, p);
  double ld = 0;
  const bool good = ldlt_logdet(w, p, &ld);
  const double* s = S + static_cast<size_t>(e) * kNumScal;
  if (!good) {
    if (lane == 0) {
      Fo[e] = 1e300;
      Go[e] = 0;
      Ho[e] = 0;
    }
    return;
  }
  auto Ai = [&](int a, int c) { return w.Ai[c * kLD + a]; };
  // beta = A^-1 b
  if (lane < p) {
    double t = 0;
    for (int c = 0; c < p; ++c) t += Ai(lane, c) * vb[0][c];
    beta[lane] = t;
  }
  __syncthreads();
  // q = c0 - 2 b^T beta + beta^T A beta: stationary in beta, so the explicit
  // inverse's error in beta enters q only to second order.
  double bAb = 0;
  if (lane < p) {
    const double* tri = CA + static_cast<size_t>(e) * TA;
    double t = 0;
    for (int c = 0; c < p; ++c) t += tri[tri_index(lane, c)] * beta[c];
    bAb = beta[lane] * t;
  }
  bAb = wave_sum_d(bAb);
  double q = s[0];
  for (int a = 0; a < p; ++a) q -= 2.0 * vb[0][a] * beta[a];
  q += bAb;
  if (q <= 0) q = 1e-12;
  const double F = 0.5 * (df * log(q / df) + s[3] + ld);
  // u = b1 - A1 beta, wv = A^-1 u
  if (lane < p) {
    double t = vb[1][lane];
    for (int c = 0; c < p; ++c) t -= A1[c * kLD + lane] * beta[c];
    u[lane] = t;
  }
  __syncthreads();
  if (lane < p) {
    double t = 0;
    for (int c = 0; c < p; ++c) t += Ai(lane, c) * u[c];
    wv[lane] = t;
  }
  // traces: lane a takes row a of M = A^-1 A1 (kept in registers) and of A^-1 A2
  double tr1 = 0, tr2 = 0, tr11 = 0;
  double bA1b = 0, bA2b = 0;
  if (lane < p) {
    const int a = lane;
    for (int c = 0; c < p; ++c) {
      double mac = 0, m2 = 0;
      for (int j = 0; j < p; ++j) {
        mac += Ai(a, j) * A1[c * kLD + j];
        m2 += Ai(a, j) * A2[c * kLD + j];
      }
      if (c == a) { tr1 += mac; tr2 += m2; }
      // tr(M M) = sum_{a,c} M_ac M_ca; M_ca = sum_j Ai(c, j) A1(j, a)
      double mca = 0;
      for (int j = 0; j < p; ++j) mca += Ai(c, j) * A1[a * kLD + j];
      tr11 += mac * mca;
    }
    double r1 = 0, r2 = 0;
    for (int c = 0; c < p; ++c) {
      r1 += A1[c * kLD + a] * beta[c];
      r2 += A2[c * kLD + a] * beta[c];
    }
    bA1b = beta[a] * r1;
    bA2b = beta[a] * r2;
  }
  tr1 = wave_sum_d(tr1);
  tr2 = wave_sum_d(tr2);
  tr11 = wave_sum_d(tr11);
  bA1b = wave_sum_d(bA1b);
  bA2b = wave_sum_d(bA2b);
  __syncthreads();
  if (lane == 0) {
    double bb1 = 0, bb2 = 0, uw = 0;
    for (int a = 0; a < p; ++a) {
      bb1 += beta[a] * vb[1][a];
      bb2 += beta[a] * vb[2][a];
      uw += u[a] * wv[a];
    }
    const double q1 = s[1] - 2.0 * bb1 + bA1b;
    const double q2 = s[2] - 2.0 * bb2 + bA2b - 2.0 * uw;
    Fo[e] = F;
    Go[e] = 0.5 * (df * q1 / q + s[4] + tr1);
    Ho[e] = 0.5 * (df * (q2 / q - (q1 / q) * (q1 / q)) + s[5] + tr2 - tr11);
  }
}

}  // namespace eqtl_amd
