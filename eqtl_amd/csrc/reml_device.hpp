// Device-resident pieces of the batched REML search (optimize_delta,
// model_lmm.cpp:39-86): an FP64 matrix-core GEMM for the weighted Grams, and
// the per-gene state machine of the reference search (grid, 40 golden-section
// steps, midpoint, at most 8 central-difference Newton steps), so that the
// whole search for a batch of genes is a fixed sequence of launches with no
// host round trip. That sequence is captured once in a hipGraph and replayed
// per batch (eqtl_gpu.hip, reml_ref_device).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mfma.hpp"

namespace eqtl_amd {

// ---- C = A^T B in FP64 on matrix cores ----------------------------------------
// A: K x M (column m at A + m * lda, K-contiguous), B: K x N (column j at
// B + j * ldb), C: M x N column-major (C[j * ldc + m]). Workgroup tile 64 x 64,
// wave (wr, wc) owns the 32 x 32 quadrant as 2 x 2 blocks of
// v_mfma_f64_16x16x4f64 (fragment grouping of mfma_f64x8: lane l supplies
// k = 8 (l >> 4) + t of row l & 15; result element r is row (l >> 4) + 4 r,
// column l & 15). Operands are read straight from L2 (the GEMMs here are
// skinny: M <= 528, K = n samples); entries past M, N or K read as zero.
constexpr int kDgT = 64;

__device__ inline f64x8 dg_frag(const double* col, bool in, int k0, int K) {
  f64x8 v;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int k = k0 + t;
    v.v[t] = (in && k < K) ? col[k] : 0.0;
  }
  return v;
}

__global__ __launch_bounds__(256) void k_dgemm_tn(int M, int N, int K, const double* __restrict__ A, int lda,
                                                  const double* __restrict__ B, int ldb, double* __restrict__ C,
                                                  int ldc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int m0 = blockIdx.x * kDgT + wr * 32, n0 = blockIdx.y * kDgT + wc * 32;
  const int kk = 8 * (lane >> 4);
  const double* acol[2];
  const double* bcol[2];
  bool ain[2], bin[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int m = m0 + 16 * b + (lane & 15), j = n0 + 16 * b + (lane & 15);
    ain[b] = m < M;
    bin[b] = j < N;
    acol[b] = A + static_cast<size_t>(ain[b] ? m : 0) * lda;
    bcol[b] = B + static_cast<size_t>(bin[b] ? j : 0) * ldb;
  }
  f64x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f64x4{0, 0, 0, 0};
  for (int k0 = 0; k0 < K; k0 += 32) {
    f64x8 af[2], bf[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      af[b] = dg_frag(acol[b], ain[b], k0 + kk, K);
      bf[b] = dg_frag(bcol[b], bin[b], k0 + kk, K);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = mfma_f64x8(af[a], bf[b], acc[a][b]);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = n0 + 16 * b + (lane & 15);
      if (j >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 16 * a + (lane >> 4) + 4 * r;
        if (m < M) C[static_cast<size_t>(j) * ldc + m] = acc[a][b][r];
      }
    }
}

// ---- reference search state, one entry per gene ------------------------------
// The arithmetic below restates the host form of the search step for step
// (same expressions, no contraction into FMAs), so the device search gives
// the same delta bit for bit.
struct RemlRefState {
  double lo, hi, d, cur, h, dn, step;
  int32_t active, pending;
};

// bracket [best / 4, best * 4] clipped to [1e-8, 1e6] around the grid's best point
__global__ void k_ref_init(const double* best, int G, RemlRefState* st) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  RemlRefState s;
  s.lo = fmax(best[k] / 4.0, 1e-8);
  s.hi = fmin(best[k] * 4.0, 1e6);
  s.d = s.cur = s.h = s.dn = s.step = 0.0;
  s.active = s.pending = 0;
  st[k] = s;
}

// golden-section points: ed[k] = hi - (hi - lo) / phi, ed[G + k] = lo + (hi - lo) / phi
__global__ void k_ref_golden_points(const RemlRefState* st, int G, double phi, double* ed) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  const double lo = st[k].lo, hi = st[k].hi;
  ed[k] = hi - (hi - lo) / phi;
  ed[G + k] = lo + (hi - lo) / phi;
}

__global__ void k_ref_golden_update(RemlRefState* st, int G, const double* ed, const double* F) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  if (F[k] < F[G + k]) st[k].hi = ed[G + k];
  else st[k].lo = ed[k];
}

// midpoint of the final bracket (evaluated next into cur)
__global__ void k_ref_mid(RemlRefState* st, int G, double* ed) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  const double d = 0.5 * (st[k].lo + st[k].hi);
  st[k].d = d;
  ed[k] = d;
}

__global__ void k_ref_mid_done(RemlRefState* st, int G, const double* F) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  st[k].cur = F[k];
  st[k].active = 1;
  st[k].pending = 0;
}

// central-difference points d -+ h, h = max(d 1e-4, 1e-9) (inactive genes: d)
__global__ void k_ref_newton_points(RemlRefState* st, int G, double* ed) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  const double d = st[k].d;
  if (st[k].active) {
    const double h = fmax(d * 1e-4, 1e-9);
    st[k].h = h;
    ed[k] = d - h;
    ed[G + k] = d + h;
  } else {
    ed[k] = d;
    ed[G + k] = d;
  }
}

// Newton proposal from the differences; a gene whose curvature is not
// positive or whose step leaves (1e-8, 1e6) stops. ed[k] <- the trial point.
__global__ void k_ref_newton_propose(RemlRefState* st, int G, const double* F, double* ed) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  RemlRefState s = st[k];
  s.pending = 0;
  if (s.active) {
    const double h = s.h, fm = F[k], fp = F[G + k];
    const double g1 = (fp - fm) / (2.0 * h);
    const double g2 = (fp - 2.0 * s.cur + fm) / (h * h);
    if (!(g2 > 1e-12)) {
      s.active = 0;
    } else {
      const double step = g1 / g2;
      const double dn = s.d - step;
      if (!(dn > 1e-8 && dn < 1e6)) {
        s.active = 0;
      } else {
        s.pending = 1;
        s.dn = dn;
        s.step = step;
      }
    }
  }
  ed[k] = s.pending ? s.dn : s.d;
  st[k] = s;
}

// accept a trial point that lowers f; stop when the step is below d 1e-8
__global__ void k_ref_newton_accept(RemlRefState* st, int G, const double* F) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G) return;
  RemlRefState s = st[k];
  if (s.pending) {
    if (!(F[k] < s.cur)) {
      s.active = 0;
    } else {
      s.d = s.dn;
      s.cur = F[k];
      if (fabs(s.step) < s.d * 1e-8) s.active = 0;
    }
    s.pending = 0;
  }
  st[k] = s;
}

__global__ void k_ref_out(const RemlRefState* st, int G, double* delta) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < G) delta[k] = st[k].d;
}

__global__ void k_iota(int32_t* v, int G, int reps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < G * reps) v[i] = i % G;
}

}  // namespace eqtl_amd
