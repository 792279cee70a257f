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

// Split over k when gridDim.z > 1: workgroup z sums samples [z kchunk,
// (z + 1) kchunk) into its own output C + z cstride (k_dgemm_sum_splits adds
// them up).
__global__ __launch_bounds__(256) void k_dgemm_tn(int M, int N_arg, int K, const double* __restrict__ A, int lda,
                                                  const double* __restrict__ B, int ldb, double* __restrict__ C,
                                                  int ldc, const int* N_dev, int kchunk, int64_t cstride) {
  const int N = N_dev ? *N_dev : N_arg;  // columns counted on the device (N_arg: the grid's bound)
  if (static_cast<int>(blockIdx.y) * kDgT >= N) return;
  {
    const int kb = static_cast<int>(blockIdx.z) * kchunk;
    A += kb;
    B += kb;
    K = min(K - kb, kchunk);
    C += static_cast<int64_t>(blockIdx.z) * cstride;
  }
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
  // operands of the next 32-sample step are loaded before this step's MFMAs
  f64x8 af[2], bf[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    af[b] = dg_frag(acol[b], ain[b], kk, K);
    bf[b] = dg_frag(bcol[b], bin[b], kk, K);
  }
  for (int k0 = 0; k0 < K; k0 += 32) {
    f64x8 an[2], bn[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      an[b] = dg_frag(acol[b], ain[b], k0 + 32 + kk, K);
      bn[b] = dg_frag(bcol[b], bin[b], k0 + 32 + kk, K);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = mfma_f64x8(af[a], bf[b], acc[a][b]);
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      af[b] = an[b];
      bf[b] = bn[b];
    }
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

// C (M x N, ldc) = sum over the S split outputs P + z M N (leading dimension M).
__global__ void k_dgemm_sum_splits(const double* __restrict__ P, int S, int M, int N, double* __restrict__ C,
                                   int ldc) {
  const int64_t MN = static_cast<int64_t>(M) * N;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < MN;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double v = 0.0;
    for (int z = 0; z < S; ++z) v += P[z * MN + t];
    const int64_t j = t / M, m = t - j * M;
    C[j * ldc + m] = v;
  }
}

// ---- fused f-only evaluation inputs on matrix cores ----------------------------
// For evaluation e (gene ek[e], delta ed[e]) with d_i = 1 / max(delta lam_i + 1,
// 1e-12): CA[e][t] = sum_i P[t][i] d_i (the packed X~^T D X~), CB[e][a] =
// sum_i X~[i][a] d_i y~_i (b), S[e][0] = sum_i y~_i^2 d_i (c0) and S[e][3] =
// sum_i ln v_i. The weights are formed in registers as the B operand of
// v_mfma_f64_16x16x4f64 (lane: evaluation l & 15 of its wave's 16, samples
// 8 (l >> 4) + [0..7] of the 32-sample step), against MB blocks of 16 packed
// entries of P and PB blocks of 16 covariates of X~ as A, so no weight matrix
// is written. The A operands come fragment-major (Af, k_frag_pack): per
// 32-sample step, MB + PB blocks of 4 KiB, each 4 x 1 KiB (lane l's values
// 2h, 2h + 1 of its 8 in chunk h), staged in LDS by direct 16-byte copies one
// step ahead (double-buffered) and shared by the four waves of the
// workgroup (4 x 16 evaluations). The count of evaluations is E, or *E_dev
// when given (read on the device, so that a captured launch sequence can
// evaluate a varying number).
template <int MB, int PB>
__host__ __device__ constexpr size_t gram_stage_bytes() {
  return static_cast<size_t>(MB + PB) * 4096;
}

// Af[((s * nb + b) * 4 + h) * 128 + 2 lane + j] = M[16 b + (lane & 15)][32 s + 8 (lane >> 4) + 2 h + j]
// for block b < nb (rows of M from row0: M = src rows, column-major with
// leading dimension n), zero past the rows or the samples.
__global__ void k_frag_pack(const double* __restrict__ src, int rows, int n, int nb, int b0, int nb_total,
                            double* __restrict__ Af) {
  const int nks = (n + 31) / 32;
  const size_t total = static_cast<size_t>(nks) * nb * 4 * 128;
  for (size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int j = static_cast<int>(idx & 1), lane = static_cast<int>((idx >> 1) & 63);
    const int h = static_cast<int>((idx >> 7) & 3);
    const size_t q = idx >> 9;
    const int b = static_cast<int>(q % nb), st = static_cast<int>(q / nb);
    const int r = 16 * b + (lane & 15), k = 32 * st + 8 * (lane >> 4) + 2 * h + j;
    const double v = (r < rows && k < n) ? src[static_cast<size_t>(r) * n + k] : 0.0;
    Af[((static_cast<size_t>(st) * nb_total + b0 + b) * 4 + h) * 128 + 2 * lane + j] = v;
  }
}

template <int MB, int PB>
__global__ __launch_bounds__(256) void k_reml_gram_f(const double* __restrict__ Af, const double* __restrict__ Yt,
                                                     const double* __restrict__ lam, int n, int p, int TA, int E_arg,
                                                     const int* E_dev, const int32_t* __restrict__ ek,
                                                     const double* __restrict__ ed, double* __restrict__ CA,
                                                     double* __restrict__ CB, double* __restrict__ S, int nscal) {
  extern __shared__ __attribute__((aligned(16))) unsigned char gsm[];
  constexpr int NB = MB + PB;
  constexpr size_t kStage = gram_stage_bytes<MB, PB>();
  const int E = E_dev ? *E_dev : E_arg;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (static_cast<int>(blockIdx.x) * 64 >= E) return;  // whole workgroup: no barrier is left waiting
  const int e_blk = blockIdx.x * 64 + wave * 16;
  const int e = min(e_blk + (lane & 15), E - 1);
  const double dl = ed[e];
  const double* y = Yt + static_cast<size_t>(ek[e]) * n;
  const int kk = 8 * (lane >> 4);
  const int nks = (n + 31) / 32;
  // this wave's copies of a stage: chunks q = wave + 4 i (NB per wave)
  auto issue = [&](int st, unsigned char* buf) {
    const unsigned char* src = reinterpret_cast<const unsigned char*>(Af) + static_cast<size_t>(st) * kStage;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int q = wave + 4 * i;
      __builtin_amdgcn_global_load_lds(src + q * 1024 + lane * 16,
                                       (__attribute__((address_space(3))) void*)(buf + q * 1024), 16, 0, 0);
    }
  };
  f64x4 a1[MB], a2[PB];
#pragma unroll
  for (int b = 0; b < MB; ++b) a1[b] = f64x4{0, 0, 0, 0};
#pragma unroll
  for (int b = 0; b < PB; ++b) a2[b] = f64x4{0, 0, 0, 0};
  double c0 = 0, lv = 0;
  issue(0, gsm);
  for (int st = 0; st < nks; ++st) {
    unsigned char* cur = gsm + (st & 1) * kStage;
    // this step's sample values first (older than the next stage's copies,
    // so waiting for them leaves the copies in flight)
    double yv[8], lv8[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int i = 32 * st + kk + t;
      yv[t] = i < n ? y[i] : 0.0;
      lv8[t] = i < n ? lam[i] : 0.0;
    }
    if (st + 1 < nks) issue(st + 1, gsm + ((st + 1) & 1) * kStage);
    f64x8 wd, wy;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int i = 32 * st + kk + t;
      if (i < n) {
        double v = dl * lv8[t] + 1.0;
        if (v < 1e-12) v = 1e-12;
        const double d = 1.0 / v, yi = yv[t];
        wd.v[t] = d;
        wy.v[t] = d * yi;
        c0 += yi * (d * yi);
        lv += log(v);
      } else {
        wd.v[t] = 0.0;
        wy.v[t] = 0.0;
      }
    }
    // the stage of this step (copied one step ago by all four waves) has landed
    if (st + 1 < nks) __builtin_amdgcn_s_waitcnt((NB & 0xF) | ((NB >> 4) << 14) | (0x7 << 4) | (0xF << 8));
    else __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      f64x8 af;
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const double2 v = *reinterpret_cast<const double2*>(cur + (b * 4 + h) * 1024 + lane * 16);
        af.v[2 * h] = v.x;
        af.v[2 * h + 1] = v.y;
      }
      if (b < MB) a1[b] = mfma_f64x8(af, wd, a1[b]);
      else a2[b - MB] = mfma_f64x8(af, wy, a2[b - MB]);
    }
    __syncthreads();  // every wave is done with cur before the copies of step st + 2 overwrite it
  }
  // lanes l, l ^ 16, l ^ 32, l ^ 48 hold the same evaluation
  c0 += __shfl_xor(c0, 16);
  c0 += __shfl_xor(c0, 32);
  lv += __shfl_xor(lv, 16);
  lv += __shfl_xor(lv, 32);
  const int eo = e_blk + (lane & 15);
  if (eo >= E) return;
  if (lane < 16) {
    S[static_cast<size_t>(eo) * nscal + 0] = c0;
    S[static_cast<size_t>(eo) * nscal + 3] = lv;
  }
#pragma unroll
  for (int b = 0; b < MB; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int t = 16 * b + (lane >> 4) + 4 * r;
      if (t < TA) CA[static_cast<size_t>(eo) * TA + t] = a1[b][r];
    }
#pragma unroll
  for (int b = 0; b < PB; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int c = 16 * b + (lane >> 4) + 4 * r;
      if (c < p) CB[static_cast<size_t>(eo) * p + c] = a2[b][r];
    }
}

// ---- reference search state, one entry per gene ------------------------------
// The kernels below restate the search's steps expression for expression,
// with no contraction into FMAs.
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

// Central-difference points of the active genes, compacted: list[q] = gene,
// evaluations 2q (d - h) and 2q + 1 (d + h), h = max(d 1e-4, 1e-9); *count =
// active genes (zeroed before the launch). Positions follow atomic order; each
// evaluation depends on its gene only, so the results do not.
__global__ void k_ref_newton_points(RemlRefState* st, int G, int* count, int* list, int32_t* ek, double* ed) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G || !st[k].active) return;
  const double d = st[k].d;
  const double h = fmax(d * 1e-4, 1e-9);
  st[k].h = h;
  const int q = atomicAdd(count, 1);
  list[q] = k;
  ek[2 * q] = k;
  ek[2 * q + 1] = k;
  ed[2 * q] = d - h;
  ed[2 * q + 1] = d + h;
}

// count2 = 2 * count (the evaluations of the points above)
__global__ void k_ref_twice(const int* count, int* count2) {
  if (threadIdx.x == 0) *count2 = 2 * *count;
}

// Newton proposal of active gene list[q] from its differences; a gene whose
// curvature is not positive or whose step leaves (1e-8, 1e6) stops.
// Evaluation q (gene list[q]) is its trial point (d itself when none).
__global__ void k_ref_newton_propose(RemlRefState* st, const int* count, const int* list, const double* F,
                                     int32_t* ek, double* ed) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= *count) return;
  const int k = list[q];
  RemlRefState s = st[k];
  s.pending = 0;
  const double h = s.h, fm = F[2 * q], fp = F[2 * q + 1];
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
  ek[q] = k;
  ed[q] = s.pending ? s.dn : s.d;
  st[k] = s;
}

// accept a trial point that lowers f; stop when the step is below d 1e-8
__global__ void k_ref_newton_accept(RemlRefState* st, const int* count, const int* list, const double* F) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= *count) return;
  const int k = list[q];
  RemlRefState s = st[k];
  if (s.pending) {
    if (!(F[q] < s.cur)) {
      s.active = 0;
    } else {
      s.d = s.dn;
      s.cur = F[q];
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
