// K7, trans LMM in the split form with the covariate projections interpolated
// in ln delta (scan_trans_lmm.cpp:11-40,160-233 per test; per-gene statistics
// scan_trans.hpp:87-110).
//
// A test needs gg = g~^T D_k g~, num = (s_k o r_k)^T g~ and the projection
// W_k^T h = Z_k^T g~ onto the gene's whitened covariates, Z_k = D_k X~ Lc_k^-T
// (D_k = diag(1 / (delta_k lam + 1)), Lc_k Lc_k^T = X~^T D_k X~). Z_k is an
// analytic function of u = ln delta_k with the poles of 1 / (1 + lam e^u) at
// distance pi from the real axis, so on an interval of u of width 6.5 its
// Chebyshev interpolant at 32 nodes u_r is exact to ~1e-12 (64 / 96 / 128
// nodes for widths 13 / 23 / 33). Genes are sorted by ln delta into
// super-groups of 64 (four gene groups of 16) and each super-group takes a
// segment of u with nodes u_r; then
//     Z_k^T g~ = sum_r l_r(u_k) T_r,   T_r = Z(u_r)^T g~ = (Q Z(u_r))^T g,
// and the T_r of every SNP are one more column block of the rotation (the
// same A x [Q | Q Z] product), shared by all genes of the segment. The test's
// p projections then cost R (nodes) multiply-adds instead of n each.
//
// One workgroup = one super-group (64 genes) x a strided set of 128-row SNP
// tiles, eight waves: wave w = 4 c + r covers rows 32 r .. 32 r + 31 of a tile
// against groups 2c, 2c + 1 of the super-group.
//  * Main loop over n (32-sample k-steps): per group two column blocks, d_k
//    (against the squared rows -> gg) and s_k o r_k (-> num), split bf16,
//    three products each. The super-group's gene tile (8 blocks x hi/lo, 16
//    KiB per k-step) is copied global -> LDS by the half-0 waves
//    (double-buffered), the row pair's next stage by the half-1 waves (a ring
//    of three stages).
//  * Projections: per covariate a and 32-node chunk, the row block's T values
//    (fp32, from global memory) split three ways (hi, mid, lo: exact) against
//    the genes' interpolation weights l_r(u_k) split three ways: six bf16
//    products (errors ~2^-24 relative), FP32 accumulation over the nodes, then
//    the square is added to the test's FP64 sum of squares.
//  * Epilogue per group: gg, sum of squares, num folded into LDS, then the FP64
//    test of the eight-wave LMM kernel (refinement, polish hand-off, hits,
//    ACAT, best) through w8_epilogue.
#pragma once
#include "wald_lmm.hpp"

namespace eqtl_amd {

struct LrArgs {
  WaldArgs w;
  const float* Tp;           // m_pad x ldT node projections of the block's SNPs
  int64_t ldT;
  const int32_t* sg_col;     // per super-group: first T column of its segment
  const int32_t* sg_R;       // per super-group: nodes of its segment (32, 64, 96 or 128)
  const int32_t* slot_gene;  // slot (ln-delta order) -> gene, -1 for padding slots
  const __bf16* lcoef;       // interpolation weights, see lr_coef_index
  const __bf16* gb;          // gene tiles, see lr_tile_index
  int32_t n_sg, p;
};

constexpr int kLrStages = 3;
constexpr size_t kLrTileBytes = 16 * 1024;  // one k-step of a super-group's gene tile
constexpr size_t kLrRegion0 = 2 * kLrTileBytes + kLrStages * kAStageBytes;
constexpr size_t kLrFoldBytes = sizeof(double) * 3 * kW8Tests;
constexpr size_t kLrLdsBytes = kLrRegion0 + sizeof(RowInfo) + sizeof(double) * kTailHead + sizeof(GeneAcc) * 64;
static_assert(kLrFoldBytes + kW8MergeBytes <= kLrRegion0, "LR fold and merge scratch alias the staging buffers");
static_assert(kLrLdsBytes + 64 <= kLdsCap, "LR Wald LDS layout");

// Element (slot, block b, sample i) of the gene tiles: per super-group sg and
// k-step s one 16 KiB tile, the hi plane (8 blocks x 64 lanes x 8 values)
// then the lo plane; block 2 q + b holds quantity b (0: d, 1: s o r) of group
// q = (slot >> 4) & 3, lane = (slot & 15) + 16 * ((i & 31) >> 3), value i & 7.
__host__ __device__ inline size_t lr_tile_index(int slot, int b, int i, int nks, int plane) {
  const int sg = slot >> 6, q = (slot >> 4) & 3, col = slot & 15;
  const int s = i >> 5, lane = col + 16 * ((i & 31) >> 3);
  return ((static_cast<size_t>(sg) * nks + s) * 2 + plane) * 4096 + (2 * q + b) * 512 + lane * 8 + (i & 7);
}

// Interpolation weight r of slot: per group of 16 slots and 32-node chunk a
// 3 x 512 block (hi, mid, lo planes), lane (slot & 15) + 16 * ((r & 31) >> 3),
// value r & 7 — the B fragment of v_mfma_f32_16x16x32_bf16 (genes as columns).
__host__ __device__ inline size_t lr_coef_index(int slot, int r, int plane) {
  const int lane = (slot & 15) + 16 * ((r & 31) >> 3);
  return (static_cast<size_t>(slot >> 4) * 4 + (r >> 5)) * 1536 + plane * 512 + lane * 8 + (r & 7);
}

// x = hi + mid + lo exactly (fp32 -> three bf16 pieces)
__device__ inline void split3_bf16(const f32x8& x, bf16x8& hi, bf16x8& mid, bf16x8& lo) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const __bf16 h = to_bf16(x.v[t]);
    const float r1 = x.v[t] - static_cast<float>(h);
    const __bf16 m = to_bf16(r1);
    hi[t] = h;
    mid[t] = m;
    lo[t] = to_bf16(r1 - static_cast<float>(m));
  }
}

// One 32-sample k-step of wave (c, r): its row pair from the current stage
// (a_rd), the gene tile (b_cur); half 0 copies its share of the next gene
// tile, half 1 the row pair's stage three k-steps on. The LDS pointers are
// __restrict__ so the compiler does not drain the copies in flight before the
// reads.
__device__ __attribute__((always_inline)) inline void lr_kstep(
    f32x4 (&ag)[2][kRB], f32x4 (&an)[2][kRB], const unsigned char* __restrict__ a_rd,
    unsigned char* __restrict__ a_wr, const float* const (&a_src)[kRB], const unsigned char* __restrict__ b_cur,
    unsigned char* __restrict__ b_nxt, const unsigned char* b_src, int c, int r, int lane) {
  f32x8 ac[kRB];
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb) {
    const float4 lo = *reinterpret_cast<const float4*>(a_rd + (rb * 2 + 0) * 1024 + lane * 16);
    const float4 hi = *reinterpret_cast<const float4*>(a_rd + (rb * 2 + 1) * 1024 + lane * 16);
    ac[rb].v[0] = lo.x; ac[rb].v[1] = lo.y; ac[rb].v[2] = lo.z; ac[rb].v[3] = lo.w;
    ac[rb].v[4] = hi.x; ac[rb].v[5] = hi.y; ac[rb].v[6] = hi.z; ac[rb].v[7] = hi.w;
  }
  if (c == 0) {
    if (b_nxt) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int ch = r + 4 * k;
        lds_direct16(b_src + ch * 1024 + lane * 16, b_nxt + ch * 1024);
      }
    }
  } else if (a_wr) {
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) {
      lds_direct16(a_src[rb], a_wr + rb * 2048);
      lds_direct16(a_src[rb] + 4, a_wr + rb * 2048 + 1024);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  bf16x8 ah[kRB], al[kRB], sh[kRB], sl[kRB];
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb) {
    split_bf16(ac[rb], ah[rb], al[rb]);
    f32x8 sv;
#pragma unroll
    for (int t = 0; t < 8; ++t) sv.v[t] = ac[rb].v[t] * ac[rb].v[t];
    split_bf16(sv, sh[rb], sl[rb]);
  }
#pragma unroll
  for (int gq = 0; gq < 2; ++gq) {
    const int q = 2 * c + gq;
    const bf16x8 dh = *reinterpret_cast<const bf16x8*>(b_cur + (2 * q) * 1024 + lane * 16);
    const bf16x8 dl = *reinterpret_cast<const bf16x8*>(b_cur + 8192 + (2 * q) * 1024 + lane * 16);
    const bf16x8 vh = *reinterpret_cast<const bf16x8*>(b_cur + (2 * q + 1) * 1024 + lane * 16);
    const bf16x8 vl = *reinterpret_cast<const bf16x8*>(b_cur + 8192 + (2 * q + 1) * 1024 + lane * 16);
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) {
      ag[gq][rb] = mfma_bf16(sh[rb], dh, ag[gq][rb]);
      an[gq][rb] = mfma_bf16(ah[rb], vh, an[gq][rb]);
    }
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) {
      ag[gq][rb] = mfma_bf16(sh[rb], dl, ag[gq][rb]);
      an[gq][rb] = mfma_bf16(ah[rb], vl, an[gq][rb]);
    }
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) {
      ag[gq][rb] = mfma_bf16(sl[rb], dh, ag[gq][rb]);
      an[gq][rb] = mfma_bf16(al[rb], vh, an[gq][rb]);
    }
  }
}

// Gene tiles of the low-rank scan, one workgroup per gene k at slot
// slot_of[k]: block 0 d_i = 1 / max(delta_k lam_i + 1, 1e-12), block 1 the
// whitened null residual s o r (u64, as written by k_lmm_basis), fp32 then
// split bf16 hi / lo (samples past n stay zero).
__global__ __launch_bounds__(256) void k_lr_gene_tiles(const double* lam, const double* delta, const double* u64,
                                                       const int32_t* slot_of, int n, int nks, int G, __bf16* gb) {
  const int k = blockIdx.x;
  if (k >= G) return;
  const int slot = slot_of[k];
  const double dk = delta[k];
  for (int i = threadIdx.x; i < n; i += 256) {
    double v = dk * lam[i] + 1.0;
    if (v < 1e-12) v = 1e-12;
    const double q[2] = {1.0 / v, u64[static_cast<size_t>(k) * n + i]};
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const float f = static_cast<float>(q[b]);
      const __bf16 h = to_bf16(f);
      gb[lr_tile_index(slot, b, i, nks, 0)] = h;
      gb[lr_tile_index(slot, b, i, nks, 1)] = to_bf16(f - static_cast<float>(h));
    }
  }
}

// fp64 column-major (leading dimension ld) -> fp32 -> bf16 hi / lo, same layout
__global__ void k_split_bf16_f64(const double* x, int64_t count, __bf16* hi, __bf16* lo) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < count;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float f = static_cast<float>(x[t]);
    const __bf16 h = to_bf16(f);
    hi[t] = h;
    lo[t] = to_bf16(f - static_cast<float>(h));
  }
}

__device__ inline W8Gene lr_gene(const LrArgs& la, int group) {
  const WaldArgs& a = la.w;
  W8Gene g;
  g.gl = threadIdx.x & 15;
  g.gene = la.slot_gene[group * 16 + g.gl];
  g.ok = g.gene >= 0 && g.gene < a.G && a.gene_valid[g.gene];
  g.gscal = g.ok ? a.gene_scal[g.gene] : 0.0;
  const bool excl = g.ok && a.gene_chrom && a.gene_chrom[g.gene] >= 0;
  g.gch = excl ? a.gene_chrom[g.gene] : -1;
  g.gcs = excl ? a.gene_cs[g.gene] : 0;
  g.gce = excl ? a.gene_ce[g.gene] : 0;
  g.win_known = excl && a.snp_chrom;
  return g;
}

__global__ __launch_bounds__(kW8Threads) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_wald_lr(LrArgs la) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const WaldArgs& a = la.w;
  // XCD-aware: the n_splits consecutive workgroups of one XCD share a
  // super-group, whose gene tiles are then served from that XCD's L2
  const int wg = blockIdx.x;
  const int jx = wg >> 3;
  const int sg = (jx / a.n_splits) * 8 + (wg & 7);
  const int split = jx % a.n_splits;
  if (sg >= la.n_sg) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = wave >> 2, r = wave & 3;
  const int gl = lane & 15, kk = 8 * (lane >> 4);
  const int nks = a.n_pad >> 5;

  unsigned char* const buf0 = smem;
  unsigned char* const buf1 = smem + kLrTileBytes;
  unsigned char* const abuf = smem + 2 * kLrTileBytes;
  RowInfo& ri = *reinterpret_cast<RowInfo*>(smem + kLrRegion0);
  double* thead = reinterpret_cast<double*>(smem + kLrRegion0 + sizeof(RowInfo));
  GeneAcc* gstate = reinterpret_cast<GeneAcc*>(smem + kLrRegion0 + sizeof(RowInfo) + sizeof(double) * kTailHead);
  double* fold = reinterpret_cast<double*>(smem);
  GeneAcc* wpart = reinterpret_cast<GeneAcc*>(smem + kLrFoldBytes);
  for (int i = threadIdx.x; i < kTailHead; i += kW8Threads) thead[i] = (&a.tail->df)[i];
  if (threadIdx.x < 64) acc_init(gstate[threadIdx.x]);
  const TailView tv{thead, a.tail->coef_b};
  const unsigned char* gsrc = reinterpret_cast<const unsigned char*>(la.gb) + static_cast<size_t>(sg) * nks * kLrTileBytes;
  const int pair_off = r * 4096;
  const int segR = la.sg_R[sg], segc = la.sg_col[sg], nqr = segR >> 5;
  const int n_tiles = (a.n_tiles + 1) >> 1;

  for (int tile = split; tile < n_tiles; tile += a.n_splits) {
    const int64_t row0 = static_cast<int64_t>(tile) * kTileRows;
    __syncthreads();  // the previous tile's epilogue is done with ri, fold and the gene partials
    w8_row_info<EQTL_MODEL_LMM>(a, ri, row0);
    const int64_t rowP = row0 + 32 * r;
    const float* a_row[kRB];
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) a_row[rb] = a.A + (rowP + 16 * rb + gl) * a.n_pad + kk;
    f32x4 ag[2][kRB], an[2][kRB];
#pragma unroll
    for (int gq = 0; gq < 2; ++gq)
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) {
        ag[gq][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
        an[gq][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
    // prologue: gene tile of k-step 0 (half 0), the first row stages (half 1)
    if (c == 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int ch = r + 4 * k;
        lds_direct16(gsrc + ch * 1024 + lane * 16, buf0 + ch * 1024);
      }
    } else {
      for (int s = 0; s < kLrStages - 1 && s < nks; ++s) {
        unsigned char* dst = abuf + s * kAStageBytes + pair_off;
#pragma unroll
        for (int rb = 0; rb < kRB; ++rb) {
          lds_direct16(a_row[rb] + s * 32, dst + rb * 2048);
          lds_direct16(a_row[rb] + s * 32 + 4, dst + rb * 2048 + 1024);
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int s = 0; s < nks; ++s) {
      const int sr = s + kLrStages - 1;  // k-step whose rows are fetched now
      const float* src[kRB];
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) src[rb] = a_row[rb] + sr * 32;
      unsigned char* const a_wr = sr < nks ? abuf + (sr % kLrStages) * kAStageBytes + pair_off : nullptr;
      const unsigned char* const a_rd = abuf + (s % kLrStages) * kAStageBytes + pair_off;
      unsigned char* const b_nxt = s + 1 < nks ? (((s + 1) & 1) ? buf1 : buf0) : nullptr;
      lr_kstep(ag, an, a_rd, a_wr, src, (s & 1) ? buf1 : buf0, b_nxt, gsrc + static_cast<size_t>(s + 1) * kLrTileBytes,
               c, r, lane);
      // the next gene tile (half 0's copies) and the next row stage (issued
      // two k-steps ago by half 1) must have landed; half 1 leaves the stage
      // it issued now (its four youngest operations) in flight
      if (c == 1 && s + 2 < nks) __builtin_amdgcn_s_waitcnt(kWaitVm4Lgkm0);
      else __builtin_amdgcn_s_waitcnt(0);
      lds_barrier();
    }
    // projections through the interpolation weights, squared and summed
    double ssq[2][kRB][4];
#pragma unroll
    for (int gq = 0; gq < 2; ++gq)
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
        for (int e = 0; e < 4; ++e) ssq[gq][rb][e] = 0.0;
    const float* trow[kRB];
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb)
      trow[rb] = la.Tp + static_cast<size_t>(rowP + 16 * rb + gl) * la.ldT + segc + kk;
    const __bf16* lbase = la.lcoef + static_cast<size_t>((sg * 4 + 2 * c) * 4) * 1536 + lane * 8;
#pragma unroll 1
    for (int ca = 0; ca < la.p; ++ca) {
      f32x4 pj[2][kRB];
#pragma unroll
      for (int gq = 0; gq < 2; ++gq)
#pragma unroll
        for (int rb = 0; rb < kRB; ++rb) pj[gq][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
      for (int qr = 0; qr < nqr; ++qr) {
        f32x8 xv[kRB];
#pragma unroll
        for (int rb = 0; rb < kRB; ++rb) xv[rb] = load8(trow[rb] + ca * segR + 32 * qr);
        bf16x8 Lh[2], Lm[2], Ll[2];
#pragma unroll
        for (int gq = 0; gq < 2; ++gq) {
          const __bf16* lb = lbase + static_cast<size_t>(gq * 4 + qr) * 1536;
          Lh[gq] = load8_bf16(lb);
          Lm[gq] = load8_bf16(lb + 512);
          Ll[gq] = load8_bf16(lb + 1024);
        }
#pragma unroll
        for (int rb = 0; rb < kRB; ++rb) {
          bf16x8 xh, xm, xl;
          split3_bf16(xv[rb], xh, xm, xl);
#pragma unroll
          for (int gq = 0; gq < 2; ++gq) {
            f32x4 acc = pj[gq][rb];
            acc = mfma_bf16(xh, Lh[gq], acc);
            acc = mfma_bf16(xh, Lm[gq], acc);
            acc = mfma_bf16(xm, Lh[gq], acc);
            acc = mfma_bf16(xh, Ll[gq], acc);
            acc = mfma_bf16(xm, Lm[gq], acc);
            acc = mfma_bf16(xl, Lh[gq], acc);
            pj[gq][rb] = acc;
          }
        }
      }
#pragma unroll
      for (int gq = 0; gq < 2; ++gq)
#pragma unroll
        for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const double v = pj[gq][rb][e];
            ssq[gq][rb][e] = fma(v, v, ssq[gq][rb][e]);
          }
    }
    // epilogue, one gene group at a time: its (gg, sum of squares, num) into
    // LDS from the four waves holding it, then the FP64 tests
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
      __syncthreads();  // the previous group's epilogue is done with fold and wpart
      if (c == (q >> 1)) {
        const int gq = q & 1;
#pragma unroll
        for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int t = (32 * r + 16 * rb + 4 * (lane >> 4) + e) * 16 + gl;
            fold[t] = (gq == 0 ? ag[0][rb][e] : ag[1][rb][e]);
            fold[kW8Tests + t] = (gq == 0 ? ssq[0][rb][e] : ssq[1][rb][e]);
            fold[2 * kW8Tests + t] = (gq == 0 ? an[0][rb][e] : an[1][rb][e]);
          }
      }
      __syncthreads();
      const W8Gene gi = lr_gene(la, sg * 4 + q);
      w8_epilogue<EQTL_MODEL_LMM>(a, gi, ri, tv, gstate + 16 * q, wpart, row0,
                                  [&](int t, double& gg, double& S, double& num) {
                                    gg = fold[t];
                                    S = gg - fold[kW8Tests + t];
                                    num = fold[2 * kW8Tests + t];
                                  });
    }
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const int gene = la.slot_gene[sg * 64 + threadIdx.x];
    if (gene >= 0) a.partial[static_cast<size_t>(split) * a.G_pad + gene] = gstate[threadIdx.x];
  }
}

}  // namespace eqtl_amd
