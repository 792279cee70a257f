// K7, trans LMM in the split form with the covariate projections and the
// weighted sum of squares interpolated in ln delta (scan_trans_lmm.cpp:11-40,
// 160-233 per test; per-gene statistics scan_trans.hpp:87-110).
//
// A test needs gg = g~^T D_k g~, num = (s_k o r_k)^T g~ and the projection
// W_k^T h = Z_k^T g~ onto the gene's whitened covariates, Z_k = D_k X~ Lc_k^-T
// (D_k = diag(1 / (delta_k lam + 1)), Lc_k Lc_k^T = X~^T D_k X~). Z_k and D_k
// are analytic in u = ln delta_k with the poles of 1 / (1 + lam e^u) at
// distance pi from the real axis, so on an interval of u of width 6.5 their
// Chebyshev interpolants at 32 nodes u_r are exact to ~1e-12 (relative, for
// Z^T g and gg; 64 / 96 / 128 nodes for widths 13 / 23 / 33). Genes are sorted
// by ln delta into super-groups of 64 (four gene groups of 16) and each
// super-group takes a segment of u with nodes u_r; then
//     Z_k^T g~ = sum_r l_r(u_k) T_r,   T_r = Z(u_r)^T g~ = (Q Z(u_r))^T g,
//     gg = d(u_k, lbar) sum_r l_r(u_k) U_r,  U_r = sum_c d(u_r, lam_c) / d(u_r, lbar) g~_c^2,
// where T (one more column block of the A x [Q | Q Z] product) and U (the
// squared rotated rows against the node weights) are shared by all genes of
// the segment (lbar, the mean eigenvalue, keeps the U_r of a segment within
// a small factor of each other; unnormalised they span up to 1e6 over a wide
// segment and their FP32 rounding would cancel). A test then contracts one
// block over the samples (num) and R node weights per quantity instead of
// p + 1 blocks over the samples.
//
// Three kernels per block:
//  * k_lr_num: num = (s_k o r_k)^T g~ over the samples. One workgroup = a
//    pair of super-groups (128 genes) x a strided set of 128-row SNP tiles,
//    eight waves: wave w = 4 c + r covers rows 32 r .. 32 r + 31 of a tile
//    against the four groups of super-group c of the pair. Per 32-sample
//    k-step the pair's gene tiles (2 x 8 KiB, s_k o r_k split bf16 hi / lo)
//    and the tile's rows (hi / lo planes written by the rotation, 16 KiB) are
//    copied global -> LDS into two-stage rings (four direct copies per wave
//    and k-step; two workgroups per CU), three bf16 products per MAC; num is
//    stored to the fold.
//    With 128 genes per workgroup a k-step moves 32 KiB for 192 MFMAs (one
//    super-group: 24 KiB for 96).
//  * k_wald_lr: the interpolation. One workgroup = one super-group x the
//    same tiles; per quantity (p projections, then gg) and 32-node chunk the
//    tile's T or U values (written by the rotation as the exact three-way
//    bf16 split hi + mid + lo of the fp32 value) are staged through an LDS
//    ring and contracted against the genes' weights l_r(u_k) split the same
//    way: six bf16 products (errors ~2^-24 relative), FP32 accumulation over
//    the nodes; a projection's square is added to the test's FP64 sum of
//    squares, gg is kept. gg and S = gg - sum of squares go to the fold.
//  * k_lr_tests: the FP64 tests of the eight-wave LMM kernel (refinement,
//    polish hand-off, hits, ACAT, best) over the fold.
#pragma once
#include "wald_lmm.hpp"

namespace eqtl_amd {

struct LrArgs {
  WaldArgs w;
  const __bf16* Ghi;         // rotated rows, bf16 hi plane (m_pad x n_pad, fragment-major: frag_index);
  int64_t lo_off;            //   lo plane at Ghi + lo_off
  const __bf16* Tp;          // m_pad x ldT node projections (T) and node sums of squares (U), fragment-major:
  int64_t tpl;               //   hi plane, then the mid plane at Tp + tpl and the lo plane at Tp + 2 tpl
  int64_t ldT;
  const int32_t* sg_col;     // per super-group: first T column of its segment
  const int32_t* sg_gcol;    // per super-group: first U column of its segment
  const int32_t* sg_R;       // per super-group: nodes of its segment (32, 64, 96 or 128)
  const int32_t* slot_gene;  // slot (ln-delta order) -> gene, -1 for padding slots
  const __bf16* lcoef;       // interpolation weights, see lr_coef_index
  const double* gscale;      // per slot d(delta_k, lbar): the scale of the interpolated gg
  const __bf16* gb;          // gene tiles, see lr_tile_index
  int32_t n_sg, p;
  int32_t sg_per_xcd;        // super-groups per XCD and round (the launch has n_splits splits of each)
  int32_t pair_per_xcd;      // k_lr_num: super-group pairs per XCD and round (n_splits splits of each)
  float* fold;               // per (tile, super-group): gg, S, num planes of the 4 x 2048 tests, see lr_fold_base
};

// First float of the fold block of (tile, super-group): three planes (gg, S,
// num), each 4 groups x 2048 tests in the eight-wave epilogue's test order
// (t = row * 16 + gene).
__host__ __device__ inline size_t lr_fold_base(int tile, int sg, int n_sg) {
  return (static_cast<size_t>(tile) * n_sg + sg) * 3 * 4 * kW8Tests;
}

// k_lr_num: both copy rings (the pair's gene tiles, the row planes) hold
// kLrStages k-steps and are filled kLrStages - 1 k-steps ahead. Two stages
// (64 KiB) let two workgroups share a CU (at most 128 registers per lane),
// so one workgroup's k-step barrier is covered by the other's MFMAs (four
// stages and one workgroup per CU: 13% slower).
constexpr int kLrStages = 2;
constexpr size_t kLrTileBytes = 8 * 1024;  // one k-step of a super-group's gene tile
constexpr size_t kLrPairTileBytes = 2 * kLrTileBytes;
constexpr size_t kLrNumLdsBytes = kLrStages * (kLrPairTileBytes + kAStageBytes);
// k_wald_lr: the interpolation's ring of kMixStages steps of 24 KiB (128 rows
// x 32 nodes x three bf16 planes); after it, per group and test gg and S
// (fp32), copied out to the fold.
constexpr int kMixStages = 4;
constexpr size_t kMixStageBytes = 24 * 1024;
constexpr size_t kLrMixFoldBytes = 2 * 4 * kW8Tests * sizeof(float);
constexpr size_t kLrLdsBytes =
    kMixStages * kMixStageBytes > kLrMixFoldBytes ? kMixStages * kMixStageBytes : kLrMixFoldBytes;
// s_waitcnt immediates (gfx9 encoding): vmcnt <= v, expcnt at its maximum, lgkmcnt 0
constexpr int lr_wait_vm(int v) { return (v & 0xF) | ((v >> 4) << 14) | (0x7 << 4); }
// k_wald_lr's ring: waits until at most the three copies of each of the
// `ahead` younger steps (0 <= ahead <= kMixStages - 2) are in flight
__device__ __attribute__((always_inline)) inline void lr_wait3(int ahead) {
  static_assert(kMixStages <= 6, "lr_wait3 covers up to four younger steps");
  if (ahead >= 4) __builtin_amdgcn_s_waitcnt(lr_wait_vm(12));
  else if (ahead == 3) __builtin_amdgcn_s_waitcnt(lr_wait_vm(9));
  else if (ahead == 2) __builtin_amdgcn_s_waitcnt(lr_wait_vm(6));
  else if (ahead == 1) __builtin_amdgcn_s_waitcnt(lr_wait_vm(3));
  else __builtin_amdgcn_s_waitcnt(0);
}
static_assert(kLrLdsBytes + 64 <= kLdsCap, "LR interpolation LDS layout");
static_assert(kLrNumLdsBytes + 64 <= kLdsCap, "LR num LDS layout");
// the tests kernel: RowInfo, the t tail's padded LDS copy (tail_to_lds), the
// 64 genes' states per wave, the genes' data
constexpr size_t kLrTestsLdsBytes = sizeof(RowInfo) + sizeof(double) * (kTailLdsHead + kTailLdsMid) +
                                    sizeof(GeneAcc) * 64 * kW8Waves + sizeof(W8Gene) * 64;
static_assert(kLrTestsLdsBytes + 64 <= kLdsCap, "LR tests LDS layout");

// Element (slot, sample i) of the gene tiles (s o r): per super-group sg and
// k-step s one 8 KiB tile, the hi plane (4 groups x 64 lanes x 8 values) then
// the lo plane; group q = (slot >> 4) & 3, lane = (slot & 15) + 16 * ((i & 31)
// >> 3), value i & 7.
__host__ __device__ inline size_t lr_tile_index(int slot, int i, int nks, int plane) {
  const int sg = slot >> 6, q = (slot >> 4) & 3, col = slot & 15;
  const int s = i >> 5, lane = col + 16 * ((i & 31) >> 3);
  return ((static_cast<size_t>(sg) * nks + s) * 2 + plane) * 2048 + q * 512 + lane * 8 + (i & 7);
}

// Interpolation weight r of slot: per group of 16 slots and 32-node chunk a
// 3 x 512 block (hi, mid, lo planes), lane (slot & 15) + 16 * ((r & 31) >> 3),
// value r & 7 — the B fragment of v_mfma_f32_16x16x32_bf16 (genes as columns).
__host__ __device__ inline size_t lr_coef_index(int slot, int r, int plane) {
  const int lane = (slot & 15) + 16 * ((r & 31) >> 3);
  return (static_cast<size_t>(slot >> 4) * 4 + (r >> 5)) * 1536 + plane * 512 + lane * 8 + (r & 7);
}

// One 32-sample k-step of k_lr_num's wave (c, r): its row pair's hi / lo
// fragments from the current stage (a_rd) against the four gene groups of
// super-group c of the pair (b_cur), 24 MFMAs; it issues its four copies of
// the k-step kLrStages - 1 on (c = 0: gene-tile chunks 4 r .. 4 r + 3 of the
// pair from b_src into b_nxt; c = 1: the row pair's planes into a_wr). The LDS
// pointers are __restrict__ so the compiler does not drain the copies in
// flight before the reads.
__device__ __attribute__((always_inline)) inline void lr_kstep(
    f32x4 (&an)[4][kRB], const unsigned char* __restrict__ a_rd, unsigned char* __restrict__ a_wr,
    const __bf16* const (&a_src)[kRB], int64_t lo_off, const unsigned char* __restrict__ b_cur,
    unsigned char* __restrict__ b_nxt, const unsigned char* b_src, int c, int r, int lane) {
  bf16x8 ah[kRB], al[kRB];
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb) {
    ah[rb] = *reinterpret_cast<const bf16x8*>(a_rd + (rb * 2 + 0) * 1024 + lane * 16);
    al[rb] = *reinterpret_cast<const bf16x8*>(a_rd + (rb * 2 + 1) * 1024 + lane * 16);
  }
  if (c == 0) {
    if (b_nxt) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int ch = 4 * r + k;
        lds_direct16(b_src + (ch & 7) * 1024 + lane * 16, b_nxt + ch * 1024);
      }
    }
  } else if (a_wr) {
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) {
      lds_direct16(a_src[rb], a_wr + rb * 2048);
      lds_direct16(a_src[rb] + lo_off, a_wr + rb * 2048 + 1024);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  const unsigned char* bt = b_cur + c * kLrTileBytes;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bf16x8 vh = *reinterpret_cast<const bf16x8*>(bt + q * 1024 + lane * 16);
    const bf16x8 vl = *reinterpret_cast<const bf16x8*>(bt + 4096 + q * 1024 + lane * 16);
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) an[q][rb] = mfma_bf16(ah[rb], vh, an[q][rb]);
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) an[q][rb] = mfma_bf16(ah[rb], vl, an[q][rb]);
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) an[q][rb] = mfma_bf16(al[rb], vh, an[q][rb]);
  }
}

// Gene tiles of the low-rank scan, one workgroup per gene k at slot
// slot_of[k]: the whitened null residual s o r (u64, as written by
// k_lmm_basis), fp32 then split bf16 hi / lo (samples past n stay zero).
__global__ __launch_bounds__(256) void k_lr_gene_tiles(const double* u64, const int32_t* slot_of, int n, int nks,
                                                       int G, __bf16* gb) {
  const int k = blockIdx.x;
  if (k >= G) return;
  const int slot = slot_of[k];
  for (int i = threadIdx.x; i < n; i += 256) {
    const float f = static_cast<float>(u64[static_cast<size_t>(k) * n + i]);
    const __bf16 h = to_bf16(f);
    gb[lr_tile_index(slot, i, nks, 0)] = h;
    gb[lr_tile_index(slot, i, nks, 1)] = to_bf16(f - static_cast<float>(h));
  }
}

// fp64 columns (leading dimension np, ncols of them) -> fp32 -> bf16 hi / lo
// in the rotation kernel's fragment-major B order (rot_b_index), zero for the
// padding columns up to rot_b_coords(ncols)
__global__ void k_split_rot_b(const double* x, int np, int ncols, __bf16* hi, __bf16* lo) {
  const int64_t count = rot_b_coords(ncols) * np;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < count;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t col = t / np, i = t - col * np;
    const float f = col < ncols ? static_cast<float>(x[t]) : 0.f;
    const __bf16 h = to_bf16(f);
    const int64_t o = rot_b_index(col, i, np / 32);
    hi[o] = h;
    lo[o] = to_bf16(f - static_cast<float>(h));
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

// The genes' interpolation weights for one 32-node chunk, both groups of a wave.
struct LrW {
  bf16x8 h[2], m[2], l[2];
};
__device__ inline LrW lr_weights(const __bf16* lb) {
  LrW w;
#pragma unroll
  for (int gq = 0; gq < 2; ++gq) {
    const __bf16* l = lb + static_cast<size_t>(gq) * 4 * 1536;
    w.h[gq] = load8_bf16(l);
    w.m[gq] = load8_bf16(l + 512);
    w.l[gq] = load8_bf16(l + 1024);
  }
  return w;
}

// One quantity's values for the wave's two row blocks (hi / mid / lo planes).
struct LrX {
  bf16x8 h[kRB], m[kRB], l[kRB];
};

// Accumulates one quantity's interpolation for both groups of a wave:
// pj[gq][rb] += sum over the chunk's 32 nodes of l_r x value_r, six products.
__device__ __attribute__((always_inline)) inline void lr_mix(f32x4 (&pj)[2][kRB], const LrX& x, const LrW& w) {
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb) {
    const bf16x8& xh = x.h[rb];
    const bf16x8& xm = x.m[rb];
    const bf16x8& xl = x.l[rb];
    const bf16x8* Lh = w.h;
    const bf16x8* Lm = w.m;
    const bf16x8* Ll = w.l;
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

// num of every test of a pair of super-groups: the contraction over the
// samples, stored to the fold's num plane (fp32).
__global__ __launch_bounds__(kW8Threads) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_lr_num(LrArgs la) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const WaldArgs& a = la.w;
  // XCD-aware: workgroup wg runs on XCD wg % 8. Per round each XCD takes
  // pair_per_xcd consecutive super-group pairs x n_splits row splits; the
  // workgroups of one split sweep the same SNP tiles in the same order (and
  // those of one pair read the same gene tiles), so both are served from the
  // XCD's L2 instead of each pair streaming the whole block.
  const int wg = blockIdx.x, xcd = wg & 7, jx = wg >> 3;
  const int per_xcd = la.pair_per_xcd * a.n_splits;
  const int l = jx % per_xcd;
  const int pr = ((jx / per_xcd) * 8 + xcd) * la.pair_per_xcd + l / a.n_splits;
  const int split = l % a.n_splits;
  if (2 * pr >= la.n_sg) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = wave >> 2, r = wave & 3;  // super-group 2 pr + c of the pair, row quarter r
  const int sg = 2 * pr + c;
  const bool sg_ok = sg < la.n_sg;  // an odd count leaves the last pair's second half empty
  const int gl = lane & 15;
  const int nks = a.n_pad >> 5;

  unsigned char* const bbuf = smem;
  unsigned char* const abuf = smem + kLrStages * kLrPairTileBytes;
  // gene-tile chunks this wave copies (c = 0): 4 r .. 4 r + 3 of the pair's
  // 16, all of super-group 2 pr + (r >> 1) (the first one again for an empty
  // second half: its products are not stored)
  const int sg_src = min(2 * pr + (r >> 1), la.n_sg - 1);
  const unsigned char* gsrc =
      reinterpret_cast<const unsigned char*>(la.gb) + static_cast<size_t>(sg_src) * nks * kLrTileBytes;
  const int pair_off = r * 4096;
  const int n_tiles = (a.n_tiles + 1) >> 1;

  for (int tile = split; tile < n_tiles; tile += a.n_splits) {
    const int64_t row0 = static_cast<int64_t>(tile) * kTileRows;
    __syncthreads();  // every wave is done with the previous tile's stages
    // this wave's two 16-row blocks in the fragment-major planes (one 1 KiB
    // block per k-step; the lo plane lo_off further on)
    const __bf16* a_row[kRB];
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb)
      a_row[rb] = la.Ghi + ((row0 >> 4) + 2 * r + rb) * static_cast<int64_t>(nks) * 512 + lane * 8;
    f32x4 an[4][kRB];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) an[q][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
    // prologue: the first kLrStages - 1 k-steps, four copies per wave each
    for (int s = 0; s < kLrStages - 1 && s < nks; ++s) {
      if (c == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int ch = 4 * r + k;
          lds_direct16(gsrc + static_cast<size_t>(s) * kLrTileBytes + (ch & 7) * 1024 + lane * 16,
                       bbuf + s * kLrPairTileBytes + ch * 1024);
        }
      } else {
        unsigned char* dst = abuf + s * kAStageBytes + pair_off;
#pragma unroll
        for (int rb = 0; rb < kRB; ++rb) {
          lds_direct16(a_row[rb] + s * 512, dst + rb * 2048);
          lds_direct16(a_row[rb] + s * 512 + la.lo_off, dst + rb * 2048 + 1024);
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int s = 0; s < nks; ++s) {
      const int sf = s + kLrStages - 1;  // k-step whose operands are fetched now
      const __bf16* src[kRB];
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) src[rb] = a_row[rb] + sf * 512;
      const bool fetch = sf < nks;
      unsigned char* const a_wr = fetch ? abuf + (sf % kLrStages) * kAStageBytes + pair_off : nullptr;
      const unsigned char* const a_rd = abuf + (s % kLrStages) * kAStageBytes + pair_off;
      unsigned char* const b_nxt = fetch ? bbuf + (sf % kLrStages) * kLrPairTileBytes : nullptr;
      lr_kstep(an, a_rd, a_wr, src, la.lo_off, bbuf + (s % kLrStages) * kLrPairTileBytes, b_nxt,
               gsrc + static_cast<size_t>(sf) * kLrTileBytes, c, r, lane);
      // k-step s + 1's operands must have landed: of this wave's copies (four
      // per k-step) those of the younger k-steps may stay in flight
      const int ahead = min(nks - 1, s + kLrStages - 1) - (s + 1);
      if (ahead >= 2) __builtin_amdgcn_s_waitcnt(lr_wait_vm(8));
      else if (ahead == 1) __builtin_amdgcn_s_waitcnt(lr_wait_vm(4));
      else __builtin_amdgcn_s_waitcnt(0);
      lds_barrier();
    }
    // num plane of (tile, sg): test (row, gene) of group q at q * 2048 + row * 16 + gene;
    // streaming stores (read back only by k_lr_tests)
    if (sg_ok) {
      float* fnum = la.fold + lr_fold_base(tile, sg, la.n_sg) + 2 * 4 * kW8Tests;
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            __builtin_nontemporal_store(an[q][rb][e],
                                        fnum + q * kW8Tests + (32 * r + 16 * rb + 4 * (lane >> 4) + e) * 16 + gl);
    }
  }
}

// gg and S of every test of one super-group: the interpolation over the
// segment's nodes, stored to the fold's gg and S planes.
__global__ __launch_bounds__(kW8Threads) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_wald_lr(LrArgs la) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const WaldArgs& a = la.w;
  // XCD-aware: workgroup wg runs on XCD wg % 8. Per round each XCD takes
  // sg_per_xcd consecutive super-groups x n_splits row splits; the workgroups
  // of one split sweep the same SNP tiles in the same order (and those of one
  // super-group read the same gene tiles), so both are served from the XCD's
  // L2 instead of each super-group streaming the whole block.
  const int wg = blockIdx.x, xcd = wg & 7, jx = wg >> 3;
  const int per_xcd = la.sg_per_xcd * a.n_splits;
  const int l = jx % per_xcd;
  const int sg = ((jx / per_xcd) * 8 + xcd) * la.sg_per_xcd + l / a.n_splits;
  const int split = l % a.n_splits;
  if (sg >= la.n_sg) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = wave >> 2, r = wave & 3;
  const int gl = lane & 15;

  // the per-gene scale of gg for this wave's two groups
  const double gsc0 = la.gscale[(sg * 4 + 2 * c) * 16 + gl];
  const double gsc1 = la.gscale[(sg * 4 + 2 * c + 1) * 16 + gl];
  const int segR = la.sg_R[sg], segc = la.sg_col[sg], segg = la.sg_gcol[sg], nqr = segR >> 5;
  const int n_tiles = (a.n_tiles + 1) >> 1;
  const __bf16* lbase = la.lcoef + static_cast<size_t>((sg * 4 + 2 * c) * 4) * 1536 + lane * 8;

  for (int tile = split; tile < n_tiles; tile += a.n_splits) {
    const int64_t row0 = static_cast<int64_t>(tile) * kTileRows;
    __syncthreads();  // the previous tile's fold has been copied out
    // interpolation: quantities 0 .. p-1 (projections, squared and summed),
    // then p (gg); each quantity's values are read one quantity ahead
    double ssq[2][kRB][4];
    float ggv[2][kRB][4];
#pragma unroll
    for (int gq = 0; gq < 2; ++gq)
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          ssq[gq][rb][e] = 0.0;
          ggv[gq][rb][e] = 0.f;
        }
    // column of chunk qr of quantity ca (ca == p: the node sums of squares)
    auto qcol = [&](int ca, int qr) { return static_cast<int64_t>((ca < la.p ? segc + ca * segR : segg) + 32 * qr); };
    f32x4 pj[2][kRB];
#pragma unroll
    for (int gq = 0; gq < 2; ++gq)
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) pj[gq][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
    auto finish = [&](bool is_gg) {
#pragma unroll
      for (int gq = 0; gq < 2; ++gq)
#pragma unroll
        for (int rb = 0; rb < kRB; ++rb) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float v = pj[gq][rb][e];
            if (is_gg) ggv[gq][rb][e] = v;
            else ssq[gq][rb][e] = fma(static_cast<double>(v), static_cast<double>(v), ssq[gq][rb][e]);
          }
          pj[gq][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    };
    // Step st = (quantity ca, chunk qr) stages the tile's 128 rows x 32 node
    // values in LDS (ring slot st % kMixStages): chunk ch = (rr * 2 + rb) * 3
    // + plane of 1 KiB holds rows 32 rr + 16 rb + (l & 15), nodes 8 (l >> 4)
    // .. + 7 of that plane in lane l's 16 bytes. Wave w copies chunks 3w ..
    // 3w + 2; the two waves of a row quarter read the same chunks.
    const int nsteps = (la.p + 1) * nqr;
    const int ch0 = 3 * wave;
    auto issue = [&](int st) {
      const int ca = st / nqr, qr = st - ca * nqr;
      const int64_t chunk = qcol(ca, qr) >> 5;
      unsigned char* dst = smem + (st % kMixStages) * kMixStageBytes;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int ch = ch0 + k, rr = ch / 6, rb = (ch / 3) & 1, plane = ch % 3;
        const __bf16* src =
            la.Tp + plane * la.tpl + (((row0 >> 4) + 2 * rr + rb) * (la.ldT >> 5) + chunk) * 512 + lane * 8;
        lds_direct16(src, dst + ch * 1024);
      }
    };
    auto values = [&](int st) {
      const unsigned char* slot = smem + (st % kMixStages) * kMixStageBytes + lane * 16;
      LrX x;
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) {
        const int ch = (r * 2 + rb) * 3;
        x.h[rb] = *reinterpret_cast<const bf16x8*>(slot + (ch + 0) * 1024);
        x.m[rb] = *reinterpret_cast<const bf16x8*>(slot + (ch + 1) * 1024);
        x.l[rb] = *reinterpret_cast<const bf16x8*>(slot + (ch + 2) * 1024);
      }
      return x;
    };
    if (nqr == 1) {
      // one chunk: the weights are the same for every quantity (loaded before
      // the copies, so the waits below cover them too)
      const LrW w = lr_weights(lbase);
      __builtin_amdgcn_sched_barrier(0);
      for (int st = 0; st < kMixStages - 1 && st < nsteps; ++st) issue(st);
#pragma unroll 1
      for (int st = 0; st < nsteps; ++st) {
        // step st's chunks have landed: the copies of the up to two later
        // steps already issued may stay in flight
        lr_wait3(min(nsteps - 1, st + kMixStages - 2) - st);
        lds_barrier();  // every wave's chunks of step st are in; step st - 1's slot is free
        if (st + kMixStages - 1 < nsteps) issue(st + kMixStages - 1);
        __builtin_amdgcn_sched_barrier(0);
        lr_mix(pj, values(st), w);
        finish(st == la.p);
      }
    } else {
      // per chunk weights, loaded one step ahead and issued before that
      // step's copies, so waiting for them also retires the older copies
      LrW w = lr_weights(lbase);
      __builtin_amdgcn_sched_barrier(0);
      for (int st = 0; st < kMixStages - 1 && st < nsteps; ++st) issue(st);
#pragma unroll 1
      for (int st = 0; st < nsteps; ++st) {
        const int ca = st / nqr, qr = st - ca * nqr;
        if (st == 0) {
          lr_wait3(min(nsteps - 1, kMixStages - 2));
        } else if (st + kMixStages - 2 < nsteps) {
          __builtin_amdgcn_s_waitcnt(lr_wait_vm(3));
        } else {
          __builtin_amdgcn_s_waitcnt(0);
        }
        lds_barrier();
        LrW wn;
        if (st + 1 < nsteps) {
          const int qr1 = (st + 1) % nqr;
          wn = lr_weights(lbase + static_cast<size_t>(qr1) * 1536);
        }
        __builtin_amdgcn_sched_barrier(0);
        if (st + kMixStages - 1 < nsteps) issue(st + kMixStages - 1);
        __builtin_amdgcn_sched_barrier(0);
        lr_mix(pj, values(st), w);
        if (qr == nqr - 1) finish(ca == la.p);
        w = wn;
      }
    }
    // fold: gg = d(delta, lbar) x the interpolated sum, S = gg - ||W^T h||^2
    // (FP64); into LDS planes in the tests' order, then copied out whole
    float* fgg = reinterpret_cast<float*>(smem);
    float* fS = fgg + 4 * kW8Tests;
    __syncthreads();  // the staging buffers are free (every wave is past the last step)
#pragma unroll
    for (int gq = 0; gq < 2; ++gq) {
      const int q = 2 * c + gq;
      const double gsc = gq ? gsc1 : gsc0;
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int t = q * kW8Tests + (32 * r + 16 * rb + 4 * (lane >> 4) + e) * 16 + gl;
          const double gg = gsc * ggv[gq][rb][e];
          fgg[t] = static_cast<float>(gg);
          fS[t] = static_cast<float>(gg - ssq[gq][rb][e]);
        }
    }
    __syncthreads();
    {
      const float4* src = reinterpret_cast<const float4*>(smem);
      float4* dst = reinterpret_cast<float4*>(la.fold + lr_fold_base(tile, sg, la.n_sg));
#pragma unroll 4
      // streaming stores: the fold is read back only by k_lr_tests, so it
      // should not displace the rows and node values this kernel re-reads
      for (int i = threadIdx.x; i < static_cast<int>(kLrMixFoldBytes / 16); i += kW8Threads) {
        const float4 v = src[i];
        __builtin_nontemporal_store(v.x, &dst[i].x);
        __builtin_nontemporal_store(v.y, &dst[i].y);
        __builtin_nontemporal_store(v.z, &dst[i].z);
        __builtin_nontemporal_store(v.w, &dst[i].w);
      }
    }
  }
}

// The FP64 tests of the interpolated form over the fold buffer: one
// workgroup per (super-group, split), laid out as in k_wald_lr, takes the
// split's tiles; per tile and gene group the tests of the eight-wave epilogue
// (refinement, polish hand-off, hits, ACAT, best) and, at the end, the 64
// genes' partial states.
__global__ __launch_bounds__(kW8Threads) void k_lr_tests(LrArgs la) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const WaldArgs& a = la.w;
  // the workgroup layout of k_wald_lr: the workgroups of one XCD that share a
  // split test the same SNP rows (whose FP64 rows the refinement reads), those
  // that share a super-group the same genes (their FP64 residual vectors)
  const int wg = blockIdx.x, xcd = wg & 7, jx = wg >> 3;
  const int per_xcd = la.sg_per_xcd * a.n_splits;
  const int l = jx % per_xcd;
  const int sg = ((jx / per_xcd) * 8 + xcd) * la.sg_per_xcd + l / a.n_splits;
  const int split = l % a.n_splits;
  if (sg >= la.n_sg) return;
  RowInfo& ri = *reinterpret_cast<RowInfo*>(smem);
  double* thead = reinterpret_cast<double*>(smem + sizeof(RowInfo));
  double* tmid = thead + kTailLdsHead;
  // per wave the running states of the super-group's 64 genes: a wave adds
  // its own totals without waiting for the others; merged at the end
  GeneAcc* wstate = reinterpret_cast<GeneAcc*>(tmid + kTailLdsMid);
  W8Gene* ginfo = reinterpret_cast<W8Gene*>(wstate + 64 * kW8Waves);  // the super-group's 64 genes
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  tail_to_lds(*a.tail, thead, tmid, threadIdx.x, kW8Threads);
  acc_init(wstate[threadIdx.x]);
  if (threadIdx.x < 64) ginfo[threadIdx.x] = lr_gene(la, sg * 4 + (threadIdx.x >> 4));
  const TailView tv{thead, tmid, a.tail->coef_b, kTailLdsStride, kTailLdsStride};
  const int n_tiles = (a.n_tiles + 1) >> 1;
  // the SNP data of the next tile, loaded while this tile's tests run
  uint8_t r_keep = 0;
  double r_af = 0.0;
  int32_t r_chrom = -1;
  int64_t r_pos = 0;
  auto load_rows = [&](int tl) {
    if (threadIdx.x < kTileRows) {
      const int64_t j = static_cast<int64_t>(tl) * kTileRows + threadIdx.x;
      const bool in = j < a.m, ann = in && a.snp_chrom;
      r_keep = in ? a.keep[j] : 0;
      r_af = in ? a.af[j] : 0.0;
      r_chrom = ann ? a.snp_chrom[a.snp0 + j] : -1;
      r_pos = ann ? a.snp_pos[a.snp0 + j] : 0;
    }
  };
  if (split < n_tiles) load_rows(split);
  // Thread x takes tests x + 512 i (i = 0..3) of each group, in that order.
  // Their gg, S, num are loaded one group ahead (the next group's loads are
  // in flight while this group's tests run).
  static_assert(kW8Tests == 4 * kW8Threads, "four tests per thread and group");
  float c0[4], c1[4], c2[4], n0[4], n1[4], n2[4];
  auto load = [&](int tl, int q, float (&g0)[4], float (&g1)[4], float (&g2)[4]) {
    const float* fb = la.fold + lr_fold_base(tl, sg, la.n_sg) + q * kW8Tests + threadIdx.x;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      g0[i] = fb[kW8Threads * i];
      g1[i] = fb[4 * kW8Tests + kW8Threads * i];
      g2[i] = fb[8 * kW8Tests + kW8Threads * i];
    }
  };
  if (split < n_tiles) load(split, 0, c0, c1, c2);
  for (int tile = split; tile < n_tiles; tile += a.n_splits) {
    const int64_t row0 = static_cast<int64_t>(tile) * kTileRows;
    __syncthreads();  // the previous tile's tests are done with ri
    if (threadIdx.x < kTileRows) {
      const int t = threadIdx.x;
      ri.keep[t] = r_keep;
      ri.af[t] = r_af;
      ri.gtg[t] = 0.0;
      ri.chrom[t] = r_chrom;
      ri.pos[t] = r_pos;
    }
    if (tile + a.n_splits < n_tiles) load_rows(tile + a.n_splits);
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
      const int nt = q < 3 ? tile : tile + a.n_splits, nq = q < 3 ? q + 1 : 0;
      if (nt < n_tiles) load(nt, nq, n0, n1, n2);
      if (q == 0) __syncthreads();  // ri is in
      W8Gene gi = ginfo[q * 16 + (threadIdx.x & 15)];
      gi.gl = threadIdx.x & 15;
      const GeneAcc st = w8_tests<EQTL_MODEL_LMM>(a, gi, ri, tv, row0,
                                  [&](int, double& gg, double& S, double& num) {
                                    gg = c0[0];
                                    S = c1[0];
                                    num = c2[0];
#pragma unroll
                                    for (int i = 0; i < 3; ++i) {
                                      c0[i] = c0[i + 1];
                                      c1[i] = c1[i + 1];
                                      c2[i] = c2[i + 1];
                                    }
                                  });
      if (lane < 16) acc_merge(wstate[wave * 64 + q * 16 + lane], st);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        c0[i] = n0[i];
        c1[i] = n1[i];
        c2[i] = n2[i];
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    GeneAcc g = wstate[threadIdx.x];
    for (int w = 1; w < kW8Waves; ++w) acc_merge(g, wstate[w * 64 + threadIdx.x]);
    const int gene = la.slot_gene[sg * 64 + threadIdx.x];
    if (gene >= 0) a.partial[static_cast<size_t>(split) * a.G_pad + gene] = g;
  }
}

}  // namespace eqtl_amd
