// K7 for the LMM: fused Wald contraction + FP64 epilogue with eight waves per
// workgroup, two per SIMD (scan_trans_lmm.cpp:11-40,160-233 per test; per-gene
// statistics scan_trans.hpp:87-110).
//
// A workgroup owns one group of 16 genes and a strided set of 128-row SNP
// tiles, as the four-wave kernel does. Wave w = 4 c + r covers rows 32 r ..
// 32 r + 31 of a tile (two 16-row blocks) against column half c of the
// group's NQ quantity blocks: half 0 holds blocks [0, NQ0), half 1 blocks
// [NQ0, NQ). The two waves sharing a SIMD (r, c = 0 and 1) together hold one
// row pair against all blocks, so each wave keeps half the accumulators of
// the four-wave kernel and the SIMD always has a second wave to issue from
// while one waits on LDS, a barrier or its FP64 epilogue.
//
// Per 32-sample k-step the group's gene tile (NQ x 2 KiB) is copied global ->
// LDS (double-buffered) and each row pair into a ring of row stages, all by
// the four half-1 waves; the k-step ends with one workgroup barrier. After the last k-step each wave folds its accumulators to
// partial (gg, sum of squares, num) per test into LDS; the epilogue then runs
// on all 512 threads (4 tests each), and per-gene states are merged through
// lane shuffles and LDS into a running state per gene kept in LDS.
#pragma once
#include "scan_kernels.hpp"

namespace eqtl_amd {

constexpr int kW8Waves = 8;
constexpr int kW8Threads = kW8Waves * 64;
constexpr int kW8Tests = kTileRows * 16;                                  // tests per tile
constexpr size_t kW8FoldBytes = sizeof(double) * 6 * kW8Tests;            // 2 halves x (gg, ss, num)
constexpr size_t kW8MergeBytes = sizeof(GeneAcc) * kW8Waves * 16;         // per-wave gene partials
constexpr long kW8TailBytes = static_cast<long>(sizeof(RowInfo)) + static_cast<long>(sizeof(double)) * kTailHead +
                              static_cast<long>(sizeof(GeneAcc)) * 16;

__host__ __device__ constexpr int w8_a_stages(int nq) {
  return (2L * nq * 2048 + 3L * kAStageBytes + kW8TailBytes <= kLdsCap)   ? 3
         : (2L * nq * 2048 + 2L * kAStageBytes + kW8TailBytes <= kLdsCap) ? 2
                                                                           : 0;
}
__host__ __device__ constexpr size_t w8_region0_bytes(int nq) {
  const size_t b = 2 * static_cast<size_t>(nq) * 2048 + static_cast<size_t>(w8_a_stages(nq)) * kAStageBytes;
  const size_t f = kW8FoldBytes + kW8MergeBytes;
  return b > f ? b : f;
}
__host__ __device__ constexpr size_t w8_lds_bytes(int nq) {
  return w8_region0_bytes(nq) + static_cast<size_t>(kW8TailBytes);
}
static_assert(w8_lds_bytes(34) <= kLdsCap && w8_a_stages(23) == 3 && w8_a_stages(26) == 2, "W8 LDS layout");

// One k-step of one wave: this wave's rows (a_rd, its row pair in the current
// stage) and the gene tile (b_cur) are read from LDS, the copies of later
// operands are started, and the MFMAs cover blocks [B0, B0 + NQH). Block 0
// (the weights d, against the squared rows) is split by row block so that
// both halves square and split the same number of rows: half 0 runs it for
// row block 0 (in acc[0][0]; acc[1][0] stays zero), half 1 for row block 1
// (in acc0).
// The four half-1 waves issue all copies: their share of the next gene tile
// (b_nxt, chunks r, r + 4, ...), then their row pair of a later stage (a_wr,
// from a_src[rb]); the half-0 waves, which also compute the squared rows,
// issue none. (Measured: this beats splitting the copy kinds between the
// halves or spreading them over the k-step.)
// The LDS pointers are __restrict__ (disjoint regions), so the compiler does
// not drain the copies in flight before this k-step's LDS reads.
template <int PREC, int NQ, int NQH, int B0, int kAS>
__device__ __attribute__((always_inline)) inline void w8_kstep(
    f32x4 (&acc)[kRB][NQH], f32x4& acc0, f32x8 (&an)[kRB], const unsigned char* __restrict__ a_rd,
    unsigned char* __restrict__ a_wr, const float* const (&a_src)[kRB], bool a_reg_next,
    const unsigned char* __restrict__ b_cur, unsigned char* __restrict__ b_nxt, const unsigned char* b_src, int r,
    int lane) {
  constexpr int chunks = (NQ * 2048) >> 10;
  f32x8 ac[kRB];
  if constexpr (kAS > 0) {
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) {
      const float4 lo = *reinterpret_cast<const float4*>(a_rd + (rb * 2 + 0) * 1024 + lane * 16);
      const float4 hi = *reinterpret_cast<const float4*>(a_rd + (rb * 2 + 1) * 1024 + lane * 16);
      ac[rb].v[0] = lo.x; ac[rb].v[1] = lo.y; ac[rb].v[2] = lo.z; ac[rb].v[3] = lo.w;
      ac[rb].v[4] = hi.x; ac[rb].v[5] = hi.y; ac[rb].v[6] = hi.z; ac[rb].v[7] = hi.w;
    }
  } else {
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) ac[rb] = an[rb];
  }
  if constexpr (B0 != 0) {
    if (b_nxt)
      for (int c = r; c < chunks; c += 4) lds_direct16(b_src + c * 1024 + lane * 16, b_nxt + c * 1024);
  }
  if constexpr (B0 != 0 && kAS > 0) {
    if (a_wr) {
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) {
        lds_direct16(a_src[rb], a_wr + rb * 2048);
        lds_direct16(a_src[rb] + 4, a_wr + rb * 2048 + 1024);
      }
    }
  }
  if constexpr (kAS == 0) {
    if (a_reg_next) {
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) an[rb] = load8(a_src[rb]);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (PREC == EQTL_PREC_FP32) {
    constexpr int rs = B0 == 0 ? 0 : 1;  // the row block whose squares this half contracts
    f32x8 sq;
#pragma unroll
    for (int t = 0; t < 8; ++t) sq.v[t] = ac[rs].v[t] * ac[rs].v[t];
    if constexpr (B0 != 0) acc0 = mfma_f32x8(sq, load8(reinterpret_cast<const float*>(b_cur + lane * 32)), acc0);
    f32x8 bn = load8(reinterpret_cast<const float*>(b_cur + (B0 * 64 + lane) * 32));
#pragma unroll
    for (int b = 0; b < NQH; ++b) {
      const f32x8 bv = bn;
      if (b + 1 < NQH) bn = load8(reinterpret_cast<const float*>(b_cur + ((B0 + b + 1) * 64 + lane) * 32));
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) {
        if (B0 == 0 && b == 0) {
          if (rb == 0) acc[0][0] = mfma_f32x8(sq, bv, acc[0][0]);
        } else {
          acc[rb][b] = mfma_f32x8(ac[rb], bv, acc[rb][b]);
        }
      }
    }
  } else {
    // gene fragments in groups of kCB blocks, one group ahead; the explicit
    // lgkm wait sits before the next group's reads go out (the direct-to-LDS
    // copies make the compiler's own lgkm waits all-or-nothing)
    constexpr int kCB = 3;
    constexpr int kGroups = (NQH + kCB - 1) / kCB;
    bf16x8 fh[2][kCB], fl[2][kCB];
    auto load_group = [&](int g, int slot) {
#pragma unroll
      for (int c = 0; c < kCB; ++c) {
        const int b = g * kCB + c;
        if (b < NQH) {
          fh[slot][c] = *reinterpret_cast<const bf16x8*>(b_cur + ((B0 + b) * 64 + lane) * 16);
          fl[slot][c] = *reinterpret_cast<const bf16x8*>(b_cur + NQ * 1024 + ((B0 + b) * 64 + lane) * 16);
        }
      }
    };
    load_group(0, 0);
    bf16x8 f0h, f0l;  // half 1: block 0's fragments
    if constexpr (B0 != 0) {
      f0h = *reinterpret_cast<const bf16x8*>(b_cur + lane * 16);
      f0l = *reinterpret_cast<const bf16x8*>(b_cur + NQ * 1024 + lane * 16);
    }
    constexpr int rs = B0 == 0 ? 0 : 1;  // the row block whose squares this half contracts
    bf16x8 ah[kRB], al[kRB], sh, sl;
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) split_bf16(ac[rb], ah[rb], al[rb]);
    {
      f32x8 sv;
#pragma unroll
      for (int t = 0; t < 8; ++t) sv.v[t] = ac[rs].v[t] * ac[rs].v[t];
      split_bf16(sv, sh, sl);
    }
    if constexpr (B0 != 0) {
      __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
      acc0 = mfma_bf16(sh, f0h, acc0);
      acc0 = mfma_bf16(sh, f0l, acc0);
      acc0 = mfma_bf16(sl, f0h, acc0);
    }
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
      __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
      __builtin_amdgcn_sched_barrier(0);
      if (g + 1 < kGroups) load_group(g + 1, (g + 1) & 1);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int term = 0; term < 3; ++term)
#pragma unroll
        for (int c = 0; c < kCB; ++c) {
          const int b = g * kCB + c;
          if (b < NQH) {
            const bf16x8& y = term == 1 ? fl[g & 1][c] : fh[g & 1][c];
#pragma unroll
            for (int rb = 0; rb < kRB; ++rb) {
              if (B0 == 0 && b == 0) {
                if (rb == 0) acc[0][0] = mfma_bf16(term == 2 ? sl : sh, y, acc[0][0]);
              } else {
                acc[rb][b] = mfma_bf16(term == 2 ? al[rb] : ah[rb], y, acc[rb][b]);
              }
            }
          }
        }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// The contraction of one tile by one wave, then its fold into LDS:
// fold[(3 c + k) * kW8Tests + row * 16 + gene], k = 0 gg, 1 sum of squares of
// the covariate projections, 2 num, each partial over this half's blocks.
// gsrc: the group's gene tiles (k-step s at gsrc + s * NQ * 2 KiB).
template <int PREC, int NQ, int NQH, int B0>
__device__ __attribute__((always_inline)) inline void w8_tile(const WaldArgs& a, unsigned char* smem,
                                                              const unsigned char* gsrc, int64_t row0, int c, int r,
                                                              int lane) {
  constexpr int kAS = w8_a_stages(NQ);
  constexpr int tile_bytes = NQ * 2048;
  constexpr int chunks = tile_bytes >> 10;
  const int gl = lane & 15;
  const int kk = 8 * (lane >> 4);
  const int nks = a.n_pad >> 5;
  unsigned char* const buf0 = smem;
  unsigned char* const buf1 = smem + tile_bytes;
  unsigned char* const abuf = smem + 2 * tile_bytes;
  const int64_t rowP = row0 + 32 * r;  // this wave's row pair
  const float* a_row[kRB];
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb) a_row[rb] = a.A + (rowP + 16 * rb + gl) * a.n_pad + kk;
  const int pair_off = r * 4096;  // row pair r in a stage: blocks rb = 0, 1, 2 x 1 KiB each
  (void)c;

  f32x4 acc[kRB][NQH];
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
    for (int b = 0; b < NQH; ++b) acc[rb][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 acc0 = f32x4{0.f, 0.f, 0.f, 0.f};  // half 1: block 0 of row block 1
  f32x8 an[kRB];
  // prologue (half 1): gene tile of k-step 0 and the first row stages
  if constexpr (B0 != 0) {
    for (int ch = r; ch < chunks; ch += 4) lds_direct16(gsrc + ch * 1024 + lane * 16, buf0 + ch * 1024);
  }
  if constexpr (B0 != 0 && kAS > 0) {
    for (int s = 0; s < kAS - 1 && s < nks; ++s) {
      unsigned char* dst = abuf + s * kAStageBytes + pair_off;
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) {
        lds_direct16(a_row[rb] + s * 32, dst + rb * 2048);
        lds_direct16(a_row[rb] + s * 32 + 4, dst + rb * 2048 + 1024);
      }
    }
  }
  if constexpr (kAS == 0) {
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) an[rb] = load8(a_row[rb]);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int s = 0; s < nks; ++s) {
    const int sr = kAS > 0 ? s + kAS - 1 : s + 1;  // k-step whose rows are fetched now
    const float* src[kRB];
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) src[rb] = a_row[rb] + sr * 32;
    unsigned char* const a_wr =
        (kAS > 0 && sr < nks) ? abuf + (sr % (kAS > 0 ? kAS : 1)) * kAStageBytes + pair_off : nullptr;
    const unsigned char* const a_rd = kAS > 0 ? abuf + (s % (kAS > 0 ? kAS : 1)) * kAStageBytes + pair_off : nullptr;
    unsigned char* const b_nxt = s + 1 < nks ? (((s + 1) & 1) ? buf1 : buf0) : nullptr;
    w8_kstep<PREC, NQ, NQH, B0, kAS>(acc, acc0, an, a_rd, a_wr, src, kAS == 0 && sr < nks, (s & 1) ? buf1 : buf0,
                                         b_nxt, gsrc + static_cast<size_t>(s + 1) * tile_bytes, r, lane);
    // every wave's copies for the next k-step must have landed (hence the
    // barrier); with three stages a half-1 wave leaves its row copies of
    // k-step s+2 (its four youngest operations) in flight
    if (B0 != 0 && kAS == 3 && s + 2 < nks) __builtin_amdgcn_s_waitcnt(kWaitVm4Lgkm0);
    else __builtin_amdgcn_s_waitcnt(0);
    lds_barrier();
  }
  // fold (the last barrier has released the staging buffers)
  double* fold = reinterpret_cast<double*>(smem);
  const int nq = a.nq;
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double gg = 0, ss = 0, num = 0;
#pragma unroll
      for (int b = 0; b < NQH; ++b) {
        const int gb = B0 + b;
        const double v = acc[rb][b][q];
        if (gb == 0) gg = v;
        else if (gb < nq - 1) ss += v * v;
        else if (gb == nq - 1) num = v;
      }
      if (B0 != 0 && rb == 1) gg = acc0[q];
      const int row = 32 * r + 16 * rb + 4 * (lane >> 4) + q;
      const int t = row * 16 + gl;
      fold[(3 * c + 0) * kW8Tests + t] = gg;
      fold[(3 * c + 1) * kW8Tests + t] = ss;
      fold[(3 * c + 2) * kW8Tests + t] = num;
    }
}

__device__ inline GeneAcc shfl_xor_acc(const GeneAcc& x, int m) {
  GeneAcc y;
  y.acat_hi = __shfl_xor(x.acat_hi, m);
  y.acat_lo = __shfl_xor(x.acat_lo, m);
  y.acat_pol = __shfl_xor(x.acat_pol, m);
  y.best_p = __shfl_xor(x.best_p, m);
  y.best_snp = __shfl_xor(x.best_snp, m);
  y.best_af = __shfl_xor(x.best_af, m);
  y.best_beta = __shfl_xor(x.best_beta, m);
  y.best_se = __shfl_xor(x.best_se, m);
  y.best_stat = __shfl_xor(x.best_stat, m);
  y.best_r2 = __shfl_xor(x.best_r2, m);
  y.used = __shfl_xor(x.used, m);
  y.n_tested = __shfl_xor(x.n_tested, m);
  y.n_sig = __shfl_xor(x.n_sig, m);
  y.pad = 0;
  return y;
}

// Work of one workgroup: gene group and SNP tiles. Trans scope, XCD-aware:
// the n_splits consecutive workgroups of one XCD share a gene group, so its
// column blocks are served from that XCD's L2. Cis scope: work item it of the
// queue (its partial state goes to slot it).
struct W8Work {
  int group, split, t_lo, t_hi, t_step;
};
__device__ inline W8Work w8_work_trans(const WaldArgs& a) {
  W8Work w;
  const int wg = blockIdx.x;
  const int xcd = wg & 7;
  const int jx = wg >> 3;
  w.group = (jx / a.n_splits) * 8 + xcd;
  w.split = jx % a.n_splits;
  w.t_lo = w.split;
  w.t_hi = (a.n_tiles + 1) >> 1;
  w.t_step = a.n_splits;
  return w;
}
__device__ inline W8Work w8_work_item(const WaldArgs& a, int it) {
  W8Work w;
  w.group = a.work[3 * it];
  w.t_lo = a.work[3 * it + 1];
  w.t_hi = a.work[3 * it + 2];
  w.split = it;
  w.t_step = 1;
  return w;
}

// The epilogue's gene: thread -> gene (tid & 15) of the group, tests of rows
// (tid >> 4) + 32 i, with the gene's scalar and window.
struct W8Gene {
  int gl, gene;
  bool ok, win_known;
  double gscal;
  int32_t gch;
  int64_t gcs, gce;
};
__device__ inline W8Gene w8_gene(const WaldArgs& a, int group) {
  W8Gene g;
  g.gl = threadIdx.x & 15;
  g.gene = group * 16 + g.gl;
  g.ok = g.gene < a.G && a.gene_valid[g.gene];
  g.gscal = g.ok ? a.gene_scal[g.gene] : 0.0;
  const bool excl = g.ok && a.gene_chrom && a.gene_chrom[g.gene] >= 0;
  g.gch = excl ? a.gene_chrom[g.gene] : -1;
  g.gcs = excl ? a.gene_cs[g.gene] : 0;
  g.gce = excl ? a.gene_ce[g.gene] : 0;
  g.win_known = excl && a.snp_chrom;
  return g;
}

// FP64 tests of one tile and gene group (2048 tests, 4 per thread, in the
// order i = 0..3): fold(t, gg, S, num) gives test t = row * 16 + gene. Returns
// the thread's gene state merged over the four lanes holding its gene (lanes
// l, l ^ 16, l ^ 32, l ^ 48 then hold the wave's total for gene l & 15).
template <int MODEL, class Fold>
__device__ inline GeneAcc w8_tests(const WaldArgs& a, const W8Gene& gi, const RowInfo& ri, const TailView& tv,
                                   int64_t row0, Fold fold) {
  GeneAcc st;
  acc_init(st);
  dd asum{0.0, 0.0};
#pragma unroll 1
  for (int i = 0; i < kW8Tests / kW8Threads; ++i) {
    const int jl = (threadIdx.x >> 4) + (kW8Threads / 16) * i;
    const int t = jl * 16 + gi.gl;
    const int64_t j = row0 + jl;
    const bool inwin = gi.win_known && ri.chrom[jl] == gi.gch && ri.pos[jl] >= gi.gcs && ri.pos[jl] <= gi.gce;
    const bool ok = gi.ok && j < a.m && ri.keep[jl] && (a.cis ? inwin : !inwin);
    double gg, S, num;
    fold(t, gg, S, num);
    const bool cand = refine_wanted<MODEL>(a, ok, gg, S, num, gi.gscal, ri.gtg[jl]);
    const double n64 = wave_refine(a, cand, j, gi.gene);
    if (cand) num = n64;
    if (ok) wald_test<MODEL>(a, ri, jl, j, gi.gene, gi.gscal, gg, S, num, tv, st, asum);
  }
  st.acat_hi = asum.hi;
  st.acat_lo = asum.lo;
  // lanes l, l ^ 16, l ^ 32, l ^ 48 hold the same gene
  acc_merge(st, shfl_xor_acc(st, 16));
  acc_merge(st, shfl_xor_acc(st, 32));
  return st;
}

// w8_tests, then the eight waves' totals (wpart) merged into the running
// states of the group's 16 genes in gstate.
template <int MODEL, class Fold>
__device__ inline void w8_epilogue(const WaldArgs& a, const W8Gene& gi, const RowInfo& ri, const TailView& tv,
                                   GeneAcc* gstate, GeneAcc* wpart, int64_t row0, Fold fold) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const GeneAcc st = w8_tests<MODEL>(a, gi, ri, tv, row0, fold);
  if (lane < 16) wpart[wave * 16 + lane] = st;
  __syncthreads();
  if (threadIdx.x < 16) {
    GeneAcc g = gstate[threadIdx.x];
    for (int w = 0; w < kW8Waves; ++w) acc_merge(g, wpart[w * 16 + threadIdx.x]);
    gstate[threadIdx.x] = g;
  }
}

// Loads one tile's SNP data (keep, af, annotation; the SNP mean for GLM) into LDS.
template <int MODEL>
__device__ inline void w8_row_info(const WaldArgs& a, RowInfo& ri, int64_t row0) {
  if (threadIdx.x < kTileRows) {
    const int t = threadIdx.x;
    const int64_t j = row0 + t;
    const bool in = j < a.m;
    ri.keep[t] = in ? a.keep[j] : 0;
    ri.af[t] = in ? a.af[j] : 0.0;
    ri.gtg[t] = (MODEL == EQTL_MODEL_GLM && in) ? a.gtg[j] : 0.0;
    const bool ann = in && a.snp_chrom;
    ri.chrom[t] = ann ? a.snp_chrom[a.snp0 + j] : -1;
    ri.pos[t] = ann ? a.snp_pos[a.snp0 + j] : 0;
  }
}

// One piece of work of the kernel below. MODEL: EQTL_MODEL_LMM (A = rotated
// dosage, Wald t test) or EQTL_MODEL_GLM (A = centred dosage, NB score test);
// both contract the same column blocks.
template <int PREC, int NQ, int MODEL>
__device__ __attribute__((always_inline)) inline void w8_item(const WaldArgs& a, unsigned char* smem,
                                                              const W8Work& wk) {
  constexpr int NQ0 = (NQ + 1) / 2, NQ1 = NQ - NQ0;
  if (wk.group >= a.n_groups) return;
  const int group = wk.group;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = wave >> 2, r = wave & 3;
  const W8Gene gi = w8_gene(a, group);

  const size_t r0b = w8_region0_bytes(NQ);
  RowInfo& ri = *reinterpret_cast<RowInfo*>(smem + r0b);
  double* thead = reinterpret_cast<double*>(smem + r0b + sizeof(RowInfo));
  GeneAcc* gstate = reinterpret_cast<GeneAcc*>(smem + r0b + sizeof(RowInfo) + sizeof(double) * kTailHead);
  const double* fold = reinterpret_cast<const double*>(smem);
  GeneAcc* wpart = reinterpret_cast<GeneAcc*>(smem + kW8FoldBytes);
  for (int i = threadIdx.x; i < kTailHead; i += kW8Threads) thead[i] = (&a.tail->df)[i];
  if (threadIdx.x < 16) acc_init(gstate[threadIdx.x]);
  const TailView tv{thead, &a.tail->u_m, a.tail->coef_b};
  const unsigned char* gsrc =
      PREC == EQTL_PREC_FP32 ? reinterpret_cast<const unsigned char*>(a.gc.f32 + a.gc.tile_f32(group, 0))
                             : reinterpret_cast<const unsigned char*>(a.gc.bf + a.gc.tile_bf(group, 0));

  for (int tile = wk.t_lo; tile < wk.t_hi; tile += wk.t_step) {
    const int64_t row0 = static_cast<int64_t>(tile) * kTileRows;
    __syncthreads();  // the previous tile's epilogue is done with ri, fold and the gene partials
    w8_row_info<MODEL>(a, ri, row0);
    if (c == 0) w8_tile<PREC, NQ, NQ0, 0>(a, smem, gsrc, row0, c, r, lane);
    else w8_tile<PREC, NQ, NQ1, NQ0>(a, smem, gsrc, row0, c, r, lane);
    __syncthreads();  // fold complete
    w8_epilogue<MODEL>(a, gi, ri, tv, gstate, wpart, row0, [&](int t, double& gg, double& S, double& num) {
      gg = fold[0 * kW8Tests + t] + fold[3 * kW8Tests + t];
      S = gg - (fold[1 * kW8Tests + t] + fold[4 * kW8Tests + t]);
      num = fold[2 * kW8Tests + t] + fold[5 * kW8Tests + t];
    });
  }
  if (threadIdx.x < 16) {
    const size_t slot =
        a.work ? static_cast<size_t>(wk.split) * 16 : static_cast<size_t>(wk.split) * a.G_pad + group * 16;
    a.partial[slot + threadIdx.x] = gstate[threadIdx.x];
  }
}

template <int PREC, int NQ, int MODEL>
__global__ __launch_bounds__(kW8Threads) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_wald_lmm(WaldArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  if (a.work) {  // cis: persistent workgroups over the item queue
    __shared__ int slot;
    for (int it = next_item(a, &slot); it >= 0; it = next_item(a, &slot))
      w8_item<PREC, NQ, MODEL>(a, smem, w8_work_item(a, it));
    return;
  }
  w8_item<PREC, NQ, MODEL>(a, smem, w8_work_trans(a));
}

}  // namespace eqtl_amd
