// Genotype-side kernels of the trans scans:
//   K1 k_decode   — genotype rows (2-bit PLINK or float dosage) → centred,
//                   mean-imputed dosages + per-SNP filters/stats
//                   (plink_bed.cpp:263-320, scan_common.cpp:133-151)
//   K6 k_rotate_lds / k_rotate_f32 — G~ = (G - mu)^T Q on MFMA (scan_trans_lmm.cpp:168)
//   K7/K8 k_wald  — fused contraction against the per-gene column blocks +
//                   FP64 Wald epilogue + per-gene reductions + hit emission
//                   (scan_trans_lmm.cpp:11-40,170-188; scan_trans_lm.cpp:21-36,
//                   161-177; scan_trans.hpp:53-110)
//   k_merge / k_finalize — fold per-split gene partials, finish ACAT.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "geno_rows.hpp"
#include "mfma.hpp"
#include "numerics.hpp"
#include "../../include/eqtl_gpu.h"

namespace eqtl_amd {

constexpr int kWave = 64;

// Per-gene running summary (device side of eqtl_gene_state).
struct GeneAcc {
  double acat_hi, acat_lo;
  double acat_pol;  // ACAT terms of tests polished in FP64 after the Wald kernel
  double best_p;
  int64_t best_snp;
  double best_af, best_beta, best_se, best_stat, best_r2;
  int32_t used, n_tested, n_sig, pad;
};

__host__ __device__ inline void acc_init(GeneAcc& a) {
  a.acat_hi = 0; a.acat_lo = 0; a.acat_pol = 0; a.best_p = 2.0; a.best_snp = -1;
  a.best_af = a.best_beta = a.best_se = a.best_stat = a.best_r2 = 0;
  a.used = a.n_tested = a.n_sig = 0; a.pad = 0;
}

// (p, snp) lexicographic minimum reproduces "strict < in SNP order" for best.
__host__ __device__ inline void acc_merge(GeneAcc& a, const GeneAcc& b) {
  const dd s = dd_add(dd{a.acat_hi, a.acat_lo}, dd{b.acat_hi, b.acat_lo});
  a.acat_hi = s.hi; a.acat_lo = s.lo;
  a.acat_pol += b.acat_pol;
  a.used += b.used; a.n_tested += b.n_tested; a.n_sig += b.n_sig;
  if (b.best_snp >= 0 && (b.best_p < a.best_p || (b.best_p == a.best_p && (a.best_snp < 0 || b.best_snp < a.best_snp)))) {
    a.best_p = b.best_p; a.best_snp = b.best_snp; a.best_af = b.best_af; a.best_beta = b.best_beta;
    a.best_se = b.best_se; a.best_stat = b.best_stat; a.best_r2 = b.best_r2;
  }
}

// ---------------------------------------------------------------------------
// K1: decode. One wave per SNP row. Rows >= m (padding) are written as zeros.
// BED rows get the source's rules here (decode_row, plink_bed.cpp:263-320);
// DOSE32 rows had them applied when they were written (a rejected row is all
// missing), so only n_ok > 0 is left. Then subset_maf_or_nan
// (scan_common.cpp:133-151) on the mean-imputed row.
// ---------------------------------------------------------------------------
struct DecodeArgs {
  const uint8_t* rows;
  int64_t m, m_pad;
  int32_t bps, n, n_file, n_pad;
  int32_t n_src;              // analysis samples (source rules, imputation)
  const int32_t* sample_col;  // n_src file columns
  const int32_t* test_col;    // nullable: n file columns of the test subset (else sample_col, n == n_src)
  int32_t impute;
  double max_miss, maf_min;
  float* A;            // m_pad x n_pad, centred imputed dosage (fp32)
  double* A64;         // nullable, same centred values in fp64
  double* dosage;      // nullable, m x n imputed dosage (uncentred)
  uint8_t* keep;       // m_pad
  double* af;          // m_pad
  double* mean;        // nullable, m_pad: the centring mean of each kept SNP (0 otherwise)
};

template <int FMT>
__global__ __launch_bounds__(256) void k_decode(DecodeArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (j >= a.m_pad) return;
  float* Arow = a.A + j * a.n_pad;
  double* A64row = a.A64 ? a.A64 + j * a.n_pad : nullptr;
  if (j >= a.m) {
    for (int i = lane; i < a.n_pad; i += kWave) {
      Arow[i] = 0.f;
      if (A64row) A64row[i] = 0.0;
    }
    if (lane == 0) {
      a.keep[j] = 0;
      a.af[j] = __builtin_nan("");
      if (a.mean) a.mean[j] = 0.0;
    }
    return;
  }
  const uint8_t* row = a.rows + j * a.bps;
  int n_miss = 0, bad = 0;
  double sum = 0.0, sum2 = 0.0;  // integer-valued for BED rows (exact)
  for (int i = lane; i < a.n_src; i += kWave) {
    const int col = a.sample_col[i];
    if (col < 0 || col >= a.n_file) { bad = 1; continue; }
    bool miss;
    const double g = geno_at<FMT>(row, col, miss);
    if (miss) { ++n_miss; continue; }
    sum += g;
    sum2 += g * g;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    n_miss += __shfl_xor(n_miss, o);
    sum += __shfl_xor(sum, o);
    sum2 += __shfl_xor(sum2, o);
    bad |= __shfl_xor(bad, o);
  }
  const int n_src = a.n_src;
  const int n_ok = n_src - n_miss;
  bool keep = !bad && n_ok > 0;
  double mu = 0, muc = 0, afs = __builtin_nan("");
  if constexpr (FMT == EQTL_GENO_BED) {
    if (keep) {
      const double miss_frac = static_cast<double>(n_miss) / static_cast<double>(n_src);
      if (miss_frac > a.max_miss + 1e-15) keep = false;
      if (!a.impute && n_miss > 0) keep = false;
    }
    if (keep) {
      double afd = (sum / n_ok) / 2.0;
      afd = fmin(1.0, fmax(0.0, afd));
      if (afd < 1e-12 || afd > 1.0 - 1e-12) keep = false;
      if (a.maf_min > 0.0 && (afd < a.maf_min || afd > 1.0 - a.maf_min)) keep = false;  // pass_maf
    }
  }
  const int n = a.n;
  const int32_t* tcol = a.test_col ? a.test_col : a.sample_col;
  if (keep) {
    mu = sum / n_ok;
    // subset_maf_or_nan on the imputed vector (restricted to the test subset)
    double sI = sum + n_miss * mu, s2I = sum2 + n_miss * mu * mu;
    if (a.test_col) {
      double s1 = 0.0, s2 = 0.0;
      for (int i = lane; i < n; i += kWave) {
        bool miss;
        const double g = geno_at<FMT>(row, tcol[i], miss);
        const double v = miss ? mu : g;
        s1 += v;
        s2 += v * v;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
      }
      sI = s1;
      s2I = s2;
    }
    double af = (sI / n) / 2.0;
    af = fmin(1.0, fmax(0.0, af));
    const double maf = af > 0.5 ? 1.0 - af : af;
    const double var = s2I / n - (sI / n) * (sI / n);
    if (maf < 1e-12 || var < 1e-12) keep = false;
    afs = af;
    muc = sI / n;
  }
  for (int i = lane; i < a.n_pad; i += kWave) {
    double v = 0.0, dose = __builtin_nan("");
    if (i < n) {
      const int col = tcol[i];
      if (col >= 0 && col < a.n_file) {
        bool miss;
        const double g = geno_at<FMT>(row, col, miss);
        dose = miss ? (a.impute ? mu : __builtin_nan("")) : g;
        if (keep) v = (miss ? mu : g) - muc;
      }
      if (a.dosage) a.dosage[j * n + i] = dose;
    }
    Arow[i] = static_cast<float>(v);
    if (A64row) A64row[i] = v;
  }
  if (lane == 0) {
    a.keep[j] = keep ? 1 : 0;
    a.af[j] = keep ? afs : __builtin_nan("");
    if (a.mean) a.mean[j] = keep ? muc : 0.0;
  }
}

// ---------------------------------------------------------------------------
// K6, FP32 form: rotation Gt[j, c] = sum_i A[j, i] Q[i, c]. WG tile 64 rows x
// 64 coords, 4 waves x (16 rows x 4 blocks of 16 coords), operands read from
// L2. Q stored coordinate-major (Q[c * n_pad + i] = Q(i, c)), i.e.
// K-contiguous for the B operand.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rotate_f32(const float* __restrict__ A, const float* __restrict__ Q32,
                                                    float* __restrict__ Gt, int n_pad) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t r0 = static_cast<int64_t>(blockIdx.y) * 64 + wave * 16;
  const int c0 = blockIdx.x * 64;
  const int kk = 8 * (lane >> 4);
  f32x4 acc[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  const float* arow = A + (r0 + (lane & 15)) * n_pad + kk;
  for (int k0 = 0; k0 < n_pad; k0 += 32) {
    const f32x8 av = load8(arow + k0);
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const size_t bo = static_cast<size_t>(c0 + cb * 16 + (lane & 15)) * n_pad + k0 + kk;
      acc[cb] = mfma_f32x8(av, load8(Q32 + bo), acc[cb]);
    }
  }
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t row = r0 + 4 * (lane >> 4) + r;
      Gt[row * n_pad + c0 + cb * 16 + (lane & 15)] = acc[cb][r];
    }
}

// K6, split form, staged through LDS: 128 SNP rows x 128 coordinates per
// workgroup. Per 32-sample step the rows' FP32 values (8 chunks of 16 B per
// row) and the coordinates' Bhi / Blo (4 chunks each) are copied global -> LDS
// with direct 16-byte loads, double-buffered. Each copy instruction reads
// whole cache lines: a row copy takes 8 rows x 128 B (lane l: row l >> 3,
// slot l & 7 holding chunk (l & 7) ^ ((row >> 1) & 7), so the 16 rows an MFMA
// fragment reads lie in distinct LDS banks), a coordinate copy one 1 KiB block
// of the fragment-major B planes (rot_b_index; stored chunk-major in LDS, slot
// (c, r) at (c * 128 + r) * 16). Wave (wr, wc) owns the 64 x 64 quadrant
// (rows wr*64.., coordinates wc*64..) as 4 x 4 blocks, three bf16 products
// per block (hi*hi, hi*lo, lo*hi, in that order). The B planes hold whole
// 128-coordinate blocks (zero past n_cols); coordinates past n_cols are not
// stored.
constexpr int kRotT = 128;
// Element (coordinate, sample k) of the rotation's B planes: per 64-coordinate
// block and 32-sample step one contiguous 4 KiB block, chunk (k & 31) >> 3
// then coordinate then k & 7 (nks = samples / 32).
__host__ __device__ inline int64_t rot_b_index(int64_t coord, int64_t k, int64_t nks) {
  return ((coord >> 6) * nks + (k >> 5)) * 2048 + ((k & 31) >> 3) * 512 + (coord & 63) * 8 + (k & 7);
}
// coordinates the B planes hold for n_cols columns
__host__ __device__ inline int64_t rot_b_coords(int64_t n_cols) { return (n_cols + kRotT - 1) / kRotT * kRotT; }
constexpr size_t kRotA = 8 * kRotT * 16;       // FP32 rows: 16 KiB per step
constexpr size_t kRotStage = kRotA + 8 * kRotT * 16;  // + Bhi and Blo: 32 KiB
constexpr size_t kRotLds = 2 * kRotStage;

// Element (row, col) of a bf16 plane in fragment-major order: per 16-row block
// and 32-column chunk one contiguous 1 KiB block of 64 x 16 bytes, lane l
// holding row (l & 15), columns 8 (l >> 4) .. + 7 of the chunk — the operand
// fragment of v_mfma_f32_16x16x32_bf16, so a 64-lane direct copy of 16 bytes
// per lane moves one whole block. nch = chunks per row block.
__host__ __device__ inline int64_t frag_index(int64_t row, int64_t col, int64_t nch) {
  return ((row >> 4) * nch + (col >> 5)) * 512 + ((row & 15) + 16 * ((col & 31) >> 3)) * 8 + (col & 7);
}

// Also used for any B of n_cols coordinate-major columns (output leading
// dimension ld_out); the rotation is n_cols = ld_out = n_pad. SQUARE: the A
// values are squared before the split (sum_i A_ji^2 B_ic). Outputs (each
// nullable): Gt fp32 (row-major, leading dimension ld_out); Ohi / Olo the
// bf16 hi / lo split of the fp32 value, with Omid the exact three-way split
// hi / mid / lo, in fragment-major order (frag_index, ld_out / 32 chunks per
// row block) at columns frag_col0 + col.
template <bool SQUARE>
__global__ __launch_bounds__(256) void k_rotate_lds(const float* __restrict__ A, const __bf16* __restrict__ Bhi,
                                                    const __bf16* __restrict__ Blo, float* __restrict__ Gt,
                                                    int n_pad, int n_cols, int64_t ld_out, __bf16* __restrict__ Ohi,
                                                    __bf16* __restrict__ Olo, __bf16* __restrict__ Omid,
                                                    int frag_col0) {
  extern __shared__ __attribute__((aligned(16))) unsigned char rsm[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t R0 = static_cast<int64_t>(blockIdx.y) * kRotT;
  const int C0 = blockIdx.x * kRotT;
  // this wave's 8 of the step's 32 copy instructions (q = wave + 4 i):
  // q < 16 rows 8q .. 8q + 7 of A, then Bhi, then Blo (chunk (q & 7) >> 1 of
  // the coordinates' half h = q & 1)
  const int nks = n_pad / 32;
  const unsigned char* src[8];
  int step_bytes[8], dst_off[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int q = wave + 4 * i;
    const int h = q & 1;
    if (q < 16) {
      const int row = 8 * q + (lane >> 3);
      const int c = (lane & 7) ^ ((row >> 1) & 7);
      src[i] = reinterpret_cast<const unsigned char*>(A + (R0 + row) * n_pad + 4 * c);
      step_bytes[i] = 32 * static_cast<int>(sizeof(float));
      dst_off[i] = q * 1024;
    } else {
      const int c = (q & 7) >> 1;
      const __bf16* base = q < 24 ? Bhi : Blo;
      src[i] = reinterpret_cast<const unsigned char*>(base + rot_b_index(C0 + h * 64, 0, nks) + c * 512 + lane * 8);
      step_bytes[i] = 2048 * static_cast<int>(sizeof(__bf16));
      dst_off[i] = static_cast<int>(kRotA) + (q < 24 ? 0 : 4 * kRotT * 16) + (c * kRotT + h * 64) * 16;
    }
  }
  auto issue = [&](int s, unsigned char* buf) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
      __builtin_amdgcn_global_load_lds(src[i] + s * step_bytes[i],
                                       (__attribute__((address_space(3))) void*)(buf + dst_off[i]), 16, 0, 0);
  };
  f32x4 acc[4][4];
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  issue(0, rsm);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  const int g = lane >> 4;
  for (int s = 0; s < nks; ++s) {
    const unsigned char* cur = rsm + (s & 1) * kRotStage;
    if (s + 1 < nks) issue(s + 1, rsm + ((s + 1) & 1) * kRotStage);
    bf16x8 bh[4], bl[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const int rc = wc * 64 + cb * 16 + (lane & 15);
      bh[cb] = *reinterpret_cast<const bf16x8*>(cur + kRotA + (g * kRotT + rc) * 16);
      bl[cb] = *reinterpret_cast<const bf16x8*>(cur + kRotA + 4 * kRotT * 16 + (g * kRotT + rc) * 16);
    }
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      const int ra = wr * 64 + rb * 16 + (lane & 15);
      const int sw = (ra >> 1) & 7;
      const float4 lo = *reinterpret_cast<const float4*>(cur + (ra * 8 + ((2 * g) ^ sw)) * 16);
      const float4 hi = *reinterpret_cast<const float4*>(cur + (ra * 8 + ((2 * g + 1) ^ sw)) * 16);
      f32x8 av;
      av.v[0] = lo.x; av.v[1] = lo.y; av.v[2] = lo.z; av.v[3] = lo.w;
      av.v[4] = hi.x; av.v[5] = hi.y; av.v[6] = hi.z; av.v[7] = hi.w;
      if constexpr (SQUARE) {
#pragma unroll
        for (int t = 0; t < 8; ++t) av.v[t] *= av.v[t];
      }
      bf16x8 ah, al;
      split_bf16(av, ah, al);
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        acc[rb][cb] = mfma_bf16(ah, bh[cb], acc[rb][cb]);
        acc[rb][cb] = mfma_bf16(ah, bl[cb], acc[rb][cb]);
        acc[rb][cb] = mfma_bf16(al, bh[cb], acc[rb][cb]);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const int col = C0 + wc * 64 + cb * 16 + (lane & 15);
      if (col >= n_cols) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = R0 + wr * 64 + rb * 16 + 4 * (lane >> 4) + r;
        const float v = acc[rb][cb][r];
        if (Gt) Gt[row * ld_out + col] = v;
        if (Ohi) {
          const int64_t o = frag_index(row, frag_col0 + col, ld_out >> 5);
          const __bf16 h = to_bf16(v);
          const float r1 = v - static_cast<float>(h);
          Ohi[o] = h;
          if (Omid) {
            const __bf16 m = to_bf16(r1);
            Omid[o] = m;
            Olo[o] = to_bf16(r1 - static_cast<float>(m));
          } else {
            Olo[o] = to_bf16(r1);
          }
        }
      }
    }
}

// ---------------------------------------------------------------------------
// Fused Wald scans: one workgroup = one group of 16 genes x a strided set of
// 128-SNP tiles. The gene-side operand is stored in column blocks: block b of a
// group holds quantity b for genes 0..15 (lane & 15), so after the MFMA
// contraction each lane holds all quantities of one (SNP row, gene) pair per
// output register:
//   LMM: b=0 d (against A^2 -> gg), b=1..p whitened orthonormal covariates
//        (-> W^T h), b=p+1 whitened null residual (-> r^T h); k_wald_lmm
//   LM : b=0 residualised phenotype y_s (-> g^T y_s); k_wald_lm below
// Both stage the group's k-step gene tile and a ring of SNP-row stages in LDS
// with direct global->LDS 16-byte loads, and run the FP64 epilogue (t tail,
// p, ACAT, hits) with the accumulators parked in LDS.
// ---------------------------------------------------------------------------
struct PolishItem {
  int32_t j;     // block-local SNP row
  int32_t gene;
  int64_t pad;
};

struct WaldArgs {
  const float* A;          // m_pad x n_pad
  GeneCols gc;             // gene-side operand (fragment-major tiles)
  int32_t n_pad, nq, G, G_pad, n_groups, n_groups_pad, n_splits, n_tiles;
  int64_t m, snp0;
  const uint8_t* keep;     // block-local
  const double* af;        // block-local
  const double* gtg;       // LM: per SNP ||M_X g||^2 (block-local)
  const double* gene_scal; // LMM rss_null, LM yty
  const int32_t* gene_valid;
  const int32_t* snp_chrom;   // global SNP annotation (nullable)
  const int64_t* snp_pos;
  const int32_t* gene_chrom;  // nullable: no cis exclusion
  const int64_t* gene_cs;
  const int64_t* gene_ce;
  double df, pthr;
  const TTail* tail;          // t tail table for df
  GeneAcc* partial;           // [n_splits][G_pad]
  eqtl_hit* hits;
  unsigned long long* hit_count;  // [0] hits, [1] polish items
  int64_t hit_cap;
  // FP64 numerator refinement (tau = 0 and tau_hi = inf disable it)
  const double* A64;          // m_pad x n_pad centred dosages, FP64, analysis order
  const double* vref;         // G x n FP64 residual vectors in sample space
  int32_t n;
  double tau, tau_hi;
  // LMM / GLM tests recomputed entirely in FP64 afterwards: |t| > tau_hi, or
  // S < kappa gg (the covariates explain more than 1 - kappa of the SNP's
  // weighted sum of squares, so S = gg - ||W^T h||^2 cancels)
  double kappa;
  PolishItem* polish;
  int64_t polish_cap;
  // scope: 0 trans (a gene's cis window is excluded), 1 cis (only the window
  // is tested). Cis launches run over work items {group, tile_lo, tile_hi}
  // (128-row tiles of the block) instead of (group, split) pairs; partial
  // states are then written per item. The workgroups of a cis launch are
  // persistent: each takes the next queue position from *queue (zeroed before
  // the launch) until n_items are taken; position q runs item order[q]
  // (longest items first).
  int32_t cis;
  const int32_t* work;
  const int32_t* order;
  unsigned int* queue;
  int32_t n_items;
};

// Next cis work item of a persistent workgroup (-1: the queue is drained).
// Every thread of the workgroup must call it; slot is a workgroup-shared int.
__device__ inline int next_item(const struct WaldArgs& a, int* slot) {
  __syncthreads();  // the previous item is done with slot and with LDS
  if (threadIdx.x == 0) *slot = static_cast<int>(atomicAdd(a.queue, 1u));
  __syncthreads();
  const int q = *slot;
  return q < a.n_items ? a.order[q] : -1;
}

constexpr int kRB = 2;          // 16-row blocks per wave
// s_waitcnt immediate (gfx9 encoding): vmcnt <= 2 * kRB (bits 3:0, 15:14),
// expcnt and lgkmcnt at their maximum (no wait).
constexpr int kWaitVm4 = (2 * kRB) | (0x7 << 4) | (0xF << 8);
constexpr int kWaitVm4Lgkm0 = (2 * kRB) | (0x7 << 4);  // same, and lgkmcnt 0
constexpr int kWaitLgkm0 = 0xF | (0x7 << 4) | (0x3 << 14);  // lgkmcnt 0 only
static_assert(2 * kRB < 16, "vmcnt field");

constexpr int kTileRows = 128;  // SNP rows per workgroup tile

// Per-tile SNP data staged in LDS.
struct RowInfo {
  double af[kTileRows];
  double gtg[kTileRows];
  int64_t pos[kTileRows];
  int32_t chrom[kTileRows];
  uint8_t keep[kTileRows];
};

struct TailView {
  const double* head;    // LDS copy of the header + piece A
  const double* mid;     // piece M (global, or an LDS copy)
  const double* coef_b;  // piece B (global)
  int sa = kTailDeg;     // coefficient row strides of pieces A and M (p_tail_xy)
  int sm = kTailDeg;
};

// FP64 numerator refinement at both ends of the t range.
//  * near null: for p near 1 the ACAT term is about -1/(pi (1 - p)), and
//    1 - p is proportional to |t| there, so the contraction's absolute error in
//    t turns into a large relative error of that term (and of the gene's ACAT
//    p); tests with |t| < tau are refined.
//  * extreme t: the relative p error grows like |t| * dt, so tests with
//    |t| > tau_hi are refined too (and, for LMM, polished after the kernel).
// Refinement recomputes the numerator in FP64: num = v_k^T g_c with g_c the
// centred FP64 dosage and v_k the gene's residual vector mapped back to sample
// space (LMM: v_k = Q (s o r_k), so v_k^T g_c = (s o r_k)^T Q^T g_c; LM:
// v_k = r_k, and with the FP64 g^T g the LM test is then entirely FP64).
template <int MODEL>
__device__ inline bool refine_wanted(const WaldArgs& a, bool ok, double gg, double S, double num, double gscal,
                                     double gtg) {
  if (!ok || (MODEL != EQTL_MODEL_GLM && !(a.df > 0))) return false;
  const double tau2 = a.tau * a.tau, hi2 = a.tau_hi * a.tau_hi;
  if constexpr (MODEL == EQTL_MODEL_GLM) {
    // score test: U = num + mean * sum(r) (gtg carries the SNP mean, gscal sum r)
    // (tests with |z| > tau_hi are polished in FP64 after the kernel)
    if (S < 1e-12) return false;
    const double U = num + gtg * gscal;
    return U * U < tau2 * S;  // z^2 = U^2 / S
  } else if constexpr (MODEL == EQTL_MODEL_LMM) {
    if (gg < 1e-12 || S <= 1e-15) return false;
    // t^2 = num^2 df / (S qq) with S qq = max(S gscal - num^2, 0)
    const double n2 = num * num;
    const double sq = fmax(S * gscal - n2, 0.0);
    const double t2s = n2 * a.df;
    return t2s < tau2 * sq || t2s > hi2 * sq;
  } else {
    if (gtg < 1e-12) return false;
    const double beta = num / gtg;
    const double rss = fmax(gscal - beta * num, 0.0);
    const double t2r = beta * beta * gtg * a.df;  // t^2 = t2r / rss
    return t2r < tau2 * rss || t2r > hi2 * rss;
  }
}

// The whole wave computes each flagged lane's FP64 dot product together: one
// 2n-double read per refined test and no divergence. Every lane of the wave
// must call this (wave-uniform control flow). The operands are read in rounds
// of kRefineRound samples with every load of a round issued before the first
// FMA, so a refined test costs about one memory latency per round (n = 2,000:
// two rounds) instead of one per 64 samples.
constexpr int kRefineUnroll = 16;
constexpr int kRefineRound = 64 * kRefineUnroll;
__device__ inline double wave_refine(const WaldArgs& a, bool cand, int64_t j, int gene) {
  unsigned long long mask = __ballot(cand);
  const int lane = threadIdx.x & 63;
  double mine = 0.0;
  while (mask) {
    const int src = __builtin_ctzll(mask);
    mask &= mask - 1;
    const int js = __shfl(static_cast<int>(j), src);
    const int gs = __shfl(gene, src);
    const double* x = a.A64 + static_cast<size_t>(js) * a.n_pad;
    const double* v = a.vref + static_cast<size_t>(gs) * a.n;
    double s = 0.0;
    for (int base = 0; base < a.n; base += kRefineRound) {
      double xv[kRefineUnroll], vv[kRefineUnroll];
#pragma unroll
      for (int u = 0; u < kRefineUnroll; ++u) {
        const int i = base + lane + 64 * u;
        const bool in = i < a.n;
        xv[u] = in ? x[i] : 0.0;
        vv[u] = in ? v[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kRefineUnroll; ++u) s = fma(xv[u], vv[u], s);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == src) mine = s;
  }
  return mine;
}

// One (SNP row j, gene) test: FP64 Wald (test_lmm_gtil, scan_trans_lmm.cpp:11-40
// / hit_from_gty, scan_trans_lm.cpp:21-36), then the per-gene statistics
// (apply_snp_hit_stats, scan_trans.hpp:87-102). The caller has checked the
// gene, the row bound, the SNP's keep flag and the window rule of the scope.
template <int MODEL>
__device__ inline void wald_test(const WaldArgs& a, const RowInfo& ri, int jl, int64_t j, int gene, double gscal,
                                 double gg, double S, double num, const TailView& tv, GeneAcc& st, dd& asum) {
  const int64_t snp = a.snp0 + j;
  double beta = 0, se = 0, stat = 0, p = 1.0, r2 = 0;
  const double df = a.df;
  bool big_t = false;  // |t| > tau_hi (LMM, GLM: polished after the kernel)
  bool lazy = false;   // LMM: beta, se, t, r2 still to be formed
  if constexpr (MODEL == EQTL_MODEL_GLM) {
    // test_glm_nb (model_glm.cpp:117-131): U = g^T r with the uncentred dosage
    // = (g - mean)^T r + mean * sum(r); info = S; two-sided normal p
    if (!(S < 1e-12)) {
      const double U = num + ri.gtg[jl] * gscal;
      const double rs = sqrt(S);
      beta = U / S;
      se = 1.0 / rs;
      stat = U / rs;
      p = isfinite(stat) ? fmin(1.0, fmax(0.0, erfc(fabs(stat) * 0.70710678118654752440))) : 1.0;
      big_t = fabs(stat) > a.tau_hi;
    }
  } else if constexpr (MODEL == EQTL_MODEL_LMM) {
    if (!(gg < 1e-12) && !(S <= 1e-15)) {
      // p depends on t only through y = t^2 / (df + t^2) = num^2 / (S gscal)
      // (t^2 = num^2 df / (S qq), S qq = S gscal - num^2): one division per
      // test. qq clamped to 0 gives t = 0 and df <= 0 gives p = 1, as in the
      // reference; beta, se, t and r2 are formed its way below, only for the
      // tests that report them (hits, best).
      const double n2 = num * num, sgs = S * gscal, sq = sgs - n2;
      const double th2 = a.tau_hi * a.tau_hi;
      if (df > 0 && sq > 0) {
        const double y = n2 / sgs;
        p = isfinite(y) ? p_tail_xy(tv.head, tv.mid, tv.coef_b, 1.0 - y, y, tv.sa, tv.sm) : 1.0;
        big_t = n2 * df > th2 * sq;
      } else {
        big_t = !(df > 0) && n2 > th2 * S;
      }
      lazy = true;
    }
  } else {
    const double gtg = ri.gtg[jl];
    const double gty = num;
    if (!(gtg < 1e-12) && df > 0) {
      beta = gty / gtg;
      const double rss = gscal - beta * gty;
      const double s2 = fmax(rss / df, 0.0);
      se = sqrt(s2 / gtg);
      stat = se > 0 ? beta / se : 0.0;
      p = p_from_t_tab2(tv.head, tv.mid, tv.coef_b, stat, tv.sa, tv.sm);
      r2 = gscal > 0 ? 1.0 - rss / gscal : 0.0;
    } else if (!(gtg < 1e-12)) {
      beta = gty / gtg;  // df <= 0: p = 1 with beta set (model_lm.cpp:44-48)
    }
  }
  if (!isfinite(p)) return;
  ++st.n_tested;
  // LMM / GLM tests with |t| > tau_hi or a cancelling S are polished in FP64
  // after the kernel, which then does their ACAT term, n_sig, hit and best
  // (k_polish, k_pbest_*); here they only count as tested
  const bool polish = MODEL != EQTL_MODEL_LM && (big_t || (S < a.kappa * gg && !(S <= 1e-15)));
  if (polish) {
    const unsigned long long pi = atomicAdd(a.hit_count + 1, 1ull);
    if (static_cast<int64_t>(pi) < a.polish_cap) a.polish[pi] = PolishItem{static_cast<int32_t>(j), gene, -1};
    return;
  }
  double term;
  if (acat_accept(p, &term)) {
    asum = dd_add_d(asum, term);
    ++st.used;
  }
  const double afj = ri.af[jl];
  const bool report = p <= a.pthr || p < st.best_p || (p == st.best_p && snp < st.best_snp);
  if (MODEL == EQTL_MODEL_LMM && lazy && report) {
    // test_lmm_gtil's own arithmetic (scan_trans_lmm.cpp:28-37)
    beta = num / S;
    double qq = gscal - S * beta * beta;
    if (qq < 0) qq = 0;
    const double sigma2 = df > 0 ? qq / df : 1.0;
    se = sqrt(fmax(sigma2 / S, 0.0));
    stat = se > 0 ? beta / se : 0.0;
    r2 = gscal > 1e-15 ? fmax(0.0, 1.0 - qq / gscal) : 0.0;
  }
  if (p <= a.pthr) {
    ++st.n_sig;
    const unsigned long long idx = atomicAdd(a.hit_count, 1ull);
    if (static_cast<int64_t>(idx) < a.hit_cap) {
      eqtl_hit h;
      h.snp = snp; h.gene = gene; h.pad = 0;
      h.af = afj; h.beta = beta; h.se = se; h.stat = stat; h.p = p; h.r2 = r2;
      a.hits[idx] = h;
    }
  }
  if (p < st.best_p || (p == st.best_p && snp < st.best_snp)) {
    st.best_p = p; st.best_snp = snp; st.best_af = afj; st.best_beta = beta;
    st.best_se = se; st.best_stat = stat; st.best_r2 = r2;
  }
}

__device__ inline void lds_direct16(const void* gsrc, unsigned char* lds_wave_base) {
  __builtin_amdgcn_global_load_lds(gsrc, (__attribute__((address_space(3))) void*)(lds_wave_base), 16, 0, 0);
}

// LDS layout of the LM Wald kernel. Region 0 holds, during the contraction,
// the two k-step buffers of the gene tile (2 KiB each) and a ring of
// kLmStages SNP-row stages; then gty of the tile's tests during the epilogue; then the
// per-gene partials of the final reduction. RowInfo and the t tail's padded
// LDS copy (tail_to_lds) follow it.
constexpr int kFoldDoubles = kRB * 4 * 256;  // one gty per test of the tile
constexpr int kAStageBytes = kTileRows * 32 * 4;  // 128 rows x 32 samples fp32
constexpr int kLmStages = 2;
constexpr long kLdsCap = 160 * 1024;
__host__ __device__ constexpr size_t lm_region0_bytes() {
  const size_t b = 2 * 2048 + static_cast<size_t>(kLmStages) * kAStageBytes;
  const size_t f = sizeof(double) * kFoldDoubles, r = 256 * sizeof(GeneAcc);
  return b > f ? (b > r ? b : r) : (f > r ? f : r);
}
__host__ __device__ constexpr size_t lm_lds_bytes() {
  return lm_region0_bytes() + sizeof(RowInfo) + sizeof(double) * (kTailLdsHead + kTailLdsMid);
}
static_assert(lm_lds_bytes() <= kLdsCap, "LM Wald LDS layout");

// Direct-to-LDS copy of one k-step of this wave's SNP rows (kRB blocks x 2
// chunks of 1 KiB).
__device__ inline void wald_stage_rows(const float* const (&src)[kRB], unsigned char* dst, int lane) {
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
    for (int h = 0; h < 2; ++h) lds_direct16(src[rb] + 4 * h, dst + (rb * 2 + h) * 1024);
  (void)lane;
}

// One 32-sample k-step of the LM contraction for one wave: read this
// k-step's rows (a_rd) and gene tile (b_cur: one block, y_s of the 16 genes)
// from LDS, start the copies of the next gene tile (b_nxt) and of a later row
// stage (a_wr), and run the MFMAs. The LDS pointers are __restrict__: the four
// regions are disjoint, which lets the compiler see that the copies in flight
// do not feed this k-step's LDS reads (otherwise it drains vmcnt before them).
template <int PREC>
__device__ __attribute__((always_inline)) inline void lm_kstep(
    f32x4 (&acc)[kRB], const unsigned char* __restrict__ a_rd, unsigned char* __restrict__ a_wr,
    const float* const (&a_src)[kRB], const unsigned char* __restrict__ b_cur, unsigned char* __restrict__ b_nxt,
    const unsigned char* b_src, int wave, int lane) {
  f32x8 ac[kRB];
#pragma unroll
  for (int rb = 0; rb < kRB; ++rb) {
    const float4 lo = *reinterpret_cast<const float4*>(a_rd + (rb * 2 + 0) * 1024 + lane * 16);
    const float4 hi = *reinterpret_cast<const float4*>(a_rd + (rb * 2 + 1) * 1024 + lane * 16);
    ac[rb].v[0] = lo.x; ac[rb].v[1] = lo.y; ac[rb].v[2] = lo.z; ac[rb].v[3] = lo.w;
    ac[rb].v[4] = hi.x; ac[rb].v[5] = hi.y; ac[rb].v[6] = hi.z; ac[rb].v[7] = hi.w;
  }
  // the gene tile is two 1 KiB chunks (either precision): waves 0 and 1
  if (b_nxt && wave < 2) lds_direct16(b_src + wave * 1024 + lane * 16, b_nxt + wave * 1024);
  // the row stage goes out after the tile load, so the caller's wait can
  // leave it (the youngest 2 kRB operations) in flight
  __builtin_amdgcn_sched_barrier(0);
  if (a_wr) wald_stage_rows(a_src, a_wr, lane);
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (PREC == EQTL_PREC_FP32) {
    const f32x8 bv = load8(reinterpret_cast<const float*>(b_cur + lane * 32));
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) acc[rb] = mfma_f32x8(ac[rb], bv, acc[rb]);
  } else {
    const bf16x8 bh = *reinterpret_cast<const bf16x8*>(b_cur + lane * 16);
    const bf16x8 bl = *reinterpret_cast<const bf16x8*>(b_cur + 1024 + lane * 16);
    bf16x8 ah[kRB], al[kRB];
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) split_bf16(ac[rb], ah[rb], al[rb]);
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) acc[rb] = mfma_bf16(ah[rb], bh, acc[rb]);
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) acc[rb] = mfma_bf16(ah[rb], bl, acc[rb]);
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) acc[rb] = mfma_bf16(al[rb], bh, acc[rb]);
  }
}

// K8: fused trans/cis LM scan (gty = g^T y_s per test, FP64 Wald from it and
// the per-SNP FP64 g^T g). One workgroup = one group of 16 genes x a strided
// set of 128-SNP tiles, 4 waves x 32 rows; the group's one-block gene tile and
// the SNP rows are staged in LDS as in the LMM kernel. lm_item runs one
// (group, tiles) piece of work.
template <int PREC>
__device__ __attribute__((always_inline)) inline void lm_item(const WaldArgs& a, unsigned char* smem, int group,
                                                              int split, int t_lo, int t_hi, int t_step) {
  constexpr int kAS = kLmStages;
  if (group >= a.n_groups) return;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int gl = lane & 15;
  const int gene = group * 16 + gl;
  const bool gene_ok = gene < a.G && a.gene_valid[gene];
  const int kk = 8 * (lane >> 4);
  const int nks = a.n_pad >> 5;
  const double gscal = gene_ok ? a.gene_scal[gene] : 0.0;
  // the gene's window (trans: excluded; cis: the only SNPs tested)
  const bool excl = gene_ok && a.gene_chrom && a.gene_chrom[gene] >= 0;
  const int32_t gch = excl ? a.gene_chrom[gene] : -1;
  const int64_t gcs = excl ? a.gene_cs[gene] : 0, gce = excl ? a.gene_ce[gene] : 0;
  const bool win_known = excl && a.snp_chrom;

  constexpr int tile_bytes = 2048;  // one k-step tile of the group (either precision)
  unsigned char* const buf0 = smem;
  unsigned char* const buf1 = smem + tile_bytes;
  unsigned char* const abuf = smem + 2 * tile_bytes;
  const size_t r0b = lm_region0_bytes();
  RowInfo& ri = *reinterpret_cast<RowInfo*>(smem + r0b);
  double* thead = reinterpret_cast<double*>(smem + r0b + sizeof(RowInfo));
  double* tmid = thead + kTailLdsHead;
  double* fold = reinterpret_cast<double*>(smem);
  tail_to_lds(*a.tail, thead, tmid, threadIdx.x, 256);
  const TailView tv{thead, tmid, a.tail->coef_b, kTailLdsStride, kTailLdsStride};
  const unsigned char* gsrc =
      PREC == EQTL_PREC_FP32 ? reinterpret_cast<const unsigned char*>(a.gc.f32 + a.gc.tile_f32(group, 0))
                             : reinterpret_cast<const unsigned char*>(a.gc.bf + a.gc.tile_bf(group, 0));
  // this wave's part of a row stage: chunk (rb, h) = 64 lanes x 16 B, so every
  // later ds_read_b128 of it is conflict-free
  const int wrow = wave * kRB * 2 * 1024;

  GeneAcc st;
  acc_init(st);
  dd asum{0.0, 0.0};
  for (int tile = t_lo; tile < t_hi; tile += t_step) {
    const int64_t row0 = static_cast<int64_t>(tile) * kTileRows;
    const int64_t rowA = row0 + wave * 32;
    __syncthreads();  // the previous tile's epilogue is done with ri and fold
    if (threadIdx.x < kTileRows) {
      const int t = threadIdx.x;
      const int64_t j = row0 + t;
      const bool in = j < a.m;
      ri.keep[t] = in ? a.keep[j] : 0;
      ri.af[t] = in ? a.af[j] : 0.0;
      ri.gtg[t] = in ? a.gtg[j] : 0.0;
      const bool ann = in && a.snp_chrom;
      ri.chrom[t] = ann ? a.snp_chrom[a.snp0 + j] : -1;
      ri.pos[t] = ann ? a.snp_pos[a.snp0 + j] : 0;
    }
    f32x4 acc[kRB];
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) acc[rb] = f32x4{0.f, 0.f, 0.f, 0.f};
    // SNP-row operand: lane l of row block rb holds row rowA + 16 rb + (l & 15),
    // samples 32 s + 8 (l >> 4) + [0..7] of k-step s
    const float* arow[kRB];
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb) arow[rb] = a.A + (rowA + rb * 16 + gl) * a.n_pad + kk;
    // prologue: gene tile of k-step 0, rows of the first stages
    if (wave < 2) lds_direct16(gsrc + wave * 1024 + lane * 16, buf0 + wave * 1024);
    for (int s = 0; s < kAS - 1 && s < nks; ++s) {
      const float* src[kRB];
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) src[rb] = arow[rb] + s * 32;
      wald_stage_rows(src, abuf + s * kAStageBytes + wrow, lane);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int s = 0; s < nks; ++s) {
      const int sr = s + kAS - 1;  // k-step whose rows are fetched now
      const float* src[kRB];
#pragma unroll
      for (int rb = 0; rb < kRB; ++rb) src[rb] = arow[rb] + sr * 32;
      unsigned char* const a_wr = sr < nks ? abuf + (sr % kAS) * kAStageBytes + wrow : nullptr;
      const unsigned char* const a_rd = abuf + (s % kAS) * kAStageBytes + wrow;
      unsigned char* const b_nxt = s + 1 < nks ? (((s + 1) & 1) ? buf1 : buf0) : nullptr;
      lm_kstep<PREC>(acc, a_rd, a_wr, src, (s & 1) ? buf1 : buf0, b_nxt,
                     gsrc + static_cast<size_t>(s + 1) * tile_bytes, wave, lane);
      // the next k-step's gene tile and rows must have landed (with three
      // stages the rows of k-step s+2, issued last, could stay in flight).
      // The barrier is a bare s_barrier: a workgroup release fence would
      // drain vmcnt to zero.
      if (kAS == 3 && s + 2 < nks) __builtin_amdgcn_s_waitcnt(kWaitVm4Lgkm0);
      else __builtin_amdgcn_s_waitcnt(0);
      lds_barrier();
    }
    // park gty per test in LDS (the last k-step's barrier has released the
    // staging buffers), then the FP64 epilogue, one test per iteration
#pragma unroll
    for (int rb = 0; rb < kRB; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) fold[(rb * 4 + r) * 256 + threadIdx.x] = acc[rb][r];
#pragma unroll 1
    for (int t = 0; t < kRB * 4; ++t) {
      const int rb = t >> 2, r = t & 3;
      const int jl = wave * 32 + rb * 16 + 4 * (lane >> 4) + r;
      const int64_t j = row0 + jl;
      const bool inwin = win_known && ri.chrom[jl] == gch && ri.pos[jl] >= gcs && ri.pos[jl] <= gce;
      const bool ok = gene_ok && j < a.m && ri.keep[jl] && (a.cis ? inwin : !inwin);
      double num = fold[t * 256 + threadIdx.x];
      const bool cand = refine_wanted<EQTL_MODEL_LM>(a, ok, 0.0, 0.0, num, gscal, ri.gtg[jl]);
      const double n64 = wave_refine(a, cand, j, gene);
      if (cand) num = n64;
      if (ok) wald_test<EQTL_MODEL_LM>(a, ri, jl, j, gene, gscal, 0.0, 0.0, num, tv, st, asum);
    }
  }
  st.acat_hi = asum.hi;
  st.acat_lo = asum.lo;
  // reduce the 16 partial states of each gene; the staging buffers are free now
  GeneAcc* red = reinterpret_cast<GeneAcc*>(smem);
  __syncthreads();
  red[(wave * 4 + (lane >> 4)) * 16 + gl] = st;
  __syncthreads();
  if (threadIdx.x < 16) {
    GeneAcc t = red[threadIdx.x];
    for (int q = 1; q < 16; ++q) acc_merge(t, red[q * 16 + threadIdx.x]);
    const size_t slot = a.work ? static_cast<size_t>(split) * 16 : static_cast<size_t>(split) * a.G_pad + group * 16;
    a.partial[slot + threadIdx.x] = t;
  }
}

template <int PREC>
// Three waves per SIMD (registers capped at 168, two row stages so that three
// workgroups fit in LDS): the FP64 epilogue bounds this kernel and needs the
// extra waves to hide its latency (measured 10% faster than two waves with
// three stages).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_wald_lm(WaldArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  if (a.work) {  // cis: persistent workgroups over the item queue
    __shared__ int slot;
    for (int it = next_item(a, &slot); it >= 0; it = next_item(a, &slot))
      lm_item<PREC>(a, smem, a.work[3 * it], it, a.work[3 * it + 1], a.work[3 * it + 2], 1);
    return;
  }
  // XCD-aware mapping: workgroups are dealt round-robin to the 8 XCDs; the
  // n_splits consecutive workgroups of one XCD share a gene group, so the
  // group's column blocks are served from that XCD's L2.
  const int wg = blockIdx.x;
  const int xcd = wg & 7;
  const int jx = wg >> 3;
  const int split = jx % a.n_splits;
  lm_item<PREC>(a, smem, (jx / a.n_splits) * 8 + xcd, split, split, (a.n_tiles + 1) >> 1, a.n_splits);
}

// ---------------------------------------------------------------------------
// FP64 polish of LMM / GLM tests with |t| > tau_hi (strong signals, p far
// below any reporting threshold, where the contraction's relative error in S
// is amplified by |t|) or with S < kappa gg (covariates that explain most of
// the SNP, where S = gg - |W^T h|^2 cancels): test_lmm_gtil recomputed from
// the FP64 dosage, one workgroup per test:
//   g~ = Q^T g_c,  gg = sum d g~^2,  num = (s o r)^T g~,  a = X~^T (d o g~),
//   W^T h = Lc^-1 a (Lc Lc^T = X~^T D X~),  S = gg - |Lc^-1 a|^2,
// then beta/se/stat/p (continued-fraction tail) and r^2 as in the kernel.
// The Wald kernel only counted these tests as tested; here each one adds its
// ACAT term, its n_sig count and its hit record (appended to the block's
// list), and the k_pbest_* kernels merge the polished p into the gene's best.
// ---------------------------------------------------------------------------
struct PolishArgs {
  const PolishItem* items;
  int64_t count;
  const double* A64;
  int32_t n_pad, n, p;
  const double* Q;      // n x n col-major
  const double* lam;
  const double* delta;
  const double* U;      // G x n: s o r in the rotated space
  const double* Xt;     // n x p col-major (rotated covariates)
  const double* Lc;     // G x p x p, lower, col-major
  const double* w;      // GLM: G x n null working weights
  const double* mean;   // GLM: block-local SNP centring means
  const double* gene_scal;
  const double* af;
  double df, lbeta;
  int64_t snp0;
  double* gt;           // count x n scratch
  double pthr;
  eqtl_hit* hits;       // the block's hit list, appended to at hit_count[0]
  unsigned long long* hit_count;
  int64_t hit_cap;
  eqtl_hit* out;        // count
  GeneAcc* state;       // running per-gene state (ACAT term of the polished p)
};

// GLM (test_glm_nb, model_glm.cpp:100-132): the same quantities without the
// rotation — weights w_k, covariates X, residual r_k — and U = num + mean * sum r,
// info = S, two-sided normal p.
// Q^T g_c for each polished LMM test, spread over the chip: grid (items,
// ceil(n / 64)); wave w of block (item, cb) forms columns cb * 64 + 16 w ..
// + 15, four at a time (independent accumulators keep the column loads in
// flight), lanes striding the samples and a butterfly sum per column.
__global__ __launch_bounds__(256) void k_polish_rotate(PolishArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const PolishItem it = a.items[blockIdx.x];
  const int n = a.n;
  const double* gc = a.A64 + static_cast<size_t>(it.j) * a.n_pad;
  double* gr = a.gt + static_cast<size_t>(blockIdx.x) * n;
  const int cbeg = blockIdx.y * 64 + wave * 16;
  const int cend = min(cbeg + 16, n);
  for (int c0 = cbeg; c0 < cend; c0 += 4) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = lane; i < n; i += 64) {
      const double g = gc[i];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (c0 + q < cend) s[q] = fma(a.Q[static_cast<size_t>(c0 + q) * n + i], g, s[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double v = s[q];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0 && c0 + q < cend) gr[c0 + q] = v;
    }
  }
}

template <int MODEL>
__global__ __launch_bounds__(256) void k_polish(PolishArgs a) {
  __shared__ double red[4];
  __shared__ double va[34];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const PolishItem it = a.items[blockIdx.x];
  const int n = a.n, p = a.p, k = it.gene;
  // LMM: the rotated dosage Q^T g_c comes from k_polish_rotate
  const double* gt = MODEL == EQTL_MODEL_LMM ? a.gt + static_cast<size_t>(blockIdx.x) * n
                                             : a.A64 + static_cast<size_t>(it.j) * a.n_pad;
  const double dk = MODEL == EQTL_MODEL_LMM ? a.delta[k] : 0.0;
  const double* u = a.U + static_cast<size_t>(k) * n;
  const double* wk = MODEL == EQTL_MODEL_GLM ? a.w + static_cast<size_t>(k) * n : nullptr;
  double part[34];
#pragma unroll
  for (int c = 0; c < 34; ++c) part[c] = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    double d;
    if constexpr (MODEL == EQTL_MODEL_GLM) {
      d = wk[i];
    } else {
      double v = dk * a.lam[i] + 1.0;
      if (v < 1e-12) v = 1e-12;
      d = 1.0 / v;
    }
    const double g = gt[i], dg = d * g;
    part[0] += dg * g;
    part[1] += u[i] * g;
#pragma unroll
    for (int c = 0; c < 32; ++c)
      if (c < p) part[2 + c] += a.Xt[static_cast<size_t>(c) * n + i] * dg;
  }
#pragma unroll
  for (int c = 0; c < 34; ++c) {
    if (c < p + 2) {
      double s = part[c];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      __syncthreads();
      if (lane == 0) red[wave] = s;
      __syncthreads();
      if (threadIdx.x == 0) va[c] = red[0] + red[1] + red[2] + red[3];
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double* L = a.Lc + static_cast<size_t>(k) * p * p;
  double w[32];
  double ss = 0;
  for (int i = 0; i < p; ++i) {
    double t = va[2 + i];
    for (int c = 0; c < i; ++c) t -= L[c * p + i] * w[c];
    w[i] = L[i * p + i] > 0 ? t / L[i * p + i] : 0.0;  // dependent covariate: no direction
    ss += w[i] * w[i];
  }
  const double gg = va[0], num = va[1], S = gg - ss, gscal = a.gene_scal[k], df = a.df;
  double beta = 0, se = 0, stat = 0, pv = 1.0, r2 = 0;
  if constexpr (MODEL == EQTL_MODEL_GLM) {
    if (!(S < 1e-12)) {
      const double U = num + a.mean[it.j] * gscal;
      const double rs = sqrt(S);
      beta = U / S;
      se = 1.0 / rs;
      stat = U / rs;
      pv = isfinite(stat) ? fmin(1.0, fmax(0.0, erfc(fabs(stat) * 0.70710678118654752440))) : 1.0;
    }
  } else if (!(gg < 1e-12) && !(S <= 1e-15)) {
    beta = num / S;
    double qq = gscal - S * beta * beta;
    if (qq < 0) qq = 0;
    const double sigma2 = df > 0 ? qq / df : 1.0;
    se = sqrt(fmax(sigma2 / S, 0.0));
    stat = se > 0 ? beta / se : 0.0;
    pv = p_from_t_dev(stat, df, a.lbeta);
    r2 = gscal > 1e-15 ? fmax(0.0, 1.0 - qq / gscal) : 0.0;
  }
  eqtl_hit h;
  h.snp = a.snp0 + it.j; h.gene = k; h.pad = 0;
  h.af = a.af[it.j]; h.beta = beta; h.se = se; h.stat = stat; h.p = pv; h.r2 = r2;
  a.out[blockIdx.x] = h;
  double term;
  if (acat_accept(pv, &term)) {
    atomicAdd(&a.state[k].acat_pol, term);
    atomicAdd(&a.state[k].used, 1);
  }
  if (pv <= a.pthr) {
    atomicAdd(&a.state[k].n_sig, 1);
    const unsigned long long idx = atomicAdd(a.hit_count, 1ull);
    if (static_cast<int64_t>(idx) < a.hit_cap) a.hits[idx] = h;
  }
}

// Best entry over the kernel's best and the polished tests, lexicographic in
// (p, SNP index): key[2g] = min p bits (non-negative doubles order as their
// bit patterns), key[2g + 1] = smallest SNP at that p; the winning polished
// test (if any) writes the gene's best fields.
__device__ inline unsigned long long pbits(double p) { return static_cast<unsigned long long>(__double_as_longlong(p)); }
__global__ void k_pbest_init(const GeneAcc* state, unsigned long long* key, int G) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  key[2 * g] = pbits(state[g].best_p);
  key[2 * g + 1] = state[g].best_snp >= 0 ? static_cast<unsigned long long>(state[g].best_snp) : ~0ull;
}
__global__ void k_pbest_p(const eqtl_hit* out, int64_t count, unsigned long long* key) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) atomicMin(&key[2 * out[i].gene], pbits(out[i].p));
}
__global__ void k_pbest_state_snp(const GeneAcc* state, unsigned long long* key, int G) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  if (key[2 * g] != pbits(state[g].best_p)) key[2 * g + 1] = ~0ull;  // the kernel's best lost on p
}
__global__ void k_pbest_snp(const eqtl_hit* out, int64_t count, unsigned long long* key) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const eqtl_hit h = out[i];
  if (key[2 * h.gene] == pbits(h.p)) atomicMin(&key[2 * h.gene + 1], static_cast<unsigned long long>(h.snp));
}
__global__ void k_pbest_write(const eqtl_hit* out, int64_t count, const unsigned long long* key, GeneAcc* state) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const eqtl_hit h = out[i];
  if (key[2 * h.gene] != pbits(h.p) || key[2 * h.gene + 1] != static_cast<unsigned long long>(h.snp)) return;
  GeneAcc& s = state[h.gene];
  s.best_p = h.p; s.best_snp = h.snp; s.best_af = h.af; s.best_beta = h.beta;
  s.best_se = h.se; s.best_stat = h.stat; s.best_r2 = h.r2;
}

// ---------------------------------------------------------------------------
// FP64 minimum p over a gene's window, for the exact permutation counts
// (p_emp_count, stats_extra.cpp:39-46, compares T = -log10(max(p, 1e-300))
// of each permutation with the observed one; a permutation whose split-
// precision T lies near the observed T is decided again here, from the FP64
// tests the reference computes: test_lm, model_lm.cpp:32-57; test_lmm,
// model_lmm.cpp:160-223, in the bordered-Schur form of k_polish).
// A task (gene-side column e, SNP range of its window) is cut into pieces of
// kExactPiece SNPs, one workgroup per piece; each wave takes every fourth SNP
// of the piece and runs its test in FP64 over the exact rows R (LM: the
// residualised dosage g'; LMM: Q^T g'), the lanes striding the samples; the
// window rule (chromosome, [cs, ce]) and the SNP's keep flag decide which
// SNPs are tested. The tail probability comes from the scan's t-tail table
// (2e-12 relative to the continued fraction). out[task] receives the smallest
// p as an atomic minimum of its bit pattern (non-negative doubles order as
// their bits); the caller sets every slot to the bits of 2.0 first.
// ---------------------------------------------------------------------------
constexpr int kExactPiece = 64;

struct ExactTask {
  int32_t e;         // gene-side column (gene or pseudo-gene)
  int32_t chrom;     // window chromosome id
  int64_t jlo, jhi;  // global SNP index range holding the window
  int64_t cs, ce;    // window bounds (positions)
};

struct ExactPieceRef {
  int32_t task;
  int32_t pad;
  int64_t jlo, jhi;  // this piece's SNP range
};

struct ExactArgs {
  const ExactTask* tasks;
  const ExactPieceRef* pieces;
  const double* R;        // exact rows, n doubles per SNP, SNP r0 + q at R + q * n
  int64_t r0;             // global index of R's first SNP
  const uint8_t* keep;    // per SNP of R
  const double* gtg;      // LM: ||g'||^2 per SNP of R
  const int32_t* snp_chrom;  // global annotation
  const int64_t* snp_pos;
  int32_t n, p;
  const double* lam;      // LMM
  const double* delta;    // LMM, per gene-side column
  const double* U;        // LMM: s o r (rotated space); LM: y_s; n per column
  const double* Xt;       // LMM: Q^T X, n x p col-major
  const double* Lc;       // LMM: p x p per column
  const double* gene_scal;  // LMM rss_null, LM yty
  const TTail* tail;
  double df;
  unsigned long long* out;  // per task: bits of the minimum p
};

template <int MODEL>
__global__ __launch_bounds__(256) void k_exact_minp(ExactArgs a) {
  constexpr int kP = 32;
  __shared__ double Ls[kP * kP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const ExactPieceRef pc = a.pieces[blockIdx.x];
  const ExactTask tk = a.tasks[pc.task];
  const int n = a.n, p = a.p, e = tk.e;
  const double* u = a.U + static_cast<size_t>(e) * n;
  const double gscal = a.gene_scal[e];
  const double dk = MODEL == EQTL_MODEL_LMM ? a.delta[e] : 0.0;
  if constexpr (MODEL == EQTL_MODEL_LMM) {
    const double* L = a.Lc + static_cast<size_t>(e) * p * p;
    for (int i = threadIdx.x; i < p * p; i += 256) Ls[i] = L[i];
    __syncthreads();
  }
  double best = 2.0;
  for (int64_t j = pc.jlo + wave; j < pc.jhi; j += 4) {
    const int64_t q = j - a.r0;
    if (!a.keep[q] || a.snp_chrom[j] != tk.chrom || a.snp_pos[j] < tk.cs || a.snp_pos[j] > tk.ce) continue;
    const double* g = a.R + static_cast<size_t>(q) * n;
    double pv = 1.0;
    if constexpr (MODEL == EQTL_MODEL_LM) {
      double num = 0.0;
      for (int i = lane; i < n; i += 64) num = fma(u[i], g[i], num);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) num += __shfl_xor(num, o);
      const double gtg = a.gtg[q];
      if (!(gtg < 1e-12) && a.df > 0) {
        const double beta = num / gtg;
        const double rss = gscal - beta * num;
        const double s2 = fmax(rss / a.df, 0.0);
        const double se = sqrt(s2 / gtg);
        const double t = se > 0 ? beta / se : 0.0;
        pv = p_from_t_tab2(&a.tail->df, &a.tail->u_m, a.tail->coef_b, t);
      }
    } else {
      double gg = 0.0, num = 0.0, av[kP];
#pragma unroll
      for (int c = 0; c < kP; ++c) av[c] = 0.0;
      for (int i = lane; i < n; i += 64) {
        double v = dk * a.lam[i] + 1.0;
        if (v < 1e-12) v = 1e-12;
        const double gi = g[i], dg = gi / v;
        gg = fma(dg, gi, gg);
        num = fma(u[i], gi, num);
#pragma unroll
        for (int c = 0; c < kP; ++c)
          if (c < p) av[c] = fma(a.Xt[static_cast<size_t>(c) * n + i], dg, av[c]);
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        gg += __shfl_xor(gg, o);
        num += __shfl_xor(num, o);
#pragma unroll
        for (int c = 0; c < kP; ++c)
          if (c < p) av[c] += __shfl_xor(av[c], o);
      }
      // W^T h = Lc^-1 a (forward substitution), S = gg - |W^T h|^2
      double ss = 0.0;
#pragma unroll
      for (int i = 0; i < kP; ++i) {
        if (i < p) {
          double t = av[i];
#pragma unroll
          for (int c = 0; c < kP; ++c)
            if (c < i) t -= Ls[c * p + i] * av[c];
          av[i] = Ls[i * p + i] > 0 ? t / Ls[i * p + i] : 0.0;  // dependent covariate: no direction
          ss += av[i] * av[i];
        }
      }
      const double S = gg - ss;
      if (!(gg < 1e-12) && !(S <= 1e-15)) {
        const double beta = num / S;
        double qq = gscal - S * beta * beta;
        if (qq < 0) qq = 0;
        const double sigma2 = a.df > 0 ? qq / a.df : 1.0;
        const double se = sqrt(fmax(sigma2 / S, 0.0));
        const double t = se > 0 ? beta / se : 0.0;
        pv = p_from_t_tab2(&a.tail->df, &a.tail->u_m, a.tail->coef_b, t);
      }
    }
    if (isfinite(pv)) best = fmin(best, pv);
  }
  if (lane == 0 && best < 2.0) atomicMin(&a.out[pc.task], static_cast<unsigned long long>(__double_as_longlong(best)));
}

// ---------------------------------------------------------------------------
// Self-test of the hardware conventions the kernels rely on (one workgroup):
//  bit 0: bf16 16x16x32 MFMA operand/result lane layout
//  bit 1: fp32 16x16x4 MFMA layout with the 8-k-slot grouping of mfma_f32x8
//  bit 2: direct global→LDS 16-byte loads land at wave base + 16 * lane
//  bit 3: fp64 16x16x4 with the grouping of mfma_f64x8 (result rows (l>>4)+4r)
// Small integers make every product and sum exact, so equality is exact.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_selftest(int* fail) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[4096];
  __shared__ int src_pattern[1024];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // A[r][k] = (r + 2k) % 5 - 2, B[k][c] = (3c + k) % 7 - 3, C = A B (16 x 16, K = 32)
  auto Av = [](int r, int k) { return static_cast<float>((r + 2 * k) % 5 - 2); };
  auto Bv = [](int k, int c) { return static_cast<float>((3 * c + k) % 7 - 3); };
  if (wave == 0) {
    bf16x8 a8, b8;
    f32x8 af, bfv;
    for (int t = 0; t < 8; ++t) {
      const int k = 8 * (lane >> 4) + t;
      a8[t] = static_cast<__bf16>(Av(lane & 15, k));
      b8[t] = static_cast<__bf16>(Bv(k, lane & 15));
      af.v[t] = Av(lane & 15, k);
      bfv.v[t] = Bv(k, lane & 15);
    }
    f64x8 ad, bd;
    for (int t = 0; t < 8; ++t) { ad.v[t] = af.v[t]; bd.v[t] = bfv.v[t]; }
    f32x4 c1 = mfma_bf16(a8, b8, f32x4{0, 0, 0, 0});
    f32x4 c2 = mfma_f32x8(af, bfv, f32x4{0, 0, 0, 0});
    f64x4 c3 = mfma_f64x8(ad, bd, f64x4{0, 0, 0, 0});
    int bad1 = 0, bad2 = 0, bad3 = 0;
    for (int r = 0; r < 4; ++r) {
      const int row = 4 * (lane >> 4) + r, col = lane & 15;
      float e = 0;
      for (int k = 0; k < 32; ++k) e += Av(row, k) * Bv(k, col);
      bad1 |= c1[r] != e;
      bad2 |= c2[r] != e;
      double e3 = 0;
      for (int k = 0; k < 32; ++k) e3 += Av((lane >> 4) + 4 * r, k) * Bv(k, col);
      bad3 |= c3[r] != e3;
    }
    if (bad1) atomicOr(fail, 1);
    if (bad2) atomicOr(fail, 2);
    if (bad3) atomicOr(fail, 8);
  }
  // LDS-direct layout: global int pattern → LDS via 16-byte per-lane loads
  int* g = reinterpret_cast<int*>(fail + 16);  // caller provides 1024 ints after the flag area
  for (int i = threadIdx.x; i < 1024; i += 256) src_pattern[i] = 0;
  __syncthreads();
  for (int c = wave; c < 4; c += 4) lds_direct16(g + c * 256 + lane * 4, lds + c * 1024);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  const int* li = reinterpret_cast<const int*>(lds);
  for (int i = threadIdx.x; i < 1024; i += 256)
    if (li[i] != g[i]) atomicOr(fail, 4);
}

// Fold one block's partials into the running per-gene state.
__global__ void k_merge(GeneAcc* state, const GeneAcc* partial, int G, int G_pad, int n_splits) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  GeneAcc t = state[g];
  for (int s = 0; s < n_splits; ++s) acc_merge(t, partial[static_cast<size_t>(s) * G_pad + g]);
  state[g] = t;
}

// Cis: fold each gene's work-item partials (items sorted by group; item_first
// has n_groups + 1 entries).
__global__ void k_merge_items(GeneAcc* state, const GeneAcc* partial, const int32_t* item_first, int G) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int grp = g >> 4;
  GeneAcc t = state[g];
  for (int it = item_first[grp]; it < item_first[grp + 1]; ++it) acc_merge(t, partial[static_cast<size_t>(it) * 16 + (g & 15)]);
  state[g] = t;
}

__global__ void k_state_init(GeneAcc* state, int G) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G) acc_init(state[g]);
}

__global__ void k_finalize(const GeneAcc* state, eqtl_gene_state* out, int G) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const GeneAcc a = state[g];
  eqtl_gene_state o;
  o.acat_p = acat_p_final(dd_add_d(dd{a.acat_hi, a.acat_lo}, a.acat_pol), a.used);
  o.best_p = a.best_p;
  o.best_snp = a.best_snp;
  o.best_af = a.best_af; o.best_beta = a.best_beta; o.best_se = a.best_se;
  o.best_stat = a.best_stat; o.best_r2 = a.best_r2;
  o.n_tested = a.n_tested;
  o.n_sig = a.n_sig;
  out[g] = o;
}

// Covariate residualisation of the decoded block (LM and LMM scans), FP64,
// one wave per SNP row: g' = g_c - Wx (Wx^T g_c), written over the FP64 and
// fp32 rows, and gtg = ||g'||^2 (residualize_g, scan_trans_lm.cpp:11-19, with
// an orthonormal basis Wx of the covariates; LM uses gtg as g_s^T g_s). Both
// Wald statistics are invariant to adding covariate combinations to g, so
// the tests are unchanged; but the contraction's rounding then scales with
// ||g'|| instead of ||g||, so a SNP that the covariates explain almost fully
// (a conditioning SNP in LD, genotype PCs) keeps its relative accuracy in S
// and in the numerator.
// kResRows SNP rows per wave: each covariate value loaded serves all of them
// (one row per wave re-read the n x p basis from L2 for every SNP), and the
// basis is staged in LDS 64 samples at a time for the workgroup's four waves
// (16 rows per load of the basis from L2). Per row the sums and FMAs run in
// the same order as for one row alone.
constexpr int kResRows = 4;
__global__ __launch_bounds__(256) void k_residualize(double* A64, float* A32, int64_t m_pad, int n, int n_pad,
                                                     const double* Wx, int p, double* gtg) {
  constexpr int kPCov = 32;  // covariate columns at most (intercept included)
  constexpr int kCh = 8;     // columns per pass (independent reductions in flight)
  __shared__ double cfs[4][kResRows][kPCov];  // per wave: the rows' coefficients Wx^T g_c
  __shared__ double wsh[kPCov][64];           // the basis at 64 samples
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // every wave of the workgroup has rows (m_pad is a multiple of 4 * kResRows
  // and the grid covers it exactly), so the workgroup barriers below are
  // reached by all of its waves
  const int64_t j0 = (static_cast<int64_t>(blockIdx.x) * 4 + w) * kResRows;
  if (p > kPCov) return;
  double* g[kResRows];
  float* g32[kResRows];
#pragma unroll
  for (int r = 0; r < kResRows; ++r) {
    g[r] = A64 + (j0 + r) * n_pad;
    g32[r] = A32 + (j0 + r) * n_pad;
  }
  // wsh[c][t] = Wx[c][i0 + t] for columns c0 .. c0 + nc - 1 (zero past n)
  auto stage = [&](int c0, int nc, int i0) {
    __syncthreads();  // every wave is done with the previous chunk
    for (int q = threadIdx.x; q < nc * 64; q += 256) {
      const int c = c0 + (q >> 6), i = i0 + (q & 63);
      wsh[c][q & 63] = i < n ? Wx[static_cast<size_t>(c) * n + i] : 0.0;
    }
    __syncthreads();
  };
#pragma unroll
  for (int c0 = 0; c0 < kPCov; c0 += kCh) {
    if (c0 >= p) break;
    const int nc = min(kCh, p - c0);
    double sm[kResRows][kCh];
#pragma unroll
    for (int r = 0; r < kResRows; ++r)
#pragma unroll
      for (int u = 0; u < kCh; ++u) sm[r][u] = 0.0;
    for (int i0 = 0; i0 < n; i0 += 64) {
      stage(c0, nc, i0);
      const int i = i0 + lane;
      if (i < n) {
        double gi[kResRows];
#pragma unroll
        for (int r = 0; r < kResRows; ++r) gi[r] = g[r][i];
#pragma unroll
        for (int u = 0; u < kCh; ++u)
          if (c0 + u < p) {
            const double wv = wsh[c0 + u][lane];
#pragma unroll
            for (int r = 0; r < kResRows; ++r) sm[r][u] = fma(wv, gi[r], sm[r][u]);
          }
      }
    }
#pragma unroll
    for (int r = 0; r < kResRows; ++r)
#pragma unroll
      for (int u = 0; u < kCh; ++u) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sm[r][u] += __shfl_xor(sm[r][u], o);
        if (lane == 0 && c0 + u < kPCov) cfs[w][r][c0 + u] = sm[r][u];
      }
  }
  double ss[kResRows];
#pragma unroll
  for (int r = 0; r < kResRows; ++r) ss[r] = 0.0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    stage(0, p, i0);  // its first barrier also publishes cfs
    const int i = i0 + lane;
    if (i < n) {
      double v[kResRows];
#pragma unroll
      for (int r = 0; r < kResRows; ++r) v[r] = g[r][i];
      for (int c = 0; c < p; ++c) {
        const double wv = wsh[c][lane];
#pragma unroll
        for (int r = 0; r < kResRows; ++r) v[r] = fma(-wv, cfs[w][r][c], v[r]);
      }
#pragma unroll
      for (int r = 0; r < kResRows; ++r) {
        g[r][i] = v[r];
        g32[r][i] = static_cast<float>(v[r]);
        ss[r] = fma(v[r], v[r], ss[r]);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kResRows; ++r) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ss[r] += __shfl_xor(ss[r], o);
    if (lane == 0 && gtg) gtg[j0 + r] = ss[r];
  }
}

// ---------------------------------------------------------------------------
// Synthetic PLINK rows (benchmark input generated in HBM). Counter-based hash
// so SNP j depends only on (seed, j).
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t mix64(uint64_t x) {
  x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27; x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}
__host__ __device__ inline double u01(uint64_t h) { return (h >> 11) * (1.0 / 9007199254740992.0); }

__global__ void k_synth_bed(uint8_t* rows, int64_t m, int n, int bps, uint64_t seed, double maf_lo,
                            double maf_hi, double miss) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= m * bps) return;
  const int64_t j = idx / bps;
  const int byte = static_cast<int>(idx % bps);
  const uint64_t hs = mix64(seed * 0x9E3779B97F4A7C15ull + static_cast<uint64_t>(j));
  const double maf = maf_lo + (maf_hi - maf_lo) * u01(hs);
  const double f = (mix64(hs ^ 0x5555) & 1) ? 1.0 - maf : maf;
  uint8_t out = 0;
  for (int q = 0; q < 4; ++q) {
    const int i = byte * 4 + q;
    int code = 1;  // padding / missing
    if (i < n) {
      const uint64_t h = mix64(hs + 0x632BE59BD9B4E019ull * static_cast<uint64_t>(i + 1));
      const double um = u01(h), u1 = u01(mix64(h ^ 0xA5A5A5A5ull)), u2 = u01(mix64(h ^ 0x3C3C3C3Cull));
      if (um < miss) {
        code = 1;
      } else {
        const int g = (u1 < f) + (u2 < f);
        code = g == 2 ? 0 : (g == 1 ? 2 : 3);
      }
    }
    out |= static_cast<uint8_t>(code << (2 * q));
  }
  rows[idx] = out;
}

}  // namespace eqtl_amd
