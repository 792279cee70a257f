// K2: GRM (compute_grm_from, grm.cpp:112-224).
//   k_grm_stats / k_grm_scan / k_grm_emit — decode each genotype row, apply
//       the GRM's SNP rules (the source's filters, then p = mean/2 in
//       [1e-8, 1-1e-8], sd >= 1e-12) and route each SNP: rows with every call
//       present and integral go, as 2-bit dosage codes, to the exact buffer
//       of the grouped int8 Gram; the others are written as z = (g - mean)/sd
//       (missing → 0) in fp64 with their 0/1 missing indicator in bf16,
//       sample-major (SNP-contiguous) for the MFMA operands, counting each
//       sample's missing calls;
//   k_grm_hist / k_grm_groups / k_grm_scatter / k_grm_expand / k_grm_i8 —
//       the exact SNPs grouped by allele count (one weight per group): the
//       int8 Gram of each group (exact int32) added into FP64 times the
//       group's weight, with the centring as FP64 rank-1 terms (u, C);
//   k_grm_syrk — lower block-triangle of Z^T Z (fp64 MFMA, LDS-staged
//       128 x 128 tiles), accumulated into an FP64 tile array;
//   k_grm_joint — the same triangle of M^T M (joint missing counts, bf16
//       MFMA on 0/1, exact), when the block has a missing call;
//   k_grm_gram — both on 64 x 64 tiles for small n, where 128-wide tiles
//       would leave most CUs idle.
//   k_grm_finish — (A - u_i - u_j + C) over the per-pair GCTA divisor
//       m - miss_i - miss_j + joint_ij (diagonal m - miss_i), symmetric K and
//       packed N.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "geno_rows.hpp"
#include "mfma.hpp"

namespace eqtl_amd {

// Per-SNP class of a GRM pass (k_grm_stats):
//   0 skipped (the source's rules or the GRM's p / sd rules),
//   1 exact: every call present and integral (0/1/2) — goes to the grouped
//     int8 Gram (k_grm_i8), exact in integer arithmetic,
//   2 FP64 path, no missing call; 3 FP64 path with a missing call.
constexpr int8_t kGrmSkip = 0, kGrmExact = 1, kGrmFloat = 2, kGrmFloatMiss = 3;

// IDENT: BED rows whose analysis samples are the file's columns in order
// (sample_col[i] == i), read a byte (four calls) at a time.
template <int FMT, bool IDENT>
__global__ __launch_bounds__(256) void k_grm_stats(const uint8_t* rows, int64_t m, int bps, int n,
                                                   const int32_t* sample_col, int impute, double max_miss,
                                                   double maf_min, int exact_ok, int8_t* cls, int32_t* ac,
                                                   double* mean_out, double* sd_out) {
  const int lane = threadIdx.x & 63;
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (j >= m) return;
  bool use = true, integral = true;
  int n_miss = 0;
  double sum = 0.0;  // integer-valued for BED rows (exact)
  const uint8_t* row = rows + j * bps;
  if constexpr (IDENT) {
    static_assert(FMT == EQTL_GENO_BED, "byte-wise path reads PLINK-1 codes");
    int isum = 0;
    for (int b = lane; 4 * b < n; b += 64) {
      const uint32_t byte = row[b];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t c = (byte >> (2 * k)) & 3u;
        if (4 * b + k >= n) break;
        if (c == 1u) ++n_miss;
        else isum += 2 - static_cast<int>(c >> 1) - static_cast<int>(c & 1u);  // {00: 2, 10: 1, 11: 0}
      }
    }
    sum = static_cast<double>(isum);
  } else {
    for (int i = lane; i < n; i += 64) {
      bool miss;
      const double g = geno_at<FMT>(row, sample_col[i], miss);
      if (miss) { ++n_miss; continue; }
      sum += g;
      if constexpr (FMT != EQTL_GENO_BED) integral = integral && g == rint(g) && g >= 0.0 && g <= 2.0;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    n_miss += __shfl_xor(n_miss, o);
    sum += __shfl_xor(sum, o);
  }
  if constexpr (FMT != EQTL_GENO_BED) integral = __all(integral);
  double mu = 0, mean = 0, sd = 1;
  const int n_ok = n - n_miss;
  if (n_ok == 0) use = false;
  // the source's rules (decode_row) for BED rows; DOSE32 rows carry them
  if constexpr (FMT == EQTL_GENO_BED) {
    if (use && static_cast<double>(n_miss) / n > max_miss + 1e-15) use = false;
    if (use && !impute && n_miss > 0) use = false;
  }
  if (use) {
    mu = sum / n_ok;
    if constexpr (FMT == EQTL_GENO_BED) {
      double af = fmin(1.0, fmax(0.0, mu / 2.0));
      if (af < 1e-12 || af > 1.0 - 1e-12) use = false;
      if (maf_min > 0.0 && (af < maf_min || af > 1.0 - maf_min)) use = false;
    }
  }
  if (use) {
    mean = (sum + n_miss * mu) / n;
    const double pf = mean / 2.0;
    if (pf < 1e-8 || pf > 1.0 - 1e-8) use = false;
    sd = sqrt(2.0 * pf * (1.0 - pf));
    if (sd < 1e-12) use = false;
  }
  if (lane == 0) {
    int8_t k = kGrmSkip;
    if (use) k = n_miss > 0 ? kGrmFloatMiss : (exact_ok && integral ? kGrmExact : kGrmFloat);
    cls[j] = k;
    ac[j] = static_cast<int32_t>(sum);  // the allele count for class 1 (exact integer sum)
    mean_out[j] = mean;
    sd_out[j] = sd;
  }
}

// Positions of a block's SNPs (one workgroup): pos[j] = rank of j among the
// exact SNPs (class 1) or among the FP64-path SNPs (classes 2, 3), in SNP
// order, so the FP64 Gram sums in a fixed order. counts: [used, FP64-path
// with a missing call, FP64-path, exact].
__global__ __launch_bounds__(1024) void k_grm_scan(const int8_t* cls, int64_t m, int32_t* pos, int64_t* counts) {
  __shared__ int32_t s1[1024], s2[1024], s3[1024];
  const int t = threadIdx.x;
  const int64_t per = (m + 1023) / 1024;
  const int64_t a = min(m, t * per), b = min(m, a + per);
  int c1 = 0, c2 = 0, c3 = 0;
  for (int64_t j = a; j < b; ++j) {
    const int8_t k = cls[j];
    c1 += k == kGrmExact;
    c2 += k >= kGrmFloat;
    c3 += k == kGrmFloatMiss;
  }
  s1[t] = c1; s2[t] = c2; s3[t] = c3;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
    const int v1 = t >= off ? s1[t - off] : 0, v2 = t >= off ? s2[t - off] : 0, v3 = t >= off ? s3[t - off] : 0;
    __syncthreads();
    s1[t] += v1; s2[t] += v2; s3[t] += v3;
    __syncthreads();
  }
  int p1 = s1[t] - c1, p2 = s2[t] - c2;
  for (int64_t j = a; j < b; ++j) {
    const int8_t k = cls[j];
    if (k == kGrmExact) pos[j] = p1++;
    else if (k >= kGrmFloat) pos[j] = p2++;
    else pos[j] = -1;
  }
  if (t == 1023) {
    counts[0] = s1[t] + s2[t];
    counts[1] = s3[t];
    counts[2] = s2[t];
    counts[3] = s1[t];
  }
}

// Writes a classified block: FP64-path SNPs as z = (g - mean)/sd (missing → 0)
// in fp64 and their 0/1 missing indicator in bf16, sample-major (column
// pos[j] of a zstride-wide matrix) for the MFMA operands, counting each
// sample's missing calls; columns counts[2] .. zstride - 1 are zeroed by the
// waves past m. Exact SNPs as 2-bit minor-allele dosage codes in
// analysis-sample order (rowq dwords per row, 16 samples per dword, padding
// samples 0) at row ex_base + pos[j] of the exact buffer, with their minor
// allele count (1 .. n) as key.
template <int FMT, bool IDENT>
__global__ __launch_bounds__(256) void k_grm_emit(const uint8_t* rows, int64_t m, int bps, int n, int n_pad,
                                                  const int32_t* sample_col, const int8_t* cls, const int32_t* pos,
                                                  const int32_t* ac, const int64_t* counts, const double* mean_in,
                                                  const double* sd_in, double* z64, __bf16* mis, int64_t zstride,
                                                  int32_t* miss_cnt, uint32_t* ex_rows, int32_t* ex_key,
                                                  int64_t ex_base, int rowq) {
  const int lane = threadIdx.x & 63;
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (j >= m) {
    const int64_t col = counts[2] + (j - m);
    if (col < zstride)
      for (int i = lane; i < n_pad; i += 64) {
        z64[static_cast<size_t>(i) * zstride + col] = 0.0;
        mis[static_cast<size_t>(i) * zstride + col] = static_cast<__bf16>(0.f);
      }
    return;
  }
  const int8_t k = cls[j];
  if (k == kGrmSkip) return;
  const uint8_t* row = rows + j * bps;
  if (k == kGrmExact) {
    // minor-allele coding: (g_i - mu)(g_j - mu) is unchanged by g -> 2 - g,
    // mu -> 2 - mu, and the uncentred Gram of the minor allele cancels least
    const int64_t r = ex_base + pos[j];
    const int32_t a = ac[j];
    const bool flip = a > n;
    for (int d = lane; d < rowq; d += 64) {
      uint32_t w = 0;
      if constexpr (IDENT) {
        // 16 PLINK-1 codes (no missing call in an exact row) -> dosage codes:
        // 00 -> 2, 10 -> 1, 11 -> 0 (flipped: 0, 1, 2), bit-parallel
        uint32_t in = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (16 * d + 4 * q < n) in |= static_cast<uint32_t>(row[4 * d + q]) << (8 * q);
        const uint32_t hi = (in >> 1) & 0x55555555u, lo = in & 0x55555555u;
        w = flip ? (((hi & lo) << 1) | (hi & ~lo)) : (((~hi & ~lo & 0x55555555u) << 1) | (hi & ~lo));
        const int valid = n - 16 * d;
        if (valid < 16) w = valid > 0 ? (w & ((1u << (2 * valid)) - 1u)) : 0u;  // padding samples: 0
      } else {
        for (int c = 0; c < 16; ++c) {
          const int i = 16 * d + c;
          if (i < n) {
            bool miss;
            const uint32_t g = static_cast<uint32_t>(geno_at<FMT>(row, sample_col[i], miss));
            w |= (flip ? 2u - g : g) << (2 * c);
          }
        }
      }
      ex_rows[static_cast<size_t>(r) * rowq + d] = w;
    }
    if (lane == 0) ex_key[r] = flip ? 2 * n - a : a;
    return;
  }
  const double mean = mean_in[j], sd = sd_in[j];
  const int64_t col = pos[j];
  for (int i = lane; i < n_pad; i += 64) {
    double z = 0.0;
    float mi = 0.f;
    if (i < n) {
      bool miss;
      const double g = geno_at<FMT>(row, sample_col[i], miss);
      if (miss) {
        mi = 1.f;
        atomicAdd(&miss_cnt[i], 1);
      } else {
        z = (g - mean) / sd;
      }
    }
    const size_t o = static_cast<size_t>(i) * zstride + col;
    z64[o] = z;
    mis[o] = static_cast<__bf16>(mi);
  }
}

// ---- grouped int8 Gram of the exact SNPs --------------------------------------
// For an exact SNP (every call present, minor allele count a, n samples;
// minor-allele dosages) the reference's mean is a/n and its weight
// 1/sd^2 = 1/(2 pf (1 - pf)), pf = a/(2n), depend on a alone. With K_ij N_ij = sum_s w_s (g_si - mu_s)
// (g_sj - mu_s) = A_ij - u_i - u_j + C, A = sum_s w_s g_s g_s^T,
// u_i = sum_s w_s mu_s g_si and C = sum_s w_s mu_s^2, SNPs of equal allele
// count form a group whose Gram sum_{s in group} g_s g_s^T is exact in int32
// on the int8 matrix cores; A adds w_group times each group's Gram in FP64,
// u its per-sample dosage sums times w mu. Groups are padded to 64-SNP
// steps (zero columns), so a step never mixes weights.

// grid-stride fill
__global__ void k_fill_i32_gs(int32_t* p, int64_t count, int32_t v) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void k_grm_hist(const int32_t* key, int64_t cnt, int32_t* hist) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r < cnt) atomicAdd(&hist[key[r]], 1);
}

// One workgroup: group starts (in SNP slots, groups in allele-count order,
// each padded to whole 64-slot steps), per-step weights (wstep: the group's w,
// negated except on the group's last step) and w mu, and C += sum_a
// count_a w_a mu_a^2 (sequential over a: a fixed order).
__global__ __launch_bounds__(1024) void k_grm_groups(const int32_t* hist, int nkeys, int n, int64_t* gstart,
                                                     int32_t* fill, double* wstep, double* wmu, int64_t* nsteps,
                                                     double* C_acc) {
  __shared__ int64_t sc[1024];
  const int t = threadIdx.x;
  const int per = (nkeys + 1023) / 1024;
  const int a = min(nkeys, t * per), b = min(nkeys, a + per);
  int64_t c = 0;
  for (int k = a; k < b; ++k) c += (hist[k] + 63) / 64 * 64;
  sc[t] = c;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int64_t v = t >= off ? sc[t - off] : 0;
    __syncthreads();
    sc[t] += v;
    __syncthreads();
  }
  int64_t p = sc[t] - c;
  for (int k = a; k < b; ++k) {
    gstart[k] = p;
    fill[k] = 0;
    const int64_t padded = (hist[k] + 63) / 64 * 64;
    if (padded) {
      const double mean = static_cast<double>(k) / n, pf = mean / 2.0;
      const double w = 1.0 / (2.0 * pf * (1.0 - pf));
      const int64_t s0 = p / 64, s1 = (p + padded) / 64;
      for (int64_t s = s0; s < s1; ++s) {
        wstep[s] = s == s1 - 1 ? w : -w;
        wmu[s] = w * mean;
      }
    }
    p += padded;
  }
  if (t == 1023) *nsteps = sc[t] / 64;
  if (t == 0) {
    double C = 0.0;
    for (int k = 0; k < nkeys; ++k)
      if (hist[k]) {
        const double mean = static_cast<double>(k) / n, pf = mean / 2.0;
        const double w = 1.0 / (2.0 * pf * (1.0 - pf));
        C += static_cast<double>(hist[k]) * (w * mean * mean);
      }
    *C_acc += C;
  }
}

__global__ void k_grm_scatter(const int32_t* key, int64_t cnt, const int64_t* gstart, int32_t* fill,
                              int32_t* order) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= cnt) return;
  const int k = key[r];
  order[gstart[k] + atomicAdd(&fill[k], 1)] = static_cast<int32_t>(r);
}

// Expands steps [s0, s1) of the grouped order into the int8 operand X, and
// each sample's sum over the workgroup's steps of w mu times its dosage sum in
// the step (one group per step) into upart[blockIdx.x * n_pad + i]. X is
// blocked for the Gram's copies: per block of 16 samples and step one
// contiguous KiB, chunk c (slots 16c .. 16c + 15) of sample 16 b + r at byte
// c * 256 + r * 16, blocks nst = s1 - s0 steps apart. Workgroup = 16 steps x
// 256 samples: thread (sg, rq) reads one dword (16 samples) of four order
// slots' rows and writes them transposed into LDS; thread t then writes
// sample i0 + t's 64 bytes.
__global__ __launch_bounds__(256) void k_grm_expand(const uint32_t* ex_rows, int rowq, const int32_t* order,
                                                    int64_t s0, int64_t s1, const double* wmu, int n_pad, int8_t* X,
                                                    double* upart) {
  __shared__ uint32_t T[256 * 17];
  const int t = threadIdx.x, sg = t & 15, rq = t >> 4;
  const int i0 = blockIdx.y * 256;
  const int d = (i0 >> 4) + sg;
  double up = 0.0;
  for (int ss = 0; ss < 16; ++ss) {
    const int64_t s = s0 + static_cast<int64_t>(blockIdx.x) * 16 + ss;
    if (s >= s1) break;
    uint32_t w[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int o = order[s * 64 + rq * 4 + r];
      w[r] = (o >= 0 && d < rowq) ? ex_rows[static_cast<size_t>(o) * rowq + d] : 0u;
    }
#pragma unroll
    for (int c = 0; c < 16; ++c)
      T[(sg * 16 + c) * 17 + rq] = ((w[0] >> (2 * c)) & 3u) | (((w[1] >> (2 * c)) & 3u) << 8) |
                                   (((w[2] >> (2 * c)) & 3u) << 16) | (((w[3] >> (2 * c)) & 3u) << 24);
    __syncthreads();
    if (i0 + t < n_pad) {
      uint32_t v[16];
      uint32_t bsum = 0;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        v[q] = T[t * 17 + q];
        bsum += (v[q] * 0x01010101u) >> 24;  // byte sum (each byte <= 2)
      }
      const int i = i0 + t;
      unsigned char* dst = reinterpret_cast<unsigned char*>(X) +
                           (static_cast<size_t>(i >> 4) * (s1 - s0) + (s - s0)) * 1024 + (i & 15) * 16;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        *reinterpret_cast<uint4*>(dst + q * 256) = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
      up += wmu[s] * static_cast<double>(bsum);
    }
    __syncthreads();
  }
  if (i0 + t < n_pad) upart[static_cast<size_t>(blockIdx.x) * n_pad + i0 + t] = up;
}

// u_i += sum_b upart[b][i] in b order
__global__ void k_grm_reduce_u(const double* upart, int nb, int n_pad, double* u) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pad) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += upart[static_cast<size_t>(b) * n_pad + i];
  u[i] += s;
}

// A = X^T diag(w) X over steps with per-group weights, lower block triangle in
// 128 x 128 tiles, one workgroup per tile (1-D grid over a tile list whose
// chunks of 32 tiles are blocks of 4 tile rows x 8 tile columns, so the 32
// workgroups an XCD runs at once read 12 row blocks of X, served from its
// L2, instead of 64). Per 64-SNP step the tile's 128 A rows and 128 B rows of
// X — 8 + 8 blocks of 16 samples, one contiguous KiB each in X's blocked
// layout — are copied global -> LDS by direct 16-byte loads (one instruction
// per KiB, block b of a matrix at LDS offset b KiB, so lane l of an MFMA
// fragment reads the 16 bytes at l * 16 of its block) into a ring of
// kI8Stages stages, kI8Stages - 1 steps ahead (a step's MFMA work is short
// against the copy latency). Wave (wr, wc) owns a 64 x 64 quadrant as
// 4 x 4 v_mfma_i32_16x16x64_i8 blocks accumulating in int32 (exact: |sum| <= 4
// x group size); at a group's last step (wstep > 0) and at the launch's last
// step each int32 block is added to the FP64 quadrant times the group's weight
// and cleared. Result element r of lane l is row 4 (l >> 4) + r, column
// l & 15 of its block.
constexpr int kI8T = 128;
constexpr size_t kI8MatBytes = static_cast<size_t>(kI8T) * 64;  // 8 KiB per operand per step
constexpr int kI8Stages = 8;
constexpr size_t kI8StageBytes = 2 * kI8MatBytes;
constexpr size_t kI8LdsBytes = kI8Stages * kI8StageBytes;  // 128 KiB
// waves per workgroup: wave (wr, wc) owns rows wr * kI8RW .. + kI8RW - 1 and
// columns wc * 64 .. + 63 of the tile
constexpr int kI8Waves = 8;  // two per SIMD (measured 16% faster than four waves of 64 x 64)
constexpr int kI8RW = 2 * kI8T / kI8Waves;  // 64 (4 waves) or 32 (8 waves)
constexpr int kI8RB = kI8RW / 16;
constexpr int kI8Copies = 16 / kI8Waves;  // 1 KiB copies per wave per step
// s_waitcnt immediates (gfx9 encoding): vmcnt <= kI8Copies (kI8Stages - 1)
// with expcnt / lgkmcnt at their maximum (no wait)
constexpr int kI8VmAhead = kI8Copies * (kI8Stages - 1);
constexpr int kI8WaitAhead = (kI8VmAhead & 0xF) | ((kI8VmAhead >> 4) << 14) | (0x7 << 4) | (0xF << 8);
static_assert(kI8VmAhead < 64, "vmcnt field");
typedef int i32x4 __attribute__((ext_vector_type(4)));

__device__ inline void i8_step(i32x4 (&a8)[kI8RB][4], const unsigned char* __restrict__ A,
                               const unsigned char* __restrict__ Bm, int wr, int wc, int lane) {
  i32x4 bf[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) bf[cb] = *reinterpret_cast<const i32x4*>(Bm + (wc * 4 + cb) * 1024 + lane * 16);
#pragma unroll
  for (int rb = 0; rb < kI8RB; ++rb) {
    const i32x4 af = *reinterpret_cast<const i32x4*>(A + (wr * kI8RB + rb) * 1024 + lane * 16);
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
      a8[rb][cb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, bf[cb], a8[rb][cb], 0, 0, 0);
  }
}

__global__ __launch_bounds__(64 * kI8Waves) void k_grm_i8(const int8_t* X, int nsteps, int n_pad,
                                                         const int32_t* __restrict__ tiles, int ntiles,
                                                         const double* __restrict__ wstep, double* acc) {
  extern __shared__ __attribute__((aligned(16))) unsigned char sm[];
  // tile of this workgroup: workgroups are dispatched round-robin over the
  // 8 XCDs; XCD x = id % 8 takes chunks x, x + 8, ... of 32 consecutive
  // entries of the tile list (ordered so a chunk covers few row blocks)
  const int x = blockIdx.x & 7, l = blockIdx.x >> 3;
  const int t = (8 * (l >> 5) + x) * 32 + (l & 31);
  if (t >= ntiles) return;
  const int by = tiles[t] >> 16, bx = tiles[t] & 0xFFFF;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  // a wave whose block lies wholly above a diagonal tile's diagonal idles
  const bool active = !(bx == by && wr * kI8RW + kI8RW - 1 < wc * 64);
  const int R0 = by * kI8T, C0 = bx * kI8T;
  // this wave's copies per step: blocks q = wave + kI8Waves i (matrix q >> 3,
  // block q & 7 of its 8); blocks past n_pad read the last block (never stored)
  const int8_t* src[kI8Copies];
  int dst_off[kI8Copies];
#pragma unroll
  for (int i = 0; i < kI8Copies; ++i) {
    const int q = wave + kI8Waves * i;
    const int mat = q >> 3, b = q & 7;
    const int blk = min(((mat ? C0 : R0) >> 4) + b, (n_pad >> 4) - 1);
    src[i] = X + static_cast<size_t>(blk) * nsteps * 1024 + lane * 16;
    dst_off[i] = static_cast<int>(mat * kI8MatBytes) + b * 1024;
  }
  auto issue = [&](int s) {
    unsigned char* buf = sm + (s % kI8Stages) * kI8StageBytes;
#pragma unroll
    for (int i = 0; i < kI8Copies; ++i)
      __builtin_amdgcn_global_load_lds(src[i] + static_cast<size_t>(s) * 1024,
                                       (__attribute__((address_space(3))) void*)(buf + dst_off[i]), 16, 0, 0);
  };
  double f[kI8RB][4][4];
#pragma unroll
  for (int rb = 0; rb < kI8RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) f[rb][cb][r] = 0.0;
  for (int p = 0; p < kI8Stages - 1 && p < nsteps; ++p) issue(p);
  // one pass of the outer loop per group segment (the group's steps, or the
  // part of them in this launch): int32 blocks start at zero, run to the
  // segment's last step, then go into the FP64 block
  int s = 0;
  while (s < nsteps) {
    i32x4 a8[kI8RB][4];
#pragma unroll
    for (int rb = 0; rb < kI8RB; ++rb)
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) a8[rb][cb] = i32x4{0, 0, 0, 0};
    double w;
    while (true) {
      // the stage written now was read in step s - 1, which every wave has
      // finished (the barrier at the end of that step)
      if (s + kI8Stages - 1 < nsteps) {
        issue(s + kI8Stages - 1);
        __builtin_amdgcn_s_waitcnt(kI8WaitAhead);  // own copies of step s landed
      } else {
        __builtin_amdgcn_s_waitcnt(0);
      }
      lds_barrier();  // everyone's copies of step s landed
      if (active) {
        const unsigned char* st = sm + (s % kI8Stages) * kI8StageBytes;
        i8_step(a8, st, st + kI8MatBytes, wr, wc, lane);
      }
      const double ws = wstep[__builtin_amdgcn_readfirstlane(s)];
      // this step's LDS reads are complete (the MFMAs consumed them) before
      // any wave passes, so the next step may overwrite the stage
      __builtin_amdgcn_s_waitcnt(kI8WaitAhead & ~(0xF << 8));  // lgkmcnt 0, vmcnt left alone
      lds_barrier();
      ++s;
      if (ws > 0.0 || s == nsteps) {
        w = fabs(ws);
        break;
      }
    }
    if (active) {
#pragma unroll
      for (int rb = 0; rb < kI8RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
          for (int r = 0; r < 4; ++r) f[rb][cb][r] = fma(w, static_cast<double>(a8[rb][cb][r]), f[rb][cb][r]);
    }
  }
  // n_pad is a multiple of 64: a wave's 32- or 64-row block is wholly inside
  // or outside in rows; columns come in 64-blocks, n_pad a multiple of 64
  if (!active || R0 + wr * kI8RW >= n_pad || C0 + wc * 64 >= n_pad) return;
  double* base = acc + static_cast<size_t>(R0 + wr * kI8RW + 4 * (lane >> 4)) * n_pad + C0 + wc * 64 + (lane & 15);
  double cur[kI8RB][4][4];
#pragma unroll
  for (int rb = 0; rb < kI8RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) cur[rb][cb][r] = base[static_cast<size_t>(rb * 16 + r) * n_pad + cb * 16];
#pragma unroll
  for (int rb = 0; rb < kI8RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) base[static_cast<size_t>(rb * 16 + r) * n_pad + cb * 16] = cur[rb][cb][r] + f[rb][cb][r];
}

// Small n (few 128-wide tiles): one workgroup per 64x64 tile (by >= bx, lower
// block triangle), operands straight from L2. Z^T Z on the
// fp64 matrix cores (v_mfma_f64_16x16x4f64: operands as in the fp32 16x16x4
// form — per instruction t, lane group l>>4 supplies k = 8*(l>>4)+t of the
// current 32-SNP step — while result element r of lane l is row (l>>4) + 4r),
// so K carries no reduced-precision rounding.
__global__ __launch_bounds__(256) void k_grm_gram(const double* z64, const __bf16* mis, int m_pad, int n_pad,
                                                  double* acc, int32_t* joint, int any_missing) {
  const int bx = blockIdx.x, by = blockIdx.y;
  if (bx > by) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = by * 64 + wave * 16;  // sample rows i
  const int c0 = bx * 64;              // sample cols k
  const int kk = 8 * (lane >> 4);
  f64x4 a64[4];
  f32x4 accm[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) { a64[cb] = f64x4{0, 0, 0, 0}; accm[cb] = f32x4{0, 0, 0, 0}; }
  const size_t arow = static_cast<size_t>(r0 + (lane & 15)) * m_pad + kk;
  for (int k0 = 0; k0 < m_pad; k0 += 32) {
    const f64x8 az = load8_f64(z64 + arow + k0);
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const size_t bo = static_cast<size_t>(c0 + cb * 16 + (lane & 15)) * m_pad + kk + k0;
      a64[cb] = mfma_f64x8(az, load8_f64(z64 + bo), a64[cb]);
    }
    if (any_missing) {
      const bf16x8 am = load8_bf16(mis + arow + k0);
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const size_t bo = static_cast<size_t>(c0 + cb * 16 + (lane & 15)) * m_pad + kk + k0;
        accm[cb] = mfma_bf16(am, load8_bf16(mis + bo), accm[cb]);  // 0/1 products: exact below 2^24 SNPs
      }
    }
  }
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // fp64 result rows are interleaved ((l >> 4) + 4 r); the bf16 count tile uses 4 (l >> 4) + r
      const size_t o = static_cast<size_t>(r0 + (lane >> 4) + 4 * r) * n_pad + c0 + cb * 16 + (lane & 15);
      acc[o] += a64[cb][r];
      const size_t om = static_cast<size_t>(r0 + 4 * (lane >> 4) + r) * n_pad + c0 + cb * 16 + (lane & 15);
      joint[om] += static_cast<int32_t>(accm[cb][r]);
    }
}

// Joint missing counts M^T M for the GRM divisor, one workgroup per 64x64 tile
// (by >= bx, lower block triangle): 0/1 products on the bf16 matrix cores
// (exact below 2^24 SNPs), result element r of lane l is row 4 (l>>4) + r.
__global__ __launch_bounds__(256) void k_grm_joint(const __bf16* mis, int m_pad, int n_pad, int32_t* joint) {
  const int bx = blockIdx.x, by = blockIdx.y;
  if (bx > by) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = by * 64 + wave * 16;  // sample rows i
  const int c0 = bx * 64;              // sample cols k
  const int kk = 8 * (lane >> 4);
  f32x4 accm[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) accm[cb] = f32x4{0, 0, 0, 0};
  const size_t arow = static_cast<size_t>(r0 + (lane & 15)) * m_pad + kk;
  for (int k0 = 0; k0 < m_pad; k0 += 32) {
    const bf16x8 am = load8_bf16(mis + arow + k0);
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const size_t bo = static_cast<size_t>(c0 + cb * 16 + (lane & 15)) * m_pad + kk + k0;
      accm[cb] = mfma_bf16(am, load8_bf16(mis + bo), accm[cb]);
    }
  }
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const size_t om = static_cast<size_t>(r0 + 4 * (lane >> 4) + r) * n_pad + c0 + cb * 16 + (lane & 15);
      joint[om] += static_cast<int32_t>(accm[cb][r]);
    }
}

// Z^T Z for the GRM (unique lower block triangle), FP64 matrix cores, 128 x 128
// tiles of K per workgroup (grid ceil(n_pad / 128)^2, by >= bx). Per 32-SNP
// step the tile's 128 A rows (samples R0..) and 128 B rows (samples C0..) of Z
// are copied global -> LDS (direct 16-byte loads, double-buffered), each row's
// 256 bytes as 16 chunks of 16 B stored chunk-major (slot (c, r) at
// (c * 128 + r) * 16), so the 16 lanes reading one MFMA fragment hit 16
// consecutive slots. Wave (wr, wc) owns the 64 x 64 quadrant (rows wr*64..,
// cols wc*64..) as 4 x 4 blocks of v_mfma_f64_16x16x4f64 (the fragment
// grouping of mfma_f64x8: lane l supplies k = 8*(l>>4) + t, result element r is
// row (l>>4) + 4r). The quadrant above the diagonal of a diagonal tile is
// skipped. Rows past n_pad are read clamped and never stored.
constexpr int kGramT = 128;
constexpr size_t kGramMatBytes = static_cast<size_t>(kGramT) * 32 * sizeof(double);  // 32 KiB per matrix per step
constexpr size_t kGramLdsBytes = 4 * kGramMatBytes;                                   // 2 buffers x (A, B)
// sample count (padded) from which the 128-wide tiles are used (measured: 64x64
// L2-fed tiles are faster at n = 1000, 128x128 LDS tiles from n = 2000 on)
constexpr int kGramLargeN = 1536;

__global__ __launch_bounds__(256) void k_grm_syrk(const double* z64, int m_pad, int n_pad, double* acc) {
  extern __shared__ __attribute__((aligned(16))) unsigned char gsm[];
  const int bx = blockIdx.x, by = blockIdx.y;
  if (bx > by) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const bool active = !(bx == by && wr < wc);
  const int R0 = by * kGramT, C0 = bx * kGramT;
  // this wave's copies: 64 instructions of 1 KiB per step (matrix, chunk c, row half h), q = wave + 4 i
  const double* src[16];
  int dst_off[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int q = wave + 4 * i;
    const int mat = q >> 5, c = (q & 31) >> 1, h = q & 1;
    const int row = min((mat ? C0 : R0) + h * 64 + lane, n_pad - 1);
    src[i] = z64 + static_cast<size_t>(row) * m_pad + 2 * c;
    dst_off[i] = static_cast<int>(mat * kGramMatBytes) + (c * kGramT + h * 64) * 16;
  }
  auto issue = [&](int k0, unsigned char* buf) {
#pragma unroll
    for (int i = 0; i < 16; ++i)
      __builtin_amdgcn_global_load_lds(src[i] + k0, (__attribute__((address_space(3))) void*)(buf + dst_off[i]), 16,
                                       0, 0);
  };
  f64x4 a64[4][4];
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) a64[rb][cb] = f64x4{0, 0, 0, 0};
  const int nks = m_pad / 32;
  issue(0, gsm);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int s = 0; s < nks; ++s) {
    unsigned char* cur = gsm + (s & 1) * 2 * kGramMatBytes;
    if (s + 1 < nks) issue((s + 1) * 32, gsm + ((s + 1) & 1) * 2 * kGramMatBytes);
    if (active) {
      const unsigned char* A = cur;
      const unsigned char* Bm = cur + kGramMatBytes;
      const int cg = 4 * (lane >> 4);
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        const int ra = wr * 64 + rb * 16 + (lane & 15);
        f64x8 az;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const double2 v = *reinterpret_cast<const double2*>(A + ((cg + t) * kGramT + ra) * 16);
          az.v[2 * t] = v.x;
          az.v[2 * t + 1] = v.y;
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
          const int rbm = wc * 64 + cb * 16 + (lane & 15);
          f64x8 bz;
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const double2 v = *reinterpret_cast<const double2*>(Bm + ((cg + t) * kGramT + rbm) * 16);
            bz.v[2 * t] = v.x;
            bz.v[2 * t + 1] = v.y;
          }
          a64[rb][cb] = mfma_f64x8(az, bz, a64[rb][cb]);
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  if (!active) return;
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = R0 + wr * 64 + rb * 16 + (lane >> 4) + 4 * r;
        const int k = C0 + wc * 64 + cb * 16 + (lane & 15);
        if (i < n_pad && k < n_pad) acc[static_cast<size_t>(i) * n_pad + k] += a64[rb][cb][r];
      }
}

// u (per sample) and C_acc (one value) hold the exact path's rank-1 centring
// terms; both are zero when no exact SNP was accumulated.
__global__ void k_grm_finish(const double* acc, const int32_t* joint, const int32_t* miss_cnt, const double* u,
                             const double* C_acc, int n, int n_pad, int m, double* K, int32_t* Npk) {
  const int i = blockIdx.y * 16 + threadIdx.y;
  const int j = blockIdx.x * 16 + threadIdx.x;
  if (i >= n || j >= n || j > i) return;
  const size_t o = static_cast<size_t>(i) * n_pad + j;  // i >= j: stored in the lower block triangle
  int subN;
  if (i == j) subN = m - miss_cnt[i];
  else subN = m - miss_cnt[i] - miss_cnt[j] + joint[o];
  const double a = acc[o] - u[i] - u[j] + *C_acc;
  const double v = subN > 0 ? a / static_cast<double>(subN) : 0.0;
  K[static_cast<size_t>(j) * n + i] = v;
  K[static_cast<size_t>(i) * n + j] = v;
  Npk[static_cast<size_t>(i) * (i + 1) / 2 + j] = subN;
}

}  // namespace eqtl_amd
