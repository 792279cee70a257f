// K2: GRM (compute_grm_from, grm.cpp:112-224).
//   k_grm_standardize — decode each genotype row, apply the GRM's SNP rules
//       (the source's filters, then p = mean/2 in [1e-8, 1-1e-8], sd >= 1e-12),
//       write z = (g - mean)/sd (missing → 0) in fp64 and the 0/1 missing
//       indicator in bf16, sample-major (SNP-contiguous) for the MFMA
//       operands; count per-sample missing calls of used SNPs.
//   k_grm_syrk — lower block-triangle of Z^T Z (fp64 MFMA, LDS-staged
//       128 x 128 tiles), accumulated into an FP64 tile array;
//   k_grm_joint — the same triangle of M^T M (joint missing counts, bf16
//       MFMA on 0/1, exact), when the block has a missing call;
//   k_grm_gram — both on 64 x 64 tiles for small n, where 128-wide tiles
//       would leave most CUs idle.
//   k_grm_finish — per-pair GCTA divisor m - miss_i - miss_j + joint_ij
//       (diagonal m - miss_i), symmetric K and packed N.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "geno_rows.hpp"
#include "mfma.hpp"

namespace eqtl_amd {

template <int FMT>
__global__ __launch_bounds__(256) void k_grm_standardize(const uint8_t* rows, int64_t m, int64_t m_pad, int bps,
                                                         int n, int n_pad, int n_file, const int32_t* sample_col,
                                                         double* z64, __bf16* mis, int32_t* miss_cnt,
                                                         unsigned long long* used, int impute, double max_miss,
                                                         double maf_min) {
  const int lane = threadIdx.x & 63;
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (j >= m_pad) return;
  bool use = j < m;
  int n_miss = 0;
  double sum = 0.0;  // integer-valued for BED rows (exact)
  const uint8_t* row = use ? rows + j * bps : nullptr;
  if (use) {
    for (int i = lane; i < n; i += 64) {
      bool miss;
      const double g = geno_at<FMT>(row, sample_col[i], miss);
      if (miss) { ++n_miss; continue; }
      sum += g;
    }
    for (int o = 32; o > 0; o >>= 1) {
      n_miss += __shfl_xor(n_miss, o);
      sum += __shfl_xor(sum, o);
    }
  }
  double mu = 0, mean = 0, sd = 1;
  if (use) {
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
  }
  for (int i = lane; i < n_pad; i += 64) {
    double z = 0.0;
    float mi = 0.f;
    if (use && i < n) {
      bool miss;
      const double g = geno_at<FMT>(row, sample_col[i], miss);
      if (miss) {
        mi = 1.f;
        atomicAdd(&miss_cnt[i], 1);
      } else {
        z = (g - mean) / sd;
      }
    }
    const size_t o = static_cast<size_t>(i) * m_pad + j;
    z64[o] = z;
    mis[o] = static_cast<__bf16>(mi);
  }
  if (use && lane == 0) {
    atomicAdd(used, 1ull);
    if (n_miss > 0) atomicAdd(used + 1, 1ull);
  }
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

__global__ void k_grm_finish(const double* acc, const int32_t* joint, const int32_t* miss_cnt, int n, int n_pad,
                             int m, double* K, int32_t* Npk) {
  const int i = blockIdx.y * 16 + threadIdx.y;
  const int j = blockIdx.x * 16 + threadIdx.x;
  if (i >= n || j >= n || j > i) return;
  const size_t o = static_cast<size_t>(i) * n_pad + j;  // i >= j: stored in the lower block triangle
  int subN;
  if (i == j) subN = m - miss_cnt[i];
  else subN = m - miss_cnt[i] - miss_cnt[j] + joint[o];
  const double v = subN > 0 ? acc[o] / static_cast<double>(subN) : 0.0;
  K[static_cast<size_t>(j) * n + i] = v;
  K[static_cast<size_t>(i) * n + j] = v;
  Npk[static_cast<size_t>(i) * (i + 1) / 2 + j] = subN;
}

}  // namespace eqtl_amd
