// Permutation index streams of the reference's cis/trans permutations
// (scan_cis.hpp:106-111, scan_perm.cpp:84-88):
//   std::mt19937 rng(seed' + 9973*b + fnv1a(gene));   (32-bit wrap)
//   std::iota(idx, 0); std::shuffle(idx.begin(), idx.end(), rng);
// reproduced on the device so the streams are bit-identical to libstdc++
// (GCC 11) without shipping the index arrays from the host. The steps below
// follow the published MT19937 (Matsumoto & Nishimura 1998, standard seeding
// with multiplier 1812433253) and libstdc++'s shuffle: one swap up front when
// n is even, then pairs of swaps from one draw on [0, r(r+1)) split as
// (x / (r+1), x % (r+1)); draws use Lemire's multiply-and-reject reduction of a
// 32-bit output (uniform_int_distribution with a 32-bit generator range).
// Valid for n <= 65535 (the two-at-a-time branch, n*n <= 2^32 - 1).
#pragma once
#include <stdint.h>
#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define EQ_PHD __host__ __device__ inline
#else
#define EQ_PHD inline
#endif

namespace eqtl_amd {

constexpr int kMtN = 624;

// MT state lives in caller-provided storage (global scratch on the device).
struct Mt19937 {
  uint32_t* x;
  int idx;

  EQ_PHD void seed(uint32_t s) {
    x[0] = s;
    for (int i = 1; i < kMtN; ++i) x[i] = 1812433253u * (x[i - 1] ^ (x[i - 1] >> 30)) + static_cast<uint32_t>(i);
    idx = kMtN;
  }
  EQ_PHD void twist() {
    for (int k = 0; k < kMtN; ++k) {
      const uint32_t y = (x[k] & 0x80000000u) | (x[k + 1 < kMtN ? k + 1 : 0] & 0x7fffffffu);
      const int k2 = k + 397 < kMtN ? k + 397 : k + 397 - kMtN;
      x[k] = x[k2] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    idx = 0;
  }
  EQ_PHD uint32_t next() {
    if (idx >= kMtN) twist();
    uint32_t y = x[idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  // uniform integer on [0, range) for range < 2^32 (Lemire, rejection on the low word)
  EQ_PHD uint32_t below(uint32_t range) {
    uint64_t prod = static_cast<uint64_t>(next()) * range;
    uint32_t low = static_cast<uint32_t>(prod);
    if (low < range) {
      const uint32_t thr = static_cast<uint32_t>(0u - range) % range;
      while (low < thr) {
        prod = static_cast<uint64_t>(next()) * range;
        low = static_cast<uint32_t>(prod);
      }
    }
    return static_cast<uint32_t>(prod >> 32);
  }
};

EQ_PHD uint32_t perm_seed(int32_t seed, int32_t b, uint32_t gene_hash) {
  return static_cast<uint32_t>(seed >= 0 ? seed : 1) + static_cast<uint32_t>(b) * 9973u + gene_hash;
}

// idx[0..n) := shuffle(iota(n)) with the generator already seeded.
EQ_PHD void shuffle_iota(Mt19937& g, int32_t* idx, int n) {
  for (int i = 0; i < n; ++i) idx[i] = i;
  if (n <= 1) return;
  int i = 1;
  if ((n % 2) == 0) {
    const int j = static_cast<int>(g.below(2));
    const int t = idx[i]; idx[i] = idx[j]; idx[j] = t;
    ++i;
  }
  while (i < n) {
    const uint32_t r = static_cast<uint32_t>(i) + 1;  // swap range for position i
    const uint32_t x = g.below(r * (r + 1));
    const int p1 = static_cast<int>(x / (r + 1)), p2 = static_cast<int>(x % (r + 1));
    int t = idx[i]; idx[i] = idx[p1]; idx[p1] = t;
    ++i;
    t = idx[i]; idx[i] = idx[p2]; idx[p2] = t;
    ++i;
  }
}

EQ_PHD uint32_t fnv1a32(const char* s) {  // util.hpp:51-58 (low 32 bits of 64-bit FNV-1a)
  unsigned long long h = 14695981039346656037ULL;
  for (; *s; ++s) {
    h ^= static_cast<unsigned char>(*s);
    h *= 1099511628211ULL;
  }
  return static_cast<uint32_t>(h & 0xFFFFFFFFu);
}

}  // namespace eqtl_amd

#if defined(__HIPCC__) || defined(__HIP__)
namespace eqtl_amd {

// Permuted phenotypes of one batch of (gene, b) pairs, one thread per pair
// e = local_gene * B + b (scan_cis.hpp:104-148): the index stream of
// (seed, b, gene) permutes the gene's base vector.
//   LM : y_e[i]  = y_s[idx[i]]                       (sample space)
//   LMM: y~_e[i] = y~[i] + (w[idx[i]] - w[i]) / sqrt(max(d_i, 1e-12))
//        with w = r~ o sqrt(d) the whitened null residual, i.e. the rotated
//        X~ b0 + r~_perm of the reference's y = Q (X~ b0 + r~_perm).
// mt: E x kMtN words of generator state, idx: E x n scratch.
__global__ void k_perm_pheno(int E, int B, int n, int32_t seed, const int32_t* gene_of, const uint32_t* hash_of,
                             uint32_t* mt, int32_t* idx, int lmm, const double* base, const double* yt,
                             const double* delta, const double* lam, double* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int lk = e / B, b = e % B;
  const int k = gene_of[lk];
  Mt19937 g{mt + static_cast<size_t>(e) * kMtN, kMtN};
  g.seed(perm_seed(seed, b, hash_of[lk]));
  int32_t* ix = idx + static_cast<size_t>(e) * n;
  shuffle_iota(g, ix, n);
  const double* bk = base + static_cast<size_t>(k) * n;
  double* o = out + static_cast<size_t>(e) * n;
  if (!lmm) {
    for (int i = 0; i < n; ++i) o[i] = bk[ix[i]];
  } else {
    const double* yk = yt + static_cast<size_t>(k) * n;
    const double dk = delta[k];
    for (int i = 0; i < n; ++i) {
      double v = dk * lam[i] + 1.0;
      if (v < 1e-12) v = 1e-12;
      const double d = 1.0 / v;
      o[i] = yk[i] + (bk[ix[i]] - bk[i]) / sqrt(fmax(d, 1e-12));
    }
  }
}

// w = U / sqrt(d): the whitened null residual from the kernel's s o (s o r~)
// column (k_lmm_basis stores U = s o r with r = s o r~).
__global__ void k_whitened_residual(const double* U, const double* delta, const double* lam, int n, int G,
                                    double* w) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(n) * G) return;
  const int k = static_cast<int>(i / n), r = static_cast<int>(i % n);
  double v = delta[k] * lam[r] + 1.0;
  if (v < 1e-12) v = 1e-12;
  w[i] = U[i] * sqrt(v);
}

}  // namespace eqtl_amd
#endif
