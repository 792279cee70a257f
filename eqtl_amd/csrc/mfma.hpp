// CDNA4 (gfx950) matrix-core fragments used by the contraction kernels.
//
// 16x16 output tiles, wave64. Operand/result lane layout (one place only):
//   A (16 rows x K): lane l holds row (l & 15), k-slots 8*(l >> 4) + [0..7]
//   B (K x 16 cols): lane l holds col (l & 15), k-slots 8*(l >> 4) + [0..7]
//   C (16 x 16)    : lane l holds col (l & 15), rows 4*(l >> 4) + [0..3]
//                    (fp64 16x16x4: rows (l >> 4) + 4*[0..3])
// Both operands are K-contiguous in memory, so each lane loads 8 consecutive
// k values (16 B of bf16, 32 B of fp32) for its row/column.
//
// bf16:  v_mfma_f32_16x16x32_bf16 consumes the 8 k-slots in one instruction.
// fp32:  v_mfma_f32_16x16x4_f32 takes one k-slot per lane; MFMA t (t=0..7)
//        uses each lane's element t, so lane groups cover k = 8*(l>>4)+t and
//        eight instructions cover the same 32 k values (the k order inside a
//        sum is irrelevant as long as A and B agree).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eqtl_amd {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

typedef __attribute__((__vector_size__(4 * sizeof(double)))) double f64x4;

struct f32x8 {
  float v[8];
};

struct f64x8 {
  double v[8];
};

__device__ inline f32x4 mfma_bf16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline f32x4 mfma_f32x8(const f32x8& a, const f32x8& b, f32x4 c) {
#pragma unroll
  for (int t = 0; t < 8; ++t) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.v[t], b.v[t], c, 0, 0, 0);
  return c;
}

// fp64 16x16x4: operands laid out as in the fp32 form (8 instructions per 32 k);
// the result differs: lane l holds col (l & 15), rows (l >> 4) + 4 * [0..3].
__device__ inline f64x4 mfma_f64x8(const f64x8& a, const f64x8& b, f64x4 c) {
#pragma unroll
  for (int t = 0; t < 8; ++t) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a.v[t], b.v[t], c, 0, 0, 0);
  return c;
}

// Workgroup barrier without the memory-model fences of __syncthreads(): the
// caller has already waited for the counters it needs. The "memory" clobber
// keeps the compiler from moving memory operations across it.
__device__ inline void lds_barrier() { asm volatile("s_barrier" ::: "memory"); }

// Round-to-nearest-even fp32 -> bf16 and the exact residual split x = hi + lo.
__device__ inline __bf16 to_bf16(float x) { return static_cast<__bf16>(x); }

__device__ inline void split_bf16(const f32x8& x, bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const __bf16 h = to_bf16(x.v[t]);
    hi[t] = h;
    lo[t] = to_bf16(x.v[t] - static_cast<float>(h));
  }
}

__device__ inline f32x8 load8(const float* p) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  const float4 b = *reinterpret_cast<const float4*>(p + 4);
  f32x8 r;
  r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
  r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
  return r;
}

__device__ inline f64x8 load8_f64(const double* p) {
  f64x8 r;
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const double2 v = reinterpret_cast<const double2*>(p)[h];
    r.v[2 * h] = v.x;
    r.v[2 * h + 1] = v.y;
  }
  return r;
}

__device__ inline bf16x8 load8_bf16(const __bf16* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}

// Gene-side operand, fragment-major: for gene group g (16 genes) and k-step s
// (32 samples) one contiguous tile holds, per quantity block b, the 64 lanes'
// 8-value B fragments (lane = gene_in_group + 16*((k % 32) / 8), value k % 8).
// bf16 tiles hold the hi plane then the lo plane (2 * nq KiB), fp32 tiles
// nq * 2 KiB, so a whole tile is one contiguous copy into LDS.
struct GeneCols {
  float* f32;
  __bf16* bf;
  int nq, nks;
  __host__ __device__ size_t tile_f32(int g, int s) const { return (static_cast<size_t>(g) * nks + s) * nq * 512; }
  __host__ __device__ size_t tile_bf(int g, int s) const { return (static_cast<size_t>(g) * nks + s) * nq * 1024; }
};

// value v of quantity b for gene `gene` at sample i, in all precision copies
__device__ inline void put_val(const GeneCols& gc, int gene, int b, int i, double v) {
  const int g = gene >> 4, col = gene & 15, s = i >> 5, lane = col + 16 * ((i & 31) >> 3), e = i & 7;
  const float f = static_cast<float>(v);
  gc.f32[gc.tile_f32(g, s) + (static_cast<size_t>(b) * 64 + lane) * 8 + e] = f;
  const __bf16 h = static_cast<__bf16>(f);
  const size_t ob = gc.tile_bf(g, s) + (static_cast<size_t>(b) * 64 + lane) * 8 + e;
  gc.bf[ob] = h;
  gc.bf[ob + static_cast<size_t>(gc.nq) * 512] = static_cast<__bf16>(f - static_cast<float>(h));
}

}  // namespace eqtl_amd
