// VCF text -> DOSE32 genotype rows on the device.
//   k_vcf_decode — one wave per record. The wave walks the record's sample
//     columns 256 bytes at a time: lane l holds dword l of the step (coalesced
//     loads, prefetched two steps ahead), counts its tabs, and a wave prefix
//     sum over the lanes gives the sample index of every column that starts
//     in its dword. The lane decodes those columns' GT or DS subfield
//     (VcfSession::parse_record, vcf_session.cpp:160-286) from a 16-byte
//     register window (its dword and the next three lanes' dwords, through
//     shuffles): branch-free for the common GT / GT:DS forms, byte by byte
//     otherwise, and from the line in memory for a subfield that runs past
//     the window. It writes the float ALT dosage (NaN = missing) and
//     accumulates, over the analysis samples, the missing count, the dosage
//     sum and the called allele copies. The record-level rules (n_ok,
//     miss_frac, AF in (1e-12, 1-1e-12), Filter, pass_maf) then decide the
//     row; a rejected row is rewritten as all NaN. No LDS, no barriers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/eqtl_gpu.h"
#include "vcf_parse.hpp"

namespace eqtl_amd {

constexpr int kVcfThreads = 256;
constexpr int kVcfWaves = kVcfThreads / 64;  // records per workgroup

struct VcfArgs {
  const unsigned char* text;
  const eqtl_vcf_record* recs;
  int64_t m;
  int32_t n_file, n;
  const uint8_t* used;  // n_file: number of analysis samples reading the column
  int32_t used_all;     // every column read exactly once (used[] is all ones)
  int32_t impute;
  double max_miss, maf_min;
  float* rows;                 // m x n_file
  unsigned long long* bad;     // lowest malformed record (initialised to m)
};

// The general parse of a column (the uncommon forms): from the register
// window, else from the line in memory. Kept out of line so that the main
// loop's registers stay few.
__device__ __attribute__((noinline)) VcfCall vcf_column_slow(const unsigned char* line, int L, uint64_t lo,
                                                             uint64_t hi, int o, int wbase, int gt_key, int ds_key,
                                                             bool hap) {
  VcfCall c;
  if (parse_sample_window(lo, hi, o, L - wbase, gt_key, ds_key, hap, c)) return c;
  auto at_line = [&](int q) -> unsigned { return line[q]; };
  return parse_sample(at_line, wbase + o, L, gt_key, ds_key, hap);
}

// ANY_DS: some record of the launch reads DS (else the DS code is compiled out).
template <bool ANY_DS>
__global__ __launch_bounds__(kVcfThreads) void k_vcf_decode(VcfArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kVcfWaves + (threadIdx.x >> 6);
  if (r >= a.m) return;  // wave-uniform
  const eqtl_vcf_record rec = a.recs[r];
  const int L = rec.length;
  const unsigned char* line = a.text + rec.offset;
  const int s0 = static_cast<int>(reinterpret_cast<uintptr_t>(line) & 3);
  const unsigned* gw = reinterpret_cast<const unsigned*>(line - s0);
  const int D = (s0 + L + 3) >> 2;  // dwords of the aligned span
  const int n_file = a.n_file;
  float* row = a.rows + r * static_cast<int64_t>(n_file);
  const bool hap = rec.haploid != 0;
  const int gt_key = rec.gt_key, ds_key = ANY_DS ? rec.ds_key : -1;
  int n_miss = 0, copies = 0;
  double sum = 0.0;
  bool over = false;
  // column k starting at byte o (0..4) of the window that begins at line position wbase
  auto take = [&](int k, int o, int wbase, uint64_t lo, uint64_t hi) {
    if (k >= n_file) {
      over = true;
      return;
    }
    VcfCall c;
    const int sh = 8 * o;  // o <= 4: 16 - o >= 12 window bytes from the column start
    const uint64_t x0 = o == 0 ? lo : ((lo >> sh) | (hi << (64 - sh))), x1 = o == 0 ? hi : (hi >> sh);
    if (!parse_sample_fast(x0, x1, 16 - o, L - wbase - o, gt_key, ds_key, hap, c))
      c = vcf_column_slow(line, L, lo, hi, o, wbase, gt_key, ds_key, hap);
    row[k] = c.miss ? __builtin_nanf("") : c.dose;
    const int u = a.used_all ? 1 : a.used[k];
    if (u) {
      if (c.miss) n_miss += u;
      else { sum += u * static_cast<double>(c.dose); copies += u * c.copies; }
    }
  };
  int carry = 0;  // tabs before this step (wave-uniform)
  unsigned cur = lane < D ? gw[lane] : 0u;
  unsigned nxt = lane + 64 < D ? gw[lane + 64] : 0u;
  const int d_end = D > 0 ? D : 1;  // an empty line still has its one (empty) column
  for (int d0 = 0; d0 < d_end; d0 += 64) {
    const int d = d0 + lane;
    const unsigned nx2 = d + 128 < D ? gw[d + 128] : 0u;  // two steps ahead
    unsigned mk = d < D ? (tab_mask(cur) & line_mask(d, s0, L)) : 0u;
    const int c = __popc(mk);
    int incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o);
      if (lane >= o) incl += t;
    }
    // 16-byte window: this lane's dword and the next three (the next step's
    // dwords past lane 63)
    const unsigned c1 = __shfl(cur, (lane + 1) & 63), n1 = __shfl(nxt, (lane + 1) & 63);
    const unsigned c2 = __shfl(cur, (lane + 2) & 63), n2 = __shfl(nxt, (lane + 2) & 63);
    const unsigned c3 = __shfl(cur, (lane + 3) & 63), n3 = __shfl(nxt, (lane + 3) & 63);
    const unsigned w1 = lane + 1 < 64 ? c1 : n1, w2 = lane + 2 < 64 ? c2 : n2, w3 = lane + 3 < 64 ? c3 : n3;
    const uint64_t lo = cur | (static_cast<uint64_t>(w1) << 32), hi = w2 | (static_cast<uint64_t>(w3) << 32);
    const int wbase = 4 * d - s0;
    int k = carry + incl - c;  // tabs before this dword
    if (d == 0) take(0, s0, wbase, lo, hi);  // the first column
    while (mk) {
      const int b = __builtin_ctz(mk);
      mk &= mk - 1;
      ++k;
      take(k, b + 1, wbase, lo, hi);
    }
    carry += __shfl(incl, 63);
    cur = nxt;
    nxt = nx2;
  }
  if (carry != n_file - 1 || __any(over)) {
    if (lane == 0) atomicMin(a.bad, static_cast<unsigned long long>(r));
    return;
  }
  // record-level rules over the analysis samples
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    n_miss += __shfl_xor(n_miss, o);
    copies += __shfl_xor(copies, o);
    sum += __shfl_xor(sum, o);
  }
  const int n_ok = a.n - n_miss;
  bool keep = n_ok > 0;
  if (keep && static_cast<double>(n_miss) / static_cast<double>(a.n) > a.max_miss + 1e-15) keep = false;
  double af = copies > 0 ? sum / static_cast<double>(copies) : 0.0;
  af = fmin(1.0, fmax(0.0, af));
  if (af < 1e-12 || af > 1.0 - 1e-12) keep = false;
  if (!a.impute && n_miss > 0) keep = false;
  if (a.maf_min > 0.0 && (af < a.maf_min || af > 1.0 - a.maf_min)) keep = false;  // pass_maf
  if (!keep) {
    __threadfence_block();  // the column stores above complete before the rewrite
    for (int i = lane; i < n_file; i += 64) row[i] = __builtin_nanf("");
  }
}

}  // namespace eqtl_amd
