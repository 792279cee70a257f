// VCF text -> DOSE32 genotype rows on the device.
//   k_vcf_decode — one workgroup per record. The record's sample columns are
//     staged in LDS with coalesced dword loads when they fit; each thread
//     counts the tabs of a contiguous chunk, a workgroup prefix sum gives every
//     thread the sample index of the columns that start in its chunk, and the
//     thread parses the GT or DS subfield of those columns
//     (VcfSession::parse_record, vcf_session.cpp:160-286): it writes the float
//     ALT dosage (NaN = missing) and accumulates, over the analysis samples,
//     the missing count, the dosage sum and the called allele copies. The
//     record-level rules (n_ok, miss_frac, AF in (1e-12, 1-1e-12), Filter,
//     pass_maf) then decide the row; a rejected row is rewritten as all NaN.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/eqtl_gpu.h"
#include "vcf_parse.hpp"

namespace eqtl_amd {

constexpr int kVcfThreads = 256;
constexpr int kVcfLdsMax = 60 * 1024;  // largest staged record span (bytes), within the default 64 KiB

struct VcfArgs {
  const unsigned char* text;
  const eqtl_vcf_record* recs;
  int64_t m;
  int32_t n_file, n;
  const uint8_t* used;  // n_file: number of analysis samples reading the column
  int32_t impute;
  double max_miss, maf_min;
  float* rows;                 // m x n_file
  unsigned long long* bad;     // lowest malformed record (initialised to m)
};

template <bool STAGED>
__global__ __launch_bounds__(kVcfThreads) void k_vcf_decode(VcfArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char vsm[];
  __shared__ int s_tabs[kVcfThreads / 64];
  __shared__ double s_sum[kVcfThreads / 64];
  __shared__ int s_cnt[2][kVcfThreads / 64];
  __shared__ int s_keep;
  const int64_t r = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const eqtl_vcf_record rec = a.recs[r];
  const int L = rec.length;
  const unsigned char* line = a.text + rec.offset;
  const int s0 = static_cast<int>(reinterpret_cast<uintptr_t>(line) & 3);
  const unsigned* gw = reinterpret_cast<const unsigned*>(line - s0);
  const int D = (s0 + L + 3) >> 2;  // dwords of the aligned span
  const unsigned* lw = reinterpret_cast<const unsigned*>(vsm);
  if constexpr (STAGED) {
    unsigned* dst = reinterpret_cast<unsigned*>(vsm);
    for (int d = tid; d < D; d += kVcfThreads) dst[d] = gw[d];
    __syncthreads();
  }
  auto word = [&](int d) -> unsigned { return STAGED ? lw[d] : gw[d]; };
  auto at = [&](int q) -> unsigned { return STAGED ? vsm[s0 + q] : line[q]; };

  // pass 1: tabs per chunk, workgroup exclusive prefix
  const int per = (D + kVcfThreads - 1) / kVcfThreads;
  const int d0 = min(D, tid * per), d1 = min(D, d0 + per);
  int cnt = 0;
  for (int d = d0; d < d1; ++d) cnt += __popc(tab_mask(word(d)) & line_mask(d, s0, L));
  int incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) s_tabs[wave] = incl;
  __syncthreads();
  int before = incl - cnt, total = 0;
#pragma unroll
  for (int w = 0; w < kVcfThreads / 64; ++w) {
    if (w < wave) before += s_tabs[w];
    total += s_tabs[w];
  }
  if (total != a.n_file - 1) {
    if (tid == 0) atomicMin(a.bad, static_cast<unsigned long long>(r));
    return;  // uniform over the workgroup
  }

  // pass 2: parse the columns that start in this chunk
  float* row = a.rows + r * static_cast<int64_t>(a.n_file);
  const bool hap = rec.haploid != 0;
  int n_miss = 0, copies = 0;
  double sum = 0.0;
  auto take = [&](int k, int q) {
    const VcfCall c = parse_sample(at, q, L, rec.gt_key, rec.ds_key, hap);
    row[k] = c.miss ? __builtin_nanf("") : c.dose;
    const int u = a.used[k];
    if (u) {
      if (c.miss) n_miss += u;
      else { sum += u * static_cast<double>(c.dose); copies += u * c.copies; }
    }
  };
  if (tid == 0) take(0, 0);
  int k = before;
  for (int d = d0; d < d1; ++d) {
    unsigned mk = tab_mask(word(d)) & line_mask(d, s0, L);
    while (mk) {
      const int b = __builtin_ctz(mk);
      mk &= mk - 1;
      ++k;
      take(k, 4 * d - s0 + b + 1);
    }
  }
  // record-level rules over the analysis samples
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    n_miss += __shfl_xor(n_miss, o);
    copies += __shfl_xor(copies, o);
    sum += __shfl_xor(sum, o);
  }
  if (lane == 0) {
    s_cnt[0][wave] = n_miss;
    s_cnt[1][wave] = copies;
    s_sum[wave] = sum;
  }
  __syncthreads();
  if (tid == 0) {
    int nm = 0, cp = 0;
    double sm = 0.0;
    for (int w = 0; w < kVcfThreads / 64; ++w) {
      nm += s_cnt[0][w];
      cp += s_cnt[1][w];
      sm += s_sum[w];
    }
    const int n_ok = a.n - nm;
    bool keep = n_ok > 0;
    if (keep && static_cast<double>(nm) / static_cast<double>(a.n) > a.max_miss + 1e-15) keep = false;
    double af = cp > 0 ? sm / static_cast<double>(cp) : 0.0;
    af = fmin(1.0, fmax(0.0, af));
    if (af < 1e-12 || af > 1.0 - 1e-12) keep = false;
    if (!a.impute && nm > 0) keep = false;
    if (a.maf_min > 0.0 && (af < a.maf_min || af > 1.0 - a.maf_min)) keep = false;  // pass_maf
    s_keep = keep ? 1 : 0;
  }
  __syncthreads();  // also orders this workgroup's row stores before the rewrite
  if (!s_keep)
    for (int i = tid; i < a.n_file; i += kVcfThreads) row[i] = __builtin_nanf("");
}

}  // namespace eqtl_amd
