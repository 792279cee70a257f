// Scalar VCF text parsing shared by k_vcf_decode and its host-side unit check:
// tab masks of a dword, the decimal -> float conversion of DS values, and the
// GT / DS subfield of one sample column (VcfSession::parse_record,
// vcf_session.cpp:186-263, over the VCF text form of the fields).
#pragma once
#include <math.h>
#include <stdint.h>
#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define EQ_HD __host__ __device__ inline
#else
#define EQ_HD inline  // host-only build of the same scalar code (unit checks)
#endif

namespace eqtl_amd {

// Bytes of a dword equal to '\t', as a 4-bit mask (bit b = byte b). The
// zero-byte test ((x & 0x7f..) + 0x7f..) | x has its high bit clear exactly for
// the zero bytes of x, with no carries between bytes.
EQ_HD unsigned tab_mask(unsigned v) {
  const unsigned x = v ^ 0x09090909u;
  const unsigned y = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
  return ((y >> 7) & 1u) | ((y >> 14) & 2u) | ((y >> 21) & 4u) | ((y >> 28) & 8u);
}

// Bytes b of span dword d that lie on the line: line position 4d - s0 + b in [0, L).
EQ_HD unsigned line_mask(int d, int s0, int L) {
  const int lo = s0 - 4 * d, hi = L + s0 - 4 * d;
  unsigned mk = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) mk |= (b >= lo && b < hi) ? (1u << b) : 0u;
  return mk;
}

EQ_HD bool is_digit(unsigned c) { return c - '0' < 10u; }

// Powers of ten that are exact doubles.
EQ_HD double exact_p10(int k) {
  double r = 1.0;
  for (int i = 0; i < k; ++i) r *= 10.0;  // exact for k <= 22
  return r;
}

// Decimal text -> float, as a float cast of the nearest double: M * 10^e with
// M < 2^53 and |e| <= 22 is one correctly rounded product or quotient; longer
// mantissas or exponents are scaled in steps. ok = false when no digit is read.
template <class Byte>
EQ_HD float parse_decimal(const Byte& at, int q, int L, bool& ok) {
  bool neg = false;
  if (q < L && (at(q) == '-' || at(q) == '+')) {
    neg = at(q) == '-';
    ++q;
  }
  unsigned long long M = 0;
  int e10 = 0;
  ok = false;
  while (q < L && is_digit(at(q))) {
    ok = true;
    if (M < 100000000000000000ull) M = M * 10 + (at(q) - '0');
    else ++e10;
    ++q;
  }
  if (q < L && at(q) == '.') {
    ++q;
    while (q < L && is_digit(at(q))) {
      ok = true;
      if (M < 100000000000000000ull) {
        M = M * 10 + (at(q) - '0');
        --e10;
      }
      ++q;
    }
  }
  if (!ok) return 0.f;
  if (q < L && (at(q) == 'e' || at(q) == 'E')) {
    int q2 = q + 1;
    bool eneg = false;
    if (q2 < L && (at(q2) == '-' || at(q2) == '+')) {
      eneg = at(q2) == '-';
      ++q2;
    }
    if (q2 < L && is_digit(at(q2))) {
      int ex = 0;
      while (q2 < L && is_digit(at(q2))) {
        if (ex < 100000) ex = ex * 10 + (at(q2) - '0');
        ++q2;
      }
      e10 += eneg ? -ex : ex;
    }
  }
  double v;
  if (M == 0) {
    v = 0.0;
  } else if (M < (1ull << 53) && e10 >= -22 && e10 <= 22) {
    v = e10 >= 0 ? static_cast<double>(M) * exact_p10(e10) : static_cast<double>(M) / exact_p10(-e10);
  } else {
    v = static_cast<double>(M);
    int e = e10;
    while (e > 22) { v *= 1e22; e -= 22; }
    while (e < -22) { v /= 1e22; e += 22; }
    v = e >= 0 ? v * exact_p10(e) : v / exact_p10(-e);
  }
  return static_cast<float>(neg ? -v : v);
}

struct VcfCall {
  float dose;
  bool miss;
  int copies;  // allele copies the call adds to the AF denominator
};

// One sample column starting at line position q: its GT or DS subfield.
template <class Byte>
EQ_HD VcfCall parse_sample(const Byte& at, int q, int L, int gt_key, int ds_key, bool hap) {
  VcfCall r{0.f, true, 0};
  const bool use_ds = ds_key >= 0;
  const int key = use_ds ? ds_key : gt_key;
  if (key < 0) return r;
  for (int f = 0; f < key; ++f) {  // skip to the key-th ':'-separated subfield
    while (q < L && at(q) != ':' && at(q) != '\t') ++q;
    if (q >= L || at(q) == '\t') return r;  // the column has fewer subfields: missing
    ++q;
  }
  if (use_ds) {
    if (q >= L || at(q) == ':' || at(q) == '\t') return r;
    if (at(q) == '.' && !(q + 1 < L && is_digit(at(q + 1)))) return r;
    bool ok;
    const float v = parse_decimal(at, q, L, ok);
    if (!ok || !isfinite(v)) return r;
    r.dose = v;
    r.miss = false;
    r.copies = hap ? 1 : 2;
    return r;
  }
  // GT: allele tokens separated by '/' or '|'; '.' = missing allele
  int na = 0, a0 = 0, a1 = 0;
  bool m0 = false, m1 = false;
  while (q < L) {
    const unsigned c = at(q);
    if (c == ':' || c == '\t') break;
    if (c == '/' || c == '|') { ++q; continue; }
    bool mis = false;
    int val = 0;
    if (c == '.') {
      mis = true;
      ++q;
    } else if (is_digit(c)) {
      while (q < L && is_digit(at(q))) {
        if (val < 10000000) val = val * 10 + (at(q) - '0');
        ++q;
      }
    } else {
      ++q;  // not part of a genotype
      continue;
    }
    if (na == 0) { a0 = val; m0 = mis; }
    else if (na == 1) { a1 = val; m1 = mis; }
    ++na;
  }
  if (na == 0 || m0 || (na >= 2 && m1)) return r;
  r.miss = false;
  r.dose = static_cast<float>(na >= 2 ? a0 + a1 : a0);
  r.copies = (hap || na == 1) ? 1 : 2;
  return r;
}

}  // namespace eqtl_amd
