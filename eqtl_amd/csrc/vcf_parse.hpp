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

// Extent [qs, qe) of the key-th ':'-separated subfield of the sample column
// starting at q, scanning no further than lim (lim_is_end: lim is the end of
// the line). Returns 1 found, 0 absent (the column ends first: a missing
// value), -1 undecided (lim was reached inside the column).
template <class Byte>
EQ_HD int locate_subfield(const Byte& at, int q, int lim, bool lim_is_end, int key, int& qs, int& qe) {
  for (int f = 0; f < key; ++f) {
    while (q < lim && at(q) != ':' && at(q) != '\t') ++q;
    if (q >= lim) return lim_is_end ? 0 : -1;
    if (at(q) == '\t') return 0;
    ++q;
  }
  qs = q;
  while (q < lim && at(q) != ':' && at(q) != '\t') ++q;
  if (q >= lim && !lim_is_end) return -1;
  qe = q;
  return 1;
}

// DS subfield [q, e): a float; '.' (not followed by a digit), empty, no
// digits or non-finite = missing.
template <class Byte>
EQ_HD VcfCall parse_ds(const Byte& at, int q, int e, bool hap) {
  VcfCall r{0.f, true, 0};
  if (q >= e) return r;
  if (at(q) == '.' && !(q + 1 < e && is_digit(at(q + 1)))) return r;
  bool ok;
  const float v = parse_decimal(at, q, e, ok);
  if (!ok || !isfinite(v)) return r;
  r.dose = v;
  r.miss = false;
  r.copies = hap ? 1 : 2;
  return r;
}

// GT subfield [q, e): allele tokens separated by '/' or '|', '.' = missing
// allele. Missing call when there is no allele or either of the first two is
// missing; dosage = sum of the first two allele indices; a single allele is a
// haploid call (one copy in the AF denominator).
template <class Byte>
EQ_HD VcfCall parse_gt(const Byte& at, int q, int e, bool hap) {
  VcfCall r{0.f, true, 0};
  int na = 0, a0 = 0, a1 = 0;
  bool m0 = false, m1 = false;
  while (q < e) {
    const unsigned c = at(q);
    if (c == '/' || c == '|') { ++q; continue; }
    bool mis = false;
    int val = 0;
    if (c == '.') {
      mis = true;
      ++q;
    } else if (is_digit(c)) {
      while (q < e && is_digit(at(q))) {
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

// One sample column starting at line position q (line length L): its DS
// subfield when ds_key >= 0, else its GT subfield.
template <class Byte>
EQ_HD VcfCall parse_sample(const Byte& at, int q, int L, int gt_key, int ds_key, bool hap) {
  VcfCall r{0.f, true, 0};
  const int key = ds_key >= 0 ? ds_key : gt_key;
  if (key < 0) return r;
  int qs = 0, qe = 0;
  if (locate_subfield(at, q, L, true, key, qs, qe) != 1) return r;
  return ds_key >= 0 ? parse_ds(at, qs, qe, hap) : parse_gt(at, qs, qe, hap);
}

// The sample column starting at byte o of a 16-byte window of the line (two
// little-endian 64-bit words lo, hi; room = line bytes from the window's first
// byte on). Returns false when its subfield runs past the window: the caller
// then parses the column from the line itself.
EQ_HD bool parse_sample_window(uint64_t lo, uint64_t hi, int o, int room, int gt_key, int ds_key, bool hap,
                               VcfCall& out) {
  out = VcfCall{0.f, true, 0};
  const int key = ds_key >= 0 ? ds_key : gt_key;
  if (key < 0) return true;
  auto at = [&](int j) -> unsigned {
    return static_cast<unsigned>((j < 8 ? lo >> (8 * j) : hi >> (8 * (j - 8))) & 0xffu);
  };
  const int lim = room < 16 ? room : 16;
  int qs = 0, qe = 0;
  const int st = locate_subfield(at, o, lim, room <= 16, key, qs, qe);
  if (st < 0) return false;
  if (st == 1) out = ds_key >= 0 ? parse_ds(at, qs, qe, hap) : parse_gt(at, qs, qe, hap);
  return true;
}

// 0x80 in every byte of x that is zero (exact: no carries between bytes).
EQ_HD uint64_t swar_zero_bytes(uint64_t x) {
  return ~(((x & 0x7f7f7f7f7f7f7f7full) + 0x7f7f7f7f7f7f7f7full) | x) & 0x8080808080808080ull;
}

// Branch-free decode of the commonest columns, GT first in FORMAT: a diploid
// call of one-digit alleles ("a|b", "a/b", '.' alleles) or a haploid one ("a"),
// and, when DS is the second key, a DS value of at most seven characters,
// digits and one '.' ("0.123", "1", ".5"), right after such a call. x0, x1 hold 16 bytes from the
// column start, of which win are window bytes and avail are on the line.
// Returns false when the column has another form or does not end inside the
// window (the caller then uses the general parse); otherwise the result is the
// general parse's for the same column.
EQ_HD bool parse_sample_fast(uint64_t x0, uint64_t x1, int win, int avail, int gt_key, int ds_key, bool hap,
                             VcfCall& out) {
  if (gt_key != 0 || (ds_key >= 0 && ds_key != 1)) return false;
  auto byte = [&](int i) -> unsigned {
    return static_cast<unsigned>(((i < 8) ? (x0 >> (8 * i)) : (x1 >> (8 * (i - 8)))) & 0xffu);
  };
  // 1: terminator (':', tab or end of line), 0: other byte, -1: past the window
  auto term = [&](int i) -> int {
    if (i >= avail) return 1;
    if (i >= win) return -1;
    const unsigned ch = byte(i);
    return (ch == ':' || ch == '\t') ? 1 : 0;
  };
  const unsigned b0 = byte(0), b1 = byte(1), b2 = byte(2);
  const bool al0 = avail >= 1 && (is_digit(b0) || b0 == '.');
  const bool al2 = avail >= 3 && (is_digit(b2) || b2 == '.');
  const bool dip = al0 && avail >= 2 && (b1 == '/' || b1 == '|') && al2 && term(3) == 1;
  const bool hpl = al0 && term(1) == 1;
  if (!dip && !hpl) return false;
  const int gt_len = dip ? 3 : 1;
  if (ds_key < 0) {
    out.miss = b0 == '.' || (dip && b2 == '.');
    out.dose = out.miss ? 0.f : static_cast<float>(static_cast<int>(b0) - '0' + (dip ? static_cast<int>(b2) - '0' : 0));
    out.copies = out.miss ? 0 : ((hap || !dip) ? 1 : 2);
    return true;
  }
  // DS after "GT:"; the column must continue with ':'
  if (gt_len >= avail || byte(gt_len) != ':') return false;
  const int q0 = gt_len + 1;  // 2 or 4: bytes q0..q0+7 are window bytes (win >= 12)
  const int sh = 8 * q0;
  const uint64_t v = (x0 >> sh) | (x1 << (64 - sh));
  // first terminator (':', tab, end of line) among the 8 bytes
  const uint64_t zt = swar_zero_bytes(v ^ 0x3a3a3a3a3a3a3a3aull) | swar_zero_bytes(v ^ 0x0909090909090909ull);
  int t = zt ? (__builtin_ctzll(zt) >> 3) : 8;
  if (avail - q0 < t) t = avail - q0;
  if (t >= 8) return false;  // longer value: general parse
  const uint64_t low = t == 0 ? 0ull : (~0ull >> (64 - 8 * t));
  const uint64_t dg = v ^ 0x3030303030303030ull;  // digits become 0..9
  const uint64_t not_digit = ((((dg & 0x7f7f7f7f7f7f7f7full) + 0x7676767676767676ull) | dg) & 0x8080808080808080ull) & low;
  const uint64_t dots = swar_zero_bytes(v ^ 0x2e2e2e2e2e2e2e2eull) & low;
  if ((not_digit & ~dots) != 0 || __builtin_popcountll(dots) > 1) return false;  // sign, exponent, ...
  const int nd = t - __builtin_popcountll(dots);
  if (nd == 0) {  // empty, or '.' alone: missing
    out = VcfCall{0.f, true, 0};
    return true;
  }
  uint64_t dd = dg & low;
  int frac = 0;
  if (dots) {  // drop the '.' byte
    const int p = __builtin_ctzll(dots) >> 3;
    frac = t - p - 1;
    const uint64_t keep = p == 0 ? 0ull : (~0ull >> (64 - 8 * p));
    dd = (dd & keep) | ((dd >> 8) & ~keep);
  }
  dd <<= 8 * (8 - nd);  // right-align: leading zero digits
  // eight ASCII-free digits, first in the lowest byte -> integer (the SWAR
  // reduction of Lemire's "parse_eight_digits_unrolled")
  dd = ((dd & 0x0f0f0f0f0f0f0f0full) * 2561) >> 8;
  dd = ((dd & 0x00ff00ff00ff00ffull) * 6553601) >> 16;
  const unsigned M = static_cast<unsigned>(((dd & 0x0000ffff0000ffffull) * 42949672960001ull) >> 32);
  double p10 = 1.0;  // 10^frac, frac <= 6: exact
  p10 = (frac & 1) ? p10 * 10.0 : p10;
  p10 = (frac & 2) ? p10 * 100.0 : p10;
  p10 = (frac & 4) ? p10 * 10000.0 : p10;
  out.miss = false;
  out.copies = hap ? 1 : 2;
  out.dose = static_cast<float>(static_cast<double>(M) / p10);
  return true;
}

}  // namespace eqtl_amd
