// Host check of the scalar VCF parsing shared with k_vcf_decode
// (vcf_parse.hpp): tab masks against a byte loop, and the decimal -> float
// conversion of DS values against strtod -> float (htslib's float fields are
// the float cast of the parsed double).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "../eqtl_amd/csrc/vcf_parse.hpp"

using namespace eqtl_amd;

int main() {
  std::mt19937_64 r(7);
  long fails = 0;
  // tab_mask: bytes drawn from a set that stresses carries between bytes
  const unsigned pick[] = {0x09, 0x08, 0x0a, 0x89, 0x00, 0xff, 0x7f, 0x80, 0x30, 0x3a};
  for (int it = 0; it < 2000000; ++it) {
    unsigned v = 0, want = 0;
    for (int b = 0; b < 4; ++b) {
      const unsigned c = (it & 1) ? pick[r() % 10] : static_cast<unsigned>(r() & 0xff);
      v |= c << (8 * b);
      if (c == 0x09) want |= 1u << b;
    }
    if (tab_mask(v) != want) ++fails;
  }
  // line_mask: brute force
  for (int d = 0; d < 40; ++d)
    for (int s0 = 0; s0 < 4; ++s0)
      for (int L = 0; L < 160; ++L) {
        unsigned want = 0;
        for (int b = 0; b < 4; ++b) {
          const int pos = 4 * d - s0 + b;
          if (pos >= 0 && pos < L) want |= 1u << b;
        }
        if (line_mask(d, s0, L) != want) ++fails;
      }
  if (fails) {
    std::printf("mask failures: %ld\n", fails);
    return 1;
  }
  // decimals: the forms DS takes, plus long mantissas and exponents
  std::uniform_real_distribution<double> u(0, 1);
  char buf[96];
  long n = 0, long_mis = 0;
  for (int it = 0; it < 3000000; ++it) {
    const double v = u(r) * ((it % 7 == 0) ? 1e6 : 2.0);
    const int form = it % 9;
    switch (form) {
      case 0: std::snprintf(buf, sizeof buf, "%.1f", v); break;
      case 1: std::snprintf(buf, sizeof buf, "%.3f", v); break;
      case 2: std::snprintf(buf, sizeof buf, "%.6f", v); break;
      case 3: std::snprintf(buf, sizeof buf, "%.9f", v); break;
      case 4: std::snprintf(buf, sizeof buf, "%.4e", v); break;
      case 5: std::snprintf(buf, sizeof buf, "%.3g", v * 1e-5); break;
      case 6: std::snprintf(buf, sizeof buf, "-%.5f", v); break;
      case 7: std::snprintf(buf, sizeof buf, "%.17g", v); break;
      default: std::snprintf(buf, sizeof buf, "%.25f", v * 1e-3); break;
    }
    const std::string s(buf);
    const int L = static_cast<int>(s.size());
    auto at = [&](int q) -> unsigned { return static_cast<unsigned char>(s[q]); };
    bool ok = false;
    const float got = parse_decimal(at, 0, L, ok);
    const float want = static_cast<float>(std::strtod(buf, nullptr));
    int sig = 0;
    bool lead = true;
    for (char c : s) {
      if (c >= '1' && c <= '9') lead = false;
      if (c == 'e' || c == 'E') break;
      if (c >= '0' && c <= '9' && !lead) ++sig;
    }
    ++n;
    if (!ok) { ++fails; continue; }
    if (sig <= 15) {
      if (std::memcmp(&got, &want, sizeof got) != 0) {
        if (fails < 10) std::printf("mismatch %s: %.9g vs %.9g\n", buf, got, want);
        ++fails;
      }
    } else if (got != want) {
      // more digits than an exact double mantissa: at most one float ulp
      if (std::fabs(got - want) > std::ldexp(std::fabs(want), -23)) ++fails;
      else ++long_mis;
    }
  }
  // non-numbers
  for (const char* bad : {"", ".", "nan", "inf", "-", "+", "e5", "."}) {
    const std::string s(bad);
    auto at = [&](int q) -> unsigned { return static_cast<unsigned char>(s[q]); };
    bool ok = true;
    parse_decimal(at, 0, static_cast<int>(s.size()), ok);
    if (ok) ++fails;
  }
  std::printf("decimals %ld, long-mantissa one-ulp differences %ld, failures %ld\n", n, long_mis, fails);
  // register-window parse of a column (the kernel's fast path) against the
  // whole-line parse, for every column start and every alignment of the line
  const char* cells[] = {"0|1", "1/1", "./.", ".", "1", "0/1/1", ".|0", "10|2", "0|0:0.25", "1|1:.", "0/1:1.5e-1",
                         "0|1:0.333333333:9", "./.:0.1", ":", "", "0|0:123456.78912", "1/1:7:8:9:0.5",
                         "0|1:12345678901234567", "0/0:-0.5", "0|1:0.123", "1:1.5", "0/0:", "1|1:2", "0|1:.5",
                         "1|0:5.", "0|1:0.1.2", "0|0:.", ".|.:0.000", "1/1:1.999:7", "0|1x", "1|", "|1",
                         "0|1:00012.5000", "0/1:0.12345678901"};
  long cols = 0, fallback = 0, wfail = 0, fast = 0;
  for (int it = 0; it < 20000; ++it) {
    const int ncol = 1 + static_cast<int>(r() % 12);
    std::string line;
    for (int k = 0; k < ncol; ++k) {
      if (k) line += '\t';
      line += cells[r() % (sizeof(cells) / sizeof(cells[0]))];
    }
    const int L = static_cast<int>(line.size());
    const int gt_key = (r() & 1) ? 0 : static_cast<int>(r() % 3) - 1;  // mostly 0, also -1, 1
    const int ds_key = (r() & 1) ? ((r() & 1) ? 1 : static_cast<int>(r() % 4) - 1) : -1;
    const bool hap = r() & 1;
    for (int s0 = 0; s0 < 4; ++s0) {
      // the line at byte s0 of an aligned buffer, zero padded
      std::string buf(static_cast<size_t>(s0), 'Z');
      buf += line;
      buf.append(32, '\0');
      auto word = [&](int d) -> uint64_t {
        uint32_t w;
        std::memcpy(&w, buf.data() + 4 * d, 4);
        return w;
      };
      auto at = [&](int q) -> unsigned { return static_cast<unsigned char>(line[q]); };
      for (int p = 0; p < L + 1; ++p) {
        if (!(p == 0 || line[p - 1] == '\t')) continue;
        const int anchor = p == 0 ? s0 : p - 1 + s0;  // aligned-buffer byte of the column start / its tab
        const int d = anchor / 4;
        const int o = p == 0 ? s0 : (anchor % 4) + 1;
        const uint64_t lo = word(d) | (word(d + 1) << 32), hi = word(d + 2) | (word(d + 3) << 32);
        const int wbase = 4 * d - s0;
        VcfCall w;
        const VcfCall ref = parse_sample(at, p, L, gt_key, ds_key, hap);
        ++cols;
        {  // the fast path on the same window bytes, when it takes the column
          const int sh = 8 * o;
          const uint64_t x0 = o == 0 ? lo : ((lo >> sh) | (hi << (64 - sh))), x1 = o == 0 ? hi : (hi >> sh);
          VcfCall f;
          if (parse_sample_fast(x0, x1, 16 - o, L - wbase - o, gt_key, ds_key, hap, f)) {
            ++fast;
            const bool same = f.miss == ref.miss && f.copies == ref.copies &&
                              (f.miss || std::memcmp(&f.dose, &ref.dose, sizeof(float)) == 0);
            if (!same) {
              if (wfail < 5) std::printf("fast mismatch: '%s' p=%d s0=%d\n", line.c_str(), p, s0);
              ++wfail;
            }
          }
        }
        if (!parse_sample_window(lo, hi, o, L - wbase, gt_key, ds_key, hap, w)) {
          ++fallback;
          continue;
        }
        const bool same = w.miss == ref.miss && w.copies == ref.copies &&
                          (w.miss || std::memcmp(&w.dose, &ref.dose, sizeof(float)) == 0);
        if (!same) {
          if (wfail < 5) std::printf("window mismatch: '%s' p=%d s0=%d\n", line.c_str(), p, s0);
          ++wfail;
        }
      }
    }
  }
  std::printf("window columns %ld, fast %ld, fallbacks %ld, mismatches %ld\n", cols, fast, fallback, wfail);
  if (wfail || fallback == 0 || fallback == cols || fast == 0) ++fails;
  return fails ? 1 : 0;
}
