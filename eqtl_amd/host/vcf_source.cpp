// VCF genotype source of the host (VcfSession, vcf_session.cpp:51-306): file
// inflation, header, fixed columns of each record; the sample columns are
// decoded on the device (eqtl_gpu_vcf_decode, VcfSession::parse_record).
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <fstream>
#include <thread>
#include <unordered_set>

#include "eqtl_host.hpp"

namespace eqtl {

bool is_haploid_chrom(const std::string& chrom) {  // vcf_session.cpp:149-158
  std::string k = chrom;
  if (k.size() > 3 && (k[0] == 'c' || k[0] == 'C') && (k[1] == 'h' || k[1] == 'H') && (k[2] == 'r' || k[2] == 'R'))
    k = k.substr(3);
  return k == "X" || k == "x" || k == "Y" || k == "y" || k == "M" || k == "m" || k == "MT" || k == "mt" || k == "Mt";
}

namespace {

uint32_t le16(const unsigned char* p) { return p[0] | (static_cast<uint32_t>(p[1]) << 8); }
uint32_t le32(const unsigned char* p) {
  return p[0] | (static_cast<uint32_t>(p[1]) << 8) | (static_cast<uint32_t>(p[2]) << 16) |
         (static_cast<uint32_t>(p[3]) << 24);
}

struct Bgzf {
  size_t payload, plen, out;
  uint32_t isize, crc;
};

// BGZF blocks of a gzip file (each a gzip member carrying the "BC" extra
// subfield with its size); empty when the file is not BGZF.
std::vector<Bgzf> bgzf_blocks(const std::vector<unsigned char>& raw) {
  std::vector<Bgzf> v;
  size_t p = 0, out = 0;
  while (p < raw.size()) {
    if (p + 18 > raw.size() || raw[p] != 0x1f || raw[p + 1] != 0x8b || raw[p + 2] != 8 || !(raw[p + 3] & 4))
      return {};
    const size_t xlen = le16(&raw[p + 10]);
    if (p + 12 + xlen > raw.size()) return {};
    size_t blen = 0;
    for (size_t q = p + 12; q + 4 <= p + 12 + xlen;) {
      const size_t slen = le16(&raw[q + 2]);
      if (raw[q] == 'B' && raw[q + 1] == 'C' && slen == 2) blen = le16(&raw[q + 4]) + 1;
      q += 4 + slen;
    }
    if (blen == 0 || blen < 12 + xlen + 8 || p + blen > raw.size()) return {};
    Bgzf b;
    b.payload = p + 12 + xlen;
    b.plen = blen - (12 + xlen) - 8;
    b.crc = le32(&raw[p + blen - 8]);
    b.isize = le32(&raw[p + blen - 4]);
    b.out = out;
    out += b.isize;
    v.push_back(b);
    p += blen;
  }
  return v;
}

std::string inflate_members(const std::vector<unsigned char>& raw, const std::string& path) {
  std::string out;
  z_stream zs{};
  if (inflateInit2(&zs, 15 + 16) != Z_OK) die("zlib init failed");
  zs.next_in = const_cast<unsigned char*>(raw.data());
  zs.avail_in = static_cast<uInt>(std::min<size_t>(raw.size(), 1u << 30));
  size_t consumed = 0;
  std::vector<char> buf(1 << 20);
  for (;;) {
    zs.next_out = reinterpret_cast<unsigned char*>(buf.data());
    zs.avail_out = static_cast<uInt>(buf.size());
    const int rc = inflate(&zs, Z_NO_FLUSH);
    out.append(buf.data(), buf.size() - zs.avail_out);
    if (rc == Z_STREAM_END) {
      consumed += zs.total_in;
      if (consumed + 2 <= raw.size() && raw[consumed] == 0x1f && raw[consumed + 1] == 0x8b) {
        inflateReset(&zs);
        zs.next_in = const_cast<unsigned char*>(raw.data() + consumed);
        zs.avail_in = static_cast<uInt>(std::min<size_t>(raw.size() - consumed, 1u << 30));
        continue;
      }
      break;
    }
    if (rc != Z_OK) {
      inflateEnd(&zs);
      die("corrupt gzip stream: " + path);
    }
    if (zs.avail_in == 0) {
      const size_t done = consumed + zs.total_in;
      if (done >= raw.size()) {
        inflateEnd(&zs);
        die("truncated gzip stream: " + path);
      }
      zs.next_in = const_cast<unsigned char*>(raw.data() + done);
      zs.avail_in = static_cast<uInt>(std::min<size_t>(raw.size() - done, 1u << 30));
    }
  }
  inflateEnd(&zs);
  return out;
}

}  // namespace

std::string read_maybe_gzip(const std::string& path, int threads) {
  std::ifstream in(path, std::ios::binary);
  if (!in) die("cannot open VCF/BCF: " + path);
  std::vector<unsigned char> raw((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  if (raw.size() < 2 || raw[0] != 0x1f || raw[1] != 0x8b) return std::string(raw.begin(), raw.end());
  const std::vector<Bgzf> blocks = bgzf_blocks(raw);
  if (blocks.empty()) return inflate_members(raw, path);
  std::string out(blocks.back().out + blocks.back().isize, '\0');
  std::atomic<long> bad{-1};
  const int T = std::max(1, std::min<int>(threads, static_cast<int>(blocks.size())));
  auto work = [&](int t) {
    z_stream zs{};
    if (inflateInit2(&zs, -15) != Z_OK) { bad = 0; return; }
    const size_t per = (blocks.size() + T - 1) / T;
    for (size_t i = t * per; i < std::min(blocks.size(), (t + 1) * per); ++i) {
      const Bgzf& b = blocks[i];
      inflateReset(&zs);
      zs.next_in = const_cast<unsigned char*>(raw.data() + b.payload);
      zs.avail_in = static_cast<uInt>(b.plen);
      unsigned char* dst = reinterpret_cast<unsigned char*>(&out[0]) + b.out;
      zs.next_out = dst;
      zs.avail_out = b.isize;
      const int rc = inflate(&zs, Z_FINISH);
      if ((rc != Z_STREAM_END && !(rc == Z_BUF_ERROR && b.isize == 0)) || zs.total_out != b.isize ||
          crc32(0L, dst, b.isize) != b.crc) {
        bad = static_cast<long>(i);
        break;
      }
    }
    inflateEnd(&zs);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  if (bad >= 0) die("corrupt BGZF block " + std::to_string(bad.load()) + ": " + path);
  return out;
}

void VcfSource::open(const std::string& path, int threads) {  // vcf_session.cpp:51-104
  path_ = path;
  text_ = read_maybe_gzip(path, threads);
  if (text_.compare(0, 3, "BCF") == 0) die("BCF input is not supported by this engine (use VCF text): " + path);
  if (text_.compare(0, 16, "##fileformat=VCF") != 0) die("cannot read VCF header: " + path);
  parse_header_and_records();
  std::ifstream tbi(path + ".tbi"), csi(path + ".csi");
  indexed_ = static_cast<bool>(tbi) || static_cast<bool>(csi);
  info("vcf: " + std::to_string(samples_.size()) + " samples, " + std::to_string(sites_.size()) + " biallelic sites" +
       (indexed_ ? " (indexed)" : " (no index)"));
  if (n_multi_skip_ > 0) warn("VCF skipped multi-allelic sites: " + std::to_string(n_multi_skip_));
}

void VcfSource::parse_header_and_records() {
  const char* s = text_.data();
  const size_t N = text_.size();
  size_t pos = 0;
  bool have_chrom_line = false, any_ds = false;
  while (pos < N) {
    size_t nl = text_.find('\n', pos);
    if (nl == std::string::npos) nl = N;
    std::string line(s + pos, nl - pos);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    pos = nl + 1;
    if (line.compare(0, 2, "##") == 0) {
      const bool id_ds = line.find("<ID=DS,") != std::string::npos || line.find("<ID=DS>") != std::string::npos;
      any_ds |= id_ds;
      if (id_ds && line.compare(0, 10, "##FORMAT=<") == 0 && line.find("Type=Float") != std::string::npos)
        ds_float_ = true;
      continue;
    }
    if (line.compare(0, 6, "#CHROM") == 0) {
      auto cols = split_char(line, '\t');
      for (size_t i = 9; i < cols.size(); ++i) samples_.push_back(cols[i]);
      std::unordered_set<std::string> seen;
      for (const auto& x : samples_)
        if (!seen.insert(x).second) die("duplicate ID in VCF samples: " + x);
      have_chrom_line = true;
      break;
    }
    die("cannot read VCF header (line before #CHROM): " + path_);
  }
  if (!have_chrom_line) die("cannot read VCF header (no #CHROM line): " + path_);
  if (any_ds) info("VCF: FORMAT/DS present → use dosage field when available");
  const size_t ns = samples_.size();
  // data lines: fixed columns on the host, sample columns left for the device
  size_t line_no = 0;
  while (pos < N) {
    const char* hit = static_cast<const char*>(memchr(s + pos, '\n', N - pos));
    const size_t nl = hit ? static_cast<size_t>(hit - s) : N;
    size_t end = nl;
    if (end > pos && s[end - 1] == '\r') --end;
    const size_t beg = pos;
    pos = nl + 1;
    ++line_no;
    if (end == beg || s[beg] == '#') continue;
    size_t tab[9];
    int nt = 0;
    for (size_t q = beg; q < end && nt < 9;) {
      const char* t = static_cast<const char*>(memchr(s + q, '\t', end - q));
      if (!t) break;
      tab[nt++] = static_cast<size_t>(t - s);
      q = tab[nt - 1] + 1;
    }
    if (nt < 7 || (ns > 0 && nt < 9)) die("VCF data line " + std::to_string(line_no) + " has too few columns: " + path_);
    const std::string chrom(s + beg, tab[0] - beg);
    const std::string pos_s(s + tab[0] + 1, tab[1] - tab[0] - 1);
    const std::string id(s + tab[1] + 1, tab[2] - tab[1] - 1);
    const std::string ref(s + tab[2] + 1, tab[3] - tab[2] - 1);
    const std::string alt(s + tab[3] + 1, tab[4] - tab[3] - 1);
    const int n_allele = alt == "." ? 1 : 2 + static_cast<int>(std::count(alt.begin(), alt.end(), ','));
    if (n_allele != 2) {
      ++n_multi_skip_;
      continue;
    }
    char* e = nullptr;
    const long long p1 = std::strtoll(pos_s.c_str(), &e, 10);
    if (pos_s.empty() || !e || *e) die("VCF data line " + std::to_string(line_no) + ": bad POS '" + pos_s + "'");
    Rec r{0, 0, -1, -1, is_haploid_chrom(chrom) ? 1 : 0};
    if (ns > 0) {
      const std::string fmt(s + tab[7] + 1, tab[8] - tab[7] - 1);
      const auto keys = split_char(fmt, ':');
      for (size_t k = 0; k < keys.size() && k < 32767; ++k) {
        if (keys[k] == "GT" && r.gt_key < 0) r.gt_key = static_cast<int16_t>(k);
        if (keys[k] == "DS" && ds_float_ && r.ds_key < 0) r.ds_key = static_cast<int16_t>(k);
      }
      r.offset = static_cast<int64_t>(tab[8] + 1);
      if (end - (tab[8] + 1) > 0x7fffffffu) die("VCF data line " + std::to_string(line_no) + " is too long");
      r.length = static_cast<int32_t>(end - (tab[8] + 1));
    }
    BimSite st;
    st.chrom = chrom;
    st.pos = p1;
    st.a1 = alt;
    st.a2 = ref;
    // fill_snp_id (scan_common.cpp:153-156)
    st.id = (!id.empty() && id != ".") ? id : chrom + ":" + std::to_string(p1) + ":" + ref + ":" + alt;
    recs_.push_back(r);
    sites_.push_back(std::move(st));
  }
  if (sites_.empty()) die("no biallelic records in VCF: " + path_);
  // one contiguous, position-sorted block per chromosome (bgzip + tabix order)
  std::unordered_set<std::string> seen;
  for (size_t i = 0; i < sites_.size(); ++i) {
    const bool first = i == 0 || !chrom_equal(sites_[i - 1].chrom, sites_[i].chrom);
    if (first) {
      if (!seen.insert(chrom_key(sites_[i].chrom)).second)
        die("VCF: chromosome " + sites_[i].chrom + " appears in more than one block (sort the file)");
    } else if (sites_[i - 1].pos > sites_[i].pos) {
      die("VCF: positions not sorted within chromosome " + sites_[i].chrom);
    }
  }
}

const uint8_t* VcfSource::rows(Engine& eng, size_t j0, size_t m, int32_t* on_dev) {
  if (m == 0 || j0 + m > recs_.size()) die("VCF rows out of range");
  const int64_t base = recs_[j0].offset;
  const int64_t end = recs_[j0 + m - 1].offset + recs_[j0 + m - 1].length;
  desc_.resize(m);
  for (size_t i = 0; i < m; ++i) {
    const Rec& r = recs_[j0 + i];
    desc_[i] = eqtl_vcf_record{r.offset - base, r.length, r.gt_key, r.ds_key, r.haploid, 0};
  }
  const uint8_t* dev = nullptr;
  eng.check(eqtl_gpu_vcf_decode(eng.ctx(), text_.data() + base, 0, end - base, desc_.data(),
                                static_cast<int64_t>(m), &dev),
            "vcf_decode");
  *on_dev = 1;
  return dev;
}

const uint8_t* VcfSource::gather(Engine& eng, const std::vector<int64_t>& idx, int32_t* on_dev) {
  if (idx.empty()) die("VCF gather of no rows");
  gather_text_.clear();
  desc_.resize(idx.size());
  for (size_t i = 0; i < idx.size(); ++i) {
    const Rec& r = recs_.at(static_cast<size_t>(idx[i]));
    desc_[i] = eqtl_vcf_record{static_cast<int64_t>(gather_text_.size()), r.length, r.gt_key, r.ds_key, r.haploid, 0};
    gather_text_.append(text_, static_cast<size_t>(r.offset), static_cast<size_t>(r.length));
  }
  const uint8_t* dev = nullptr;
  eng.check(eqtl_gpu_vcf_decode(eng.ctx(), gather_text_.data(), 0, static_cast<int64_t>(gather_text_.size()),
                                desc_.data(), static_cast<int64_t>(idx.size()), &dev),
            "vcf_decode");
  *on_dev = 1;
  return dev;
}

}  // namespace eqtl
