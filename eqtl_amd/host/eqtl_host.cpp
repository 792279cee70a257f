// Host side of the engine (see eqtl_host.hpp). Each function names the
// reference code whose behaviour it mirrors.
#include "eqtl_host.hpp"
#include "pheno_norm.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <map>
#include <numeric>
#include <sstream>
#include <unordered_set>

namespace eqtl {

// ============================================================ util
void die(const std::string& msg) {
  std::cerr << "[E] " << msg << "\n";
  throw std::runtime_error(msg);
}
void warn(const std::string& msg) { std::cerr << "[W] " << msg << "\n"; }
void info(const std::string& msg) { std::cerr << "[I] " << msg << "\n"; }

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
  return s.substr(a, b - a);
}

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream iss(s);
  std::string t;
  while (iss >> t) out.push_back(t);
  return out;
}

std::vector<std::string> split_char(const std::string& s, char c) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == c) {
      out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(ch);
    }
  }
  out.push_back(cur);
  return out;
}

std::string chrom_key(const std::string& chrom) {  // util.cpp:68-83
  std::string s = chrom;
  if (s.size() >= 3 && std::tolower(static_cast<unsigned char>(s[0])) == 'c' &&
      std::tolower(static_cast<unsigned char>(s[1])) == 'h' && std::tolower(static_cast<unsigned char>(s[2])) == 'r')
    s = s.substr(3);
  size_t start = 0;
  while (start + 1 < s.size() && s[start] == '0') ++start;
  return s.substr(start);
}

bool chrom_equal(const std::string& a, const std::string& b) { return a == b || chrom_key(a) == chrom_key(b); }

std::vector<std::string> intersect_order(const std::vector<std::string>& prefer,
                                         const std::vector<std::string>& other) {
  std::unordered_set<std::string> m(other.begin(), other.end());
  std::vector<std::string> out;
  for (const auto& id : prefer)
    if (m.count(id)) out.push_back(id);
  return out;
}

static void assert_unique(const std::vector<std::string>& ids, const std::string& what) {
  std::unordered_set<std::string> seen;
  for (const auto& id : ids)
    if (!seen.insert(id).second) die(what + ": duplicate id: " + id);
}

std::string model_str(Model m) {
  switch (m) {
    case Model::Lm: return "lm";
    case Model::Glm: return "glm";
    case Model::Lmm: return "lmm";
    case Model::Glmm: return "glmm";
  }
  return "?";
}

// ============================================================ PLINK bed
static std::vector<std::string> read_fam(const std::string& path) {  // plink_bed.cpp:26-40
  std::ifstream in(path);
  if (!in) die("cannot open " + path);
  std::vector<std::string> s;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    auto t = split_ws(line);
    if (t.size() < 2) die("malformed fam line (need FID IID ...): " + path);
    s.push_back(t[1]);
  }
  if (s.empty()) die("empty fam: " + path);
  assert_unique(s, "plink fam IID");
  return s;
}

static std::vector<BimSite> read_bim(const std::string& path) {  // plink_bed.cpp:42-160
  std::ifstream in(path);
  if (!in) die("cannot open " + path);
  std::vector<BimSite> v;
  std::string line;
  while (std::getline(in, line)) {
    auto t = split_ws(line);
    if (t.empty()) continue;
    if (t.size() < 6) die("malformed bim line in: " + path);
    BimSite s;
    s.chrom = t[0];
    s.id = t[1];
    char* end = nullptr;
    s.pos = std::strtoll(t[3].c_str(), &end, 10);
    if (!end || *end) die("malformed bim pos in: " + path);
    s.a1 = t[4];
    s.a2 = t[5];
    v.push_back(std::move(s));
  }
  if (v.empty()) die("empty bim: " + path);
  // one contiguous, position-sorted block per chromosome (plink_bed.cpp:143-159)
  std::unordered_set<std::string> seen;
  for (size_t i = 0; i < v.size(); ++i) {
    const bool first = i == 0 || !chrom_equal(v[i - 1].chrom, v[i].chrom);
    if (first) {
      if (!seen.insert(chrom_key(v[i].chrom)).second)
        die("plink bim: chromosome " + v[i].chrom + " appears in more than one block");
    } else if (v[i - 1].pos > v[i].pos) {
      die("plink bim: positions not sorted within chromosome " + v[i].chrom);
    }
  }
  return v;
}

void PlinkBed::open(const std::string& prefix) {  // plink_bed.cpp:187-230
  prefix_ = prefix;
  samples_ = read_fam(prefix + ".fam");
  sites_ = read_bim(prefix + ".bim");
  bed_ = std::make_unique<std::ifstream>(prefix + ".bed", std::ios::binary);
  if (!*bed_) die("cannot open " + prefix + ".bed");
  unsigned char magic[3];
  bed_->read(reinterpret_cast<char*>(magic), 3);
  if (bed_->gcount() != 3) die("cannot read bed magic: " + prefix + ".bed");
  if (magic[0] != 0x6c || magic[1] != 0x1b) die("not a PLINK1 .bed (bad magic): " + prefix + ".bed");
  if (magic[2] != 0x01) die("only SNP-major PLINK .bed is supported: " + prefix + ".bed");
  bps_ = static_cast<int32_t>((samples_.size() + 3) / 4);
  bed_->seekg(0, std::ios::end);
  const uint64_t fsz = static_cast<uint64_t>(bed_->tellg());
  if (fsz < 3ull + static_cast<uint64_t>(bps_) * sites_.size()) die("bed size too small for bim/fam: " + prefix + ".bed");
  info("bfile: " + std::to_string(samples_.size()) + " samples, " + std::to_string(sites_.size()) + " SNPs [" +
       prefix + "]");
}

std::vector<std::string> GenoSource::chromosomes() const {
  std::vector<std::string> out;
  const auto& st = sites();
  for (size_t i = 0; i < st.size(); ++i)
    if (i == 0 || !chrom_equal(st[i].chrom, out.back())) out.push_back(st[i].chrom);
  return out;
}

const uint8_t* PlinkBed::rows(Engine&, size_t j0, size_t m, int32_t* on_dev) {
  read_rows(j0, m, buf_);
  *on_dev = 0;
  return buf_.data();
}

const uint8_t* PlinkBed::gather(Engine&, const std::vector<int64_t>& idx, int32_t* on_dev) {
  buf_.clear();
  for (int64_t j : idx) {
    read_rows(static_cast<size_t>(j), 1, row_);
    buf_.insert(buf_.end(), row_.begin(), row_.begin() + bps_);
  }
  *on_dev = 0;
  return buf_.data();
}

void PlinkBed::read_rows(size_t j0, size_t m, std::vector<uint8_t>& buf) {
  buf.resize(m * static_cast<size_t>(bps_));
  bed_->clear();
  bed_->seekg(static_cast<std::streamoff>(3 + j0 * static_cast<size_t>(bps_)), std::ios::beg);
  bed_->read(reinterpret_cast<char*>(buf.data()), static_cast<std::streamsize>(buf.size()));
  if (static_cast<size_t>(bed_->gcount()) != buf.size()) die("bed fread short at SNP " + std::to_string(j0));
}

// ============================================================ GFF3 genes
static std::string attr_value(const std::string& attrs, const std::string& key) {
  for (const auto& kv : split_char(attrs, ';')) {
    const auto eq = kv.find('=');
    if (eq == std::string::npos) continue;
    if (trim(kv.substr(0, eq)) == key) {
      std::string v = trim(kv.substr(eq + 1));
      const auto comma = v.find(',');
      return comma == std::string::npos ? v : v.substr(0, comma);
    }
  }
  return {};
}

std::unordered_map<std::string, GeneLoc> load_gff_tss(const std::string& path,
                                                      const std::string& id_key) {  // annot.cpp:26-71
  std::ifstream in(path);
  if (!in) die("cannot parse gff: " + path);
  std::unordered_map<std::string, GeneLoc> out;
  std::unordered_set<std::string> multi;
  std::string line;
  int n_gene = 0;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    auto f = split_char(line, '\t');
    if (f.size() < 9 || f[2] != "gene") continue;
    ++n_gene;
    std::string id;
    if (!id_key.empty()) id = attr_value(f[8], id_key);
    else {
      id = attr_value(f[8], "ID");
      if (id.empty()) id = attr_value(f[8], "Name");
      if (id.empty()) id = attr_value(f[8], "gene_id");
    }
    if (id.empty()) continue;
    if (id.rfind("gene:", 0) == 0) id = id.substr(5);
    const char strand = (f[6] == "-" || f[6] == "+") ? f[6][0] : '+';
    const int64_t start = std::atoll(f[3].c_str()), end = std::atoll(f[4].c_str());
    if (out.count(id) || multi.count(id)) {
      if (out.count(id)) {
        warn("GFF multi-hit gene skipped: " + id);
        out.erase(id);
        multi.insert(id);
      }
      continue;
    }
    GeneLoc g;
    g.id = id;
    g.chrom = f[0];
    g.start = start;
    g.end = end;
    g.tss = strand == '-' ? end : start;
    g.strand = strand;
    g.ok = true;
    out.emplace(id, std::move(g));
  }
  info("gff: " + std::to_string(n_gene) + " gene features, " + std::to_string(out.size()) + " unique ids with TSS");
  return out;
}

// ============================================================ outputs (output.cpp:8-110)
void write_pairs_header(std::ostream& os, Model m) {  // output.cpp:8-13
  os << "gene\tsnp\tchrom\tpos\tref\talt\taf\tbeta\tse\tstat\tp\tr2\tn\ttss_dist\tscope";
  if (m == Model::Glm) os << "\tphi\tglm_converged";
  os << '\n';
}

void write_pair_line(std::ostream& os, const AssocHit& h, Model m, const std::string& scope) {  // output.cpp:27-76
  char buf[2048];
  int n;
  if (h.has_tss_dist)
    n = std::snprintf(buf, sizeof(buf), "%s\t%s\t%s\t%lld\t%s\t%s\t%.10g\t%.10g\t%.10g\t%.10g\t%.10g\t%.10g\t%d\t%.10g\t%s",
                      h.gene.c_str(), h.snp.c_str(), h.chrom.c_str(), static_cast<long long>(h.pos), h.ref.c_str(),
                      h.alt.c_str(), h.maf, h.beta, h.se, h.stat, h.p, h.r2, h.n, h.tss_dist, scope.c_str());
  else
    n = std::snprintf(buf, sizeof(buf), "%s\t%s\t%s\t%lld\t%s\t%s\t%.10g\t%.10g\t%.10g\t%.10g\t%.10g\t%.10g\t%d\tNA\t%s",
                      h.gene.c_str(), h.snp.c_str(), h.chrom.c_str(), static_cast<long long>(h.pos), h.ref.c_str(),
                      h.alt.c_str(), h.maf, h.beta, h.se, h.stat, h.p, h.r2, h.n, scope.c_str());
  if (n > 0 && m == Model::Glm && static_cast<size_t>(n) < sizeof(buf))
    n += std::snprintf(buf + n, sizeof(buf) - static_cast<size_t>(n), "\t%.10g\t%d", h.phi, h.glm_converged ? 1 : 0);
  if (n > 0 && static_cast<size_t>(n) + 1 < sizeof(buf)) {
    buf[n++] = '\n';
    os.write(buf, n);
  } else {
    die("pair line too long");
  }
}

void emit_line(ScopeOut& out, bool top, int64_t k1, int64_t k2, const AssocHit& h, Model m, const std::string& scope) {
  if (!out.buffered) {
    write_pair_line(top ? out.top : out.pairs, h, m, scope);
    return;
  }
  std::ostringstream os;
  write_pair_line(os, h, m, scope);
  (top ? out.top_buf : out.pair_buf).emplace_back(std::make_pair(k1, k2), os.str());
}

void flush_lines(ScopeOut& out) {
  for (auto* b : {&out.pair_buf, &out.top_buf}) {
    std::stable_sort(b->begin(), b->end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    std::ostream& os = b == &out.pair_buf ? static_cast<std::ostream&>(out.pairs) : static_cast<std::ostream&>(out.top);
    for (const auto& kv : *b) os << kv.second;
    b->clear();
  }
}

// looks_like_counts (model_lm.cpp:8-17): finite values are non-negative integers.
static bool looks_like_counts(const double* y, size_t n) {
  size_t seen = 0;
  for (size_t i = 0; i < n; ++i) {
    if (!std::isfinite(y[i])) continue;
    ++seen;
    if (y[i] < 0 || std::fabs(y[i] - std::floor(y[i])) > 1e-8) return false;
  }
  return seen > 0;
}

void write_region_header(std::ostream& os) {
  os << "gene\tchrom\ttss\tn_tested\tn_sig\tacat_p\tq_bh\tp_emp\tp_beta\tbeta_shape1\tbeta_shape2\n";
}

static void fmt_na_or(char* out, size_t n, double v) {
  if (!std::isfinite(v)) std::snprintf(out, n, "NA");
  else std::snprintf(out, n, "%.10g", v);
}

void write_region_line(std::ostream& os, const GeneSummary& g) {
  char p_emp[32], p_beta[32], q_bh[32], buf[1024];
  fmt_na_or(p_emp, sizeof(p_emp), g.p_emp);
  fmt_na_or(p_beta, sizeof(p_beta), g.p_beta);
  fmt_na_or(q_bh, sizeof(q_bh), g.q_bh);
  const int n = std::snprintf(buf, sizeof(buf), "%s\t%s\t%lld\t%d\t%d\t%.10g\t%s\t%s\t%s\t%.10g\t%.10g\n", g.gene.c_str(),
                              g.chrom.c_str(), static_cast<long long>(g.tss), g.n_tested, g.n_sig, g.acat_p, q_bh,
                              p_emp, p_beta, g.beta_shape1, g.beta_shape2);
  if (n > 0 && static_cast<size_t>(n) < sizeof(buf)) os.write(buf, n);
}

void bh_adjust(std::vector<GeneSummary>& gs) {  // scan_common.cpp:201-217
  const int m = static_cast<int>(gs.size());
  if (m == 0) return;
  std::vector<int> idx(m);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return gs[a].acat_p < gs[b].acat_p; });
  double prev = 1.0;
  for (int rank = m; rank >= 1; --rank) {
    auto& g = gs[idx[rank - 1]];
    double q = g.acat_p * static_cast<double>(m) / static_cast<double>(rank);
    if (q > prev) q = prev;
    if (q > 1.0) q = 1.0;
    g.q_bh = q;
    prev = q;
  }
}

// ============================================================ GRM files (grm.cpp:10-108)
Grm load_grm_gcta(const std::string& prefix) {
  Grm g;
  std::ifstream idf(prefix + ".grm.id");
  if (!idf) die("cannot open " + prefix + ".grm.id");
  std::string line;
  while (std::getline(idf, line)) {
    auto t = split_ws(line);
    if (t.size() >= 2) g.ids.push_back(t[1]);
    else if (t.size() == 1) g.ids.push_back(t[0]);
  }
  const int n = g.n();
  if (n == 0) die("empty GRM id file");
  const uint64_t ntri = static_cast<uint64_t>(n) * (n + 1) / 2;
  std::ifstream bf(prefix + ".grm.bin", std::ios::binary);
  if (!bf) die("cannot open " + prefix + ".grm.bin");
  std::vector<float> buf(ntri);
  bf.read(reinterpret_cast<char*>(buf.data()), static_cast<std::streamsize>(ntri * sizeof(float)));
  if (static_cast<uint64_t>(bf.gcount()) != ntri * sizeof(float)) die("GRM bin size mismatch");
  g.K.assign(static_cast<size_t>(n) * n, 0.0);
  size_t k = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      const double v = buf[k++];
      g.K[static_cast<size_t>(j) * n + i] = v;
      g.K[static_cast<size_t>(i) * n + j] = v;
    }
  info("grm: " + std::to_string(n) + " samples from " + prefix);
  return g;
}

void write_grm_gcta(const std::string& prefix, const Grm& g) {
  const int n = g.n();
  std::ofstream idf(prefix + ".grm.id");
  if (!idf) die("cannot write " + prefix + ".grm.id");
  for (const auto& id : g.ids) idf << id << '\t' << id << '\n';
  const uint64_t ntri = static_cast<uint64_t>(n) * (n + 1) / 2;
  std::vector<float> buf(ntri);
  size_t k = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) buf[k++] = static_cast<float>(g.K[static_cast<size_t>(j) * n + i]);
  std::ofstream bf(prefix + ".grm.bin", std::ios::binary);
  if (!bf) die("cannot write " + prefix + ".grm.bin");
  bf.write(reinterpret_cast<const char*>(buf.data()), static_cast<std::streamsize>(ntri * sizeof(float)));
  // .grm.N.bin only when the pair counts are known (a loaded GRM has none)
  if (g.N_packed.empty()) {
    info("wrote GRM " + prefix + ".grm.id/.grm.bin");
    return;
  }
  if (g.N_packed.size() != ntri) die("GRM N_packed size mismatch on write");
  for (size_t t = 0; t < ntri; ++t) buf[t] = static_cast<float>(g.N_packed[t]);
  std::ofstream nf(prefix + ".grm.N.bin", std::ios::binary);
  if (!nf) die("cannot write " + prefix + ".grm.N.bin");
  nf.write(reinterpret_cast<const char*>(buf.data()), static_cast<std::streamsize>(ntri * sizeof(float)));
  info("wrote GRM " + prefix + ".grm.id/.grm.bin/.grm.N.bin");
}

Grm slice_grm(const Grm& g, const std::vector<std::string>& sample_ids) {
  std::unordered_map<std::string, int> m;
  for (int i = 0; i < g.n(); ++i) m[g.ids[i]] = i;
  Grm o;
  o.ids = sample_ids;
  const int n = static_cast<int>(sample_ids.size());
  std::vector<int> ix(n);
  for (int i = 0; i < n; ++i) {
    auto it = m.find(sample_ids[i]);
    if (it == m.end()) die("GRM missing sample: " + sample_ids[i]);
    ix[i] = it->second;
  }
  o.K.resize(static_cast<size_t>(n) * n);
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < n; ++r) o.K[static_cast<size_t>(c) * n + r] = g.K[static_cast<size_t>(ix[c]) * g.n() + ix[r]];
  return o;
}

// ============================================================ engine
Engine::Engine(int device) { check(eqtl_gpu_open(device, &ctx_), "eqtl_gpu_open"); }
Engine::~Engine() {
  if (ctx_) eqtl_gpu_close(ctx_);
}
void Engine::check(int rc, const char* what) const {
  if (rc != EQTL_OK) die(std::string(what) + ": " + eqtl_gpu_last_error());
}

void Engine::scan_block(const uint8_t* rows, int32_t on_dev, int64_t snp0, int64_t m, int32_t bps,
                        std::vector<eqtl_hit>& hits) const {
  if (hits.capacity() < (1u << 16)) hits.reserve(1u << 16);
  hits.resize(hits.capacity());
  int64_t total = 0;
  check(eqtl_gpu_scan_block(ctx_, rows, on_dev, snp0, m, bps, hits.data(), static_cast<int64_t>(hits.size()), &total),
        "scan_block");
  const int64_t first = std::min<int64_t>(total, static_cast<int64_t>(hits.size()));
  hits.resize(static_cast<size_t>(total));
  if (total > first) {
    int64_t got = 0;
    check(eqtl_gpu_scan_hits(ctx_, first, hits.data() + first, total - first, &got), "scan_hits");
    if (got != total - first) die("scan_hits returned fewer hits than the block holds");
  }
}

// set_sample_order (plink_bed.cpp:245-255, vcf_session.cpp:127-138) on the
// engine, with the source's row format.
static void set_sample_order(Engine& eng, const GenoSource& geno, const std::vector<std::string>& ids) {
  std::unordered_map<std::string, int> m;
  for (size_t i = 0; i < geno.samples().size(); ++i) m[geno.samples()[i]] = static_cast<int>(i);
  std::vector<int32_t> col;
  for (const auto& id : ids) {
    auto it = m.find(id);
    if (it == m.end()) die("sample not in genotype file: " + id);
    col.push_back(it->second);
  }
  eng.check(eqtl_gpu_set_samples(eng.ctx(), static_cast<int32_t>(geno.n_file()), col.data(),
                                 static_cast<int32_t>(col.size())), "set_samples");
  eng.check(eqtl_gpu_set_geno_format(eng.ctx(), geno.row_format()), "set_geno_format");
}

static constexpr size_t kBlockSnps = 16384;

Grm compute_grm(Engine& eng, GenoSource& bed, const std::vector<std::string>& sample_ids, MissHand miss,
                double max_miss, double maf) {
  set_sample_order(eng, bed, sample_ids);
  eng.check(eqtl_gpu_set_policy(eng.ctx(), miss == MissHand::Impute, max_miss, maf), "set_policy");
  eng.check(eqtl_gpu_grm_begin(eng.ctx()), "grm_begin");
  for (size_t j0 = 0; j0 < bed.sites().size(); j0 += kBlockSnps) {
    const size_t m = std::min(kBlockSnps, bed.sites().size() - j0);
    int32_t on_dev = 0;
    const uint8_t* rows = bed.rows(eng, j0, m, &on_dev);
    eng.check(eqtl_gpu_grm_accumulate(eng.ctx(), rows, on_dev, static_cast<int64_t>(m), bed.bytes_per_snp()),
              "grm_accumulate");
  }
  Grm g;
  g.ids = sample_ids;
  const int n = static_cast<int>(sample_ids.size());
  g.K.resize(static_cast<size_t>(n) * n);
  g.N_packed.resize(static_cast<size_t>(n) * (n + 1) / 2);
  int64_t used = 0;
  const int rc = eqtl_gpu_grm_finish(eng.ctx(), g.K.data(), g.N_packed.data(), &used);
  if (rc == EQTL_E_ARG) die("no SNPs for GRM");
  eng.check(rc, "grm_finish");
  info("computed GRM from " + std::to_string(used) + " SNPs");
  return g;
}

// Gene-side preparation for the model (prep_lm / prep_lmm / prep_glm_nb for
// every job). glm: phi and converged per gene for the output columns.
static void prep_genes(const Options& opt, Engine& eng, Model model, const std::vector<double>& Y, int G,
                       const CovData& cov, std::vector<double>* phi, std::vector<int32_t>* conv) {
  eqtl_gpu_ctx* c = eng.ctx();
  if (model == Model::Lmm) {
    eng.check(eqtl_gpu_lmm_prep(c, Y.data(), G, cov.X.data(), cov.p, opt.delta_mode, nullptr, nullptr, nullptr),
              "lmm_prep");
  } else if (model == Model::Glm) {
    phi->assign(G, 0.0);
    conv->assign(G, 0);
    eng.check(eqtl_gpu_glm_prep(c, Y.data(), G, cov.X.data(), cov.p, phi->data(), conv->data()), "glm_prep");
  } else {
    eng.check(eqtl_gpu_lm_prep(c, Y.data(), G, cov.X.data(), cov.p, nullptr), "lm_prep");
  }
}

static const char* model_tag(Model m) {
  return m == Model::Lmm ? "LMM" : (m == Model::Glm ? "GLM" : (m == Model::Glmm ? "GLMM" : "LM"));
}

static int32_t engine_model(Model m) {
  return m == Model::Lmm ? EQTL_MODEL_LMM : (m == Model::Glm ? EQTL_MODEL_GLM : EQTL_MODEL_LM);
}

static uint32_t fnv1a(const std::string& s);

// Trans stage 2 (stage2_perm_topk, scan_perm.cpp:11-142): genes whose best p
// is below --perm-trans-thr are permuted over their top-K SNPs (the K smallest
// p below the threshold, earliest SNP first on ties). The engine's cis
// permutation stage runs them on a virtual genotype set: gene k's top-K rows
// in a block of their own, annotated as "chromosome" k with positions 1..K, so
// that the block is gene k's window.
static void stage2_perm(const Options& opt, Engine& eng, GenoSource& bed, Model model, std::vector<TransJob>& jobs,
                        std::vector<std::vector<std::pair<double, int64_t>>>& cand,
                        const std::vector<eqtl_gene_state>& st, const std::vector<double>& Y, const CovData& cov) {
  const int G = static_cast<int>(jobs.size());
  const size_t n = Y.size() / std::max(G, 1);
  std::vector<int> q;  // qualifying genes
  for (int k = 0; k < G; ++k) {
    jobs[k].summary.p_emp = std::numeric_limits<double>::quiet_NaN();
    if (st[k].n_tested > 0 && st[k].best_p < opt.perm_trans_thr && !cand[k].empty()) q.push_back(k);
  }
  if (q.empty()) return;
  const int bps = bed.bytes_per_snp();
  std::vector<int64_t> vidx;
  std::vector<int32_t> vchrom, gch;
  std::vector<int64_t> vpos, gcs, gce;
  std::vector<double> Yq(n * q.size());
  std::vector<uint32_t> hash;
  std::vector<double> best;
  for (size_t qi = 0; qi < q.size(); ++qi) {
    const int k = q[qi];
    auto& cs = cand[k];
    std::sort(cs.begin(), cs.end());
    if (static_cast<int>(cs.size()) > opt.perm_trans_top) cs.resize(opt.perm_trans_top);
    std::sort(cs.begin(), cs.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
    for (size_t i = 0; i < cs.size(); ++i) {
      vidx.push_back(cs[i].second);
      vchrom.push_back(static_cast<int32_t>(qi));
      vpos.push_back(static_cast<int64_t>(i) + 1);
    }
    gch.push_back(static_cast<int32_t>(qi));
    gcs.push_back(1);
    gce.push_back(static_cast<int64_t>(cs.size()));
    std::copy(Y.begin() + static_cast<size_t>(k) * n, Y.begin() + static_cast<size_t>(k + 1) * n,
              Yq.begin() + qi * n);
    hash.push_back(fnv1a(jobs[k].gene));
    best.push_back(st[k].best_p);
  }
  const int Gq = static_cast<int>(q.size());
  eqtl_gpu_ctx* c = eng.ctx();
  const int64_t mv = static_cast<int64_t>(vpos.size());
  eng.check(eqtl_gpu_set_snp_annot(c, vchrom.data(), vpos.data(), mv), "set_snp_annot");
  int32_t on_dev = 0;
  const uint8_t* vrows = bed.gather(eng, vidx, &on_dev);
  eng.check(eqtl_gpu_upload_bed(c, vrows, on_dev, mv, bps), "upload_bed");
  if (model == Model::Lmm)
    eng.check(eqtl_gpu_lmm_prep(c, Yq.data(), Gq, cov.X.data(), cov.p, opt.delta_mode, nullptr, nullptr, nullptr),
              "lmm_prep");
  else
    eng.check(eqtl_gpu_lm_prep(c, Yq.data(), Gq, cov.X.data(), cov.p, nullptr), "lm_prep");
  eng.check(eqtl_gpu_set_gene_windows(c, gch.data(), gcs.data(), gce.data(), Gq), "set_gene_windows");
  info(std::string("trans stage 2: ") + std::to_string(Gq) + " genes x " + std::to_string(opt.perm) +
       " permutations over their top-K SNPs");
  std::vector<double> p_emp(Gq), perm_min_p(static_cast<size_t>(Gq) * opt.perm);
  eng.check(eqtl_gpu_cis_perm(c, opt.precision, opt.perm, opt.seed, hash.data(), best.data(), opt.delta_mode,
                              perm_min_p.data(), p_emp.data()),
            "cis_perm");
  for (int qi = 0; qi < Gq; ++qi) jobs[q[qi]].summary.p_emp = p_emp[qi];
  if (!opt.disable_beta_approx) {  // scan_perm.cpp:136-138
    std::vector<double> pb(Gq), s1(Gq), s2(Gq);
    eng.check(eqtl_gpu_beta_approx(c, perm_min_p.data(), best.data(), Gq, opt.perm, pb.data(), s1.data(), s2.data()),
              "beta_approx");
    for (int qi = 0; qi < Gq; ++qi) {
      GeneSummary& s = jobs[q[qi]].summary;
      s.p_beta = pb[qi];
      s.beta_shape1 = s1[qi];
      s.beta_shape2 = s2[qi];
    }
  }
}

void scan_snp_outer(const Options& opt, Engine& eng, GenoSource& bed, Model model, const std::string& scope,
                    ScopeOut& out, double pthr, std::vector<TransJob>& jobs, const PhenoData& ph,
                    const CovData& cov, const std::vector<double>* Q, const std::vector<double>* lambda) {
  if (jobs.empty()) return;
  const int n = static_cast<int>(ph.n());
  const int G = static_cast<int>(jobs.size());
  std::vector<double> Y(static_cast<size_t>(n) * G);
  for (int k = 0; k < G; ++k)
    std::copy(ph.Y.begin() + static_cast<size_t>(jobs[k].col) * n, ph.Y.begin() + static_cast<size_t>(jobs[k].col + 1) * n,
              Y.begin() + static_cast<size_t>(k) * n);
  // gene windows for the trans exclusion (in_cis_window, scan_common.hpp:65-69)
  std::vector<std::string> chroms = bed.chromosomes();
  auto chrom_id = [&](const std::string& c) -> int32_t {
    for (size_t i = 0; i < chroms.size(); ++i)
      if (chrom_equal(chroms[i], c)) return static_cast<int32_t>(i);
    return -2;  // chromosome not in the bed: never matches
  };
  std::vector<int32_t> snp_chrom(bed.sites().size());
  std::vector<int64_t> snp_pos(bed.sites().size());
  for (size_t j = 0, ci = 0; j < bed.sites().size(); ++j) {
    while (!chrom_equal(bed.sites()[j].chrom, chroms[ci])) ++ci;
    snp_chrom[j] = static_cast<int32_t>(ci);
    snp_pos[j] = bed.sites()[j].pos;
  }
  std::vector<int32_t> gch(G, -1);
  std::vector<int64_t> gcs(G, 0), gce(G, 0);
  for (int k = 0; k < G; ++k) {
    if (scope == "trans" && jobs[k].has_loc) {
      gch[k] = chrom_id(jobs[k].loc.chrom);
      gcs[k] = std::max<int64_t>(1, jobs[k].loc.start - opt.window);
      gce[k] = jobs[k].loc.end + opt.window;
    }
  }
  eqtl_gpu_ctx* c = eng.ctx();
  eng.check(eqtl_gpu_set_snp_annot(c, snp_chrom.data(), snp_pos.data(), static_cast<int64_t>(snp_pos.size())),
            "set_snp_annot");
  (void)Q;
  (void)lambda;
  std::vector<double> phi;
  std::vector<int32_t> conv;
  prep_genes(opt, eng, model, Y, G, cov, &phi, &conv);
  eng.check(eqtl_gpu_set_gene_windows(c, gch.data(), gcs.data(), gce.data(), G), "set_gene_windows");
  // glm runs gene by gene in the reference (scan.cpp:228-300): its pairs are
  // written gene-major, SNP order within a gene
  const bool gene_major = model == Model::Glm;
  std::vector<eqtl_hit> gm_hits;
  // with --perm the scan also keeps every pair below perm_trans_thr: the
  // candidates of each gene's stage-2 top-K set (topk_consider, scan_trans.hpp:53-70)
  const bool stage2 = opt.perm > 0;
  const double pthr_dev = stage2 ? std::max(pthr, opt.perm_trans_thr) : pthr;
  std::vector<std::vector<std::pair<double, int64_t>>> cand(stage2 ? G : 0);
  std::vector<int> n_sig(G, 0);
  eng.check(eqtl_gpu_scan_begin(c, engine_model(model), opt.precision, EQTL_SCOPE_TRANS, pthr_dev), "scan_begin");
  std::vector<eqtl_hit> hits;
  const auto& sites = bed.sites();
  auto fill = [&](AssocHit& h, const eqtl_hit& e) {
    const TransJob& job = jobs[e.gene];
    const BimSite& s = sites[e.snp];
    h.gene = job.gene;
    h.snp = s.id;
    h.chrom = s.chrom;
    h.pos = s.pos;
    h.ref = s.a2;
    h.alt = s.a1;
    h.maf = e.af;
    h.beta = e.beta;
    h.se = e.se;
    h.stat = e.stat;
    h.p = e.p;
    h.r2 = e.r2;
    h.n = n;
    if (model == Model::Glm) {
      h.phi = phi[e.gene];
      h.glm_converged = conv[e.gene] != 0;
    }
    if (job.has_loc && job.loc.ok) {
      h.has_tss_dist = true;
      h.tss_dist = static_cast<double>(s.pos - job.loc.tss);
    }
  };
  size_t snp_cnt = 0;
  for (size_t j0 = 0; j0 < sites.size(); j0 += kBlockSnps) {
    const size_t m = std::min(kBlockSnps, sites.size() - j0);
    int32_t on_dev = 0;
    const uint8_t* rows = bed.rows(eng, j0, m, &on_dev);
    eng.scan_block(rows, on_dev, static_cast<int64_t>(j0), static_cast<int64_t>(m), bed.bytes_per_snp(), hits);
    const int64_t nh = static_cast<int64_t>(hits.size());
    for (int64_t i = 0; i < nh; ++i) {  // SNP-major, gene order (scan_trans_lmm.cpp:185-188)
      const eqtl_hit& e = hits[i];
      if (stage2 && e.p < opt.perm_trans_thr) cand[e.gene].emplace_back(e.p, e.snp);
      if (e.p > pthr) continue;
      ++n_sig[e.gene];
      if (gene_major) {
        gm_hits.push_back(e);
        continue;
      }
      AssocHit h;
      fill(h, e);
      emit_line(out, false, e.snp, jobs[e.gene].order, h, model, out.tag);
    }
    const size_t before = snp_cnt;
    snp_cnt += m;
    if (before / 100000 != snp_cnt / 100000)
      info(std::string("trans/gw ") + model_tag(model) + ": SNP " + std::to_string(snp_cnt));
  }
  if (gene_major) {
    std::stable_sort(gm_hits.begin(), gm_hits.end(), [](const eqtl_hit& a, const eqtl_hit& b) {
      return a.gene != b.gene ? a.gene < b.gene : a.snp < b.snp;
    });
    for (const auto& e : gm_hits) {
      AssocHit h;
      fill(h, e);
      emit_line(out, false, jobs[e.gene].order, e.snp, h, model, out.tag);
    }
  }
  std::vector<eqtl_gene_state> st(G);
  eng.check(eqtl_gpu_scan_end(c, st.data()), "scan_end");
  for (int k = 0; k < G; ++k) {
    auto& job = jobs[k];
    job.summary.gene = job.gene;
    if (job.has_loc && job.loc.ok) {
      job.summary.chrom = job.loc.chrom;
      job.summary.tss = job.loc.tss;
    }
    job.summary.n_tested = st[k].n_tested;
    job.summary.n_sig = n_sig[k];
    job.summary.acat_p = st[k].acat_p;
    if (st[k].best_snp >= 0 && st[k].best_p <= pthr && st[k].best_p <= 1.0) {
      eqtl_hit e;
      e.snp = st[k].best_snp;
      e.gene = k;
      e.af = st[k].best_af;
      e.beta = st[k].best_beta;
      e.se = st[k].best_se;
      e.stat = st[k].best_stat;
      e.p = st[k].best_p;
      e.r2 = st[k].best_r2;
      AssocHit h;
      fill(h, e);
      emit_line(out, true, job.order, 0, h, model, scope);
    }
  }
  if (stage2) stage2_perm(opt, eng, bed, model, jobs, cand, st, Y, cov);
}

// FNV-1a of a gene id, low 32 bits (util.hpp:51-58): permutation seeds.
static uint32_t fnv1a(const std::string& s) {
  unsigned long long h = 14695981039346656037ULL;
  for (unsigned char ch : s) {
    h ^= ch;
    h *= 1099511628211ULL;
  }
  return static_cast<uint32_t>(h & 0xFFFFFFFFu);
}

// Cis scan of every gene's window with optional gene-level permutations
// (run_cis_parallel, scan_cis.cpp:18-163, and scan_gene_snps,
// scan_cis.hpp:15-185): genes in (chromosome, TSS) order, pairs written
// gene by gene in SNP order, top only when best p <= pthr, p_emp from the
// engine's permutation stage.
void scan_cis(const Options& opt, Engine& eng, GenoSource& bed, Model model, ScopeOut& out, double pthr,
              std::vector<TransJob>& jobs, const PhenoData& ph, const CovData& cov) {
  if (jobs.empty()) return;
  const int n = static_cast<int>(ph.n());
  const int G = static_cast<int>(jobs.size());
  std::vector<double> Y(static_cast<size_t>(n) * G);
  for (int k = 0; k < G; ++k)
    std::copy(ph.Y.begin() + static_cast<size_t>(jobs[k].col) * n,
              ph.Y.begin() + static_cast<size_t>(jobs[k].col + 1) * n, Y.begin() + static_cast<size_t>(k) * n);
  const auto& sites = bed.sites();
  std::vector<std::string> chroms = bed.chromosomes();
  auto chrom_id = [&](const std::string& c) -> int32_t {
    for (size_t i = 0; i < chroms.size(); ++i)
      if (chrom_equal(chroms[i], c)) return static_cast<int32_t>(i);
    return -1;  // chromosome not in the bed: empty window
  };
  std::vector<int32_t> snp_chrom(sites.size());
  std::vector<int64_t> snp_pos(sites.size());
  for (size_t j = 0, ci = 0; j < sites.size(); ++j) {
    while (!chrom_equal(sites[j].chrom, chroms[ci])) ++ci;
    snp_chrom[j] = static_cast<int32_t>(ci);
    snp_pos[j] = sites[j].pos;
  }
  // windows (cis_window_bounds, scan_common.hpp:58-63) and their SNP index ranges
  // chromosomes are contiguous runs of the bim with sorted positions (checked when the bim is read)
  std::vector<size_t> cbeg(chroms.size(), 0), cend(chroms.size(), 0);
  for (size_t j = 0; j < sites.size(); ++j) {
    if (j == 0 || snp_chrom[j] != snp_chrom[j - 1]) cbeg[snp_chrom[j]] = j;
    cend[snp_chrom[j]] = j + 1;
  }
  std::vector<int32_t> gch(G);
  std::vector<int64_t> gcs(G), gce(G);
  int64_t lo = INT64_MAX, hi = -1;
  for (int k = 0; k < G; ++k) {
    gch[k] = chrom_id(jobs[k].loc.chrom);
    gcs[k] = std::max<int64_t>(1, jobs[k].loc.start - opt.window);
    gce[k] = jobs[k].loc.end + opt.window;
    if (gch[k] < 0) continue;
    auto b = snp_pos.begin() + cbeg[gch[k]], e = snp_pos.begin() + cend[gch[k]];
    const auto a0 = std::lower_bound(b, e, gcs[k]), a1 = std::upper_bound(b, e, gce[k]);
    if (a0 < a1) {
      lo = std::min<int64_t>(lo, a0 - snp_pos.begin());
      hi = std::max<int64_t>(hi, a1 - snp_pos.begin());
    }
  }
  eqtl_gpu_ctx* c = eng.ctx();
  eng.check(eqtl_gpu_set_snp_annot(c, snp_chrom.data(), snp_pos.data(), static_cast<int64_t>(snp_pos.size())),
            "set_snp_annot");
  if (opt.perm > 0) {
    int32_t on_dev = 0;
    const uint8_t* all = bed.rows(eng, 0, sites.size(), &on_dev);
    eng.check(eqtl_gpu_upload_bed(c, all, on_dev, static_cast<int64_t>(sites.size()), bed.bytes_per_snp()),
              "upload_bed");
  }
  const int gm = engine_model(model);
  std::vector<double> phi;
  std::vector<int32_t> conv;
  prep_genes(opt, eng, model, Y, G, cov, &phi, &conv);
  eng.check(eqtl_gpu_set_gene_windows(c, gch.data(), gcs.data(), gce.data(), G), "set_gene_windows");
  eng.check(eqtl_gpu_scan_begin(c, gm, opt.precision, EQTL_SCOPE_CIS, pthr), "scan_begin");
  std::vector<eqtl_hit> hits, all_hits;
  for (int64_t j0 = std::max<int64_t>(lo, 0); j0 < hi; j0 += static_cast<int64_t>(kBlockSnps)) {
    const size_t m = std::min<size_t>(kBlockSnps, static_cast<size_t>(hi - j0));
    int32_t on_dev = 0;
    const uint8_t* rows = bed.rows(eng, static_cast<size_t>(j0), m, &on_dev);
    eng.scan_block(rows, on_dev, j0, static_cast<int64_t>(m), bed.bytes_per_snp(), hits);
    all_hits.insert(all_hits.end(), hits.begin(), hits.end());
  }
  std::vector<eqtl_gene_state> st(G);
  eng.check(eqtl_gpu_scan_end(c, st.data()), "scan_end");
  auto fill = [&](AssocHit& h, const eqtl_hit& e) {
    const TransJob& job = jobs[e.gene];
    const BimSite& s = sites[e.snp];
    h.gene = job.gene;
    h.snp = s.id;
    h.chrom = s.chrom;
    h.pos = s.pos;
    h.ref = s.a2;
    h.alt = s.a1;
    h.maf = e.af;
    h.beta = e.beta;
    h.se = e.se;
    h.stat = e.stat;
    h.p = e.p;
    h.r2 = e.r2;
    h.n = n;
    if (model == Model::Glm) {
      h.phi = phi[e.gene];
      h.glm_converged = conv[e.gene] != 0;
    }
    if (job.loc.ok) {
      h.has_tss_dist = true;
      h.tss_dist = static_cast<double>(s.pos - job.loc.tss);
    }
  };
  std::stable_sort(all_hits.begin(), all_hits.end(), [](const eqtl_hit& a, const eqtl_hit& b) {
    return a.gene != b.gene ? a.gene < b.gene : a.snp < b.snp;
  });
  for (const auto& e : all_hits) {
    AssocHit h;
    fill(h, e);
    emit_line(out, false, jobs[e.gene].order, e.snp, h, model, "cis");
  }
  std::vector<double> p_emp(G, std::numeric_limits<double>::quiet_NaN());
  std::vector<double> p_beta(G, std::numeric_limits<double>::quiet_NaN()), shape1(G, 0.0), shape2(G, 0.0);
  if (opt.perm > 0) {
    std::vector<uint32_t> hash(G);
    std::vector<double> best(G);
    for (int k = 0; k < G; ++k) {
      hash[k] = fnv1a(jobs[k].gene);
      best[k] = st[k].n_tested > 0 ? st[k].best_p : std::numeric_limits<double>::quiet_NaN();
    }
    info(std::string("cis ") + model_tag(model) + ": " + std::to_string(opt.perm) +
         " gene-level permutations");
    std::vector<double> perm_min_p(static_cast<size_t>(G) * opt.perm);
    eng.check(eqtl_gpu_cis_perm(c, opt.precision, opt.perm, opt.seed, hash.data(), best.data(), opt.delta_mode,
                                perm_min_p.data(), p_emp.data()),
              "cis_perm");
    if (!opt.disable_beta_approx) {  // scan_cis.hpp:172-174; genes with no tested SNP keep NA / 0 / 0
      std::vector<double> pb(G), s1(G), s2(G);
      eng.check(eqtl_gpu_beta_approx(c, perm_min_p.data(), best.data(), G, opt.perm, pb.data(), s1.data(), s2.data()),
                "beta_approx");
      for (int k = 0; k < G; ++k)
        if (st[k].n_tested > 0) {
          p_beta[k] = pb[k];
          shape1[k] = s1[k];
          shape2[k] = s2[k];
        }
    }
  }
  for (int k = 0; k < G; ++k) {
    auto& job = jobs[k];
    job.summary.gene = job.gene;
    job.summary.chrom = job.loc.chrom;
    job.summary.tss = job.loc.tss;
    job.summary.n_tested = st[k].n_tested;
    job.summary.n_sig = st[k].n_sig;
    job.summary.acat_p = st[k].acat_p;
    job.summary.p_emp = p_emp[k];
    job.summary.p_beta = p_beta[k];
    job.summary.beta_shape1 = shape1[k];
    job.summary.beta_shape2 = shape2[k];
    if (st[k].best_snp >= 0 && st[k].best_p <= pthr && st[k].best_p <= 1.0) {
      eqtl_hit e;
      e.snp = st[k].best_snp;
      e.gene = k;
      e.af = st[k].best_af;
      e.beta = st[k].best_beta;
      e.se = st[k].best_se;
      e.stat = st[k].best_stat;
      e.p = st[k].best_p;
      e.r2 = st[k].best_r2;
      AssocHit h;
      fill(h, e);
      emit_line(out, true, job.order, 0, h, model, "cis");
    }
  }
}

// ============================================================ orchestrator (scan.cpp:18-340)
// The genotype source named by -b or -v (scan.cpp:20-45, 321-339).
static std::unique_ptr<GenoSource> open_geno(const Options& opt) {
  if (opt.use_bfile()) {
    auto bed = std::make_unique<PlinkBed>();
    bed->open(opt.bfile);
    return bed;
  }
  auto vcf = std::make_unique<VcfSource>();
  vcf->open(opt.vcf, opt.threads);
  return vcf;
}

int run_make_grm(const Options& opt) {
  std::unique_ptr<GenoSource> geno = open_geno(opt);
  std::vector<std::string> samples = geno->samples();
  if (!opt.pheno.empty()) {
    PhenoData ph = load_pheno(opt.pheno);
    samples = intersect_order(ph.sample_ids, geno->samples());
    if (samples.empty()) die("no overlapping samples for GRM");
  }
  Engine eng(opt.device);
  Grm g = compute_grm(eng, *geno, samples, opt.miss, opt.max_miss, opt.maf);
  write_grm_gcta(opt.out, g);
  return 0;
}

// sparsify_grm (model_lmm.cpp:88-104): off-diagonal |K_ij| < thr set to 0.
static void sparsify_grm(std::vector<double>& K, int n, double thr) {
  size_t n_zero = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j)
      if (std::fabs(K[static_cast<size_t>(j) * n + i]) < thr) {
        K[static_cast<size_t>(j) * n + i] = 0.0;
        K[static_cast<size_t>(i) * n + j] = 0.0;
        n_zero += 2;
      }
  info("fast: GRM sparse approx thr=" + std::to_string(thr) + " zeroed " + std::to_string(n_zero) +
       " off-diagonal entries");
}

// Analysis samples with a finite phenotype and finite covariates
// (build_gene_ready, scan_common.cpp:11-25).
static std::vector<int32_t> finite_keep(const PhenoData& ph, int col, const CovData& cov) {
  const size_t n = ph.n();
  std::vector<int32_t> keep;
  for (size_t i = 0; i < n; ++i) {
    if (!std::isfinite(ph.Y[static_cast<size_t>(col) * n + i])) continue;
    bool ok = true;
    for (int c = 0; c < cov.p && ok; ++c) ok = std::isfinite(cov.X[static_cast<size_t>(c) * n + i]);
    if (ok) keep.push_back(static_cast<int32_t>(i));
  }
  return keep;
}

// The phenotypes of jobs (rows keep) and the covariates on keep.
static void subset_inputs(const PhenoData& ph, const CovData& cov, const std::vector<int32_t>& keep,
                          std::vector<TransJob>& jobs, PhenoData& ph_g, CovData& cov_g) {
  const size_t n = ph.n(), nk = keep.size();
  ph_g.sample_ids.clear();
  for (int32_t i : keep) ph_g.sample_ids.push_back(ph.sample_ids[i]);
  ph_g.gene_ids.clear();
  ph_g.Y.assign(nk * jobs.size(), 0.0);
  for (size_t k = 0; k < jobs.size(); ++k) {
    ph_g.gene_ids.push_back(jobs[k].gene);
    for (size_t r = 0; r < nk; ++r) ph_g.Y[k * nk + r] = ph.Y[static_cast<size_t>(jobs[k].col) * n + keep[r]];
    jobs[k].col = static_cast<int>(k);
  }
  cov_g.sample_ids = ph_g.sample_ids;
  cov_g.cov_names = cov.cov_names;
  cov_g.p = cov.p;
  cov_g.X.assign(nk * cov.p, 0.0);
  for (int c = 0; c < cov.p; ++c)
    for (size_t r = 0; r < nk; ++r) cov_g.X[static_cast<size_t>(c) * nk + r] = cov.X[static_cast<size_t>(c) * n + keep[r]];
}

int run_eqtl(const Options& opt) {
  PhenoData ph = load_pheno(opt.pheno);
  std::unique_ptr<GenoSource> geno = open_geno(opt);
  GenoSource& bed = *geno;
  std::vector<std::string> order = intersect_order(ph.sample_ids, bed.samples());
  if (order.empty()) die(opt.use_bfile() ? "no overlapping samples between pheno and fam"
                                         : "no overlapping samples between pheno and VCF");
  info("overlap samples: " + std::to_string(order.size()));
  slice_pheno(ph, order);
  if (opt.pheno_int)  // apply_pheno_norm per gene (scan_common.hpp:124-126)
    for (size_t k = 0; k < ph.G(); ++k) inverse_normal_transform(&ph.Y[k * ph.n()], ph.n());
  CovData cov = load_covar(opt.covar, order);
  std::unordered_map<std::string, GeneLoc> annot;
  const bool have_gff = !opt.gff.empty();
  if (have_gff) annot = load_gff_tss(opt.gff, opt.gff_id_key);
  else info("no GFF: mode ignored; genome-wide all-pairs (gw)");
  std::vector<std::string> scopes;
  if (!have_gff) scopes = {"gw"};
  else if (opt.mode == Mode::Cis) scopes = {"cis"};
  else if (opt.mode == Mode::Trans) scopes = {"trans"};
  else if (opt.mode == Mode::Gw) scopes = {"gw"};
  else scopes = {"cis", "trans"};
  {
    std::ofstream probe(opt.out + ".probe");
    if (!probe) die("cannot write to output prefix: " + opt.out);
    probe.close();
    std::remove((opt.out + ".probe").c_str());
  }
  if (std::find(scopes.begin(), scopes.end(), "cis") != scopes.end() && !bed.has_index())  // scan.cpp:93-99
    die("cis mode requires an indexed genotype file; for VCF run: tabix -p vcf <file>.vcf.gz");
  if (have_gff) {
    size_t matched = 0;
    for (const auto& g : ph.gene_ids) matched += annot.count(g);
    info("pheno genes matched to GFF: " + std::to_string(matched) + "/" + std::to_string(ph.gene_ids.size()));
    if (matched == 0) die("no phenotype gene IDs match the GFF annotation (check --gff-id-key and ID format)");
  }
  Engine eng(opt.device);
  bool any_lmm = false;
  for (Model m : opt.models) any_lmm |= (m == Model::Lmm);
  Grm grm;
  if (any_lmm) {
    if (!opt.grm.empty()) grm = slice_grm(load_grm_gcta(opt.grm), order);
    else {
      info("computing GRM from genotypes (overlap samples)");
      grm = compute_grm(eng, bed, order, opt.miss, opt.max_miss, opt.maf);
    }
  }
  set_sample_order(eng, bed, order);
  eng.check(eqtl_gpu_set_policy(eng.ctx(), opt.miss == MissHand::Impute, opt.max_miss, opt.maf), "set_policy");
  for (Model model : opt.models)
    if (model == Model::Glm && opt.perm > 0)
      die("--perm: gene-level permutations are provided for lm and lmm (the NB glm's parametric resampling is not)");
  for (Model model : opt.models) {
    for (const std::string& scope : scopes) {
      const std::string prefix = opt.out + "." + model_str(model) + "." + scope;
      ScopeOut so;
      so.tag = scope;
      so.pairs.open(prefix + ".pairs.tsv");
      so.top.open(prefix + ".top.tsv");
      so.region.open(prefix + ".region.tsv");
      if (!so.pairs || !so.top || !so.region) die("cannot open output for " + prefix);
      write_pairs_header(so.pairs, model);
      write_pairs_header(so.top, model);
      write_region_header(so.region);
      const double pthr = (scope == "cis") ? opt.pval_cis : opt.pval_trans;
      std::vector<TransJob> jobs;
      for (size_t gi = 0; gi < ph.gene_ids.size(); ++gi) {
        if (model == Model::Glm && !looks_like_counts(&ph.Y[gi * ph.n()], ph.n())) {  // scan.cpp:246-249
          warn("skip non-count gene for glm: " + ph.gene_ids[gi]);
          continue;
        }
        TransJob job;
        job.gene = ph.gene_ids[gi];
        job.col = static_cast<int>(gi);
        if (have_gff) {
          auto it = annot.find(job.gene);
          if (it != annot.end()) {
            job.loc = it->second;
            job.has_loc = true;
          } else if (scope != "gw") {
            continue;
          }
        }
        if ((scope == "trans" || scope == "cis") && !job.has_loc) continue;
        // build_gene_ready (scan_common.cpp:11-35): >= 3 samples with finite
        // phenotype and covariates, and a non-zero y on them
        job.keep = finite_keep(ph, job.col, cov);
        double ymax = 0;
        for (int32_t i : job.keep) ymax = std::max(ymax, std::fabs(ph.Y[gi * ph.n() + i]));
        if (job.keep.size() < 3 || ymax < 1e-15) continue;
        jobs.push_back(std::move(job));
      }
      // run_cis_parallel (threads > 1) orders cis genes by (chromosome, TSS);
      // the single-thread gene loop keeps the phenotype order (scan.cpp:217-300)
      if (scope == "cis" && opt.threads > 1)
        std::stable_sort(jobs.begin(), jobs.end(), [](const TransJob& a, const TransJob& b) {
          if (a.loc.chrom != b.loc.chrom) return a.loc.chrom < b.loc.chrom;
          return a.loc.tss < b.loc.tss;
        });
      for (size_t k = 0; k < jobs.size(); ++k) jobs[k].order = static_cast<int>(k);
      // genes grouped by sample subset, in order of first appearance; one
      // group (the usual case: no missing values) streams its output
      std::vector<std::vector<int>> groups;
      {
        std::map<std::vector<int32_t>, int> gid;
        for (size_t k = 0; k < jobs.size(); ++k) {
          auto it = gid.find(jobs[k].keep);
          if (it == gid.end()) {
            it = gid.emplace(jobs[k].keep, static_cast<int>(groups.size())).first;
            groups.emplace_back();
          }
          groups[it->second].push_back(static_cast<int>(k));
        }
      }
      so.buffered = groups.size() > 1;
      if (groups.size() > 1)
        info(std::to_string(groups.size()) + " distinct sample subsets (missing phenotype/covariate values)");
      for (const auto& grp : groups) {
        std::vector<TransJob> gjobs;
        for (int k : grp) gjobs.push_back(jobs[k]);
        const std::vector<int32_t> keep = gjobs[0].keep;
        const bool full = keep.size() == ph.n();
        eng.check(eqtl_gpu_set_test_subset(eng.ctx(), full ? nullptr : keep.data(), static_cast<int32_t>(keep.size())),
                  "set_test_subset");
        PhenoData ph_g;
        CovData cov_g;
        subset_inputs(ph, cov, keep, gjobs, ph_g, cov_g);
        if (model == Model::Lmm) {
          // the GRM on the subset (slice of K); --fast sparsifies it where the
          // reference does: a basis shared by all genes, and per-gene cis bases
          const int nk = static_cast<int>(keep.size());
          std::vector<double> Kg(static_cast<size_t>(nk) * nk);
          for (int b = 0; b < nk; ++b)
            for (int a = 0; a < nk; ++a)
              Kg[static_cast<size_t>(b) * nk + a] = grm.K[static_cast<size_t>(keep[b]) * ph.n() + keep[a]];
          if (opt.fast && (groups.size() == 1 || scope == "cis")) sparsify_grm(Kg, nk, 1e-4);
          const int rc = eqtl_gpu_eigh(eng.ctx(), Kg.data(), nk, nullptr, nullptr);
          if (rc == EQTL_E_SOLVER) die("GRM eigen decomposition failed");
          eng.check(rc, "eigh");
        }
        if (scope == "cis") {
          info(std::string("cis ") + model_tag(model) + ": " + std::to_string(gjobs.size()) + " genes");
          scan_cis(opt, eng, bed, model, so, pthr, gjobs, ph_g, cov_g);
        } else {
          info(std::string("trans/gw ") + model_tag(model) + ": SNP-outer (" + std::to_string(gjobs.size()) +
               " genes)");
          scan_snp_outer(opt, eng, bed, model, scope, so, pthr, gjobs, ph_g, cov_g, nullptr, nullptr);
        }
        for (size_t i = 0; i < grp.size(); ++i) jobs[grp[i]].summary = gjobs[i].summary;
      }
      eng.check(eqtl_gpu_set_test_subset(eng.ctx(), nullptr, 0), "set_test_subset");
      if (so.buffered) flush_lines(so);
      std::vector<GeneSummary> summaries;
      for (auto& j : jobs) summaries.push_back(j.summary);
      bh_adjust(summaries);
      for (const auto& s : summaries) write_region_line(so.region, s);
      info("finished " + prefix);
    }
  }
  return 0;
}

}  // namespace eqtl
