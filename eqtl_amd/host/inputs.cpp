// Command line and the two sample-keyed input tables (phenotypes,
// covariates), written from the format specification below. The flag names,
// the table formats and the user-visible messages are the reference's
// interface (src/options.cpp, src/pheno.cpp) and are kept; the code is this
// engine's own.
//
// Command line: every flag has a long form (--name VALUE or --name=VALUE) and
// some a one-letter form (-x VALUE or -xVALUE). A flag either takes exactly one
// value or none. An unknown flag, or a value-taking flag at the end of the
// line, is reported and the parse fails (status 1); -h / --help asks for the
// usage text (status 2); no arguments at all also gives the usage text.
//
// Tables (both): text lines, blank lines and lines whose first non-blank
// character is '#' skipped (except the phenotype header, below); fields split
// on tabs when the line has a tab, else on runs of whitespace; fields trimmed.
//  * Phenotypes: the first line is the header (a leading '#' is dropped):
//    column 0 names the sample column, columns 1.. are gene ids (unique).
//    Each further line: sample id, then one value per gene; "NA", "NaN" and
//    "." are missing. A line with a different field count is an error. A
//    sample id seen again with the same values (missing equal to missing) is
//    dropped with a warning; with different values it is an error.
//  * Covariates: lines with fewer than two fields are skipped. The first
//    remaining line is a header when its second field is not a number (then
//    fields 1.. name the covariates; else they are cov1, cov2, ...). Each
//    data line: sample id and one number per covariate. Repeated ids: as for
//    phenotypes (exact equality). Every analysis sample must be present. If
//    no column equals 1 (within 1e-12) for all analysis samples, an intercept
//    column is put first. Without a covariate file the design is the
//    intercept alone.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <unordered_map>

#include "eqtl_host.hpp"

namespace eqtl {

// ============================================================ text tables
namespace {

std::vector<std::string> fields_of(const std::string& line) {
  std::vector<std::string> f = line.find('\t') != std::string::npos ? split_char(line, '\t') : split_ws(line);
  for (auto& x : f) x = trim(x);
  return f;
}

bool parses_as_number(const std::string& s) {
  if (s.empty()) return false;
  const char* b = s.c_str();
  char* e = nullptr;
  (void)std::strtod(b, &e);
  return e != b && *e == '\0';
}

double value_or_missing(const std::string& s) {
  if (s == "NA" || s == "NaN" || s == ".") return std::numeric_limits<double>::quiet_NaN();
  return std::stod(s);
}

// Rows keyed by sample id, in first-seen order; a repeated id must carry the
// same values (same(a, b)) and is then dropped with a warning.
struct KeyedRows {
  std::vector<std::string> ids;
  std::vector<std::vector<double>> vals;
  std::unordered_map<std::string, size_t> where;

  void add(const std::string& id, std::vector<double> v, const std::string& what,
           const std::function<bool(double, double)>& same) {
    auto it = where.find(id);
    if (it == where.end()) {
      where.emplace(id, ids.size());
      ids.push_back(id);
      vals.push_back(std::move(v));
      return;
    }
    const auto& prev = vals[it->second];
    bool equal = prev.size() == v.size();
    for (size_t j = 0; equal && j < v.size(); ++j) equal = same(prev[j], v[j]);
    if (!equal) die(what + ": duplicate id with conflicting values: " + id);
    warn(what + ": duplicate id row with identical values (kept first): " + id);
  }
};

}  // namespace

PhenoData load_pheno(const std::string& path) {
  std::ifstream in(path);
  if (!in) die("cannot open pheno: " + path);
  std::string line;
  if (!std::getline(in, line)) die("empty pheno: " + path);
  std::string head = trim(line);
  if (!head.empty() && head.front() == '#') head.erase(0, 1);
  const std::vector<std::string> cols = fields_of(head);
  if (cols.size() < 2) die("pheno header needs sample + >=1 gene");
  PhenoData P;
  P.gene_ids.assign(cols.begin() + 1, cols.end());
  {
    std::unordered_map<std::string, int> seen;
    for (const auto& g : P.gene_ids)
      if (seen[g]++) die("pheno genes: duplicate id: " + g);
  }
  KeyedRows rows;
  const auto nan_equal = [](double a, double b) { return (std::isnan(a) && std::isnan(b)) || a == b; };
  while (std::getline(in, line)) {
    const std::string t = trim(line);
    if (t.empty() || t.front() == '#') continue;
    const std::vector<std::string> f = fields_of(t);
    if (f.size() != cols.size()) die("pheno row column mismatch for sample " + (f.empty() ? "?" : f[0]));
    std::vector<double> v;
    v.reserve(f.size() - 1);
    for (size_t j = 1; j < f.size(); ++j) v.push_back(value_or_missing(f[j]));
    rows.add(f[0], std::move(v), "pheno samples", nan_equal);
  }
  if (rows.ids.empty()) die("no samples in pheno");
  const size_t n = rows.ids.size(), G = P.gene_ids.size();
  P.sample_ids = rows.ids;
  P.Y.assign(n * G, 0.0);
  for (size_t g = 0; g < G; ++g)
    for (size_t i = 0; i < n; ++i) P.Y[g * n + i] = rows.vals[i][g];
  info("pheno: " + std::to_string(n) + " samples, " + std::to_string(G) + " genes");
  return P;
}

void slice_pheno(PhenoData& p, const std::vector<std::string>& sample_ids) {
  std::unordered_map<std::string, size_t> at;
  for (size_t i = 0; i < p.sample_ids.size(); ++i) at.emplace(p.sample_ids[i], i);
  const size_t n_old = p.sample_ids.size(), n_new = sample_ids.size(), G = p.G();
  std::vector<size_t> src(n_new);
  for (size_t i = 0; i < n_new; ++i) {
    auto it = at.find(sample_ids[i]);
    if (it == at.end()) die("internal: sample missing in pheno slice");
    src[i] = it->second;
  }
  std::vector<double> Y(n_new * G);
  for (size_t g = 0; g < G; ++g) {
    const double* col = p.Y.data() + g * n_old;
    for (size_t i = 0; i < n_new; ++i) Y[g * n_new + i] = col[src[i]];
  }
  p.sample_ids = sample_ids;
  p.Y = std::move(Y);
}

CovData load_covar(const std::string& path, const std::vector<std::string>& sample_ids) {
  const size_t n = sample_ids.size();
  CovData C;
  C.sample_ids = sample_ids;
  if (path.empty()) {
    C.cov_names = {"intercept"};
    C.p = 1;
    C.X.assign(n, 1.0);
    return C;
  }
  std::ifstream in(path);
  if (!in) die("cannot open covar: " + path);
  std::string line;
  std::vector<std::string> names;
  KeyedRows rows;
  bool header_possible = true;
  while (std::getline(in, line)) {
    const std::string t = trim(line);
    if (t.empty() || t.front() == '#') continue;
    const std::vector<std::string> f = fields_of(t);
    if (f.size() < 2) continue;
    if (header_possible) {
      header_possible = false;
      if (!parses_as_number(f[1])) {
        names.assign(f.begin() + 1, f.end());
        continue;
      }
    }
    std::vector<double> v;
    v.reserve(f.size() - 1);
    for (size_t j = 1; j < f.size(); ++j) v.push_back(std::stod(f[j]));
    rows.add(f[0], std::move(v), "covar samples", [](double a, double b) { return a == b; });
  }
  if (rows.ids.empty()) die("no covariates in " + path);
  const size_t nc = rows.vals.front().size();
  for (size_t j = names.size(); j < nc; ++j) names.push_back("cov" + std::to_string(j + 1));
  names.resize(nc);
  // analysis samples x covariates, column-major
  std::vector<double> raw(n * nc);
  for (size_t i = 0; i < n; ++i) {
    auto it = rows.where.find(sample_ids[i]);
    if (it == rows.where.end()) die("sample " + sample_ids[i] + " missing in covar");
    const auto& v = rows.vals[it->second];
    if (v.size() != nc) die("covar row width mismatch");
    for (size_t j = 0; j < nc; ++j) raw[j * n + i] = v[j];
  }
  bool has_intercept = false;
  for (size_t j = 0; j < nc && !has_intercept; ++j) {
    const double* col = raw.data() + j * n;
    has_intercept = std::all_of(col, col + n, [](double x) { return std::fabs(x - 1.0) <= 1e-12; });
  }
  if (has_intercept) {
    C.p = static_cast<int>(nc);
    C.X = std::move(raw);
    C.cov_names = names;
  } else {
    C.p = static_cast<int>(nc) + 1;
    C.X.assign(n, 1.0);
    C.X.insert(C.X.end(), raw.begin(), raw.end());
    C.cov_names.push_back("intercept");
    C.cov_names.insert(C.cov_names.end(), names.begin(), names.end());
  }
  info("covar: " + std::to_string(C.p) + " columns (incl intercept if added)");
  return C;
}

// ============================================================ command line
namespace {

std::vector<Model> models_from(const std::string& list) {
  std::vector<Model> out;
  for (const auto& piece : split_char(list, ',')) {
    const std::string t = trim(piece);
    if (t.empty()) continue;
    static const std::pair<const char*, Model> known[] = {{"lm", Model::Lm}, {"lmm", Model::Lmm}, {"glm", Model::Glm}};
    bool found = false;
    for (const auto& k : known)
      if (t == k.first) {
        out.push_back(k.second);
        found = true;
      }
    if (found) continue;
    if (t == "glmm") die("--model glmm is not provided by this engine (lm, glm, lmm)");
    die("invalid --model value: " + t + " (lm|glm|lmm|glmm)");
  }
  if (out.empty()) die("empty --model");
  return out;
}

Mode mode_from(const std::string& s) {
  static const std::pair<const char*, Mode> known[] = {
      {"cis", Mode::Cis}, {"trans", Mode::Trans}, {"all", Mode::All}, {"gw", Mode::Gw}, {"gwas", Mode::Gw}};
  for (const auto& k : known)
    if (s == k.first) return k.second;
  die("invalid --mode: " + s + " (cis|trans|all|gw)");
}

// One of a fixed set of words -> its value, or die(message).
template <class T>
T choice(const std::string& s, std::initializer_list<std::pair<const char*, T>> words, const std::string& message) {
  for (const auto& w : words)
    if (s == w.first) return w.second;
  die(message);
}

double number_in(const std::string& s, double lo, double hi, bool lo_open, const std::string& message) {
  const double v = std::atof(s.c_str());
  if ((lo_open ? !(v > lo) : !(v >= lo)) || !(v <= hi)) die(message);
  return v;
}

struct Flag {
  const char* name;  // long form, without the dashes
  char letter;       // one-letter form, 0 if none
  bool value;        // takes a value
  std::function<void(Options&, const std::string&)> apply;
};

const std::vector<Flag>& flag_table() {
  static const std::vector<Flag> table = {
      {"vcf", 'v', true, [](Options& o, const std::string& v) { o.vcf = v; }},
      {"bfile", 'b', true, [](Options& o, const std::string& v) { o.bfile = v; }},
      {"pheno", 'e', true, [](Options& o, const std::string& v) { o.pheno = v; }},
      {"out", 'o', true, [](Options& o, const std::string& v) { o.out = v; }},
      {"gff", 'g', true, [](Options& o, const std::string& v) { o.gff = v; }},
      {"covar", 'c', true, [](Options& o, const std::string& v) { o.covar = v; }},
      {"grm", 'k', true, [](Options& o, const std::string& v) { o.grm = v; }},
      {"gff-id-key", 0, true, [](Options& o, const std::string& v) { o.gff_id_key = v; }},
      {"make-grm", 0, false, [](Options& o, const std::string&) { o.make_grm = true; }},
      {"mode", 'm', true, [](Options& o, const std::string& v) { o.mode = mode_from(v); }},
      {"model", 0, true, [](Options& o, const std::string& v) { o.models = models_from(v); }},
      {"window", 'w', true, [](Options& o, const std::string& v) { o.window = std::atoi(v.c_str()); }},
      {"pval-cis", 0, true, [](Options& o, const std::string& v) { o.pval_cis = std::atof(v.c_str()); }},
      {"pval-trans", 0, true, [](Options& o, const std::string& v) { o.pval_trans = std::atof(v.c_str()); }},
      {"miss-hand", 0, true,
       [](Options& o, const std::string& v) {
         o.miss = choice<MissHand>(v, {{"filter", MissHand::Filter}, {"impute", MissHand::Impute}},
                                   "invalid --miss-hand (filter|impute)");
       }},
      {"max-miss", 0, true,
       [](Options& o, const std::string& v) { o.max_miss = number_in(v, 0.0, 1.0, false, "--max-miss must be in [0,1]"); }},
      {"maf", 0, true,
       [](Options& o, const std::string& v) { o.maf = number_in(v, 0.0, 0.5, false, "--maf must be in [0,0.5]"); }},
      {"thread", 't', true, [](Options& o, const std::string& v) { o.threads = std::max(1, std::atoi(v.c_str())); }},
      {"threads", 0, true, [](Options& o, const std::string& v) { o.threads = std::max(1, std::atoi(v.c_str())); }},
      {"perm", 0, true, [](Options& o, const std::string& v) { o.perm = std::atoi(v.c_str()); }},
      {"permutations", 0, true, [](Options& o, const std::string& v) { o.perm = std::atoi(v.c_str()); }},
      {"perm-trans-thr", 0, true,
       [](Options& o, const std::string& v) {
         o.perm_trans_thr = number_in(v, 0.0, 1.0, true, "--perm-trans-thr must be in (0,1]");
       }},
      {"perm-trans-top", 0, true,
       [](Options& o, const std::string& v) {
         o.perm_trans_top = std::atoi(v.c_str());
         if (o.perm_trans_top < 1) die("--perm-trans-top must be >= 1");
       }},
      {"seed", 0, true, [](Options& o, const std::string& v) { o.seed = std::atoi(v.c_str()); }},
      {"disable-beta-approx", 0, false, [](Options& o, const std::string&) { o.disable_beta_approx = true; }},
      {"fast", 0, false, [](Options& o, const std::string&) { o.fast = true; }},
      {"pheno-norm", 0, true,
       [](Options& o, const std::string& v) {
         o.pheno_int = choice<bool>(v, {{"none", false}, {"int", true}}, "invalid --pheno-norm (none|int)");
       }},
      {"precision", 0, true,
       [](Options& o, const std::string& v) {
         o.precision = choice<int>(v, {{"fp32", EQTL_PREC_FP32}, {"bf16x3", EQTL_PREC_BF16X3}},
                                   "invalid --precision (fp32|bf16x3)");
       }},
      {"delta", 0, true,
       [](Options& o, const std::string& v) {
         o.delta_mode = choice<int>(v, {{"ref", EQTL_DELTA_REF}, {"newton", EQTL_DELTA_NEWTON}},
                                    "invalid --delta (ref|newton)");
       }},
      {"device", 0, true, [](Options& o, const std::string& v) { o.device = std::atoi(v.c_str()); }},
  };
  return table;
}

}  // namespace

int parse_options(int argc, char** argv, Options& opt) {
  if (argc <= 1) return 2;
  const auto& table = flag_table();
  for (int i = 1; i < argc; ++i) {
    const std::string arg = argv[i];
    if (arg == "-h" || arg == "--help") return 2;
    const Flag* f = nullptr;
    std::string value;
    bool have_value = false;
    if (arg.size() > 2 && arg.compare(0, 2, "--") == 0) {
      std::string name = arg.substr(2);
      const size_t eq = name.find('=');
      if (eq != std::string::npos) {
        value = name.substr(eq + 1);
        name.erase(eq);
        have_value = true;
      }
      for (const auto& cand : table)
        if (name == cand.name) f = &cand;
    } else if (arg.size() >= 2 && arg[0] == '-' && arg[1] != '-') {
      for (const auto& cand : table)
        if (cand.letter == arg[1]) f = &cand;
      if (f && arg.size() > 2) {
        value = arg.substr(2);
        have_value = true;
      }
    }
    if (!f) {
      std::cerr << "eqtl: unrecognized option '" << arg << "'\n";
      return 1;
    }
    if (f->value && !have_value) {
      if (i + 1 >= argc) {
        std::cerr << "eqtl: option '" << arg << "' requires an argument\n";
        return 1;
      }
      value = argv[++i];
    } else if (!f->value && have_value) {
      std::cerr << "eqtl: option '--" << f->name << "' doesn't allow an argument\n";
      return 1;
    }
    f->apply(opt, value);
  }
  if (opt.vcf.empty() && opt.bfile.empty()) {
    std::cerr << "[E] need exactly one of -v/--vcf or -b/--bfile\n";
    return 1;
  }
  if (!opt.vcf.empty() && !opt.bfile.empty()) {
    std::cerr << "[E] use only one of -v/--vcf or -b/--bfile\n";
    return 1;
  }
  if (!opt.make_grm && opt.pheno.empty()) {
    std::cerr << "[E] missing -e/--pheno (or use --make-grm)\n";
    return 1;
  }
  if (opt.window < 0) die("--window must be >= 0");
  if (opt.pheno_int) {
    for (Model m : opt.models)
      if (m == Model::Glm || m == Model::Glmm)
        die("--pheno-norm int is incompatible with count model " + model_str(m) +
            " (apply INT only for lm/lmm; count models need raw integer counts)");
  }
  return 0;
}

}  // namespace eqtl
