// C++ host side of the engine: the reference's scan/model seams (same names,
// argument meaning and error behaviour) over the C ABI in include/eqtl_gpu.h.
// The compute behind every seam runs in libeqtl_gpu.so; this layer parses
// inputs, orders samples/genes/SNPs and writes the reference's output files.
#pragma once
#include <cstdint>
#include <fstream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/eqtl_gpu.h"

namespace eqtl {

// ---- options (options.hpp:8-64, engine subset) -----------------------------
enum class Mode { Cis, Trans, All, Gw };
enum class Model { Lm, Glm, Lmm, Glmm };
enum class MissHand { Filter, Impute };

struct Options {
  std::string bfile, vcf, pheno, out = "eqtl_out", gff, covar, grm, gff_id_key;
  bool make_grm = false;
  Mode mode = Mode::All;
  std::vector<Model> models{Model::Lmm};
  int window = 1000000;
  double pval_cis = 1e-5, pval_trans = 1e-5;
  MissHand miss = MissHand::Impute;
  double max_miss = 0.8, maf = 0.0;
  int threads = 1;
  int perm = 0;
  double perm_trans_thr = 1e-5;
  int perm_trans_top = 1000;
  int seed = -1;
  bool disable_beta_approx = false;
  bool fast = false;  // --fast: sparsified GRM (|K_ij| < 1e-4 -> 0) for the LMM bases
  bool pheno_int = false;  // --pheno-norm int: per-gene rank inverse normal transform
  // engine-specific
  int precision = EQTL_PREC_BF16X3;
  int delta_mode = EQTL_DELTA_REF;
  int device = 0;
  bool use_bfile() const { return !bfile.empty(); }
};

int parse_options(int argc, char** argv, Options& opt);
std::string model_str(Model m);

// ---- util (util.cpp:15-25, 50-105) ----------------------------------------
[[noreturn]] void die(const std::string& msg);
void warn(const std::string& msg);
void info(const std::string& msg);
std::string trim(const std::string& s);
std::vector<std::string> split_ws(const std::string& s);
std::vector<std::string> split_char(const std::string& s, char c);
std::string chrom_key(const std::string& chrom);
bool chrom_equal(const std::string& a, const std::string& b);
std::vector<std::string> intersect_order(const std::vector<std::string>& prefer,
                                         const std::vector<std::string>& other);

// ---- genotype sources (plink_bed.hpp:48-73 concept) --------------------------
// A site as the outputs name it: a1 = effect (ALT) allele, a2 = other (REF).
struct BimSite {
  std::string chrom, id, a1, a2;
  int64_t pos = 0;
};

class Engine;

// Samples, sites in file order, and blocks of genotype rows in the engine's
// row format (EQTL_GENO_*) for the samples and policy already set on the
// engine.
class GenoSource {
 public:
  virtual ~GenoSource() = default;
  virtual const std::vector<std::string>& samples() const = 0;
  virtual const std::vector<BimSite>& sites() const = 0;
  virtual int32_t row_format() const = 0;
  virtual int32_t bytes_per_snp() const = 0;
  // has_index (plink_bed.hpp:69, vcf_session.hpp): region queries possible
  virtual bool has_index() const = 0;
  // Rows [j0, j0 + m); *on_dev = 1 when the pointer is device memory. Valid
  // until the next call.
  virtual const uint8_t* rows(Engine& eng, size_t j0, size_t m, int32_t* on_dev) = 0;
  // Rows of the listed SNPs, in list order.
  virtual const uint8_t* gather(Engine& eng, const std::vector<int64_t>& idx, int32_t* on_dev) = 0;
  size_t n_file() const { return samples().size(); }
  std::vector<std::string> chromosomes() const;
};

// PLINK bed/bim/fam (plink_bed.cpp:26-230).
class PlinkBed : public GenoSource {
 public:
  void open(const std::string& prefix);
  const std::vector<std::string>& samples() const override { return samples_; }
  const std::vector<BimSite>& sites() const override { return sites_; }
  int32_t row_format() const override { return EQTL_GENO_BED; }
  int32_t bytes_per_snp() const override { return bps_; }
  bool has_index() const override { return true; }
  const uint8_t* rows(Engine& eng, size_t j0, size_t m, int32_t* on_dev) override;
  const uint8_t* gather(Engine& eng, const std::vector<int64_t>& idx, int32_t* on_dev) override;
  // Reads SNP rows [j0, j0+m) into buf (m * bytes_per_snp bytes).
  void read_rows(size_t j0, size_t m, std::vector<uint8_t>& buf);

 private:
  std::string prefix_;
  std::vector<std::string> samples_;
  std::vector<BimSite> sites_;
  int32_t bps_ = 0;
  std::unique_ptr<std::ifstream> bed_;
  std::vector<uint8_t> buf_, row_;
};

// VCF (plain or gzip/BGZF-compressed) text (VcfSession, vcf_session.cpp:51-306).
// The host inflates the file (BGZF blocks in parallel), reads the header and
// the fixed columns of each record, and keeps the text; the engine decodes the
// sample columns of a block of records on the device (eqtl_gpu_vcf_decode).
class VcfSource : public GenoSource {
 public:
  void open(const std::string& path, int threads);
  const std::vector<std::string>& samples() const override { return samples_; }
  const std::vector<BimSite>& sites() const override { return sites_; }
  int32_t row_format() const override { return EQTL_GENO_DOSE32; }
  int32_t bytes_per_snp() const override { return static_cast<int32_t>(4 * samples_.size()); }
  bool has_index() const override { return indexed_; }
  const uint8_t* rows(Engine& eng, size_t j0, size_t m, int32_t* on_dev) override;
  const uint8_t* gather(Engine& eng, const std::vector<int64_t>& idx, int32_t* on_dev) override;
  size_t multi_allelic_skipped() const { return n_multi_skip_; }

  struct Rec {
    int64_t offset;  // sample columns in text_
    int32_t length;
    int16_t gt_key, ds_key;
    int32_t haploid;
  };
  const std::vector<Rec>& records() const { return recs_; }
  const std::string& text() const { return text_; }

 private:
  void parse_header_and_records();
  std::string path_;
  std::string text_;
  std::vector<std::string> samples_;
  std::vector<BimSite> sites_;
  std::vector<Rec> recs_;
  bool ds_float_ = false, indexed_ = false;
  size_t n_multi_skip_ = 0;
  std::string gather_text_;
  std::vector<eqtl_vcf_record> desc_;
};

// is_haploid_chrom (vcf_session.cpp:149-158): X, Y, MT (any case, optional "chr").
bool is_haploid_chrom(const std::string& chrom);
// Decompressed bytes of a file: gzip members (BGZF blocks inflated on
// `threads` threads) or the raw bytes when not gzip-compressed.
std::string read_maybe_gzip(const std::string& path, int threads);

// ---- phenotypes / covariates (pheno.cpp:17-196) ----------------------------
struct PhenoData {
  std::vector<std::string> sample_ids, gene_ids;
  std::vector<double> Y;  // samples x genes, column-major
  size_t n() const { return sample_ids.size(); }
  size_t G() const { return gene_ids.size(); }
};
struct CovData {
  std::vector<std::string> sample_ids, cov_names;
  std::vector<double> X;  // samples x p, column-major
  int p = 0;
};
PhenoData load_pheno(const std::string& path);
void slice_pheno(PhenoData& p, const std::vector<std::string>& sample_ids);
CovData load_covar(const std::string& path, const std::vector<std::string>& sample_ids);

// ---- annotation (annot.cpp:26-71, own GFF3 gene-line parser) ---------------
struct GeneLoc {
  std::string id, chrom;
  int64_t start = 0, end = 0, tss = 0;
  char strand = '+';
  bool ok = false;
};
std::unordered_map<std::string, GeneLoc> load_gff_tss(const std::string& path, const std::string& id_key);

// ---- outputs (output.cpp:8-110) -------------------------------------------
struct AssocHit {
  std::string gene, snp, chrom, ref, alt;
  int64_t pos = 0;
  double maf = 0, beta = 0, se = 0, stat = 0, p = 1, r2 = 0, tss_dist = 0;
  double phi = 0;              // NB dispersion of the gene's null fit (glm)
  bool glm_converged = true;
  int n = 0;
  bool has_tss_dist = false;
};
struct GeneSummary {
  std::string gene, chrom;
  int64_t tss = 0;
  double acat_p = 1, p_emp = std::numeric_limits<double>::quiet_NaN(),
         p_beta = std::numeric_limits<double>::quiet_NaN(), beta_shape1 = 0, beta_shape2 = 0,
         q_bh = std::numeric_limits<double>::quiet_NaN();
  int n_tested = 0, n_sig = 0;
};
void write_pairs_header(std::ostream& os, Model m);
void write_pair_line(std::ostream& os, const AssocHit& h, Model m, const std::string& scope);
void write_region_header(std::ostream& os);
void write_region_line(std::ostream& os, const GeneSummary& g);
void bh_adjust(std::vector<GeneSummary>& gs);

struct ScopeOut {
  std::ofstream pairs, top, region;
  std::string tag;
  // per-gene sample subsets run group by group: lines are then buffered with
  // their sort keys and written in the reference's order at the end
  bool buffered = false;
  std::vector<std::pair<std::pair<int64_t, int64_t>, std::string>> pair_buf, top_buf;
};
// One pair / top line keyed (k1, k2) for the buffered order.
void emit_line(ScopeOut& out, bool top, int64_t k1, int64_t k2, const AssocHit& h, Model m, const std::string& scope);
void flush_lines(ScopeOut& out);

// ---- GRM (grm.cpp:10-108) --------------------------------------------------
struct Grm {
  std::vector<std::string> ids;
  std::vector<double> K;  // n x n column-major
  std::vector<int32_t> N_packed;
  int n() const { return static_cast<int>(ids.size()); }
};
Grm load_grm_gcta(const std::string& prefix);
void write_grm_gcta(const std::string& prefix, const Grm& g);
Grm slice_grm(const Grm& g, const std::vector<std::string>& sample_ids);

// ---- engine seams -----------------------------------------------------------
// RAII owner of one GPU context.
class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  eqtl_gpu_ctx* ctx() const { return ctx_; }
  void check(int rc, const char* what) const;
  // eqtl_gpu_scan_block, then every hit of the block into hits (resized to
  // the block's count; its capacity is reused across blocks)
  void scan_block(const uint8_t* rows, int32_t on_dev, int64_t snp0, int64_t m, int32_t bps,
                  std::vector<eqtl_hit>& hits) const;

 private:
  eqtl_gpu_ctx* ctx_ = nullptr;
};

// compute_grm over the genotype source for the given sample order (grm.cpp:112-242).
Grm compute_grm(Engine& eng, GenoSource& geno, const std::vector<std::string>& sample_ids, MissHand miss,
                double max_miss, double maf);

struct TransJob {
  std::string gene;
  GeneLoc loc;
  bool has_loc = false;
  int col = -1;    // phenotype column
  int order = 0;   // position in the scope's output order
  std::vector<int32_t> keep;  // analysis samples with finite phenotype and covariates
  GeneSummary summary;
  AssocHit best;
};

// scan_lm_snp_outer / scan_lmm_snp_outer (scan_trans_lm.cpp:100-222,
// scan_trans_lmm.cpp:82-251) for jobs that share the analysis sample set.
void scan_snp_outer(const Options& opt, Engine& eng, GenoSource& bed, Model model, const std::string& scope,
                    ScopeOut& out, double pthr, std::vector<TransJob>& jobs, const PhenoData& ph,
                    const CovData& cov, const std::vector<double>* Q, const std::vector<double>* lambda);

int run_eqtl(const Options& opt);
int run_make_grm(const Options& opt);

}  // namespace eqtl
