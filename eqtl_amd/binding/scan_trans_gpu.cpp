// Reference-side binding: the SNP-outer trans / gw scans of WWz33/eqtl
// (scan_lm_snp_outer, src/scan_trans_lm.cpp:41-163; scan_lmm_snp_outer,
// src/scan_trans_lmm.cpp:82-251; declared in include/eqtl/scan_trans.hpp:115-123)
// implemented over this engine's C ABI (include/eqtl_gpu.h). A maintainer
// builds the reference with this translation unit in place of
// src/scan_trans_lm.cpp and src/scan_trans_lmm.cpp and links libeqtl_gpu.so;
// everything else (options, inputs, job set-up in scan.cpp, output writers,
// stage-2 permutations in scan_perm.cpp) stays the reference's own.
//
// Contract kept: the jobs' GeneReady (samples kept, y, X, basis or K) in,
// pair lines written SNP-major in gene order for p <= pthr
// (write_pair_line), per-gene n_tested / n_sig / acat_p / best, the top line
// when best p <= pthr, top-K g vectors for stage 2 when --perm is on, then
// stage2_perm_topk. Genotypes arrive as the source's SnpRec (already filtered
// and mean-imputed by PlinkBed / VcfSession) and go to the engine as float32
// dosage rows (EQTL_GENO_DOSE32), one block of kBlock SNPs at a time, with
// the block's chromosome ids and positions as the annotation the cis
// exclusion reads. Genes are grouped by their sample set (mixed-keep genes:
// one group per distinct keep, whose rows carry only the kept samples); the
// LMM uses the group's basis (gr.basis / basis_ref) or decomposes its K on
// the GPU.
#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "eqtl/plink_bed.hpp"
#include "eqtl/scan_trans.hpp"
#include "eqtl/vcf_session.hpp"
#include "eqtl_gpu.h"

namespace eqtl {
namespace {

constexpr int64_t kBlock = 16384;

void gpu_check(int rc, const char* what) {
  if (rc != EQTL_OK) die(std::string(what) + ": " + eqtl_gpu_last_error());
}

struct GpuContext {
  eqtl_gpu_ctx* c = nullptr;
  GpuContext() { gpu_check(eqtl_gpu_open(0, &c), "eqtl_gpu_open"); }
  ~GpuContext() { eqtl_gpu_close(c); }
};

// chromosome string -> engine chromosome id (chrom_equal classes share an id)
struct ChromIds {
  std::map<std::string, int32_t> id;
  int32_t of(const std::string& chrom) {
    auto it = id.emplace(chrom_key(chrom), static_cast<int32_t>(id.size()));
    return it.first->second;
  }
};

// One SNP of the current block: what the pair lines and stage 2 need.
struct BlockSnp {
  SnpRec meta;             // dosage emptied after packing
  Eigen::VectorXd dosage;  // kept only when top-K vectors are collected
};

template <class Job>
void set_summary_start(Job& j) {
  j.best.p = 2.0;
  j.summary.gene = j.gene;
  j.summary.n_tested = 0;
  j.summary.n_sig = 0;
  if (j.has_loc && j.loc.ok) {
    j.summary.chrom = j.loc.chrom;
    j.summary.tss = j.loc.tss;
  }
}

// The group's basis for the LMM: the job's own, a shared one, or K on the GPU.
template <class Job>
void set_group_basis(eqtl_gpu_ctx* c, const Options& opt, const Job& j) {
  const LmmBasis* b = j.gr.has_basis ? &j.gr.basis : j.gr.basis_ref;
  if (b && b->Q.rows() > 0) {
    gpu_check(eqtl_gpu_set_basis(c, b->Q.data(), b->lambda.data(), static_cast<int32_t>(b->Q.rows())),
              "eqtl_gpu_set_basis");
    return;
  }
  Eigen::MatrixXd K = j.gr.K_ref ? *j.gr.K_ref : j.gr.K;
  if (K.rows() == 0) die("LMM job without a basis or GRM");
  if (opt.fast) sparsify_grm(K, 1e-4);
  gpu_check(eqtl_gpu_eigh(c, K.data(), static_cast<int32_t>(K.rows()), nullptr, nullptr), "eqtl_gpu_eigh");
}

template <class Job>
void scan_group(const Options& opt, eqtl_gpu_ctx* c, bool lmm,
                const std::string& scope, ScopeOut& out, double pthr, std::vector<Job*>& grp,
                const std::function<void(const std::function<bool(const SnpRec&)>&)>& stream, Model model) {
  const std::vector<int>& keep = grp[0]->gr.keep;
  const int n = static_cast<int>(keep.size());
  const int G = static_cast<int>(grp.size());
  const int p = static_cast<int>(grp[0]->gr.X.cols());
  // genes: Y n x G, X n x p (column-major, the group's kept samples)
  std::vector<double> Y(static_cast<size_t>(n) * G);
  for (int k = 0; k < G; ++k)
    for (int i = 0; i < n; ++i) Y[static_cast<size_t>(k) * n + i] = grp[k]->gr.y(i);
  const Eigen::MatrixXd& X = grp[0]->gr.X;
  if (lmm) {
    set_group_basis(c, opt, *grp[0]);
    gpu_check(eqtl_gpu_lmm_prep(c, Y.data(), G, X.data(), p, EQTL_DELTA_REF, nullptr, nullptr, nullptr),
              "eqtl_gpu_lmm_prep");
  } else {
    gpu_check(eqtl_gpu_lm_prep(c, Y.data(), G, X.data(), p, nullptr), "eqtl_gpu_lm_prep");
  }
  // cis exclusion windows (scope "trans"); "gw" tests every SNP
  ChromIds chroms;
  std::vector<int32_t> gch(G, -1);
  std::vector<int64_t> gcs(G, 0), gce(G, 0);
  for (int k = 0; k < G; ++k)
    if (scope == "trans" && grp[k]->has_loc) {
      const auto w = cis_window_bounds(grp[k]->loc, opt.window);
      gch[k] = chroms.of(grp[k]->loc.chrom);
      gcs[k] = w.first;
      gce[k] = w.second;
    }
  std::vector<int32_t> ann_chrom(1, -2);
  std::vector<int64_t> ann_pos(1, 0);
  gpu_check(eqtl_gpu_set_snp_annot(c, ann_chrom.data(), ann_pos.data(), 1), "eqtl_gpu_set_snp_annot");
  gpu_check(eqtl_gpu_set_gene_windows(c, gch.data(), gcs.data(), gce.data(), G), "eqtl_gpu_set_gene_windows");
  const double scan_thr = opt.perm > 0 ? std::max(pthr, opt.perm_trans_thr) : pthr;
  gpu_check(eqtl_gpu_scan_begin(c, lmm ? EQTL_MODEL_LMM : EQTL_MODEL_LM, EQTL_PREC_BF16X3, EQTL_SCOPE_TRANS,
                                scan_thr),
            "eqtl_gpu_scan_begin");
  const int topK = opt.perm > 0 ? opt.perm_trans_top : 0;
  std::vector<float> rows;
  std::vector<BlockSnp> block;
  std::vector<eqtl_hit> hits(1 << 16);
  std::vector<AssocHit> best(G);
  std::vector<char> have_best(G, 0);
  auto flush = [&]() {
    const int64_t m = static_cast<int64_t>(block.size());
    if (m == 0) return;
    ann_chrom.resize(m);
    ann_pos.resize(m);
    for (int64_t j = 0; j < m; ++j) {
      ann_chrom[j] = chroms.of(block[j].meta.chrom);
      ann_pos[j] = block[j].meta.pos;
    }
    gpu_check(eqtl_gpu_set_snp_annot(c, ann_chrom.data(), ann_pos.data(), m), "eqtl_gpu_set_snp_annot");
    int64_t total = 0;
    gpu_check(eqtl_gpu_scan_block(c, reinterpret_cast<const uint8_t*>(rows.data()), 0, 0, m,
                                  static_cast<int32_t>(sizeof(float) * n), hits.data(),
                                  static_cast<int64_t>(hits.size()), &total),
              "eqtl_gpu_scan_block");
    if (total > static_cast<int64_t>(hits.size())) {
      const int64_t first = static_cast<int64_t>(hits.size());
      hits.resize(static_cast<size_t>(total));
      int64_t got = 0;
      gpu_check(eqtl_gpu_scan_hits(c, first, hits.data() + first, total - first, &got), "eqtl_gpu_scan_hits");
    }
    for (int64_t i = 0; i < std::min<int64_t>(total, static_cast<int64_t>(hits.size())); ++i) {
      const eqtl_hit& e = hits[static_cast<size_t>(i)];
      Job& job = *grp[e.gene];
      const SnpRec& s = block[static_cast<size_t>(e.snp)].meta;
      AssocHit h;
      h.beta = e.beta;
      h.se = e.se;
      h.stat = e.stat;
      h.p = e.p;
      h.r2 = e.r2;
      h.maf = e.af;
      h.n = n;
      fill_hit_meta(h, job, s);
      if (topK > 0 && e.p < opt.perm_trans_thr)
        topk_consider(job.top, topK, e.p, block[static_cast<size_t>(e.snp)].dosage);
      if (e.p < best[e.gene].p || !have_best[e.gene]) {
        best[e.gene] = h;
        have_best[e.gene] = 1;
      }
      if (e.p <= pthr) write_pair_line(out.pairs, h, model, scope);  // SNP-major, gene order
    }
    rows.clear();
    block.clear();
  };
  stream([&](const SnpRec& s) {
    const size_t base = rows.size();
    rows.resize(base + static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) rows[base + i] = static_cast<float>(s.dosage[static_cast<size_t>(keep[i])]);
    BlockSnp b;
    b.meta.chrom = s.chrom;
    b.meta.pos = s.pos;
    b.meta.id = s.id;
    b.meta.ref = s.ref;
    b.meta.alt = s.alt;
    b.meta.maf = s.maf;
    if (topK > 0) {
      b.dosage.resize(n);
      for (int i = 0; i < n; ++i) b.dosage(i) = s.dosage[static_cast<size_t>(keep[i])];
    }
    block.push_back(std::move(b));
    if (static_cast<int64_t>(block.size()) == kBlock) flush();
    return true;
  });
  flush();
  std::vector<eqtl_gene_state> st(G);
  gpu_check(eqtl_gpu_scan_end(c, st.data()), "eqtl_gpu_scan_end");
  for (int k = 0; k < G; ++k) {
    Job& job = *grp[k];
    job.summary.n_tested = st[k].n_tested;
    job.summary.n_sig = st[k].n_sig;
    job.summary.acat_p = st[k].acat_p;
    if (have_best[k] && best[k].p == st[k].best_p) job.best = best[k];
    else job.best.p = st[k].best_p <= 1.0 ? st[k].best_p : 2.0;  // best above the reporting threshold
  }
}

template <class Job, class G>
void snp_outer_gpu(const Options& opt, G& geno, const MissPolicy& mp, double maf, const std::string& scope,
                   ScopeOut& out, double pthr, std::vector<Job>& jobs, Model model) {
  if (jobs.empty()) return;
  const bool lmm = model == Model::Lmm;
  GpuContext ctx;
  eqtl_gpu_ctx* c = ctx.c;
  // analysis samples = the source's sample order; rows are their dosages
  const int n_src = static_cast<int>(geno.samples().size());
  std::vector<int32_t> cols(n_src);
  for (int i = 0; i < n_src; ++i) cols[i] = i;  // identity column map (rows are packed per group)
  gpu_check(eqtl_gpu_set_geno_format(c, EQTL_GENO_DOSE32), "eqtl_gpu_set_geno_format");
  gpu_check(eqtl_gpu_set_policy(c, 1, 1.0, 0.0), "eqtl_gpu_set_policy");  // the SnpRec is final
  for (auto& j : jobs) set_summary_start(j);
  // genes grouped by their sample set, in order of first appearance
  std::map<std::vector<int>, std::vector<Job*>> groups;
  std::vector<const std::vector<int>*> order;
  for (auto& j : jobs) {
    auto it = groups.find(j.gr.keep);
    if (it == groups.end()) {
      it = groups.emplace(j.gr.keep, std::vector<Job*>{}).first;
      order.push_back(&it->first);
    }
    it->second.push_back(&j);
  }
  for (const auto* key : order) {
    auto& grp = groups[*key];
    // the rows handed over hold the group's kept samples only (the source has
    // imputed over all analysis samples; subset_maf_or_nan then runs on these)
    const int32_t nk = static_cast<int32_t>(grp[0]->gr.keep.size());
    gpu_check(eqtl_gpu_set_samples(c, nk, cols.data(), nk), "eqtl_gpu_set_samples");
    auto stream = [&](const std::function<bool(const SnpRec&)>& fn) { geno.for_each_snp(mp, maf, fn); };
    scan_group<Job>(opt, c, lmm, scope, out, pthr, grp, stream, model);
  }
  for (auto& job : jobs) {
    if (job.best.p <= pthr && job.best.p <= 1.0) write_pair_line(out.top, job.best, model, scope);
    if (opt.perm > 0) {
      stage2_perm_topk(opt, model, job, nullptr);
    } else {
      job.summary.p_emp = std::numeric_limits<double>::quiet_NaN();
      job.summary.p_beta = std::numeric_limits<double>::quiet_NaN();
    }
  }
}

}  // namespace

template <typename G>
void scan_lm_snp_outer(const Options& opt, G& geno, const MissPolicy& mp, double maf, const std::string& scope,
                       ScopeOut& out, double pthr, std::vector<GeneLmJob>& jobs) {
  snp_outer_gpu(opt, geno, mp, maf, scope, out, pthr, jobs, Model::Lm);
}

template <typename G>
void scan_lmm_snp_outer(const Options& opt, G& geno, const MissPolicy& mp, double maf, const std::string& scope,
                        ScopeOut& out, double pthr, std::vector<GeneLmmJob>& jobs) {
  snp_outer_gpu(opt, geno, mp, maf, scope, out, pthr, jobs, Model::Lmm);
}

template void scan_lm_snp_outer<PlinkBed>(const Options&, PlinkBed&, const MissPolicy&, double, const std::string&,
                                          ScopeOut&, double, std::vector<GeneLmJob>&);
template void scan_lm_snp_outer<VcfSession>(const Options&, VcfSession&, const MissPolicy&, double,
                                            const std::string&, ScopeOut&, double, std::vector<GeneLmJob>&);
template void scan_lmm_snp_outer<PlinkBed>(const Options&, PlinkBed&, const MissPolicy&, double, const std::string&,
                                           ScopeOut&, double, std::vector<GeneLmmJob>&);
template void scan_lmm_snp_outer<VcfSession>(const Options&, VcfSession&, const MissPolicy&, double,
                                             const std::string&, ScopeOut&, double, std::vector<GeneLmmJob>&);

}  // namespace eqtl
