/* eqtl_gpu.h — C ABI of the MI355X (gfx950) eqtl association engine.
 *
 * Plain C types only; every call returns 0 (EQTL_OK) or a negative EQTL_E_*
 * code, and eqtl_gpu_last_error() returns the calling thread's last message.
 * One host thread drives one context; all device work of a context runs in
 * order on that context's single HIP stream (default priority).
 *
 * Layouts: Q n x n column-major (column c = eigenvector c), lambda ascending;
 * Y n x G column-major (gene k contiguous); X n x p column-major with the
 * intercept included by the caller; genotype rows SNP-major, bytes_per_snp
 * each (PLINK-1 2-bit codes without the magic header, or float32 dosages: see
 * eqtl_gpu_set_geno_format).
 *
 * The reference (WWz33/eqtl, C++17) has no FFI; its seams are C++ functions.
 * Each entry point below names the reference function(s) whose work it takes
 * over; the C++ host mirror in eqtl_amd/host/ calls these in place of them.
 */
#ifndef EQTL_GPU_H
#define EQTL_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct eqtl_gpu_ctx eqtl_gpu_ctx;

enum {
  EQTL_OK = 0,
  EQTL_E_ARG = -1,      /* bad argument / shape */
  EQTL_E_HIP = -2,      /* HIP runtime error */
  EQTL_E_NOMEM = -3,    /* device allocation failed */
  EQTL_E_STATE = -4,    /* call order (e.g. scan before prep) */
  EQTL_E_SOLVER = -5,   /* eigensolver failure (make_lmm_basis dies) */
  EQTL_E_CAPACITY = -6, /* a fixed-size device buffer is too small (the message names it) */
  EQTL_E_FORMAT = -7    /* malformed input record (the message names it) */
};

/* LM: OLS Wald t test (model_lm.cpp); LMM: spectral LMM Wald (model_lmm.cpp);
 * GLM: negative-binomial (NB2, log link) score test at the null fit
 * (model_glm.cpp; --model glm). */
enum { EQTL_MODEL_LM = 0, EQTL_MODEL_LMM = 1, EQTL_MODEL_GLM = 2 };
/* Contraction precision of the fused Wald kernel (and the genotype rotation).
 * FP32: fp32 MFMA on fp32 operands. BF16X3: operands split into bf16 hi+lo,
 * three bf16 MFMA products (hi*hi + hi*lo + lo*hi) with fp32 accumulation. */
enum { EQTL_PREC_FP32 = 0, EQTL_PREC_BF16X3 = 1 };
/* REML delta: REF = restated grid + golden + numeric-Newton search
 * (model_lmm.cpp:39-86); NEWTON = grid start + safeguarded Newton on log delta
 * with the analytic adjoint gradient/Hessian; GIVEN = caller's delta_in. */
enum { EQTL_DELTA_REF = 0, EQTL_DELTA_NEWTON = 1, EQTL_DELTA_GIVEN = 2 };
/* Scan scope: TRANS tests every SNP outside the gene's window (in_cis_window,
 * scan_common.hpp:65-69; every SNP when no windows are set); CIS tests only
 * the SNPs inside it (for_each_snp_region over cis_window_bounds,
 * scan_cis.cpp:122-129). */
enum { EQTL_SCOPE_TRANS = 0, EQTL_SCOPE_CIS = 1 };

/* One significant (p <= pthr) pair — the fields of write_pair_line
 * (output.cpp:27-76) that the engine computes; text columns come from the bim. */
typedef struct {
  int64_t snp;   /* global SNP index (bim order) */
  int32_t gene;  /* gene index (order of Y's columns) */
  int32_t pad;
  double af, beta, se, stat, p, r2;
} eqtl_hit;

/* Per-gene summary of a trans scan (scan_trans.hpp:87-110 + AcatAcc). */
typedef struct {
  double acat_p;    /* AcatAcc::p() */
  double best_p;    /* 2.0 when nothing tested */
  int64_t best_snp; /* first SNP (bim order) reaching best_p; -1 if none */
  double best_af, best_beta, best_se, best_stat, best_r2;
  int32_t n_tested;
  int32_t n_sig;
} eqtl_gene_state;

/* ---- context ------------------------------------------------------------- */
int eqtl_gpu_open(int32_t device, eqtl_gpu_ctx** out);
int eqtl_gpu_close(eqtl_gpu_ctx* ctx);
const char* eqtl_gpu_last_error(void);
/* Device properties for callers that size their own work. */
int eqtl_gpu_device_info(eqtl_gpu_ctx* ctx, char* name, int32_t name_len, int32_t* n_cu,
                         int64_t* hbm_bytes);

/* ---- genotype source (PlinkBed, plink_bed.cpp:245-357) -------------------- */
/* Analysis sample order: sample_col[i] = fam column of analysis sample i
 * (PlinkBed::set_sample_order, plink_bed.cpp:245-255). */
int eqtl_gpu_set_samples(eqtl_gpu_ctx* ctx, int32_t n_file, const int32_t* sample_col, int32_t n);
/* Test subset (build_gene_ready, scan_common.cpp:11-63: the samples whose
 * phenotype and covariates are finite): keep[0..nk) are strictly increasing
 * indices into the analysis samples of eqtl_gpu_set_samples. Genotypes are
 * still decoded, filtered and mean-imputed over all analysis samples (the
 * source's SnpRec), then restricted to the subset (subset_dosage), where
 * subset_maf_or_nan and every test run; genes prepared afterwards, the
 * eigenbasis (K of nk x nk) and the scans use the nk subset samples.
 * keep = NULL (or the full identity) restores the analysis set. The GRM
 * entry points need the full set. */
int eqtl_gpu_set_test_subset(eqtl_gpu_ctx* ctx, const int32_t* keep, int32_t nk);
/* Missing policy and MAF gate (MissPolicy + maf_min). impute: 1 = mean-impute,
 * 0 = drop any SNP with a missing call. */
int eqtl_gpu_set_policy(eqtl_gpu_ctx* ctx, int32_t impute, double max_miss, double maf_min);
/* Per-SNP chromosome id (any int, equal = same chromosome) and position for
 * the trans cis-exclusion (in_cis_window, scan_common.hpp:65-69). */
int eqtl_gpu_set_snp_annot(eqtl_gpu_ctx* ctx, const int32_t* chrom, const int64_t* pos, int64_t m);
/* Decode m bed rows (K1; decode_row plink_bed.cpp:263-320 followed by
 * subset_maf_or_nan scan_common.cpp:133-151). rows_on_device: rows is a device
 * pointer. Outputs (host, nullable): dosage m x n imputed (NaN for missing when
 * not imputing), keep[m] (1 = SNP tested), af[m] (subset allele frequency). */
int eqtl_gpu_decode_bed(eqtl_gpu_ctx* ctx, const uint8_t* rows, int32_t rows_on_device, int64_t m,
                        int32_t bytes_per_snp, double* dosage, uint8_t* keep, double* af);

/* ---- genotype row formats ----------------------------------------------------
 * Rows handed to eqtl_gpu_decode_bed, _grm_accumulate, _scan_block and
 * _upload_bed are in the context's row format:
 *   BED    PLINK-1 2-bit codes, bytes_per_snp >= ceil(n_file / 4) (default);
 *   DOSE32 n_file float32 ALT dosages in file-sample order, NaN = missing call;
 *          a row the genotype source rejected (VcfSession::parse_record false,
 *          or the MAF gate of for_each_snp) is all NaN, so it is never tested.
 *          bytes_per_snp >= 4 * n_file. eqtl_gpu_vcf_decode writes this format. */
enum { EQTL_GENO_BED = 0, EQTL_GENO_DOSE32 = 1 };
int eqtl_gpu_set_geno_format(eqtl_gpu_ctx* ctx, int32_t format);

/* One VCF data line, located in a text buffer by the host (which parses the
 * fixed columns CHROM..FORMAT itself and skips records with n_allele != 2). */
typedef struct {
  int64_t offset;  /* first byte of the first sample column */
  int32_t length;  /* bytes of the sample columns, to the end of the line (no newline) */
  int16_t gt_key;  /* index of GT among the record's FORMAT keys, -1 if absent */
  int16_t ds_key;  /* index of DS among the FORMAT keys when the header declares
                    * FORMAT/DS as Float, else -1; DS is then used for every sample */
  int32_t haploid; /* 1: chromosome X, Y or MT (one allele copy per call in the AF) */
  int32_t pad;
} eqtl_vcf_record;

/* VCF sample columns -> DOSE32 rows on the device (VcfSession::parse_record,
 * vcf_session.cpp:160-286, then pass_maf, util.hpp:61-64, as for_each_snp,
 * vcf_session.cpp:288-306, applies it) for the analysis samples and policy of
 * eqtl_gpu_set_samples / _set_policy: GT (a call with a missing allele is
 * missing; one allele = haploid call) or DS (float, '.' or non-finite =
 * missing), miss_frac / AF (ALT count over the called allele copies) /
 * Filter / MAF rules. text: host memory, or 4-byte aligned device memory
 * readable for 4 bytes past text_len (text_on_device = 1). *rows_dev receives a device buffer of m
 * rows of 4 * n_file bytes owned by the context and valid until the next call.
 * EQTL_E_FORMAT if a record's sample column count differs from n_file. */
int eqtl_gpu_vcf_decode(eqtl_gpu_ctx* ctx, const char* text, int32_t text_on_device, int64_t text_len,
                        const eqtl_vcf_record* recs, int64_t m, const uint8_t** rows_dev);

/* ---- GRM (compute_grm_from, grm.cpp:112-224) ----------------------------- */
/* Accumulate m bed rows into the GRM; call repeatedly, then eqtl_gpu_grm_finish.
 * SNPs whose calls are all present and integral are held (2 bits per sample)
 * until the buffer limit or eqtl_gpu_grm_finish and then summed exactly by
 * allele-count group on the int8 matrix cores; the others are summed per call
 * in FP64. Same K either way to FP64 rounding. */
int eqtl_gpu_grm_begin(eqtl_gpu_ctx* ctx);
int eqtl_gpu_grm_accumulate(eqtl_gpu_ctx* ctx, const uint8_t* rows, int32_t rows_on_device,
                            int64_t m, int32_t bytes_per_snp);
/* Device memory of the exact-SNP path (no counterpart in the reference):
 * bytes of held 2-bit rows before the grouped Gram runs (default 2 GiB) and
 * bytes of the int8 operand expanded per Gram launch (default 2 GiB; at least
 * 16 steps of 64 SNPs). 0 keeps a value. Smaller limits cost speed, not
 * accuracy. */
int eqtl_gpu_grm_set_limits(eqtl_gpu_ctx* ctx, int64_t exact_buffer_bytes, int64_t chunk_bytes);
/* eqtl_gpu_grm_accumulate without any device allocation, for callers that
 * cannot allocate (tight serving loops, code under HIP graph capture of the
 * caller's own work, harnesses that forbid hipMalloc). It has no counterpart
 * in the reference. The block's standardised dosages, missing flags,
 * counters and (for host rows) the staged rows live in a fixed device scratch
 * array of 32 MiB that each device holds for the life of the process: a call
 * needs 256 + (host rows: m * bytes_per_snp rounded up to 256) + m_pad * n_pad
 * * 10 bytes (m_pad, n_pad = m, n rounded up to 128, 64), and a larger block
 * is refused with EQTL_E_CAPACITY (at n = 2000, up to 1,536 SNPs per call).
 * Same results as eqtl_gpu_grm_accumulate. All calls on one device share the
 * array: calls that overlap on one device (any context) race. */
int eqtl_gpu_grm_accumulate_noalloc(eqtl_gpu_ctx* ctx, const uint8_t* rows, int32_t rows_on_device,
                                    int64_t m, int32_t bytes_per_snp);
/* K n x n column-major FP64, N_packed n(n+1)/2 (GCTA .grm.N.bin order), m_used. */
int eqtl_gpu_grm_finish(eqtl_gpu_ctx* ctx, double* K, int32_t* N_packed, int64_t* m_used);

/* ---- eigenbasis (make_lmm_basis, model_lmm.cpp:106-113) ------------------ */
/* K (n x n, col-major) -> Q, lambda (ascending, clamped >= 0), FP64 on the GPU
 * (rocSOLVER dsyevd). The basis stays resident in the context. Q_out/lambda_out
 * nullable. */
int eqtl_gpu_eigh(eqtl_gpu_ctx* ctx, const double* K, int32_t n, double* Q_out, double* lambda_out);
/* Inject a basis (test hook: lets LMM permutation streams match an external Q). */
int eqtl_gpu_set_basis(eqtl_gpu_ctx* ctx, const double* Q, const double* lambda, int32_t n);

/* ---- gene preparation ------------------------------------------------------ */
/* prep_lmm for all genes at once (model_lmm.cpp:124-153; scan_trans_lmm.cpp:110-135):
 * Y~ = Q^T Y, X~ = Q^T X (once, FP64), per-gene REML delta (delta_mode), then
 * the per-gene whitened orthonormal covariate basis and null residual that the
 * fused Wald kernel contracts against. delta_in used when mode is GIVEN.
 * delta_out / rss_null_out (G each) nullable. Y (n x G) and X (n x p),
 * column-major, may be host or device memory. */
int eqtl_gpu_lmm_prep(eqtl_gpu_ctx* ctx, const double* Y, int32_t G, const double* X, int32_t p,
                      int32_t delta_mode, const double* delta_in, double* delta_out,
                      double* rss_null_out);
/* REML adjoint: f(theta), df/dtheta, d2f/dtheta2 at theta = log(delta) for all
 * G genes in one launch (reml_negll, model_lmm.cpp:10-37, differentiated).
 * Needs a prior eqtl_gpu_lmm_prep (uses its Y~, X~). */
int eqtl_gpu_reml_grad_hess(eqtl_gpu_ctx* ctx, const double* log_delta, int32_t G, double* f,
                            double* g, double* h);
/* prep_lm for all genes (model_lm.cpp:19-30): residualised y_s, yty. yty_out nullable. */
int eqtl_gpu_lm_prep(eqtl_gpu_ctx* ctx, const double* Y, int32_t G, const double* X, int32_t p,
                     double* yty_out);

/* prep_glm_nb for all genes at once (model_glm.cpp:28-98): per gene, IRLS for
 * the NB2 log-link null model with the dispersion phi re-estimated by moments
 * between rounds (nb_irls / nb_mom_phi), then the null working weights
 * w = mu / (1 + phi mu), the Cholesky factor of X^T W X and the column blocks
 * the fused kernel contracts for the score test (test_glm_nb, :100-132).
 * Y must hold non-negative integer counts (the caller skips other genes, as
 * scan.cpp:246-249 does). Genes whose fit fails (X^T W X not positive
 * definite) are not tested (the reference's p = NaN). phi_out / converged_out
 * (G each) nullable. */
int eqtl_gpu_glm_prep(eqtl_gpu_ctx* ctx, const double* Y, int32_t G, const double* X, int32_t p, double* phi_out,
                      int32_t* converged_out);

/* Cis windows for the trans exclusion, per gene: chrom id (< 0: gene has no
 * location, never excluded) and [cstart, cend] (cis_window_bounds). Optional:
 * without it nothing is excluded (scope "gw"). */
int eqtl_gpu_set_gene_windows(eqtl_gpu_ctx* ctx, const int32_t* chrom, const int64_t* cstart,
                              const int64_t* cend, int32_t G);

/* ---- SNP-block scans (scan_lm_snp_outer / scan_lmm_snp_outer; the cis scope
 * runs the same tests per gene window, scan_gene_snps without permutations) --- */
int eqtl_gpu_scan_begin(eqtl_gpu_ctx* ctx, int32_t model, int32_t precision, int32_t scope, double pthr);
/* One block of m SNPs with global indices snp0 .. snp0+m-1 (bim order). Every
 * pair with p <= pthr is kept (the reference writes each one,
 * scan_trans_lmm.cpp:185-188, scan_trans_lm.cpp:115-118): *n_hits is the
 * block's total, the first min(total, hit_cap) hits in (snp, gene) order are
 * copied to hits, and eqtl_gpu_scan_hits returns the rest. The device lists
 * grow as the block needs (EQTL_E_NOMEM only when device memory runs out). */
int eqtl_gpu_scan_block(eqtl_gpu_ctx* ctx, const uint8_t* rows, int32_t rows_on_device,
                        int64_t snp0, int64_t m, int32_t bytes_per_snp, eqtl_hit* hits,
                        int64_t hit_cap, int64_t* n_hits);
/* Hits first .. first+cap-1 (same order) of the last scanned block; *n_out =
 * how many were copied (0 past the end). Valid until the next scan_block. */
int eqtl_gpu_scan_hits(eqtl_gpu_ctx* ctx, int64_t first, eqtl_hit* hits, int64_t cap, int64_t* n_out);
/* Work split chosen by the cis autotuner (no counterpart in the reference,
 * which runs one gene window per thread, scan_cis.cpp:18-163): cis launches
 * cut a gene group's tiles of 128 SNPs into work items that persistent
 * workgroups take from a device queue, longest first. For the window-size
 * classes (group spans of <= 4, <= 32 and more tiles) tiles[3] gets the tiles
 * per item in use and timed[3] whether the class was timed (1) or still has
 * the default of 4 (0) for the current scan configuration. */
int eqtl_gpu_cis_tuning(eqtl_gpu_ctx* ctx, int32_t* tiles, int32_t* timed);
/* Final per-gene summaries (G entries). */
int eqtl_gpu_scan_end(eqtl_gpu_ctx* ctx, eqtl_gene_state* state);

/* ---- cis permutations (scan_gene_snps, scan_cis.hpp:69-180) ------------------ */
/* Resident genotypes (all m SNPs of the annotation, bim order) for the
 * permutation stage, which revisits every gene's window B times. */
int eqtl_gpu_upload_bed(eqtl_gpu_ctx* ctx, const uint8_t* rows, int32_t rows_on_device, int64_t m,
                        int32_t bytes_per_snp);
/* Gene-level permutations after an observed cis scan (scope CIS) of the
 * prepared genes: for every gene with a tested SNP and b < B, the index
 * stream mt19937((seed >= 0 ? seed : 1) + 9973 b + gene_hash[k]) + std::shuffle
 * permutes the gene's null residual (LM: y_s; LMM: the whitened spectral
 * residual, un-whitened and re-fit with its own REML delta, delta_mode REF or
 * NEWTON), and the minimum p over the gene's window is recorded.
 * gene_hash[k] = fnv1a of the gene id (util.hpp:51-58); best_p_obs = the
 * observed scan's best_p. Outputs (host, nullable): perm_min_p G x B (row k =
 * gene k; NaN for genes not permuted) and p_emp (p_emp_count,
 * stats_extra.cpp:39-46; NaN when not permuted). The gene side is consumed:
 * prepare genes again before the next scan. */
int eqtl_gpu_cis_perm(eqtl_gpu_ctx* ctx, int32_t precision, int32_t B, int32_t seed, const uint32_t* gene_hash,
                      const double* best_p_obs, int32_t delta_mode, double* perm_min_p, double* p_emp);

/* Beta-approximated permutation p per gene (beta_approx_p, stats_extra.cpp:79-158):
 * moment fit of Beta(a, b) to the gene's permutation minima in (0, 1) (>= 10
 * needed), maximum-likelihood refinement with a Nelder-Mead simplex (bounds
 * a in [0.1, 10], b in [5, 1e6]; moment fit kept when it fails), p_beta =
 * I_obs(a, b). perm_min_p: G x B host array (row k = gene k, NaN allowed),
 * obs_min_p: G. Outputs host arrays of G (shape1/shape2 nullable); NaN where
 * the reference reports NaN. */
int eqtl_gpu_beta_approx(eqtl_gpu_ctx* ctx, const double* perm_min_p, const double* obs_min_p, int32_t G, int32_t B,
                         double* p_beta, double* shape1, double* shape2);

/* ---- instrumentation ------------------------------------------------------- */
/* When enabled, the library brackets each launch of the named kernel family
 * ("wald", "num", "tests", "rotate", "decode", "reml", "vcf") with HIP events
 * on its stream. In the interpolated trans LMM form "num" is the contraction
 * over the samples, "wald" the interpolation and "tests" the FP64 test kernel;
 * otherwise "wald" is the fused Wald kernel. */
int eqtl_gpu_set_timing(eqtl_gpu_ctx* ctx, int32_t enable);
int eqtl_gpu_get_timing(eqtl_gpu_ctx* ctx, const char* family, double* total_ms, int64_t* launches);
/* Plan of the trans LMM scan prepared by eqtl_gpu_scan_begin (no counterpart
 * in the reference): *low_rank 1 when the interpolated form runs, the number
 * of 64-gene super-groups and ln-delta segments, and the mean number of
 * interpolation nodes per gene (32 to 128; 0 when the form is off). */
int eqtl_gpu_scan_lr_plan(eqtl_gpu_ctx* ctx, int32_t* low_rank, int32_t* n_super_groups, int32_t* n_segments,
                          double* mean_nodes);
/* Synthetic PLINK rows generated on the device (benchmark inputs resident in
 * HBM): m SNPs x n samples, MAF ~ U(maf_lo, maf_hi), A1 flipped by a fair
 * coin, HWE genotypes, miss rate; counter-based hash, so row j depends only
 * on (seed, j). Returns a device pointer owned by the context. */
int eqtl_gpu_synth_bed(eqtl_gpu_ctx* ctx, uint64_t seed, int64_t m, int32_t n, double maf_lo,
                       double maf_hi, double miss, uint8_t** rows_dev);
/* Copy m rows starting at row j0 of a device bed buffer back to the host. */
int eqtl_gpu_copy_rows(eqtl_gpu_ctx* ctx, const uint8_t* rows_dev, int64_t j0, int64_t m,
                       int32_t bytes_per_snp, uint8_t* out);
int eqtl_gpu_synchronize(eqtl_gpu_ctx* ctx);
/* Checks the device conventions the kernels rely on (matrix-core lane layouts,
 * direct global-to-LDS loads). *failed_mask = 0 when all hold. */
int eqtl_gpu_selftest(eqtl_gpu_ctx* ctx, int32_t* failed_mask);

#ifdef __cplusplus
}
#endif
#endif /* EQTL_GPU_H */
