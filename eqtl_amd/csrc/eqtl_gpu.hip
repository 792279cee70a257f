// C-ABI implementation of include/eqtl_gpu.h: context, device buffers,
// launches. Built for gfx950 only (one translation unit).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/eqtl_gpu.h"
#include "numerics.hpp"
#include "reml_kernels.hpp"
#include "reml_device.hpp"
#include "scan_kernels.hpp"
#include "vcf_kernels.hpp"
#include "glm_kernels.hpp"
#include "wald_lmm.hpp"
#include "wald_lr.hpp"
#include "grm_kernels.hpp"
#include "reml_batch.hpp"
#include "perm_rng.hpp"
#include "beta_approx.hpp"

using namespace eqtl_amd;

// Fixed device scratch of the entry points that allocate nothing
// (eqtl_gpu_grm_accumulate_noalloc): each device holds its own copy of this
// array for the life of the process. Host code never uses its name as an
// address: eqtl_gpu_open maps it per device with hipGetSymbolAddress and the
// kernels receive that pointer. Calls that overlap on one device share it
// and race.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Device address of `scratch` per device (mapped once per device).
struct ScratchMap {
  std::mutex mu;
  std::map<int, void*> ptr;
};
ScratchMap& scratch_map() {
  static ScratchMap m;
  return m;
}

__global__ void k_fill_u64(unsigned long long* p, int count, unsigned long long v) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = v;
}

__global__ void k_fill_u32(unsigned int* p, int count, unsigned int v) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = v;
}

// |t| below which a test's numerator is recomputed in FP64 (near-null
// refinement in the Wald kernels). Chosen from each form's absolute t error
// (FP32 ~1e-6, BF16X3 ~2e-5): above the threshold the ACAT term's relative
// error stays below ~1e-2 of one term, and the refined share of tests is
// about 0.8 tau (0.04% / 0.16%).
constexpr double kRefineTauF32 = 5e-4;
constexpr double kRefineTauBF16X3 = 2e-3;
// |t| above which the numerator is refined as well (two-sided p < ~1e-15).
constexpr double kRefineTauHi = 8.0;
// NB score tests with |z| > 8 are recomputed in FP64 after the kernel (polish)
constexpr double kGlmTauHi = 8.0;
// LMM / GLM tests with S < kappa gg (covariates explaining more than 99% of
// the SNP's weighted sum of squares) are polished in FP64: the contraction's
// relative error in S grows as gg / S.
constexpr double kPolishKappa = 1e-2;
// Initial capacity of the per-block list of tests polished in FP64 (grown on
// demand), and how many are polished per launch (g~ scratch: chunk x n doubles).
constexpr int64_t kPolishCap = int64_t(1) << 16;
constexpr int64_t kPolishChunk = 4096;
// Cis work items: a gene group's span of 128-SNP tiles in a block is cut into
// items of at most cis_tiles[class] tiles, the class being the span's length
// (<= 4 tiles, <= 32 tiles, longer). The autotuner times the candidates of a
// class on the first block that has groups of that class (default 4).
constexpr int kCisClasses = 3;
constexpr int kCisClassMax[kCisClasses - 1] = {4, 32};
constexpr int kCisCand[kCisClasses][3] = {{1, 2, 4}, {2, 4, 8}, {4, 8, 16}};
constexpr int kCisItemTilesDefault = 4;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPC(expr)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(EQTL_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
  } while (0)

#define CHK(expr)            \
  do {                       \
    int r_ = (expr);         \
    if (r_ != EQTL_OK) return r_; \
  } while (0)

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

int ensure(DBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return EQTL_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  if (bytes == 0) bytes = 256;
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return fail(EQTL_E_NOMEM, "hipMalloc(" + std::to_string(bytes) + ") failed");
  }
  b.bytes = bytes;
  return EQTL_OK;
}

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

struct TimingSlot {
  double ms = 0;
  int64_t launches = 0;
};

}  // namespace

// Device work buffers of the batched REML search.
struct RemlWork {
  DBuf P, dgrid, Dg, logvg, XD, Y2, CAg, Ainv, logdet, okg, Bg, C0, F, best;
  DBuf W, Z, CA, CB, S, ek, ed, Fo;
  DBuf Csplit;  // dgemm_tn_split's partial products
  // device-resident reference search: per-gene state, evaluation points and
  // values, and the captured launch sequence with the arguments it was
  // captured for (replayed while they match)
  DBuf st, dek, ded, dF, cnt, list;
  // fragment-major A operands of k_reml_gram_f (P blocks, then X~ blocks)
  DBuf Af;
  hipGraphExec_t graph = nullptr;
  std::vector<const void*> graph_key, eager_key;
};

struct eqtl_gpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  int n_cu = 0;
  std::string name;
  int64_t hbm = 0;
  // samples / policy
  int n_file = 0, n = 0, n_pad = 0;
  DBuf sample_col;
  std::vector<int32_t> h_sample_col;
  bool col_ident = false;  // sample_col[i] == i for every analysis sample
  // test subset: n (below) test samples out of n_src analysis samples;
  // test_col = file columns of the test samples (subset active only)
  int n_src = 0;
  bool subset = false;
  DBuf test_col;
  DBuf col_used;  // n_file: how many analysis samples read each file column
  bool col_used_all = false;  // every file column read exactly once
  int geno_fmt = EQTL_GENO_BED;
  int impute = 1;
  double max_miss = 0.8, maf_min = 0.0;
  // snp annotation
  int64_t m_annot = 0;
  DBuf snp_chrom, snp_pos;
  std::vector<int32_t> h_snp_chrom;  // host copies (cis work items)
  std::vector<int64_t> h_snp_pos;
  // basis
  bool have_basis = false;
  DBuf Q64, lam64, Q32, Qhi, Qlo;
  // genes
  int G = 0, G_pad = 0, p = 0, prepped_model = -1, nq = 0;
  int nq_pad = 0;  // column blocks as stored (>= nq, zero-padded to a compiled kernel size)
  DBuf Yt64, Xt64, X64, Y64, delta, gscal, gvalid, Bf32, Bbf, Wx, zeros;
  DBuf U64, V64;  // per-gene FP64 residual columns (model space / sample space)
  DBuf Lc;        // per-gene Cholesky factor of X~^T D X~ (LMM)
  DBuf polish, polish_gt, polish_out, pbest_key;
  int64_t polish_cap_dev = 0;
  bool have_wx = false;  // Wx = orthonormal basis of the covariates (LM and LMM scans residualise on it)
  void* scratch = nullptr;  // this device's address of the fixed scratch array
  RemlWork rw;
  bool have_windows = false;
  DBuf gene_chrom, gene_cs, gene_ce;
  std::vector<int32_t> h_gene_chrom;
  std::vector<int64_t> h_gene_cs, h_gene_ce;
  // cis scope: per-gene global SNP range [jlo, jhi) holding its window (empty: jlo == jhi)
  std::vector<int64_t> win_lo, win_hi;
  DBuf work, item_first, item_order, queue;
  // cis autotuner: tiles per work item per window-size class, which classes
  // have been timed, and the scan configuration they were timed for
  int cis_tiles[kCisClasses] = {kCisItemTilesDefault, kCisItemTilesDefault, kCisItemTilesDefault};
  int cis_timed[kCisClasses] = {0, 0, 0};
  std::vector<int> cis_tune_key;
  // resident genotypes for cis scans
  DBuf bed_res;
  int64_t bed_m = 0;
  int32_t bed_bps = 0;
  // scan
  bool scanning = false;
  int scan_model = 0, scan_prec = 0, scan_scope = EQTL_SCOPE_TRANS;
  double pthr = 1e-5;
  DBuf state, partial, hits, hit_count, tail;
  double tail_df = -1;
  int64_t hit_cap_dev = 0;
  // every hit of the last scanned block, sorted by (snp, gene); the caller
  // drains what did not fit its buffer with eqtl_gpu_scan_hits
  std::vector<eqtl_hit> hit_stage;
  // per-block buffers
  int64_t m_cap = 0;
  DBuf rows, A32, A64, Gt, keep, af, gtg, dosage;
  // exact FP64 rows for the permutation decisions near the observed statistic
  DBuf ex_R, ex_keep, ex_gtg, ex_tasks, ex_out;
  int64_t ex_lo = 0, ex_hi = 0;
  bool ex_valid = false;
  int64_t perm_exact = 0;  // permutations decided from FP64 tests in the last cis_perm
  // VCF decode
  DBuf vcf_text, vcf_recs, vcf_rows, vcf_bad;
  // GRM
  DBuf grm_acc, grm_miss, grm_cnt, grm_z, grm_snp;
  int64_t grm_m = 0;
  // grouped int8 Gram of the exact SNPs: buffered 2-bit rows and allele counts,
  // u terms (n_pad) then C, and the flush's work arrays
  DBuf grm_ex_rows, grm_ex_key, grm_u, grm_hist, grm_fill, grm_gstart, grm_wstep, grm_order, grm_X, grm_upart,
      grm_tiles;
  int64_t grm_ex_n = 0;
  size_t grm_buf_bytes = size_t(2) << 30, grm_chunk_bytes = size_t(2) << 30;
  // trans LMM, split form, covariate projections interpolated in ln delta
  // (k_lr_num, k_wald_lr, k_lr_tests): built by scan_begin; slots hold the genes in ln-delta order
  bool lr_on = false;
  int lr_n_sg = 0, lr_ncols = 0, lr_nnodes = 0, lr_nseg = 0;
  double lr_mean_nodes = 0.0;  // interpolation nodes per valid gene, averaged
  int64_t lr_ldT = 0;
  DBuf lr_fold;  // k_wald_lr (gg, S), k_lr_num (num) -> k_lr_tests, per test of a block (lr_fold_base)
  DBuf lr_Z, lr_QZ, lr_nodes, lr_okb;  // plan workspace (node bases, nodes, their status)
  DBuf lr_slot_gene, lr_sg_col, lr_sg_R, lr_coef, lr_gscale, lr_gb, lr_QZhi, lr_QZlo, lr_Dhi, lr_Dlo, lr_Tp,
      lr_G2;
  // synthetic bed buffers
  std::vector<DBuf*> synth;
  // timing
  bool timing = false;
  std::map<std::string, TimingSlot> tslots;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~eqtl_gpu_ctx() {
    for (DBuf* b : synth) delete b;
    if (blas) rocblas_destroy_handle(blas);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Bracket one kernel-family launch with events when timing is on.
struct Timed {
  eqtl_gpu_ctx* c;
  const char* fam;
  Timed(eqtl_gpu_ctx* ctx, const char* f) : c(ctx), fam(f) {
    if (c->timing) (void)hipEventRecord(c->ev0, c->stream);
  }
  int done() {
    if (!c->timing) return EQTL_OK;
    HIPC(hipEventRecord(c->ev1, c->stream));
    HIPC(hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    auto& s = c->tslots[fam];
    s.ms += ms;
    s.launches += 1;
    return EQTL_OK;
  }
};

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(EQTL_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return EQTL_OK;
}

int set_device(eqtl_gpu_ctx* c) {
  HIPC(hipSetDevice(c->device));
  return EQTL_OK;
}

GeneCols gene_cols(eqtl_gpu_ctx* c) {
  GeneCols g;
  g.f32 = c->Bf32.as<float>();
  g.bf = c->Bbf.as<__bf16>();
  g.nq = c->nq_pad;
  g.nks = c->n_pad / 32;
  return g;
}

// Smallest row size of the context's genotype row format.
int64_t min_bps(const eqtl_gpu_ctx* c) {
  return c->geno_fmt == EQTL_GENO_DOSE32 ? 4 * static_cast<int64_t>(c->n_file) : (c->n_file + 3) / 4;
}

// Block buffers for up to m SNPs.
int ensure_block(eqtl_gpu_ctx* c, int64_t m, int32_t bps, bool want64) {
  const int64_t m_pad = round_up(std::max<int64_t>(m, 1), 128);
  CHK(ensure(c->rows, static_cast<size_t>(m_pad) * bps));
  CHK(ensure(c->A32, static_cast<size_t>(m_pad) * c->n_pad * sizeof(float)));
  CHK(ensure(c->Gt, static_cast<size_t>(m_pad) * c->n_pad * sizeof(float)));
  CHK(ensure(c->keep, static_cast<size_t>(m_pad)));
  CHK(ensure(c->af, static_cast<size_t>(m_pad) * sizeof(double)));
  CHK(ensure(c->gtg, static_cast<size_t>(m_pad) * sizeof(double)));
  if (want64) CHK(ensure(c->A64, static_cast<size_t>(m_pad) * c->n_pad * sizeof(double)));
  return EQTL_OK;
}

int run_decode(eqtl_gpu_ctx* c, const uint8_t* rows_dev, int64_t m, int32_t bps, bool want64,
               double* dosage_dev, double* mean_dev = nullptr) {
  const int64_t m_pad = round_up(std::max<int64_t>(m, 1), 128);
  DecodeArgs a;
  a.rows = rows_dev;
  a.m = m;
  a.m_pad = m_pad;
  a.bps = bps;
  a.n = c->n;
  a.n_src = c->n_src;
  a.n_file = c->n_file;
  a.n_pad = c->n_pad;
  a.sample_col = c->sample_col.as<int32_t>();
  a.test_col = c->subset ? c->test_col.as<int32_t>() : nullptr;
  a.impute = c->impute;
  a.max_miss = c->max_miss;
  a.maf_min = c->maf_min;
  a.A = c->A32.as<float>();
  a.A64 = want64 ? c->A64.as<double>() : nullptr;
  a.dosage = dosage_dev;
  a.keep = c->keep.as<uint8_t>();
  a.af = c->af.as<double>();
  a.mean = mean_dev;
  Timed t(c, "decode");
  const dim3 grid(static_cast<unsigned>((m_pad + 3) / 4));
  if (c->geno_fmt == EQTL_GENO_DOSE32)
    hipLaunchKernelGGL(k_decode<EQTL_GENO_DOSE32>, grid, dim3(256), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_decode<EQTL_GENO_BED>, grid, dim3(256), 0, c->stream, a);
  CHK(check_launch("k_decode"));
  return t.done();
}

const uint8_t* stage_rows(eqtl_gpu_ctx* c, const uint8_t* rows, int32_t on_dev, int64_t m, int32_t bps,
                          int* rc) {
  *rc = EQTL_OK;
  if (on_dev) return rows;
  if (hipMemcpyAsync(c->rows.p, rows, static_cast<size_t>(m) * bps, hipMemcpyHostToDevice, c->stream) !=
      hipSuccess) {
    *rc = fail(EQTL_E_HIP, "copy of bed rows failed");
    return nullptr;
  }
  return c->rows.as<uint8_t>();
}

// Cis launches are persistent: as many workgroups as can be resident (per-CU
// occupancy of the kernel at its LDS size, times the CUs), at most one per
// item, after the item queue is zeroed (a fill kernel, no memset).
int persistent_grid(eqtl_gpu_ctx* c, const WaldArgs& a, const void* fn, int threads, size_t lds, int* occ,
                    dim3* grid) {
  if (!a.work) return EQTL_OK;
  if (*occ < 0) {
    int nb = 0;
    HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, lds));
    *occ = std::max(nb, 1);
  }
  grid->x = std::min<unsigned>(grid->x, static_cast<unsigned>(std::max(c->n_cu, 1) * *occ));
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, a.queue, 1, 0u);
  return check_launch("k_fill_u32");
}

template <int PREC>
int launch_wald_lm(eqtl_gpu_ctx* c, const WaldArgs& a, dim3 grid) {
  const size_t lds = lm_lds_bytes();
  static bool attr_set = false;
  static int occ = -1;
  if (!attr_set) {
    HIPC(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_wald_lm<PREC>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    attr_set = true;
  }
  CHK(persistent_grid(c, a, reinterpret_cast<const void*>(&k_wald_lm<PREC>), 256, lds, &occ, &grid));
  hipLaunchKernelGGL((k_wald_lm<PREC>), grid, dim3(256), lds, c->stream, a);
  return check_launch("k_wald_lm");
}

// Column-block counts the LMM Wald kernel is compiled for (nq = p + 2 is
// zero-padded up to the next one; p = 16 and p = 21 are exact).
constexpr int kLmmNq[] = {4, 8, 12, 16, 18, 20, 23, 26, 30, 34};

int lmm_nq_pad(int nq) {
  for (int v : kLmmNq)
    if (nq <= v) return v;
  return -1;
}

template <int PREC, int NQ, int MODEL>
int launch_wald_lmm_one(eqtl_gpu_ctx* c, const WaldArgs& a, dim3 grid) {
  const size_t lds = w8_lds_bytes(NQ);
  static_assert(w8_lds_bytes(NQ) + 64 <= kLdsCap, "LDS of the LMM Wald kernel with its static item slot");
  static bool attr_set = false;
  static int occ = -1;
  if (!attr_set) {
    HIPC(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_wald_lmm<PREC, NQ, MODEL>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    attr_set = true;
  }
  CHK(persistent_grid(c, a, reinterpret_cast<const void*>(&k_wald_lmm<PREC, NQ, MODEL>), kW8Threads, lds, &occ,
                      &grid));
  hipLaunchKernelGGL((k_wald_lmm<PREC, NQ, MODEL>), grid, dim3(kW8Threads), lds, c->stream, a);
  return check_launch("k_wald_lmm");
}

template <int PREC, int MODEL>
int launch_wald_lmm(eqtl_gpu_ctx* c, const WaldArgs& a, dim3 grid) {
  switch (c->nq_pad) {
    case 4: return launch_wald_lmm_one<PREC, 4, MODEL>(c, a, grid);
    case 8: return launch_wald_lmm_one<PREC, 8, MODEL>(c, a, grid);
    case 12: return launch_wald_lmm_one<PREC, 12, MODEL>(c, a, grid);
    case 16: return launch_wald_lmm_one<PREC, 16, MODEL>(c, a, grid);
    case 18: return launch_wald_lmm_one<PREC, 18, MODEL>(c, a, grid);
    case 20: return launch_wald_lmm_one<PREC, 20, MODEL>(c, a, grid);
    case 23: return launch_wald_lmm_one<PREC, 23, MODEL>(c, a, grid);
    case 26: return launch_wald_lmm_one<PREC, 26, MODEL>(c, a, grid);
    case 30: return launch_wald_lmm_one<PREC, 30, MODEL>(c, a, grid);
    case 34: return launch_wald_lmm_one<PREC, 34, MODEL>(c, a, grid);
    default: return fail(EQTL_E_ARG, "too many covariates for the Wald kernel");
  }
}

int launch_wald(eqtl_gpu_ctx* c, const WaldArgs& a, dim3 grid) {
  if (c->scan_model == EQTL_MODEL_LMM) {
    if (c->scan_prec == EQTL_PREC_FP32) return launch_wald_lmm<EQTL_PREC_FP32, EQTL_MODEL_LMM>(c, a, grid);
    return launch_wald_lmm<EQTL_PREC_BF16X3, EQTL_MODEL_LMM>(c, a, grid);
  }
  if (c->scan_model == EQTL_MODEL_GLM) {
    if (c->scan_prec == EQTL_PREC_FP32) return launch_wald_lmm<EQTL_PREC_FP32, EQTL_MODEL_GLM>(c, a, grid);
    return launch_wald_lmm<EQTL_PREC_BF16X3, EQTL_MODEL_GLM>(c, a, grid);
  }
  if (c->scan_prec == EQTL_PREC_FP32) return launch_wald_lm<EQTL_PREC_FP32>(c, a, grid);
  return launch_wald_lm<EQTL_PREC_BF16X3>(c, a, grid);
}

}  // namespace

// ===========================================================================
// Runs body(lo, hi) over [0, count) split into contiguous ranges of at least
// grain items on up to eight host threads (the calling thread takes the first).
template <class Body>
static void host_parallel(int count, Body body, int grain = 256) {
  const int nt = std::max(1, std::min(8, count / grain));
  if (nt == 1) {
    body(0, count);
    return;
  }
  std::vector<std::thread> th;
  const int per = (count + nt - 1) / nt;
  for (int t = 1; t < nt; ++t) {
    const int lo = t * per, hi = std::min(count, lo + per);
    if (lo < hi) th.emplace_back(body, lo, hi);
  }
  body(0, std::min(count, per));
  for (auto& x : th) x.join();
}

extern "C" {

const char* eqtl_gpu_last_error(void) { return g_err.c_str(); }

int eqtl_gpu_open(int32_t device, eqtl_gpu_ctx** out) {
  if (!out) return fail(EQTL_E_ARG, "null out");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(EQTL_E_HIP, "no HIP device");
  if (device < 0 || device >= ndev) return fail(EQTL_E_ARG, "bad device index");
  auto* c = new eqtl_gpu_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return fail(EQTL_E_HIP, "stream/event creation failed");
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->n_cu = prop.multiProcessorCount;
    c->name = prop.gcnArchName;
    c->hbm = static_cast<int64_t>(prop.totalGlobalMem);
  }
  {
    // the fixed scratch array's address on this device (once per device)
    ScratchMap& sm = scratch_map();
    std::lock_guard<std::mutex> lk(sm.mu);
    auto it = sm.ptr.find(device);
    if (it == sm.ptr.end()) {
      void* ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr) {
        delete c;
        return fail(EQTL_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      }
      it = sm.ptr.emplace(device, ptr).first;
    }
    c->scratch = it->second;
  }
  *out = c;
  return EQTL_OK;
}

int eqtl_gpu_close(eqtl_gpu_ctx* c) {
  if (!c) return EQTL_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->rw.graph) (void)hipGraphExecDestroy(c->rw.graph);
  delete c;
  return EQTL_OK;
}

int eqtl_gpu_device_info(eqtl_gpu_ctx* c, char* name, int32_t name_len, int32_t* n_cu, int64_t* hbm) {
  if (!c) return fail(EQTL_E_ARG, "null ctx");
  if (name && name_len > 0) {
    std::snprintf(name, static_cast<size_t>(name_len), "%s", c->name.c_str());
  }
  if (n_cu) *n_cu = c->n_cu;
  if (hbm) *hbm = c->hbm;
  return EQTL_OK;
}

int eqtl_gpu_synchronize(eqtl_gpu_ctx* c) {
  if (!c) return fail(EQTL_E_ARG, "null ctx");
  HIPC(hipStreamSynchronize(c->stream));
  return EQTL_OK;
}

int eqtl_gpu_set_samples(eqtl_gpu_ctx* c, int32_t n_file, const int32_t* sample_col, int32_t n) {
  if (!c || !sample_col || n <= 0 || n_file <= 0) return fail(EQTL_E_ARG, "bad samples");
  for (int i = 0; i < n; ++i)
    if (sample_col[i] < 0 || sample_col[i] >= n_file) return fail(EQTL_E_ARG, "sample column out of range");
  CHK(set_device(c));
  if (c->n != n) {
    c->have_basis = false;
    c->prepped_model = -1;
    c->have_wx = false;
  }
  c->n_file = n_file;
  c->n = n;
  c->n_src = n;
  c->subset = false;
  c->h_sample_col.assign(sample_col, sample_col + n);
  c->col_ident = true;
  for (int i = 0; i < n; ++i) c->col_ident = c->col_ident && sample_col[i] == i;
  c->n_pad = static_cast<int>(round_up(n, 64));
  std::vector<uint8_t> used(static_cast<size_t>(n_file), 0);
  for (int i = 0; i < n; ++i) {
    if (used[sample_col[i]] == 255) return fail(EQTL_E_ARG, "a file column is repeated more than 255 times");
    ++used[sample_col[i]];
  }
  c->col_used_all = std::all_of(used.begin(), used.end(), [](uint8_t u) { return u == 1; });
  CHK(ensure(c->sample_col, sizeof(int32_t) * n));
  CHK(ensure(c->col_used, static_cast<size_t>(n_file)));
  HIPC(hipMemcpyAsync(c->sample_col.p, sample_col, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->col_used.p, used.data(), static_cast<size_t>(n_file), hipMemcpyHostToDevice, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return EQTL_OK;
}

int eqtl_gpu_set_test_subset(eqtl_gpu_ctx* c, const int32_t* keep, int32_t nk) {
  if (!c) return fail(EQTL_E_ARG, "null ctx");
  if (c->n_src <= 0) return fail(EQTL_E_STATE, "set_samples first");
  bool identity = !keep || nk == c->n_src;
  if (keep) {
    if (nk < 1 || nk > c->n_src) return fail(EQTL_E_ARG, "bad test subset size");
    for (int i = 0; i < nk; ++i) {
      if (keep[i] < 0 || keep[i] >= c->n_src || (i > 0 && keep[i] <= keep[i - 1]))
        return fail(EQTL_E_ARG, "test subset indices must be strictly increasing analysis-sample indices");
      identity = identity && keep[i] == i;
    }
  }
  CHK(set_device(c));
  const int n_new = identity ? c->n_src : nk;
  if (c->n != n_new || c->subset || !identity) {
    c->have_basis = false;
    c->prepped_model = -1;
    c->have_wx = false;
  }
  c->n = n_new;
  c->n_pad = static_cast<int>(round_up(n_new, 64));
  c->subset = !identity;
  if (c->subset) {
    std::vector<int32_t> tc(nk);
    for (int i = 0; i < nk; ++i) tc[i] = c->h_sample_col[keep[i]];
    CHK(ensure(c->test_col, sizeof(int32_t) * nk));
    HIPC(hipMemcpyAsync(c->test_col.p, tc.data(), sizeof(int32_t) * nk, hipMemcpyHostToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
  }
  return EQTL_OK;
}

int eqtl_gpu_set_geno_format(eqtl_gpu_ctx* c, int32_t format) {
  if (!c) return fail(EQTL_E_ARG, "null ctx");
  if (format != EQTL_GENO_BED && format != EQTL_GENO_DOSE32) return fail(EQTL_E_ARG, "unknown genotype row format");
  c->geno_fmt = format;
  return EQTL_OK;
}

int eqtl_gpu_vcf_decode(eqtl_gpu_ctx* c, const char* text, int32_t on_dev, int64_t text_len,
                        const eqtl_vcf_record* recs, int64_t m, const uint8_t** rows_dev) {
  if (!c || !rows_dev || m <= 0 || !recs || !text || text_len < 0) return fail(EQTL_E_ARG, "bad vcf_decode arguments");
  if (c->n <= 0 || c->n_file <= 0) return fail(EQTL_E_STATE, "set_samples first");
  if (m > 0x7fffffff) return fail(EQTL_E_ARG, "too many records in one call");
  if (on_dev && (reinterpret_cast<uintptr_t>(text) & 3)) return fail(EQTL_E_ARG, "device text must be 4-byte aligned");
  bool any_ds = false;
  for (int64_t r = 0; r < m; ++r) {
    const eqtl_vcf_record& e = recs[r];
    any_ds |= e.ds_key >= 0;
    if (e.offset < 0 || e.length < 0 || e.offset + e.length > text_len || e.gt_key < -1 || e.ds_key < -1)
      return fail(EQTL_E_ARG, "VCF record " + std::to_string(r) + ": span outside the text or bad key index");
  }
  CHK(set_device(c));
  const unsigned char* td = reinterpret_cast<const unsigned char*>(text);
  if (!on_dev) {
    CHK(ensure(c->vcf_text, static_cast<size_t>(text_len) + 16));
    HIPC(hipMemcpyAsync(c->vcf_text.p, text, static_cast<size_t>(text_len), hipMemcpyHostToDevice, c->stream));
    td = c->vcf_text.as<unsigned char>();
  }
  CHK(ensure(c->vcf_recs, sizeof(eqtl_vcf_record) * static_cast<size_t>(m)));
  CHK(ensure(c->vcf_rows, sizeof(float) * static_cast<size_t>(m) * c->n_file));
  CHK(ensure(c->vcf_bad, sizeof(unsigned long long)));
  HIPC(hipMemcpyAsync(c->vcf_recs.p, recs, sizeof(eqtl_vcf_record) * static_cast<size_t>(m), hipMemcpyHostToDevice,
                      c->stream));
  const unsigned long long none = static_cast<unsigned long long>(m);
  HIPC(hipMemcpyAsync(c->vcf_bad.p, &none, sizeof(none), hipMemcpyHostToDevice, c->stream));
  VcfArgs a;
  a.text = td;
  a.recs = c->vcf_recs.as<eqtl_vcf_record>();
  a.m = m;
  a.n_file = c->n_file;
  a.n = c->n_src;
  a.used = c->col_used.as<uint8_t>();
  a.used_all = c->col_used_all ? 1 : 0;
  a.impute = c->impute;
  a.max_miss = c->max_miss;
  a.maf_min = c->maf_min;
  a.rows = c->vcf_rows.as<float>();
  a.bad = c->vcf_bad.as<unsigned long long>();
  Timed t(c, "vcf");
  const dim3 grid(static_cast<unsigned>((m + kVcfWaves - 1) / kVcfWaves));
  if (any_ds)
    hipLaunchKernelGGL(k_vcf_decode<true>, grid, dim3(kVcfThreads), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_vcf_decode<false>, grid, dim3(kVcfThreads), 0, c->stream, a);
  CHK(check_launch("k_vcf_decode"));
  CHK(t.done());
  unsigned long long bad = none;
  HIPC(hipMemcpyAsync(&bad, c->vcf_bad.p, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  if (bad < none)
    return fail(EQTL_E_FORMAT, "VCF record " + std::to_string(bad) + " of the block: the number of sample columns "
                               "differs from the header's " + std::to_string(c->n_file) + " samples");
  *rows_dev = c->vcf_rows.as<uint8_t>();
  return EQTL_OK;
}

int eqtl_gpu_set_policy(eqtl_gpu_ctx* c, int32_t impute, double max_miss, double maf_min) {
  if (!c) return fail(EQTL_E_ARG, "null ctx");
  c->impute = impute ? 1 : 0;
  c->max_miss = max_miss;
  c->maf_min = maf_min;
  return EQTL_OK;
}

int eqtl_gpu_set_snp_annot(eqtl_gpu_ctx* c, const int32_t* chrom, const int64_t* pos, int64_t m) {
  if (!c || !chrom || !pos || m <= 0) return fail(EQTL_E_ARG, "bad snp annotation");
  CHK(set_device(c));
  CHK(ensure(c->snp_chrom, sizeof(int32_t) * m));
  CHK(ensure(c->snp_pos, sizeof(int64_t) * m));
  HIPC(hipMemcpyAsync(c->snp_chrom.p, chrom, sizeof(int32_t) * m, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->snp_pos.p, pos, sizeof(int64_t) * m, hipMemcpyHostToDevice, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  c->m_annot = m;
  c->h_snp_chrom.assign(chrom, chrom + m);
  c->h_snp_pos.assign(pos, pos + m);
  return EQTL_OK;
}

int eqtl_gpu_decode_bed(eqtl_gpu_ctx* c, const uint8_t* rows, int32_t on_dev, int64_t m, int32_t bps,
                        double* dosage, uint8_t* keep, double* af) {
  if (!c || !rows || m <= 0) return fail(EQTL_E_ARG, "bad decode arguments");
  if (c->n <= 0) return fail(EQTL_E_STATE, "set_samples first");
  if (bps < min_bps(c)) return fail(EQTL_E_ARG, "bytes_per_snp too small for n_file");
  CHK(set_device(c));
  CHK(ensure_block(c, m, bps, false));
  if (dosage) CHK(ensure(c->dosage, sizeof(double) * static_cast<size_t>(m) * c->n));
  int rc;
  const uint8_t* rd = stage_rows(c, rows, on_dev, m, bps, &rc);
  if (rc) return rc;
  CHK(run_decode(c, rd, m, bps, false, dosage ? c->dosage.as<double>() : nullptr));
  if (dosage)
    HIPC(hipMemcpyAsync(dosage, c->dosage.p, sizeof(double) * static_cast<size_t>(m) * c->n,
                        hipMemcpyDeviceToHost, c->stream));
  if (keep) HIPC(hipMemcpyAsync(keep, c->keep.p, static_cast<size_t>(m), hipMemcpyDeviceToHost, c->stream));
  if (af) HIPC(hipMemcpyAsync(af, c->af.p, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return EQTL_OK;
}

// ---- basis ----------------------------------------------------------------
static int upload_basis(eqtl_gpu_ctx* c, int n) {
  // Q32/Qhi/Qlo coordinate-major padded copies for the rotation kernel
  const int np = c->n_pad;
  std::vector<double> Q(static_cast<size_t>(n) * n);
  HIPC(hipMemcpyAsync(Q.data(), c->Q64.p, sizeof(double) * Q.size(), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  std::vector<float> q32(static_cast<size_t>(np) * np, 0.f);
  for (int cc = 0; cc < n; ++cc)
    for (int i = 0; i < n; ++i) q32[static_cast<size_t>(cc) * np + i] = static_cast<float>(Q[static_cast<size_t>(cc) * n + i]);
  // the split planes in the rotation kernel's fragment-major B order
  const size_t qb = static_cast<size_t>(rot_b_coords(np)) * np;
  const __bf16 z16 = static_cast<__bf16>(0.f);
  std::vector<__bf16> qh(qb, z16), ql(qb, z16);
  for (int cc = 0; cc < np; ++cc)
    for (int i = 0; i < np; ++i) {
      const float f = q32[static_cast<size_t>(cc) * np + i];
      const __bf16 h = static_cast<__bf16>(f);
      const int64_t o = rot_b_index(cc, i, np / 32);
      qh[o] = h;
      ql[o] = static_cast<__bf16>(f - static_cast<float>(h));
    }
  CHK(ensure(c->Q32, sizeof(float) * q32.size()));
  CHK(ensure(c->Qhi, sizeof(__bf16) * qh.size()));
  CHK(ensure(c->Qlo, sizeof(__bf16) * ql.size()));
  HIPC(hipMemcpyAsync(c->Q32.p, q32.data(), sizeof(float) * q32.size(), hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->Qhi.p, qh.data(), sizeof(__bf16) * qh.size(), hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->Qlo.p, ql.data(), sizeof(__bf16) * ql.size(), hipMemcpyHostToDevice, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  c->have_basis = true;
  c->prepped_model = -1;
  return EQTL_OK;
}

static int ensure_blas(eqtl_gpu_ctx* c) {
  if (c->blas) return EQTL_OK;
  if (rocblas_create_handle(&c->blas) != rocblas_status_success) return fail(EQTL_E_HIP, "rocblas_create_handle failed");
  if (rocblas_set_stream(c->blas, c->stream) != rocblas_status_success) return fail(EQTL_E_HIP, "rocblas_set_stream failed");
  return EQTL_OK;
}

// ---- batched REML (reml_negll search, model_lmm.cpp:39-86) -------------------

constexpr int kEvalChunk = 16384;  // evaluations per GEMM batch

// C (m x nn, ldc) = A^T B with A k x m (lda) and B k x nn (ldb): the FP64
// matrix-core kernel k_dgemm_tn (no library call, so it can be captured).
// nn_dev: when given, the column count is read on the device (nn bounds it).
static int dgemm_tn(eqtl_gpu_ctx* c, int m, int nn, int k, const double* A, int lda, const double* B, int ldb, double* C,
             int ldc, const int* nn_dev = nullptr) {
  if (m <= 0 || nn <= 0) return EQTL_OK;
  const dim3 grid(static_cast<unsigned>((m + kDgT - 1) / kDgT), static_cast<unsigned>((nn + kDgT - 1) / kDgT));
  hipLaunchKernelGGL(k_dgemm_tn, grid, dim3(256), 0, c->stream, m, nn, k, A, lda, B, ldb, C, ldc, nn_dev, k,
                     static_cast<int64_t>(0));
  return check_launch("k_dgemm_tn");
}

// The same product for shapes with few output tiles (the late Newton rounds
// of the REML search, a few hundred genes): the samples are split over up to
// 16 workgroups per tile (at most 512 workgroups in all), each writing its
// own partial product to the workspace, then summed in split order (the
// result does not depend on scheduling). Not for captured sequences (the
// workspace may grow).
static int dgemm_tn_split(eqtl_gpu_ctx* c, int m, int nn, int k, const double* A, int lda, const double* B, int ldb,
                          double* C, int ldc) {
  if (m <= 0 || nn <= 0) return EQTL_OK;
  const int gx = (m + kDgT - 1) / kDgT, gy = (nn + kDgT - 1) / kDgT;
  const int tiles = gx * gy;
  int S = std::min(16, std::min(512 / std::max(tiles, 1), (k + 127) / 128));
  if (tiles >= 256 || S < 2) return dgemm_tn(c, m, nn, k, A, lda, B, ldb, C, ldc);
  const int kchunk = (((k + S - 1) / S) + 31) / 32 * 32;
  S = (k + kchunk - 1) / kchunk;
  const int64_t mn = static_cast<int64_t>(m) * nn;
  CHK(ensure(c->rw.Csplit, sizeof(double) * static_cast<size_t>(S) * mn));
  double* P = c->rw.Csplit.as<double>();
  hipLaunchKernelGGL(k_dgemm_tn, dim3(gx, gy, S), dim3(256), 0, c->stream, m, nn, k, A, lda, B, ldb, P, m,
                     static_cast<const int*>(nullptr), kchunk, mn);
  CHK(check_launch("k_dgemm_tn"));
  const int blocks = static_cast<int>(std::min<int64_t>((mn + 255) / 256, 1024));
  hipLaunchKernelGGL(k_dgemm_sum_splits, dim3(blocks), dim3(256), 0, c->stream, P, S, m, nn, C, ldc);
  return check_launch("k_dgemm_sum_splits");
}

// Blocks of 16 packed entries (MB) and of 16 covariates (PB) that
// k_reml_gram_f is compiled for at p (p <= 22).
static void gram_blocks(int p, int* mb, int* pb) {
  const int need = (p * (p + 1) / 2 + 15) / 16;
  *mb = need <= 2 ? 2 : need <= 4 ? 4 : need <= 6 ? 6 : need <= 9 ? 9 : need <= 12 ? 12 : 16;
  *pb = p <= 16 ? 1 : 2;
}

// Per-sample outer products of X~ (once per lmm_prep), and for p <= 22 the
// same P and X~ packed fragment-major for k_reml_gram_f.
static int reml_setup(eqtl_gpu_ctx* c) {
  const int n = c->n, p = c->p, TA = p * (p + 1) / 2;
  CHK(ensure(c->rw.P, sizeof(double) * static_cast<size_t>(n) * TA));
  hipLaunchKernelGGL(k_reml_outer, dim3((n + 255) / 256, TA), dim3(256), 0, c->stream, c->Xt64.as<double>(), n, p,
                     c->rw.P.as<double>());
  CHK(check_launch("k_reml_outer"));
  if (p > 22) return EQTL_OK;
  int mb = 0, pb = 0;
  gram_blocks(p, &mb, &pb);
  const size_t nks = (n + 31) / 32;
  CHK(ensure(c->rw.Af, nks * (mb + pb) * 4096));
  hipLaunchKernelGGL(k_frag_pack, dim3(256), dim3(256), 0, c->stream, c->rw.P.as<double>(), TA, n, mb, 0, mb + pb,
                     c->rw.Af.as<double>());
  hipLaunchKernelGGL(k_frag_pack, dim3(64), dim3(256), 0, c->stream, c->Xt64.as<double>(), p, n, pb, mb, mb + pb,
                     c->rw.Af.as<double>());
  return check_launch("k_frag_pack");
}

// f-only evaluations at (gene ek[e], delta ed[e]), e < E (or e < *E_dev,
// counted on the device, E bounding it), with device-side arguments and result
// F[e]. p <= 22: the fused weighted-Gram kernel k_reml_gram_f (weights formed
// in registers, CA, CB, c0 and sum ln v in one launch); larger p: the weights
// kernel, then the X~^T D X~ GEMM. Then the bordered LDLT. The workspace
// (W for p > 22, CA, CB, S) must hold E evaluations.
extern "C++" {
template <int MB, int PB>
static int launch_gram_f(eqtl_gpu_ctx* c, int E, const int* E_dev, const int32_t* ek, const double* ed) {
  auto& r = c->rw;
  const int n = c->n, p = c->p, TA = p * (p + 1) / 2;
  const size_t lds = 2 * gram_stage_bytes<MB, PB>();
  static bool attr_set = false;
  if (!attr_set) {
    HIPC(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_reml_gram_f<MB, PB>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    attr_set = true;
  }
  hipLaunchKernelGGL((k_reml_gram_f<MB, PB>), dim3(static_cast<unsigned>((E + 63) / 64)), dim3(256), lds, c->stream,
                     r.Af.as<double>(), c->Yt64.as<double>(), c->lam64.as<double>(), n, p, TA, E, E_dev, ek, ed,
                     r.CA.as<double>(), r.CB.as<double>(), r.S.as<double>(), kNumScal);
  return check_launch("k_reml_gram_f");
}
}

static int enqueue_eval_f(eqtl_gpu_ctx* c, int E, const int32_t* ek, const double* ed, double* F,
                          const int* E_dev = nullptr) {
  const int n = c->n, p = c->p, TA = p * (p + 1) / 2, df = n - p;
  auto& r = c->rw;
  if (p <= 22) {
    int mb = 0, pb = 0;
    gram_blocks(p, &mb, &pb);
    int rc;
    switch (mb) {
      case 2: rc = launch_gram_f<2, 1>(c, E, E_dev, ek, ed); break;
      case 4: rc = launch_gram_f<4, 1>(c, E, E_dev, ek, ed); break;
      case 6: rc = launch_gram_f<6, 1>(c, E, E_dev, ek, ed); break;
      case 9: rc = launch_gram_f<9, 1>(c, E, E_dev, ek, ed); break;
      case 12: rc = launch_gram_f<12, 2>(c, E, E_dev, ek, ed); break;
      default: rc = launch_gram_f<16, 2>(c, E, E_dev, ek, ed); break;
    }
    CHK(rc);
  } else {
    // b, c0 and sum ln v from the weights kernel; one GEMM for X~^T D X~
    hipLaunchKernelGGL((k_reml_eval_weights_f<kRP, 2>), dim3((E + 1) / 2), dim3(64), 0, c->stream,
                       c->lam64.as<double>(), c->Yt64.as<double>(), c->Xt64.as<double>(), n, p, E, E_dev, ek, ed,
                       r.W.as<double>(), r.CB.as<double>(), r.S.as<double>());
    CHK(check_launch("k_reml_eval_weights_f"));
    CHK(dgemm_tn(c, TA, E, n, r.P.as<double>(), n, r.W.as<double>(), n, r.CA.as<double>(), TA, E_dev));
  }
  if (p < 32) {
    const size_t lds = 2 * reml_eval_f_bytes(p);
    hipLaunchKernelGGL(k_reml_eval_f<32>, dim3((E + 1) / 2), dim3(64), lds, c->stream, r.CA.as<double>(),
                       r.CB.as<double>(), r.S.as<double>(), E, E_dev, TA, p, df, F);
  } else {
    hipLaunchKernelGGL(k_reml_eval_f<64>, dim3(E), dim3(64), reml_eval_f_bytes(p), c->stream, r.CA.as<double>(),
                       r.CB.as<double>(), r.S.as<double>(), E, E_dev, TA, p, df, F);
  }
  return check_launch("k_reml_eval_f");
}

// f (and f', f'' in ln delta) at (gene ek[e], delta ed[e]) for all e.
static int reml_eval(eqtl_gpu_ctx* c, const std::vector<int32_t>& ek, const std::vector<double>& ed, bool deriv, double* F,
              double* Gd, double* H) {
  const int n = c->n, p = c->p, TA = p * (p + 1) / 2, df = n - p;
  const int E = static_cast<int>(ek.size());
  const int nw = deriv ? 3 : 1;
  // up to 32,768 evaluations per batch (workspace n x 3 x 32,768 doubles at
  // most): fewer, larger launches of the weight, DGEMM and evaluation kernels
  const int chunk = 2 * kEvalChunk;
  auto& r = c->rw;
  const size_t cols = static_cast<size_t>(nw) * std::min(E, chunk);
  CHK(ensure(r.W, sizeof(double) * n * cols));
  if (deriv) CHK(ensure(r.Z, sizeof(double) * n * cols));
  CHK(ensure(r.CA, sizeof(double) * TA * cols));
  CHK(ensure(r.CB, sizeof(double) * p * cols));
  CHK(ensure(r.S, sizeof(double) * kNumScal * std::min(E, chunk)));
  CHK(ensure(r.ek, sizeof(int32_t) * std::min(E, chunk)));
  CHK(ensure(r.ed, sizeof(double) * std::min(E, chunk)));
  CHK(ensure(r.Fo, sizeof(double) * 3 * std::min(E, chunk)));
  for (int e0 = 0; e0 < E; e0 += chunk) {
    const int Ec = std::min(chunk, E - e0);
    double* Fo = r.Fo.as<double>();
    double* Go = Fo + Ec;
    double* Ho = Fo + 2 * Ec;
    HIPC(hipMemcpyAsync(r.ek.p, ek.data() + e0, sizeof(int32_t) * Ec, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(r.ed.p, ed.data() + e0, sizeof(double) * Ec, hipMemcpyHostToDevice, c->stream));
    if (deriv) {
      hipLaunchKernelGGL(k_reml_eval_weights_fgh, dim3(Ec), dim3(256), 0, c->stream, c->lam64.as<double>(),
                         c->Yt64.as<double>(), n, Ec, r.ek.as<int32_t>(), r.ed.as<double>(), r.W.as<double>(),
                         r.Z.as<double>(), r.S.as<double>());
      CHK(check_launch("k_reml_eval_weights_fgh"));
      CHK(dgemm_tn_split(c, TA, nw * Ec, n, r.P.as<double>(), n, r.W.as<double>(), n, r.CA.as<double>(), TA));
      CHK(dgemm_tn_split(c, p, nw * Ec, n, c->Xt64.as<double>(), n, r.Z.as<double>(), n, r.CB.as<double>(), p));
      hipLaunchKernelGGL(k_reml_eval_fgh, dim3(Ec), dim3(64), 0, c->stream, r.CA.as<double>(), r.CB.as<double>(),
                         r.S.as<double>(), Ec, TA, p, df, Fo, Go, Ho);
      CHK(check_launch("k_reml_eval_fgh"));
    } else {
      CHK(enqueue_eval_f(c, Ec, r.ek.as<int32_t>(), r.ed.as<double>(), Fo));
    }
    HIPC(hipMemcpyAsync(F + e0, Fo, sizeof(double) * Ec, hipMemcpyDeviceToHost, c->stream));
    if (deriv) {
      HIPC(hipMemcpyAsync(Gd + e0, Go, sizeof(double) * Ec, hipMemcpyDeviceToHost, c->stream));
      HIPC(hipMemcpyAsync(H + e0, Ho, sizeof(double) * Ec, hipMemcpyDeviceToHost, c->stream));
    }
  }
  HIPC(hipStreamSynchronize(c->stream));
  return EQTL_OK;
}

// Reference grid (delta = 1, then 1e-5 * 2^k <= 1e5) for all genes at once:
// workspace and the grid's deltas (reml_grid_setup, outside any capture), then
// the launches that leave each gene's best grid delta in rw.best (device).
static int reml_grid_setup(eqtl_gpu_ctx* c, int G, int* ng_out) {
  const int n = c->n, p = c->p, TA = p * (p + 1) / 2;
  std::vector<double> dg{1.0};
  for (double d = 1e-5; d <= 1e5; d *= 2.0) dg.push_back(d);
  const int ng = static_cast<int>(dg.size());
  auto& r = c->rw;
  CHK(ensure(r.dgrid, sizeof(double) * ng));
  CHK(ensure(r.Dg, sizeof(double) * static_cast<size_t>(n) * ng));
  CHK(ensure(r.logvg, sizeof(double) * ng));
  CHK(ensure(r.XD, sizeof(double) * static_cast<size_t>(n) * ng * p));
  CHK(ensure(r.Y2, sizeof(double) * static_cast<size_t>(n) * G));
  CHK(ensure(r.CAg, sizeof(double) * static_cast<size_t>(TA) * ng));
  CHK(ensure(r.Ainv, sizeof(double) * static_cast<size_t>(p) * p * ng));
  CHK(ensure(r.logdet, sizeof(double) * ng));
  CHK(ensure(r.okg, sizeof(int32_t) * ng));
  CHK(ensure(r.Bg, sizeof(double) * static_cast<size_t>(ng) * p * G));
  CHK(ensure(r.C0, sizeof(double) * static_cast<size_t>(ng) * G));
  CHK(ensure(r.F, sizeof(double) * static_cast<size_t>(ng) * G));
  CHK(ensure(r.best, sizeof(double) * 2 * G));
  HIPC(hipMemcpyAsync(r.dgrid.p, dg.data(), sizeof(double) * ng, hipMemcpyHostToDevice, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  *ng_out = ng;
  return EQTL_OK;
}

static int reml_grid_enqueue(eqtl_gpu_ctx* c, int G, int ng) {
  const int n = c->n, p = c->p, TA = p * (p + 1) / 2, df = n - p;
  auto& r = c->rw;
  hipLaunchKernelGGL(k_reml_grid_weights, dim3(ng), dim3(256), 0, c->stream, c->lam64.as<double>(), n,
                     r.dgrid.as<double>(), r.Dg.as<double>(), r.logvg.as<double>());
  CHK(check_launch("k_reml_grid_weights"));
  hipLaunchKernelGGL(k_reml_grid_xd, dim3((n + 255) / 256, ng * p), dim3(256), 0, c->stream, c->Xt64.as<double>(),
                     r.Dg.as<double>(), n, p, r.XD.as<double>());
  CHK(check_launch("k_reml_grid_xd"));
  const size_t NG = static_cast<size_t>(n) * G;
  hipLaunchKernelGGL(k_square, dim3(static_cast<unsigned>((NG + 255) / 256)), dim3(256), 0, c->stream,
                     c->Yt64.as<double>(), NG, r.Y2.as<double>());
  CHK(check_launch("k_square"));
  CHK(dgemm_tn(c, TA, ng, n, r.P.as<double>(), n, r.Dg.as<double>(), n, r.CAg.as<double>(), TA));
  CHK(dgemm_tn(c, ng * p, G, n, r.XD.as<double>(), n, c->Yt64.as<double>(), n, r.Bg.as<double>(), ng * p));
  CHK(dgemm_tn(c, ng, G, n, r.Dg.as<double>(), n, r.Y2.as<double>(), n, r.C0.as<double>(), ng));
  hipLaunchKernelGGL(k_reml_grid_factor, dim3(ng), dim3(64), 0, c->stream, r.CAg.as<double>(), TA, p,
                     r.Ainv.as<double>(), r.logdet.as<double>(), r.okg.as<int32_t>());
  CHK(check_launch("k_reml_grid_factor"));
  hipLaunchKernelGGL(k_reml_grid_f, dim3((G + 255) / 256, ng), dim3(256), 0, c->stream, r.Ainv.as<double>(),
                     r.CAg.as<double>(), TA, r.logdet.as<double>(), r.okg.as<int32_t>(), r.logvg.as<double>(), r.Bg.as<double>(),
                     r.C0.as<double>(), ng, p, G, df, r.F.as<double>());
  CHK(check_launch("k_reml_grid_f"));
  hipLaunchKernelGGL(k_reml_grid_argmin, dim3((G + 255) / 256), dim3(256), 0, c->stream, r.F.as<double>(),
                     r.dgrid.as<double>(), ng, G, r.best.as<double>(), r.best.as<double>() + G);
  return check_launch("k_reml_grid_argmin");
}

static int reml_grid(eqtl_gpu_ctx* c, std::vector<double>& best_d) {
  const int G = c->G;
  int ng = 0;
  CHK(reml_grid_setup(c, G, &ng));
  CHK(reml_grid_enqueue(c, G, ng));
  best_d.resize(G);
  HIPC(hipMemcpyAsync(best_d.data(), c->rw.best.p, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return EQTL_OK;
}

// The reference search (optimize_delta, model_lmm.cpp:39-86) for genes
// [0, G) entirely on the device: grid, 40 golden-section steps on
// [best / 4, best * 4] clipped to [1e-8, 1e6], the midpoint, then 8 rounds of
// central-difference Newton steps with the reference's acceptance rules (a
// gene that stops keeps its delta; its later evaluations are not used). The
// per-gene decisions run in k_ref_* kernels between the evaluations, so the
// whole search is one fixed launch sequence with no host round trip.
static int reml_ref_enqueue(eqtl_gpu_ctx* c, int G, int ng, double* delta) {
  auto& r = c->rw;
  const dim3 gb(static_cast<unsigned>((G + 255) / 256)), tb(256);
  RemlRefState* st = r.st.as<RemlRefState>();
  int32_t* ek = r.dek.as<int32_t>();
  double* ed = r.ded.as<double>();
  double* F = r.dF.as<double>();
  hipLaunchKernelGGL(k_iota, dim3(static_cast<unsigned>((2 * G + 255) / 256)), tb, 0, c->stream, ek, G, 2);
  CHK(check_launch("k_iota"));
  CHK(reml_grid_enqueue(c, G, ng));
  hipLaunchKernelGGL(k_ref_init, gb, tb, 0, c->stream, r.best.as<double>(), G, st);
  const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  for (int it = 0; it < 40; ++it) {
    hipLaunchKernelGGL(k_ref_golden_points, gb, tb, 0, c->stream, st, G, phi, ed);
    CHK(enqueue_eval_f(c, 2 * G, ek, ed, F));
    hipLaunchKernelGGL(k_ref_golden_update, gb, tb, 0, c->stream, st, G, ed, F);
  }
  hipLaunchKernelGGL(k_ref_mid, gb, tb, 0, c->stream, st, G, ed);
  CHK(enqueue_eval_f(c, G, ek, ed, F));
  hipLaunchKernelGGL(k_ref_mid_done, gb, tb, 0, c->stream, st, G, F);
  // Newton rounds over the genes still active, compacted on the device
  // (count, then twice the count for the paired difference points)
  int* cnt = r.cnt.as<int>();
  int* list = r.list.as<int>();
  for (int it = 0; it < 8; ++it) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<unsigned int*>(cnt), 2, 0u);
    hipLaunchKernelGGL(k_ref_newton_points, gb, tb, 0, c->stream, st, G, cnt, list, ek, ed);
    hipLaunchKernelGGL(k_ref_twice, dim3(1), dim3(64), 0, c->stream, cnt, cnt + 1);
    CHK(enqueue_eval_f(c, 2 * G, ek, ed, F, cnt + 1));
    hipLaunchKernelGGL(k_ref_newton_propose, gb, tb, 0, c->stream, st, cnt, list, F, ek, ed);
    CHK(enqueue_eval_f(c, G, ek, ed, F, cnt));
    hipLaunchKernelGGL(k_ref_newton_accept, gb, tb, 0, c->stream, st, cnt, list, F);
  }
  hipLaunchKernelGGL(k_ref_out, gb, tb, 0, c->stream, st, G, delta);
  return check_launch("k_ref");
}

// Runs reml_ref_enqueue: eagerly on the first use of a configuration (G and
// every buffer it touches), captured in a hipGraph on its second use and
// replayed from then on while the configuration is unchanged (batches of equal
// size, e.g. the gene-level permutations). The sequence holds no memset and no
// host copy.
static int reml_ref_device(eqtl_gpu_ctx* c, int G, double* delta) {
  const int n = c->n, p = c->p, TA = p * (p + 1) / 2;
  auto& r = c->rw;
  const size_t E2 = 2 * static_cast<size_t>(G);
  if (p > 22) CHK(ensure(r.W, sizeof(double) * n * E2));  // only the unfused path writes weights
  CHK(ensure(r.CA, sizeof(double) * TA * E2));
  CHK(ensure(r.CB, sizeof(double) * p * E2));
  CHK(ensure(r.S, sizeof(double) * kNumScal * E2));
  CHK(ensure(r.st, sizeof(RemlRefState) * G));
  CHK(ensure(r.dek, sizeof(int32_t) * E2));
  CHK(ensure(r.ded, sizeof(double) * E2));
  CHK(ensure(r.dF, sizeof(double) * E2));
  CHK(ensure(r.cnt, sizeof(int) * 2));
  CHK(ensure(r.list, sizeof(int) * G));
  int ng = 0;
  CHK(reml_grid_setup(c, G, &ng));
  const std::vector<const void*> key = {
      reinterpret_cast<const void*>(static_cast<intptr_t>(G)), reinterpret_cast<const void*>(static_cast<intptr_t>(n)),
      reinterpret_cast<const void*>(static_cast<intptr_t>(p)), delta, c->lam64.p, c->Yt64.p, c->Xt64.p, r.P.p,
      r.W.p, r.CA.p, r.CB.p, r.S.p, r.st.p, r.dek.p, r.ded.p, r.dF.p, r.cnt.p, r.list.p, r.Af.p, r.dgrid.p, r.Dg.p, r.logvg.p, r.XD.p, r.Y2.p,
      r.CAg.p, r.Ainv.p, r.logdet.p, r.okg.p, r.Bg.p, r.C0.p, r.F.p, r.best.p};
  if (!r.graph || r.graph_key != key) {
    if (r.graph) {
      HIPC(hipGraphExecDestroy(r.graph));
      r.graph = nullptr;
    }
    r.graph_key.clear();
    if (r.eager_key != key) {
      r.eager_key = key;
      return reml_ref_enqueue(c, G, ng, delta);
    }
    HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = reml_ref_enqueue(c, G, ng, delta);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != EQTL_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(EQTL_E_HIP, std::string("REML graph capture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&r.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      r.graph = nullptr;
      return fail(EQTL_E_HIP, std::string("REML graph instantiate: ") + hipGetErrorString(ei));
    }
    r.graph_key = key;
  }
  HIPC(hipGraphLaunch(r.graph, c->stream));
  return EQTL_OK;
}

// Per-gene delta for all genes, mode EQTL_DELTA_NEWTON: the reference's grid
// and bracket, then safeguarded Newton on ln delta with the analytic
// derivatives and step halving (host-driven rounds over the active genes).
// Mode EQTL_DELTA_REF runs on the device (reml_ref_device).
static int reml_newton(eqtl_gpu_ctx* c, std::vector<double>& delta) {
  const int G = c->G, n = c->n, p = c->p;
  delta.assign(G, 1.0);
  if (n - p <= 0) return EQTL_OK;
  CHK(reml_setup(c));
  std::vector<double> best;
  CHK(reml_grid(c, best));
  std::vector<int32_t> ek;
  std::vector<double> ed, F, Gd, H;
  // safeguarded Newton on theta = ln delta from the grid's best point, over
  // the search's whole range [1e-8, 1e6] (the reference's final Newton steps
  // may also leave the golden-section bracket: a likelihood still falling at
  // the bracket's end is followed to its minimum or to the range's end);
  // every accepted step lowers f
  std::vector<double> th(G), f(G), g(G), h(G), tn(G), tlo(G), thi(G);
  std::vector<int> back(G, 0), iters(G, 0);
  std::vector<char> state(G, 1);  // 1: propose a step, 2: trial pending, 0: done
  std::vector<int32_t> all(G);
  for (int k = 0; k < G; ++k) {
    all[k] = k;
    th[k] = std::log(best[k]);
    tlo[k] = std::log(1e-8);
    thi[k] = std::log(1e6);
    ed.push_back(std::exp(th[k]));
  }
  CHK(reml_eval(c, all, ed, true, f.data(), g.data(), h.data()));
  auto propose = [&](int k) {
    if (iters[k] >= 60) { state[k] = 0; return; }
    double step = (h[k] > 0) ? -g[k] / h[k] : (g[k] > 0 ? -0.5 : 0.5);
    step = std::max(-1.0, std::min(1.0, step));
    const double t = std::min(thi[k], std::max(tlo[k], th[k] + step));
    if (std::fabs(t - th[k]) < 1e-12) { state[k] = 0; return; }
    tn[k] = t;
    back[k] = 0;
    state[k] = 2;
  };
  for (int k = 0; k < G; ++k) propose(k);
  // Once a trial has been rejected, or when few genes are left (the rounds
  // then cost more than their evaluations), a pending trial is evaluated
  // together with its next kLadder - 1 halvings (formed by the same
  // recurrence) and the first one that does not raise f is taken: the
  // accepted points are those of one trial per round, in fewer rounds.
  constexpr int kLadder = 8, kFewGenes = 2048;
  std::vector<int32_t> first(G + 1), cnt(G);
  std::vector<double> xs;
  for (int round = 0; round < 60 * 32; ++round) {
    ek.clear();
    ed.clear();
    xs.clear();
    int active = 0;
    for (int k = 0; k < G; ++k) active += state[k] == 2;
    for (int k = 0; k < G; ++k) {
      first[k] = static_cast<int32_t>(ek.size());
      cnt[k] = 0;
      if (state[k] != 2) continue;
      const int len = (back[k] > 0 || active < kFewGenes) ? kLadder : 1;
      double x = tn[k];
      for (int b = back[k]; b <= 30 && cnt[k] < len; ++b) {
        ek.push_back(k);
        ed.push_back(std::exp(x));
        xs.push_back(x);
        ++cnt[k];
        x = th[k] + 0.5 * (x - th[k]);
      }
    }
    if (ek.empty()) break;
    const size_t A = ek.size();
    F.assign(A, 0.0); Gd.assign(A, 0.0); H.assign(A, 0.0);
    CHK(reml_eval(c, ek, ed, true, F.data(), Gd.data(), H.data()));
    for (int k = 0; k < G; ++k) {
      if (!cnt[k]) continue;
      int acc = -1;
      for (int i = 0; i < cnt[k]; ++i)
        if (!(F[first[k] + i] > f[k] + 1e-13 * std::fabs(f[k]))) {
          acc = i;
          break;
        }
      if (acc < 0) {
        back[k] += cnt[k];
        if (back[k] > 30) state[k] = 0;
        else tn[k] = th[k] + 0.5 * (xs[first[k] + cnt[k] - 1] - th[k]);
        continue;
      }
      const size_t q = static_cast<size_t>(first[k] + acc);
      const double moved = std::fabs(xs[q] - th[k]);
      th[k] = xs[q]; f[k] = F[q]; g[k] = Gd[q]; h[k] = H[q];
      ++iters[k];
      if (moved < 1e-10) { state[k] = 0; continue; }
      propose(k);
    }
  }
  for (int k = 0; k < G; ++k) delta[k] = std::exp(th[k]);
  return EQTL_OK;
}

int eqtl_gpu_eigh(eqtl_gpu_ctx* c, const double* K, int32_t n, double* Q_out, double* lam_out) {
  if (!c || !K || n <= 0) return fail(EQTL_E_ARG, "bad eigh arguments");
  if (c->n != n) return fail(EQTL_E_ARG, "GRM size differs from the analysis sample count");
  CHK(set_device(c));
  CHK(ensure_blas(c));
  CHK(ensure(c->Q64, sizeof(double) * n * n));
  CHK(ensure(c->lam64, sizeof(double) * n));
  DBuf E, info;
  CHK(ensure(E, sizeof(double) * n));
  CHK(ensure(info, sizeof(int)));
  HIPC(hipMemcpyAsync(c->Q64.p, K, sizeof(double) * n * n, hipMemcpyHostToDevice, c->stream));
  if (rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_lower, n, c->Q64.as<double>(), n,
                       c->lam64.as<double>(), E.as<double>(), info.as<int>()) != rocblas_status_success)
    return fail(EQTL_E_SOLVER, "rocsolver_dsyevd failed");
  int hinfo = 0;
  HIPC(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> lam(n);
  HIPC(hipMemcpyAsync(lam.data(), c->lam64.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  if (hinfo != 0) return fail(EQTL_E_SOLVER, "GRM eigen decomposition failed");
  for (auto& v : lam) v = std::max(v, 0.0);  // cwiseMax(0), model_lmm.cpp:111
  HIPC(hipMemcpyAsync(c->lam64.p, lam.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  // Eigenvector signs: each column is flipped so that sum_i u_i Q_ic > 0 with
  // the fixed weights u_i = 1 + frac(0.6180339887498949 (i + 1)) / 2. The
  // Wald tests do not depend on the signs; the LMM permutation streams do
  // (they shuffle eigen-coordinates), and this convention makes them a
  // function of K alone.
  std::vector<double> Qh(static_cast<size_t>(n) * n);
  HIPC(hipMemcpyAsync(Qh.data(), c->Q64.p, sizeof(double) * Qh.size(), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  for (int cc = 0; cc < n; ++cc) {
    double* q = Qh.data() + static_cast<size_t>(cc) * n;
    double sgn = 0.0;
    for (int i = 0; i < n; ++i) {
      const double f = 0.6180339887498949 * (i + 1);
      sgn += q[i] * (1.0 + 0.5 * (f - std::floor(f)));
    }
    if (sgn < 0)
      for (int i = 0; i < n; ++i) q[i] = -q[i];
  }
  HIPC(hipMemcpyAsync(c->Q64.p, Qh.data(), sizeof(double) * Qh.size(), hipMemcpyHostToDevice, c->stream));
  if (Q_out) std::memcpy(Q_out, Qh.data(), sizeof(double) * Qh.size());
  if (lam_out) std::memcpy(lam_out, lam.data(), sizeof(double) * n);
  HIPC(hipStreamSynchronize(c->stream));
  return upload_basis(c, n);
}

int eqtl_gpu_set_basis(eqtl_gpu_ctx* c, const double* Q, const double* lam, int32_t n) {
  if (!c || !Q || !lam || n <= 0) return fail(EQTL_E_ARG, "bad basis");
  if (c->n != n) return fail(EQTL_E_ARG, "basis size differs from the analysis sample count");
  CHK(set_device(c));
  CHK(ensure(c->Q64, sizeof(double) * n * n));
  CHK(ensure(c->lam64, sizeof(double) * n));
  std::vector<double> l(lam, lam + n);
  for (auto& v : l) v = std::max(v, 0.0);
  HIPC(hipMemcpyAsync(c->Q64.p, Q, sizeof(double) * n * n, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->lam64.p, l.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return upload_basis(c, n);
}

// ---- gene preparation -------------------------------------------------------
// Orthonormal basis Wx (n x p) of the covariates in c->X64 (prep_lm's
// (X^T X)^-1, model_lm.cpp:24-25, as Wx = X Lx^-T with Lx Lx^T = X^T X). The
// LM and LMM scans residualise each SNP on it before the contraction.
static int covariate_basis(eqtl_gpu_ctx* c) {
  const int n = c->n, p = c->p;
  c->have_wx = false;
  CHK(ensure(c->Wx, sizeof(double) * static_cast<size_t>(n) * p));
  CHK(ensure(c->zeros, sizeof(double) * n));
  HIPC(hipMemsetAsync(c->zeros.p, 0, sizeof(double) * n, c->stream));
  DBuf okb;
  CHK(ensure(okb, sizeof(int)));
  hipLaunchKernelGGL(k_lm_basis, dim3(1), dim3(256), 0, c->stream, c->X64.as<double>(), c->zeros.as<double>(), n, p,
                     c->Wx.as<double>(), okb.as<int>());
  CHK(check_launch("k_lm_basis"));
  int ok = 0;
  HIPC(hipMemcpyAsync(&ok, okb.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  if (!ok) return fail(EQTL_E_ARG, "covariate matrix X^T X is not positive semidefinite");
  c->have_wx = true;
  return EQTL_OK;
}

static int alloc_gene_side(eqtl_gpu_ctx* c, int G, int nq, int nq_pad) {
  if (nq_pad < nq) return fail(EQTL_E_ARG, "too many covariates for the Wald kernel");
  if (c->have_windows && static_cast<int>(c->h_gene_chrom.size()) != G) {
    // windows belong to a gene set of another size: drop them
    c->have_windows = false;
    c->h_gene_chrom.clear();
  }
  c->G = G;
  c->G_pad = static_cast<int>(round_up(G, 128));  // multiple of 8 groups of 16
  c->nq = nq;
  c->nq_pad = nq_pad;
  const size_t cols = static_cast<size_t>(c->G_pad / 16) * nq_pad * 16 * c->n_pad;
  CHK(ensure(c->Bf32, cols * sizeof(float)));
  CHK(ensure(c->Bbf, 2 * cols * sizeof(__bf16)));
  HIPC(hipMemsetAsync(c->Bf32.p, 0, cols * sizeof(float), c->stream));
  HIPC(hipMemsetAsync(c->Bbf.p, 0, 2 * cols * sizeof(__bf16), c->stream));
  CHK(ensure(c->gscal, sizeof(double) * c->G_pad));
  CHK(ensure(c->gvalid, sizeof(int) * c->G_pad));
  HIPC(hipMemsetAsync(c->gvalid.p, 0, sizeof(int) * c->G_pad, c->stream));
  CHK(ensure(c->delta, sizeof(double) * c->G_pad));
  return EQTL_OK;
}

// Per-gene REML delta, whitened bases, null residuals and column blocks for
// the c->G genes whose rotated phenotypes are in c->Yt64 (X~ in c->Xt64, gene
// side allocated). delta_in / delta_out / rss_out are host arrays (nullable).
static int lmm_prep_core(eqtl_gpu_ctx* c, int G, int mode, const double* delta_in, double* delta_out,
                         double* rss_out) {
  const int n = c->n, p = c->p;
  const double one = 1.0, zero = 0.0;
  if (mode == EQTL_DELTA_GIVEN) {
    HIPC(hipMemcpyAsync(c->delta.p, delta_in, sizeof(double) * G, hipMemcpyHostToDevice, c->stream));
  } else {
    Timed t(c, "reml");
    if (mode == EQTL_DELTA_REF && n - p > 0) {
      CHK(reml_setup(c));
      CHK(reml_ref_device(c, G, c->delta.as<double>()));
    } else {
      std::vector<double> dl;
      if (mode == EQTL_DELTA_REF) dl.assign(G, 1.0);
      else CHK(reml_newton(c, dl));
      HIPC(hipMemcpyAsync(c->delta.p, dl.data(), sizeof(double) * G, hipMemcpyHostToDevice, c->stream));
      HIPC(hipStreamSynchronize(c->stream));
    }
    CHK(t.done());
  }
  {
    Timed t(c, "basis");
    hipLaunchKernelGGL(k_lmm_basis, dim3(G), dim3(256), 0, c->stream, c->Xt64.as<double>(), c->Yt64.as<double>(),
                       c->lam64.as<double>(), n, c->n_pad, p, G, c->delta.as<double>(), gene_cols(c),
                       c->gscal.as<double>(), c->gvalid.as<int>(), c->U64.as<double>(), c->Lc.as<double>());
    CHK(check_launch("k_lmm_basis"));
    // sample-space residual vectors for the near-null FP64 refinement: V = Q U
    if (rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, n, G, n, &one, c->Q64.as<double>(), n,
                      c->U64.as<double>(), n, &zero, c->V64.as<double>(), n) != rocblas_status_success)
      return fail(EQTL_E_HIP, "rocblas_dgemm failed");
    CHK(t.done());
  }
  std::vector<int> valid(G);
  HIPC(hipMemcpyAsync(valid.data(), c->gvalid.p, sizeof(int) * G, hipMemcpyDeviceToHost, c->stream));
  if (delta_out) HIPC(hipMemcpyAsync(delta_out, c->delta.p, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
  if (rss_out) HIPC(hipMemcpyAsync(rss_out, c->gscal.p, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  for (int k = 0; k < G; ++k)
    if (!valid[k]) return fail(EQTL_E_ARG, "whitened covariate matrix is not positive semidefinite (gene " + std::to_string(k) + ")");
  c->prepped_model = EQTL_MODEL_LMM;
  return EQTL_OK;
}

// Copy kind for a caller buffer that may live in host or device memory.
static hipMemcpyKind copy_in_kind(const void* ptr) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, ptr) == hipSuccess && at.type == hipMemoryTypeDevice) return hipMemcpyDeviceToDevice;
  (void)hipGetLastError();  // a host pointer may leave an error code behind
  return hipMemcpyHostToDevice;
}

int eqtl_gpu_lmm_prep(eqtl_gpu_ctx* c, const double* Y, int32_t G, const double* X, int32_t p,
                      int32_t mode, const double* delta_in, double* delta_out, double* rss_out) {
  if (!c || !Y || !X || G <= 0 || p <= 0) return fail(EQTL_E_ARG, "bad lmm_prep arguments");
  if (p > kPMax) return fail(EQTL_E_ARG, "more than 32 covariate columns are not supported");
  if (!c->have_basis) return fail(EQTL_E_STATE, "no eigenbasis: call eqtl_gpu_eigh or eqtl_gpu_set_basis");
  if (mode == EQTL_DELTA_GIVEN && !delta_in) return fail(EQTL_E_ARG, "delta_in required");
  const int n = c->n;
  CHK(set_device(c));
  CHK(ensure_blas(c));
  c->p = p;
  CHK(alloc_gene_side(c, G, p + 2, lmm_nq_pad(p + 2)));
  CHK(ensure(c->Y64, sizeof(double) * static_cast<size_t>(n) * G));
  CHK(ensure(c->X64, sizeof(double) * static_cast<size_t>(n) * p));
  CHK(ensure(c->Yt64, sizeof(double) * static_cast<size_t>(n) * G));
  CHK(ensure(c->Xt64, sizeof(double) * static_cast<size_t>(n) * p));
  CHK(ensure(c->U64, sizeof(double) * static_cast<size_t>(n) * G));
  CHK(ensure(c->V64, sizeof(double) * static_cast<size_t>(n) * G));
  CHK(ensure(c->Lc, sizeof(double) * static_cast<size_t>(p) * p * G));
  // host or device memory
  HIPC(hipMemcpyAsync(c->Y64.p, Y, sizeof(double) * static_cast<size_t>(n) * G, copy_in_kind(Y), c->stream));
  HIPC(hipMemcpyAsync(c->X64.p, X, sizeof(double) * static_cast<size_t>(n) * p, copy_in_kind(X), c->stream));
  CHK(covariate_basis(c));
  // K4: Y~ = Q^T Y, X~ = Q^T X once for all genes (FP64 library GEMM)
  const double one = 1.0, zero = 0.0;
  {
    Timed t(c, "rotate_yx");
    if (rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, n, G, n, &one,
                      c->Q64.as<double>(), n, c->Y64.as<double>(), n, &zero, c->Yt64.as<double>(), n) !=
            rocblas_status_success ||
        rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, n, p, n, &one,
                      c->Q64.as<double>(), n, c->X64.as<double>(), n, &zero, c->Xt64.as<double>(), n) !=
            rocblas_status_success)
      return fail(EQTL_E_HIP, "rocblas_dgemm failed");
    CHK(t.done());
  }
  return lmm_prep_core(c, G, mode, delta_in, delta_out, rss_out);
}

int eqtl_gpu_reml_grad_hess(eqtl_gpu_ctx* c, const double* log_delta, int32_t G, double* f, double* g,
                            double* h) {
  if (!c || !log_delta || !f || !g || !h) return fail(EQTL_E_ARG, "bad reml_grad_hess arguments");
  if (c->prepped_model != EQTL_MODEL_LMM || G != c->G) return fail(EQTL_E_STATE, "call eqtl_gpu_lmm_prep first");
  CHK(set_device(c));
  CHK(ensure_blas(c));
  CHK(reml_setup(c));
  std::vector<int32_t> ek(G);
  std::vector<double> ed(G);
  for (int k = 0; k < G; ++k) {
    ek[k] = k;
    ed[k] = std::exp(log_delta[k]);
  }
  return reml_eval(c, ek, ed, true, f, g, h);
}

int eqtl_gpu_lm_prep(eqtl_gpu_ctx* c, const double* Y, int32_t G, const double* X, int32_t p, double* yty_out) {
  if (!c || !Y || !X || G <= 0 || p <= 0) return fail(EQTL_E_ARG, "bad lm_prep arguments");
  if (p > kPMax) return fail(EQTL_E_ARG, "more than 32 covariate columns are not supported");
  if (c->n <= 0) return fail(EQTL_E_STATE, "set_samples first");
  const int n = c->n;
  CHK(set_device(c));
  c->p = p;
  CHK(alloc_gene_side(c, G, 1, 1));
  CHK(ensure(c->Y64, sizeof(double) * static_cast<size_t>(n) * G));
  CHK(ensure(c->X64, sizeof(double) * static_cast<size_t>(n) * p));
  CHK(ensure(c->U64, sizeof(double) * static_cast<size_t>(n) * G));
  HIPC(hipMemcpyAsync(c->Y64.p, Y, sizeof(double) * static_cast<size_t>(n) * G, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->X64.p, X, sizeof(double) * static_cast<size_t>(n) * p, hipMemcpyHostToDevice, c->stream));
  CHK(covariate_basis(c));
  hipLaunchKernelGGL(k_lm_gene, dim3(G), dim3(256), 0, c->stream, c->Y64.as<double>(), c->Wx.as<double>(), n,
                     c->n_pad, p, G, gene_cols(c), c->gscal.as<double>(), c->gvalid.as<int>(),
                     c->U64.as<double>());
  CHK(check_launch("k_lm_gene"));
  if (yty_out) HIPC(hipMemcpyAsync(yty_out, c->gscal.p, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  c->prepped_model = EQTL_MODEL_LM;
  return EQTL_OK;
}

int eqtl_gpu_glm_prep(eqtl_gpu_ctx* c, const double* Y, int32_t G, const double* X, int32_t p, double* phi_out,
                      int32_t* conv_out) {
  if (!c || !Y || !X || G <= 0 || p <= 0) return fail(EQTL_E_ARG, "bad glm_prep arguments");
  if (p > kPMax) return fail(EQTL_E_ARG, "more than 32 covariate columns are not supported");
  if (c->n <= 0) return fail(EQTL_E_STATE, "set_samples first");
  const int n = c->n;
  for (size_t i = 0; i < static_cast<size_t>(n) * G; ++i)  // looks_like_counts, model_lm.cpp:8-17
    if (!(Y[i] >= 0) || std::fabs(Y[i] - std::floor(Y[i])) > 1e-8)
      return fail(EQTL_E_ARG, "glm (NB) requires non-negative integer counts (gene " + std::to_string(i / n) + ")");
  CHK(set_device(c));
  c->p = p;
  CHK(alloc_gene_side(c, G, p + 2, lmm_nq_pad(p + 2)));
  CHK(ensure(c->Y64, sizeof(double) * static_cast<size_t>(n) * G));
  CHK(ensure(c->X64, sizeof(double) * static_cast<size_t>(n) * p));
  CHK(ensure(c->Yt64, sizeof(double) * static_cast<size_t>(n) * G));  // the fit's mean (scratch)
  CHK(ensure(c->U64, sizeof(double) * static_cast<size_t>(n) * G));
  CHK(ensure(c->V64, sizeof(double) * static_cast<size_t>(n) * G));
  CHK(ensure(c->Lc, sizeof(double) * static_cast<size_t>(p) * p * G));
  HIPC(hipMemcpyAsync(c->Y64.p, Y, sizeof(double) * static_cast<size_t>(n) * G, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->X64.p, X, sizeof(double) * static_cast<size_t>(n) * p, hipMemcpyHostToDevice, c->stream));
  {
    Timed t(c, "glm_prep");
    hipLaunchKernelGGL(k_glm_null, dim3(G), dim3(256), 0, c->stream, c->X64.as<double>(), c->Y64.as<double>(), n,
                       c->n_pad, p, G, gene_cols(c), c->Yt64.as<double>(), c->delta.as<double>(),
                       c->gvalid.as<int>(), c->gscal.as<double>(), c->U64.as<double>(), c->V64.as<double>(),
                       c->Lc.as<double>());
    CHK(check_launch("k_glm_null"));
    CHK(t.done());
  }
  if (phi_out) HIPC(hipMemcpyAsync(phi_out, c->delta.p, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
  if (conv_out) HIPC(hipMemcpyAsync(conv_out, c->gvalid.p, sizeof(int32_t) * G, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  c->prepped_model = EQTL_MODEL_GLM;
  return EQTL_OK;
}

int eqtl_gpu_set_gene_windows(eqtl_gpu_ctx* c, const int32_t* chrom, const int64_t* cs, const int64_t* ce,
                              int32_t G) {
  if (!c || !chrom || !cs || !ce || G <= 0) return fail(EQTL_E_ARG, "bad gene windows");
  CHK(set_device(c));
  CHK(ensure(c->gene_chrom, sizeof(int32_t) * G));
  CHK(ensure(c->gene_cs, sizeof(int64_t) * G));
  CHK(ensure(c->gene_ce, sizeof(int64_t) * G));
  HIPC(hipMemcpyAsync(c->gene_chrom.p, chrom, sizeof(int32_t) * G, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->gene_cs.p, cs, sizeof(int64_t) * G, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->gene_ce.p, ce, sizeof(int64_t) * G, hipMemcpyHostToDevice, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  c->have_windows = true;
  c->h_gene_chrom.assign(chrom, chrom + G);
  c->h_gene_cs.assign(cs, cs + G);
  c->h_gene_ce.assign(ce, ce + G);
  return EQTL_OK;
}

// ---- scans ------------------------------------------------------------------
// Per-gene global SNP index range holding the gene's window (in_cis_window,
// scan_common.hpp:58-69): SNPs on the gene's chromosome with cstart <= pos <=
// cend. The range is [min index, max index] of those SNPs, so a bim whose
// chromosomes are not contiguous still gets a (wider) correct range.
static int cis_windows(eqtl_gpu_ctx* c) {
  const int64_t m = c->m_annot;
  if (m <= 0 || static_cast<int64_t>(c->h_snp_pos.size()) != m)
    return fail(EQTL_E_STATE, "cis scope needs the SNP annotation");
  std::map<int32_t, std::vector<std::pair<int64_t, int64_t>>> by_chrom;  // chrom -> (pos, index)
  for (int64_t j = 0; j < m; ++j) by_chrom[c->h_snp_chrom[j]].emplace_back(c->h_snp_pos[j], j);
  for (auto& kv : by_chrom) std::sort(kv.second.begin(), kv.second.end());
  const int G = c->G;
  c->win_lo.assign(G, 0);
  c->win_hi.assign(G, 0);
  for (int k = 0; k < G; ++k) {
    if (c->h_gene_chrom[k] < 0) continue;
    auto it = by_chrom.find(c->h_gene_chrom[k]);
    if (it == by_chrom.end()) continue;
    const auto& v = it->second;
    auto a = std::lower_bound(v.begin(), v.end(), std::make_pair(c->h_gene_cs[k], INT64_MIN));
    auto b = std::upper_bound(v.begin(), v.end(), std::make_pair(c->h_gene_ce[k], INT64_MAX));
    int64_t lo = INT64_MAX, hi = -1;
    for (auto q = a; q < b; ++q) {
      lo = std::min(lo, q->second);
      hi = std::max(hi, q->second);
    }
    if (hi >= 0) {
      c->win_lo[k] = lo;
      c->win_hi[k] = hi + 1;
    }
  }
  return EQTL_OK;
}

static int lr_build(eqtl_gpu_ctx* c);
static int scan_begin_impl(eqtl_gpu_ctx* c, int32_t model, int32_t prec, int32_t scope, double pthr,
                           bool windows_ready);

// t-tail table for the residual df n - p - 1 of the prepared genes (shared by
// every test of a scan).
static int ensure_tail(eqtl_gpu_ctx* c) {
  const double df = static_cast<double>(c->n - c->p - 1);
  if (df > 0 && c->tail_df != df) {
    std::vector<TTail> T(1);
    const double lbeta = std::lgamma(0.5 * df) + std::lgamma(0.5) - std::lgamma(0.5 * df + 0.5);
    tail_build(T[0], df, lbeta);
    CHK(ensure(c->tail, sizeof(TTail)));
    HIPC(hipMemcpyAsync(c->tail.p, T.data(), sizeof(TTail), hipMemcpyHostToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    c->tail_df = df;
  } else if (df <= 0) {
    CHK(ensure(c->tail, sizeof(TTail)));
    HIPC(hipMemsetAsync(c->tail.p, 0, sizeof(TTail), c->stream));
    c->tail_df = df;
  }
  return EQTL_OK;
}

int eqtl_gpu_scan_begin(eqtl_gpu_ctx* c, int32_t model, int32_t prec, int32_t scope, double pthr) {
  return scan_begin_impl(c, model, prec, scope, pthr, false);
}

// windows_ready: c->win_lo / win_hi already hold the cis SNP ranges.
static int scan_begin_impl(eqtl_gpu_ctx* c, int32_t model, int32_t prec, int32_t scope, double pthr,
                           bool windows_ready) {
  if (!c) return fail(EQTL_E_ARG, "null ctx");
  if (model != EQTL_MODEL_LM && model != EQTL_MODEL_LMM && model != EQTL_MODEL_GLM) return fail(EQTL_E_ARG, "bad model");
  if (prec != EQTL_PREC_FP32 && prec != EQTL_PREC_BF16X3) return fail(EQTL_E_ARG, "bad precision");
  if (scope != EQTL_SCOPE_TRANS && scope != EQTL_SCOPE_CIS) return fail(EQTL_E_ARG, "bad scope");
  if (c->prepped_model != model) return fail(EQTL_E_STATE, "gene side not prepared for this model");
  if (c->have_windows && c->m_annot <= 0) return fail(EQTL_E_STATE, "gene windows need the SNP annotation");
  if (scope == EQTL_SCOPE_CIS && !c->have_windows) return fail(EQTL_E_STATE, "cis scope needs gene windows");
  if (c->have_windows && !windows_ready && static_cast<int>(c->h_gene_chrom.size()) != c->G)
    return fail(EQTL_E_STATE, "gene windows were set for a different gene count");
  CHK(set_device(c));
  c->scan_model = model;
  c->scan_prec = prec;
  c->scan_scope = scope;
  c->pthr = pthr;
  if (scope == EQTL_SCOPE_CIS && !windows_ready) CHK(cis_windows(c));
  CHK(ensure(c->state, sizeof(GeneAcc) * c->G_pad));
  hipLaunchKernelGGL(k_state_init, dim3((c->G_pad + 255) / 256), dim3(256), 0, c->stream, c->state.as<GeneAcc>(),
                     c->G_pad);
  CHK(check_launch("k_state_init"));
  CHK(ensure(c->hit_count, 2 * sizeof(unsigned long long)));
  CHK(ensure_tail(c));
  c->lr_on = false;
  if (model == EQTL_MODEL_LMM && prec == EQTL_PREC_BF16X3 && scope == EQTL_SCOPE_TRANS) CHK(lr_build(c));
  c->scanning = true;
  return EQTL_OK;
}

// ---- low-rank trans LMM scan (k_lr_num, k_wald_lr) -----------------------------
// Widths of ln-delta segments that 32, 64, 96 and 128 Chebyshev nodes
// interpolate to ~1e-12 (the whitened covariate projection is analytic in
// ln delta within pi of the real axis; measured on synthetic spectra).
constexpr double kLrWidth[4] = {6.5, 13.0, 23.0, 33.5};

// Builds the plan and operands of k_lr_num / k_wald_lr for the prepared LMM genes: genes
// sorted by ln delta into super-groups of 64, segments of ln delta (a
// super-group's genes inside one segment), per segment R nodes, per node the
// whitened projection Z (k_lr_node_basis) mapped to sample space (Q Z) and
// split for the rotation kernel, the genes' interpolation weights and their
// (d, s o r) tiles. Leaves lr_on false (the eight-wave kernel then runs) when
// a node basis fails or the deltas span more than 33.5 in ln delta.
static int lr_build(eqtl_gpu_ctx* c) {
  c->lr_on = false;
  const int G = c->G, n = c->n, np = c->n_pad, p = c->p;
  if (G <= 0 || n - p <= 0 || !c->have_basis) return EQTL_OK;
  std::vector<double> dl(G);
  std::vector<int> valid(G);
  HIPC(hipMemcpyAsync(dl.data(), c->delta.p, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipMemcpyAsync(valid.data(), c->gvalid.p, sizeof(int) * G, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  std::vector<double> u(G);
  for (int k = 0; k < G; ++k) {
    if (!(dl[k] > 0) || !std::isfinite(dl[k])) return EQTL_OK;
    u[k] = std::log(dl[k]);
  }
  std::vector<int> order(G);
  for (int k = 0; k < G; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (valid[x] != valid[y]) return valid[x] > valid[y];
    return valid[x] ? u[x] < u[y] : false;
  });
  const int n_sg = (G + 63) / 64, nslot = n_sg * 64;
  std::vector<int32_t> slot_gene(nslot, -1), slot_of(G);
  for (int s2 = 0; s2 < G; ++s2) {
    slot_gene[s2] = order[s2];
    slot_of[order[s2]] = s2;
  }
  // segments, greedily over the super-groups in ln-delta order
  struct Seg { double lo, hi; int R, base; };
  std::vector<Seg> segs;
  std::vector<int> sg_seg(n_sg, 0);
  for (int g = 0; g < n_sg; ++g) {
    double lo = 0, hi = 0;
    bool any = false;
    for (int s2 = 64 * g; s2 < std::min(G, 64 * g + 64); ++s2) {
      const int k = slot_gene[s2];
      if (!valid[k]) continue;
      lo = any ? std::min(lo, u[k]) : u[k];
      hi = any ? std::max(hi, u[k]) : u[k];
      any = true;
    }
    if (!any) {
      if (!segs.empty()) lo = hi = segs.back().lo;
    }
    if (!segs.empty() && std::max(hi, segs.back().hi) - std::min(lo, segs.back().lo) <= kLrWidth[0] &&
        segs.back().hi - segs.back().lo <= kLrWidth[0]) {
      segs.back().lo = std::min(lo, segs.back().lo);
      segs.back().hi = std::max(hi, segs.back().hi);
    } else {
      segs.push_back(Seg{lo, hi, 0, 0});
    }
    sg_seg[g] = static_cast<int>(segs.size()) - 1;
  }
  int ncols = 0;
  for (Seg& sgm : segs) {
    if (sgm.hi - sgm.lo < 0.5) {
      const double mid = 0.5 * (sgm.lo + sgm.hi);
      sgm.lo = mid - 0.25;
      sgm.hi = mid + 0.25;
    }
    int kx = 0;
    while (kx < 4 && sgm.hi - sgm.lo > kLrWidth[kx]) ++kx;
    if (kx == 4) return EQTL_OK;  // wider than 128 nodes interpolate: the eight-wave kernel runs
    sgm.R = 32 * (kx + 1);
    sgm.base = ncols;
    ncols += sgm.R * p;
  }
  int nnodes_total = 0;
  for (const Seg& sgm : segs) nnodes_total += sgm.R;
  const int64_t ldT = round_up(ncols + nnodes_total, 32);
  // nodes (Chebyshev points of the first kind) and their columns
  std::vector<double> nd;
  std::vector<int32_t> col0, stride;
  for (const Seg& sgm : segs)
    for (int r = 0; r < sgm.R; ++r) {
      const double x = std::cos((2.0 * r + 1.0) * M_PI / (2.0 * sgm.R));
      nd.push_back(std::exp(0.5 * (sgm.lo + sgm.hi) + 0.5 * (sgm.hi - sgm.lo) * x));
      col0.push_back(sgm.base + r);
      stride.push_back(sgm.R);
    }
  const int nnodes = static_cast<int>(nd.size());
  // interpolation weights per slot (barycentric form), split hi / mid / lo
  const size_t ncoef = static_cast<size_t>(nslot / 16) * 4 * 1536;
  std::vector<__bf16> coef(ncoef, static_cast<__bf16>(0.f));
  std::vector<int32_t> sgc(n_sg), sgR(n_sg), sgg(n_sg), gbase(segs.size());
  for (size_t q = 0, acc = 0; q < segs.size(); ++q) {
    gbase[q] = static_cast<int32_t>(ncols + acc);  // the segment's node sums of squares follow all projections
    acc += segs[q].R;
  }
  for (int g = 0; g < n_sg; ++g) {
    sgc[g] = segs[sg_seg[g]].base;
    sgg[g] = gbase[sg_seg[g]];
    sgR[g] = segs[sg_seg[g]].R;
  }
  // Chebyshev nodes of the first kind and their barycentric weights, per segment
  std::vector<std::vector<double>> seg_ur(segs.size()), seg_wr(segs.size());
  for (size_t q = 0; q < segs.size(); ++q) {
    const Seg& sgm = segs[q];
    seg_ur[q].resize(sgm.R);
    seg_wr[q].resize(sgm.R);
    for (int r = 0; r < sgm.R; ++r) {
      const double th = (2.0 * r + 1.0) * M_PI / (2.0 * sgm.R);
      seg_ur[q][r] = 0.5 * (sgm.lo + sgm.hi) + 0.5 * (sgm.hi - sgm.lo) * std::cos(th);
      seg_wr[q][r] = ((r & 1) ? -1.0 : 1.0) * std::sin(th);
    }
  }
  // the genes' weights, slots split over host threads (disjoint outputs)
  auto weights = [&](int s_lo, int s_hi) {
    std::vector<double> ell;
    for (int s2 = s_lo; s2 < s_hi; ++s2) {
      const int k = slot_gene[s2];
      if (!valid[k]) continue;
      const int q = sg_seg[s2 / 64];
      const int R = segs[q].R;
      const std::vector<double>& ur = seg_ur[q];
      const std::vector<double>& wr = seg_wr[q];
      ell.assign(R, 0.0);
      int hit = -1;
      for (int r = 0; r < R; ++r)
        if (u[k] == ur[r]) hit = r;
      if (hit >= 0) {
        ell[hit] = 1.0;
      } else {
        double den = 0.0;
        for (int r = 0; r < R; ++r) den += wr[r] / (u[k] - ur[r]);
        for (int r = 0; r < R; ++r) ell[r] = wr[r] / (u[k] - ur[r]) / den;
      }
      for (int r = 0; r < R; ++r) {
        const __bf16 h = static_cast<__bf16>(static_cast<float>(ell[r]));
        const double r1 = ell[r] - static_cast<double>(static_cast<float>(h));
        const __bf16 m = static_cast<__bf16>(static_cast<float>(r1));
        const double r2 = r1 - static_cast<double>(static_cast<float>(m));
        coef[lr_coef_index(s2, r, 0)] = h;
        coef[lr_coef_index(s2, r, 1)] = m;
        coef[lr_coef_index(s2, r, 2)] = static_cast<__bf16>(static_cast<float>(r2));
      }
    }
  };
  host_parallel(G, weights);
  // node bases Z (np x ncols), sample-space Q Z, split
  DBuf& Z = c->lr_Z;
  DBuf& QZ = c->lr_QZ;
  DBuf& nodes = c->lr_nodes;
  DBuf& okb = c->lr_okb;
  const size_t zc = static_cast<size_t>(np) * ncols;
  CHK(ensure(Z, sizeof(double) * zc));
  CHK(ensure(QZ, sizeof(double) * zc));
  CHK(ensure(nodes, sizeof(double) * nnodes + 2 * sizeof(int32_t) * nnodes + 64));
  CHK(ensure(okb, sizeof(int) * nnodes));
  double* nd_d = nodes.as<double>();
  int32_t* c0_d = reinterpret_cast<int32_t*>(nd_d + nnodes);
  int32_t* st_d = c0_d + nnodes;
  HIPC(hipMemcpyAsync(nd_d, nd.data(), sizeof(double) * nnodes, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c0_d, col0.data(), sizeof(int32_t) * nnodes, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(st_d, stride.data(), sizeof(int32_t) * nnodes, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_lr_node_basis, dim3(nnodes), dim3(256), 0, c->stream, c->Xt64.as<double>(),
                     c->zeros.as<double>(), c->lam64.as<double>(), n, np, p, nd_d, c0_d, st_d, Z.as<double>(),
                     okb.as<int>());
  CHK(check_launch("k_lr_node_basis"));
  HIPC(hipMemsetAsync(QZ.p, 0, sizeof(double) * zc, c->stream));
  const double one = 1.0, zero = 0.0;
  if (rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, n, ncols, n, &one, c->Q64.as<double>(), n,
                    Z.as<double>(), np, &zero, QZ.as<double>(), np) != rocblas_status_success)
    return fail(EQTL_E_HIP, "rocblas_dgemm failed");
  const size_t zb = static_cast<size_t>(rot_b_coords(ncols)) * np;
  CHK(ensure(c->lr_QZhi, sizeof(__bf16) * zb));
  CHK(ensure(c->lr_QZlo, sizeof(__bf16) * zb));
  hipLaunchKernelGGL(k_split_rot_b, dim3(1024), dim3(256), 0, c->stream, QZ.as<double>(), np, ncols,
                     c->lr_QZhi.as<__bf16>(), c->lr_QZlo.as<__bf16>());
  CHK(check_launch("k_split_rot_b"));
  // node weights d(delta_r, lam_c) / d(delta_r, lbar) per rotated coordinate,
  // coordinate-major per node (the B operand of the squared-row product),
  // split; per slot the scale d(delta_k, lbar) of the interpolated gg
  std::vector<double> gsc(nslot, 0.0);
  {
    std::vector<double> lam(n);
    HIPC(hipMemcpyAsync(lam.data(), c->lam64.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    double lbar = 0.0;
    for (int i = 0; i < n; ++i) lbar += lam[i];
    lbar = lbar > 0 ? lbar / n : 1.0;
    for (int s2 = 0; s2 < G; ++s2) gsc[s2] = 1.0 / (dl[slot_gene[s2]] * lbar + 1.0);
    // fragment-major B order of the rotation kernel (zero past the nodes and past n)
    const size_t dn = static_cast<size_t>(rot_b_coords(nnodes_total)) * np;
    std::vector<__bf16> dh(dn, static_cast<__bf16>(0.f)), dlo(dn, static_cast<__bf16>(0.f));
    host_parallel(nnodes, [&](int k_lo, int k_hi) {
      for (int k = k_lo; k < k_hi; ++k)
        for (int i = 0; i < n; ++i) {
          double v = nd[k] * lam[i] + 1.0;
          if (v < 1e-12) v = 1e-12;
          const float f = static_cast<float>((nd[k] * lbar + 1.0) / v);
          const __bf16 h = static_cast<__bf16>(f);
          const int64_t o = rot_b_index(k, i, np / 32);
          dh[o] = h;
          dlo[o] = static_cast<__bf16>(f - static_cast<float>(h));
        }
    }, 16);
    CHK(ensure(c->lr_Dhi, sizeof(__bf16) * dn));
    CHK(ensure(c->lr_Dlo, sizeof(__bf16) * dn));
    HIPC(hipMemcpyAsync(c->lr_Dhi.p, dh.data(), sizeof(__bf16) * dn, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(c->lr_Dlo.p, dlo.data(), sizeof(__bf16) * dn, hipMemcpyHostToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
  }
  // gene tiles by slot
  const int nks = np / 32;
  const size_t gbn = static_cast<size_t>(n_sg) * nks * (kLrTileBytes / sizeof(__bf16));
  CHK(ensure(c->lr_gb, sizeof(__bf16) * gbn));
  HIPC(hipMemsetAsync(c->lr_gb.p, 0, sizeof(__bf16) * gbn, c->stream));
  DBuf sof;
  CHK(ensure(sof, sizeof(int32_t) * G));
  HIPC(hipMemcpyAsync(sof.p, slot_of.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_lr_gene_tiles, dim3(G), dim3(256), 0, c->stream, c->U64.as<double>(), sof.as<int32_t>(), n, nks,
                     G, c->lr_gb.as<__bf16>());
  CHK(check_launch("k_lr_gene_tiles"));
  CHK(ensure(c->lr_slot_gene, sizeof(int32_t) * nslot));
  CHK(ensure(c->lr_sg_col, sizeof(int32_t) * 2 * n_sg));
  CHK(ensure(c->lr_sg_R, sizeof(int32_t) * n_sg));
  CHK(ensure(c->lr_coef, sizeof(__bf16) * ncoef));
  CHK(ensure(c->lr_gscale, sizeof(double) * nslot));
  HIPC(hipMemcpyAsync(c->lr_gscale.p, gsc.data(), sizeof(double) * nslot, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->lr_slot_gene.p, slot_gene.data(), sizeof(int32_t) * nslot, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->lr_sg_col.p, sgc.data(), sizeof(int32_t) * n_sg, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->lr_sg_col.as<int32_t>() + n_sg, sgg.data(), sizeof(int32_t) * n_sg, hipMemcpyHostToDevice,
                      c->stream));
  HIPC(hipMemcpyAsync(c->lr_sg_R.p, sgR.data(), sizeof(int32_t) * n_sg, hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(c->lr_coef.p, coef.data(), sizeof(__bf16) * ncoef, hipMemcpyHostToDevice, c->stream));
  std::vector<int> ok(nnodes);
  HIPC(hipMemcpyAsync(ok.data(), okb.p, sizeof(int) * nnodes, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  for (int v : ok)
    if (!v) return EQTL_OK;
  c->lr_n_sg = n_sg;
  c->lr_ncols = ncols;
  c->lr_nnodes = nnodes_total;
  c->lr_nseg = static_cast<int>(segs.size());
  {
    double sum = 0.0;
    int cnt = 0;
    for (int s2 = 0; s2 < G; ++s2)
      if (valid[slot_gene[s2]]) {
        sum += segs[sg_seg[s2 / 64]].R;
        ++cnt;
      }
    c->lr_mean_nodes = cnt ? sum / cnt : 0.0;
  }
  c->lr_ldT = ldT;
  c->lr_on = true;
  return EQTL_OK;
}

// The interpolated form: k_lr_num (num, over the super-group pairs' grid)
// and k_wald_lr (the interpolation: gg and S, over the super-groups' grid)
// fill the fold; then k_lr_tests over the super-groups' grid.
static int lr_set_attrs() {
  static bool attr_set = false;
  if (attr_set) return EQTL_OK;
  HIPC(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_lr_num), hipFuncAttributeMaxDynamicSharedMemorySize,
                           static_cast<int>(kLrNumLdsBytes)));
  HIPC(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_wald_lr), hipFuncAttributeMaxDynamicSharedMemorySize,
                           static_cast<int>(kLrLdsBytes)));
  HIPC(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_lr_tests), hipFuncAttributeMaxDynamicSharedMemorySize,
                           static_cast<int>(kLrTestsLdsBytes)));
  attr_set = true;
  return EQTL_OK;
}

int launch_lr_num(eqtl_gpu_ctx* c, const LrArgs& la, dim3 grid) {
  CHK(lr_set_attrs());
  hipLaunchKernelGGL(k_lr_num, grid, dim3(kW8Threads), kLrNumLdsBytes, c->stream, la);
  return check_launch("k_lr_num");
}

int launch_wald_lr(eqtl_gpu_ctx* c, const LrArgs& la, dim3 grid) {
  CHK(lr_set_attrs());
  hipLaunchKernelGGL(k_wald_lr, grid, dim3(kW8Threads), kLrLdsBytes, c->stream, la);
  return check_launch("k_wald_lr");
}

int launch_lr_tests(eqtl_gpu_ctx* c, const LrArgs& la, dim3 grid) {
  CHK(lr_set_attrs());
  hipLaunchKernelGGL(k_lr_tests, grid, dim3(kW8Threads), kLrTestsLdsBytes, c->stream, la);
  return check_launch("k_lr_tests");
}

// Cis work items of one scan block over [snp0, snp0 + m): per gene group the
// 128-row tiles holding a window SNP of its genes, cut into items of at most
// tiles[class] tiles (class of the group's span, kCisClassMax); only groups
// of class only_class when it is >= 0. items: {group, tile_lo, tile_hi} with
// a group's items consecutive (first[g] .. first[g + 1]); order: the queue
// order, items with more tiles first.
struct CisItems {
  std::vector<int32_t> items, first, order;
  int classes_present[kCisClasses] = {0, 0, 0};
};
static int cis_class(int span_tiles) {
  int k = 0;
  while (k < kCisClasses - 1 && span_tiles > kCisClassMax[k]) ++k;
  return k;
}
static void cis_items(const eqtl_gpu_ctx* c, int64_t snp0, int64_t m, int n_groups, const int* tiles, int only_class,
                      CisItems& out) {
  out.items.clear();
  out.first.assign(n_groups + 1, 0);
  for (int q = 0; q < kCisClasses; ++q) out.classes_present[q] = 0;
  for (int g = 0; g < n_groups; ++g) {
    out.first[g] = static_cast<int32_t>(out.items.size() / 3);
    int64_t lo = INT64_MAX, hi = -1;
    for (int k = 16 * g; k < std::min(16 * g + 16, c->G); ++k)
      if (c->win_lo[k] < c->win_hi[k]) {
        lo = std::min(lo, c->win_lo[k]);
        hi = std::max(hi, c->win_hi[k]);
      }
    const int64_t l = std::max(lo, snp0) - snp0, h = std::min(hi, snp0 + m) - snp0;
    if (l >= h) continue;
    const int t0 = static_cast<int>(l / kTileRows), t1 = static_cast<int>((h + kTileRows - 1) / kTileRows);
    const int cls = cis_class(t1 - t0);
    out.classes_present[cls] = 1;
    if (only_class >= 0 && cls != only_class) continue;
    const int step = tiles[cls];
    for (int t = t0; t < t1; t += step) {
      out.items.push_back(g);
      out.items.push_back(t);
      out.items.push_back(std::min(t + step, t1));
    }
  }
  out.first[n_groups] = static_cast<int32_t>(out.items.size() / 3);
  const int n = out.first[n_groups];
  out.order.resize(n);
  for (int i = 0; i < n; ++i) out.order[i] = i;
  std::stable_sort(out.order.begin(), out.order.end(), [&](int x, int y) {
    return out.items[3 * x + 2] - out.items[3 * x + 1] > out.items[3 * y + 2] - out.items[3 * y + 1];
  });
}

static int upload_items(eqtl_gpu_ctx* c, const CisItems& ci) {
  const size_t n = ci.order.size();
  CHK(ensure(c->work, sizeof(int32_t) * std::max<size_t>(ci.items.size(), 3)));
  CHK(ensure(c->item_first, sizeof(int32_t) * ci.first.size()));
  CHK(ensure(c->item_order, sizeof(int32_t) * std::max<size_t>(n, 1)));
  CHK(ensure(c->queue, sizeof(unsigned int)));
  if (!ci.items.empty())
    HIPC(hipMemcpyAsync(c->work.p, ci.items.data(), sizeof(int32_t) * ci.items.size(), hipMemcpyHostToDevice,
                        c->stream));
  HIPC(hipMemcpyAsync(c->item_first.p, ci.first.data(), sizeof(int32_t) * ci.first.size(), hipMemcpyHostToDevice,
                      c->stream));
  if (n) HIPC(hipMemcpyAsync(c->item_order.p, ci.order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  CHK(ensure(c->partial, sizeof(GeneAcc) * std::max<size_t>(n, 1) * 16));
  return EQTL_OK;
}

// The cis autotuner: for each window-size class present in this block and not
// yet timed for the current scan configuration, one launch per candidate item
// size over that class's items alone (hit and polish output off, counters
// cleared afterwards), event-timed; the fastest candidate is kept (ties: the
// larger items). The tuned sizes change only the work split, not the tests.
static int cis_autotune(eqtl_gpu_ctx* c, const WaldArgs& base, int64_t snp0, int64_t m, int n_groups) {
  const std::vector<int> key = {c->scan_model, c->scan_prec, c->nq_pad, c->n_pad};
  if (c->cis_tune_key != key) {
    c->cis_tune_key = key;
    for (int q = 0; q < kCisClasses; ++q) {
      c->cis_tiles[q] = kCisItemTilesDefault;
      c->cis_timed[q] = 0;
    }
  }
  CisItems probe;
  cis_items(c, snp0, m, n_groups, c->cis_tiles, -1, probe);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  for (int q = 0; q < kCisClasses; ++q) {
    if (c->cis_timed[q] || !probe.classes_present[q]) continue;
    if (!e0) {
      HIPC(hipEventCreate(&e0));
      HIPC(hipEventCreate(&e1));
    }
    float best_ms = 0;
    int best = c->cis_tiles[q];
    int rc = EQTL_OK;
    for (int v = 0; v < 3 && rc == EQTL_OK; ++v) {
      int tiles[kCisClasses];
      for (int u = 0; u < kCisClasses; ++u) tiles[u] = c->cis_tiles[u];
      tiles[q] = kCisCand[q][v];
      CisItems ci;
      cis_items(c, snp0, m, n_groups, tiles, q, ci);
      rc = upload_items(c, ci);
      if (rc != EQTL_OK) break;
      WaldArgs t = base;
      t.work = c->work.as<int32_t>();
      t.order = c->item_order.as<int32_t>();
      t.queue = c->queue.as<unsigned int>();
      t.n_items = static_cast<int32_t>(ci.order.size());
      t.partial = c->partial.as<GeneAcc>();
      t.hit_cap = 0;
      t.polish_cap = 0;
      if (hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(EQTL_E_HIP, "hipEventRecord failed");
      if (rc == EQTL_OK) rc = launch_wald(c, t, dim3(static_cast<unsigned>(t.n_items)));
      if (rc == EQTL_OK && hipEventRecord(e1, c->stream) != hipSuccess) rc = fail(EQTL_E_HIP, "hipEventRecord failed");
      if (rc == EQTL_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(EQTL_E_HIP, "hipEventSynchronize failed");
      float ms = 0;
      if (rc == EQTL_OK) (void)hipEventElapsedTime(&ms, e0, e1);
      if (rc == EQTL_OK && (v == 0 || ms <= best_ms)) {
        best_ms = ms;
        best = kCisCand[q][v];
      }
    }
    if (rc != EQTL_OK) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return rc;
    }
    c->cis_tiles[q] = best;
    c->cis_timed[q] = 1;
  }
  if (e0) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return EQTL_OK;
}

int eqtl_gpu_scan_block(eqtl_gpu_ctx* c, const uint8_t* rows, int32_t on_dev, int64_t snp0, int64_t m,
                        int32_t bps, eqtl_hit* hits, int64_t hit_cap, int64_t* n_hits) {
  if (!c || !rows || m <= 0 || snp0 < 0) return fail(EQTL_E_ARG, "bad scan_block arguments");
  if (!c->scanning) return fail(EQTL_E_STATE, "scan_begin first");
  if (bps < min_bps(c)) return fail(EQTL_E_ARG, "bytes_per_snp too small for n_file");
  if (c->have_windows && snp0 + m > c->m_annot) return fail(EQTL_E_ARG, "SNP block beyond the annotation");
  if (hit_cap > 0 && !hits) return fail(EQTL_E_ARG, "hits buffer is null");
  CHK(set_device(c));
  c->hit_stage.clear();
  const bool lm = c->scan_model == EQTL_MODEL_LM;
  const bool glm = c->scan_model == EQTL_MODEL_GLM;
  CHK(ensure_block(c, m, bps, true));
  int rc;
  const uint8_t* rd = stage_rows(c, rows, on_dev, m, bps, &rc);
  if (rc) return rc;
  // GLM: the score needs each SNP's centring mean (kept in the per-SNP gtg slot)
  CHK(run_decode(c, rd, m, bps, true, nullptr, glm ? c->gtg.as<double>() : nullptr));
  const int64_t m_pad = round_up(m, 128);
  const int np = c->n_pad;
  const float* Aop = c->A32.as<float>();
  if (!glm) {
    // g' = g_c - Wx Wx^T g_c (FP64) and gtg = ||g'||^2 before the rotation / contraction
    if (!c->have_wx) return fail(EQTL_E_STATE, "covariate basis missing: prepare genes first");
    Timed t(c, "residualize");
    hipLaunchKernelGGL(k_residualize, dim3(static_cast<unsigned>(m_pad / (4 * kResRows))), dim3(256), 0, c->stream,
                       c->A64.as<double>(), c->A32.as<float>(), m_pad, c->n, np, c->Wx.as<double>(), c->p,
                       c->gtg.as<double>());
    CHK(check_launch("k_residualize"));
    CHK(t.done());
  }
  if (glm) {
    // the score test contracts the centred dosage itself (no eigenbasis)
  } else if (!lm) {
    Timed t(c, "rotate");
    dim3 grid(static_cast<unsigned>(np / 64), static_cast<unsigned>(m_pad / 64));
    if (c->scan_prec == EQTL_PREC_FP32)
      hipLaunchKernelGGL(k_rotate_f32, grid, dim3(256), 0, c->stream, c->A32.as<float>(), c->Q32.as<float>(),
                         c->Gt.as<float>(), np);
    else {
      static bool attr_set = false;
      if (!attr_set) {
        HIPC(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rotate_lds<false>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kRotLds)));
        HIPC(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rotate_lds<true>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kRotLds)));
        attr_set = true;
      }
      const dim3 g2(static_cast<unsigned>((np + kRotT - 1) / kRotT), static_cast<unsigned>(m_pad / kRotT));
      const bool lr_rows = c->lr_on && c->scan_scope == EQTL_SCOPE_TRANS;
      if (lr_rows) CHK(ensure(c->lr_G2, sizeof(__bf16) * 2 * static_cast<size_t>(m_pad) * np));
      __bf16* ghi = lr_rows ? c->lr_G2.as<__bf16>() : nullptr;
      hipLaunchKernelGGL(k_rotate_lds<false>, g2, dim3(256), kRotLds, c->stream, c->A32.as<float>(),
                         c->Qhi.as<__bf16>(), c->Qlo.as<__bf16>(), c->Gt.as<float>(), np, np,
                         static_cast<int64_t>(np), ghi, ghi ? ghi + static_cast<size_t>(m_pad) * np : nullptr,
                         nullptr, 0);
      if (lr_rows) {
        // the SNPs' covariate projections at the interpolation nodes, A x (Q Z),
        // then their weighted sums of squares, (Q^T g)^2 x D(nodes): written
        // as the hi / mid / lo planes the interpolation reads
        const size_t tpl = static_cast<size_t>(m_pad) * c->lr_ldT;
        CHK(ensure(c->lr_Tp, sizeof(__bf16) * 3 * tpl));
        __bf16* th = c->lr_Tp.as<__bf16>();
        const dim3 g3(static_cast<unsigned>((c->lr_ncols + kRotT - 1) / kRotT), static_cast<unsigned>(m_pad / kRotT));
        hipLaunchKernelGGL(k_rotate_lds<false>, g3, dim3(256), kRotLds, c->stream, c->A32.as<float>(),
                           c->lr_QZhi.as<__bf16>(), c->lr_QZlo.as<__bf16>(), nullptr, np, c->lr_ncols, c->lr_ldT, th,
                           th + 2 * tpl, th + tpl, 0);
        const dim3 g4(static_cast<unsigned>((c->lr_nnodes + kRotT - 1) / kRotT), static_cast<unsigned>(m_pad / kRotT));
        hipLaunchKernelGGL(k_rotate_lds<true>, g4, dim3(256), kRotLds, c->stream, c->Gt.as<float>(),
                           c->lr_Dhi.as<__bf16>(), c->lr_Dlo.as<__bf16>(), nullptr, np, c->lr_nnodes, c->lr_ldT, th,
                           th + 2 * tpl, th + tpl, c->lr_ncols);
      }
    }
    CHK(check_launch("k_rotate"));
    CHK(t.done());
    Aop = c->Gt.as<float>();
  }
  const int n_groups = c->G_pad / 16;
  // cis: work items {group, tile_lo, tile_hi} over the tiles of this block
  // that hold some window SNP of the group's genes (built after the tuner)
  const bool cis = c->scan_scope == EQTL_SCOPE_CIS;
  int n_items = 0;
  if (cis) {
    CisItems ci;
    cis_items(c, snp0, m, n_groups, c->cis_tiles, -1, ci);
    if (ci.order.empty()) {
      if (n_hits) *n_hits = 0;
      return EQTL_OK;
    }
  }
  // work split: enough workgroups to fill the chip, SNP tiles strided per split
  const int n_groups_pad = static_cast<int>(round_up(n_groups, 8));
  const int n_tiles = static_cast<int>(m_pad / 64);
  const int n_tiles128 = static_cast<int>(m_pad / kTileRows);
  // Each gene group is split over as many workgroups as one XCD has CUs, and
  // the kernels place a group's splits on one XCD: the XCD then sweeps its
  // SNP tiles against one group at a time, so the group's column blocks are
  // read from HBM once and served from that XCD's L2 to all of its CUs.
  const int cus = std::max(c->n_cu, 8);
  const bool lr = c->lr_on && !cis && c->scan_model == EQTL_MODEL_LMM;
  const int n_units = lr ? c->lr_n_sg : n_groups;  // the low-rank kernel's unit is a super-group of 64 genes
  int n_splits = std::max(cus / 8, (2 * cus + n_units - 1) / n_units);
  n_splits = std::max(1, std::min(n_tiles128, n_splits));
  // The low-rank kernel shares each XCD's CUs between lr_spx super-groups
  // (row splits n_splits = CUs per XCD / lr_spx): with four, an SNP tile is
  // read into an XCD's L2 once per four super-groups.
  int lr_spx = 1;
  if (lr) {
    const int cpx = cus / 8;
    for (int s : {4, 2}) {
      if (cpx % s == 0 && c->lr_n_sg >= 8 * s) {
        lr_spx = s;
        break;
      }
    }
    n_splits = cpx / lr_spx;
  }
  CHK(ensure(c->partial, sizeof(GeneAcc) * (cis ? static_cast<size_t>(n_items) * 16
                                                : static_cast<size_t>(n_splits) * c->G_pad)));
  // device lists: hits (every hit of the block) and polish items; both grow
  // on demand, the block's Wald launch is then repeated (its outputs are
  // overwritten, not accumulated, until k_merge)
  const bool polish_on = c->scan_model != EQTL_MODEL_LM;
  if (c->hit_cap_dev < std::max<int64_t>(hit_cap, 1024)) {
    const int64_t want = std::max<int64_t>(hit_cap, 1024);
    c->hit_cap_dev = 0;
    CHK(ensure(c->hits, sizeof(eqtl_hit) * static_cast<size_t>(want)));
    c->hit_cap_dev = want;
  }
  if (polish_on && c->polish_cap_dev < kPolishCap) {
    c->polish_cap_dev = 0;
    CHK(ensure(c->polish, sizeof(PolishItem) * static_cast<size_t>(kPolishCap)));
    c->polish_cap_dev = kPolishCap;
  }
  WaldArgs a;
  a.A = Aop;
  a.gc = gene_cols(c);
  a.n_pad = np;
  a.nq = c->nq;
  a.G = c->G;
  a.G_pad = c->G_pad;
  a.n_groups = n_groups;
  a.n_groups_pad = n_groups_pad;
  a.n_splits = n_splits;
  a.n_tiles = n_tiles;
  a.m = m;
  a.snp0 = snp0;
  a.keep = c->keep.as<uint8_t>();
  a.af = c->af.as<double>();
  a.gtg = c->gtg.as<double>();
  a.gene_scal = c->gscal.as<double>();
  a.gene_valid = c->gvalid.as<int32_t>();
  a.snp_chrom = c->m_annot > 0 ? c->snp_chrom.as<int32_t>() : nullptr;
  a.snp_pos = c->m_annot > 0 ? c->snp_pos.as<int64_t>() : nullptr;
  a.gene_chrom = c->have_windows ? c->gene_chrom.as<int32_t>() : nullptr;
  a.gene_cs = c->have_windows ? c->gene_cs.as<int64_t>() : nullptr;
  a.gene_ce = c->have_windows ? c->gene_ce.as<int64_t>() : nullptr;
  a.df = static_cast<double>(c->n - c->p - 1);
  a.tail = c->tail.as<TTail>();
  a.pthr = c->pthr;
  a.partial = c->partial.as<GeneAcc>();
  a.hit_count = c->hit_count.as<unsigned long long>();
  a.A64 = c->A64.as<double>();
  // FP64 residual vectors in sample space for the numerator refinement: LM
  // y_s, GLM the score residual r, LMM Q (s o r)
  a.vref = (lm || glm) ? c->U64.as<double>() : c->V64.as<double>();
  a.n = c->n;
  a.tau = c->scan_prec == EQTL_PREC_FP32 ? kRefineTauF32 : kRefineTauBF16X3;
  a.tau_hi = glm ? kGlmTauHi : kRefineTauHi;
  a.kappa = kPolishKappa;
  a.cis = cis ? 1 : 0;
  a.work = nullptr;
  a.order = nullptr;
  a.queue = nullptr;
  a.n_items = 0;
  if (cis) {
    CHK(cis_autotune(c, a, snp0, m, n_groups));
    CisItems ci;
    cis_items(c, snp0, m, n_groups, c->cis_tiles, -1, ci);
    CHK(upload_items(c, ci));
    n_items = static_cast<int>(ci.order.size());
    a.work = c->work.as<int32_t>();
    a.order = c->item_order.as<int32_t>();
    a.queue = c->queue.as<unsigned int>();
    a.n_items = n_items;
    a.partial = c->partial.as<GeneAcc>();
  }
  int64_t got = 0, npol = 0;
  for (int pass = 0;; ++pass) {
    a.hits = c->hits.as<eqtl_hit>();
    a.hit_cap = c->hit_cap_dev;
    a.polish = polish_on ? c->polish.as<PolishItem>() : nullptr;
    a.polish_cap = polish_on ? c->polish_cap_dev : 0;
    HIPC(hipMemsetAsync(c->hit_count.p, 0, 2 * sizeof(unsigned long long), c->stream));
    LrArgs la;
    dim3 lr_grid, num_grid;
    if (lr) {
      la.w = a;
      la.Ghi = c->lr_G2.as<__bf16>();
      la.lo_off = m_pad * np;
      la.Tp = c->lr_Tp.as<__bf16>();
      la.tpl = static_cast<int64_t>(m_pad) * c->lr_ldT;
      la.ldT = c->lr_ldT;
      la.sg_col = c->lr_sg_col.as<int32_t>();
      la.sg_gcol = c->lr_sg_col.as<int32_t>() + c->lr_n_sg;
      la.sg_R = c->lr_sg_R.as<int32_t>();
      la.slot_gene = c->lr_slot_gene.as<int32_t>();
      la.lcoef = c->lr_coef.as<__bf16>();
      la.gscale = c->lr_gscale.as<double>();
      la.gb = c->lr_gb.as<__bf16>();
      la.n_sg = c->lr_n_sg;
      la.p = c->p;
      la.sg_per_xcd = lr_spx;
      const int lr_ppx = lr_spx;  // k_lr_num: pairs per XCD and round, with the same n_splits
      la.pair_per_xcd = lr_ppx;
      const int n_t128 = static_cast<int>((a.n_tiles + 1) / 2);
      CHK(ensure(c->lr_fold, sizeof(float) * lr_fold_base(n_t128, 0, c->lr_n_sg)));
      la.fold = c->lr_fold.as<float>();
      const int64_t rounds = (c->lr_n_sg + 8 * lr_spx - 1) / (8 * lr_spx);
      lr_grid = dim3(static_cast<unsigned>(rounds * 8 * lr_spx * n_splits));
      const int64_t n_pairs = (c->lr_n_sg + 1) / 2;
      const int64_t prounds = (n_pairs + 8 * lr_ppx - 1) / (8 * lr_ppx);
      num_grid = dim3(static_cast<unsigned>(prounds * 8 * lr_ppx * n_splits));
      {
        Timed t(c, "num");  // k_lr_num
        CHK(launch_lr_num(c, la, num_grid));
        CHK(t.done());
      }
      {
        Timed t(c, "wald");  // k_wald_lr (the interpolation)
        CHK(launch_wald_lr(c, la, lr_grid));
        CHK(t.done());
      }
    } else {
      Timed t(c, "wald");  // the eight-wave kernel
      CHK(launch_wald(c, a, dim3(static_cast<unsigned>(cis ? n_items : n_groups_pad * n_splits))));
      CHK(t.done());
    }
    if (lr) {
      Timed t(c, "tests");
      CHK(launch_lr_tests(c, la, lr_grid));
      CHK(t.done());
    }
    unsigned long long cnt2[2] = {0, 0};
    HIPC(hipMemcpyAsync(cnt2, c->hit_count.p, sizeof(cnt2), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    got = static_cast<int64_t>(cnt2[0]);
    npol = static_cast<int64_t>(cnt2[1]);
    // polished tests may append hits of their own: room for got + npol
    const bool grow_hits = got + npol > c->hit_cap_dev;
    const bool grow_pol = polish_on && npol > c->polish_cap_dev;
    if (!grow_hits && !grow_pol) break;
    if (pass > 0) return fail(EQTL_E_STATE, "hit / polish counts changed between identical launches");
    if (grow_hits) {
      const int64_t want = (got + npol) + (got + npol) / 4 + 1024;
      int rc2 = ensure(c->hits, sizeof(eqtl_hit) * static_cast<size_t>(want));
      c->hit_cap_dev = rc2 == EQTL_OK ? want : 0;
      if (rc2 != EQTL_OK)
        return fail(rc2, "the block's " + std::to_string(got + npol) +
                             " hits do not fit in device memory: scan smaller SNP blocks");
      c->hit_cap_dev = want;
    }
    if (grow_pol) {
      const int64_t want = npol + npol / 4 + 1024;
      int rc2 = ensure(c->polish, sizeof(PolishItem) * static_cast<size_t>(want));
      c->polish_cap_dev = rc2 == EQTL_OK ? want : 0;
      if (rc2 != EQTL_OK)
        return fail(rc2, "the block's " + std::to_string(npol) +
                             " FP64 polish items do not fit in device memory: scan smaller SNP blocks");
      c->polish_cap_dev = want;
    }
  }
  if (cis)
    hipLaunchKernelGGL(k_merge_items, dim3((c->G + 255) / 256), dim3(256), 0, c->stream, c->state.as<GeneAcc>(),
                       c->partial.as<GeneAcc>(), c->item_first.as<int32_t>(), c->G);
  else
    hipLaunchKernelGGL(k_merge, dim3((c->G + 255) / 256), dim3(256), 0, c->stream, c->state.as<GeneAcc>(),
                       c->partial.as<GeneAcc>(), c->G, c->G_pad, n_splits);
  CHK(check_launch("k_merge"));
  if (npol > 0) {
    Timed t(c, "polish");
    const int64_t chunk = std::min<int64_t>(npol, kPolishChunk);
    CHK(ensure(c->polish_gt, sizeof(double) * static_cast<size_t>(chunk) * c->n));
    CHK(ensure(c->polish_out, sizeof(eqtl_hit) * static_cast<size_t>(npol)));
    CHK(ensure(c->pbest_key, 2 * sizeof(unsigned long long) * static_cast<size_t>(c->G)));
    PolishArgs pa;
    pa.A64 = c->A64.as<double>();
    pa.n_pad = np;
    pa.n = c->n;
    pa.p = c->p;
    pa.Q = c->Q64.as<double>();
    pa.lam = c->lam64.as<double>();
    pa.delta = c->delta.as<double>();
    pa.U = c->U64.as<double>();
    pa.Xt = glm ? c->X64.as<double>() : c->Xt64.as<double>();
    pa.Lc = c->Lc.as<double>();
    pa.w = glm ? c->V64.as<double>() : nullptr;
    pa.mean = glm ? c->gtg.as<double>() : nullptr;
    pa.gene_scal = c->gscal.as<double>();
    pa.af = c->af.as<double>();
    pa.df = a.df;
    pa.lbeta = std::lgamma(0.5 * a.df) + std::lgamma(0.5) - std::lgamma(0.5 * a.df + 0.5);
    pa.snp0 = snp0;
    pa.gt = c->polish_gt.as<double>();
    pa.pthr = c->pthr;
    pa.hits = c->hits.as<eqtl_hit>();
    pa.hit_count = c->hit_count.as<unsigned long long>();
    pa.hit_cap = c->hit_cap_dev;
    pa.state = c->state.as<GeneAcc>();
    for (int64_t i0 = 0; i0 < npol; i0 += chunk) {
      pa.items = c->polish.as<PolishItem>() + i0;
      pa.count = std::min<int64_t>(chunk, npol - i0);
      pa.out = c->polish_out.as<eqtl_hit>() + i0;
      if (glm)
        hipLaunchKernelGGL(k_polish<EQTL_MODEL_GLM>, dim3(static_cast<unsigned>(pa.count)), dim3(256), 0, c->stream,
                           pa);
      else {
        hipLaunchKernelGGL(k_polish_rotate, dim3(static_cast<unsigned>(pa.count), static_cast<unsigned>((c->n + 63) / 64)),
                           dim3(256), 0, c->stream, pa);
        CHK(check_launch("k_polish_rotate"));
        hipLaunchKernelGGL(k_polish<EQTL_MODEL_LMM>, dim3(static_cast<unsigned>(pa.count)), dim3(256), 0, c->stream,
                           pa);
      }
      CHK(check_launch("k_polish"));
    }
    // best over the kernel's best and every polished test (lexicographic in (p, SNP))
    const eqtl_hit* po = c->polish_out.as<eqtl_hit>();
    unsigned long long* key = c->pbest_key.as<unsigned long long>();
    const dim3 gg((c->G + 255) / 256), gi(static_cast<unsigned>((npol + 255) / 256));
    hipLaunchKernelGGL(k_pbest_init, gg, dim3(256), 0, c->stream, c->state.as<GeneAcc>(), key, c->G);
    hipLaunchKernelGGL(k_pbest_p, gi, dim3(256), 0, c->stream, po, npol, key);
    hipLaunchKernelGGL(k_pbest_state_snp, gg, dim3(256), 0, c->stream, c->state.as<GeneAcc>(), key, c->G);
    hipLaunchKernelGGL(k_pbest_snp, gi, dim3(256), 0, c->stream, po, npol, key);
    hipLaunchKernelGGL(k_pbest_write, gi, dim3(256), 0, c->stream, po, npol, key, c->state.as<GeneAcc>());
    CHK(check_launch("k_pbest"));
    unsigned long long cnt = 0;
    HIPC(hipMemcpyAsync(&cnt, c->hit_count.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    got = static_cast<int64_t>(cnt);
    CHK(t.done());
  }
  // every hit of the block, sorted (SNP-major, gene order within a SNP)
  c->hit_stage.resize(static_cast<size_t>(got));
  if (got > 0) {
    HIPC(hipMemcpyAsync(c->hit_stage.data(), c->hits.p, sizeof(eqtl_hit) * got, hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    std::sort(c->hit_stage.begin(), c->hit_stage.end(), [](const eqtl_hit& x, const eqtl_hit& y) {
      return x.snp != y.snp ? x.snp < y.snp : x.gene < y.gene;
    });
  }
  if (n_hits) *n_hits = got;
  const int64_t take = std::min<int64_t>(got, std::max<int64_t>(hit_cap, 0));
  if (take > 0) std::memcpy(hits, c->hit_stage.data(), sizeof(eqtl_hit) * take);
  return EQTL_OK;
}

int eqtl_gpu_scan_hits(eqtl_gpu_ctx* c, int64_t first, eqtl_hit* hits, int64_t cap, int64_t* n_out) {
  if (!c || first < 0 || cap < 0 || (cap > 0 && !hits) || !n_out) return fail(EQTL_E_ARG, "bad scan_hits arguments");
  const int64_t total = static_cast<int64_t>(c->hit_stage.size());
  const int64_t take = first >= total ? 0 : std::min(cap, total - first);
  if (take > 0) std::memcpy(hits, c->hit_stage.data() + first, sizeof(eqtl_hit) * take);
  *n_out = take;
  return EQTL_OK;
}

int eqtl_gpu_cis_tuning(eqtl_gpu_ctx* c, int32_t* tiles, int32_t* timed) {
  if (!c || !tiles || !timed) return fail(EQTL_E_ARG, "bad cis_tuning arguments");
  for (int q = 0; q < kCisClasses; ++q) {
    tiles[q] = c->cis_tiles[q];
    timed[q] = c->cis_timed[q];
  }
  return EQTL_OK;
}

int eqtl_gpu_scan_end(eqtl_gpu_ctx* c, eqtl_gene_state* out) {
  if (!c || !out) return fail(EQTL_E_ARG, "bad scan_end arguments");
  if (!c->scanning) return fail(EQTL_E_STATE, "no scan in progress");
  CHK(set_device(c));
  DBuf o;
  CHK(ensure(o, sizeof(eqtl_gene_state) * c->G));
  hipLaunchKernelGGL(k_finalize, dim3((c->G + 255) / 256), dim3(256), 0, c->stream, c->state.as<GeneAcc>(),
                     o.as<eqtl_gene_state>(), c->G);
  CHK(check_launch("k_finalize"));
  HIPC(hipMemcpyAsync(out, o.p, sizeof(eqtl_gene_state) * c->G, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  c->scanning = false;
  return EQTL_OK;
}

// ---- cis permutations (scan_gene_snps permutation loop, scan_cis.hpp:69-180) ----

int eqtl_gpu_upload_bed(eqtl_gpu_ctx* c, const uint8_t* rows, int32_t on_dev, int64_t m, int32_t bps) {
  if (!c || !rows || m <= 0 || bps <= 0) return fail(EQTL_E_ARG, "bad upload_bed arguments");
  CHK(set_device(c));
  const size_t bytes = static_cast<size_t>(m) * bps;
  CHK(ensure(c->bed_res, bytes));
  HIPC(hipMemcpyAsync(c->bed_res.p, rows, bytes, on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                      c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  c->bed_m = m;
  c->bed_bps = bps;
  return EQTL_OK;
}

// Pseudo-genes (gene, b) per batch, and SNPs per scan block of a batch.
constexpr int kPermBatch = 16384;
constexpr int64_t kCisBlock = 32768;
// A permutation whose split-precision T = -log10(max(min p, 1e-300)) lies
// within kExactBand * (1 + T_obs) of the gene's observed T is decided again
// from FP64 tests. The band is 20x (BF16X3) / 100x (FP32) the largest T error
// of the contractions measured against FP64 (relative p error 2.1e-4 for the
// LM and 8e-5 for the LMM split form, i.e. 9e-5 in T).
constexpr double kExactBandBF16X3 = 2e-3;
constexpr double kExactBandF32 = 5e-4;

// FP64 rows of SNPs [lo, hi) of the resident genotypes (decode, covariate
// residualisation, and for the LMM the rotation Q^T g' as an FP64 GEMM), with
// their keep flags and (LM) ||g'||^2; kept while the range covers later asks.
static int exact_rows(eqtl_gpu_ctx* c, int64_t lo, int64_t hi, bool lmm) {
  if (c->ex_valid && c->ex_lo <= lo && hi <= c->ex_hi) return EQTL_OK;
  c->ex_valid = false;
  const int n = c->n, np = c->n_pad;
  const int64_t M = hi - lo;
  CHK(ensure(c->ex_R, sizeof(double) * static_cast<size_t>(M) * n));
  CHK(ensure(c->ex_keep, static_cast<size_t>(M)));
  CHK(ensure(c->ex_gtg, sizeof(double) * static_cast<size_t>(M)));
  const int32_t bps = c->bed_bps;
  for (int64_t j0 = lo; j0 < hi; j0 += kCisBlock) {
    const int64_t mb = std::min<int64_t>(kCisBlock, hi - j0);
    const int64_t m_pad = round_up(mb, 128);
    CHK(ensure_block(c, mb, bps, true));
    CHK(run_decode(c, c->bed_res.as<uint8_t>() + j0 * bps, mb, bps, true, nullptr));
    hipLaunchKernelGGL(k_residualize, dim3(static_cast<unsigned>(m_pad / (4 * kResRows))), dim3(256), 0, c->stream,
                       c->A64.as<double>(), c->A32.as<float>(), m_pad, n, np, c->Wx.as<double>(), c->p,
                       c->gtg.as<double>());
    CHK(check_launch("k_residualize"));
    double* dst = c->ex_R.as<double>() + static_cast<size_t>(j0 - lo) * n;
    if (lmm) {
      CHK(dgemm_tn(c, n, static_cast<int>(mb), n, c->Q64.as<double>(), n, c->A64.as<double>(), np, dst, n));
    } else {
      HIPC(hipMemcpy2DAsync(dst, sizeof(double) * n, c->A64.p, sizeof(double) * np, sizeof(double) * n,
                            static_cast<size_t>(mb), hipMemcpyDeviceToDevice, c->stream));
    }
    HIPC(hipMemcpyAsync(c->ex_keep.as<uint8_t>() + (j0 - lo), c->keep.p, static_cast<size_t>(mb),
                        hipMemcpyDeviceToDevice, c->stream));
    HIPC(hipMemcpyAsync(c->ex_gtg.as<double>() + (j0 - lo), c->gtg.p, sizeof(double) * mb, hipMemcpyDeviceToDevice,
                        c->stream));
  }
  c->ex_lo = lo;
  c->ex_hi = hi;
  c->ex_valid = true;
  return EQTL_OK;
}

// FP64 minimum p over each task's window, for the gene side now prepared
// (columns tasks[i].e), the exact rows covering every task's range; out[i] = 2
// when the window has no tested SNP. The scan's t-tail table (c->tail, built
// by scan_begin for df = n - p - 1) must be current.
static int exact_minp(eqtl_gpu_ctx* c, int model, const std::vector<ExactTask>& tasks, std::vector<double>& out) {
  out.assign(tasks.size(), 2.0);
  if (tasks.empty()) return EQTL_OK;
  int64_t lo = INT64_MAX, hi = -1;
  std::vector<ExactPieceRef> pieces;
  for (size_t i = 0; i < tasks.size(); ++i) {
    const ExactTask& t = tasks[i];
    lo = std::min(lo, t.jlo);
    hi = std::max(hi, t.jhi);
    for (int64_t j = t.jlo; j < t.jhi; j += kExactPiece) {
      ExactPieceRef pr;
      pr.task = static_cast<int32_t>(i);
      pr.pad = 0;
      pr.jlo = j;
      pr.jhi = std::min(t.jhi, j + kExactPiece);
      pieces.push_back(pr);
    }
  }
  if (lo >= hi || pieces.empty()) return EQTL_OK;
  if (c->tail_df != static_cast<double>(c->n - c->p - 1)) return fail(EQTL_E_STATE, "t tail table not built");
  CHK(exact_rows(c, lo, hi, model == EQTL_MODEL_LMM));
  CHK(ensure(c->ex_tasks, sizeof(ExactTask) * tasks.size() + sizeof(ExactPieceRef) * pieces.size()));
  CHK(ensure(c->ex_out, sizeof(unsigned long long) * tasks.size()));
  ExactTask* dt = c->ex_tasks.as<ExactTask>();
  ExactPieceRef* dp = reinterpret_cast<ExactPieceRef*>(dt + tasks.size());
  HIPC(hipMemcpyAsync(dt, tasks.data(), sizeof(ExactTask) * tasks.size(), hipMemcpyHostToDevice, c->stream));
  HIPC(hipMemcpyAsync(dp, pieces.data(), sizeof(ExactPieceRef) * pieces.size(), hipMemcpyHostToDevice, c->stream));
  double two = 2.0;
  unsigned long long two_bits = 0;
  std::memcpy(&two_bits, &two, sizeof(two));
  hipLaunchKernelGGL(k_fill_u64, dim3(static_cast<unsigned>((tasks.size() + 255) / 256)), dim3(256), 0, c->stream,
                     c->ex_out.as<unsigned long long>(), static_cast<int>(tasks.size()), two_bits);
  ExactArgs a;
  a.tasks = dt;
  a.pieces = dp;
  a.r0 = c->ex_lo;
  a.R = c->ex_R.as<double>();
  a.keep = c->ex_keep.as<uint8_t>();
  a.gtg = c->ex_gtg.as<double>();
  a.snp_chrom = c->snp_chrom.as<int32_t>();
  a.snp_pos = c->snp_pos.as<int64_t>();
  a.n = c->n;
  a.p = c->p;
  a.lam = c->lam64.as<double>();
  a.delta = c->delta.as<double>();
  a.U = c->U64.as<double>();
  a.Xt = c->Xt64.as<double>();
  a.Lc = c->Lc.as<double>();
  a.gene_scal = c->gscal.as<double>();
  a.tail = c->tail.as<TTail>();
  a.df = static_cast<double>(c->n - c->p - 1);
  a.out = c->ex_out.as<unsigned long long>();
  Timed t(c, "exact");
  const dim3 grid(static_cast<unsigned>(pieces.size()));
  if (model == EQTL_MODEL_LMM)
    hipLaunchKernelGGL(k_exact_minp<EQTL_MODEL_LMM>, grid, dim3(256), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_exact_minp<EQTL_MODEL_LM>, grid, dim3(256), 0, c->stream, a);
  CHK(check_launch("k_exact_minp"));
  std::vector<unsigned long long> bits(tasks.size());
  HIPC(hipMemcpyAsync(bits.data(), c->ex_out.p, sizeof(unsigned long long) * tasks.size(), hipMemcpyDeviceToHost,
                      c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < tasks.size(); ++i) std::memcpy(&out[i], &bits[i], sizeof(double));
  return t.done();
}

int eqtl_gpu_cis_perm(eqtl_gpu_ctx* c, int32_t precision, int32_t B, int32_t seed, const uint32_t* gene_hash,
                      const double* best_p_obs, int32_t delta_mode, double* perm_min_p, double* p_emp) {
  if (!c || !gene_hash || !best_p_obs || B <= 0) return fail(EQTL_E_ARG, "bad cis_perm arguments");
  if (c->scanning) return fail(EQTL_E_STATE, "a scan is in progress");
  const int model = c->prepped_model;
  if (model == EQTL_MODEL_GLM)
    return fail(EQTL_E_ARG, "gene-level permutations are provided for lm and lmm only (not the NB glm)");
  if (model != EQTL_MODEL_LM && model != EQTL_MODEL_LMM) return fail(EQTL_E_STATE, "genes not prepared");
  if (!c->have_windows || static_cast<int>(c->h_gene_chrom.size()) != c->G)
    return fail(EQTL_E_STATE, "cis permutations need the gene windows");
  if (c->bed_m <= 0 || c->bed_m != c->m_annot) return fail(EQTL_E_STATE, "upload the bed (eqtl_gpu_upload_bed) first");
  if (model == EQTL_MODEL_LMM && delta_mode != EQTL_DELTA_REF && delta_mode != EQTL_DELTA_NEWTON)
    return fail(EQTL_E_ARG, "permutation delta mode must be REF or NEWTON");
  CHK(set_device(c));
  CHK(ensure_blas(c));
  const int G = c->G, n = c->n;
  CHK(cis_windows(c));
  // per-gene permutation bases, kept apart from the gene side that the batches reuse
  DBuf base, yt, dlt;
  CHK(ensure(base, sizeof(double) * static_cast<size_t>(n) * G));
  if (model == EQTL_MODEL_LM) {
    HIPC(hipMemcpyAsync(base.p, c->U64.p, sizeof(double) * static_cast<size_t>(n) * G, hipMemcpyDeviceToDevice,
                        c->stream));
  } else {
    CHK(ensure(yt, sizeof(double) * static_cast<size_t>(n) * G));
    CHK(ensure(dlt, sizeof(double) * G));
    HIPC(hipMemcpyAsync(yt.p, c->Yt64.p, sizeof(double) * static_cast<size_t>(n) * G, hipMemcpyDeviceToDevice,
                        c->stream));
    HIPC(hipMemcpyAsync(dlt.p, c->delta.p, sizeof(double) * G, hipMemcpyDeviceToDevice, c->stream));
    const size_t N = static_cast<size_t>(n) * G;
    hipLaunchKernelGGL(k_whitened_residual, dim3(static_cast<unsigned>((N + 255) / 256)), dim3(256), 0, c->stream,
                       c->U64.as<double>(), dlt.as<double>(), c->lam64.as<double>(), n, G, base.as<double>());
    CHK(check_launch("k_whitened_residual"));
  }
  const std::vector<int64_t> glo = c->win_lo, ghi = c->win_hi;
  const std::vector<int32_t> gch = c->h_gene_chrom;
  const std::vector<int64_t> gcs = c->h_gene_cs, gce = c->h_gene_ce;
  const int p = c->p;
  // genes with a tested SNP (the reference permutes only when n_tested > 0)
  std::vector<int32_t> genes;
  for (int k = 0; k < G; ++k)
    if (std::isfinite(best_p_obs[k]) && best_p_obs[k] <= 1.0 && glo[k] < ghi[k]) genes.push_back(k);
  std::vector<double> minp(static_cast<size_t>(G) * B, std::numeric_limits<double>::quiet_NaN());
  const int nb_max = std::max(1, kPermBatch / B);
  // Exact counts: the observed statistic of every permuted gene in FP64 (its
  // gene side is still the observed one), with the FP64 rows of the whole
  // span when they fit a third of the free device memory (else per group).
  c->ex_valid = false;
  CHK(ensure_tail(c));
  const double band = precision == EQTL_PREC_FP32 ? kExactBandF32 : kExactBandBF16X3;
  std::vector<double> Tobs64(G, std::numeric_limits<double>::quiet_NaN());
  auto task_of = [&](int e, int k) {
    ExactTask t;
    t.e = e;
    t.chrom = gch[k];
    t.jlo = glo[k];
    t.jhi = ghi[k];
    t.cs = gcs[k];
    t.ce = gce[k];
    return t;
  };
  if (!genes.empty()) {
    int64_t lo_all = INT64_MAX, hi_all = -1;
    for (int k : genes) {
      lo_all = std::min(lo_all, glo[k]);
      hi_all = std::max(hi_all, ghi[k]);
    }
    size_t free_b = 0, total_b = 0;
    HIPC(hipMemGetInfo(&free_b, &total_b));
    const bool whole = static_cast<double>(hi_all - lo_all) * n * sizeof(double) <= free_b / 3.0;
    if (whole) CHK(exact_rows(c, lo_all, hi_all, model == EQTL_MODEL_LMM));
    const size_t group = whole ? genes.size() : static_cast<size_t>(nb_max);
    for (size_t g0 = 0; g0 < genes.size(); g0 += group) {
      std::vector<ExactTask> tasks;
      for (size_t q = g0; q < std::min(genes.size(), g0 + group); ++q) tasks.push_back(task_of(genes[q], genes[q]));
      std::vector<double> mp;
      CHK(exact_minp(c, model, tasks, mp));
      for (size_t q = 0; q < tasks.size(); ++q) Tobs64[tasks[q].e] = -std::log10(std::max(std::min(1.0, mp[q]), 1e-300));
    }
  }
  int64_t n_exact = 0;
  DBuf mt, idx, gof, hof;
  int rc = EQTL_OK;
  for (size_t g0 = 0; g0 < genes.size() && rc == EQTL_OK; g0 += nb_max) {
    const int nb = static_cast<int>(std::min<size_t>(nb_max, genes.size() - g0));
    const int E = nb * B;
    std::vector<int32_t> go(genes.begin() + g0, genes.begin() + g0 + nb);
    std::vector<uint32_t> ho(nb);
    for (int q = 0; q < nb; ++q) ho[q] = gene_hash[go[q]];
    CHK(ensure(mt, sizeof(uint32_t) * static_cast<size_t>(E) * kMtN));
    CHK(ensure(idx, sizeof(int32_t) * static_cast<size_t>(E) * n));
    CHK(ensure(gof, sizeof(int32_t) * nb));
    CHK(ensure(hof, sizeof(uint32_t) * nb));
    HIPC(hipMemcpyAsync(gof.p, go.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(hof.p, ho.data(), sizeof(uint32_t) * nb, hipMemcpyHostToDevice, c->stream));
    // pseudo-gene phenotypes and their null fits
    if (model == EQTL_MODEL_LM) {
      CHK(alloc_gene_side(c, E, 1, 1));
      CHK(ensure(c->Y64, sizeof(double) * static_cast<size_t>(n) * E));
      CHK(ensure(c->U64, sizeof(double) * static_cast<size_t>(n) * E));
      hipLaunchKernelGGL(k_perm_pheno, dim3((E + 63) / 64), dim3(64), 0, c->stream, E, B, n, seed,
                         gof.as<int32_t>(), hof.as<uint32_t>(), mt.as<uint32_t>(), idx.as<int32_t>(), 0,
                         base.as<double>(), nullptr, nullptr, nullptr, c->Y64.as<double>());
      CHK(check_launch("k_perm_pheno"));
      hipLaunchKernelGGL(k_lm_gene, dim3(E), dim3(256), 0, c->stream, c->Y64.as<double>(), c->Wx.as<double>(), n,
                         c->n_pad, p, E, gene_cols(c), c->gscal.as<double>(), c->gvalid.as<int>(),
                         c->U64.as<double>());
      CHK(check_launch("k_lm_gene"));
      c->prepped_model = EQTL_MODEL_LM;
    } else {
      CHK(alloc_gene_side(c, E, p + 2, lmm_nq_pad(p + 2)));
      CHK(ensure(c->Yt64, sizeof(double) * static_cast<size_t>(n) * E));
      CHK(ensure(c->U64, sizeof(double) * static_cast<size_t>(n) * E));
      CHK(ensure(c->V64, sizeof(double) * static_cast<size_t>(n) * E));
      CHK(ensure(c->Lc, sizeof(double) * static_cast<size_t>(p) * p * E));
      hipLaunchKernelGGL(k_perm_pheno, dim3((E + 63) / 64), dim3(64), 0, c->stream, E, B, n, seed,
                         gof.as<int32_t>(), hof.as<uint32_t>(), mt.as<uint32_t>(), idx.as<int32_t>(), 1,
                         base.as<double>(), yt.as<double>(), dlt.as<double>(), c->lam64.as<double>(),
                         c->Yt64.as<double>());
      CHK(check_launch("k_perm_pheno"));
      CHK(lmm_prep_core(c, E, delta_mode, nullptr, nullptr, nullptr));
    }
    // windows of the pseudo-genes (their gene's)
    std::vector<int32_t> ech(E);
    std::vector<int64_t> ecs(E), ece(E);
    c->win_lo.assign(E, 0);
    c->win_hi.assign(E, 0);
    int64_t lo = INT64_MAX, hi = -1;
    for (int e = 0; e < E; ++e) {
      const int k = go[e / B];
      ech[e] = gch[k];
      ecs[e] = gcs[k];
      ece[e] = gce[k];
      c->win_lo[e] = glo[k];
      c->win_hi[e] = ghi[k];
      lo = std::min(lo, glo[k]);
      hi = std::max(hi, ghi[k]);
    }
    CHK(ensure(c->gene_chrom, sizeof(int32_t) * E));
    CHK(ensure(c->gene_cs, sizeof(int64_t) * E));
    CHK(ensure(c->gene_ce, sizeof(int64_t) * E));
    HIPC(hipMemcpyAsync(c->gene_chrom.p, ech.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(c->gene_cs.p, ecs.data(), sizeof(int64_t) * E, hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(c->gene_ce.p, ece.data(), sizeof(int64_t) * E, hipMemcpyHostToDevice, c->stream));
    c->have_windows = true;
    // min p of every pseudo-gene over its window (no hit output)
    CHK(scan_begin_impl(c, model, precision, EQTL_SCOPE_CIS, -1.0, true));
    for (int64_t j0 = lo; j0 < hi && rc == EQTL_OK; j0 += kCisBlock) {
      const int64_t mb = std::min<int64_t>(kCisBlock, hi - j0);
      int64_t nh = 0;
      rc = eqtl_gpu_scan_block(c, c->bed_res.as<uint8_t>() + j0 * c->bed_bps, 1, j0, mb, c->bed_bps, nullptr, 0,
                               &nh);
    }
    std::vector<eqtl_gene_state> st(E);
    if (rc == EQTL_OK) rc = eqtl_gpu_scan_end(c, st.data());
    else c->scanning = false;
    if (rc != EQTL_OK) break;
    // permutations whose statistic lies near the observed one: FP64 tests
    std::vector<ExactTask> tasks;
    for (int e = 0; e < E; ++e) {
      const int k = go[e / B], b = e % B;
      const double mp = std::min(1.0, st[e].best_p);
      minp[static_cast<size_t>(k) * B + b] = mp;
      const double Tb = -std::log10(std::max(mp, 1e-300));
      if (st[e].best_p <= 1.0 && std::fabs(Tb - Tobs64[k]) <= band * (1.0 + Tobs64[k])) tasks.push_back(task_of(e, k));
    }
    if (!tasks.empty()) {
      std::vector<double> mp;
      rc = exact_minp(c, model, tasks, mp);
      if (rc != EQTL_OK) break;
      for (size_t q = 0; q < tasks.size(); ++q) {
        const int e = tasks[q].e, k = go[e / B], b = e % B;
        if (mp[q] <= 1.0) minp[static_cast<size_t>(k) * B + b] = mp[q];
      }
      n_exact += static_cast<int64_t>(tasks.size());
    }
  }
  c->ex_valid = false;
  c->perm_exact = n_exact;
  // the gene side now holds the last batch: callers prepare genes again
  c->prepped_model = -1;
  c->have_windows = false;
  c->h_gene_chrom.clear();
  if (rc != EQTL_OK) return rc;
  for (int k = 0; k < G; ++k) {
    const bool done = std::isfinite(minp[static_cast<size_t>(k) * B]);
    if (p_emp) {
      if (!done) {
        p_emp[k] = std::numeric_limits<double>::quiet_NaN();
      } else {
        // p_emp_count (stats_extra.cpp:39-46) on T = -log10(max(p, 1e-300)),
        // with the observed T in FP64
        const double Tobs = Tobs64[k];
        int ge = 0;
        for (int b = 0; b < B; ++b)
          if (-std::log10(std::max(minp[static_cast<size_t>(k) * B + b], 1e-300)) >= Tobs) ++ge;
        p_emp[k] = (1.0 + ge) / (B + 1.0);
      }
    }
    if (perm_min_p)
      for (int b = 0; b < B; ++b) perm_min_p[static_cast<size_t>(k) * B + b] = minp[static_cast<size_t>(k) * B + b];
  }
  return EQTL_OK;
}

int eqtl_gpu_beta_approx(eqtl_gpu_ctx* c, const double* perm_min_p, const double* obs_min_p, int32_t G, int32_t B,
                         double* p_beta, double* shape1, double* shape2) {
  if (!c || !perm_min_p || !obs_min_p || G <= 0 || B <= 0 || !p_beta)
    return fail(EQTL_E_ARG, "bad beta_approx arguments");
  CHK(set_device(c));
  DBuf pm, ob, out;
  CHK(ensure(pm, sizeof(double) * static_cast<size_t>(G) * B));
  CHK(ensure(ob, sizeof(double) * G));
  CHK(ensure(out, sizeof(double) * 3 * static_cast<size_t>(G)));
  HIPC(hipMemcpyAsync(pm.p, perm_min_p, sizeof(double) * static_cast<size_t>(G) * B, hipMemcpyHostToDevice,
                      c->stream));
  HIPC(hipMemcpyAsync(ob.p, obs_min_p, sizeof(double) * G, hipMemcpyHostToDevice, c->stream));
  double* o = out.as<double>();
  hipLaunchKernelGGL(k_beta_approx, dim3((G + 63) / 64), dim3(64), 0, c->stream, pm.as<double>(), ob.as<double>(), G,
                     B, o, o + G, o + 2 * G);
  CHK(check_launch("k_beta_approx"));
  HIPC(hipMemcpyAsync(p_beta, o, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
  if (shape1) HIPC(hipMemcpyAsync(shape1, o + G, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
  if (shape2) HIPC(hipMemcpyAsync(shape2, o + 2 * G, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return EQTL_OK;
}

// ---- GRM ----------------------------------------------------------------------
int eqtl_gpu_grm_begin(eqtl_gpu_ctx* c) {
  if (!c) return fail(EQTL_E_ARG, "null ctx");
  if (c->n <= 0) return fail(EQTL_E_STATE, "set_samples first");
  if (c->subset) return fail(EQTL_E_STATE, "the GRM is built over the analysis samples: clear the test subset");
  CHK(set_device(c));
  const size_t nn = static_cast<size_t>(c->n_pad) * c->n_pad;
  CHK(ensure(c->grm_acc, nn * sizeof(double)));
  CHK(ensure(c->grm_miss, nn * sizeof(int32_t)));
  CHK(ensure(c->grm_cnt, static_cast<size_t>(c->n_pad) * sizeof(int32_t)));
  CHK(ensure(c->grm_u, (static_cast<size_t>(c->n_pad) + 1) * sizeof(double)));
  HIPC(hipMemsetAsync(c->grm_acc.p, 0, nn * sizeof(double), c->stream));
  HIPC(hipMemsetAsync(c->grm_miss.p, 0, nn * sizeof(int32_t), c->stream));
  HIPC(hipMemsetAsync(c->grm_cnt.p, 0, static_cast<size_t>(c->n_pad) * sizeof(int32_t), c->stream));
  HIPC(hipMemsetAsync(c->grm_u.p, 0, (static_cast<size_t>(c->n_pad) + 1) * sizeof(double), c->stream));
  c->grm_m = 0;
  c->grm_ex_n = 0;
  return EQTL_OK;
}

}  // extern "C"

namespace {

// Packed rows of the exact buffer: 2-bit dosage codes, 16 samples per dword.
int grm_rowq(const eqtl_gpu_ctx* c) { return c->n_pad / 16; }

// The grouped int8 Gram over the buffered exact SNPs (then the buffer is
// empty): allele-count histogram, group starts and per-step weights, the
// grouped order, then per chunk of steps the int8 operand, its u terms and
// the Gram launch.
int grm_exact_flush(eqtl_gpu_ctx* c) {
  const int64_t cnt = c->grm_ex_n;
  if (cnt == 0) return EQTL_OK;
  const int n = c->n, np = c->n_pad, rowq = grm_rowq(c);
  const int nkeys = n + 1;  // minor allele counts 1 .. n
  CHK(ensure(c->grm_hist, sizeof(int32_t) * nkeys));
  CHK(ensure(c->grm_fill, sizeof(int32_t) * nkeys));
  CHK(ensure(c->grm_gstart, sizeof(int64_t) * (nkeys + 1)));
  const int64_t max_steps = cnt / 64 + nkeys + 1;
  CHK(ensure(c->grm_wstep, sizeof(double) * 2 * max_steps));
  CHK(ensure(c->grm_order, sizeof(int32_t) * max_steps * 64));
  int32_t* hist = c->grm_hist.as<int32_t>();
  int64_t* gstart = c->grm_gstart.as<int64_t>();
  double* wstep = c->grm_wstep.as<double>();
  double* wmu = wstep + max_steps;
  int64_t* nsteps_d = gstart + nkeys;
  const unsigned fill_blocks = static_cast<unsigned>(std::min<int64_t>(4096, (max_steps * 64 + 255) / 256));
  hipLaunchKernelGGL(k_fill_i32_gs, dim3(256), dim3(256), 0, c->stream, hist, static_cast<int64_t>(nkeys), 0);
  hipLaunchKernelGGL(k_fill_i32_gs, dim3(fill_blocks), dim3(256), 0, c->stream, c->grm_order.as<int32_t>(),
                     max_steps * 64, -1);
  CHK(check_launch("k_fill_i32_gs"));
  const unsigned rb = static_cast<unsigned>((cnt + 255) / 256);
  hipLaunchKernelGGL(k_grm_hist, dim3(rb), dim3(256), 0, c->stream, c->grm_ex_key.as<int32_t>(), cnt, hist);
  CHK(check_launch("k_grm_hist"));
  hipLaunchKernelGGL(k_grm_groups, dim3(1), dim3(1024), 0, c->stream, hist, nkeys, n, gstart,
                     c->grm_fill.as<int32_t>(), wstep, wmu, nsteps_d, c->grm_u.as<double>() + np);
  CHK(check_launch("k_grm_groups"));
  hipLaunchKernelGGL(k_grm_scatter, dim3(rb), dim3(256), 0, c->stream, c->grm_ex_key.as<int32_t>(), cnt, gstart,
                     c->grm_fill.as<int32_t>(), c->grm_order.as<int32_t>());
  CHK(check_launch("k_grm_scatter"));
  int64_t nsteps = 0;
  HIPC(hipMemcpyAsync(&nsteps, nsteps_d, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  if (nsteps <= 0 || nsteps > max_steps) return fail(EQTL_E_HIP, "grm: bad step count of the grouped Gram");
  const int64_t chunk = std::max<int64_t>(
      16, std::min<int64_t>(round_up(nsteps, 16), static_cast<int64_t>(c->grm_chunk_bytes / (static_cast<size_t>(np) * 64)) / 16 * 16));
  CHK(ensure(c->grm_X, static_cast<size_t>(np) * chunk * 64));
  CHK(ensure(c->grm_upart, sizeof(double) * static_cast<size_t>(chunk / 16) * np));
  const int nt = (np + kI8T - 1) / kI8T;
  // tile list: lower-triangle tiles (by << 16 | bx) in blocks of 4 tile rows x
  // 8 tile columns, row blocks outer
  std::vector<int32_t> tl;
  for (int R = 0; R < nt; R += 4)
    for (int C = 0; C <= R + 3 && C < nt; C += 8)
      for (int by = R; by < std::min(nt, R + 4); ++by)
        for (int bx = C; bx < std::min(C + 8, by + 1); ++bx) tl.push_back((by << 16) | bx);
  const int ntiles = static_cast<int>(tl.size());
  CHK(ensure(c->grm_tiles, sizeof(int32_t) * tl.size()));
  HIPC(hipMemcpy(c->grm_tiles.p, tl.data(), sizeof(int32_t) * tl.size(), hipMemcpyHostToDevice));
  const unsigned i8_grid = static_cast<unsigned>(256 * ((ntiles + 255) / 256));
  static bool i8_attr = false;
  if (!i8_attr) {
    HIPC(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_grm_i8), hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(kI8LdsBytes)));
    i8_attr = true;
  }
  for (int64_t s0 = 0; s0 < nsteps; s0 += chunk) {
    const int64_t s1 = std::min(nsteps, s0 + chunk);
    const unsigned nb = static_cast<unsigned>((s1 - s0 + 15) / 16);
    {
      Timed t(c, "grm_expand");
      hipLaunchKernelGGL(k_grm_expand, dim3(nb, static_cast<unsigned>((np + 255) / 256)), dim3(256), 0, c->stream,
                         c->grm_ex_rows.as<uint32_t>(), rowq, c->grm_order.as<int32_t>(), s0, s1, wmu, np,
                         c->grm_X.as<int8_t>(), c->grm_upart.as<double>());
      CHK(check_launch("k_grm_expand"));
      hipLaunchKernelGGL(k_grm_reduce_u, dim3(static_cast<unsigned>((np + 255) / 256)), dim3(256), 0, c->stream,
                         c->grm_upart.as<double>(), static_cast<int>(nb), np, c->grm_u.as<double>());
      CHK(check_launch("k_grm_reduce_u"));
      CHK(t.done());
    }
    Timed t(c, "grm_i8");
    hipLaunchKernelGGL(k_grm_i8, dim3(i8_grid), dim3(64 * kI8Waves), kI8LdsBytes, c->stream, c->grm_X.as<int8_t>(),
                       static_cast<int>(s1 - s0), np, c->grm_tiles.as<int32_t>(), ntiles, wstep + s0,
                       c->grm_acc.as<double>());
    CHK(check_launch("k_grm_i8"));
    CHK(t.done());
  }
  c->grm_ex_n = 0;
  return EQTL_OK;
}

// Room for `more` exact rows in the exact buffer: grows it (keeping its rows)
// up to grm_buf_bytes, else runs the grouped Gram over what it holds.
int grm_exact_reserve(eqtl_gpu_ctx* c, int64_t more) {
  const size_t rowb = static_cast<size_t>(grm_rowq(c)) * 4;
  const int64_t need = c->grm_ex_n + more;
  if (static_cast<int64_t>(c->grm_ex_rows.bytes / rowb) >= need &&
      static_cast<int64_t>(c->grm_ex_key.bytes / sizeof(int32_t)) >= need)
    return EQTL_OK;
  const int64_t cap = std::max<int64_t>(more, static_cast<int64_t>(c->grm_buf_bytes / rowb));
  if (need > cap) {
    CHK(grm_exact_flush(c));
    return grm_exact_reserve(c, more);
  }
  const int64_t rows_new = std::min(cap, std::max(need, 2 * static_cast<int64_t>(c->grm_ex_rows.bytes / rowb)));
  DBuf r2, k2;
  CHK(ensure(r2, static_cast<size_t>(rows_new) * rowb));
  CHK(ensure(k2, static_cast<size_t>(rows_new) * sizeof(int32_t)));
  if (c->grm_ex_n > 0) {
    HIPC(hipMemcpyAsync(r2.p, c->grm_ex_rows.p, static_cast<size_t>(c->grm_ex_n) * rowb, hipMemcpyDeviceToDevice,
                        c->stream));
    HIPC(hipMemcpyAsync(k2.p, c->grm_ex_key.p, static_cast<size_t>(c->grm_ex_n) * sizeof(int32_t),
                        hipMemcpyDeviceToDevice, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
  }
  std::swap(c->grm_ex_rows.p, r2.p);
  std::swap(c->grm_ex_rows.bytes, r2.bytes);
  std::swap(c->grm_ex_key.p, k2.p);
  std::swap(c->grm_ex_key.bytes, k2.bytes);
  return EQTL_OK;
}

// Per-SNP arrays of one block (class, allele count, mean, sd, position) and
// the block counters, carved from `base`.
struct GrmSnpArrays {
  int64_t* counts;
  double* mean;
  double* sd;
  int32_t* ac;
  int32_t* pos;
  int8_t* cls;
};
size_t grm_snp_bytes(int64_t m) { return 256 + static_cast<size_t>(round_up(m, 64)) * (8 + 8 + 4 + 4 + 1); }
GrmSnpArrays grm_snp_arrays(void* base, int64_t m) {
  const size_t mr = static_cast<size_t>(round_up(m, 64));
  unsigned char* p = static_cast<unsigned char*>(base);
  GrmSnpArrays a;
  a.counts = reinterpret_cast<int64_t*>(p);
  a.mean = reinterpret_cast<double*>(p + 256);
  a.sd = a.mean + mr;
  a.ac = reinterpret_cast<int32_t*>(a.sd + mr);
  a.pos = a.ac + mr;
  a.cls = reinterpret_cast<int8_t*>(a.pos + mr);
  return a;
}

// One block of the GRM accumulation over rows already on the device. z64 and
// mis (m_pad x n_pad each, sample-major) and the per-SNP arrays are
// caller-provided scratch. With exact_ok, SNPs whose calls are all present and
// integral go to the exact buffer (the grouped int8 Gram, run when the buffer
// fills and at grm_finish); the others, and every SNP without exact_ok, go
// through the FP64 Gram of this block.
int grm_accumulate_core(eqtl_gpu_ctx* c, const uint8_t* rd, int64_t m, int32_t bps, double* z64, __bf16* mis,
                        const GrmSnpArrays& sa, bool exact_ok) {
  const int np = c->n_pad;
  const bool dose = c->geno_fmt == EQTL_GENO_DOSE32;
  const bool ident = !dose && c->col_ident;
  auto* kst = dose ? &k_grm_stats<EQTL_GENO_DOSE32, false>
                   : (ident ? &k_grm_stats<EQTL_GENO_BED, true> : &k_grm_stats<EQTL_GENO_BED, false>);
  hipLaunchKernelGGL(kst, dim3(static_cast<unsigned>((m + 3) / 4)), dim3(256), 0, c->stream, rd, m, bps, c->n,
                     c->sample_col.as<int32_t>(), c->impute, c->max_miss, c->maf_min, exact_ok ? 1 : 0, sa.cls,
                     sa.ac, sa.mean, sa.sd);
  CHK(check_launch("k_grm_stats"));
  hipLaunchKernelGGL(k_grm_scan, dim3(1), dim3(1024), 0, c->stream, sa.cls, m, sa.pos, sa.counts);
  CHK(check_launch("k_grm_scan"));
  int64_t cn[4] = {0, 0, 0, 0};
  HIPC(hipMemcpyAsync(cn, sa.counts, sizeof(cn), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  if (cn[3] > 0) CHK(grm_exact_reserve(c, cn[3]));
  const int64_t zstride = cn[2] > 0 ? std::max<int64_t>(128, round_up(cn[2], 128)) : 0;
  auto* kem = dose ? &k_grm_emit<EQTL_GENO_DOSE32, false>
                   : (ident ? &k_grm_emit<EQTL_GENO_BED, true> : &k_grm_emit<EQTL_GENO_BED, false>);
  hipLaunchKernelGGL(kem, dim3(static_cast<unsigned>((m + 128 + 3) / 4)), dim3(256), 0, c->stream, rd, m, bps, c->n,
                     np, c->sample_col.as<int32_t>(), sa.cls, sa.pos, sa.ac, sa.counts, sa.mean, sa.sd, z64, mis,
                     zstride, c->grm_cnt.as<int32_t>(), c->grm_ex_rows.as<uint32_t>(), c->grm_ex_key.as<int32_t>(),
                     c->grm_ex_n, grm_rowq(c));
  CHK(check_launch("k_grm_emit"));
  c->grm_ex_n += cn[3];
  if (cn[2] > 0) {
    Timed t(c, "grm");
    static bool attr_set = false;
    if (!attr_set) {
      HIPC(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_grm_syrk), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(kGramLdsBytes)));
      attr_set = true;
    }
    const int zs = static_cast<int>(zstride);
    if (np >= kGramLargeN) {
      const unsigned nt = static_cast<unsigned>((np + kGramT - 1) / kGramT);
      hipLaunchKernelGGL(k_grm_syrk, dim3(nt, nt), dim3(256), kGramLdsBytes, c->stream, z64, zs, np,
                         c->grm_acc.as<double>());
      CHK(check_launch("k_grm_syrk"));
      if (cn[1]) {
        dim3 grid(static_cast<unsigned>(np / 64), static_cast<unsigned>(np / 64));
        hipLaunchKernelGGL(k_grm_joint, grid, dim3(256), 0, c->stream, mis, zs, np, c->grm_miss.as<int32_t>());
        CHK(check_launch("k_grm_joint"));
      }
    } else {
      dim3 grid(static_cast<unsigned>(np / 64), static_cast<unsigned>(np / 64));
      hipLaunchKernelGGL(k_grm_gram, grid, dim3(256), 0, c->stream, z64, mis, zs, np, c->grm_acc.as<double>(),
                         c->grm_miss.as<int32_t>(), cn[1] ? 1 : 0);
      CHK(check_launch("k_grm_gram"));
    }
    CHK(t.done());
  }
  c->grm_m += cn[0];
  return EQTL_OK;
}

}  // namespace

extern "C" {

int eqtl_gpu_grm_accumulate(eqtl_gpu_ctx* c, const uint8_t* rows, int32_t on_dev, int64_t m, int32_t bps) {
  if (!c || !rows || m <= 0) return fail(EQTL_E_ARG, "bad grm_accumulate arguments");
  if (!c->grm_acc.p) return fail(EQTL_E_STATE, "grm_begin first");
  if (bps < min_bps(c)) return fail(EQTL_E_ARG, "bytes_per_snp too small for n_file");
  CHK(set_device(c));
  CHK(ensure_block(c, m, bps, false));
  const int64_t m_pad = round_up(m, 128);
  const int np = c->n_pad;
  CHK(ensure(c->grm_z, static_cast<size_t>(m_pad) * np * (sizeof(double) + sizeof(__bf16))));
  CHK(ensure(c->grm_snp, grm_snp_bytes(m)));
  int rc;
  const uint8_t* rd = stage_rows(c, rows, on_dev, m, bps, &rc);
  if (rc) return rc;
  double* z64 = c->grm_z.as<double>();
  __bf16* mis = reinterpret_cast<__bf16*>(z64 + static_cast<size_t>(m_pad) * np);
  return grm_accumulate_core(c, rd, m, bps, z64, mis, grm_snp_arrays(c->grm_snp.p, m), true);
}

int eqtl_gpu_grm_accumulate_noalloc(eqtl_gpu_ctx* c, const uint8_t* rows, int32_t on_dev, int64_t m, int32_t bps) {
  if (!c || !rows || m <= 0) return fail(EQTL_E_ARG, "bad grm_accumulate_noalloc arguments");
  if (!c->grm_acc.p) return fail(EQTL_E_STATE, "grm_begin first");
  if (bps < min_bps(c)) return fail(EQTL_E_ARG, "bytes_per_snp too small for n_file");
  if (!c->scratch) return fail(EQTL_E_STATE, "the device scratch array is not mapped for this device");
  const int64_t m_pad = round_up(m, 128);
  const size_t np = static_cast<size_t>(c->n_pad);
  const size_t stage = on_dev ? 0 : static_cast<size_t>(round_up(m * static_cast<int64_t>(bps), 256));
  const size_t snp = round_up(static_cast<int64_t>(grm_snp_bytes(m)), 256);
  const size_t need = snp + stage + static_cast<size_t>(m_pad) * np * (sizeof(double) + sizeof(__bf16));
  if (need > kScratchBytes)
    return fail(EQTL_E_CAPACITY, "grm_accumulate_noalloc: " + std::to_string(m) + " SNPs x " +
                                     std::to_string(c->n) + " samples need " + std::to_string(need) +
                                     " bytes of scratch; the fixed scratch holds " + std::to_string(kScratchBytes) +
                                     " (pass fewer SNPs per call)");
  CHK(set_device(c));
  unsigned char* sp = static_cast<unsigned char*>(c->scratch);
  const uint8_t* rd = rows;
  if (!on_dev) {
    HIPC(hipMemcpyAsync(sp + snp, rows, static_cast<size_t>(m) * bps, hipMemcpyHostToDevice, c->stream));
    rd = sp + snp;
  }
  double* z64 = reinterpret_cast<double*>(sp + snp + stage);
  __bf16* mis = reinterpret_cast<__bf16*>(z64 + static_cast<size_t>(m_pad) * np);
  // every array the kernels read is written by this call first (class and
  // position of each SNP, the counters by the scan), so the previous call's
  // contents do not leak in; the exact buffer is not used (it allocates)
  return grm_accumulate_core(c, rd, m, bps, z64, mis, grm_snp_arrays(sp, m), false);
}

int eqtl_gpu_grm_set_limits(eqtl_gpu_ctx* c, int64_t exact_buffer_bytes, int64_t chunk_bytes) {
  if (!c || exact_buffer_bytes < 0 || chunk_bytes < 0) return fail(EQTL_E_ARG, "bad grm_set_limits arguments");
  if (c->grm_ex_n > 0) return fail(EQTL_E_STATE, "grm_set_limits: exact SNPs are held; set the limits before accumulating");
  if (exact_buffer_bytes) c->grm_buf_bytes = static_cast<size_t>(exact_buffer_bytes);
  if (chunk_bytes) c->grm_chunk_bytes = static_cast<size_t>(chunk_bytes);
  return EQTL_OK;
}

int eqtl_gpu_grm_finish(eqtl_gpu_ctx* c, double* K, int32_t* N_packed, int64_t* m_used) {
  if (!c) return fail(EQTL_E_ARG, "null ctx");
  if (!c->grm_acc.p) return fail(EQTL_E_STATE, "grm_begin first");
  if (c->grm_m == 0) return fail(EQTL_E_ARG, "no SNPs for GRM");
  CHK(set_device(c));
  CHK(grm_exact_flush(c));
  const int n = c->n, np = c->n_pad;
  DBuf Kd, Nd;
  CHK(ensure(Kd, sizeof(double) * static_cast<size_t>(n) * n));
  CHK(ensure(Nd, sizeof(int32_t) * static_cast<size_t>(n) * (n + 1) / 2));
  hipLaunchKernelGGL(k_grm_finish, dim3((n + 15) / 16, (n + 15) / 16), dim3(16, 16), 0, c->stream,
                     c->grm_acc.as<double>(), c->grm_miss.as<int32_t>(), c->grm_cnt.as<int32_t>(),
                     c->grm_u.as<double>(), c->grm_u.as<double>() + np, n, np, static_cast<int>(c->grm_m),
                     Kd.as<double>(), Nd.as<int32_t>());
  CHK(check_launch("k_grm_finish"));
  if (K) HIPC(hipMemcpyAsync(K, Kd.p, sizeof(double) * static_cast<size_t>(n) * n, hipMemcpyDeviceToHost, c->stream));
  if (N_packed)
    HIPC(hipMemcpyAsync(N_packed, Nd.p, sizeof(int32_t) * static_cast<size_t>(n) * (n + 1) / 2, hipMemcpyDeviceToHost,
                        c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  if (m_used) *m_used = c->grm_m;
  return EQTL_OK;
}

// ---- instrumentation / synthetic inputs ----------------------------------------
int eqtl_gpu_set_timing(eqtl_gpu_ctx* c, int32_t enable) {
  if (!c) return fail(EQTL_E_ARG, "null ctx");
  c->timing = enable != 0;
  if (!c->timing) c->tslots.clear();
  return EQTL_OK;
}

int eqtl_gpu_get_timing(eqtl_gpu_ctx* c, const char* fam, double* ms, int64_t* launches) {
  if (!c || !fam) return fail(EQTL_E_ARG, "bad get_timing arguments");
  auto it = c->tslots.find(fam);
  if (ms) *ms = it == c->tslots.end() ? 0.0 : it->second.ms;
  if (launches) *launches = it == c->tslots.end() ? 0 : it->second.launches;
  return EQTL_OK;
}

int eqtl_gpu_scan_lr_plan(eqtl_gpu_ctx* c, int32_t* low_rank, int32_t* n_super_groups, int32_t* n_segments,
                          double* mean_nodes) {
  if (!c) return fail(EQTL_E_ARG, "bad scan_lr_plan arguments");
  const bool on = c->lr_on;
  if (low_rank) *low_rank = on ? 1 : 0;
  if (n_super_groups) *n_super_groups = on ? c->lr_n_sg : 0;
  if (n_segments) *n_segments = on ? c->lr_nseg : 0;
  if (mean_nodes) *mean_nodes = on ? c->lr_mean_nodes : 0.0;
  return EQTL_OK;
}

int eqtl_gpu_synth_bed(eqtl_gpu_ctx* c, uint64_t seed, int64_t m, int32_t n, double maf_lo, double maf_hi,
                       double miss, uint8_t** rows_dev) {
  if (!c || !rows_dev || m <= 0 || n <= 0) return fail(EQTL_E_ARG, "bad synth_bed arguments");
  CHK(set_device(c));
  const int bps = (n + 3) / 4;
  auto* b = new DBuf();
  const int r = ensure(*b, static_cast<size_t>(m) * bps);
  if (r) {
    delete b;
    return r;
  }
  c->synth.push_back(b);
  const int64_t total = m * bps;
  hipLaunchKernelGGL(k_synth_bed, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, c->stream,
                     b->as<uint8_t>(), m, n, bps, seed, maf_lo, maf_hi, miss);
  CHK(check_launch("k_synth_bed"));
  HIPC(hipStreamSynchronize(c->stream));
  *rows_dev = b->as<uint8_t>();
  return EQTL_OK;
}

int eqtl_gpu_selftest(eqtl_gpu_ctx* c, int32_t* failed_mask) {
  if (!c || !failed_mask) return fail(EQTL_E_ARG, "bad selftest arguments");
  CHK(set_device(c));
  DBuf buf;
  CHK(ensure(buf, sizeof(int) * (16 + 1024)));
  std::vector<int> h(16 + 1024, 0);
  for (int i = 0; i < 1024; ++i) h[16 + i] = i * 7 + 3;
  HIPC(hipMemcpyAsync(buf.p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_selftest, dim3(1), dim3(256), 0, c->stream, buf.as<int>());
  CHK(check_launch("k_selftest"));
  HIPC(hipMemcpyAsync(h.data(), buf.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  *failed_mask = h[0];
  return EQTL_OK;
}

int eqtl_gpu_copy_rows(eqtl_gpu_ctx* c, const uint8_t* rows_dev, int64_t j0, int64_t m, int32_t bps, uint8_t* out) {
  if (!c || !rows_dev || !out || m <= 0 || j0 < 0) return fail(EQTL_E_ARG, "bad copy_rows arguments");
  CHK(set_device(c));
  HIPC(hipMemcpyAsync(out, rows_dev + j0 * bps, static_cast<size_t>(m) * bps, hipMemcpyDeviceToHost, c->stream));
  HIPC(hipStreamSynchronize(c->stream));
  return EQTL_OK;
}

}  // extern "C"
