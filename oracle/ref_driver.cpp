// Thin C-ABI driver over the reference's own translation units, built by
// oracle/Makefile (target `ref`) from the unmodified sources under the
// reference tree: src/model_lm.cpp, src/model_lmm.cpp, src/util.cpp,
// src/plink_bed.cpp, src/grm.cpp with its vendored Eigen 3.4 headers.
// OpenBLAS is absent, so Eigen's own GEMM/eigensolver stand in for
// -DEIGEN_USE_BLAS/-DEIGEN_USE_LAPACKE (last-bit differences only).
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle restatement and
// by bench.py's cpu_baseline leg; never linked into the product.
#include "eqtl/grm.hpp"
#include "eqtl/models.hpp"
#include "eqtl/plink_bed.hpp"
#include "eqtl/util.hpp"

#include <cstring>
#include <exception>
#include <string>
#include <vector>

using namespace eqtl;

extern "C" {

double ref_p_from_t(double t, double df) { return p_from_t(t, df); }

// make_lmm_basis (model_lmm.cpp:106-113): K n x n col-major → Q, lambda.
int ref_make_lmm_basis(const double* K, int n, double* Q, double* lam) {
  try {
    Eigen::Map<const Eigen::MatrixXd> Km(K, n, n);
    LmmBasis b = make_lmm_basis(Km);
    std::memcpy(Q, b.Q.data(), sizeof(double) * n * n);
    std::memcpy(lam, b.lambda.data(), sizeof(double) * n);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// prep_lmm (model_lmm.cpp:124-153) + test_lmm (model_lmm.cpp:160-223) for c
// raw genotype vectors G (c x n, vector-major). out: c x 5 (beta se stat p r2).
// prep_out: [delta, rss_null, y_dy, has_a00].
int ref_lmm(const double* y, const double* X, int n, int p, const double* Q, const double* lam,
            const double* G, int c, double* out, double* prep_out) {
  try {
    LmmBasis basis;
    basis.Q = Eigen::Map<const Eigen::MatrixXd>(Q, n, n);
    basis.lambda = Eigen::Map<const Eigen::VectorXd>(lam, n);
    Eigen::Map<const Eigen::VectorXd> yv(y, n);
    Eigen::Map<const Eigen::MatrixXd> Xm(X, n, p);
    GenePrepLmm P = prep_lmm(yv, Xm, basis);
    prep_out[0] = P.delta;
    prep_out[1] = P.rss_null;
    prep_out[2] = P.y_dy;
    prep_out[3] = P.has_a00 ? 1.0 : 0.0;
    for (int s = 0; s < c; ++s) {
      Eigen::Map<const Eigen::VectorXd> g(G + static_cast<size_t>(s) * n, n);
      AssocHit h = test_lmm(P, g);
      out[s * 5 + 0] = h.beta;
      out[s * 5 + 1] = h.se;
      out[s * 5 + 2] = h.stat;
      out[s * 5 + 3] = h.p;
      out[s * 5 + 4] = h.r2;
    }
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// prep_lm + test_lm (model_lm.cpp:19-57). out: c x 5.
int ref_lm(const double* y, const double* X, int n, int p, const double* G, int c, double* out) {
  try {
    Eigen::Map<const Eigen::VectorXd> yv(y, n);
    Eigen::Map<const Eigen::MatrixXd> Xm(X, n, p);
    GenePrepLm P = prep_lm(yv, Xm);
    for (int s = 0; s < c; ++s) {
      Eigen::Map<const Eigen::VectorXd> g(G + static_cast<size_t>(s) * n, n);
      AssocHit h = test_lm(P, g);
      out[s * 5 + 0] = h.beta;
      out[s * 5 + 1] = h.se;
      out[s * 5 + 2] = h.stat;
      out[s * 5 + 3] = h.p;
      out[s * 5 + 4] = h.r2;
    }
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// PlinkBed decode (plink_bed.cpp:214-357) of a bed/bim/fam prefix in fam
// order. Writes up to cap SNPs: index into the bim, af, dosage rows (n each).
// Returns the number of SNPs that passed, or -1 on error.
long ref_decode_bed(const char* prefix, int impute, double max_miss, double maf_min, long cap,
                    long* idx, double* af, double* dosage, int n) {
  try {
    PlinkBed bed;
    bed.open(prefix);
    std::vector<std::string> samples = bed.samples();
    if (static_cast<int>(samples.size()) != n) return -1;
    bed.set_sample_order(samples);
    MissPolicy mp{impute ? MissHand::Impute : MissHand::Filter, max_miss};
    const auto& sites = bed.sites();
    long k = 0;
    bed.for_each_snp(mp, maf_min, [&](const SnpRec& s) {
      if (k < cap) {
        // locate by id (ids are unique in generator output)
        long j = k == 0 ? 0 : idx[k - 1] + 1;
        while (j < static_cast<long>(sites.size()) && sites[static_cast<size_t>(j)].id != s.id) ++j;
        idx[k] = j;
        af[k] = s.maf;
        std::memcpy(dosage + static_cast<size_t>(k) * n, s.dosage.data(), sizeof(double) * n);
      }
      ++k;
      return true;
    });
    return k;
  } catch (const std::exception&) {
    return -1;
  }
}

// compute_grm over a bed prefix (grm.cpp:112-224, 236-242). K n x n col-major,
// N_packed n(n+1)/2.
int ref_compute_grm(const char* prefix, int impute, double max_miss, double maf_min, int n,
                    double* K, int* N_packed) {
  try {
    PlinkBed bed;
    bed.open(prefix);
    std::vector<std::string> samples = bed.samples();
    if (static_cast<int>(samples.size()) != n) return -1;
    MissPolicy mp{impute ? MissHand::Impute : MissHand::Filter, max_miss};
    Grm g = compute_grm(bed, samples, mp, maf_min);
    std::memcpy(K, g.K.data(), sizeof(double) * n * n);
    std::memcpy(N_packed, g.N_packed.data(), sizeof(int) * g.N_packed.size());
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

}  // extern "C"
