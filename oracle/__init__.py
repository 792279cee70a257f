"""FP64 oracle for the eqtl association engine — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker (or the timed CPU baseline). The product path
(eqtl_amd, its HIP library and host code) never imports, links or calls it.

Pinning of each oracle function (besides being a line-by-line restatement of
the cited reference code):
  p_from_t / betai   scipy.special.stdtr/betainc and mpmath at 50 digits down
                     to p=1e-300; the reference's own util.cpp via _ref.
  decode_bed         PLINK-1 code table as a known-answer fixture, brute-force
                     unpack of generator beds (all codes, n % 4 != 0, padding,
                     missing + impute/filter, max_miss edge); the reference's
                     PlinkBed via _ref.
  subset_maf_or_nan  closed form on hand-built vectors (monomorphic, var<1e-12).
  reml_negll         numpy closed form (slogdet of X^T V^-1 X etc.).
  optimize_delta     the reference's own model_lmm.cpp via _ref (exact delta);
                     scipy minimize_scalar on log delta (true argmin).
  prep/test_lmm      brute-force GLS with V = delta*K + I and a full (p+1)
                     solve in numpy (textbook); the reference via _ref.
  prep/test_lm       OLS closed form (numpy lstsq); the reference via _ref.
  trans scans        composition of the above; hit/ACAT/best semantics checked
                     on hand-built p vectors.
  acat               mpmath closed form (Liu & Xie 2020) at 40 digits.
  p_emp, bh          exact arithmetic on hand-built inputs.
  grm                GCTA closed form in numpy (eqtl_amd.synth.grm_closed_form);
                     the reference's grm.cpp via _ref.
  eigh (Jacobi)      numpy.linalg.eigh; ||QΛQ^T-K||, ||Q^TQ-I||.
  perm_index         libstdc++ std::mt19937 + std::shuffle called directly, and
                     the committed fixture tests/golden/perm_idx.json.
  lnbeta             scipy.special.betaln over the fit's shape box.
  beta_approx        moment fit: numpy closed form on cases where the simplex
                     must fail (start outside the shape bounds); ML refinement:
                     NLL at the result within the simplex tolerance of the
                     bounded optimum found by scipy.optimize; p_beta equals
                     scipy.special.betainc at the returned shapes; filtering
                     and the <10-minima NaN rule on hand-built vectors. GSL's
                     nmsimplex2 itself is absent from the image (the reference
                     links the system libgsl), so the simplex path is pinned
                     by optimality, not step-by-step.
  vcf_record         the reference's own VCF fixtures (data/smoke.vcf.gz,
                     tests/golden/smoke/): for every record the ALT dosage sum
                     and the called allele copies equal the INFO AC and AN the
                     producing tool wrote; hand-built lines with known answers
                     (haploid chromosomes and calls, '.'-alleles, DS with '.',
                     short columns, ALT '.' and multi-allelic skips, policies);
                     diploid GT lines equal orc_decode_bed on the PLINK codes
                     of the same calls. htslib (the reference's VCF reader) is
                     absent: its float parse is taken as strtod -> float.
  glm_prep/glm_test  NB2 score test (model_glm.cpp): the score U is the
                     derivative of the NB2 log-likelihood (scipy gammaln form)
                     in the SNP coefficient at the null fit, by finite
                     differences; the information equals 1 / [(J^-1)_gg] of the
                     full (p+1) x (p+1) information J = [X g]^T W [X g]
                     (Schur form checked against a plain inverse); the null fit
                     solves the NB2 score equations X^T (y - mu)/(1 + phi mu)
                     = 0 and phi is the MoM fixed point; phi -> 0 data
                     (Poisson-like) agree with a Poisson GLM fit by
                     scipy.optimize.
  cis_gene           composition of the above (LM/LMM perm bases restated from
                     scan_cis.hpp:74-160).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libeqtl_ref.so")
_lib = None
_ref = None

c_double_p = ctypes.POINTER(ctypes.c_double)


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and, when the reference tree exists, _ref/)."""
    subprocess.check_call(["make", "-s", "-C", _HERE, "all"])
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build(ref=False)
        _lib = ctypes.CDLL(_LIB)
        _lib.orc_p_from_t.restype = ctypes.c_double
        _lib.orc_betai.restype = ctypes.c_double
        _lib.orc_lnbeta.restype = ctypes.c_double
        _lib.orc_acat.restype = ctypes.c_double
        _lib.orc_p_emp.restype = ctypes.c_double
        _lib.orc_subset_maf_or_nan.restype = ctypes.c_double
        _lib.orc_reml_negll.restype = ctypes.c_double
        _lib.orc_trans_lmm_scan.restype = ctypes.c_int64
        _lib.orc_trans_lm_scan.restype = ctypes.c_int64
        _lib.orc_grm.restype = ctypes.c_int64
        _lib.orc_glm_scan.restype = ctypes.c_int64
        _lib.orc_fnv1a.restype = ctypes.c_uint32
        _lib.orc_cis_gene.restype = ctypes.c_double
        _lib.orc_num_threads.restype = ctypes.c_int32
    return _lib


def ref_lib():
    """The reference's own TUs (None when not built, e.g. on a machine without the tree)."""
    global _ref
    if _ref is None:
        if not os.path.exists(_REF):
            return None
        _ref = ctypes.CDLL(_REF, mode=os.RTLD_LAZY)
        _ref.ref_p_from_t.restype = ctypes.c_double
        _ref.ref_decode_bed.restype = ctypes.c_long
    return _ref


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


GENE_STATE = np.dtype([("delta", "f8"), ("acat_p", "f8"), ("best_p", "f8"), ("best_beta", "f8"),
                       ("best_se", "f8"), ("best_stat", "f8"), ("best_r2", "f8"), ("best_af", "f8"),
                       ("best_snp", "i8"), ("n_tested", "i4"), ("n_sig", "i4")])


def p_from_t(t, df):
    t = np.atleast_1d(_f64(t))
    out = np.empty_like(t)
    lib().orc_p_from_t_vec(_p(t), ctypes.c_int64(t.size), ctypes.c_double(df), _p(out))
    return out


def betai(a, b, x):
    return lib().orc_betai(ctypes.c_double(a), ctypes.c_double(b), ctypes.c_double(x))


def lnbeta(a, b):
    return lib().orc_lnbeta(ctypes.c_double(a), ctypes.c_double(b))


def beta_approx(perm_min_p, obs_min_p):
    """Per gene (row of perm_min_p, G x B) -> (p_beta, shape1, shape2)."""
    pm = np.atleast_2d(_f64(perm_min_p))
    obs = np.atleast_1d(_f64(obs_min_p))
    G, B = pm.shape
    out = np.empty((G, 3))
    row = np.empty(3)
    for k in range(G):
        r = np.ascontiguousarray(pm[k])
        lib().orc_beta_approx(_p(r), ctypes.c_int32(B), ctypes.c_double(obs[k]), _p(row))
        out[k] = row
    return out[:, 0].copy(), out[:, 1].copy(), out[:, 2].copy()


def acat(p):
    p = _f64(p)
    return lib().orc_acat(_p(p), ctypes.c_int64(p.size))


def p_emp(Tobs, Tperm):
    Tperm = _f64(Tperm)
    return lib().orc_p_emp(ctypes.c_double(Tobs), _p(Tperm), ctypes.c_int32(Tperm.size))


def bh(acat_p):
    a = _f64(acat_p)
    q = np.empty_like(a)
    lib().orc_bh(_p(a), ctypes.c_int32(a.size), _p(q))
    return q


def subset_maf_or_nan(g):
    g = _f64(g)
    maf = ctypes.c_double(np.nan)
    r = lib().orc_subset_maf_or_nan(_p(g), ctypes.c_int32(g.size), ctypes.byref(maf))
    return r, maf.value


def decode_bed(rows, n_file, sample_col=None, impute=True, max_miss=0.8, maf_min=0.0):
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    m, nb = rows.shape
    if sample_col is None:
        sample_col = np.arange(n_file, dtype=np.int32)
    sample_col = np.ascontiguousarray(sample_col, dtype=np.int32)
    n = sample_col.size
    dos = np.empty((m, n))
    miss = np.empty((m, n), dtype=np.int8)
    keep = np.empty(m, dtype=np.uint8)
    af = np.empty(m)
    lib().orc_decode_bed(_p(rows), ctypes.c_int64(m), ctypes.c_int32(nb), ctypes.c_int32(n_file),
                         _p(sample_col), ctypes.c_int32(n), ctypes.c_int32(1 if impute else 0),
                         ctypes.c_double(max_miss), ctypes.c_double(maf_min), _p(dos), _p(miss),
                         _p(keep), _p(af))
    return dos, miss, keep.astype(bool), af


def vcf_record(line, n_file, sample_col=None, ds_float=False, impute=True, max_miss=0.8, maf_min=0.0):
    """VcfSession::parse_record + pass_maf on one VCF data line (str or bytes,
    no newline). Returns (status, dosage[n], af, (n_miss, sum, copies)) with
    status -1 skipped (n_allele != 2), -2 column count != n_file, 0 rejected,
    1 kept, and raw[n] (the dosage before imputation, NaN = missing)."""
    if isinstance(line, str):
        line = line.encode()
    sc = np.arange(n_file, dtype=np.int32) if sample_col is None else np.ascontiguousarray(sample_col, np.int32)
    n = sc.size
    d = np.empty(n, np.float64)
    af = ctypes.c_double()
    st = np.zeros(3, np.float64)
    raw = np.empty(n, np.float64)
    rc = lib().orc_vcf_record(ctypes.c_char_p(line), ctypes.c_int64(len(line)), ctypes.c_int32(int(ds_float)),
                              ctypes.c_int32(n_file), _p(sc), ctypes.c_int32(n), ctypes.c_int32(int(impute)),
                              ctypes.c_double(max_miss), ctypes.c_double(maf_min), _p(d), ctypes.byref(af), _p(st),
                              _p(raw))
    return int(rc), d, af.value, (int(st[0]), float(st[1]), float(st[2])), raw


def eigh(K):
    K = _f64(np.asfortranarray(K)).reshape(K.shape)
    n = K.shape[0]
    Kc = np.asfortranarray(K)
    Q = np.empty((n, n), order="F")
    lam = np.empty(n)
    lib().orc_eigh(_p(Kc), ctypes.c_int32(n), _p(Q), _p(lam))
    return Q, lam


def reml_negll(delta, yt, Xt, lam):
    yt = _f64(yt)
    Xt = np.asfortranarray(Xt, dtype=np.float64)
    lam = _f64(lam)
    n, p = Xt.shape
    return lib().orc_reml_negll(ctypes.c_double(delta), _p(yt), _p(Xt), _p(lam), ctypes.c_int32(n),
                                ctypes.c_int32(p))


def optimize_delta(Yt, Xt, lam):
    Yt = np.asfortranarray(Yt, dtype=np.float64)
    Xt = np.asfortranarray(Xt, dtype=np.float64)
    lam = _f64(lam)
    n, G = Yt.shape
    p = Xt.shape[1]
    d = np.empty(G)
    lib().orc_optimize_delta(_p(Yt), ctypes.c_int32(G), _p(Xt), _p(lam), ctypes.c_int32(n),
                             ctypes.c_int32(p), _p(d))
    return d


def rotate(Q, A):
    Q = np.asfortranarray(Q, dtype=np.float64)
    A = np.asfortranarray(A, dtype=np.float64)
    n = Q.shape[0]
    A2 = A.reshape(n, -1, order="F")
    out = np.empty(A2.shape, order="F")
    lib().orc_rotate(_p(Q), ctypes.c_int32(n), _p(np.asfortranarray(A2)), ctypes.c_int64(A2.shape[1]),
                     _p(out))
    return out.reshape(A.shape, order="F")


def lmm_test(yt, Xt, lam, delta, gt):
    yt, lam, gt = _f64(yt), _f64(lam), _f64(gt)
    Xt = np.asfortranarray(Xt, dtype=np.float64)
    n, p = Xt.shape
    out = np.empty(5)
    dout = ctypes.c_double(0)
    lib().orc_lmm_test(_p(yt), _p(Xt), _p(lam), ctypes.c_int32(n), ctypes.c_int32(p),
                       ctypes.c_double(delta), _p(gt), _p(out), ctypes.byref(dout))
    return out, dout.value



def lmm_perm_pheno(Q, lam, y, X, delta, idx):
    """LMM permutation base and one permutation's pseudo-phenotype
    (scan_cis.hpp:77-87,135-148). Returns (base, y_perm), or None when the
    LDLT of X̃ᵀD⁻¹X̃ reports failure."""
    Q = np.asfortranarray(Q, dtype=np.float64)
    X = np.asfortranarray(X, dtype=np.float64)
    lam, y = _f64(lam), _f64(y)
    idx = np.ascontiguousarray(idx, np.int32)
    n, p = X.shape
    base, yp = np.empty(n), np.empty(n)
    ok = lib().orc_lmm_perm_pheno(_p(Q), _p(lam), _p(y), _p(X), ctypes.c_int32(n), ctypes.c_int32(p),
                                  ctypes.c_double(delta), _p(idx), _p(base), _p(yp))
    return (base, yp) if ok else None

def _gene_windows(gene_chrom, gene_cstart, gene_cend, G):
    if gene_chrom is None:
        return (np.full(G, -1, np.int32), np.zeros(G, np.int64), np.zeros(G, np.int64))
    return (np.ascontiguousarray(gene_chrom, np.int32), np.ascontiguousarray(gene_cstart, np.int64),
            np.ascontiguousarray(gene_cend, np.int64))


def trans_scan(model, dosage, snp_keep, snp_chrom, snp_pos, Y, X, Q=None, lam=None, gene_chrom=None,
               gene_cstart=None, gene_cend=None, delta=None, pthr=1e-5, dense=False, hit_cap=1 << 20):
    """Trans LM (model='lm') or LMM (model='lmm') scan. Returns dict."""
    dosage = np.ascontiguousarray(dosage, dtype=np.float64)
    m, n = dosage.shape
    Y = np.asfortranarray(Y, dtype=np.float64)
    X = np.asfortranarray(X, dtype=np.float64)
    G, p = Y.shape[1], X.shape[1]
    keep = np.ascontiguousarray(snp_keep, dtype=np.uint8)
    sc = np.ascontiguousarray(snp_chrom, dtype=np.int32)
    sp = np.ascontiguousarray(snp_pos, dtype=np.int64)
    gc, gs, ge = _gene_windows(gene_chrom, gene_cstart, gene_cend, G)
    state = np.zeros(G, dtype=GENE_STATE)
    dense_arrs = [np.empty((m, G)) for _ in range(4)] if dense else [None] * 4
    dp = [_p(a) if a is not None else None for a in dense_arrs]
    loop_s = ctypes.c_double(np.nan)
    hs = np.empty(hit_cap, np.int64)
    hg = np.empty(hit_cap, np.int32)
    hv = np.empty((hit_cap, 6))
    if model == "lmm":
        Qf = np.asfortranarray(Q, dtype=np.float64)
        lam = _f64(lam)
        din = None
        if delta is not None:
            dkeep = _f64(delta)
            din = _p(dkeep)
        nh = lib().orc_trans_lmm_scan(_p(dosage), _p(keep), _p(sc), _p(sp), ctypes.c_int64(m),
                                      ctypes.c_int32(n), _p(Qf), _p(lam), _p(Y), ctypes.c_int32(G), _p(X),
                                      ctypes.c_int32(p), _p(gc), _p(gs), _p(ge), din,
                                      ctypes.c_double(pthr), _p(state), *dp, _p(hs), _p(hg), _p(hv),
                                      ctypes.c_int64(hit_cap), ctypes.byref(loop_s))
    else:
        nh = lib().orc_trans_lm_scan(_p(dosage), _p(keep), _p(sc), _p(sp), ctypes.c_int64(m),
                                     ctypes.c_int32(n), _p(Y), ctypes.c_int32(G), _p(X), ctypes.c_int32(p),
                                     _p(gc), _p(gs), _p(ge), ctypes.c_double(pthr), _p(state), *dp,
                                     _p(hs), _p(hg), _p(hv), ctypes.c_int64(hit_cap))
    k = min(nh, hit_cap)
    out = {"state": state, "n_hits": int(nh), "loop_seconds": loop_s.value, "hit_snp": hs[:k].copy(), "hit_gene": hg[:k].copy(),
           "hit_vals": hv[:k].copy()}
    if dense:
        out.update(beta=dense_arrs[0], se=dense_arrs[1], stat=dense_arrs[2], p=dense_arrs[3])
    return out


def glm_prep(y, X):
    """prep_glm_nb (NB2 IRLS with MoM phi) of one gene. Returns (converged, phi,
    mu[n], beta[p], w[n])."""
    y = _f64(y)
    X = np.asfortranarray(X, dtype=np.float64)
    n, p = X.shape
    phi = ctypes.c_double()
    mu, beta, w = np.empty(n), np.empty(p), np.empty(n)
    conv = lib().orc_glm_prep(_p(y), _p(X), ctypes.c_int32(n), ctypes.c_int32(p), ctypes.byref(phi), _p(mu),
                              _p(beta), _p(w))
    return bool(conv), phi.value, mu, beta, w


def glm_test(y, X, g):
    """test_glm_nb of one SNP (imputed dosage g) for one gene: (beta, se, stat, p)."""
    y, g = _f64(y), _f64(g)
    X = np.asfortranarray(X, dtype=np.float64)
    n, p = X.shape
    out = np.empty(4)
    lib().orc_glm_test(_p(y), _p(X), ctypes.c_int32(n), ctypes.c_int32(p), _p(g), _p(out))
    return tuple(out)


def glm_scan(dosage, snp_keep, snp_chrom, snp_pos, Y, X, gene_chrom=None, gene_cstart=None, gene_cend=None,
             cis=False, pthr=1e-5, dense=False, hit_cap=1 << 20):
    """NB GLM score-test scan (trans: windows excluded; cis: only windows). Returns dict."""
    dosage = np.ascontiguousarray(dosage, dtype=np.float64)
    m, n = dosage.shape
    Y = np.asfortranarray(Y, dtype=np.float64)
    X = np.asfortranarray(X, dtype=np.float64)
    G, p = Y.shape[1], X.shape[1]
    keep = np.ascontiguousarray(snp_keep, dtype=np.uint8)
    sc = np.ascontiguousarray(snp_chrom, dtype=np.int32)
    sp = np.ascontiguousarray(snp_pos, dtype=np.int64)
    gc, gs, ge = _gene_windows(gene_chrom, gene_cstart, gene_cend, G)
    state = np.zeros(G, dtype=GENE_STATE)
    phi = np.empty(G)
    conv = np.empty(G, np.int32)
    dense_arrs = [np.empty((m, G)) for _ in range(4)] if dense else [None] * 4
    dp = [_p(a) if a is not None else None for a in dense_arrs]
    hs = np.empty(hit_cap, np.int64)
    hg = np.empty(hit_cap, np.int32)
    hv = np.empty((hit_cap, 6))
    nh = lib().orc_glm_scan(_p(dosage), _p(keep), _p(sc), _p(sp), ctypes.c_int64(m), ctypes.c_int32(n), _p(Y),
                            ctypes.c_int32(G), _p(X), ctypes.c_int32(p), _p(gc), _p(gs), _p(ge),
                            ctypes.c_int32(int(cis)), ctypes.c_double(pthr), _p(state), _p(phi), _p(conv), *dp,
                            _p(hs), _p(hg), _p(hv), ctypes.c_int64(hit_cap))
    k = min(nh, hit_cap)
    out = {"state": state, "phi": phi, "converged": conv.astype(bool), "n_hits": int(nh),
           "hit_snp": hs[:k].copy(), "hit_gene": hg[:k].copy(), "hit_vals": hv[:k].copy()}
    if dense:
        out.update(beta=dense_arrs[0], se=dense_arrs[1], stat=dense_arrs[2], p=dense_arrs[3])
    return out


def grm(dosage, miss, snp_keep):
    dosage = np.ascontiguousarray(dosage, dtype=np.float64)
    m, n = dosage.shape
    miss = np.ascontiguousarray(miss, dtype=np.int8)
    keep = np.ascontiguousarray(snp_keep, dtype=np.uint8)
    K = np.empty((n, n), order="F")
    N = np.empty(n * (n + 1) // 2, dtype=np.int32)
    used = lib().orc_grm(_p(dosage), _p(miss), _p(keep), ctypes.c_int64(m), ctypes.c_int32(n), _p(K), _p(N))
    return K, N, int(used)


def fnv1a(s: str) -> int:
    return int(lib().orc_fnv1a(s.encode()))


def perm_index(seed: int, b: int, gene: str, n: int) -> np.ndarray:
    idx = np.empty(n, dtype=np.int32)
    lib().orc_perm_index(ctypes.c_int32(seed), ctypes.c_int32(b), gene.encode(), ctypes.c_int32(n), _p(idx))
    return idx


def cis_gene(model, y, X, W, B=0, seed=-1, gene="g0", Q=None, lam=None):
    """Cis scan of one gene window with B permutations. model 'lm' | 'lmm'."""
    y = _f64(y)
    X = np.asfortranarray(X, dtype=np.float64)
    W = np.ascontiguousarray(W, dtype=np.float64)
    n, p = X.shape
    c = W.shape[0]
    Qp = _p(np.asfortranarray(Q, dtype=np.float64)) if Q is not None else None
    Qkeep = np.asfortranarray(Q, dtype=np.float64) if Q is not None else None
    lamk = _f64(lam) if lam is not None else None
    wb, ws, wt, wp = (np.empty(c) for _ in range(4))
    obs = np.empty(2)
    pm = np.empty(max(B, 1))
    tp = np.empty(max(B, 1))
    dl = ctypes.c_double(np.nan)
    pe = lib().orc_cis_gene(ctypes.c_int32(0 if model == "lm" else 1), _p(y), _p(X), ctypes.c_int32(n),
                            ctypes.c_int32(p), _p(W), ctypes.c_int32(c),
                            _p(Qkeep) if Qkeep is not None else None,
                            _p(lamk) if lamk is not None else None, ctypes.c_int32(B), ctypes.c_int32(seed),
                            gene.encode(), _p(wb), _p(ws), _p(wt), _p(wp), _p(obs), _p(pm), _p(tp),
                            ctypes.byref(dl))
    del Qp
    return {"beta": wb, "se": ws, "stat": wt, "p": wp, "n_tested": int(obs[0]), "best_p": obs[1],
            "perm_min_p": pm[:B], "T_perm": tp[:B], "p_emp": pe, "delta": dl.value}


def num_threads() -> int:
    return int(lib().orc_num_threads())
