"""NB GLM score test (prep_glm_nb / test_glm_nb, model_glm.cpp:1-132).

CPU: the oracle's restatement pinned by what the mathematics fixes — the score
is the derivative of the NB2 log-likelihood in the SNP coefficient at the null
fit (finite differences of scipy's gammaln form), the information is the
Schur complement of the full (p+1)-parameter Fisher information, the null fit
solves the NB2 score equations with phi at its MoM fixed point, and the
phi -> 0 limit is a Poisson GLM (scipy.optimize).
GPU: eqtl_gpu_glm_prep + the fused scan against the oracle (trans and cis),
and the CLI on the reference's smoke counts.
"""
import math
import os
import subprocess

import numpy as np
import pytest
from scipy.special import gammaln
from scipy.optimize import minimize

import oracle as O
import eqtl_amd
from eqtl_amd import _lib


def nb_counts(rng, n, X, beta, phi):
    mu = np.exp(X @ beta)
    if phi <= 0:
        return rng.poisson(mu).astype(float)
    lam = rng.gamma(1.0 / phi, phi * mu)
    return rng.poisson(lam).astype(float)


def design(rng, n, p):
    return np.column_stack([np.ones(n)] + [rng.standard_normal(n) for _ in range(p - 1)])


def nb2_loglik(y, mu, phi):
    r = 1.0 / phi
    return np.sum(gammaln(y + r) - gammaln(r) - gammaln(y + 1) + y * np.log(phi * mu / (1 + phi * mu))
                  - r * np.log1p(phi * mu))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_score_and_information(seed):
    rng = np.random.default_rng(seed)
    n, p = 300, 4
    X = design(rng, n, p)
    y = nb_counts(rng, n, X, np.array([1.5, 0.3, -0.2, 0.1]), 0.4)
    conv, phi, mu, beta, w = O.glm_prep(y, X)
    assert conv and phi > 0.05
    g = rng.binomial(2, 0.3, n).astype(float)
    b, se, stat, pv = O.glm_test(y, X, g)
    info, U = 1.0 / se ** 2, b / se ** 2
    # score: d loglik / d gamma at gamma = 0, mu = exp(X beta + gamma g), phi fixed
    eta = np.log(mu)
    h = 1e-6
    fd = (nb2_loglik(y, np.exp(eta + h * g), phi) - nb2_loglik(y, np.exp(eta - h * g), phi)) / (2 * h)
    assert U == pytest.approx(fd, rel=1e-6, abs=1e-6 * np.sqrt(info))
    # information: Schur complement of J = [X g]^T W [X g], W = mu / (1 + phi mu)
    Z = np.column_stack([X, g])
    J = Z.T @ (Z * (mu / (1 + phi * mu))[:, None])
    assert info == pytest.approx(1.0 / np.linalg.inv(J)[-1, -1], rel=1e-10)
    assert stat == pytest.approx(U / math.sqrt(info), rel=1e-12)
    assert pv == pytest.approx(math.erfc(abs(stat) / math.sqrt(2)), rel=1e-14)


@pytest.mark.parametrize("seed", [3, 4])
def test_oracle_null_fit_is_stationary(seed):
    rng = np.random.default_rng(seed)
    n, p = 400, 5
    X = design(rng, n, p)
    y = nb_counts(rng, n, X, np.array([2.0, 0.4, 0.0, -0.3, 0.2]), 0.25)
    conv, phi, mu, beta, w = O.glm_prep(y, X)
    assert conv
    np.testing.assert_allclose(mu, np.exp(X @ beta), rtol=1e-12)
    # NB2 score equations in beta at fixed phi
    r = (y - mu) / (1 + phi * mu)
    assert np.max(np.abs(X.T @ r)) < 1e-5 * np.sqrt(np.sum(w)) * np.sqrt(np.sum(r * r) + 1)
    # phi is the MoM fixed point at the fitted mean
    m = np.maximum(mu, 1e-8)
    mom = max(1e-8, np.sum(((y - m) ** 2 - m) / m ** 2) / max(1, n - p))
    assert mom == pytest.approx(phi, rel=1e-5)
    np.testing.assert_allclose(w, m / (1 + phi * m), rtol=1e-14)


def test_oracle_poisson_limit():
    """Poisson data: MoM phi collapses to its floor and the fit is the Poisson MLE."""
    rng = np.random.default_rng(9)
    n, p = 500, 3
    X = design(rng, n, p)
    y = nb_counts(rng, n, X, np.array([0.8, 0.25, -0.15]), 0.0)
    conv, phi, mu, beta, w = O.glm_prep(y, X)
    assert conv and phi < 0.05

    def nll(b):
        eta = X @ b
        return np.sum(np.exp(eta) - y * eta)

    def grad(b):
        return X.T @ (np.exp(X @ b) - y)

    ref = minimize(nll, np.zeros(p), jac=grad, method="BFGS", options={"gtol": 1e-10}).x
    # the NB2 fit with a small phi stays within phi of the Poisson MLE
    np.testing.assert_allclose(beta, ref, atol=5 * phi + 1e-6)


def test_oracle_nonconverged_gene_is_not_tested():
    """A covariate matrix whose X^T W X is singular: no test (p NaN), as the
    reference's failed LDLT leaves converged false."""
    rng = np.random.default_rng(5)
    n = 50
    x1 = rng.standard_normal(n)
    X = np.column_stack([np.ones(n), x1, 2 * x1])
    y = nb_counts(rng, n, np.column_stack([np.ones(n)]), np.array([1.0]), 0.3)
    conv, phi, mu, beta, w = O.glm_prep(y, X)
    assert not conv
    assert math.isnan(O.glm_test(y, X, rng.binomial(2, 0.4, n).astype(float))[3])


# ----------------------------------------------------------------------------- GPU

import functools

from _cases import dense_from_hits, rel_err, windows
from eqtl_amd import synth


@functools.lru_cache(maxsize=None)
def glm_case(seed=31, n=240, m=2400, G=40, ncov=3):
    """Counts with NB2 noise: per-gene base mean 2..60, dispersion 0 (Poisson)
    to 1.5, covariate effects, and eight genes with one planted SNP effect."""
    ds = synth.make_trans(seed, n=n, m=m, G=G, n_chrom=2, ncov=ncov, miss=0.01, spacing=1000, n_planted=0)
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    rng = np.random.default_rng(seed + 1)
    X = ds.X()
    Y = np.empty((n, G))
    ok = np.nonzero(keep)[0]
    for k in range(G):
        eta = np.log(rng.uniform(2, 60)) + X[:, 1:] @ rng.normal(0, 0.15, X.shape[1] - 1)
        if k % 5 == 0:
            j = rng.choice(ok)
            g = dos[j] - dos[j].mean()
            eta = eta + rng.choice([-1, 1]) * rng.uniform(0.2, 0.6) * g
        phi = 0.0 if k == 0 else rng.uniform(0.02, 1.5)
        Y[:, k] = nb_counts(rng, n, np.eye(n), eta, phi) if phi > 0 else rng.poisson(np.exp(eta))
    gc, gs, ge = windows(ds, 200_000)
    return ds, rows, dos, keep, Y, (gc, gs, ge)


def _glm_ref(cis):
    ds, rows, dos, keep, Y, (gc, gs, ge) = glm_case()
    return O.glm_scan(dos, keep, ds.chrom, ds.pos, Y, ds.X(), gc, gs, ge, cis=cis, pthr=1.0, dense=True,
                      hit_cap=ds.m * Y.shape[1] + 1)


_glm_ref = functools.lru_cache(maxsize=None)(_glm_ref)


@pytest.mark.gpu
def test_glm_prep_matches_oracle(gpu_ctx):
    ds, rows, dos, keep, Y, win = glm_case()
    ref = _glm_ref(False)
    gpu_ctx.set_samples(ds.n)
    phi, conv = gpu_ctx.glm_prep(Y, ds.X())
    assert np.array_equal(conv, ref["converged"]) and conv.all()
    print("GLM phi max rel err", rel_err(phi, ref["phi"]).max())
    np.testing.assert_allclose(phi, ref["phi"], rtol=1e-8, atol=1e-12)
    with pytest.raises(eqtl_amd.EqtlGpuError):
        gpu_ctx.glm_prep(Y + 0.5, ds.X())  # not counts


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [eqtl_amd.PREC_FP32, eqtl_amd.PREC_BF16X3])
@pytest.mark.parametrize("cis", [False, True])
def test_glm_scan_parity(gpu_ctx, prec, cis):
    """The fused NB score-test scan against the oracle: tested sets, counts and
    best SNP exact; z within 5e-6 (FP32) / 5e-5 (split BF16) of (1 + |z|), p
    within 1e-3 relative."""
    ds, rows, dos, keep, Y, win = glm_case()
    ref = _glm_ref(cis)
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_policy(True, 0.8, 0.0)
    gpu_ctx.set_snp_annot(ds.chrom, ds.pos)
    gpu_ctx.glm_prep(Y, ds.X())
    gpu_ctx.set_gene_windows(*win)
    scope = eqtl_amd.SCOPE_CIS if cis else eqtl_amd.SCOPE_TRANS
    hits, st = gpu_ctx.scan(eqtl_amd.MODEL_GLM, prec, rows, pthr=1.0, block=1024, hit_cap=1 << 20, scope=scope)
    m, G = ds.m, Y.shape[1]
    gd = dense_from_hits(hits, m, G)
    tr = np.isfinite(ref["p"])
    assert np.array_equal(tr, np.isfinite(gd["p"]))
    assert tr.sum() > (100 if cis else 10000)
    z_r, z_g = ref["stat"][tr], gd["stat"][tr]
    dz = np.abs(z_g - z_r) / (1 + np.abs(z_r))
    pr = rel_err(gd["p"][tr], ref["p"][tr])
    print(f"GLM prec={prec} cis={cis}: max |dz|/(1+|z|) = {dz.max():.3e}, max p rel = {pr.max():.3e}, "
          f"max |z| = {np.abs(z_r).max():.1f}")
    assert dz.max() < (5e-6 if prec == eqtl_amd.PREC_FP32 else 5e-5)
    assert pr.max() < 1e-3
    assert rel_err(gd["se"][tr], ref["se"][tr]).max() < 1e-4
    rs = ref["state"]
    assert np.array_equal(st["n_tested"], rs["n_tested"])
    assert np.array_equal(st["n_sig"], rs["n_sig"])
    assert np.array_equal(st["best_snp"], rs["best_snp"])
    print(f"ACAT vs oracle: max rel {rel_err(st['acat_p'], rs['acat_p']).max():.3e}")
    assert rel_err(st["acat_p"], rs["acat_p"]).max() < 2e-3
    assert np.all(hits["r2"] == 0)


def _read_tsv(path):
    with open(path) as f:
        head = f.readline().rstrip("\n").split("\t")
        return head, [ln.rstrip("\n").split("\t") for ln in f if ln.strip()]


def _table(path):
    with open(path) as f:
        cols = f.readline().rstrip("\n").split("\t")[1:]
        ids, vals = [], []
        for ln in f:
            t = ln.rstrip("\n").split("\t")
            ids.append(t[0])
            vals.append([float(x) for x in t[1:]])
    return ids, cols, np.array(vals)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["cis", "trans"])
def test_cli_glm_on_smoke_counts(tmp_path, mode):
    """--model glm on the reference's smoke data (VCF genotypes, count
    phenotypes, one covariate): every pair line equals the oracle's NB score
    test of the oracle-decoded dosage (se 1e-5 relative, beta and stat to 1e-5
    of se and of 1, p 1e-4),
    phi equals the oracle's null fit, pairs are gene-major in phenotype order,
    and each gene's n_tested equals the oracle's count of its SNPs."""
    from _vcf import SMOKE, read_vcf
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "eqtl_amd", "eqtl")
    vcf = os.path.join(SMOKE, "smoke.vcf.gz")
    counts, covar, gff = (os.path.join(SMOKE, f) for f in ("smoke.counts.tsv", "smoke.covar.tsv", "smoke.gff"))
    out = str(tmp_path / "g")
    w = 50000
    r = subprocess.run([cli, "-v", vcf, "-e", counts, "-c", covar, "-g", gff, "--model", "glm", "--mode", mode,
                        "-w", str(w), "--pval-cis", "1", "--pval-trans", "1", "--precision", "fp32", "-o", out],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    head, pairs = _read_tsv(f"{out}.glm.{mode}.pairs.tsv")
    assert head[-2:] == ["phi", "glm_converged"] and len(pairs) > 50
    samples, ds_float, lines = read_vcf(vcf)
    cid, genes, Y = _table(counts)
    vid, vcols, V = _table(covar)
    X = np.column_stack([np.ones(len(cid)), V[[vid.index(s) for s in cid]]])
    sc = np.array([samples.index(s) for s in cid], np.int32)
    row = {ln.split(b"\t", 3)[2].decode(): j for j, ln in enumerate(lines)}
    dos = {}
    phis = {}
    order = []
    for pr in pairs:
        gene, snp = pr[0], pr[1]
        if not order or order[-1][0] != gene:
            assert all(gene != o[0] for o in order), "pairs not gene-major"
            order.append((gene, []))
        order[-1][1].append(row[snp])
        if snp not in dos:
            st, d, *_ = O.vcf_record(lines[row[snp]], len(samples), sample_col=sc)
            assert st == 1
            dos[snp] = d
        y = Y[:, genes.index(gene)]
        if gene not in phis:
            phis[gene] = O.glm_prep(y, X)[1]
        b, se, stat, pv = O.glm_test(y, X, dos[snp])
        got = [float(v) for v in (pr[7], pr[8], pr[9], pr[10])]
        # beta near 0 is set by the score's absolute error: compare it in units of se
        assert got[0] == pytest.approx(b, rel=1e-5, abs=1e-5 * se) and got[1] == pytest.approx(se, rel=1e-5)
        assert got[2] == pytest.approx(stat, rel=1e-5, abs=1e-5) and got[3] == pytest.approx(pv, rel=1e-4)
        assert float(pr[15]) == pytest.approx(phis[gene], rel=1e-8) and pr[16] == "1"
        assert pr[11] == "0"  # r2
    gidx = [genes.index(g) for g, _ in order]
    assert gidx == sorted(gidx)  # phenotype order
    assert all(js == sorted(js) for _, js in order)
    head, region = _read_tsv(f"{out}.glm.{mode}.region.tsv")
    n_by_gene = {g: len(js) for g, js in order}
    for rg in region:
        if rg[0] in n_by_gene:
            assert int(rg[3]) == n_by_gene[rg[0]]  # --pval 1: every tested SNP is a pair
    r = subprocess.run([cli, "-v", vcf, "-e", counts, "-g", gff, "--model", "glm", "--mode", "cis", "--perm", "10",
                        "-o", out + "p"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "permutations are provided for lm and lmm" in r.stderr
