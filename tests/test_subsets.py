"""Per-gene sample subsets (build_gene_ready, scan_common.cpp:11-63: samples
with a finite phenotype and finite covariates) and --fast (sparsify_grm,
model_lmm.cpp:88-104).

GPU: eqtl_gpu_set_test_subset — the genotypes are decoded, filtered and
mean-imputed over all analysis samples, then restricted to the subset — for
decode, LM and LMM scans against the oracle on the same subset; the CLI on
phenotypes with missing values (several subsets in one run) and with --fast.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import eqtl_amd
from eqtl_amd import synth
from _cases import dense_from_hits, rel_err, windows

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "eqtl_amd", "eqtl")


def _case():
    ds = synth.make_trans(41, n=150, m=1600, G=12, n_chrom=2, ncov=2, miss=0.02, spacing=1000,
                          h2=("uniform", 0.1, 0.6), n_causal=150, n_planted=4, beta_planted=(0.4, 1.0))
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)  # imputed over all analysis samples
    return ds, rows, dos, missm, keep


@pytest.mark.gpu
def test_subset_decode(gpu_ctx):
    ds, rows, dos, missm, keep = _case()
    rng = np.random.default_rng(3)
    sub = np.sort(rng.choice(ds.n, 111, replace=False)).astype(np.int32)
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_policy(True, 0.8, 0.0)
    gpu_ctx.set_test_subset(sub)
    try:
        d, k, af = gpu_ctx.decode_bed(rows)
        for j in range(ds.m):
            ref_af, _ = O.subset_maf_or_nan(dos[j, sub]) if keep[j] else (np.nan, None)
            tested = bool(keep[j]) and np.isfinite(ref_af)
            assert bool(k[j]) == tested, j
            if tested:
                np.testing.assert_array_equal(d[j], dos[j, sub])
                assert af[j] == pytest.approx(ref_af, rel=1e-14)
    finally:
        gpu_ctx.set_test_subset(None)


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["lm", "lmm"])
def test_subset_scan_parity(gpu_ctx, model):
    ds, rows, dos, missm, keep = _case()
    rng = np.random.default_rng(4)
    sub = np.sort(rng.choice(ds.n, 120, replace=False)).astype(np.int32)
    gc, gs, ge = windows(ds, 200_000)
    Y, X = ds.Y[sub], ds.X()[sub]
    kw = {}
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_policy(True, 0.8, 0.0)
    gpu_ctx.set_snp_annot(ds.chrom, ds.pos)
    gpu_ctx.set_test_subset(sub)
    try:
        if model == "lmm":
            K, N, used = O.grm(dos, missm, keep)
            Q, lam = O.eigh(K[np.ix_(sub, sub)])
            kw = {"Q": Q, "lam": lam}
            gpu_ctx.set_basis(Q, lam)
            gpu_ctx.lmm_prep(Y, X, delta_mode=eqtl_amd.DELTA_REF)
            mdl = eqtl_amd.MODEL_LMM
        else:
            gpu_ctx.lm_prep(Y, X)
            mdl = eqtl_amd.MODEL_LM
        gpu_ctx.set_gene_windows(gc, gs, ge)
        hits, st = gpu_ctx.scan(mdl, eqtl_amd.PREC_FP32, rows, pthr=1.0, block=1024, hit_cap=1 << 20)
    finally:
        gpu_ctx.set_test_subset(None)
    ref = O.trans_scan(model, dos[:, sub], keep, ds.chrom, ds.pos, Y, X, gene_chrom=gc, gene_cstart=gs,
                       gene_cend=ge, pthr=1.0, dense=True, hit_cap=ds.m * Y.shape[1] + 1, **kw)
    gd = dense_from_hits(hits, ds.m, Y.shape[1])
    tr = np.isfinite(ref["p"])
    assert np.array_equal(tr, np.isfinite(gd["p"]))
    assert rel_err(gd["p"][tr], ref["p"][tr]).max() < 1e-3
    dt = np.abs(gd["stat"][tr] - ref["stat"][tr]) / (1 + np.abs(ref["stat"][tr]))
    assert dt.max() < 5e-5
    assert np.array_equal(st["n_tested"], ref["state"]["n_tested"])
    assert np.array_equal(st["best_snp"], ref["state"]["best_snp"])


def _read_tsv(path):
    with open(path) as f:
        head = f.readline().rstrip("\n").split("\t")
        return head, [ln.rstrip("\n").split("\t") for ln in f if ln.strip()]


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["lm", "lmm"])
def test_cli_missing_phenotypes(tmp_path, model):
    """Genes with missing phenotype values (three sample subsets, two genes
    sharing one) in one trans run: each pair equals the oracle's test on
    the gene's own subset (its own GRM slice and eigenbasis for lmm), pairs
    stay SNP-major in gene order, and each gene reports its subset size."""
    ds, rows, dos, missm, keep = _case()
    prefix = str(tmp_path / "d")
    rng = np.random.default_rng(5)
    Y = ds.Y.copy()
    Y[rng.choice(ds.n, 9, replace=False), 1] = np.nan
    Y[rng.choice(ds.n, 5, replace=False), 4] = np.nan
    Y[:, 7] = Y[:, 4]  # the same subset as gene 4
    ds.Y = Y
    ds.write(prefix)
    w = 200000
    r = subprocess.run([CLI, "-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "-g",
                        prefix + ".gff", "--model", model, "--mode", "trans", "-w", str(w), "--pval-trans", "1e-2",
                        "--precision", "fp32", "-o", str(tmp_path / "o")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "distinct sample subsets" in r.stderr
    head, pairs = _read_tsv(str(tmp_path / "o") + f".{model}.trans.pairs.tsv")
    X = ds.X()
    cs, ce = np.maximum(1, ds.gene_start - w), ds.gene_end + w
    gcw = ds.gene_chrom.astype(np.int32)
    K = None
    if model == "lmm":
        K, N, used = O.grm(dos, missm, keep)
    exp = {}
    for k in range(Y.shape[1]):
        sub = np.nonzero(np.isfinite(Y[:, k]))[0]
        kw = {}
        if model == "lmm":
            Q, lam = O.eigh(K[np.ix_(sub, sub)])
            kw = {"Q": Q, "lam": lam}
        ref = O.trans_scan(model, dos[:, sub], keep, ds.chrom, ds.pos, Y[sub, k:k + 1], X[sub],
                           gene_chrom=gcw[k:k + 1], gene_cstart=cs[k:k + 1], gene_cend=ce[k:k + 1], pthr=1e-2,
                           dense=True, hit_cap=ds.m + 1, **kw)
        for j in np.nonzero(np.isfinite(ref["p"][:, 0]) & (ref["p"][:, 0] <= 1e-2))[0]:
            exp[(f"g{k}", f"rs{j + 1}")] = (ref["beta"][j, 0], ref["p"][j, 0], sub.size)
    got = {(p[0], p[1]): p for p in pairs}
    near = {key for key, v in exp.items() if abs(np.log(v[1] / 1e-2)) < 1e-3}
    assert (set(got) ^ set(exp)) <= near
    for key in set(got) & set(exp):
        b, pv, nk = exp[key]
        p = got[key]
        assert float(p[10]) == pytest.approx(pv, rel=1e-3)
        assert int(p[12]) == nk
    # SNP-major, gene order within a SNP
    keys = [(int(p[1][2:]), int(p[0][1:])) for p in pairs]
    assert keys == sorted(keys)


@pytest.mark.gpu
def test_cli_fast_sparsified_grm(tmp_path):
    """--fast: the LMM basis comes from the GRM with |K_ij| < 1e-4 zeroed."""
    ds, rows, dos, missm, keep = _case()
    prefix = str(tmp_path / "f")
    ds.write(prefix)
    w = 200000
    out = str(tmp_path / "o")
    r = subprocess.run([CLI, "-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "-g",
                        prefix + ".gff", "--model", "lmm", "--mode", "trans", "-w", str(w), "--pval-trans", "1e-2",
                        "--precision", "fp32", "--fast", "-o", out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "fast: GRM sparse approx thr=" in r.stderr
    K, N, used = O.grm(dos, missm, keep)
    Ks = K.copy()
    small = np.abs(Ks) < 1e-4
    np.fill_diagonal(small, False)
    Ks[small] = 0.0
    assert small.sum() > 0
    Q, lam = O.eigh(Ks)
    cs, ce = np.maximum(1, ds.gene_start - w), ds.gene_end + w
    ref = O.trans_scan("lmm", dos, keep, ds.chrom, ds.pos, ds.Y, ds.X(), Q=Q, lam=lam,
                       gene_chrom=ds.gene_chrom.astype(np.int32), gene_cstart=cs, gene_cend=ce, pthr=1e-2,
                       dense=True, hit_cap=ds.m * ds.Y.shape[1] + 1)
    head, pairs = _read_tsv(out + ".lmm.trans.pairs.tsv")
    assert len(pairs) > 0
    for p in pairs:
        k, j = int(p[0][1:]), int(p[1][2:]) - 1
        assert float(p[10]) == pytest.approx(ref["p"][j, k], rel=1e-3)
