"""--pheno-norm int (inverse_normal_transform, util.cpp:132-205): rank-based
inverse normal transform with average ranks for ties, applied per gene on the
finite values. CPU: the host's transform against scipy's rankdata + ndtri;
the option rules. GPU: a trans LM run on transformed phenotypes equals the
oracle's scan of the scipy-transformed matrix."""
import os
import subprocess

import numpy as np
import pytest
from scipy.special import ndtri
from scipy.stats import rankdata

import oracle as O
from eqtl_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "eqtl_amd", "eqtl")


def scipy_int(y):
    out = y.copy()
    f = np.isfinite(y)
    if f.sum() > 1:
        out[f] = ndtri((rankdata(y[f]) - 0.5) / f.sum())
    return out


def test_host_transform_matches_scipy(tmp_path):
    exe = tmp_path / "pnc"
    subprocess.check_call(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "pheno_norm_check.cpp"), "-o",
                           str(exe)])
    rng = np.random.default_rng(0)
    ps = np.concatenate([rng.random(3000), np.logspace(-15, -1, 300), 1 - np.logspace(-15, -1, 300), [0.5, 0.0, 1.0]])
    y = np.round(rng.normal(size=400), 1)  # many ties
    y[[3, 50, 77]] = np.nan
    inp = " ".join(repr(float(p)) for p in ps) + "\n" + " ".join("nan" if np.isnan(v) else repr(float(v)) for v in y)
    out = subprocess.run([str(exe)], input=inp + "\n", capture_output=True, text=True).stdout.split("--\n")
    q = np.array([float(v) for v in out[0].split()])
    z = np.array([float(v) for v in out[1].split()])
    pc = np.clip(ps, 1e-15, 1 - 1e-15)
    ref = ndtri(pc)
    assert np.max(np.abs(q - ref) / np.maximum(np.abs(ref), 1e-3)) < 1e-13
    f = np.isfinite(y)
    assert np.max(np.abs(z[f] - scipy_int(y)[f])) < 1e-14
    assert np.isnan(z[~f]).all()


def test_pheno_norm_option_rules(tmp_path):
    r = subprocess.run([CLI, "-b", "x", "-e", "y", "--model", "glm", "--pheno-norm", "int", "-o", str(tmp_path / "o")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--pheno-norm int is incompatible with count model glm" in r.stderr
    r = subprocess.run([CLI, "-b", "x", "-e", "y", "--pheno-norm", "rank", "-o", str(tmp_path / "o")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "invalid --pheno-norm" in r.stderr


@pytest.mark.gpu
def test_cli_lm_with_int(tmp_path):
    ds = synth.make_trans(51, n=130, m=1200, G=10, n_chrom=2, ncov=2, miss=0.01, spacing=1000, n_planted=4,
                          beta_planted=(0.4, 1.0))
    ds.Y = np.exp(ds.Y)  # skewed phenotypes, so the transform matters
    prefix = str(tmp_path / "d")
    ds.write(prefix)
    w = 200000
    out = str(tmp_path / "o")
    r = subprocess.run([CLI, "-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "-g",
                        prefix + ".gff", "--model", "lm", "--mode", "trans", "-w", str(w), "--pval-trans", "1e-2",
                        "--precision", "fp32", "--pheno-norm", "int", "-o", out], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    Yn = np.column_stack([scipy_int(ds.Y[:, k]) for k in range(ds.Y.shape[1])])
    dos, missm, keep, af = O.decode_bed(ds.bed_rows(), ds.n)
    cs, ce = np.maximum(1, ds.gene_start - w), ds.gene_end + w
    ref = O.trans_scan("lm", dos, keep, ds.chrom, ds.pos, Yn, ds.X(), gene_chrom=ds.gene_chrom.astype(np.int32),
                       gene_cstart=cs, gene_cend=ce, pthr=1e-2, dense=True, hit_cap=ds.m * ds.Y.shape[1] + 1)
    with open(out + ".lm.trans.pairs.tsv") as f:
        f.readline()
        pairs = [ln.rstrip("\n").split("\t") for ln in f if ln.strip()]
    assert len(pairs) > 0
    for p in pairs:
        k, j = int(p[0][1:]), int(p[1][2:]) - 1
        assert float(p[10]) == pytest.approx(ref["p"][j, k], rel=1e-3)
    n_ref = int(np.sum(np.isfinite(ref["p"]) & (ref["p"] <= 1e-2)))
    assert abs(len(pairs) - n_ref) <= 1
