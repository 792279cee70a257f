"""The eqtl CLI (C++ host over the C ABI): failure paths on the CPU, and an
end-to-end trans LM/LMM run on the GPU whose output files are compared with
the FP64 oracle (pairs at --pval-trans 1e-3, top, region)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from eqtl_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "eqtl_amd", "eqtl")


def run(args, **kw):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300, **kw)


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    ds = synth.make_trans(31, n=120, m=1500, G=16, n_chrom=2, ncov=3, miss=0.01, spacing=1000,
                          h2=("uniform", 0.0, 0.6), n_causal=200, n_planted=4, beta_planted=(0.4, 1.0))
    prefix = str(d / "c")
    ds.write(prefix)
    return ds, prefix, d


def test_cli_bad_bed_magic(small, tmp_path):
    ds, prefix, d = small
    bad = str(tmp_path / "bad")
    for ext in (".bim", ".fam", ".pheno.tsv"):
        with open(prefix + ext) as f, open(bad + ext, "w") as g:
            g.write(f.read())
    with open(bad + ".bed", "wb") as f:
        f.write(bytes([0x6C, 0x1C, 0x01]) + bytes(10))
    r = run(["-b", bad, "-e", bad + ".pheno.tsv", "--model", "lm", "-o", str(tmp_path / "o")])
    assert r.returncode == 1 and "[E] not a PLINK1 .bed (bad magic)" in r.stderr


def test_cli_truncated_bed(small, tmp_path):
    ds, prefix, d = small
    bad = str(tmp_path / "trunc")
    for ext in (".bim", ".fam", ".pheno.tsv"):
        with open(prefix + ext) as f, open(bad + ext, "w") as g:
            g.write(f.read())
    with open(prefix + ".bed", "rb") as f, open(bad + ".bed", "wb") as g:
        g.write(f.read()[:1000])
    r = run(["-b", bad, "-e", bad + ".pheno.tsv", "--model", "lm", "-o", str(tmp_path / "o")])
    assert r.returncode == 1 and "bed size too small" in r.stderr


def test_cli_missing_pheno_and_unsorted_bim(small, tmp_path):
    ds, prefix, d = small
    r = run(["-b", prefix, "-o", str(tmp_path / "o")])
    assert r.returncode == 1 and "missing -e/--pheno" in r.stderr
    bad = str(tmp_path / "uns")
    for ext in (".bed", ".fam", ".pheno.tsv"):
        with open(prefix + ext, "rb") as f, open(bad + ext, "wb") as g:
            g.write(f.read())
    lines = open(prefix + ".bim").read().splitlines()
    lines[3], lines[4] = lines[4], lines[3]
    open(bad + ".bim", "w").write("\n".join(lines) + "\n")
    r = run(["-b", bad, "-e", bad + ".pheno.tsv", "--model", "lm", "-o", str(tmp_path / "o")])
    assert r.returncode == 1 and "positions not sorted" in r.stderr


def _read_tsv(path):
    lines = open(path).read().splitlines()
    head = lines[0].split("\t")
    return head, [l.split("\t") for l in lines[1:]]


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["lm", "lmm"])
def test_cli_trans_end_to_end(small, tmp_path, model):
    ds, prefix, d = small
    out = str(tmp_path / "run")
    w = 200000
    r = run(["-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "-g", prefix + ".gff",
             "--model", model, "--mode", "trans", "-w", str(w), "--pval-trans", "1e-3", "--precision", "fp32",
             "-o", out])
    assert r.returncode == 0, r.stderr
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    gc = ds.gene_chrom.astype(np.int32)
    cs, ce = np.maximum(1, ds.gene_start - w), ds.gene_end + w
    kw = {}
    if model == "lmm":
        K, N, used = O.grm(dos, missm, keep)
        Q, lam = O.eigh(K)
        kw = {"Q": Q, "lam": lam}
    ref = O.trans_scan(model, dos, keep, ds.chrom, ds.pos, ds.Y, ds.X(), gene_chrom=gc, gene_cstart=cs,
                       gene_cend=ce, pthr=1e-3, dense=True, hit_cap=1 << 20, **kw)
    head, pairs = _read_tsv(f"{out}.{model}.trans.pairs.tsv")
    assert head == ["gene", "snp", "chrom", "pos", "ref", "alt", "af", "beta", "se", "stat", "p", "r2", "n",
                    "tss_dist", "scope"]
    got = {(p[0], p[1]): p for p in pairs}
    exp = {(f"g{k}", f"rs{j + 1}") for j, k in zip(ref["hit_snp"], ref["hit_gene"])}
    # identical pair sets away from the threshold
    near = {(f"g{k}", f"rs{j + 1}") for j in range(ds.m) for k in range(ds.Y.shape[1])
            if np.isfinite(ref["p"][j, k]) and abs(np.log(ref["p"][j, k] / 1e-3)) < 1e-3}
    assert (set(got) ^ exp) <= near
    # file order: SNP-major then gene order
    order = [(int(p[1][2:]), int(p[0][1:])) for p in pairs]
    assert order == sorted(order)
    for (g, s), p in got.items():
        j, k = int(s[2:]) - 1, int(g[1:])
        assert p[2] == str(ds.chrom[j]) and int(p[3]) == ds.pos[j] and p[4] == "G" and p[5] == "A"
        assert abs(float(p[9]) - ref["stat"][j, k]) <= 5e-6 * (1 + abs(ref["stat"][j, k]))
        assert abs(float(p[10]) - ref["p"][j, k]) <= 1e-3 * ref["p"][j, k]
        assert int(p[12]) == ds.n
        assert float(p[13]) == ds.pos[j] - ds.gene_start[k]
        assert p[14] == "trans"
    head, region = _read_tsv(f"{out}.{model}.trans.region.tsv")
    assert head[:6] == ["gene", "chrom", "tss", "n_tested", "n_sig", "acat_p"]
    st = ref["state"]
    for row in region:
        k = int(row[0][1:])
        assert int(row[3]) == st["n_tested"][k]
        assert abs(float(row[5]) - st["acat_p"][k]) <= 1e-2 * st["acat_p"][k] + 1e-15
        assert row[7] == "NA" and row[8] == "NA"
    head, top = _read_tsv(f"{out}.{model}.trans.top.tsv")
    for row in top:
        k = int(row[0][1:])
        assert int(row[1][2:]) - 1 == st["best_snp"][k]


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["lm", "lmm"])
def test_cli_cis_with_permutations(small, tmp_path, model):
    """--mode cis --perm B: pairs (gene-major), top, region with p_emp. LM p_emp is
    compared with the oracle exactly; LMM permutations depend on the eigenvector
    signs of the engine's own GRM basis, so for LMM only the observed statistics
    and the p_emp range are checked."""
    ds, prefix, d = small
    out = str(tmp_path / "cis")
    w, B, seed = 200000, 25, 42
    r = run(["-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "-g", prefix + ".gff",
             "--model", model, "--mode", "cis", "-w", str(w), "--pval-cis", "1e-3", "--precision", "fp32",
             "--perm", str(B), "--seed", str(seed), "-o", out])
    assert r.returncode == 0, r.stderr
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    cs, ce = np.maximum(1, ds.gene_start - w), ds.gene_end + w
    kw = {}
    if model == "lmm":
        K, N, used = O.grm(dos, missm, keep)
        Q, lam = O.eigh(K)
        kw = {"Q": Q, "lam": lam}
    G = ds.Y.shape[1]
    refs = {}
    for k in range(G):
        idx = np.nonzero((ds.chrom == ds.gene_chrom[k]) & (ds.pos >= cs[k]) & (ds.pos <= ce[k]) & keep)[0]
        refs[f"g{k}"] = (idx, O.cis_gene(model, ds.Y[:, k], ds.X(), dos[idx], B=B if model == "lm" else 0,
                                         seed=seed, gene=f"g{k}", **kw))
    head, pairs = _read_tsv(f"{out}.{model}.cis.pairs.tsv")
    exp = {(g, f"rs{idx[s] + 1}") for g, (idx, ref) in refs.items() for s in range(len(idx))
           if np.isfinite(ref["p"][s]) and ref["p"][s] <= 1e-3}
    near = {(g, f"rs{idx[s] + 1}") for g, (idx, ref) in refs.items() for s in range(len(idx))
            if np.isfinite(ref["p"][s]) and abs(np.log(ref["p"][s] / 1e-3)) < 1e-3}
    got = {(p[0], p[1]) for p in pairs}
    assert (got ^ exp) <= near
    assert all(p[14] == "cis" for p in pairs)
    # gene-major, SNP order inside a gene
    genes_in_order = []
    for p in pairs:
        if not genes_in_order or genes_in_order[-1] != p[0]:
            assert p[0] not in genes_in_order
            genes_in_order.append(p[0])
    head, region = _read_tsv(f"{out}.{model}.cis.region.tsv")
    for row in region:
        idx, ref = refs[row[0]]
        assert int(row[3]) == ref["n_tested"]
        if model == "lm" and ref["n_tested"] > 0:
            # %.10g of (1 + count) / (B + 1): the count is exact
            assert round(float(row[7]) * (B + 1)) == round(ref["p_emp"] * (B + 1)), (row[0], row[7], ref["p_emp"])
        elif ref["n_tested"] > 0:
            pe = float(row[7])
            assert 1.0 / (B + 1) <= pe <= 1.0
        if ref["n_tested"] == 0:
            assert row[8:11] == ["NA", "0", "0"]
        elif model == "lm":
            _check_beta_columns(row, ref["perm_min_p"], ref["best_p"])
        else:
            assert 0.0 <= float(row[8]) <= 1.0 and float(row[9]) > 0 and float(row[10]) > 0


def _check_beta_columns(row, perm_min_p, best_p):
    """p_beta, beta_shape1, beta_shape2 of a region line against the oracle's
    beta approximation of the oracle's permutation minima (which match the
    engine's to ~1e-6 relative, so the fit agrees to a loose 1e-3)."""
    pb, a, b = O.beta_approx(np.asarray(perm_min_p)[None, :], [best_p])
    if np.isnan(pb[0]):
        assert row[8] == "NA"
        return
    assert float(row[8]) == pytest.approx(pb[0], rel=1e-3, abs=1e-12), (row, pb, a, b)
    assert float(row[9]) == pytest.approx(a[0], rel=1e-3) and float(row[10]) == pytest.approx(b[0], rel=1e-3)


@pytest.mark.gpu
def test_cli_trans_stage2_permutations(small, tmp_path):
    """--mode trans --perm B (stage2_perm_topk): LM p_emp of every gene whose best
    p is below --perm-trans-thr equals the oracle's permutation of that gene over
    its top-K SNPs; other genes report NA."""
    ds, prefix, d = small
    out = str(tmp_path / "s2")
    w, B, seed, thr, K = 200000, 30, 11, 1e-4, 20
    r = run(["-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "-g", prefix + ".gff",
             "--model", "lm", "--mode", "trans", "-w", str(w), "--pval-trans", "1e-5", "--precision", "fp32",
             "--perm", str(B), "--seed", str(seed), "--perm-trans-thr", str(thr), "--perm-trans-top", str(K),
             "-o", out])
    assert r.returncode == 0, r.stderr
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    gc = ds.gene_chrom.astype(np.int32)
    cs, ce = np.maximum(1, ds.gene_start - w), ds.gene_end + w
    ref = O.trans_scan("lm", dos, keep, ds.chrom, ds.pos, ds.Y, ds.X(), gene_chrom=gc, gene_cstart=cs,
                       gene_cend=ce, pthr=1.0, dense=True, hit_cap=1 << 20)
    head, region = _read_tsv(f"{out}.lm.trans.region.tsv")
    checked = 0
    for row in region:
        k = int(row[0][1:])
        pk = ref["p"][:, k]
        ok = np.isfinite(pk) & (pk < thr)
        if not ok.any():
            assert row[7] == "NA"
            continue
        idx = np.nonzero(ok)[0]
        top = idx[np.lexsort((idx, pk[idx]))][:K]
        top = np.sort(top)
        o = O.cis_gene("lm", ds.Y[:, k], ds.X(), dos[top], B=B, seed=seed, gene=row[0])
        Tobs = -np.log10(max(np.nanmin(pk), 1e-300))
        if np.any(np.abs(o["T_perm"] - Tobs) <= 1e-4 * Tobs):
            continue
        assert round(float(row[7]) * (B + 1)) == round(o["p_emp"] * (B + 1)), (row[0], row[7], o["p_emp"])
        _check_beta_columns(row, o["perm_min_p"], o["best_p"])
        checked += 1
    assert checked >= 1
