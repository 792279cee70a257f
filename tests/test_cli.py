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
    assert head[:7] == ["gene", "chrom", "tss", "n_tested", "n_sig", "acat_p", "q_bh"]
    st = ref["state"]
    ks, acat_f, q_f = [], [], []
    for row in region:
        k = int(row[0][1:])
        assert int(row[3]) == st["n_tested"][k]
        assert abs(float(row[5]) - st["acat_p"][k]) <= 2e-3 * st["acat_p"][k] + 1e-15
        assert row[7] == "NA" and row[8] == "NA"
        ks.append(k)
        acat_f.append(float(row[5]))
        q_f.append(float(row[6]))
    # q_bh (bh_adjust, scan_common.cpp:201-217): the BH step-up of the file's
    # own acat_p exactly (to the %.10g printing), and of the oracle's acat_p to 2e-3
    q_f = np.array(q_f)
    np.testing.assert_allclose(q_f, O.bh(np.array(acat_f)), rtol=2e-9)
    np.testing.assert_allclose(q_f, O.bh(st["acat_p"][ks]), rtol=2e-3)
    head, top = _read_tsv(f"{out}.{model}.trans.top.tsv")
    for row in top:
        k = int(row[0][1:])
        assert int(row[1][2:]) - 1 == st["best_snp"][k]


def _write_grm_file(prefix, K, ids, rng):
    """A GCTA GRM (.grm.id, .grm.bin float32 lower triangle) holding the
    samples ids in a shuffled order plus three samples not in the data, so that
    the CLI's -k path has to slice it (slice_grm, grm.cpp:80-108). Returns K as
    the file holds it (float32-rounded), in the order of ids."""
    n = len(ids)
    extra = 3
    order = rng.permutation(n + extra)  # file row r holds sample order[r] (>= n: an extra sample)
    Kx = np.eye(n + extra)
    Kx[:n, :n] = K
    Kf = Kx[np.ix_(order, order)].astype(np.float32)
    names = [ids[i] if i < n else f"extra{i}" for i in order]
    with open(prefix + ".grm.id", "w") as f:
        for nm in names:
            f.write(f"{nm}\t{nm}\n")
    Kf[np.tril_indices(n + extra)].tofile(prefix + ".grm.bin")
    inv = np.argsort(order)
    return Kf.astype(np.float64)[np.ix_(inv[:n], inv[:n])]


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["lm", "lmm"])
def test_cli_cis_with_permutations(small, tmp_path, model):
    """--mode cis --perm B: pairs (gene-major), top, region with p_emp, equal to
    the oracle's for every gene. The LMM reads its GRM from a GCTA file (-k,
    samples shuffled and three extra ones, so slice_grm runs); the engine and
    the oracle fix eigenvector signs the same way, so the LMM permutation
    streams match too."""
    ds, prefix, d = small
    out = str(tmp_path / "cis")
    w, B, seed = 200000, 200, 42
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    extra_args, kw = [], {}
    if model == "lmm":
        K, N, used = O.grm(dos, missm, keep)
        Kfile = _write_grm_file(str(tmp_path / "k"), K, ds.samples, np.random.default_rng(3))
        Q, lam = O.eigh(Kfile)
        kw = {"Q": Q, "lam": lam}
        extra_args = ["-k", str(tmp_path / "k")]
    r = run(["-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "-g", prefix + ".gff",
             "--model", model, "--mode", "cis", "-w", str(w), "--pval-cis", "1e-3", "--precision", "fp32",
             "--perm", str(B), "--seed", str(seed), "-o", out, *extra_args])
    assert r.returncode == 0, r.stderr
    cs, ce = np.maximum(1, ds.gene_start - w), ds.gene_end + w
    G = ds.Y.shape[1]
    refs = {}
    for k in range(G):
        idx = np.nonzero((ds.chrom == ds.gene_chrom[k]) & (ds.pos >= cs[k]) & (ds.pos <= ce[k]) & keep)[0]
        refs[f"g{k}"] = (idx, O.cis_gene(model, ds.Y[:, k], ds.X(), dos[idx], B=B, seed=seed, gene=f"g{k}", **kw))
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
        if ref["n_tested"] > 0:
            # %.10g of (1 + count) / (B + 1): the count is exact
            assert round(float(row[7]) * (B + 1)) == round(ref["p_emp"] * (B + 1)), (row[0], row[7], ref["p_emp"])
            _check_beta_columns(row, ref["perm_min_p"], ref["best_p"])
        else:
            assert row[8:11] == ["NA", "0", "0"]


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
@pytest.mark.parametrize("model", ["lm", "lmm"])
def test_cli_trans_stage2_permutations(small, tmp_path, model):
    """--mode trans --perm B (stage2_perm_topk, scan_perm.cpp:11-142): p_emp of
    every gene whose best p is below --perm-trans-thr equals the oracle's
    permutation of that gene over its top-K SNPs (those with p below the
    threshold); other genes report NA. The LMM reads its GRM from a GCTA file."""
    ds, prefix, d = small
    out = str(tmp_path / "s2")
    w, B, seed, thr, K = 200000, 200, 11, 1e-4, 20
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    extra_args, kw = [], {}
    if model == "lmm":
        Kg, N, used = O.grm(dos, missm, keep)
        Kfile = _write_grm_file(str(tmp_path / "k"), Kg, ds.samples, np.random.default_rng(5))
        Q, lam = O.eigh(Kfile)
        kw = {"Q": Q, "lam": lam}
        extra_args = ["-k", str(tmp_path / "k")]
    r = run(["-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "-g", prefix + ".gff",
             "--model", model, "--mode", "trans", "-w", str(w), "--pval-trans", "1e-5", "--precision", "fp32",
             "--perm", str(B), "--seed", str(seed), "--perm-trans-thr", str(thr), "--perm-trans-top", str(K),
             "-o", out, *extra_args])
    assert r.returncode == 0, r.stderr
    gc = ds.gene_chrom.astype(np.int32)
    cs, ce = np.maximum(1, ds.gene_start - w), ds.gene_end + w
    ref = O.trans_scan(model, dos, keep, ds.chrom, ds.pos, ds.Y, ds.X(), gene_chrom=gc, gene_cstart=cs,
                       gene_cend=ce, pthr=1.0, dense=True, hit_cap=1 << 20, **kw)
    head, region = _read_tsv(f"{out}.{model}.trans.region.tsv")
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
        o = O.cis_gene(model, ds.Y[:, k], ds.X(), dos[top], B=B, seed=seed, gene=row[0], **kw)
        assert round(float(row[7]) * (B + 1)) == round(o["p_emp"] * (B + 1)), (row[0], row[7], o["p_emp"])
        _check_beta_columns(row, o["perm_min_p"], o["best_p"])
        checked += 1
    assert checked >= 2


def test_cli_parser_errors(small, tmp_path):
    """Command-line errors are reported before any device work: unknown flags,
    a missing value, a bad choice and out-of-range numbers (the reference's
    messages); both --name=value and -xVALUE forms are accepted."""
    ds, prefix, d = small
    r = run(["-b", prefix, "--no-such-flag"])
    assert r.returncode == 1 and "unrecognized option '--no-such-flag'" in r.stderr
    r = run(["-b", prefix, "-e"])
    assert r.returncode == 1 and "requires an argument" in r.stderr
    r = run(["-b", prefix, "-e", "x", "--maf", "0.7"])
    assert r.returncode == 1 and "--maf must be in [0,0.5]" in r.stderr
    r = run(["-b", prefix, "-e", "x", "--model=lm,bad"])
    assert r.returncode == 1 and "invalid --model value: bad (lm|glm|lmm|glmm)" in r.stderr
    r = run(["-b" + prefix, "-e", "x", "--mode=sideways"])
    assert r.returncode == 1 and "invalid --mode: sideways (cis|trans|all|gw)" in r.stderr
    r = run(["-b", prefix, "-v", "y.vcf", "-e", "x"])
    assert r.returncode == 1 and "use only one of -v/--vcf or -b/--bfile" in r.stderr
    r = run(["-e", "x"])
    assert r.returncode == 1 and "need exactly one of -v/--vcf or -b/--bfile" in r.stderr
    r = run(["--help"])
    assert r.returncode == 0 and "Usage" in r.stdout


def test_cli_input_tables(tmp_path):
    """Phenotype and covariate tables as the loaders read them: a '#' header,
    whitespace-split rows, NA / NaN / '.', a repeated identical sample (warned,
    kept once), a conflicting repeat (error), covariates without a header and
    with an all-ones column (no intercept added)."""
    ds = synth.make_trans(3, n=40, m=50, G=3, n_chrom=1, ncov=2)
    prefix = str(tmp_path / "t")
    ds.write(prefix)
    lines = open(prefix + ".pheno.tsv").read().splitlines()
    lines[0] = "#" + lines[0]
    lines[2] = " ".join(lines[2].split("\t"))        # whitespace instead of tabs
    f = lines[3].split("\t")
    f[1] = "NA"
    lines[3] = "\t".join(f)
    lines.append(lines[4])                            # identical repeat
    open(prefix + ".pheno.tsv", "w").write("\n".join(lines) + "\n")
    cov = open(prefix + ".covar.tsv").read().splitlines()[1:]   # no header
    cov = [c + "\t1" for c in cov]                              # all-ones column
    open(prefix + ".covar.tsv", "w").write("\n".join(cov) + "\n")
    out = str(tmp_path / "o")
    r = run(["-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "--model", "lm",
             "--mode", "gw", "--pval-trans", "1", "-o", out, "--device", "99"])
    assert "duplicate id row with identical values (kept first): s4" in r.stderr
    assert "pheno: 40 samples, 3 genes" in r.stderr
    assert "covar: 3 columns (incl intercept if added)" in r.stderr
    lines.append(lines[5].replace(lines[5].split("\t")[1], "123.5", 1))
    open(prefix + ".pheno.tsv", "w").write("\n".join(lines) + "\n")
    r = run(["-b", prefix, "-e", prefix + ".pheno.tsv", "--model", "lm", "-o", out])
    assert r.returncode == 1 and "pheno samples: duplicate id with conflicting values: s5" in r.stderr


@pytest.mark.gpu
def test_cli_cis_thread_order(tmp_path):
    """-t > 1 runs run_cis_parallel's order (scan_cis.cpp:41-44): genes sorted
    by (chromosome as a string, TSS); -t 1 keeps the phenotype order. Both runs
    write the same lines."""
    ds = synth.make_trans(17, n=80, m=1100, G=30, n_chrom=11, ncov=2, spacing=1000)
    prefix = str(tmp_path / "t")
    ds.write(prefix)
    outs = {}
    for t in (1, 4):
        out = str(tmp_path / f"o{t}")
        r = run(["-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "-g", prefix + ".gff",
                 "--model", "lm", "--mode", "cis", "-w", "30000", "--pval-cis", "1", "-t", str(t), "-o", out])
        assert r.returncode == 0, r.stderr
        outs[t] = (_read_tsv(out + ".lm.cis.region.tsv")[1], _read_tsv(out + ".lm.cis.pairs.tsv")[1])
    reg1, reg4 = outs[1][0], outs[4][0]
    assert [r[0] for r in reg1] == [f"g{k}" for k in range(30) if any(r[0] == f"g{k}" for r in reg1)]
    key = [(str(ds.gene_chrom[int(r[0][1:])]), int(ds.gene_start[int(r[0][1:])])) for r in reg4]
    assert key == sorted(key)
    assert "10" in {k[0] for k in key} and "2" in {k[0] for k in key}
    assert sorted(map(tuple, reg1)) == sorted(map(tuple, reg4))
    # pairs: gene-major in each run's gene order, same lines
    g4 = []
    for p in outs[4][1]:
        if not g4 or g4[-1] != p[0]:
            g4.append(p[0])
    assert g4 == [r[0] for r in reg4 if int(r[3]) > 0]
    assert sorted(map(tuple, outs[1][1])) == sorted(map(tuple, outs[4][1]))
