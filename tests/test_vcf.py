"""VCF genotype source (VcfSession::parse_record, vcf_session.cpp:160-286).

CPU: the oracle's record restatement against the reference's own fixture
(INFO AC/AN of data/smoke.vcf.gz, copied under tests/golden/smoke/),
hand-built lines with known answers, and the PLINK decode of the same
diploid calls; the scalar parsing shared with the kernel (host build).
GPU: eqtl_gpu_vcf_decode rows bit-exact against the oracle, the float rows
through decode / GRM, and the CLI on the reference's smoke data (the steps of
scripts/run_smoke.sh, and the OLS check of scripts/gold_lm.py on every pair).
"""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import eqtl_amd
from eqtl_amd import _lib
from _vcf import SMOKE, descriptors, gpu_rows, read_vcf, synth_vcf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "eqtl_amd", "eqtl")
SMOKE_VCF = os.path.join(SMOKE, "smoke.vcf.gz")


# ----------------------------------------------------------------------------- CPU


def test_oracle_matches_fixture_ac_an():
    """Every record of the reference's smoke VCF: ALT dosage sum == INFO AC and
    called allele copies == INFO AN (the producing tool's counts over all
    samples); kept records report af == AC/AN."""
    samples, ds_float, lines = read_vcf(SMOKE_VCF)
    assert len(samples) == 100 and len(lines) == 691 and not ds_float
    n_kept = 0
    for ln in lines:
        info = dict(kv.split("=") for kv in ln.split(b"\t")[7].decode().split(";"))
        st, d, af, (nm, s, cp), raw = O.vcf_record(ln, 100)
        assert s == float(info["AC"]) and cp == float(info["AN"]), ln[:60]
        assert nm == int(np.isnan(raw).sum())
        if st == 1:
            n_kept += 1
            assert af == float(info["AC"]) / float(info["AN"])
        else:
            assert st == 0 and (int(info["AC"]) == 0 or info["AC"] == info["AN"])
    assert n_kept == 688


def _line(chrom, alt, fmt, cells):
    return "\t".join([chrom, "100", "rs1", "A", alt, ".", "PASS", ".", fmt] + cells)


def test_oracle_known_answers():
    nan = float("nan")
    # diploid / haploid calls, missing alleles, polyploid (first two alleles)
    st, d, af, (nm, s, cp), raw = O.vcf_record(_line("1", "T", "GT", ["0|1", "1/1", "./.", "1", "1/.", "0/1/1"]), 6)
    np.testing.assert_array_equal(raw, [1, 2, nan, 1, nan, 1])
    assert (nm, s, cp) == (2, 5.0, 7.0) and st == 1 and af == 5 / 7
    np.testing.assert_array_equal(d, [1, 2, 1.25, 1, 1.25, 1])  # mean imputation
    # haploid chromosome: one copy per call whatever the ploidy written
    st, d, af, (nm, s, cp), raw = O.vcf_record(_line("chrX", "T", "GT", ["0|1", "1", "0", "1/1"]), 4)
    assert (nm, s, cp) == (0, 4.0, 4.0) and st == 0  # af = 1: monomorphic
    st, d, af, (nm, s, cp), raw = O.vcf_record(_line("MT", "T", "GT", ["0", "1", "0", "0"]), 4)
    assert (s, cp, af, st) == (1.0, 4.0, 0.25, 1)
    # DS used only when the header declares FORMAT/DS as Float
    cells = ["0/0:0.5", "1/1:.", "0/1", "1/1:1.25e0"]
    st, d, af, (nm, s, cp), raw = O.vcf_record(_line("1", "T", "GT:DS", cells), 4, ds_float=True)
    np.testing.assert_array_equal(raw, [0.5, nan, nan, 1.25])
    assert (nm, s, cp, st) == (2, 1.75, 4.0, 1)
    st, d, af, (nm, s, cp), raw = O.vcf_record(_line("1", "T", "GT:DS", cells), 4, ds_float=False)
    np.testing.assert_array_equal(raw, [0, 2, 1, 2])
    # DS is a float: 0.1 -> float32(0.1)
    st, d, af, _, raw = O.vcf_record(_line("1", "T", "DS", ["0.1", "1.9", "0.3", "0.7"]), 4, ds_float=True)
    assert raw[0] == float(np.float32(0.1)) and raw[1] == float(np.float32(1.9))
    # GT key not first; sample columns shorter than FORMAT
    st, d, af, (nm, s, cp), raw = O.vcf_record(_line("1", "T", "DP:GT", ["7:0|1", "8:1|1", "9", "3:0/0"]), 4)
    np.testing.assert_array_equal(raw, [1, 2, nan, 0])
    # skips and column errors
    assert O.vcf_record(_line("1", "T,G", "GT", ["0|1", "1|1"]), 2)[0] == -1
    assert O.vcf_record(_line("1", ".", "GT", ["0|0", "0|0"]), 2)[0] == -1
    assert O.vcf_record(_line("1", "T", "GT", ["0|1", "1|1"]), 3)[0] == -2
    # policies: max_miss, Filter (no imputation), MAF gate, sample subset order
    ln = _line("1", "T", "GT", ["0|1", "./.", "0|0", "0|0", "0|0"])
    assert O.vcf_record(ln, 5, max_miss=0.19)[0] == 0 and O.vcf_record(ln, 5, max_miss=0.2)[0] == 1
    assert O.vcf_record(ln, 5, impute=False)[0] == 0
    assert O.vcf_record(ln, 5, maf_min=0.126)[0] == 0 and O.vcf_record(ln, 5, maf_min=0.125)[0] == 1
    st, d, af, _, raw = O.vcf_record(ln, 5, sample_col=[4, 0, 2])
    np.testing.assert_array_equal(raw, [0, 1, 0])
    assert af == 1 / 6


def _bed_code(a0, a1):
    # PLINK-1 2-bit code of an ALT count (A1 = ALT): {2 -> 00, 1 -> 10, 0 -> 11, missing -> 01}
    if a0 is None or a1 is None:
        return 1
    return {2: 0, 1: 2, 0: 3}[a0 + a1]


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_diploid_gt_equals_bed_decode(seed):
    """Diploid biallelic GT calls give the same dosage, keep and AF through the
    VCF restatement as through the PLINK decode of the same calls."""
    rng = np.random.default_rng(seed)
    n_file, m = 37, 200
    toks = {"0|0": (0, 0), "0/1": (0, 1), "1|0": (1, 0), "1/1": (1, 1), "./.": (None, None),
            ".|1": (None, 1), "0/.": (0, None)}
    names = list(toks)
    sc = rng.permutation(n_file)[:29].astype(np.int32)
    for impute, max_miss, maf in ((1, 0.8, 0.0), (1, 0.05, 0.1), (0, 0.8, 0.0)):
        rows = np.zeros((m, (n_file + 3) // 4), np.uint8)
        lines = []
        for j in range(m):
            p = rng.dirichlet(np.ones(len(names)) * (0.3 if j % 5 == 0 else 2.0))
            cells = [names[k] for k in rng.choice(len(names), size=n_file, p=p)]
            for i, c in enumerate(cells):
                rows[j, i // 4] |= _bed_code(*toks[c]) << (2 * (i % 4))
            lines.append(_line("2", "C", "GT", cells))
        dos, miss, keep, af = O.decode_bed(rows, n_file, sample_col=sc, impute=impute, max_miss=max_miss,
                                           maf_min=maf)
        for j, ln in enumerate(lines):
            st, d, a, _, raw = O.vcf_record(ln, n_file, sample_col=sc, impute=impute, max_miss=max_miss,
                                            maf_min=maf)
            assert (st == 1) == bool(keep[j])
            np.testing.assert_array_equal(np.isnan(raw), miss[j].astype(bool))
            if st == 1:
                np.testing.assert_array_equal(d, dos[j])
                assert a == af[j]


def test_vcf_parse_host_check(tmp_path):
    """The kernel's scalar parsing built for the host: tab masks against a byte
    loop; decimal -> float against strtod -> float."""
    exe = tmp_path / "vcf_parse_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "vcf_parse_check.cpp"), "-o",
                           str(exe)])
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr


def test_cli_needs_exactly_one_genotype_source(tmp_path):
    r = subprocess.run([CLI, "-v", SMOKE_VCF, "-b", "x", "-e", "y", "-o", str(tmp_path / "o")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "use only one of -v/--vcf or -b/--bfile" in r.stderr
    r = subprocess.run([CLI, "-e", "y", "-o", str(tmp_path / "o")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "need exactly one of -v/--vcf or -b/--bfile" in r.stderr


# ----------------------------------------------------------------------------- GPU


def _check_rows(rows, kept, lines, n_file, sc, ds_float, **pol):
    """Device rows against the oracle: bit-exact dosages of the analysis
    samples, NaN where missing, all-NaN rows for rejected records."""
    n_rej = 0
    for r, j in enumerate(kept):
        st, d, af, _, raw = O.vcf_record(lines[j], n_file, sample_col=sc, ds_float=ds_float, **pol)
        assert st in (0, 1)
        got = rows[r, sc].astype(np.float64)
        if st == 1:
            np.testing.assert_array_equal(got, raw, err_msg=f"record {j}")
        else:
            n_rej += 1
            assert np.isnan(rows[r]).all(), f"record {j}"
    return n_rej


@pytest.mark.gpu
def test_vcf_decode_smoke_fixture(gpu_ctx):
    samples, ds_float, lines = read_vcf(SMOKE_VCF)
    n_file = len(samples)
    rng = np.random.default_rng(5)
    for sc, pol in ((np.arange(n_file, dtype=np.int32), {}),
                    (rng.permutation(n_file)[:61].astype(np.int32), {"max_miss": 0.04, "maf_min": 0.05}),
                    (rng.permutation(n_file)[:40].astype(np.int32), {"impute": False})):
        print("smoke fixture", sc.size, pol, flush=True)
        gpu_ctx.set_samples(n_file, sc)
        gpu_ctx.set_policy(impute=pol.get("impute", True), max_miss=pol.get("max_miss", 0.8),
                           maf_min=pol.get("maf_min", 0.0))
        rows, kept = gpu_rows(gpu_ctx, lines, ds_float, n_file)
        assert len(kept) == len(lines)
        _check_rows(rows, kept, lines, n_file, sc, ds_float, impute=pol.get("impute", True),
                    max_miss=pol.get("max_miss", 0.8), maf_min=pol.get("maf_min", 0.0))


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gt", "ds", "wide"])
def test_vcf_decode_synthetic(gpu_ctx, case):
    """GT with haploid chromosome X, haploid and polyploid calls, '.' alleles,
    short columns, skipped multi-allelic records; GT:DS with decimals in
    several forms and '.'; a wide record (20,000 samples: not staged in LDS)."""
    rng = np.random.default_rng({"gt": 11, "ds": 12, "wide": 13}[case])
    n_file, m = {"gt": (203, 400), "ds": (150, 300), "wide": (20000, 6)}[case]
    hdr, lines = synth_vcf(rng, n_file, m, ds=(case == "ds"))
    ds_float = case == "ds"
    sc = rng.permutation(n_file)[: n_file - 7].astype(np.int32)
    for pol in ({"impute": True, "max_miss": 0.8, "maf_min": 0.0}, {"impute": True, "max_miss": 0.1, "maf_min": 0.1}):
        print("synthetic", case, pol, flush=True)
        gpu_ctx.set_samples(n_file, sc)
        gpu_ctx.set_policy(**pol)
        rows, kept = gpu_rows(gpu_ctx, lines, ds_float, n_file)
        assert len(kept) < len(lines) or case == "wide"
        n_rej = _check_rows(rows, kept, lines, n_file, sc, ds_float, **pol)
        assert 0 < n_rej < len(kept) or case == "wide"


@pytest.mark.gpu
def test_vcf_rows_through_decode_and_grm(gpu_ctx):
    """DOSE32 rows through the block decode (mean imputation, subset AF) and
    the GRM, against the oracle's records + subset_maf_or_nan + GRM."""
    rng = np.random.default_rng(21)
    n_file, m = 96, 700
    hdr, lines = synth_vcf(rng, n_file, m, ds=True, chroms=("3",))
    # records the source rejects (AF 0) and ones only subset_maf_or_nan rejects (no variance)
    for j in range(5):
        lines.insert(100 * j + 3, _line("3", "T", "GT:DS", ["0/0:0"] * n_file).encode())
        lines.insert(100 * j + 7, _line("3", "T", "GT:DS", ["0/1:1.0"] * (n_file - 3) + ["0/1:."] * 3).encode())
    sc = rng.permutation(n_file)[:90].astype(np.int32)
    gpu_ctx.set_samples(n_file, sc)
    gpu_ctx.set_policy(impute=True, max_miss=0.5, maf_min=0.01)
    gpu_ctx.set_geno_format(_lib.GENO_DOSE32)
    try:
        text, recs, kept = descriptors(lines, True)
        ptr = gpu_ctx.vcf_decode(text, recs)
        rows = gpu_ctx.copy_rows(ptr, 0, recs.size, 4 * n_file)
        dos, keep, af = gpu_ctx.decode_bed(rows)
        ref = [O.vcf_record(lines[j], n_file, sample_col=sc, ds_float=True, max_miss=0.5, maf_min=0.01) for j in kept]
        n = sc.size
        rd = np.zeros((len(kept), n))
        rmiss = np.zeros((len(kept), n), np.int8)
        src = np.zeros(len(kept), np.uint8)   # kept by the source (GRM input)
        rkeep = np.zeros(len(kept), bool)     # and by subset_maf_or_nan (tested)
        for r, (st, d, a, _, raw) in enumerate(ref):
            if st != 1:
                continue
            rd[r], src[r], rmiss[r] = d, 1, np.isnan(raw)
            saf, _ = O.subset_maf_or_nan(d)
            if np.isnan(saf):
                continue
            rkeep[r] = True
            assert keep[r]
            np.testing.assert_allclose(dos[r], d, rtol=1e-15, atol=1e-15)
            assert af[r] == pytest.approx(saf, rel=1e-14)
        assert (keep == rkeep).all()
        assert 0 < rkeep.sum() < src.sum() < len(kept)
        K, N, used = gpu_ctx.grm(rows)
        Kr, Nr, used_r = O.grm(rd, rmiss, src)
        assert used == used_r
        np.testing.assert_array_equal(N, Nr)
        np.testing.assert_allclose(K, Kr, rtol=1e-12, atol=1e-12)
    finally:
        gpu_ctx.set_geno_format(_lib.GENO_BED)


@pytest.mark.gpu
def test_vcf_decode_column_count_error(gpu_ctx):
    """A record whose sample column count differs from the header's is an
    EQTL_E_FORMAT error naming it (htslib read error in the reference)."""
    n_file = 8
    good = _line("1", "T", "GT", ["0|1"] * n_file).encode()
    bad = _line("1", "T", "GT", ["0|1"] * (n_file - 1)).encode()
    gpu_ctx.set_samples(n_file)
    gpu_ctx.set_policy()
    text, recs, kept = descriptors([good, good, bad, good], False)
    with pytest.raises(eqtl_amd.EqtlGpuError) as ei:
        gpu_ctx.vcf_decode(text, recs)
    assert ei.value.code == _lib.EQTL_E_FORMAT and "VCF record 2" in str(ei.value)
    text, recs, kept = descriptors([good, good], False)
    gpu_ctx.vcf_decode(text, recs)


def _read_tsv(path):
    with open(path) as f:
        head = f.readline().rstrip("\n").split("\t")
        return head, [ln.rstrip("\n").split("\t") for ln in f if ln.strip()]


def _pheno(path):
    with open(path) as f:
        genes = f.readline().rstrip("\n").split("\t")[1:]
        ids, Y = [], []
        for ln in f:
            t = ln.rstrip("\n").split("\t")
            ids.append(t[0])
            Y.append([float(x) for x in t[1:]])
    return ids, genes, np.array(Y)


@pytest.mark.gpu
def test_cli_smoke_grm_and_cis_lm(tmp_path):
    """scripts/run_smoke.sh on data/smoke: --make-grm from the VCF (against the
    oracle GRM of the oracle's dosages) and a cis LM scan with --pval-cis 1
    whose every pair matches numpy OLS on the mean-imputed GT dosage
    (scripts/gold_lm.py's method and tolerances: |dbeta|, |dse| <= 1e-4,
    |dlog10 p| <= 0.05; here also 1e-4 relative p)."""
    samples, ds_float, lines = read_vcf(SMOKE_VCF)
    n = len(samples)
    out = str(tmp_path / "smoke_grm")
    r = subprocess.run([CLI, "-v", SMOKE_VCF, "--make-grm", "-o", out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ref = [O.vcf_record(ln, n) for ln in lines]
    dos = np.array([d for st, d, *_ in ref])
    miss = np.array([np.isnan(raw) for *_, raw in ref]).astype(np.int8)
    keep = np.array([st == 1 for st, *_ in ref]).astype(np.uint8)
    Kr, Nr, used = O.grm(np.nan_to_num(dos), miss, keep)
    ids = [ln.split("\t")[1] for ln in open(out + ".grm.id").read().splitlines()]
    assert ids == samples
    tri = np.fromfile(out + ".grm.bin", dtype=np.float32)
    il = np.tril_indices(n)  # GCTA: lower triangle, row-major
    np.testing.assert_allclose(tri, Kr[il].astype(np.float32), rtol=2e-6, atol=2e-6)

    out = str(tmp_path / "smoke_lm")
    r = subprocess.run([CLI, "-v", SMOKE_VCF, "-e", os.path.join(SMOKE, "smoke.pheno.tsv"), "-g",
                        os.path.join(SMOKE, "smoke.gff"), "--model", "lm", "--mode", "cis", "--perm", "0",
                        "--pval-cis", "1", "--miss-hand", "impute", "--precision", "fp32", "-o", out, "-t", "1"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    head, pairs = _read_tsv(out + ".lm.cis.pairs.tsv")
    assert len(pairs) > 100
    pid, genes, Y = _pheno(os.path.join(SMOKE, "smoke.pheno.tsv"))
    order = [samples.index(s) for s in pid]
    snp_row = {}
    for j, ln in enumerate(lines):
        t = ln.split(b"\t", 3)
        snp_row[t[2].decode()] = j
    ib, ise, ip = head.index("beta"), head.index("se"), head.index("p")
    worst = 0.0
    for p in pairs:
        j = snp_row[p[1]]
        st, d, af, _, raw = O.vcf_record(lines[j], n, sample_col=np.array(order, np.int32))
        assert st == 1
        y = Y[:, genes.index(p[0])]
        A = np.column_stack([np.ones(len(y)), d])
        coef, *_ = np.linalg.lstsq(A, y, rcond=None)
        res = y - A @ coef
        s2 = res @ res / (len(y) - 2)
        se = math.sqrt(s2 * np.linalg.inv(A.T @ A)[1, 1])
        b_eq, se_eq, p_eq = float(p[ib]), float(p[ise]), float(p[ip])
        assert abs(b_eq - coef[1]) <= 1e-4 and abs(se_eq - se) <= 1e-4
        worst = max(worst, abs(b_eq - coef[1]) / se)  # error in units of the standard error
        # t tail p (the reference's) vs OLS t: same statistic
        t_ols = coef[1] / se
        p_ols = O.p_from_t(np.array([t_ols]), len(y) - 2)[0]
        assert abs(math.log10(p_eq) - math.log10(p_ols)) <= 0.05
        # the engine's bound: 1e-3 relative p against FP64 (FP32 contraction here)
        assert p_eq == pytest.approx(p_ols, rel=1e-4)
    assert worst < 1e-5
