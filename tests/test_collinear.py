"""GPU parity where the covariates explain most of a SNP, and for blocks whose
hit and polish lists outgrow their first device buffers.

Collinear cases (the conditional-analysis and genotype-PC situations):
  * covariates = intercept + a conditioning covariate c0 (a genotype g0 plus
    noise, r^2(g0, c0) = 0.999) + the top 10 eigenvectors of K + 2 N(0,1)
    columns;
  * the bed holds g0 and 47 LD proxies of it (a few calls re-drawn each, r^2
    with c0 from 0.999 down to 0.9) and SNPs whose allele frequency differs
    between two sub-populations (large r^2 with the leading eigenvectors),
    plus ordinary SNPs.
Bars (both contraction forms, LM and LMM): tested sets, n_tested, n_sig and
best SNP exact; p relative 1e-3 for p >= 1e-300; se relative 1e-4 (FP32) /
5e-4 (BF16X3); beta relative the same where |t| > 0.1.

Rank-deficient covariates (an exact copy of a column, an all-zero column, the
sum of two columns appended): where the reference's LDLT meets a zero pivot
its solves take the pseudo-inverse (model_lmm.cpp:141-151, model_lm.cpp:25)
and its REML objective is 1e300 at every delta (model_lmm.cpp:23-33). Bars:
the search's delta equal to the oracle's (1e-12 relative) for the copy and the
zero column; the scan against the oracle at the oracle's delta with the
collinear bars above, LM and LMM. The engine drops a covariate whose remaining
variance is below 1e-10 of its own. For the sum column the reference's LM
pivot is rounding noise (1e-13 of the column's scale) instead of zero, and its
(X^T X)^-1 then moves beta by orders of magnitude from the model without the
column (measured on the oracle: 1.4e4 relative), so LM is checked there
against the model without the column: beta and df-adjusted se to the bars
above (df keeps the full covariate count, as the reference's).

Overflow cases: a scan with pthr = 1 and a 100-hit caller buffer returns every
pair through eqtl_gpu_scan_hits, equal to the one-buffer run; a block with
more than 65,536 tests at |t| > 8 (the polish list's first capacity) is
polished in full and matches the oracle.
"""
import functools

import numpy as np
import pytest

import oracle as O
import eqtl_amd
from eqtl_amd import synth

from _cases import dense_from_hits, rel_err, windows

gpu = pytest.mark.gpu

TOL = {eqtl_amd.PREC_FP32: 1e-4, eqtl_amd.PREC_BF16X3: 5e-4}


def _ld_proxy(rng, g0, r2_target):
    """Copy of g0 with calls re-drawn until r^2 with g0 falls near r2_target."""
    n = g0.size
    f = g0.mean() / 2
    k = 1
    best = g0.copy()
    while True:
        g = g0.copy()
        idx = rng.choice(n, k, replace=False)
        g[idx] = rng.binomial(2, f, size=k)
        r2 = np.corrcoef(g, g0)[0, 1] ** 2
        if r2 <= r2_target or k > n // 2:
            return g if np.isfinite(r2) else best
        best = g
        k += 1


@functools.lru_cache(maxsize=None)
def collinear_case(seed=31, n=240, G=24):
    rng = np.random.default_rng(seed)
    pop = (np.arange(n) >= n // 2).astype(int)
    # sub-population SNPs: A1 frequency 0.1 vs 0.8 (strongly aligned with the top eigenvectors)
    m_pop = 300
    f_pop = np.where(pop[None, :] == 0, rng.uniform(0.05, 0.2, (m_pop, 1)), rng.uniform(0.7, 0.95, (m_pop, 1)))
    g_pop = rng.binomial(2, f_pop).astype(np.int8)
    g_rand, _ = synth.draw_genotypes(rng, 900, n, 0.05, 0.5, 0.0)
    g0 = rng.binomial(2, 0.35, size=n).astype(np.int8)
    while g0.std() == 0:
        g0 = rng.binomial(2, 0.35, size=n).astype(np.int8)
    r2s = 1 - np.geomspace(1e-2, 1e-1, 47)
    g_ld = np.stack([g0] + [_ld_proxy(rng, g0, r2) for r2 in r2s]).astype(np.int8)
    # conditioning covariate: r^2(g0, c0) = 0.999
    c0 = g0 + rng.standard_normal(n) * np.sqrt(g0.var() * (1 / 0.999 - 1))
    geno = np.concatenate([g_rand[:450], g_ld, g_pop, g_rand[450:]])
    m = geno.shape[0]
    miss = rng.random(geno.shape) < 0.005
    miss[450:450 + 48] = False  # proxies keep their exact r^2
    geno[miss] = -1
    chrom = np.repeat(np.array([1, 2], np.int32), [m // 2, m - m // 2])
    pos = np.concatenate([1 + 1000 * np.arange(m // 2), 1 + 1000 * np.arange(m - m // 2)]).astype(np.int64)
    rows = synth.pack_bed(geno)
    dos, missm, keep, af = O.decode_bed(rows, n)
    K, N, used = O.grm(dos, missm, keep)
    Q, lam = O.eigh(K)
    pcs = Q[:, ::-1][:, :10]  # top 10 eigenvectors (lambda ascending)
    C = np.hstack([c0[:, None], pcs, rng.standard_normal((n, 2))])
    X = np.hstack([np.ones((n, 1)), C])
    # phenotypes: polygenic + the conditioning SNP + a proxy effect for some genes + noise
    Z = synth.standardized(geno[rng.choice(m, 200, replace=False)])
    u = Z.T @ rng.standard_normal((200, G))
    u = (u - u.mean(0)) / u.std(0)
    h2 = rng.uniform(0.0, 0.7, G)
    Y = rng.standard_normal((n, G)) + u * np.sqrt(h2 / (1 - h2)) + 0.5 * g0[:, None]
    for k in range(0, G, 3):
        Y[:, k] += 0.4 * (g_ld[k % 48] - g_ld[k % 48].mean())
    gene_chrom = rng.integers(1, 3, G).astype(np.int32)
    gene_start = rng.integers(1, 400_000, G).astype(np.int64)
    gc, gs, ge = gene_chrom, np.maximum(1, gene_start - 50_000), gene_start + 999 + 50_000
    return dict(n=n, m=m, G=G, rows=rows, dos=dos, keep=keep, chrom=chrom, pos=pos, X=X, Y=Y, Q=Q, lam=lam,
                win=(gc, gs, ge))


@functools.lru_cache(maxsize=None)
def collinear_ref(model):
    c = collinear_case()
    kw = dict(gene_chrom=c["win"][0], gene_cstart=c["win"][1], gene_cend=c["win"][2], pthr=1e-3, dense=True,
              hit_cap=c["m"] * c["G"] + 1)
    if model == "lmm":
        kw.update(Q=c["Q"], lam=c["lam"])
    return O.trans_scan(model, c["dos"], c["keep"], c["chrom"], c["pos"], c["Y"], c["X"], **kw)


def _check(gd, st, ref, prec, what):
    tested = np.isfinite(ref["p"])
    assert np.array_equal(tested, np.isfinite(gd["p"])), what
    pr, pg = ref["p"][tested], gd["p"][tested]
    ok = pr >= 1e-300
    prel = rel_err(pg[ok], pr[ok])
    t_r = ref["stat"][tested]
    big = np.abs(t_r) > 0.1
    brel = rel_err(gd["beta"][tested][big], ref["beta"][tested][big])
    srel = rel_err(gd["se"][tested], ref["se"][tested])
    print(f"{what}: max p rel {prel.max():.2e}, beta rel {brel.max():.2e}, se rel {srel.max():.2e}, "
          f"min p {pr.min():.2e}")
    assert prel.max() < 1e-3, what
    assert brel.max() < TOL[prec], what
    assert srel.max() < TOL[prec], what
    rs = ref["state"]
    assert np.array_equal(st["n_tested"], rs["n_tested"]), what
    assert np.array_equal(st["n_sig"], rs["n_sig"]), what
    assert np.array_equal(st["best_snp"], rs["best_snp"]), what
    np.testing.assert_allclose(st["best_p"], rs["best_p"], rtol=1e-3)
    print(f"{what}: ACAT vs oracle max rel {rel_err(st['acat_p'], rs['acat_p']).max():.3e}")


@pytest.mark.parametrize("dummy", [0], ids=["cpu"])
def test_collinear_case_is_collinear(dummy):
    """The case really has SNPs that the covariates explain almost fully
    (share of the centred dosage's sum of squares in the covariates' span)."""
    c = collinear_case()
    X = c["X"]
    d = c["dos"][c["keep"].astype(bool)]
    d = d - d.mean(1, keepdims=True)
    proj = X @ np.linalg.lstsq(X, d.T, rcond=None)[0]
    r2 = (proj ** 2).sum(0) / (d ** 2).sum(1)
    assert r2.max() > 0.998 and (r2 > 0.98).sum() >= 5 and (r2 > 0.9).sum() >= 40, np.sort(r2)[-10:]


@gpu
@pytest.mark.parametrize("prec", [eqtl_amd.PREC_FP32, eqtl_amd.PREC_BF16X3])
@pytest.mark.parametrize("model", ["lm", "lmm"])
def test_collinear_parity(gpu_ctx, model, prec):
    c = collinear_case()
    ref = collinear_ref(model)
    gpu_ctx.set_samples(c["n"])
    gpu_ctx.set_policy(True, 0.8, 0.0)
    gpu_ctx.set_snp_annot(c["chrom"], c["pos"])
    if model == "lmm":
        gpu_ctx.set_basis(c["Q"], c["lam"])
        gpu_ctx.lmm_prep(c["Y"], c["X"], delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=ref["state"]["delta"])
        mc = eqtl_amd.MODEL_LMM
    else:
        gpu_ctx.lm_prep(c["Y"], c["X"])
        mc = eqtl_amd.MODEL_LM
    gpu_ctx.set_gene_windows(*c["win"])
    hits, st = gpu_ctx.scan(mc, prec, c["rows"], pthr=1.0, block=512, hit_cap=1 << 16)
    gd = dense_from_hits(hits, c["m"], c["G"])
    # n_sig at the oracle's threshold
    st2 = st.copy()
    sig = np.zeros(c["G"], np.int32)
    np.add.at(sig, hits["gene"][hits["p"] <= 1e-3], 1)
    st2["n_sig"] = sig
    _check(gd, st2, ref, prec, f"collinear {model} prec={prec}")


def _rank_deficient(X, kind):
    if kind == "copy":
        return np.column_stack([X, X[:, 1]])
    if kind == "zero":
        return np.column_stack([X, np.zeros(X.shape[0])])
    return np.column_stack([X, X[:, 1] + X[:, 2]])


@gpu
@pytest.mark.parametrize("kind", ["copy", "zero", "sum"])
@pytest.mark.parametrize("model", ["lm", "lmm"])
def test_rank_deficient_covariates(gpu_ctx, model, kind):
    c = collinear_case()
    X = _rank_deficient(c["X"], kind)
    kw = dict(gene_chrom=c["win"][0], gene_cstart=c["win"][1], gene_cend=c["win"][2], pthr=1e-3, dense=True,
              hit_cap=c["m"] * c["G"] + 1)
    if model == "lmm":
        kw.update(Q=c["Q"], lam=c["lam"])
    lm_sum = model == "lm" and kind == "sum"
    ref = O.trans_scan(model, c["dos"], c["keep"], c["chrom"], c["pos"], c["Y"], c["X"] if lm_sum else X, **kw)
    gpu_ctx.set_samples(c["n"])
    gpu_ctx.set_policy(True, 0.8, 0.0)
    gpu_ctx.set_snp_annot(c["chrom"], c["pos"])
    if model == "lmm":
        gpu_ctx.set_basis(c["Q"], c["lam"])
        if kind != "sum":
            delta, _ = gpu_ctx.lmm_prep(c["Y"], X, delta_mode=eqtl_amd.DELTA_REF)
            print(f"rank-deficient ({kind}): delta {delta[:3]} oracle {ref['state']['delta'][:3]}", flush=True)
            assert rel_err(delta, ref["state"]["delta"]).max() < 1e-12
        gpu_ctx.lmm_prep(c["Y"], X, delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=ref["state"]["delta"])
        mc = eqtl_amd.MODEL_LMM
    else:
        gpu_ctx.lm_prep(c["Y"], X)
        mc = eqtl_amd.MODEL_LM
    gpu_ctx.set_gene_windows(*c["win"])
    prec = eqtl_amd.PREC_BF16X3
    hits, st = gpu_ctx.scan(mc, prec, c["rows"], pthr=1.0, block=512, hit_cap=1 << 16)
    gd = dense_from_hits(hits, c["m"], c["G"])
    if lm_sum:
        tested = np.isfinite(ref["p"])
        assert np.array_equal(tested, np.isfinite(gd["p"]))
        assert np.array_equal(st["n_tested"], ref["state"]["n_tested"])
        df_full, df_red = c["n"] - X.shape[1] - 1, c["n"] - c["X"].shape[1] - 1
        big = tested & (np.abs(ref["stat"]) > 0.1)
        brel = rel_err(gd["beta"][big], ref["beta"][big])
        srel = rel_err(gd["se"][tested], ref["se"][tested] * np.sqrt(df_red / df_full))
        print(f"rank-deficient sum lm vs the model without the column: beta rel {brel.max():.2e}, se rel {srel.max():.2e}")
        assert brel.max() < TOL[prec] and srel.max() < TOL[prec]
        return
    st2 = st.copy()
    sig = np.zeros(c["G"], np.int32)
    np.add.at(sig, hits["gene"][hits["p"] <= 1e-3], 1)
    st2["n_sig"] = sig
    _check(gd, st2, ref, prec, f"rank-deficient {kind} {model}")


@gpu
def test_hits_drained_past_caller_buffer(gpu_ctx):
    """pthr = 1 with a 100-hit caller buffer: every pair comes back, in order."""
    c = collinear_case()
    gpu_ctx.set_samples(c["n"])
    gpu_ctx.set_snp_annot(c["chrom"], c["pos"])
    gpu_ctx.lm_prep(c["Y"], c["X"])
    gpu_ctx.set_gene_windows(*c["win"])
    full, st_full = gpu_ctx.scan(eqtl_amd.MODEL_LM, eqtl_amd.PREC_FP32, c["rows"], pthr=1.0, block=2048,
                                 hit_cap=1 << 20)
    small, st_small = gpu_ctx.scan(eqtl_amd.MODEL_LM, eqtl_amd.PREC_FP32, c["rows"], pthr=1.0, block=2048,
                                   hit_cap=100)
    assert full.size > 10_000
    assert np.array_equal(full, small)
    assert np.array_equal(st_full, st_small)
    order = np.lexsort((small["gene"], small["snp"]))
    assert np.array_equal(order, np.arange(small.size))


@gpu
def test_polish_list_grows(gpu_ctx):
    """More tests at |t| > 8 in one block than the polish list first holds."""
    rng = np.random.default_rng(41)
    n, m, G = 200, 1100, 72
    g0 = rng.binomial(2, 0.4, size=n).astype(np.int8)
    geno = np.stack([_ld_proxy(rng, g0, 0.97) if j % 50 else g0 for j in range(m)]).astype(np.int8)
    rows = synth.pack_bed(geno)
    dos, missm, keep, af = O.decode_bed(rows, n)
    # kinship from independent SNPs (a GRM of these LD copies would absorb the effect)
    gk, _ = synth.draw_genotypes(rng, 1500, n, 0.05, 0.5, 0.0)
    dk, mk, kk, _ = O.decode_bed(synth.pack_bed(gk), n)
    K, N, used = O.grm(dk, mk, kk)
    Q, lam = O.eigh(K)
    X = np.hstack([np.ones((n, 1)), rng.standard_normal((n, 2))])
    Y = 1.5 * (g0[:, None] - g0.mean()) + rng.standard_normal((n, G))
    chrom = np.ones(m, np.int32)
    pos = 1 + 1000 * np.arange(m, dtype=np.int64)
    ref = O.trans_scan("lmm", dos, keep, chrom, pos, Y, X, Q=Q, lam=lam, pthr=1e-30, dense=True, hit_cap=m * G + 1)
    assert (np.abs(ref["stat"]) > 8).sum() > 65_536
    gpu_ctx.set_samples(n)
    gpu_ctx.set_snp_annot(chrom, pos)
    gpu_ctx.set_basis(Q, lam)
    gpu_ctx.lmm_prep(Y, X, delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=ref["state"]["delta"])
    hits, st = gpu_ctx.scan(eqtl_amd.MODEL_LMM, eqtl_amd.PREC_BF16X3, rows, pthr=1.0, block=2048, hit_cap=1 << 10)
    gd = dense_from_hits(hits, m, G)
    tested = np.isfinite(ref["p"])
    assert np.array_equal(tested, np.isfinite(gd["p"]))
    bad = np.argwhere(rel_err(gd["p"], ref["p"]) > 1e-3)
    for j, k in bad[:10]:
        print("mismatch", j, k, ref["p"][j, k], gd["p"][j, k], ref["stat"][j, k], gd["stat"][j, k])
    ok = tested & (ref["p"] >= 1e-300)
    prel = rel_err(gd["p"][ok], ref["p"][ok])
    print(f"polish growth: {hits.size} hits, max p rel {prel.max():.2e}, min p {ref['p'][ok].min():.2e}")
    # the same block at a reporting threshold: exactly the pairs with p <= 1e-5
    h5, st5 = gpu_ctx.scan(eqtl_amd.MODEL_LMM, eqtl_amd.PREC_BF16X3, rows, pthr=1e-5, block=2048, hit_cap=1 << 10)
    want = hits[hits["p"] <= 1e-5]
    print("pthr 1e-5:", h5.size, "hits, expected", want.size)
    if h5.size != want.size or not np.array_equal(h5[["snp", "gene"]], want[["snp", "gene"]]):
        s5 = set(zip(h5["snp"].tolist(), h5["gene"].tolist()))
        miss = [(int(a), int(b)) for a, b in zip(want["snp"], want["gene"]) if (a, b) not in s5]
        print("missing", len(miss), miss[:20])
    assert np.array_equal(h5, want)
    assert prel.max() < 1e-3
    rs = ref["state"]
    assert np.array_equal(st["n_tested"], rs["n_tested"])
    sig = np.zeros(G, np.int32)
    np.add.at(sig, hits["gene"][hits["p"] <= 1e-30], 1)
    assert np.array_equal(sig, rs["n_sig"])
    assert np.array_equal(st["best_snp"], rs["best_snp"])
    np.testing.assert_allclose(st["best_p"], rs["best_p"], rtol=1e-3)
