"""The five benchmark configurations at their full sizes, on the GPU through
the C ABI, against the FP64 oracle (eqtl_amd.workloads builds the inputs).

  C1 trans LM  n=200 x 10k SNPs x 500 genes: every test against the oracle,
     both contraction forms: tested sets, n_tested, n_sig (p <= 1e-5) and the
     best SNP exact; p relative 1e-3; beta / se relative 1e-4 (FP32) /
     5e-4 (BF16X3), beta where |t| > 0.1.
  C2 cis LMM   n=1000 x 100k SNPs x 2000 genes (+-1 Mb windows, reference
     delta search): GRM sub-block against the oracle (1e-12, N exact), the
     eigenbasis invariants, then 20 genes spread over the gene list against the
     oracle's scan_gene_snps: delta (f no worse than the oracle's, relative
     1e-9), every window test's p relative 1e-3, n_tested and best SNP exact.
  C3 trans LMM n=2000 x 500k SNPs x 20k genes, split form, batched Newton
     delta: n_tested of every gene from the window geometry; on a sample of
     pairs (128 genes x ~400 SNPs + the planted SNPs) the tested set exact and
     p relative 1e-3 for p >= 1e-300; REML f at the Newton delta no worse than
     at the reference search's delta for all 20k genes (1e-9 relative).
  C4 cis LM + LMM, 20k windows of 50-5000 SNPs, B = 1000, split form: on 40
     genes stratified by window size (LM) and 12 genes (LMM, windows up to
     1500 SNPs, the engine's own eigenbasis given to the oracle) every
     permutation minimum relative 1e-3 and p_emp identical.
  C5 GRM n=5000 x 1M SNPs: a 48-sample sub-block of K against the oracle
     (1e-12) with N exact, also with 1% missing calls (200k SNPs); REML
     f / f' / f'' in ln delta for all 20k genes against central differences
     of f (steps 1e-4 / 3e-3, noise model below), and f against the oracle's
     restatement on 200 genes (1e-9 relative).
"""
import functools
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
import eqtl_amd
from eqtl_amd import synth, workloads

from _cases import dense_from_hits, rel_err

pytestmark = pytest.mark.gpu

BLOCK = 32768
THREADS = min(16, os.cpu_count() or 1)


def _pmap(fn, items):
    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        return list(ex.map(fn, items))


def _scan_dev(ctx, w, model, prec, pthr, scope=eqtl_amd.SCOPE_TRANS, hit_cap=1 << 20):
    ctx.scan_begin(model, prec, pthr, scope)
    out = []
    for j0 in range(0, w.m, BLOCK):
        mb = min(BLOCK, w.m - j0)
        out.append(ctx.scan_block(None, j0, hit_cap=hit_cap, rows_dev=w.rows_dev + j0 * w.bps, m=mb, bps=w.bps))
    st = ctx.scan_end()
    return np.concatenate(out), st


def _grm_dev(ctx, w, m=None):
    m = w.m if m is None else m
    ctx.grm_begin()
    for j0 in range(0, m, BLOCK):
        mb = min(BLOCK, m - j0)
        ctx.grm_accumulate_dev(w.rows_dev + j0 * w.bps, mb, w.bps)
    return ctx.grm_finish()


def _grm_subblock_check(ctx, w, K, N, m, idx):
    """K[idx, idx] and N against the GCTA closed form of that sub-block over
    all m SNPs (allele frequencies over all n samples; synth.grm_subblock,
    pinned against the oracle's GRM in the CPU tests)."""
    rows = ctx.copy_rows(w.rows_dev, 0, m, w.bps)
    Ko, No, used = synth.grm_subblock(rows, w.n, idx)
    dk = np.abs(K[np.ix_(idx, idx)] - Ko).max()
    Nfull = np.zeros((w.n, w.n), np.int64)
    Nfull[np.tril_indices(w.n)] = N  # GCTA order: row-major lower triangle (i >= j)
    Nfull = np.tril(Nfull) + np.tril(Nfull, -1).T
    return dk, np.array_equal(Nfull[np.ix_(idx, idx)], No)


# ---------------------------------------------------------------- C1 -------
@functools.lru_cache(maxsize=None)
def _c1_ref():
    w = workloads.c1()
    dos, _, keep, _ = O.decode_bed(w.rows, w.n)
    ref = O.trans_scan("lm", dos, keep, w.chrom, w.pos, w.Y, w.X, gene_chrom=w.win[0], gene_cstart=w.win[1],
                       gene_cend=w.win[2], pthr=1e-5, dense=True, hit_cap=1 << 16)
    return w, ref


@pytest.mark.parametrize("prec", [eqtl_amd.PREC_FP32, eqtl_amd.PREC_BF16X3])
def test_c1_trans_lm_full(gpu_ctx, prec):
    w, ref = _c1_ref()
    gpu_ctx.set_samples(w.n)
    gpu_ctx.set_policy(True, 0.8, 0.0)
    gpu_ctx.set_snp_annot(w.chrom, w.pos)
    gpu_ctx.lm_prep(w.Y, w.X)
    gpu_ctx.set_gene_windows(*w.win)
    hits, st = gpu_ctx.scan(eqtl_amd.MODEL_LM, prec, w.rows, pthr=1.0, block=4096, hit_cap=1 << 20)
    G = w.Y.shape[1]
    gd = dense_from_hits(hits, w.m, G)
    tested = np.isfinite(ref["p"])
    assert np.array_equal(tested, np.isfinite(gd["p"]))
    assert tested.sum() > 4e6
    pr, pg = ref["p"][tested], gd["p"][tested]
    ok = pr >= 1e-300
    prel = rel_err(pg[ok], pr[ok])
    big = np.abs(ref["stat"][tested]) > 0.1
    tol = 1e-4 if prec == eqtl_amd.PREC_FP32 else 5e-4
    brel = rel_err(gd["beta"][tested][big], ref["beta"][tested][big])
    srel = rel_err(gd["se"][tested], ref["se"][tested])
    print(f"C1 prec={prec}: {tested.sum()} tests, max p rel {prel.max():.2e}, beta {brel.max():.2e}, "
          f"se {srel.max():.2e}, min p {pr.min():.2e}")
    assert prel.max() < 1e-3 and brel.max() < tol and srel.max() < tol
    rs = ref["state"]
    assert np.array_equal(st["n_tested"], rs["n_tested"])
    sig = np.zeros(G, np.int32)
    np.add.at(sig, hits["gene"][hits["p"] <= 1e-5], 1)
    assert np.array_equal(sig, rs["n_sig"])
    assert np.array_equal(st["best_snp"], rs["best_snp"])
    # the planted trans effects are found
    for k, (j, beta) in w.extra["planted"].items():
        assert st["best_snp"][k] == j or rs["best_snp"][k] != j


# ---------------------------------------------------------------- C2 -------
def test_c2_cis_lmm_full(gpu_ctx):
    ctx = gpu_ctx
    w = workloads.c2(ctx)
    ctx.set_samples(w.n)
    ctx.set_policy(True, 0.8, 0.0)
    K, N, used = _grm_dev(ctx, w)
    idx = np.sort(np.random.default_rng(2).choice(w.n, 48, replace=False))
    dk, n_ok = _grm_subblock_check(ctx, w, K, N, w.m, idx)
    print(f"C2 GRM: {used} SNPs, 48-sample sub-block max |dK| {dk:.2e}")
    assert dk < 1e-12 and n_ok
    Q, lam = ctx.eigh(K)
    # make_lmm_basis clamps lambda at 0 (model_lmm.cpp:106-113): the pairwise
    # divisors leave K slightly indefinite, so the reconstruction is off by the
    # clamped part only
    lam_np = np.linalg.eigvalsh(K)
    np.testing.assert_allclose(lam, np.maximum(lam_np, 0), atol=1e-10 * lam_np.max())
    assert np.abs((Q * lam) @ Q.T - K).max() < 1e-9 * np.abs(K).max() + max(0.0, -lam_np.min())
    assert np.abs(Q.T @ Q - np.eye(w.n)).max() < 1e-10
    ctx.set_snp_annot(w.chrom, w.pos)
    delta, _ = ctx.lmm_prep(w.Y, w.X, delta_mode=eqtl_amd.DELTA_REF)
    ctx.set_gene_windows(*w.win)
    hits, st = _scan_dev(ctx, w, eqtl_amd.MODEL_LMM, eqtl_amd.PREC_BF16X3, 1.0, eqtl_amd.SCOPE_CIS)
    G = w.Y.shape[1]
    assert hits.size == int(st["n_tested"].sum()) > 1.5e6
    genes = list(range(0, G, G // 20))[:20]
    gch, cs, ce = w.win
    lo_hi = w.extra["window_rows"]

    def one(k):
        lo, hi = lo_hi[k]
        rows = ctx_rows[k]
        dos, _, keep, _ = O.decode_bed(rows, w.n)
        sel = np.nonzero(keep)[0]
        r = O.cis_gene("lmm", w.Y[:, k], w.X, dos[sel], B=0, gene=w.names[k], Q=Q, lam=lam)
        return lo + sel, r

    ctx_rows = {k: ctx.copy_rows(w.rows_dev, lo_hi[k][0], lo_hi[k][1] - lo_hi[k][0] + 1, w.bps) for k in genes}
    refs = _pmap(one, genes)
    Yt = Q.T @ w.Y[:, genes]
    Xt = Q.T @ w.X
    # curvature of f in ln(delta) at the oracle's delta: on flat likelihoods
    # (f'' < 1e-3, e.g. a minimum at delta -> 0) the reference search's last
    # numeric-Newton steps are driven by rounding noise, and f is compared at
    # 1e-8 instead of 1e-9; below f'' = 1e-4 (f moves by less than 1e-4 over
    # a unit of ln delta) the numeric second difference of those steps is
    # below the rounding of f and the end point is set by it: 1e-7
    th = np.log(delta)
    th[genes] = np.log([r["delta"] for _, r in refs])
    _, _, hcurv = ctx.reml_grad_hess(th)
    worst = 0.0
    for k, (snps, r) in zip(genes, refs):
        fg = O.reml_negll(delta[k], Yt[:, genes.index(k)], Xt, lam)
        fo = O.reml_negll(r["delta"], Yt[:, genes.index(k)], Xt, lam)
        tol = 1e-9 if hcurv[k] > 1e-3 else (1e-8 if hcurv[k] > 1e-4 else 1e-7)
        if fg - fo > 1e-9 * abs(fo):
            print(f"gene {k}: delta {delta[k]:.6g} vs oracle {r['delta']:.6g}, f'' {hcurv[k]:.3g}, "
                  f"df rel {(fg - fo) / abs(fo):.3g}")
        assert fg - fo <= tol * abs(fo), (k, fg, fo, hcurv[k])
        hk = hits[hits["gene"] == k]
        tested = np.isfinite(r["p"])
        assert st["n_tested"][k] == r["n_tested"] == tested.sum()
        assert np.array_equal(np.sort(hk["snp"]), snps[tested])
        h = hk[np.argsort(hk["snp"])]
        pr = r["p"][tested]
        rel = rel_err(h["p"], pr)
        worst = max(worst, rel.max())
        assert rel.max() < 1e-3, (k, rel.max())
        assert st["best_snp"][k] == snps[tested][np.argmin(pr)]
    print(f"C2: {hits.size} pairs, 20 oracle genes, max p rel {worst:.2e}")


# ---------------------------------------------------------------- C3 -------
def test_c3_trans_lmm_full(gpu_ctx):
    ctx = gpu_ctx
    w = workloads.c3(ctx)
    ctx.set_samples(w.n)
    ctx.set_policy(True, 0.8, 0.0)
    K, N, used = _grm_dev(ctx, w)
    Q, lam = ctx.eigh(K)
    ctx.set_snp_annot(w.chrom, w.pos)
    delta, _ = ctx.lmm_prep(w.Y, w.X, delta_mode=eqtl_amd.DELTA_NEWTON)
    ctx.set_gene_windows(*w.win)
    hits, st = _scan_dev(ctx, w, eqtl_amd.MODEL_LMM, eqtl_amd.PREC_BF16X3, 1e-5)
    G = w.Y.shape[1]
    # every SNP passes decode (MAF >= 0.05): n_tested = SNPs outside the gene's window
    gch, cs, ce = w.win
    per = w.m // workloads.C3_NCHROM
    pos1 = w.pos[:per]
    inwin = np.searchsorted(pos1, ce, side="right") - np.searchsorted(pos1, cs, side="left")
    assert np.array_equal(st["n_tested"], w.m - inwin)
    # REML: f at the Newton delta no worse than at the reference search's delta
    f_new, _, _ = ctx.reml_grad_hess(np.log(delta))
    d_ref, _ = ctx.lmm_prep(w.Y, w.X, delta_mode=eqtl_amd.DELTA_REF)
    f_ref, _, _ = ctx.reml_grad_hess(np.log(d_ref))
    worse = f_new - f_ref
    print(f"C3 REML: Newton f - reference-search f: max {worse.max():.2e} (f ~ {np.median(np.abs(f_ref)):.1f})")
    assert np.all(worse <= 1e-9 * np.abs(f_ref))
    # sampled pairs against the oracle (the genes' own Newton delta)
    rng = np.random.default_rng(7)
    genes = np.unique(np.concatenate([np.arange(64), rng.choice(G, 64, replace=False)]))
    planted = w.extra["planted"]
    snps = np.unique(np.concatenate([np.array([planted[k] for k in genes if k in planted]),
                                     rng.choice(w.m, 400, replace=False)]))
    rows = np.concatenate([ctx.copy_rows(w.rows_dev, int(j), 1, w.bps) for j in snps])
    # the full scan's hits on the sample
    hs = hits[np.isin(hits["snp"], snps) & np.isin(hits["gene"], genes)]
    # a dense scan of the sample (every pair reported) through the same engine
    Ys, ds = w.Y[:, genes], delta[genes]
    ctx.lmm_prep(Ys, w.X, delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=ds)
    ctx.set_snp_annot(w.chrom[snps], w.pos[snps])
    ctx.set_gene_windows(gch[genes], cs[genes], ce[genes])
    hd, _ = ctx.scan(eqtl_amd.MODEL_LMM, eqtl_amd.PREC_BF16X3, rows, pthr=1.0, block=4096,
                     hit_cap=len(snps) * len(genes) + 1)
    dos, _, keep, _ = O.decode_bed(rows, w.n)
    ref = O.trans_scan("lmm", dos, keep, w.chrom[snps], w.pos[snps], Ys, w.X, Q=Q, lam=lam, gene_chrom=gch[genes],
                       gene_cstart=cs[genes], gene_cend=ce[genes], delta=ds, pthr=1e-5, dense=True, hit_cap=1)
    gd = dense_from_hits(hd, len(snps), len(genes))
    tested = np.isfinite(ref["p"])
    assert np.array_equal(tested, np.isfinite(gd["p"]))
    ok = tested & (ref["p"] >= 1e-300)
    prel = rel_err(gd["p"][ok], ref["p"][ok])
    print(f"C3: {int(st['n_tested'].sum())} tests; sample {ok.sum()} pairs, max p rel {prel.max():.2e}, "
          f"min p {ref['p'][ok].min():.2e}")
    assert prel.max() < 1e-3
    # the full scan reported exactly the sample's pairs with p <= 1e-5 (away from the threshold)
    sig = tested & (ref["p"] <= 1e-5)
    near = tested & (np.abs(ref["p"] / 1e-5 - 1) < 1e-3)
    js = {int(s): i for i, s in enumerate(snps)}
    ks = {int(g): i for i, g in enumerate(genes)}
    got = np.zeros_like(sig)
    got[[js[int(s)] for s in hs["snp"]], [ks[int(g)] for g in hs["gene"]]] = True
    assert np.array_equal(got & ~near, sig & ~near)


# ---------------------------------------------------------------- C4 -------
def _c4_slice(cg, k_lm=40, k_lmm=12):
    """Genes stratified by window size: k_lm over the whole range, k_lmm among
    windows of at most 1500 SNPs (the oracle's LMM permutations re-fit delta
    1000 times per gene)."""
    order = np.argsort(cg, kind="stable")
    lm = order[np.linspace(0, cg.size - 1, k_lm).astype(int)]
    small = order[cg[order] <= 1500]
    lmm = small[np.linspace(0, small.size - 1, k_lmm).astype(int)]
    return np.sort(lm), np.sort(lmm)


def _say(*a):
    print(*a, flush=True)


def test_c4_cis_permutations_full(gpu_ctx):
    import time
    t0 = time.perf_counter()
    ctx = gpu_ctx
    w = workloads.c4(ctx)
    B, seed = 1000, 42
    prec = eqtl_amd.PREC_BF16X3
    ctx.set_samples(w.n)
    ctx.set_policy(True, 0.8, 0.0)
    K, N, used = _grm_dev(ctx, w)
    Q, lam = ctx.eigh(K)
    ctx.set_snp_annot(w.chrom, w.pos)
    ctx.upload_bed(rows_dev=w.rows_dev, m=w.m, bps=w.bps)
    hashes = np.array([O.fnv1a(nm) for nm in w.names], dtype=np.uint32)
    cg, first = w.extra["c_g"], w.extra["first_row"]
    g_lm, g_lmm = _c4_slice(cg)
    _say(f"C4: inputs, GRM and eigh {time.perf_counter() - t0:.1f} s")
    res = {}
    for model in (eqtl_amd.MODEL_LM, eqtl_amd.MODEL_LMM):
        if model == eqtl_amd.MODEL_LM:
            ctx.lm_prep(w.Y, w.X)
        else:
            ctx.lmm_prep(w.Y, w.X, delta_mode=eqtl_amd.DELTA_REF)
        ctx.set_gene_windows(*w.win)
        _, st = _scan_dev(ctx, w, model, prec, 1e-5, eqtl_amd.SCOPE_CIS)
        assert np.array_equal(st["n_tested"], cg)
        pm, pe = ctx.cis_perm(prec, B, seed, hashes, st["best_p"], delta_mode=eqtl_amd.DELTA_REF)
        assert np.isfinite(pe).all()
        res[model] = (pm, pe)
        _say(f"C4: model {model} observed + {B} permutations of {len(w.names)} genes done at "
             f"{time.perf_counter() - t0:.1f} s")
    rows_of = {k: ctx.copy_rows(w.rows_dev, int(first[k]), int(cg[k]), w.bps) for k in set(g_lm) | set(g_lmm)}

    def one(args):
        model, k = args
        dos, _, keep, _ = O.decode_bed(rows_of[k], w.n)
        kw = {"Q": Q, "lam": lam} if model == "lmm" else {}
        r = O.cis_gene(model, w.Y[:, k], w.X, dos[keep], B=B, seed=seed, gene=w.names[k], **kw)
        _say(f"C4 oracle: {model} gene {k} (window {cg[k]} SNPs) at {time.perf_counter() - t0:.1f} s")
        return r

    jobs = [("lm", k) for k in g_lm] + [("lmm", k) for k in g_lmm]
    refs = dict(zip(jobs, _pmap(one, jobs)))
    for model, mc in (("lm", eqtl_amd.MODEL_LM), ("lmm", eqtl_amd.MODEL_LMM)):
        pm, pe = res[mc]
        worst, near = 0.0, 0
        for k in (g_lm if model == "lm" else g_lmm):
            r = refs[(model, k)]
            rel = rel_err(pm[k], r["perm_min_p"])
            worst = max(worst, rel.max())
            assert rel.max() < 1e-3, (model, k, rel.max())
            Tobs = -np.log10(max(r["best_p"], 1e-300))
            near += int(np.any(np.abs(r["T_perm"] - Tobs) <= 1e-4 * max(Tobs, 1e-12)))
            assert pe[k] == r["p_emp"], (model, k, pe[k], r["p_emp"])
        print(f"C4 {model}: {len(g_lm) if model == 'lm' else len(g_lmm)} oracle genes at B={B}, "
              f"max perm_min_p rel {worst:.2e}, p_emp equal; genes with a near tie {near}")


# ---------------------------------------------------------------- C5 -------
def test_c5_grm_and_reml_adjoint_full(gpu_ctx):
    ctx = gpu_ctx
    w = workloads.c5(ctx)
    ctx.set_samples(w.n)
    ctx.set_policy(True, 0.8, 0.0)
    K, N, used = _grm_dev(ctx, w)
    assert used == w.m
    idx = np.sort(np.random.default_rng(5).choice(w.n, 48, replace=False))
    dk, n_ok = _grm_subblock_check(ctx, w, K, N, w.m, idx)
    print(f"C5 GRM: {used} SNPs, 48-sample sub-block max |dK| {dk:.2e}")
    assert dk < 1e-12 and n_ok
    # the 1%-missing variant (missing-count Gram and GCTA divisor)
    wm = workloads.c5(ctx, m=200_000, seed=1006, miss=0.01)
    Km, Nm, used_m = _grm_dev(ctx, wm)
    dkm, nm_ok = _grm_subblock_check(ctx, wm, Km, Nm, wm.m, idx)
    print(f"C5 GRM, 1% missing: {used_m} SNPs, sub-block max |dK| {dkm:.2e}")
    assert dkm < 1e-12 and nm_ok
    del Km, Nm
    ctx.set_samples(w.n)
    Q, lam = ctx.eigh(K)
    delta, _ = ctx.lmm_prep(w.Y, w.X, delta_mode=eqtl_amd.DELTA_NEWTON)
    th0 = np.log(delta)
    eps = np.finfo(float).eps
    for shift in (-1.0, 0.0, 1.0):
        th = th0 + shift
        f, g, h = ctx.reml_grad_hess(th)
        hg, hh = 1e-4, 3e-3
        fp, _, _ = ctx.reml_grad_hess(th + hg)
        fm, _, _ = ctx.reml_grad_hess(th - hg)
        fP, _, _ = ctx.reml_grad_hess(th + hh)
        fM, _, _ = ctx.reml_grad_hess(th - hh)
        # rounding noise of f, measured: f at ln(delta) moved by 1e-13..4e-13
        # (a true change below 1e-12 |f'|) differs from f by its rounding only;
        # on flat likelihoods at the delta range's ends (delta = 1e6) df ln(q)
        # with a cancelling q carries noise far above eps |f|
        noise = np.max([np.abs(ctx.reml_grad_hess(th + k * 1e-13)[0] - f) for k in (1, 2, 3, 4)], axis=0)
        eps_f = np.maximum(64 * eps * np.abs(f), 4 * noise)
        gfd = (fp - fm) / (2 * hg)
        hfd = (fP - 2 * f + fM) / hh ** 2
        eg = np.abs(g - gfd) - (1e-5 * (1 + np.abs(gfd)) + 2 * eps_f / hg)
        eh = np.abs(h - hfd) - (1e-3 * (1 + np.abs(hfd)) + 4 * eps_f / hh ** 2)
        print(f"C5 REML adjoint shift {shift:+.0f}: {f.size} genes, max |g - fd| {np.abs(g - gfd).max():.2e}, "
              f"max |h - fd| {np.abs(h - hfd).max():.2e}")
        assert (eg <= 0).all() and (eh <= 0).all()
    # f against the oracle's restatement on 200 genes
    sub = np.linspace(0, w.Y.shape[1] - 1, 200).astype(int)
    Yt = Q.T @ w.Y[:, sub]
    Xt = Q.T @ w.X
    f0, _, _ = ctx.reml_grad_hess(th0)
    fo = np.array(_pmap(lambda i: O.reml_negll(delta[sub[i]], Yt[:, i], Xt, lam), range(len(sub))))
    assert np.all(np.abs(f0[sub] - fo) <= 1e-9 * (1 + np.abs(fo)))


def test_c1_cli_every_pair(tmp_path):
    """C1 through the eqtl CLI with --pval-trans 1: every tested pair is written
    (4.1M lines; the reference writes each one), in SNP-major order, with p
    within 1e-3 of the oracle and the top file's best SNPs exact."""
    import subprocess
    import pandas as pd
    w, ref = _c1_ref()
    ds = w.extra["dataset"]
    prefix = str(tmp_path / "c1")
    ds.write(prefix)
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "eqtl_amd", "eqtl")
    out = str(tmp_path / "o")
    r = subprocess.run([cli, "-b", prefix, "-e", prefix + ".pheno.tsv", "-c", prefix + ".covar.tsv", "-g",
                        prefix + ".gff", "--model", "lm", "--mode", "trans", "--pval-trans", "1", "-o", out],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    df = pd.read_csv(out + ".lm.trans.pairs.tsv", sep="\t", usecols=["gene", "snp", "p"],
                     dtype={"gene": str, "snp": str, "p": np.float64})
    tested = np.isfinite(ref["p"])
    assert len(df) == tested.sum()
    k = df["gene"].str.slice(1).astype(np.int64).to_numpy()
    j = df["snp"].str.slice(2).astype(np.int64).to_numpy() - 1
    key = j * 1000 + k
    assert np.all(np.diff(key) > 0)  # SNP-major, gene order within a SNP
    pr = ref["p"][j, k]
    ok = pr >= 1e-300
    rel = np.abs(df["p"].to_numpy()[ok] - pr[ok]) / pr[ok]
    print(f"C1 CLI: {len(df)} pairs written, max p rel {rel.max():.2e}")
    assert rel.max() < 1e-3
    top = pd.read_csv(out + ".lm.trans.top.tsv", sep="\t", usecols=["gene", "snp"], dtype=str)
    for g, s in zip(top["gene"], top["snp"]):
        assert int(s[2:]) - 1 == ref["state"]["best_snp"][int(g[1:])]
