"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle.

Tolerances (stated per quantity):
  decode / keep / af / n_tested / best SNP / tested sets     exact
  Wald at a shared delta, FP32 form     |dt| <= 5e-6 (1+|t|), beta/se rel 1e-4
  Wald at a shared delta, BF16X3 form   |dt| <= 5e-5 (1+|t|), beta/se rel 5e-4
      (beta where |t| > 0.1); p relative 1e-3 for p >= 1e-300 (both forms)
  ACAT machinery                        acat_p equals the reference formula
      applied to the engine's own p values (rel 1e-12, or one 2^-54 step)
  ACAT vs oracle                        relative 2e-3 (terms -1/(pi(1-p)) for p
      near 1 carry the relative error of |t|: the FP64 numerator refinement
      below |t| = tau leaves the contraction's relative error of S)
  REML delta (reference search)         f(delta_gpu) - f(delta_oracle) <= 1e-9|f|
      and, where f'' in ln(delta) exceeds 1e-3, delta relative 1e-4 (the search
      is only defined to its own floating-point noise; on flat likelihoods that
      noise decides the golden-section path)
"""

import numpy as np
import pytest

import oracle as O
import eqtl_amd
from eqtl_amd import synth

from _cases import dense_from_hits, lmm_case, rel_err, trans_lm_case

pytestmark = pytest.mark.gpu

TOL = {eqtl_amd.PREC_FP32: (5e-6, 1e-4), eqtl_amd.PREC_BF16X3: (5e-5, 5e-4)}


def _acat_consistent(st, hits, G):
    """The engine's acat_p equals AcatAcc applied to its own p values, in SNP order."""
    order = np.lexsort((hits["snp"], hits["gene"]))
    h = hits[order]
    for k in range(G):
        pk = h["p"][h["gene"] == k]
        ref = O.acat(pk)
        got = st["acat_p"][k]
        assert abs(got - ref) <= 1.2e-16 + 1e-12 * ref, (k, got, ref)


def _p_rel(pg, pr):
    ok = np.isfinite(pr) & (pr >= 1e-300)
    return rel_err(pg[ok], pr[ok])


def test_decode_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 5, 64, 67, 130):
        geno, _ = synth.draw_genotypes(rng, 300, n, 0.0, 0.5, 0.05)
        geno[0, :] = 0           # monomorphic
        geno[1, :] = -1          # all missing
        geno[2, : max(1, n // 2)] = -1  # heavy missing (above/below max_miss)
        rows = synth.pack_bed(geno)
        for impute, max_miss, maf in ((True, 0.8, 0.0), (False, 0.8, 0.0), (True, 0.1, 0.05)):
            gpu_ctx.set_samples(n)
            gpu_ctx.set_policy(impute, max_miss, maf)
            dos, keep, af = gpu_ctx.decode_bed(rows)
            rd, rm, rk, raf = O.decode_bed(rows, n, impute=impute, max_miss=max_miss, maf_min=maf)
            # the engine additionally applies subset_maf_or_nan (scan_common.cpp:133-151)
            rk2 = rk.copy()
            for j in np.nonzero(rk)[0]:
                r, _ = O.subset_maf_or_nan(rd[j])
                rk2[j] = np.isfinite(r)
            assert np.array_equal(keep, rk2), (n, impute, max_miss, maf)
            both = rk2
            np.testing.assert_allclose(af[both], raf[both], rtol=1e-14)
            d1 = np.where(np.isnan(dos), -7, dos)
            d2 = np.where(np.isnan(rd), -7, rd)
            np.testing.assert_allclose(d1[both], d2[both], rtol=1e-14)
    gpu_ctx.set_policy(True, 0.8, 0.0)


def test_sample_subset_order(gpu_ctx):
    """set_sample_order with a permuted subset of fam columns (plink_bed.cpp:245-255)."""
    rng = np.random.default_rng(5)
    geno, _ = synth.draw_genotypes(rng, 200, 37, 0.05, 0.5, 0.02)
    rows = synth.pack_bed(geno)
    cols = rng.permutation(37)[:29].astype(np.int32)
    gpu_ctx.set_samples(37, cols)
    dos, keep, af = gpu_ctx.decode_bed(rows)
    rd, _, rk, raf = O.decode_bed(rows, 37, sample_col=cols)
    assert np.array_equal(keep[rk], rk[rk])
    np.testing.assert_allclose(dos[keep], rd[keep], rtol=1e-14)


@pytest.mark.parametrize("prec", [eqtl_amd.PREC_FP32, eqtl_amd.PREC_BF16X3])
def test_trans_lm_parity(gpu_ctx, prec):
    ds, rows, (gc, gs, ge), ref = trans_lm_case()
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_policy(True, 0.8, 0.0)
    gpu_ctx.set_snp_annot(ds.chrom, ds.pos)
    gpu_ctx.lm_prep(ds.Y, ds.X())
    gpu_ctx.set_gene_windows(gc, gs, ge)
    hits, st = gpu_ctx.scan(eqtl_amd.MODEL_LM, prec, rows, pthr=1.0, block=1024, hit_cap=1 << 20)
    m, G = ds.m, ds.Y.shape[1]
    gd = dense_from_hits(hits, m, G)
    tested_ref = np.isfinite(ref["p"])
    tested_gpu = np.isfinite(gd["p"])
    assert np.array_equal(tested_ref, tested_gpu)
    t_r, t_g = ref["stat"][tested_ref], gd["stat"][tested_ref]
    dt = np.abs(t_g - t_r) / (1 + np.abs(t_r))
    print(f"LM prec={prec}: max |dt|/(1+|t|) = {dt.max():.3e}, max p rel = {_p_rel(gd['p'][tested_ref], ref['p'][tested_ref]).max():.3e}")
    tol_t, tol_rel = TOL[prec]
    assert dt.max() < tol_t
    assert _p_rel(gd["p"][tested_ref], ref["p"][tested_ref]).max() < 1e-3
    big = np.abs(t_r) > 0.1
    assert rel_err(gd["beta"][tested_ref][big], ref["beta"][tested_ref][big]).max() < tol_rel
    assert rel_err(gd["se"][tested_ref], ref["se"][tested_ref]).max() < tol_rel
    rs = ref["state"]
    assert np.array_equal(st["n_tested"], rs["n_tested"])
    assert np.array_equal(st["n_sig"], rs["n_sig"])
    assert np.array_equal(st["best_snp"], rs["best_snp"])
    _acat_consistent(st, hits, G)
    print(f"ACAT vs oracle: max rel {rel_err(st['acat_p'], rs['acat_p']).max():.3e}")
    assert rel_err(st["acat_p"], rs["acat_p"]).max() < 2e-3


def _lmm_run(ctx, ds, rows, Q, lam, win, prec, delta_mode, delta_in=None, pthr=1.0):
    ctx.set_samples(ds.n)
    ctx.set_policy(True, 0.8, 0.0)
    ctx.set_snp_annot(ds.chrom, ds.pos)
    ctx.set_basis(Q, lam)
    delta, rss = ctx.lmm_prep(ds.Y, ds.X(), delta_mode=delta_mode, delta_in=delta_in)
    ctx.set_gene_windows(*win)
    hits, st = ctx.scan(eqtl_amd.MODEL_LMM, prec, rows, pthr=pthr, block=1024, hit_cap=1 << 20)
    return delta, rss, hits, st


def test_reml_delta_reference_search_p16(gpu_ctx):
    """15 covariates + intercept: X~^T D X~ has 136 packed entries, 128 from the
    GEMM and the last 8 summed in the weights kernel."""
    rng = np.random.default_rng(33)
    n, G = 300, 24
    K = np.corrcoef(rng.standard_normal((n, 1500)))
    lam, Q = np.linalg.eigh(K)
    lam = np.maximum(lam, 0)
    X = np.hstack([np.ones((n, 1)), rng.standard_normal((n, 15))])
    u = Q @ (np.sqrt(lam)[:, None] * rng.standard_normal((n, G)))
    Y = rng.standard_normal((n, G)) + u * rng.uniform(0, 1.5, G)[None, :]
    gpu_ctx.set_samples(n)
    gpu_ctx.set_basis(Q, lam)
    delta, _ = gpu_ctx.lmm_prep(Y, X, delta_mode=eqtl_amd.DELTA_REF)
    Yt, Xt = O.rotate(Q, Y), O.rotate(Q, X)
    dref = O.optimize_delta(Yt, Xt, lam)
    _, _, h = gpu_ctx.reml_grad_hess(np.log(dref))
    curved = h > 1e-3
    r = rel_err(delta, dref)
    print("p=16 REML ref-search delta max rel err (curved genes)", r[curved].max(), "flat genes", int((~curved).sum()))
    assert curved.sum() >= G // 2
    assert r[curved].max() < 1e-4
    for k in range(G):
        fg = O.reml_negll(delta[k], Yt[:, k], Xt, lam)
        fo = O.reml_negll(dref[k], Yt[:, k], Xt, lam)
        assert fg - fo <= 1e-9 * abs(fo), (k, fg, fo)


def test_reml_delta_reference_search(gpu_ctx):
    ds, rows, (Q, lam), win, ref = lmm_case()
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_basis(Q, lam)
    delta, rss = gpu_ctx.lmm_prep(ds.Y, ds.X(), delta_mode=eqtl_amd.DELTA_REF)
    r = rel_err(delta, ref["state"]["delta"])
    Yt = O.rotate(Q, ds.Y)
    Xt = O.rotate(Q, ds.X())
    # curvature of f in ln(delta) at the oracle's delta; below 1e-3 the likelihood
    # is flat to ~1e-9 |f| across the bracket and the search's path is set by
    # rounding noise, so only the f criterion applies there
    _, _, h = gpu_ctx.reml_grad_hess(np.log(ref["state"]["delta"]))
    curved = h > 1e-3
    print("REML ref-search delta max rel err (curved genes)", r[curved].max(), "flat genes", int((~curved).sum()))
    assert r[curved].max() < 1e-4
    for k in range(ds.Y.shape[1]):
        fg = O.reml_negll(delta[k], Yt[:, k], Xt, lam)
        fo = O.reml_negll(ref["state"]["delta"][k], Yt[:, k], Xt, lam)
        assert fg - fo <= 1e-9 * abs(fo), (k, fg, fo)


@pytest.mark.parametrize("prec", [eqtl_amd.PREC_FP32, eqtl_amd.PREC_BF16X3])
def test_trans_lmm_parity(gpu_ctx, prec):
    ds, rows, (Q, lam), win, ref = lmm_case()
    delta, rss, hits, st = _lmm_run(gpu_ctx, ds, rows, Q, lam, win, prec, eqtl_amd.DELTA_GIVEN,
                                    delta_in=ref["state"]["delta"])
    m, G = ds.m, ds.Y.shape[1]
    gd = dense_from_hits(hits, m, G)
    tested_ref = np.isfinite(ref["p"])
    assert np.array_equal(tested_ref, np.isfinite(gd["p"]))
    t_r, t_g = ref["stat"][tested_ref], gd["stat"][tested_ref]
    dt = np.abs(t_g - t_r) / (1 + np.abs(t_r))
    pr = _p_rel(gd["p"][tested_ref], ref["p"][tested_ref])
    print(f"LMM prec={prec}: max |dt|/(1+|t|) = {dt.max():.3e}, max p rel = {pr.max():.3e}, "
          f"min p = {np.nanmin(ref['p']):.3e}")
    tol_t, tol_rel = TOL[prec]
    assert dt.max() < tol_t
    assert pr.max() < 1e-3
    big = np.abs(t_r) > 0.1
    assert rel_err(gd["beta"][tested_ref][big], ref["beta"][tested_ref][big]).max() < tol_rel
    assert rel_err(gd["se"][tested_ref], ref["se"][tested_ref]).max() < tol_rel
    rs = ref["state"]
    assert np.array_equal(st["n_tested"], rs["n_tested"])
    assert np.array_equal(st["n_sig"], rs["n_sig"])
    assert np.array_equal(st["best_snp"], rs["best_snp"])
    _acat_consistent(st, hits, G)
    print(f"ACAT vs oracle: max rel {rel_err(st['acat_p'], rs['acat_p']).max():.3e}")
    assert rel_err(st["acat_p"], rs["acat_p"]).max() < 2e-3


def test_reml_grad_hess_fd(gpu_ctx):
    """Adjoint gradient/Hessian in log(delta) vs central differences of the oracle's f.

    The oracle's f is a sum of n logs plus a solve; its rounding noise is taken as
    eps_f = 64 eps |f|, which puts the noise of a central second difference at
    4 eps_f / h^2. Gradient step 1e-4; Hessian step 3e-3 (fourth-derivative
    truncation ~1e-4 for the df ~ 150 here, noise ~1e-6)."""
    ds, rows, (Q, lam), win, ref = lmm_case()
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_basis(Q, lam)
    gpu_ctx.lmm_prep(ds.Y, ds.X(), delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=ref["state"]["delta"])
    Yt = O.rotate(Q, ds.Y)
    Xt = O.rotate(Q, ds.X())
    for shift in (-1.0, 0.0, 1.0):
        th = np.log(ref["state"]["delta"]) + shift
        f, g, h = gpu_ctx.reml_grad_hess(th)
        for k in range(ds.Y.shape[1]):
            F = lambda t: O.reml_negll(np.exp(t), Yt[:, k], Xt, lam)
            fo = F(th[k])
            eps_f = 64 * np.finfo(float).eps * abs(fo)
            hg, hh = 1e-4, 3e-3
            gfd = (F(th[k] + hg) - F(th[k] - hg)) / (2 * hg)
            hfd = (F(th[k] + hh) - 2 * fo + F(th[k] - hh)) / hh ** 2
            assert abs(f[k] - fo) <= 1e-9 * (1 + abs(fo))
            assert abs(g[k] - gfd) <= 1e-5 * (1 + abs(gfd)) + 2 * eps_f / hg, (k, shift, g[k], gfd)
            assert abs(h[k] - hfd) <= 1e-3 * (1 + abs(hfd)) + 4 * eps_f / hh ** 2, (k, shift, h[k], hfd)


def test_newton_delta_is_argmin(gpu_ctx):
    """Newton delta has zero gradient (or sits on the reference bracket) and f <= f(reference delta)."""
    ds, rows, (Q, lam), win, ref = lmm_case()
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_basis(Q, lam)
    dn, _ = gpu_ctx.lmm_prep(ds.Y, ds.X(), delta_mode=eqtl_amd.DELTA_NEWTON)
    Yt = O.rotate(Q, ds.Y)
    Xt = O.rotate(Q, ds.X())
    dr = ref["state"]["delta"]
    for k in range(ds.Y.shape[1]):
        fn = O.reml_negll(dn[k], Yt[:, k], Xt, lam)
        fr = O.reml_negll(dr[k], Yt[:, k], Xt, lam)
        assert fn <= fr + 1e-9 * abs(fr), (k, dn[k], dr[k], fn, fr)
    print("Newton vs reference delta rel diff (max over interior genes):",
          np.max(rel_err(dn, dr)[(dr > 1e-4) & (dr < 1e5)]) if np.any((dr > 1e-4) & (dr < 1e5)) else None)


def test_grm_matches_closed_form(gpu_ctx):
    rng = np.random.default_rng(9)
    for n, m, miss in ((70, 900, 0.0), (130, 1500, 0.02)):
        geno, _ = synth.draw_genotypes(rng, m, n, 0.01, 0.5, miss)
        rows = synth.pack_bed(geno)
        gpu_ctx.set_samples(n)
        gpu_ctx.set_policy(True, 0.8, 0.0)
        K, N, used = gpu_ctx.grm(rows, block=512)
        dos, missm, keep, af = O.decode_bed(rows, n)
        Kr, Nr, ur = O.grm(dos, missm, keep)
        assert used == ur
        iu = np.tril_indices(n)
        Np = Nr
        assert np.array_equal(N, Np)
        err = np.abs(K - Kr).max()
        print(f"GRM n={n} m={m} miss={miss}: max abs err {err:.3e}")
        assert err < 1e-12


def test_grm_large_n_tiles(gpu_ctx):
    """n past the switch to 128 x 128 LDS-staged tiles (2 x 128 + a ragged
    third tile row), with and without missing calls, blocks of uneven length."""
    rng = np.random.default_rng(19)
    for n, m, miss in ((1601, 300, 0.02), (1700, 260, 0.0)):
        geno, _ = synth.draw_genotypes(rng, m, n, 0.01, 0.5, miss)
        rows = synth.pack_bed(geno)
        gpu_ctx.set_samples(n)
        gpu_ctx.set_policy(True, 0.8, 0.0)
        K, N, used = gpu_ctx.grm(rows, block=200)
        dos, missm, keep, af = O.decode_bed(rows, n)
        Kr, Nr, ur = O.grm(dos, missm, keep)
        assert used == ur
        assert np.array_equal(N, Nr)
        err = np.abs(K - Kr).max()
        print(f"GRM n={n} m={m} miss={miss}: max abs err {err:.3e}")
        assert err < 1e-12


def test_eigh_gpu(gpu_ctx):
    ds, rows, (Q, lam), win, ref = lmm_case()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    K, N, used = O.grm(dos, missm, keep)
    gpu_ctx.set_samples(ds.n)
    Qg, lg = gpu_ctx.eigh(K)
    w = np.maximum(np.linalg.eigvalsh(K), 0)
    assert np.abs(lg - w).max() < 1e-10
    assert np.abs(Qg.T @ Qg - np.eye(ds.n)).max() < 1e-10
    Kc = Qg @ np.diag(lg) @ Qg.T
    neg = np.linalg.eigvalsh(K).min()
    assert np.abs(Kc - K).max() < 1e-10 + abs(min(neg, 0)) * 2


def test_device_selftest(gpu_ctx):
    """Matrix-core lane layouts and direct global→LDS loads behave as the kernels assume."""
    assert gpu_ctx.selftest() == 0


@pytest.mark.parametrize("prec", [eqtl_amd.PREC_FP32, eqtl_amd.PREC_BF16X3])
def test_trans_lmm_odd_block_sizes(gpu_ctx, prec):
    """Blocks whose SNP count is not a multiple of the 128-row tile, and a threshold hit set."""
    ds, rows, (Q, lam), win, ref = lmm_case()
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_policy(True, 0.8, 0.0)
    gpu_ctx.set_snp_annot(ds.chrom, ds.pos)
    gpu_ctx.set_basis(Q, lam)
    gpu_ctx.lmm_prep(ds.Y, ds.X(), delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=ref["state"]["delta"])
    gpu_ctx.set_gene_windows(*win)
    hits, st = gpu_ctx.scan(eqtl_amd.MODEL_LMM, prec, rows, pthr=1e-3, block=333, hit_cap=1 << 16)
    assert np.array_equal(st["n_tested"], ref["state"]["n_tested"])
    assert np.array_equal(st["best_snp"], ref["state"]["best_snp"])
    got = set(zip(hits["snp"].tolist(), hits["gene"].tolist()))
    p = ref["p"]
    exp = {(j, k) for j, k in zip(*np.nonzero(np.isfinite(p) & (p <= 1e-3)))}
    near = {(j, k) for j, k in zip(*np.nonzero(np.isfinite(p) & (np.abs(np.log(p / 1e-3)) < 1e-2)))}
    assert (got ^ exp) <= near
    assert np.array_equal(st["n_sig"], np.array([sum(1 for (j, k) in got if k == g) for g in range(p.shape[1])]))
