"""Trans LMM scan in its interpolated form (covariate projections and gg
interpolated in ln delta over Chebyshev nodes, k_lr_num + k_wald_lr + k_lr_tests)
against the FP64 oracle (test_lmm_gtil, scan_trans_lmm.cpp:11-40) at given
deltas chosen so that every node count of the plan runs: segments of 32, 64,
96 and 128 nodes, several 64-gene super-groups with padding slots, and a
spread of ln delta too wide to interpolate (the eight-wave kernel then runs).
"""
import numpy as np
import pytest

import eqtl_amd
import oracle as O
from eqtl_amd import synth
from _cases import dense_from_hits, rel_err, windows

pytestmark = pytest.mark.gpu


def _case(seed, n, m, G):
    ds = synth.make_trans(seed, n=n, m=m, G=G, n_chrom=2, ncov=3, miss=0.01, spacing=1000,
                          h2=("uniform", 0.0, 0.7), n_causal=200, n_planted=6, beta_planted=(0.3, 1.5))
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    K, N, used = O.grm(dos, missm, keep)
    Q, lam = O.eigh(K)
    return ds, rows, dos, keep, Q, lam


def _deltas(rng, G, ranges):
    """ln delta evenly over each (count, lo, hi) range, genes in random order."""
    u = np.concatenate([np.linspace(lo, hi, c) for c, lo, hi in ranges])
    assert u.size == G
    return np.exp(rng.permutation(u))


def _run(ctx, ds, rows, dos, keep, Q, lam, delta):
    G = ds.Y.shape[1]
    win = windows(ds, 200_000)
    ctx.set_samples(ds.n)
    ctx.set_policy(True, 0.8, 0.0)
    ctx.set_snp_annot(ds.chrom, ds.pos)
    ctx.set_basis(Q, lam)
    ctx.lmm_prep(ds.Y, ds.X(), delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=delta)
    ctx.set_gene_windows(*win)
    hits, st = ctx.scan(eqtl_amd.MODEL_LMM, eqtl_amd.PREC_BF16X3, rows, pthr=1.0, block=1024, hit_cap=1 << 20)
    plan = ctx.lr_plan()
    ref = O.trans_scan("lmm", dos, keep, ds.chrom, ds.pos, ds.Y, ds.X(), Q=Q, lam=lam, gene_chrom=win[0],
                       gene_cstart=win[1], gene_cend=win[2], delta=delta, pthr=1.0, dense=True,
                       hit_cap=ds.m * G + 1)
    return hits, st, plan, ref


def _check(ds, hits, st, ref, label):
    m, G = ds.m, ds.Y.shape[1]
    gd = dense_from_hits(hits, m, G)
    tested = np.isfinite(ref["p"])
    assert np.array_equal(tested, np.isfinite(gd["p"]))
    t_r, t_g = ref["stat"][tested], gd["stat"][tested]
    dt = np.abs(t_g - t_r) / (1 + np.abs(t_r))
    ok = ref["p"][tested] >= 1e-300
    pr = rel_err(gd["p"][tested][ok], ref["p"][tested][ok])
    se = rel_err(gd["se"][tested], ref["se"][tested])
    print(f"{label}: max |dt|/(1+|t|) {dt.max():.3e}, max p rel {pr.max():.3e}, se rel {se.max():.3e}")
    assert dt.max() < 5e-5
    assert pr.max() < 1e-3
    assert se.max() < 5e-4
    rs = ref["state"]
    assert np.array_equal(st["n_tested"], rs["n_tested"])
    assert np.array_equal(st["n_sig"], rs["n_sig"])
    assert np.array_equal(st["best_snp"], rs["best_snp"])
    assert rel_err(st["acat_p"], rs["acat_p"]).max() < 2e-3


def test_lr_segments_32_64_and_padding(gpu_ctx):
    """150 genes: three super-groups (the last with 42 padding slots; an odd
    count, so the num kernel's second pair has an empty half) over segments
    of 32, 64 and 32 nodes."""
    ds, rows, dos, keep, Q, lam = _case(61, n=160, m=1500, G=150)
    rng = np.random.default_rng(5)
    delta = _deltas(rng, 150, [(64, -9.0, -8.6), (64, -8.5, 3.0), (22, 3.5, 4.0)])
    hits, st, plan, ref = _run(gpu_ctx, ds, rows, dos, keep, Q, lam, delta)
    low_rank, n_sg, n_seg, mean_nodes = plan
    assert low_rank and n_sg == 3 and n_seg == 3, plan
    assert abs(mean_nodes - (64 * 32 + 64 * 64 + 22 * 32) / 150) < 1e-9, plan
    _check(ds, hits, st, ref, "segments 32/64/32")


@pytest.mark.parametrize("span, nodes", [(22.0, 96), (26.0, 128)])
def test_lr_wide_segment(gpu_ctx, span, nodes):
    """One super-group whose ln delta spans 22 (96 nodes) or 26 (128 nodes)."""
    ds, rows, dos, keep, Q, lam = _case(62, n=160, m=1200, G=40)
    rng = np.random.default_rng(6)
    delta = _deltas(rng, 40, [(40, -0.5 * span - 1.0, 0.5 * span - 1.0)])
    hits, st, plan, ref = _run(gpu_ctx, ds, rows, dos, keep, Q, lam, delta)
    assert plan[0] and plan[2] == 1 and plan[3] == nodes, plan
    _check(ds, hits, st, ref, f"one segment of {nodes} nodes")


def test_lr_too_wide_runs_eight_wave_kernel(gpu_ctx):
    """ln delta spread over 35 (beyond 128 nodes): the eight-wave kernel runs."""
    ds, rows, dos, keep, Q, lam = _case(63, n=160, m=1200, G=40)
    rng = np.random.default_rng(7)
    delta = _deltas(rng, 40, [(40, -17.0, 18.0)])
    hits, st, plan, ref = _run(gpu_ctx, ds, rows, dos, keep, Q, lam, delta)
    assert not plan[0], plan
    _check(ds, hits, st, ref, "eight-wave fallback")
