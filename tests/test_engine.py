"""The device-resident pieces of the cis / permutation engine, through the C ABI.

* The reference REML search (optimize_delta, model_lmm.cpp:39-86) runs as one
  fixed launch sequence on the device: eagerly on the first use of a
  configuration, then captured in a hipGraph and replayed. Several calls with
  the same gene count, alternating two phenotype sets, with the program's own
  eager work and synchronisations between them: every replay equals the eager
  result bit for bit, and both sets match the oracle's search (delta within
  1e-4 on curved genes, f no worse than the oracle's anywhere).
* Cis scans run persistent workgroups over a device work queue whose items
  the autotuner sizes per window-size class; after a scan every class present
  has been timed, the chosen sizes are among the candidates, and a second scan
  (tuned sizes) reproduces the first one's gene states exactly and the
  oracle's per-test results.
"""
import numpy as np
import pytest

import oracle as O
import eqtl_amd

from _cases import lmm_case, rel_err

pytestmark = pytest.mark.gpu


def _oracle_delta(Y, Xt_rot, Q, lam):
    return O.optimize_delta(O.rotate(Q, Y), Xt_rot, lam)


def _check_delta(gpu_ctx, delta, oracle_delta, Y, Q, lam, X):
    Yt, Xt = O.rotate(Q, Y), O.rotate(Q, X)
    _, _, h = gpu_ctx.reml_grad_hess(np.log(oracle_delta))
    curved = h > 1e-3
    assert curved.sum() >= 5
    assert rel_err(delta, oracle_delta)[curved].max() < 1e-4
    for k in range(Y.shape[1]):
        fg = O.reml_negll(delta[k], Yt[:, k], Xt, lam)
        fo = O.reml_negll(oracle_delta[k], Yt[:, k], Xt, lam)
        assert fg - fo <= 1e-9 * abs(fo), (k, fg, fo)


def test_reml_ref_graph_replays_match_eager(gpu_ctx):
    ds, rows, (Q, lam), win, ref = lmm_case()
    X = ds.X()
    rng = np.random.default_rng(5)
    Y1 = ds.Y
    Y2 = ds.Y[rng.permutation(ds.n)] * 1.3 + 0.2
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_basis(Q, lam)
    seen = {0: [], 1: []}
    for call, which in enumerate([0, 1, 0, 1, 0, 1]):
        Y = Y1 if which == 0 else Y2
        delta, _ = gpu_ctx.lmm_prep(Y, X, delta_mode=eqtl_amd.DELTA_REF)
        print(f"call {call} (set {which}): delta[:3] = {delta[:3]}", flush=True)
        seen[which].append(delta.copy())
        # eager work and synchronisation between the replays
        gpu_ctx.reml_grad_hess(np.log(delta))
        gpu_ctx.synchronize()
    for which in (0, 1):
        first = seen[which][0]
        for d in seen[which][1:]:
            assert np.array_equal(d.view(np.uint64), first.view(np.uint64))
    assert not np.array_equal(seen[0][0], seen[1][0])
    Xt = O.rotate(Q, X)
    gpu_ctx.lmm_prep(Y1, X, delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=seen[0][0])
    _check_delta(gpu_ctx, seen[0][0], ref["state"]["delta"], Y1, Q, lam, X)
    gpu_ctx.lmm_prep(Y2, X, delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=seen[1][0])
    _check_delta(gpu_ctx, seen[1][0], _oracle_delta(Y2, Xt, Q, lam), Y2, Q, lam, X)


def _cis_states(gpu_ctx, ds, rows, win, prec):
    gpu_ctx.set_gene_windows(*win)
    return gpu_ctx.scan(eqtl_amd.MODEL_LMM, prec, rows, pthr=1.0, block=ds.m, hit_cap=1 << 20,
                        scope=eqtl_amd.SCOPE_CIS)


@pytest.mark.parametrize("prec", [eqtl_amd.PREC_FP32, eqtl_amd.PREC_BF16X3])
def test_cis_queue_and_autotuner(gpu_ctx, prec):
    ds, rows, (Q, lam), win, _ = lmm_case()
    gc, gs, ge = win
    dos, _, keep, _ = O.decode_bed(rows, ds.n)
    G = ds.Y.shape[1]
    refs, deltas = [], np.empty(G)
    for k in range(G):
        idx = np.nonzero((ds.chrom == gc[k]) & (ds.pos >= gs[k]) & (ds.pos <= ge[k]) & keep)[0]
        r = O.cis_gene("lmm", ds.Y[:, k], ds.X(), dos[idx], Q=Q, lam=lam)
        refs.append((idx, r))
        deltas[k] = r["delta"]
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_policy(True, 0.8, 0.0)
    gpu_ctx.set_snp_annot(ds.chrom, ds.pos)
    gpu_ctx.set_basis(Q, lam)
    gpu_ctx.lmm_prep(ds.Y, ds.X(), delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=deltas)
    h1, s1 = _cis_states(gpu_ctx, ds, rows, win, prec)
    tiles, timed = gpu_ctx.cis_tuning()
    print("cis autotuner: tiles per item", tiles, "timed", timed, flush=True)
    cand = [(1, 2, 4), (2, 4, 8), (4, 8, 16)]
    assert timed.sum() >= 1
    for q in range(3):
        assert tiles[q] in cand[q] or (not timed[q] and tiles[q] == 4)
    h2, s2 = _cis_states(gpu_ctx, ds, rows, win, prec)
    for f in ("n_tested", "n_sig", "best_snp"):
        assert np.array_equal(s1[f], s2[f])
    assert np.array_equal(s1["best_p"].view(np.uint64), s2["best_p"].view(np.uint64))
    assert np.allclose(s1["acat_p"], s2["acat_p"], rtol=1e-12, atol=0)
    assert len(h1) == len(h2)
    for k, (idx, ref) in enumerate(refs):
        assert s2["n_tested"][k] == ref["n_tested"]
        if ref["n_tested"] == 0:
            continue
        assert s2["best_snp"][k] == idx[np.nanargmin(ref["p"])]
        assert abs(s2["best_p"][k] - ref["best_p"]) <= 1e-3 * ref["best_p"]
