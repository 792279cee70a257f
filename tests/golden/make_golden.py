"""Writes the committed golden fixtures under tests/golden/.

  t_tail.json       two-sided Student-t p at 50 digits (mpmath.betainc,
                    regularized), p from 0.9 down to ~1e-300 — the closed form
                    the reference's p_from_t (util.cpp:259-264) evaluates.
  perm_idx.json     std::mt19937 + std::shuffle index streams for
                    (seed, b, gene, n) as seeded at scan_cis.hpp:106-111,
                    produced by libstdc++ (g++ 11.4) through oracle/liboracle.so.
  ref_cases.npz     outputs of the reference's own compiled TUs (oracle/_ref:
                    model_lm.cpp, model_lmm.cpp, util.cpp) on small seeded
                    inputs: prep_lmm delta / test_lmm and prep_lm / test_lm.

Run from the repository root: python tests/golden/make_golden.py
"""
import ctypes
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402
from eqtl_amd import synth  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def t_tail():
    import mpmath as mp

    mp.mp.dps = 50
    rows = []
    for df in (3, 17, 188, 1978):
        for t in (0.0, 0.3, 1.0, 2.5, 5.0, 10.0, 20.0, 30.0, 40.0, 44.5, 60.0):
            x = mp.mpf(df) / (df + mp.mpf(t) ** 2)
            p = mp.betainc(mp.mpf(df) / 2, mp.mpf(1) / 2, 0, x, regularized=True)
            if p < mp.mpf("1e-300") and t > 0:
                continue
            rows.append({"df": df, "t": t, "p": mp.nstr(p, 30)})
    with open(os.path.join(OUT, "t_tail.json"), "w") as f:
        json.dump(rows, f, indent=0)


def perm_idx():
    cases = []
    for seed, b, gene, n in ((42, 0, "g0", 10), (42, 1, "g0", 10), (42, 999, "ENSG000001", 50), (-1, 3, "g17", 200),
                             (7, 12, "gene_x", 1000)):
        cases.append({"seed": seed, "b": b, "gene": gene, "n": n,
                      "fnv1a": O.fnv1a(gene), "idx": O.perm_index(seed, b, gene, n).tolist()})
    with open(os.path.join(OUT, "perm_idx.json"), "w") as f:
        json.dump(cases, f)


def ref_cases():
    R = O.ref_lib()
    assert R is not None, "build oracle/_ref first (make -C oracle ref)"
    ds = synth.make_trans(101, n=90, m=400, G=6, n_chrom=2, ncov=3, h2=("uniform", 0.05, 0.7), n_causal=150,
                          n_planted=3, beta_planted=(0.4, 1.2))
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    K, N, used = O.grm(dos, missm, keep)
    n = ds.n
    Qr = np.empty((n, n), order="F")
    lr = np.empty(n)
    Kf = np.asfortranarray(K)
    assert R.ref_make_lmm_basis(Kf.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(n),
                                Qr.ctypes.data_as(ctypes.c_void_p), lr.ctypes.data_as(ctypes.c_void_p)) == 0
    X = np.asfortranarray(ds.X())
    sel = np.nonzero(keep)[0][:40]
    Gs = np.ascontiguousarray(dos[sel])
    lmm_out = np.empty((ds.Y.shape[1], len(sel), 5))
    lmm_prep = np.empty((ds.Y.shape[1], 4))
    lm_out = np.empty((ds.Y.shape[1], len(sel), 5))
    for k in range(ds.Y.shape[1]):
        y = np.ascontiguousarray(ds.Y[:, k])
        o = np.empty((len(sel), 5))
        po = np.empty(4)
        assert R.ref_lmm(y.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(n),
                         ctypes.c_int(X.shape[1]), Qr.ctypes.data_as(ctypes.c_void_p),
                         lr.ctypes.data_as(ctypes.c_void_p), Gs.ctypes.data_as(ctypes.c_void_p),
                         ctypes.c_int(len(sel)), o.ctypes.data_as(ctypes.c_void_p),
                         po.ctypes.data_as(ctypes.c_void_p)) == 0
        lmm_out[k] = o
        lmm_prep[k] = po
        o2 = np.empty((len(sel), 5))
        assert R.ref_lm(y.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(n),
                        ctypes.c_int(X.shape[1]), Gs.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(sel)),
                        o2.ctypes.data_as(ctypes.c_void_p)) == 0
        lm_out[k] = o2
    np.savez_compressed(os.path.join(OUT, "ref_cases.npz"), rows=rows, Y=ds.Y, X=np.ascontiguousarray(ds.X()),
                        sel=sel, Q=np.ascontiguousarray(Qr), lam=lr, lmm_out=lmm_out, lmm_prep=lmm_prep,
                        lm_out=lm_out, n=np.int64(n))


if __name__ == "__main__":
    t_tail()
    perm_idx()
    ref_cases()
    print("golden fixtures written to", OUT)
