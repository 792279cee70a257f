"""Shared small cases for the parity tests (seeded; sizes the FP64 oracle finishes in seconds)."""
from __future__ import annotations

import functools

import numpy as np

import oracle as O
from eqtl_amd import synth


def windows(ds, w=1_000_000):
    """cis_window_bounds (scan_common.hpp:58-63): [max(1, start-w), end+w]."""
    cs = np.maximum(1, ds.gene_start - w)
    ce = ds.gene_end + w
    return ds.gene_chrom.astype(np.int32), cs.astype(np.int64), ce.astype(np.int64)


@functools.lru_cache(maxsize=None)
def trans_lm_case(seed=11, n=200, m=3000, G=48, ncov=5, miss=0.01):
    ds = synth.make_trans(seed, n=n, m=m, G=G, n_chrom=2, ncov=ncov, miss=miss, spacing=1000,
                          n_planted=6, beta_planted=(0.3, 0.8))
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    gc, gs, ge = windows(ds, 200_000)
    ref = O.trans_scan("lm", dos, keep, ds.chrom, ds.pos, ds.Y, ds.X(), gene_chrom=gc, gene_cstart=gs,
                       gene_cend=ge, pthr=1.0, dense=True, hit_cap=m * G + 1)
    return ds, rows, (gc, gs, ge), ref


@functools.lru_cache(maxsize=None)
def lmm_case(seed=21, n=160, m=2500, G=40, ncov=3, miss=0.01):
    ds = synth.make_trans(seed, n=n, m=m, G=G, n_chrom=2, ncov=ncov, miss=miss, spacing=1000,
                          h2=("uniform", 0.0, 0.7), n_causal=300, n_planted=6, beta_planted=(0.3, 1.5))
    rows = ds.bed_rows()
    dos, missm, keep, af = O.decode_bed(rows, ds.n)
    K, N, used = O.grm(dos, missm, keep)
    Q, lam = O.eigh(K)
    gc, gs, ge = windows(ds, 200_000)
    ref = O.trans_scan("lmm", dos, keep, ds.chrom, ds.pos, ds.Y, ds.X(), Q=Q, lam=lam, gene_chrom=gc,
                       gene_cstart=gs, gene_cend=ge, pthr=1.0, dense=True, hit_cap=m * G + 1)
    return ds, rows, (Q, lam), (gc, gs, ge), ref


def dense_from_hits(hits, m, G):
    out = {k: np.full((m, G), np.nan) for k in ("beta", "se", "stat", "p")}
    for k in out:
        out[k][hits["snp"], hits["gene"]] = hits[k]
    return out


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)
