"""The five benchmark workloads' synthetic inputs, shared by bench.py and the
configuration tests (tests/test_configs.py).

Conventions (seeded numpy PCG64 per workload): A1 frequency MAF or 1 - MAF
(fair coin), HWE genotypes, each call missing at the stated rate; covariates
N(0, 1) plus the intercept; genes on the '+' strand. Genotypes of the large
workloads are generated in HBM by eqtl_gpu_synth_bed (counter-based, so SNP j
depends only on (seed, j)); the few rows the phenotypes need are copied back.

  C1  trans LM     n=200,  10k SNPs (2 chr, 1 kb apart), 500 genes, 5 cov,
                   10 genes with one planted trans SNP (beta 0.5); host arrays
  C2  cis LMM      n=1000, 100k SNPs (10 chr, 2 kb apart), 2000 genes with
                   1 kb bodies and 1 Mb flanks, polygenic h2 ~ U(0, 0.6) + 1-3
                   cis eQTLs per gene, 15 cov
  C3  trans LMM    n=2000, 500k SNPs (10 chr, 1 kb apart), 20k genes,
                   h2 ~ Beta(2, 5) over 1000 causal SNPs, 200 planted trans
                   SNPs (|beta| log-uniform on [0.05, 2]), 20 cov
  C4  cis perm     n=1000, 1M SNPs (10 chr, 100 bp apart), 20k genes with
                   windows of c_g ~ exp(U(ln 50, ln 5000)) SNPs (zero flank),
                   5% with one cis eQTL |beta| ~ U(0.1, 0.5), 15 cov
  C5  GRM + REML   n=5000, 1M SNPs (MAF ~ U(0.01, 0.5), no missing calls),
                   20k intercept-only genes, polygenic h2 ~ U(0, 0.8)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import synth


@dataclass
class Workload:
    n: int
    m: int
    bps: int
    rows_dev: int | None          # device bed rows (None: host rows in `rows`)
    chrom: np.ndarray
    pos: np.ndarray
    X: np.ndarray                 # n x p with the intercept first
    Y: np.ndarray                 # n x G
    win: tuple                    # (gene chrom, cstart, cend) per gene
    names: list
    rows: np.ndarray | None = None
    extra: dict = field(default_factory=dict)


def _row(ctx, rows_dev, j, n, bps):
    """Mean-imputed dosage of device bed row j."""
    g = synth.unpack_bed(ctx.copy_rows(rows_dev, int(j), 1, bps), n)[0].astype(float)
    g[g < 0] = g[g >= 0].mean()
    return g


def c1(seed: int = 1001) -> Workload:
    """Trans LM on host arrays (the reference's CPU-runnable case)."""
    ds = synth.make_trans(seed, n=200, m=10_000, G=500, n_chrom=2, ncov=5, maf=(0.05, 0.5), miss=0.01,
                          spacing=1000, h2=None, n_planted=10, beta_planted=(0.5, 0.5))
    cs = np.maximum(1, ds.gene_start - 1_000_000)
    ce = ds.gene_end + 1_000_000
    rows = ds.bed_rows()
    return Workload(ds.n, ds.m, rows.shape[1], None, ds.chrom, ds.pos, ds.X(), ds.Y,
                    (ds.gene_chrom.astype(np.int32), cs.astype(np.int64), ce.astype(np.int64)), list(ds.genes),
                    rows=rows, extra={"planted": ds.planted, "dataset": ds})


C2_N, C2_M, C2_G, C2_NCOV, C2_NCHROM, C2_FLANK = 1000, 100_000, 2000, 15, 10, 1_000_000


def c2(ctx, seed: int = 1002) -> Workload:
    """Cis LMM with every pair reported; genes sorted by (chromosome, TSS)."""
    n, m, G = C2_N, C2_M, C2_G
    rng = np.random.default_rng(seed)
    bps = (n + 3) // 4
    rows_dev = ctx.synth_bed(seed, m, n, 0.01, 0.5, 0.01)
    per = m // C2_NCHROM
    chrom = np.repeat(np.arange(1, C2_NCHROM + 1, dtype=np.int32), per)
    pos = np.tile(1 + 2000 * np.arange(per, dtype=np.int64), C2_NCHROM)
    gch = rng.integers(1, C2_NCHROM + 1, G).astype(np.int32)
    tss = rng.integers(1, per * 2000 - 1000, G).astype(np.int64)
    order = np.lexsort((tss, gch))
    gch, tss = gch[order], tss[order]
    cs, ce = np.maximum(1, tss - C2_FLANK), tss + 999 + C2_FLANK
    X = np.hstack([np.ones((n, 1)), rng.standard_normal((n, C2_NCOV))])
    causal = np.sort(rng.choice(m, 500, replace=False))
    Z = np.stack([synth.standardized(synth.unpack_bed(ctx.copy_rows(rows_dev, int(j), 1, bps), n))[0]
                  for j in causal])
    h2 = rng.uniform(0, 0.6, G)
    u = Z.T @ rng.standard_normal((500, G))
    u = (u - u.mean(0)) / u.std(0)
    Y = rng.standard_normal((n, G)) + u * np.sqrt(h2 / (1 - h2))
    lo_hi = []
    for k in range(G):
        first = (gch[k] - 1) * per
        lo = first + max(0, (int(cs[k]) - 1 + 1999) // 2000)
        hi = first + min(per - 1, (int(ce[k]) - 1) // 2000)
        lo_hi.append((lo, hi))
        for _ in range(int(rng.integers(1, 4))):
            j = int(rng.integers(lo, hi + 1))
            g = _row(ctx, rows_dev, j, n, bps)
            Y[:, k] += rng.uniform(0.1, 0.5) * (1 if rng.random() < 0.5 else -1) * (g - g.mean())
    return Workload(n, m, bps, rows_dev, chrom, pos, X, Y, (gch, cs, ce), [f"g{k}" for k in range(G)],
                    extra={"window_rows": lo_hi, "per": per})


C3_N, C3_M, C3_G, C3_NCOV, C3_NCHROM = 2000, 500_000, 20_000, 20, 10


def c3(ctx, seed: int = 1003) -> Workload:
    """Trans LMM (the headline metric's configuration)."""
    n, m, G = C3_N, C3_M, C3_G
    rng = np.random.default_rng(seed)
    bps = (n + 3) // 4
    rows_dev = ctx.synth_bed(seed, m, n, 0.05, 0.5, 0.01)
    per = m // C3_NCHROM
    chrom = np.repeat(np.arange(1, C3_NCHROM + 1, dtype=np.int32), per)
    pos = np.tile(1 + 1000 * np.arange(per, dtype=np.int64), C3_NCHROM)
    C = rng.standard_normal((n, C3_NCOV))
    X = np.hstack([np.ones((n, 1)), C])
    # polygenic term over 1000 causal SNPs (rows pulled back from HBM), h2 ~ Beta(2,5)
    causal = np.sort(rng.choice(m, 1000, replace=False))
    Z = np.empty((1000, n))
    for i, j in enumerate(causal):
        g = synth.unpack_bed(ctx.copy_rows(rows_dev, int(j), 1, bps), n)[0]
        Z[i] = synth.standardized(g[None, :])[0]
    h2 = np.clip(rng.beta(2, 5, G), 0, 0.95)
    u = Z.T @ rng.standard_normal((1000, G))
    u = (u - u.mean(0)) / u.std(0)
    Y = rng.standard_normal((n, G)) + u * np.sqrt(h2 / (1 - h2))
    # 200 genes with one planted trans SNP, |beta| log-uniform on [0.05, 2]
    gene_chrom = rng.integers(1, C3_NCHROM + 1, G).astype(np.int32)
    gene_start = rng.integers(1, per * 1000 - 1000, G).astype(np.int64)
    planted = {}
    for k in range(200):
        while True:
            j = int(rng.integers(0, m))
            if chrom[j] != gene_chrom[k]:
                break
        beta = float(np.exp(rng.uniform(np.log(0.05), np.log(2.0)))) * (1 if rng.random() < 0.5 else -1)
        g = _row(ctx, rows_dev, j, n, bps)
        Y[:, k] += beta * (g - g.mean())
        planted[k] = j
    cs = np.maximum(1, gene_start - 1_000_000)
    ce = gene_start + 999 + 1_000_000
    return Workload(n, m, bps, rows_dev, chrom, pos, X, Y, (gene_chrom, cs, ce), [f"g{k}" for k in range(G)],
                    extra={"planted": planted})


C4_N, C4_M, C4_G, C4_NCOV, C4_NCHROM = 1000, 1_000_000, 20_000, 15, 10


def c4(ctx, n_genes: int = C4_G, seed: int = 1004) -> Workload:
    """Cis LM + LMM with gene-level permutations: a seeded slice of n_genes of
    the 20k-gene table (all of it by default), sorted by (chromosome, start)."""
    n, m = C4_N, C4_M
    rng = np.random.default_rng(seed)
    bps = (n + 3) // 4
    rows_dev = ctx.synth_bed(seed, m, n, 0.01, 0.5, 0.01)
    per = m // C4_NCHROM
    chrom = np.repeat(np.arange(1, C4_NCHROM + 1, dtype=np.int32), per)
    pos = np.tile(1 + 100 * np.arange(per, dtype=np.int64), C4_NCHROM)
    cg = np.rint(np.exp(rng.uniform(np.log(50), np.log(5000), C4_G))).astype(np.int64)
    gch = rng.integers(1, C4_NCHROM + 1, C4_G).astype(np.int32)
    start = 1 + (rng.random(C4_G) * (per * 100 - 100 * cg)).astype(np.int64)
    end = start + 100 * cg - 1
    sel = rng.choice(C4_G, n_genes, replace=False)
    sel = sel[np.lexsort((start[sel], gch[sel]))]
    X = np.hstack([np.ones((n, 1)), rng.standard_normal((n, C4_NCOV))])
    Y = rng.standard_normal((n, n_genes))
    for k in range(n_genes):
        if rng.random() < 0.05:
            g0 = sel[k]
            first = (gch[g0] - 1) * per + (start[g0] - 1 + 99) // 100
            j = int(first + rng.integers(0, cg[g0]))
            g = _row(ctx, rows_dev, j, n, bps)
            Y[:, k] += rng.uniform(0.1, 0.5) * (1 if rng.random() < 0.5 else -1) * (g - g.mean())
    names = [f"g{int(k)}" for k in sel]
    first = (gch[sel] - 1) * per + (start[sel] - 1 + 99) // 100
    return Workload(n, m, bps, rows_dev, chrom, pos, X, Y, (gch[sel], start[sel], end[sel]), names,
                    extra={"c_g": cg[sel], "first_row": first})


C5_N, C5_M, C5_G = 5000, 1_000_000, 20_000


def c5(ctx, m: int = C5_M, seed: int = 1005, miss: float = 0.0) -> Workload:
    """GRM over n = 5000 samples (device rows), then 20k intercept-only genes
    with a polygenic term (h2 ~ U(0, 0.8)) for the REML adjoint."""
    n, G = C5_N, C5_G
    rng = np.random.default_rng(seed)
    bps = (n + 3) // 4
    rows_dev = ctx.synth_bed(seed, m, n, 0.01, 0.5, miss)
    causal = np.sort(rng.choice(m, 1000, replace=False))
    Z = np.stack([synth.standardized(synth.unpack_bed(ctx.copy_rows(rows_dev, int(j), 1, bps), n))[0]
                  for j in causal])
    h2 = rng.uniform(0, 0.8, G)
    u = Z.T @ rng.standard_normal((1000, G))
    u = (u - u.mean(0)) / u.std(0)
    Y = rng.standard_normal((n, G)) + u * np.sqrt(h2 / (1 - h2))
    X = np.ones((n, 1))
    return Workload(n, m, bps, rows_dev, np.zeros(0, np.int32), np.zeros(0, np.int64), X, Y, (None, None, None),
                    [f"g{k}" for k in range(G)], extra={"h2": h2})
