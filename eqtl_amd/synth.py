"""Seeded synthetic eQTL inputs: PLINK bed/bim/fam, phenotype/covariate TSVs, GFF3.

Conventions (shared by tests and the benchmark):
  * A1 frequency f_j = MAF_j or 1-MAF_j (fair coin), genotype ~ Binomial(2, f_j)
    (Hardy-Weinberg), each call independently missing (bed code 01) at `miss`.
  * bed is SNP-major: ceil(n/4) bytes per SNP after the 3-byte magic 6c 1b 01;
    sample i sits at byte i//4, bits 2*(i%4); codes 00=2, 01=missing, 10=1, 11=0
    copies of A1 (the PLINK-1 layout decoded at plink_bed.cpp:221-224, 263-320).
  * covariates ~ N(0,1); the loader adds the intercept column (pheno.cpp:171-193).
  * genes on the '+' strand, ID=g<k>; TSS = start.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

CODE_OF_A1 = np.array([3, 2, 0], dtype=np.uint8)  # A1 count 0,1,2 -> bed code


def pack_bed(geno: np.ndarray) -> np.ndarray:
    """geno (m, n) int8 with -1 = missing -> (m, ceil(n/4)) uint8 bed rows."""
    m, n = geno.shape
    nb = (n + 3) // 4
    codes = np.ones((m, nb * 4), dtype=np.uint8)  # padding → code 01, ignored
    obs = geno >= 0
    c = np.full(geno.shape, 1, dtype=np.uint8)
    c[obs] = CODE_OF_A1[geno[obs]]
    codes[:, :n] = c
    codes = codes.reshape(m, nb, 4)
    return (codes[..., 0] | (codes[..., 1] << 2) | (codes[..., 2] << 4) | (codes[..., 3] << 6)).astype(np.uint8)


def unpack_bed(rows: np.ndarray, n: int) -> np.ndarray:
    """Inverse of pack_bed: (m, nb) uint8 -> (m, n) int8 A1 counts, -1 missing."""
    lut = np.array([2, -1, 1, 0], dtype=np.int8)
    m = rows.shape[0]
    sh = np.array([0, 2, 4, 6], dtype=np.uint8)
    codes = (rows[:, :, None] >> sh[None, None, :]) & 3
    return lut[codes.reshape(m, -1)[:, :n]]


def draw_genotypes(rng: np.random.Generator, m: int, n: int, maf_lo: float, maf_hi: float,
                   miss: float) -> tuple[np.ndarray, np.ndarray]:
    maf = rng.uniform(maf_lo, maf_hi, size=m)
    flip = rng.random(m) < 0.5
    f = np.where(flip, 1.0 - maf, maf)
    g = rng.binomial(2, f[:, None], size=(m, n)).astype(np.int8)
    if miss > 0:
        g[rng.random((m, n)) < miss] = -1
    return g, f


def standardized(g: np.ndarray) -> np.ndarray:
    """Mean-imputed, standardised dosages (rows = SNPs) for polygenic terms."""
    x = g.astype(np.float64)
    x[g < 0] = np.nan
    mu = np.nanmean(x, axis=1, keepdims=True)
    x = np.where(np.isnan(x), mu, x)
    sd = x.std(axis=1, keepdims=True)
    sd[sd == 0] = 1.0
    return (x - mu) / sd


@dataclass
class Dataset:
    geno: np.ndarray            # (m, n) int8, -1 missing
    chrom: np.ndarray           # (m,) int32 chromosome number (1-based)
    pos: np.ndarray             # (m,) int64
    samples: list
    Y: np.ndarray               # (n, G) float64
    C: np.ndarray               # (n, c) float64 covariates without intercept
    genes: list
    gene_chrom: np.ndarray      # (G,) int32
    gene_start: np.ndarray      # (G,) int64
    gene_end: np.ndarray        # (G,) int64
    planted: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.geno.shape[1]

    @property
    def m(self):
        return self.geno.shape[0]

    def X(self) -> np.ndarray:
        """Design with the auto-added intercept first (pheno.cpp:184-189)."""
        return np.hstack([np.ones((self.n, 1)), self.C])

    def bed_rows(self) -> np.ndarray:
        return pack_bed(self.geno)

    def write(self, prefix: str) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(prefix)), exist_ok=True)
        rows = self.bed_rows()
        with open(prefix + ".bed", "wb") as f:
            f.write(bytes([0x6C, 0x1B, 0x01]))
            f.write(rows.tobytes())
        with open(prefix + ".bim", "w") as f:
            for j in range(self.m):
                f.write(f"{self.chrom[j]}\trs{j + 1}\t0\t{self.pos[j]}\tA\tG\n")
        with open(prefix + ".fam", "w") as f:
            for s in self.samples:
                f.write(f"{s}\t{s}\t0\t0\t0\t-9\n")
        with open(prefix + ".pheno.tsv", "w") as f:
            f.write("sample\t" + "\t".join(self.genes) + "\n")
            for i, s in enumerate(self.samples):
                f.write(s + "\t" + "\t".join(f"{v:.17g}" for v in self.Y[i]) + "\n")
        with open(prefix + ".covar.tsv", "w") as f:
            f.write("sample\t" + "\t".join(f"c{j + 1}" for j in range(self.C.shape[1])) + "\n")
            for i, s in enumerate(self.samples):
                f.write(s + "\t" + "\t".join(f"{v:.17g}" for v in self.C[i]) + "\n")
        with open(prefix + ".gff", "w") as f:
            f.write("##gff-version 3\n")
            for k, gname in enumerate(self.genes):
                f.write(f"{self.gene_chrom[k]}\tsynth\tgene\t{self.gene_start[k]}\t{self.gene_end[k]}"
                        f"\t.\t+\t.\tID={gname}\n")


def make_trans(seed: int, n: int, m: int, G: int, n_chrom: int, ncov: int, maf=(0.05, 0.5),
               miss: float = 0.01, spacing: int = 1000, h2=None, n_causal: int = 200,
               n_planted: int = 10, beta_planted=(0.5, 0.5), body: int = 1000) -> Dataset:
    """Trans-scan dataset: uniform SNP grid, genes with 1 kb bodies, optional
    polygenic term and planted trans effects.

    h2: None → pure N(0,1) phenotypes; ("beta", a, b) or ("uniform", lo, hi)
    draws a per-gene heritability and adds a polygenic term over n_causal SNPs.
    n_planted genes get one planted trans effect (a SNP on another chromosome);
    beta_planted=(lo, hi) draws |beta| log-uniformly with a random sign.
    """
    rng = np.random.default_rng(seed)
    geno, _ = draw_genotypes(rng, m, n, maf[0], maf[1], miss)
    per = m // n_chrom
    chrom = np.repeat(np.arange(1, n_chrom + 1, dtype=np.int32), per)
    chrom = np.concatenate([chrom, np.full(m - chrom.size, n_chrom, dtype=np.int32)])
    pos = np.empty(m, dtype=np.int64)
    for c in range(1, n_chrom + 1):
        idx = np.nonzero(chrom == c)[0]
        pos[idx] = 1 + spacing * np.arange(idx.size, dtype=np.int64)
    chrom_len = int(pos.max()) + 1
    Cv = rng.standard_normal((n, ncov))
    Y = rng.standard_normal((n, G))
    planted = {}
    if h2 is not None:
        causal = rng.choice(m, size=min(n_causal, m), replace=False)
        Z = standardized(geno[causal])  # (nc, n)
        if h2[0] == "beta":
            h = rng.beta(h2[1], h2[2], size=G)
        else:
            h = rng.uniform(h2[1], h2[2], size=G)
        b = rng.standard_normal((Z.shape[0], G))
        u = Z.T @ b
        u = (u - u.mean(0)) / np.maximum(u.std(0), 1e-12)
        scale = np.sqrt(np.clip(h, 0, 0.95) / (1 - np.clip(h, 0, 0.95)))
        Y = Y + u * scale[None, :]
    gene_chrom = rng.integers(1, n_chrom + 1, size=G).astype(np.int32)
    gene_start = rng.integers(1, max(2, chrom_len - body), size=G).astype(np.int64)
    gene_end = gene_start + body - 1
    if n_planted:
        lo, hi = beta_planted
        for k in range(min(n_planted, G)):
            cand = np.nonzero(chrom != gene_chrom[k])[0] if n_chrom > 1 else np.arange(m)
            j = int(rng.choice(cand))
            mag = float(np.exp(rng.uniform(np.log(lo), np.log(hi)))) if hi > lo else lo
            beta = mag * (1 if rng.random() < 0.5 else -1)
            g = geno[j].astype(np.float64)
            g[geno[j] < 0] = np.mean(geno[j][geno[j] >= 0])
            Y[:, k] += beta * (g - g.mean())
            planted[k] = (j, beta)
    samples = [f"s{i + 1}" for i in range(n)]
    genes = [f"g{k}" for k in range(G)]
    return Dataset(geno, chrom, pos, samples, Y, Cv, genes, gene_chrom, gene_start, gene_end, planted)


def make_cis(seed: int, n: int, G: int, n_chrom: int, snps_per_chrom: int, ncov: int,
             win_lo: int = 50, win_hi: int = 5000, spacing: int = 100, maf=(0.01, 0.5),
             miss: float = 0.01, frac_planted: float = 0.05) -> Dataset:
    """Cis dataset with log-uniform window sizes c_g in [win_lo, win_hi]: gene bodies
    hold exactly c_g SNPs (run with window 0)."""
    rng = np.random.default_rng(seed)
    m = n_chrom * snps_per_chrom
    geno, _ = draw_genotypes(rng, m, n, maf[0], maf[1], miss)
    chrom = np.repeat(np.arange(1, n_chrom + 1, dtype=np.int32), snps_per_chrom)
    pos = np.tile(spacing * np.arange(1, snps_per_chrom + 1, dtype=np.int64), n_chrom)
    c_g = np.round(np.exp(rng.uniform(np.log(win_lo), np.log(win_hi), size=G))).astype(np.int64)
    c_g = np.minimum(c_g, snps_per_chrom)
    gene_chrom = rng.integers(1, n_chrom + 1, size=G).astype(np.int32)
    span = spacing * snps_per_chrom
    gene_start = np.array([rng.integers(1, max(2, span - spacing * c + 1)) for c in c_g], dtype=np.int64)
    gene_end = gene_start + spacing * c_g - 1
    Cv = rng.standard_normal((n, ncov))
    Y = rng.standard_normal((n, G))
    planted = {}
    for k in range(G):
        if rng.random() < frac_planted:
            idx = np.nonzero((chrom == gene_chrom[k]) & (pos >= gene_start[k]) & (pos <= gene_end[k]))[0]
            if idx.size:
                j = int(rng.choice(idx))
                beta = rng.uniform(0.1, 0.5) * (1 if rng.random() < 0.5 else -1)
                g = geno[j].astype(np.float64)
                g[geno[j] < 0] = np.mean(geno[j][geno[j] >= 0])
                Y[:, k] += beta * (g - g.mean())
                planted[k] = (j, beta)
    samples = [f"s{i + 1}" for i in range(n)]
    genes = [f"g{k}" for k in range(G)]
    return Dataset(geno, chrom, pos, samples, Y, Cv, genes, gene_chrom, gene_start, gene_end, planted)


def grm_closed_form(geno: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """GCTA GRM written directly from its definition, for small inputs.

    K_ij = sum_s z_si z_sj / N_ij over SNPs with allele frequency p in [1e-8, 1-1e-8],
    z = (g - 2p)/sqrt(2p(1-p)) with p from the observed calls and missing calls
    contributing 0; N_ij = number of SNPs where both i and j are observed
    (diagonal: where i is observed). Returns (K, N).
    """
    x = geno.astype(np.float64)
    obs = geno >= 0
    x[~obs] = 0.0
    n_ok = obs.sum(1)
    keep = n_ok > 0
    mu = np.where(keep, x.sum(1) / np.maximum(n_ok, 1), 0.0)
    p = mu / 2
    use = keep & (p >= 1e-8) & (p <= 1 - 1e-8)
    x, obs, mu, p = x[use], obs[use], mu[use], p[use]
    sd = np.sqrt(2 * p * (1 - p))
    z = np.where(obs, (x - mu[:, None]) / sd[:, None], 0.0)
    A = z.T @ z
    o = obs.astype(np.float64)
    N = o.T @ o
    K = np.where(N > 0, A / np.maximum(N, 1), 0.0)
    return K, N


def grm_subblock(rows: np.ndarray, n: int, idx: np.ndarray, max_miss: float = 0.8,
                 chunk: int = 16384) -> tuple[np.ndarray, np.ndarray, int]:
    """The GCTA GRM restricted to the samples idx, from bed rows over all n
    samples, streamed over SNP chunks: K_ij = sum_s z_si z_sj / N_ij with
    z = (g - mu)/sqrt(2p(1-p)), mu the SNP's mean over its observed calls,
    p = mu/2, missing calls 0; SNPs used when they have an observed call, a
    missing share <= max_miss and p in [1e-8, 1 - 1e-8] (and AF in
    (1e-12, 1 - 1e-12)); N_ij = used SNPs where i and j are both observed.
    Per-SNP sums come from byte lookup tables, so only the idx columns are
    unpacked. Returns (K_sub, N_sub, used)."""
    idx = np.asarray(idx)
    codes = np.arange(256, dtype=np.uint8)
    dos_of_code = np.array([2, 0, 1, 0], dtype=np.int64)   # 00=2, 01=missing, 10=1, 11=0
    obs_of_code = np.array([1, 0, 1, 1], dtype=np.int64)
    sh = np.array([0, 2, 4, 6], dtype=np.uint8)
    c4 = (codes[:, None] >> sh[None, :]) & 3                  # (256, 4)
    nb = (n + 3) // 4
    tail = n - 4 * (nb - 1)                                    # samples in the last byte
    lut_sum = dos_of_code[c4].sum(1)
    lut_obs = obs_of_code[c4].sum(1)
    lut_sum_t = dos_of_code[c4[:, :tail]].sum(1)
    lut_obs_t = obs_of_code[c4[:, :tail]].sum(1)
    k = idx.size
    A = np.zeros((k, k))
    N = np.zeros((k, k))
    used = 0
    byte_i, shift_i = idx // 4, (2 * (idx % 4)).astype(np.uint8)
    for j0 in range(0, rows.shape[0], chunk):
        r = rows[j0:j0 + chunk, :nb]
        s = lut_sum[r[:, :-1]].sum(1) + lut_sum_t[r[:, -1]]
        o = lut_obs[r[:, :-1]].sum(1) + lut_obs_t[r[:, -1]]
        mu = np.where(o > 0, s / np.maximum(o, 1), 0.0)
        p = mu / 2
        af = np.clip(p, 0.0, 1.0)
        use = (o > 0) & ((n - o) / n <= max_miss + 1e-15) & (af > 1e-12) & (af < 1 - 1e-12) & \
              (p >= 1e-8) & (p <= 1 - 1e-8)
        cd = (r[:, byte_i] >> shift_i[None, :]) & 3            # (chunk, k) codes of the idx samples
        g = dos_of_code[cd].astype(np.float64)
        ob = obs_of_code[cd].astype(bool) & use[:, None]
        sd = np.sqrt(2 * p * (1 - p))
        z = np.where(ob, (g - mu[:, None]) / np.where(use, sd, 1.0)[:, None], 0.0)
        A += z.T @ z
        of = ob.astype(np.float64)
        N += of.T @ of
        used += int(use.sum())
    K = np.where(N > 0, A / np.maximum(N, 1), 0.0)
    return K, N.astype(np.int64), used
