"""VCF helpers for the tests: synthetic VCF text, and the record descriptors
the host's VcfSource hands to eqtl_gpu_vcf_decode (fixed columns parsed on
the CPU, sample columns located in a text buffer)."""
from __future__ import annotations

import gzip
import os

import numpy as np

from eqtl_amd._lib import VCF_RECORD_DTYPE

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMOKE = os.path.join(GOLDEN, "smoke")


def haploid(chrom: str) -> bool:
    k = chrom[3:] if len(chrom) > 3 and chrom[:3].lower() == "chr" else chrom
    return k in ("X", "x", "Y", "y", "M", "m", "MT", "mt", "Mt")


def read_vcf(path):
    data = gzip.open(path, "rb").read() if path.endswith(".gz") else open(path, "rb").read()
    lines = data.split(b"\n")
    samples, ds_float, body = [], False, []
    for ln in lines:
        if ln.startswith(b"##"):
            if ln.startswith(b"##FORMAT=<") and (b"<ID=DS," in ln or b"<ID=DS>" in ln) and b"Type=Float" in ln:
                ds_float = True
        elif ln.startswith(b"#CHROM"):
            samples = [s.decode() for s in ln.split(b"\t")[9:]]
        elif ln:
            body.append(ln.rstrip(b"\r"))
    return samples, ds_float, body


def descriptors(lines, ds_float):
    """Biallelic lines -> (text, VCF_RECORD_DTYPE array, kept line indices)."""
    parts, recs, kept = [], [], []
    off = 0
    for i, ln in enumerate(lines):
        cols = ln.split(b"\t", 9)
        alt = cols[4]
        n_allele = 1 if alt == b"." else 2 + alt.count(b",")
        if n_allele != 2:
            continue
        keys = cols[8].split(b":")
        gt = keys.index(b"GT") if b"GT" in keys else -1
        ds = keys.index(b"DS") if (ds_float and b"DS" in keys) else -1
        head = len(ln) - len(cols[9]) if len(cols) > 9 else len(ln)
        recs.append((off + head, len(ln) - head, gt, ds, int(haploid(cols[0].decode())), 0))
        parts.append(ln + b"\n")
        off += len(ln) + 1
        kept.append(i)
    return b"".join(parts), np.array(recs, dtype=VCF_RECORD_DTYPE), kept


def gpu_rows(ctx, lines, ds_float, n_file):
    text, recs, kept = descriptors(lines, ds_float)
    ptr = ctx.vcf_decode(text, recs)
    raw = ctx.copy_rows(ptr, 0, recs.size, 4 * n_file)
    return raw.view(np.float32).reshape(recs.size, n_file), kept


GT_TOKENS = ["0|0", "0/0", "0|1", "1|0", "0/1", "1/1", "1|1", "./.", ".|.", ".", "1/.", ".|0", "0", "1",
             "0/1/1", "1|1|0"]


def synth_vcf(rng, n_file, m, *, ds=False, chroms=("1", "X"), weights=None, ds_missing=0.03,
              short_cols=0.02, multi=0.05):
    """Synthetic VCF text lines (header, data lines): GT (with haploid and
    polyploid calls, '.' alleles) or GT:DS (decimals with 1-6 places, '.',
    exponents), a few multi-allelic / ALT '.' records and short columns."""
    hdr = [b"##fileformat=VCFv4.2", b'##FORMAT=<ID=GT,Number=1,Type=String,Description="Genotype">']
    if ds:
        hdr.append(b'##FORMAT=<ID=DS,Number=1,Type=Float,Description="Dosage">')
    hdr.append(b"#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" +
               b"\t".join(f"s{i}".encode() for i in range(n_file)))
    w = np.array(weights if weights is not None else
                 [8, 4, 3, 3, 3, 6, 6, 1, 1, 0.5, 0.3, 0.3, 0.3, 0.3, 0.1, 0.1], float)
    w /= w.sum()
    lines = []
    for j in range(m):
        chrom = chroms[(j * len(chroms)) // m]
        alt = b"T"
        u = rng.random()
        if u < multi / 2:
            alt = b"T,G"
        elif u < multi:
            alt = b"."
        # per-record genotype mix: rare-variant records make monomorphic drops
        shift = rng.random()
        wj = w.copy()
        if shift < 0.1:
            wj[[2, 3, 4, 5, 6]] *= 0.02
            wj /= wj.sum()
        toks = rng.choice(len(GT_TOKENS), size=n_file, p=wj)
        cols = []
        for i in range(n_file):
            gt = GT_TOKENS[toks[i]]
            if ds:
                v = rng.random() * 2
                r = rng.random()
                if r < ds_missing:
                    dsv = "."
                elif r < 0.1:
                    dsv = f"{v:.1f}"
                elif r < 0.2:
                    dsv = f"{v:.6f}"
                elif r < 0.25:
                    dsv = f"{v * 1e3:.3f}e-3"
                elif r < 0.3:
                    dsv = f"{v:.3g}"
                else:
                    dsv = f"{v:.3f}"
                cell = f"{gt}:{dsv}"
                if rng.random() < short_cols:
                    cell = gt  # DS subfield absent
            else:
                cell = gt
                if rng.random() < short_cols:
                    cell = gt + ":99"
            cols.append(cell)
        fmt = b"GT:DS" if ds else b"GT"
        ln = b"\t".join([chrom.encode(), str(1000 + 7 * j).encode(), f"rs{j}".encode(), b"A", alt, b".", b"PASS",
                         b".", fmt] + [c.encode() for c in cols])
        lines.append(ln)
    return hdr, lines
