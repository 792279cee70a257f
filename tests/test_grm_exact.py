"""The GRM's two paths against the oracle's compute_grm_from restatement
(grm.cpp:112-224): SNPs whose calls are all present go through the grouped
int8 Gram (exact integer Gram per allele-count group, FP64 weights and rank-1
centring), the others through the FP64 Gram of their block.

Bars as every GRM test: K within 1e-12 absolute, N and the SNP count exact.
The cases put both paths in one block, force the held-SNP buffer to run the
grouped Gram between blocks and split groups across Gram launches (small
limits), build groups far larger than a 64-SNP step (repeated rows), and
check that a run is bit-for-bit repeatable (group order and the FP64 sums
are fixed; integer sums do not depend on order).
"""
import numpy as np
import pytest

import oracle as O
from eqtl_amd import synth

pytestmark = pytest.mark.gpu


def _oracle(rows, n, maf_min=0.0):
    dos, missm, keep, _ = O.decode_bed(rows, n, maf_min=maf_min)
    return O.grm(dos, missm, keep)


def _mixed_rows(rng, m, n):
    """Complete SNPs, SNPs with a few missing calls, monomorphic and rare
    (MAF-filtered at 0.02) SNPs, in one interleaved set."""
    geno, _ = synth.draw_genotypes(rng, m, n, 0.01, 0.5, 0.0)
    with_miss = rng.random(m) < 0.3
    for j in np.flatnonzero(with_miss):
        geno[j, rng.choice(n, rng.integers(1, 6), replace=False)] = -1
    geno[rng.choice(m, 20, replace=False)] = 0  # monomorphic: skipped
    rare = rng.choice(m, 20, replace=False)
    geno[rare] = 0
    geno[rare, :2] = 1  # AF 1/n: below --maf 0.02
    return geno, with_miss


@pytest.mark.parametrize("n", [150, 1601])
def test_grm_exact_and_fp64_snps_in_one_block(gpu_ctx, n):
    rng = np.random.default_rng(400 + n)
    geno, with_miss = _mixed_rows(rng, 1800, n)
    rows = synth.pack_bed(geno)
    gpu_ctx.set_samples(n)
    gpu_ctx.set_policy(True, 0.8, 0.02)
    K, N, used = gpu_ctx.grm(rows, block=700)
    Kr, Nr, ur = _oracle(rows, n, maf_min=0.02)
    err = np.abs(K - Kr).max()
    print(f"n={n}: {used} SNPs ({int(with_miss.sum())} with missing calls), max |dK| {err:.3e}")
    assert used == ur
    assert np.array_equal(N, Nr)
    assert err < 1e-12
    K2, N2, _ = gpu_ctx.grm(rows, block=700)
    assert np.array_equal(K2, K) and np.array_equal(N2, N)


def test_grm_exact_buffer_and_chunk_limits(gpu_ctx):
    """Held-SNP buffer of 64 rows (the grouped Gram runs several times per
    block) and Gram launches of 16 steps (groups split across launches), on
    groups of up to ~700 SNPs (one SNP repeated), against the oracle and
    against the default limits."""
    n = 300
    rng = np.random.default_rng(77)
    geno, _ = synth.draw_genotypes(rng, 2400, n, 0.05, 0.5, 0.0)
    geno[300:1000] = geno[5]  # one allele count shared by 701 SNPs
    geno[1500:1530, :3] = -1  # a few SNPs through the FP64 path
    rows = synth.pack_bed(geno)
    gpu_ctx.set_samples(n)
    gpu_ctx.set_policy(True, 0.8, 0.0)
    K0, N0, u0 = gpu_ctx.grm(rows, block=500)
    rowb = ((n + 63) // 64 * 64) // 4
    try:
        gpu_ctx.grm_set_limits(exact_buffer_bytes=64 * rowb, chunk_bytes=1)
        K1, N1, u1 = gpu_ctx.grm(rows, block=500)
    finally:
        gpu_ctx.grm_set_limits(exact_buffer_bytes=2 << 30, chunk_bytes=2 << 30)
    Kr, Nr, ur = _oracle(rows, n)
    e0, e1 = np.abs(K0 - Kr).max(), np.abs(K1 - Kr).max()
    print(f"default limits max |dK| {e0:.3e}; small limits {e1:.3e}; paths differ by {np.abs(K1 - K0).max():.3e}")
    assert u0 == u1 == ur
    assert np.array_equal(N0, Nr) and np.array_equal(N1, Nr)
    assert e0 < 1e-12 and e1 < 1e-12


def test_grm_exact_with_reordered_sample_columns(gpu_ctx):
    """Analysis samples a shuffled subset of the file's columns (the per-call
    decode instead of the byte-wise one), n not a multiple of 16."""
    n_file = 421
    rng = np.random.default_rng(5)
    geno, _ = synth.draw_genotypes(rng, 900, n_file, 0.01, 0.5, 0.0)
    geno[rng.choice(900, 50, replace=False), 7] = -1  # a few SNPs through the FP64 path
    rows = synth.pack_bed(geno)
    cols = np.sort(rng.choice(n_file, 389, replace=False)).astype(np.int32)
    cols = cols[rng.permutation(cols.size)]
    gpu_ctx.set_samples(n_file, cols)
    gpu_ctx.set_policy(True, 0.8, 0.0)
    K, N, used = gpu_ctx.grm(rows, block=400)
    dos, missm, keep, _ = O.decode_bed(rows, n_file, sample_col=cols)
    Kr, Nr, ur = O.grm(dos, missm, keep)
    err = np.abs(K - Kr).max()
    print(f"reordered columns: {used} SNPs, max |dK| {err:.3e}")
    assert used == ur
    assert np.array_equal(N, Nr)
    assert err < 1e-12
