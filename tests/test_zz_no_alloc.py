"""The entry point that allocates nothing, eqtl_gpu_grm_accumulate_noalloc,
against eqtl_gpu_grm_accumulate through the C ABI.

Both run the same kernels on the same rows, so K and N must agree to the GRM
parity bar (1e-12 absolute; measured: bit-identical) and N and the SNP count
exactly. Two calls in a row, the second with a smaller input, so the second
call runs on the scratch the first left behind (its counters must be cleared,
its stale rows must not leak in). Host rows and device rows are both covered,
and a block larger than the fixed scratch is refused with EQTL_E_CAPACITY.
"""
import numpy as np
import pytest

import eqtl_amd
from eqtl_amd import synth
from eqtl_amd._lib import EQTL_E_CAPACITY, EqtlGpuError

pytestmark = pytest.mark.gpu


def _grm_pair(ctx, rows, block):
    K0, N0, u0 = ctx.grm(rows, block=block)
    K1, N1, u1 = ctx.grm(rows, block=block, noalloc=True)
    return (K0, N0, u0), (K1, N1, u1)


@pytest.mark.parametrize("n", [150, 1601])
def test_noalloc_grm_matches_allocating_entry(gpu_ctx, n):
    rng = np.random.default_rng(97 + n)
    # call 1: 1200 SNPs with missing calls (joint-count path); call 2: 300 SNPs without
    for m, miss in ((1200, 0.02), (300, 0.0)):
        geno, _ = synth.draw_genotypes(rng, m, n, 0.01, 0.5, miss)
        geno[0, :] = 0  # monomorphic: skipped by both
        rows = synth.pack_bed(geno)
        gpu_ctx.set_samples(n)
        gpu_ctx.set_policy(True, 0.8, 0.0)
        (K0, N0, u0), (K1, N1, u1) = _grm_pair(gpu_ctx, rows, block=m)
        print(f"n={n} m={m}: max |dK| = {np.abs(K1 - K0).max():.3e}, used {u1}")
        assert u1 == u0 == m - 1
        assert np.array_equal(N1, N0)
        assert np.abs(K1 - K0).max() <= 1e-12


def test_noalloc_grm_device_rows_and_capacity(gpu_ctx):
    n, m = 2000, 1536
    bps = (n + 3) // 4
    rows_dev = gpu_ctx.synth_bed(5, m + 512, n, 0.05, 0.5, 0.01)
    gpu_ctx.set_samples(n)
    gpu_ctx.set_policy(True, 0.8, 0.0)
    host = gpu_ctx.copy_rows(rows_dev, 0, m + 512, bps)
    for mm in (m, 512):  # two calls in a row, the second smaller
        K0, N0, u0 = gpu_ctx.grm(host[:mm], block=mm)
        gpu_ctx.grm_begin()
        gpu_ctx.grm_accumulate_noalloc_dev(rows_dev, mm, bps)
        K1, N1, u1 = gpu_ctx.grm_finish()
        assert u1 == u0
        assert np.array_equal(N1, N0)
        assert np.abs(K1 - K0).max() <= 1e-12
    # 1,664 SNPs at n = 2000 need more than the 32 MiB scratch
    gpu_ctx.grm_begin()
    with pytest.raises(EqtlGpuError) as ei:
        gpu_ctx.grm_accumulate_noalloc_dev(rows_dev, m + 128, bps)
    assert ei.value.code == EQTL_E_CAPACITY
