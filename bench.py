"""Trans-LMM association benchmark on one MI355X.

Workload (the metric's configuration): n = 2000 samples, m = 500,000 SNPs on
10 chromosomes, G = 20,000 genes, 20 covariates + intercept (p = 21), 1%
missing genotypes (mean-imputed), MAF ~ U(0.05, 0.5). Genotypes are generated
in HBM (counter-based hash), phenotypes/covariates on the host; the GRM is built
on the GPU from the same genotypes and eigendecomposed once (untimed set-up).

One step = the full trans-LMM pass: Q^T[Y|X] rotation, per-gene REML delta
(batched Newton), per-gene bases, then decode + rotation + fused Wald +
per-gene reductions over all 500k SNPs (hits p <= 1e-5 copied to the host).
value = Wald tests per second of that step (split-BF16 form). The plain FP32
form is timed on a SNP slice of the same data (tests/s is per-test) and the
split form's p-values are compared with the FP64 oracle on a sample of
(SNP, gene) pairs that includes the planted strong effects.

Other paths of the engine (--workload), each printing one JSON line of the
same shape with its own metric, roofline of its dominant kernel and CPU
baseline: trans-lm (LM scan, same data), trans-glm (NB score tests on counts
drawn from the same phenotypes), cis-lmm (cis LMM scan reporting every pair,
n = 1000, 100k SNPs, 2000 genes, 1 Mb flanks), cis-perm (cis LM + LMM with 1000 gene-level
permutations, n = 1000, a gene slice of a 20k-gene table), grm (the GRM at
n = 5000 x 1M SNPs, then eigh and the REML adjoint), vcf-gt / vcf-ds (device
VCF decode).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, M, G, NCOV, NCHROM = 2000, 500_000, 20_000, 20, 10
BLOCK = 32768
# CPU baseline sample: ~10-20 s of the FP64 port on the host cores
CPU_SNPS, CPU_GENES = 400, 10000
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(ctx, seed=1003):
    """The trans workload's inputs (eqtl_amd.workloads.c3)."""
    from eqtl_amd import workloads
    w = workloads.c3(ctx, seed)
    return w.rows_dev, w.bps, w.chrom, w.pos, w.X, w.Y, w.win, w.extra["planted"]


def trans_pass(ctx, rows_dev, bps, X, Y, m, prec, delta_mode, delta_in=None, pthr=1e-5, Y_dev=None):
    import eqtl_amd
    if Y_dev is not None:
        delta, _ = ctx.lmm_prep(None, X, delta_mode=delta_mode, delta_in=delta_in, Y_dev=Y_dev)
    else:
        delta, _ = ctx.lmm_prep(Y, X, delta_mode=delta_mode, delta_in=delta_in)
    ctx.scan_begin(eqtl_amd.MODEL_LMM, prec, pthr)
    nh = 0
    for j0 in range(0, m, BLOCK):
        mb = min(BLOCK, m - j0)
        hits = ctx.scan_block(None, j0, hit_cap=1 << 20, rows_dev=rows_dev + j0 * bps, m=mb, bps=bps)
        nh += hits.shape[0]
    st = ctx.scan_end()
    return delta, st, nh


def accuracy_sample(ctx, rows_dev, bps, X, Y, win, delta, planted, prec, Q, lam):
    """Max relative p error of the split form vs the FP64 oracle on sampled pairs."""
    import eqtl_amd
    import oracle as O
    from eqtl_amd import synth
    rng = np.random.default_rng(7)
    genes = np.unique(np.concatenate([np.arange(64), rng.choice(G, 64, replace=False)]))
    snps = np.unique(np.concatenate([np.array([planted[k] for k in genes if k in planted]),
                                     rng.choice(M, 400, replace=False)]))
    rows = np.concatenate([ctx.copy_rows(rows_dev, int(j), 1, bps) for j in snps])
    Ys, ds = Y[:, genes], delta[genes]
    ctx.lmm_prep(Ys, X, delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=ds)
    ctx.set_gene_windows(win[0][genes], win[1][genes], win[2][genes])
    ctx.set_snp_annot(np.zeros(len(snps), np.int32), np.arange(len(snps), dtype=np.int64))
    hits, _ = ctx.scan(eqtl_amd.MODEL_LMM, prec, rows, pthr=1.0, block=4096, hit_cap=len(snps) * len(genes) + 1)
    dos, _, keep, _ = O.decode_bed(rows, N)
    ref = O.trans_scan("lmm", dos, keep, np.zeros(len(snps)), np.arange(len(snps)), Ys, X, Q=Q, lam=lam,
                       delta=ds, pthr=1.0, dense=True, hit_cap=1)
    pg = np.full(ref["p"].shape, np.nan)
    pg[hits["snp"], hits["gene"]] = hits["p"]
    ok = np.isfinite(ref["p"]) & (ref["p"] >= 1e-300)
    rel = np.abs(pg[ok] - ref["p"][ok]) / ref["p"][ok]
    tg = np.full(ref["p"].shape, np.nan)
    tg[hits["snp"], hits["gene"]] = hits["stat"]
    bg = np.full(ref["p"].shape, np.nan)
    bg[hits["snp"], hits["gene"]] = hits["beta"]
    sg = np.full(ref["p"].shape, np.nan)
    sg[hits["snp"], hits["gene"]] = hits["se"]
    big = ok & (np.abs(ref["stat"]) > 0.1)
    return {"pairs": int(ok.sum()), "min_p": float(ref["p"][ok].min()), "max_rel_p": float(rel.max()),
            "max_rel_beta": float(np.max(np.abs(bg[big] - ref["beta"][big]) / np.abs(ref["beta"][big]))),
            "max_rel_se": float(np.max(np.abs(sg[ok] - ref["se"][ok]) / ref["se"][ok])),
            "max_abs_dt": float(np.max(np.abs(tg[ok] - ref["stat"][ok])))}


def glm_counts(Y, seed=1005):
    """NB2 counts from the workload's phenotypes (planted effects included):
    mu = base * exp(0.25 y), base ~ U(5, 50), phi ~ U(0.05, 1) per gene."""
    rng = np.random.default_rng(seed)
    Gn = Y.shape[1]
    base = rng.uniform(5, 50, Gn)
    phi = rng.uniform(0.05, 1.0, Gn)
    mu = base[None, :] * np.exp(0.25 * Y)
    lam = rng.gamma(1.0 / phi[None, :], phi[None, :] * mu)
    return rng.poisson(lam).astype(np.float64)


def glm_pass(ctx, rows_dev, bps, X, Yc, m, prec, pthr=1e-5):
    import eqtl_amd
    phi, conv = ctx.glm_prep(Yc, X)
    ctx.scan_begin(eqtl_amd.MODEL_GLM, prec, pthr)
    nh = 0
    for j0 in range(0, m, BLOCK):
        mb = min(BLOCK, m - j0)
        hits = ctx.scan_block(None, j0, hit_cap=1 << 20, rows_dev=rows_dev + j0 * bps, m=mb, bps=bps)
        nh += hits.shape[0]
    return phi, conv, ctx.scan_end(), nh


def bench_glm(args):
    """Trans NB GLM score-test scan on the metric's configuration (n = 2000,
    500k SNPs, 20k count genes, p = 21); one step = the per-gene IRLS null fits
    + decode + fused score-test kernel over all SNPs."""
    import eqtl_amd
    import oracle as O
    global BLOCK
    BLOCK = args.block
    m = min(args.snps or M, M)
    ctx = eqtl_amd.Context(0)
    t0 = time.perf_counter()
    rows_dev, bps, chrom, pos, X, Y, win, planted = make_inputs(ctx)
    Yc = glm_counts(Y)
    del Y
    log(f"inputs ready {time.perf_counter() - t0:.1f}s")
    ctx.set_samples(N)
    ctx.set_policy(True, 0.8, 0.0)
    ctx.set_snp_annot(chrom, pos)
    ctx.set_gene_windows(*win)
    prec = eqtl_amd.PREC_BF16X3
    for _ in range(args.warmup):
        glm_pass(ctx, rows_dev, bps, X, Yc, min(m, 4 * BLOCK), prec)
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    tests = 0
    for s in range(args.steps):
        phi, conv, st, nh = glm_pass(ctx, rows_dev, bps, X, Yc, m, prec)
        tests += int(st["n_tested"].sum())
        log(f"step {s}: {int(st['n_tested'].sum())} tests, {nh} hits, {int(conv.sum())} genes fitted")
    ctx.synchronize()
    dt = time.perf_counter() - t0
    wald_ms, wald_n = ctx.timing("wald")
    prep_ms, _ = ctx.timing("glm_prep")
    dec_ms, _ = ctx.timing("decode")
    ctx.set_timing(False)
    log(f"step time {dt / args.steps:.2f}s: kernel {wald_ms / args.steps:.0f} ms, null fits "
        f"{prep_ms / args.steps:.0f} ms, decode {dec_ms / args.steps:.0f} ms")
    p = X.shape[1]
    launch_ms = wald_ms / max(wald_n, 1)
    achieved = 2.0 * N * (p + 2) * (tests / max(wald_n, 1)) / (launch_ms * 1e-3) / 1e12
    extra = {}
    if not args.no_accuracy:
        # split form vs the FP64 oracle on sampled pairs (planted genes and SNPs included)
        rng = np.random.default_rng(8)
        genes = np.unique(np.concatenate([np.arange(48), rng.choice(G, 48, replace=False)]))
        snps = np.unique(np.concatenate([np.array([planted[k] for k in genes if k in planted]),
                                         rng.choice(M, 300, replace=False)]))
        rows = np.concatenate([ctx.copy_rows(rows_dev, int(j), 1, bps) for j in snps])
        Ys = Yc[:, genes]
        ctx.glm_prep(Ys, X)
        ctx.set_gene_windows(win[0][genes], win[1][genes], win[2][genes])
        ctx.set_snp_annot(np.zeros(len(snps), np.int32), np.arange(len(snps), dtype=np.int64))
        hits, _ = ctx.scan(eqtl_amd.MODEL_GLM, prec, rows, pthr=1.0, block=4096,
                           hit_cap=len(snps) * len(genes) + 1)
        dos, _, keep, _ = O.decode_bed(rows, N)
        ref = O.glm_scan(dos, keep, np.zeros(len(snps)), np.arange(len(snps)), Ys, X, pthr=1.0, dense=True,
                         hit_cap=1)
        pg = np.full(ref["p"].shape, np.nan)
        pg[hits["snp"], hits["gene"]] = hits["p"]
        ok = np.isfinite(ref["p"]) & (ref["p"] >= 1e-300)
        extra["accuracy_vs_fp64"] = {"pairs": int(ok.sum()), "min_p": float(ref["p"][ok].min()),
                                     "max_rel_p": float(np.max(np.abs(pg[ok] - ref["p"][ok]) / ref["p"][ok]))}
        log("accuracy", extra["accuracy_vs_fp64"])
    cpu = None
    if not args.no_cpu:
        ns, ng = 2000, 4000
        rows_s = ctx.copy_rows(rows_dev, 0, ns, bps)
        dos, _, keep, _ = O.decode_bed(rows_s, N)
        threads = O.num_threads()
        Ys = Yc[:, :ng]
        t1 = time.perf_counter()
        ref = O.glm_scan(dos, keep, np.zeros(ns), np.arange(ns), Ys, X, pthr=1e-5, hit_cap=1 << 16)
        dtc = time.perf_counter() - t1
        nt = int(ref["state"]["n_tested"].sum())
        cpu = {"value": nt / dtc, "unit": "NB score tests/s", "cores": threads, "kind": "port",
               "sample": f"{ns} SNPs x {ng} genes of the same data through the oracle's restatement of "
                         f"prep_glm_nb + test_glm_nb (null fits included) on {threads} threads "
                         f"({nt} tests in {dtc:.1f} s)"}
        log("cpu baseline", cpu)
    out = {
        "metric": "NB GLM score tests/s, trans (split-BF16 contraction, FP64 epilogue)",
        "value": tests / dt, "unit": "NB score tests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3 contractions, fp64 epilogue", "data": "synthetic (generated in HBM, seeded; NB2 counts)",
        "config": {"workload": "trans NB GLM n=2000 samples x 500k SNPs x 20k count genes, p=21, 1% missing",
                   "snps_per_step": m, "genes": G, "samples": N, "covariates": p, "snps_per_block": BLOCK},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None,
                     "note": "algorithmic 2n(p+2) FLOP/test over the fused kernel's event-timed launches"},
        "cpu_baseline": cpu,
        "kernel_ms_per_launch": launch_ms, "null_fit_ms_per_step": prep_ms / args.steps,
        **extra,
    }
    print(json.dumps(out))
    ctx.close()


def cpu_baseline(Q, lam, X, Y, rows, delta):
    """FP64 oracle (port of the reference's SNP loop) timed on the host cores."""
    import oracle as O
    threads = O.num_threads()
    dos, _, keep, _ = O.decode_bed(rows, N)
    n_snp = rows.shape[0]
    res = O.trans_scan("lmm", dos, keep, np.zeros(n_snp), np.arange(n_snp), Y, X, Q=Q, lam=lam, delta=delta,
                       pthr=1e-5, hit_cap=1 << 16)
    dt = res["loop_seconds"]
    tests = int(res["state"]["n_tested"].sum())
    out = {"value": tests / dt, "unit": "Wald tests/s", "cores": threads, "kind": "port",
           "sample": f"{n_snp} SNPs x {Y.shape[1]} genes of the same data (n={N}, p={X.shape[1]}, the genes' "
                     f"own delta), FP64 restatement of the reference SNP-outer loop (Q^T g per SNP, "
                     f"test_lmm_gtil per gene) on {threads} threads, timed over the SNP loop only, per-gene "
                     f"prep untimed ({tests} tests in {dt:.1f} s)"}
    cal = reference_calibration(Q, lam, X, Y[:, 0], dos[keep.astype(bool)][:CAL_SNPS])
    if cal is not None:
        out["calibration"] = cal
    return out


CAL_SNPS = 150


def reference_calibration(Q, lam, X, y, dos):
    """The port against the reference's own compiled code on identical work, one
    thread each: prep_lmm (REML search) + test_lmm (Q^T g and the bordered-Schur
    test per SNP, model_lmm.cpp:124-223) for one gene over the given SNP rows,
    as the reference's cis scan calls them. None where the reference build
    (oracle/_ref) is absent."""
    import ctypes
    import oracle as O
    R = O.ref_lib()
    if R is None:
        return None
    dos = np.ascontiguousarray(dos, dtype=np.float64)
    c, n = dos.shape
    p = X.shape[1]
    Xf = np.asfortranarray(X, dtype=np.float64)
    Qf = np.asfortranarray(Q, dtype=np.float64)
    yv = np.ascontiguousarray(y, dtype=np.float64)
    lv = np.ascontiguousarray(lam, dtype=np.float64)
    t0 = time.perf_counter()
    port = O.cis_gene("lmm", yv, Xf, dos, B=0, gene="cal", Q=Qf, lam=lv)
    t_port = time.perf_counter() - t0
    res = np.empty(c * 5)
    prep = np.empty(4)
    vp = ctypes.c_void_p
    t0 = time.perf_counter()
    rc = R.ref_lmm(yv.ctypes.data_as(vp), Xf.ctypes.data_as(vp), ctypes.c_int(n), ctypes.c_int(p),
                   Qf.ctypes.data_as(vp), lv.ctypes.data_as(vp), dos.ctypes.data_as(vp), ctypes.c_int(c),
                   res.ctypes.data_as(vp), prep.ctypes.data_as(vp))
    t_ref = time.perf_counter() - t0
    if rc != 0:
        return None
    pr = res.reshape(c, 5)[:, 3]
    ok = np.isfinite(port["p"]) & np.isfinite(pr) & (pr > 0)
    return {"port_tests_per_s": c / t_port, "reference_tests_per_s": c / t_ref,
            "port_over_reference": t_ref / t_port,
            "max_p_rel_diff": float(np.max(np.abs(port["p"][ok] - pr[ok]) / pr[ok])) if ok.any() else None,
            "sample": f"one gene x {c} SNPs (n={n}, p={p}): REML delta search + Q^T g and the test per SNP, "
                      f"the oracle's restatement vs the reference's model_lmm.cpp compiled from its own "
                      f"sources (oracle/_ref), one thread each"}


def measured_traffic(block, kernel, name="lr_kernels_trans_lmm_n2000_g20000_pmc.md"):
    """HBM bytes per launch of one kernel of this workload from the committed
    rocprofv3 PMC summary (line 'traffic_bytes_per_launch.<kernel>: B'):
    FETCH_SIZE and WRITE_SIZE runs, each scaled by the same counter on a 1 GiB
    device copy. None when the summary was taken at another block size or
    does not cover the kernel."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch.") or line.startswith("snps_per_launch:"):
            k, v = line.split(":", 1)
            vals[k.strip()] = float(v.strip())
    key = "traffic_bytes_per_launch." + kernel
    if vals.get("snps_per_launch") != block or key not in vals:
        return None
    return vals[key]


VCF_N, VCF_M = 2000, 32768


def synth_vcf_text(rng, n, m, ds):
    """Fixed-width synthetic VCF data lines (n samples, m records): GT phased
    diploid calls from per-record AF ~ U(0.05, 0.5), 1% missing; or GT:DS with a
    three-decimal dosage near the call. Returns (text uint8, record descriptors)."""
    from eqtl_amd._lib import VCF_RECORD_DTYPE
    fmt = "GT:DS" if ds else "GT"
    prefixes = "".join(f"1\t{1000 + 10 * j:09d}\trs{j:09d}\tA\tT\t.\tPASS\t.\t{fmt}\t" for j in range(m))
    pre = np.frombuffer(prefixes.encode(), np.uint8).reshape(m, -1)
    cell = 10 if ds else 4
    af = rng.uniform(0.05, 0.5, m)[:, None]
    a = (rng.random((m, n)) < af).astype(np.uint8)
    b = (rng.random((m, n)) < af).astype(np.uint8)
    body = np.empty((m, n, cell), np.uint8)
    body[:, :, 0] = 48 + a
    body[:, :, 1] = ord("|")
    body[:, :, 2] = 48 + b
    if ds:
        k = np.clip(np.rint((a + b + rng.normal(0, 0.05, (m, n))) * 1000), 0, 2000).astype(np.int64)
        body[:, :, 3] = ord(":")
        body[:, :, 4] = 48 + k // 1000
        body[:, :, 5] = ord(".")
        body[:, :, 6] = 48 + (k // 100) % 10
        body[:, :, 7] = 48 + (k // 10) % 10
        body[:, :, 8] = 48 + k % 10
    else:
        miss = rng.random((m, n)) < 0.01
        body[:, :, 0][miss] = ord(".")
        body[:, :, 2][miss] = ord(".")
    body[:, :, cell - 1] = ord("\t")
    body[:, -1, cell - 1] = ord("\n")
    lines = np.concatenate([pre, body.reshape(m, n * cell)], axis=1)
    L = lines.shape[1]
    recs = np.zeros(m, dtype=VCF_RECORD_DTYPE)
    recs["offset"] = np.arange(m, dtype=np.int64) * L + pre.shape[1]
    recs["length"] = n * cell - 1
    recs["gt_key"] = 0
    recs["ds_key"] = 1 if ds else -1
    return np.ascontiguousarray(lines).reshape(-1), recs


def vcf_traffic(ds, m, n):
    """HBM bytes per k_vcf_decode launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by a 1 GiB device copy), for the
    32,768 x 2000 workload only."""
    path = os.path.join(ROOT, "profiles", "vcf_decode_n2000_m32768.md")
    if (m, n) != (VCF_M, VCF_N) or not os.path.exists(path):
        return None
    key = "ds_traffic_bytes_per_launch:" if ds else "gt_traffic_bytes_per_launch:"
    for line in open(path):
        if line.startswith(key):
            return float(line.split(":", 1)[1])
    return None


def bench_vcf(args, ds):
    """Device VCF decode (eqtl_gpu_vcf_decode) on 32,768 records x 2000 samples
    with the text resident in HBM; one step = one call (descriptor upload,
    kernel, error flag read back)."""
    import torch
    import eqtl_amd
    rng = np.random.default_rng(77)
    n, m = VCF_N, VCF_M
    text, recs = synth_vcf_text(rng, n, m, ds)
    log(f"vcf text {text.size / 1e6:.0f} MB ({'GT:DS' if ds else 'GT'})")
    tdev = torch.empty(text.size + 64, dtype=torch.uint8, device="cuda")
    tdev[: text.size].copy_(torch.from_numpy(text))
    torch.cuda.synchronize()
    ctx = eqtl_amd.Context(0)
    ctx.set_samples(n)
    ctx.set_policy(True, 0.8, 0.0)
    call = lambda: ctx.vcf_decode(None, recs, text_dev=tdev.data_ptr(), text_len=text.size)
    for _ in range(max(args.warmup, 1)):
        call()
    # spot parity against the oracle on a few records (checker only)
    import oracle as O
    rows_dev = call()
    chk = [0, 1, m // 2, m - 1]
    got = np.vstack([ctx.copy_rows(rows_dev, j, 1, 4 * n).view(np.float32) for j in chk])
    L = int(recs["offset"][1] - recs["offset"][0])
    for i, j in enumerate(chk):
        line = text[j * L:(j + 1) * L - 1].tobytes()
        st, d, af, _, raw = O.vcf_record(line, n, ds_float=ds)
        ok = np.array_equal(got[i], raw.astype(np.float32), equal_nan=True) if st == 1 else np.isnan(got[i]).all()
        if not ok:
            raise SystemExit(f"bench: VCF record {j} differs from the oracle")
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    k_ms, k_n = ctx.timing("vcf")
    ctx.set_timing(False)
    calls = float(n) * m
    value = calls * args.steps / dt
    launch_ms = k_ms / max(k_n, 1)
    text_bytes = float(recs["length"].sum())
    alg_bytes = text_bytes + 4.0 * n * m + recs.nbytes  # text read + float rows written + descriptors
    achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
    # PCIe-inclusive rate: the same call with the text handed over from host memory
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.vcf_decode(text.tobytes(), recs)
    ctx.synchronize()
    host_rate = calls / (time.perf_counter() - t1)
    cpu = None
    if not args.no_cpu:
        t1 = time.perf_counter()
        k = 0
        while k < m and (k < 64 or time.perf_counter() - t1 < 10.0):
            O.vcf_record(text[k * L:(k + 1) * L - 1].tobytes(), n, ds_float=ds)
            k += 1
        dtc = time.perf_counter() - t1
        cpu = {"value": k * n / dtc, "unit": "genotype calls/s", "cores": 1, "kind": "port",
               "sample": f"{k} records x {n} samples of the same text through the oracle's restatement of "
                         f"VcfSession::parse_record (one thread, {dtc:.1f} s)"}
    out = {
        "metric": "VCF genotype calls decoded/s (VCF text in HBM -> float dosage rows, parse_record rules)",
        "value": value, "unit": "genotype calls/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8 text -> f32 dosage", "data": "synthetic (seeded, fixed-width lines)",
        "config": {"workload": f"VCF {'GT:DS' if ds else 'GT'} n={n} samples x {m} records per call",
                   "text_bytes_per_call": text_bytes},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": vcf_traffic(ds, m, n),
                     "note": "algorithmic bytes = sample-column text read + 4 B per call written + 24 B per "
                             "record descriptor, over the k_vcf_decode launch's HIP-event time"},
        "cpu_baseline": cpu,
        "kernel_ms_per_call": launch_ms,
        "pcie_inclusive_calls_per_s": host_rate,
    }
    print(json.dumps(out))
    ctx.close()


def lm_pass(ctx, rows_dev, bps, X, Y, m, prec, pthr=1e-5):
    import eqtl_amd
    ctx.lm_prep(Y, X)
    ctx.scan_begin(eqtl_amd.MODEL_LM, prec, pthr)
    nh = 0
    for j0 in range(0, m, BLOCK):
        mb = min(BLOCK, m - j0)
        hits = ctx.scan_block(None, j0, hit_cap=1 << 20, rows_dev=rows_dev + j0 * bps, m=mb, bps=bps)
        nh += hits.shape[0]
    return ctx.scan_end(), nh


def bench_trans_lm(args):
    """Trans LM scan on the metric's genotypes and phenotypes (n = 2000, 500k
    SNPs, 20k genes, p = 21); one step = per-gene residualisation (prep_lm) +
    decode + fused LM kernel (per SNP FP64 g^T g, per test one length-n dot
    product) + per-gene reductions over all SNPs."""
    import eqtl_amd
    import oracle as O
    global BLOCK
    BLOCK = args.block
    m = min(args.snps or M, M)
    ctx = eqtl_amd.Context(0)
    t0 = time.perf_counter()
    rows_dev, bps, chrom, pos, X, Y, win, planted = make_inputs(ctx)
    log(f"inputs ready {time.perf_counter() - t0:.1f}s")
    ctx.set_samples(N)
    ctx.set_policy(True, 0.8, 0.0)
    ctx.set_snp_annot(chrom, pos)
    ctx.set_gene_windows(*win)
    prec = eqtl_amd.PREC_BF16X3
    for _ in range(args.warmup):
        lm_pass(ctx, rows_dev, bps, X, Y, min(m, 4 * BLOCK), prec)
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    tests = 0
    for s in range(args.steps):
        st, nh = lm_pass(ctx, rows_dev, bps, X, Y, m, prec)
        tests += int(st["n_tested"].sum())
        log(f"step {s}: {int(st['n_tested'].sum())} tests, {nh} hits")
    ctx.synchronize()
    dt = time.perf_counter() - t0
    wald_ms, wald_n = ctx.timing("wald")
    dec_ms, _ = ctx.timing("decode")
    ctx.set_timing(False)
    log(f"step time {dt / args.steps:.2f}s: kernel {wald_ms / args.steps:.0f} ms, decode {dec_ms / args.steps:.0f} ms")
    launch_ms = wald_ms / max(wald_n, 1)
    achieved = 2.0 * N * (tests / max(wald_n, 1)) / (launch_ms * 1e-3) / 1e12
    extra = {}
    if not args.no_accuracy:
        rng = np.random.default_rng(9)
        genes = np.unique(np.concatenate([np.arange(64), rng.choice(G, 64, replace=False)]))
        snps = np.unique(np.concatenate([np.array([planted[k] for k in genes if k in planted]),
                                         rng.choice(M, 400, replace=False)]))
        rows = np.concatenate([ctx.copy_rows(rows_dev, int(j), 1, bps) for j in snps])
        Ys = Y[:, genes]
        ctx.lm_prep(Ys, X)
        ctx.set_gene_windows(win[0][genes], win[1][genes], win[2][genes])
        ctx.set_snp_annot(np.zeros(len(snps), np.int32), np.arange(len(snps), dtype=np.int64))
        hits, _ = ctx.scan(eqtl_amd.MODEL_LM, prec, rows, pthr=1.0, block=4096, hit_cap=len(snps) * len(genes) + 1)
        dos, _, keep, _ = O.decode_bed(rows, N)
        ref = O.trans_scan("lm", dos, keep, np.zeros(len(snps)), np.arange(len(snps)), Ys, X, pthr=1.0, dense=True,
                           hit_cap=1)
        pg = np.full(ref["p"].shape, np.nan)
        pg[hits["snp"], hits["gene"]] = hits["p"]
        ok = np.isfinite(ref["p"]) & (ref["p"] >= 1e-300)
        extra["accuracy_vs_fp64"] = {"pairs": int(ok.sum()), "min_p": float(ref["p"][ok].min()),
                                     "max_rel_p": float(np.max(np.abs(pg[ok] - ref["p"][ok]) / ref["p"][ok]))}
        log("accuracy", extra["accuracy_vs_fp64"])
    cpu = None
    if not args.no_cpu:
        ns = 1000
        rows_s = ctx.copy_rows(rows_dev, 0, ns, bps)
        dos, _, keep, _ = O.decode_bed(rows_s, N)
        threads = O.num_threads()
        t1 = time.perf_counter()
        ref = O.trans_scan("lm", dos, keep, np.zeros(ns), np.arange(ns), Y, X, pthr=1e-5, hit_cap=1 << 16)
        dtc = time.perf_counter() - t1
        nt = int(ref["state"]["n_tested"].sum())
        cpu = {"value": nt / dtc, "unit": "LM tests/s", "cores": threads, "kind": "port",
               "sample": f"{ns} SNPs x {G} genes of the same data through the oracle's restatement of the "
                         f"reference SNP-outer LM loop (prep_lm per gene, then residualize_g + GEMV + "
                         f"hit_from_gty per SNP) on {threads} threads, prep included ({nt} tests in {dtc:.1f} s)"}
        log("cpu baseline", cpu)
    out = {
        "metric": "LM tests/s, trans (split-BF16 contraction, FP64 epilogue)",
        "value": tests / dt, "unit": "LM tests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3 contractions, fp64 g^T g and epilogue", "data": "synthetic (generated in HBM, seeded)",
        "config": {"workload": "trans LM n=2000 samples x 500k SNPs x 20k genes, p=21, 1% missing",
                   "snps_per_step": m, "genes": G, "samples": N, "covariates": X.shape[1], "snps_per_block": BLOCK},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None,
                     "note": "algorithmic 2n FLOP/test over the fused LM kernel's event-timed launches"},
        "cpu_baseline": cpu,
        "kernel_ms_per_launch": launch_ms,
        **extra,
    }
    print(json.dumps(out))
    ctx.close()


CIS_N, CIS_M, CIS_G, CIS_NCOV, CIS_NCHROM = 1000, 1_000_000, 20_000, 15, 10


def cis_inputs(ctx, n_genes, seed=1004):
    """Cis permutation workload (eqtl_amd.workloads.c4): a seeded slice of
    n_genes genes of the 20k-gene table, sorted by (chromosome, start)."""
    from eqtl_amd import workloads
    w = workloads.c4(ctx, n_genes, seed)
    return w.rows_dev, w.bps, w.chrom, w.pos, w.X, w.Y, w.win, w.extra["c_g"], w.names


def cis_model_pass(ctx, model, rows_dev, bps, X, Y, win, hashes, B, seed, prec, delta_mode):
    """Observed cis scan of every gene's window, then B gene-level permutations
    (eqtl_gpu_cis_perm). Returns (observed tests, permuted tests, p_emp)."""
    import eqtl_amd
    if model == eqtl_amd.MODEL_LM:
        ctx.lm_prep(Y, X)
    else:
        ctx.lmm_prep(Y, X, delta_mode=delta_mode)
    ctx.set_gene_windows(*win)
    ctx.scan_begin(model, prec, 1e-5, eqtl_amd.SCOPE_CIS)
    for j0 in range(0, CIS_M, BLOCK):
        mb = min(BLOCK, CIS_M - j0)
        ctx.scan_block(None, j0, hit_cap=1 << 20, rows_dev=rows_dev + j0 * bps, m=mb, bps=bps)
    st = ctx.scan_end()
    pm, pe = ctx.cis_perm(prec, B, seed, hashes, st["best_p"], delta_mode=delta_mode)
    n_obs = st["n_tested"].astype(np.int64)
    return int(n_obs.sum()), int(n_obs[np.isfinite(pe)].sum()) * B, pe


def bench_cis_perm(args):
    """Cis LM + LMM scans with B = 1000 gene-level permutations (the
    --model lm,lmm --mode cis -w 0 --perm 1000 run) on a gene slice of the
    20k-gene workload; one step = for each model the per-gene prep, the observed
    window scan and all permutations (per (gene, b): mt19937 shuffle, null fit
    with a fresh REML delta for the LMM, window scan)."""
    import eqtl_amd
    import oracle as O
    global BLOCK
    BLOCK = args.block
    B, seed = args.perm, 42
    ctx = eqtl_amd.Context(0)
    t0 = time.perf_counter()
    rows_dev, bps, chrom, pos, X, Y, win, cg, names = cis_inputs(ctx, args.genes)
    ctx.set_samples(CIS_N)
    ctx.set_policy(True, 0.8, 0.0)
    import ctypes
    from eqtl_amd._lib import lib, _check
    _check(lib().eqtl_gpu_grm_begin(ctx.handle))
    for j0 in range(0, CIS_M, BLOCK):
        mb = min(BLOCK, CIS_M - j0)
        _check(lib().eqtl_gpu_grm_accumulate(ctx.handle, ctypes.c_void_p(rows_dev + j0 * bps), ctypes.c_int32(1),
                                             ctypes.c_int64(mb), ctypes.c_int32(bps)))
    K, _, _ = ctx.grm_finish()
    Q, lam = ctx.eigh(K)
    ctx.set_snp_annot(chrom, pos)
    ctx.upload_bed(rows_dev=rows_dev, m=CIS_M, bps=bps)
    hashes = np.array([O.fnv1a(nm) for nm in names], dtype=np.uint32)
    log(f"inputs ready {time.perf_counter() - t0:.1f}s: {args.genes} genes, {int(cg.sum())} window SNPs, B={B}")
    prec, dmode = eqtl_amd.PREC_BF16X3, eqtl_amd.DELTA_REF
    if args.warmup:
        w = min(args.genes, 16)
        for model in (eqtl_amd.MODEL_LM, eqtl_amd.MODEL_LMM):
            cis_model_pass(ctx, model, rows_dev, bps, X, Y[:, :w], tuple(a[:w] for a in win), hashes[:w],
                           min(B, 50), seed, prec, dmode)
    res = {}
    ctx.synchronize()
    t_all = time.perf_counter()
    for s in range(args.steps):
        for name, model in (("lm", eqtl_amd.MODEL_LM), ("lmm", eqtl_amd.MODEL_LMM)):
            ctx.set_timing(True)
            t1 = time.perf_counter()
            n_obs, n_perm, pe = cis_model_pass(ctx, model, rows_dev, bps, X, Y, win, hashes, B, seed, prec, dmode)
            ctx.synchronize()
            dtm = time.perf_counter() - t1
            fam = {f: ctx.timing(f) for f in ("wald", "reml", "rotate", "basis", "decode", "residualize", "exact", "polish")}
            ctx.set_timing(False)
            r = res.setdefault(name, {"s": 0.0, "obs_tests": 0, "perm_tests": 0, "fam": {}})
            r["s"] += dtm
            r["obs_tests"] += n_obs
            r["perm_tests"] += n_perm
            for f, (ms, nl) in fam.items():
                a = r["fam"].setdefault(f, [0.0, 0])
                a[0] += ms
                a[1] += nl
            log(f"step {s} {name}: {n_obs} observed, {n_perm} permuted tests in {dtm:.2f}s; " +
                ", ".join(f"{f} {ms:.0f} ms/{nl}" for f, (ms, nl) in fam.items()))
    dt = time.perf_counter() - t_all
    tests = sum(r["obs_tests"] + r["perm_tests"] for r in res.values())
    p = X.shape[1]
    # dominant kernel: the LMM permutations' fused Wald kernel (2n(p+2) FLOP per test)
    lw_ms, lw_n = res["lmm"]["fam"]["wald"]
    lmm_tests = res["lmm"]["obs_tests"] + res["lmm"]["perm_tests"]
    achieved = 2.0 * CIS_N * (p + 2) * lmm_tests / (lw_ms * 1e-3) / 1e12 if lw_ms > 0 else None
    cpu = None
    if not args.no_cpu:
        # the oracle's restatement of scan_gene_snps (observed + permutations) on
        # a few genes of the slice, both models, fewer permutations
        order = np.argsort(cg, kind="stable")
        small = [int(k) for k in order[max(0, len(order) // 2 - 1):][:3]]  # median window sizes
        bc = 8
        tl = tn = 0.0
        nl = nn = 0
        per = CIS_M // CIS_NCHROM
        for k in small:
            first = int((win[0][k] - 1) * per + (win[1][k] - 1 + 99) // 100)
            rows = ctx.copy_rows(rows_dev, first, int(cg[k]), bps)
            dos, _, keep, _ = O.decode_bed(rows, CIS_N)
            W = dos[keep]
            t1 = time.perf_counter()
            r_lm = O.cis_gene("lm", Y[:, k], X, W, B=bc, seed=seed, gene=names[k])
            tl += time.perf_counter() - t1
            nl += r_lm["n_tested"] * (bc + 1)
            t1 = time.perf_counter()
            r_lmm = O.cis_gene("lmm", Y[:, k], X, W, B=bc, seed=seed, gene=names[k], Q=Q, lam=lam)
            tn += time.perf_counter() - t1
            nn += r_lmm["n_tested"] * (bc + 1)
        cpu = {"value": (nl + nn) / (tl + tn), "unit": "cis tests/s (observed + permuted)", "cores": 1,
               "kind": "port",
               "sample": f"{len(small)} genes of the slice ({', '.join(str(int(cg[k])) for k in small)} window "
                         f"SNPs), both models, B={bc}, through the oracle's restatement of scan_gene_snps "
                         f"(Q^T g per LMM test, REML per permutation) on one thread: LM {nl / tl:.3g} tests/s, "
                         f"LMM {nn / tn:.3g} tests/s"}
        log("cpu baseline", cpu)
    out = {
        "metric": "cis tests/s with gene-level permutations (LM + LMM, observed + permuted)",
        "value": tests / dt, "unit": "cis tests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3 contractions, fp64 REML and epilogue", "data": "synthetic (generated in HBM, seeded)",
        "config": {"workload": f"cis LM+LMM n=1000 samples, 1M SNPs, {args.genes} of 20k genes "
                               f"(windows 50-5000 SNPs), p=16, B={B}",
                   "genes": args.genes, "window_snps": int(cg.sum()), "permutations": B, "delta": "reference search"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS if achieved else None, "traffic": None,
                     "note": "LMM permutations' fused Wald kernel: algorithmic 2n(p+2) FLOP per test over its "
                             "event-timed launches"},
        "cpu_baseline": cpu,
        "per_model": {k: {"seconds": v["s"], "observed_tests": v["obs_tests"], "permuted_tests": v["perm_tests"],
                          "tests_per_s": (v["obs_tests"] + v["perm_tests"]) / v["s"],
                          "kernel_ms": {f: a[0] for f, a in v["fam"].items()},
                          "kernel_launches": {f: a[1] for f, a in v["fam"].items()}} for k, v in res.items()},
    }
    print(json.dumps(out))
    ctx.close()


C2_N, C2_M, C2_G, C2_NCOV, C2_NCHROM, C2_FLANK = 1000, 100_000, 2000, 15, 10, 1_000_000


def bench_cis_lmm(args):
    """Cis LMM scan with every pair reported (--model lmm --mode cis -w 1000000
    --pval-cis 1): n = 1000, 100k SNPs on 10 chromosomes (2 kb apart), 2000
    genes with 1 kb bodies and a 1 Mb flank, polygenic phenotypes
    (h2 ~ U(0, 0.6)) with 1-3 cis eQTLs per gene, 15 covariates + intercept;
    GRM and eigendecomposition in the same process (untimed). One step =
    per-gene REML delta (the reference search) + bases + decode + rotation +
    fused Wald over the windows, all pairs copied to the host."""
    import ctypes
    import eqtl_amd
    import oracle as O
    from eqtl_amd import synth
    from eqtl_amd._lib import lib, _check
    global BLOCK
    BLOCK = args.block
    from eqtl_amd import workloads
    n, m, G = C2_N, C2_M, C2_G
    ctx = eqtl_amd.Context(0)
    w = workloads.c2(ctx)
    rows_dev, bps, chrom, pos, X, Y = w.rows_dev, w.bps, w.chrom, w.pos, w.X, w.Y
    gch, cs, ce = w.win
    per = w.extra["per"]
    ctx.set_samples(n)
    ctx.set_policy(True, 0.8, 0.0)
    _check(lib().eqtl_gpu_grm_begin(ctx.handle))
    for j0 in range(0, m, BLOCK):
        mb = min(BLOCK, m - j0)
        _check(lib().eqtl_gpu_grm_accumulate(ctx.handle, ctypes.c_void_p(rows_dev + j0 * bps), ctypes.c_int32(1),
                                             ctypes.c_int64(mb), ctypes.c_int32(bps)))
    K, _, _ = ctx.grm_finish()
    Q, lam = ctx.eigh(K)
    ctx.set_snp_annot(chrom, pos)
    prec = eqtl_amd.PREC_BF16X3

    def step():
        ctx.lmm_prep(Y, X, delta_mode=eqtl_amd.DELTA_REF)
        ctx.set_gene_windows(gch, cs, ce)
        ctx.scan_begin(eqtl_amd.MODEL_LMM, prec, 1.0, eqtl_amd.SCOPE_CIS)
        nh = 0
        for j0 in range(0, m, BLOCK):
            mb = min(BLOCK, m - j0)
            nh += ctx.scan_block(None, j0, hit_cap=1 << 22, rows_dev=rows_dev + j0 * bps, m=mb, bps=bps).shape[0]
        return ctx.scan_end(), nh

    for _ in range(args.warmup):
        step()
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    tests = 0
    for s_ in range(args.steps):
        st, nh = step()
        tests += int(st["n_tested"].sum())
        log(f"step {s_}: {int(st['n_tested'].sum())} tests, {nh} pairs reported")
    ctx.synchronize()
    dt = time.perf_counter() - t0
    fam = {f: ctx.timing(f) for f in ("wald", "rotate", "reml", "basis", "decode", "rotate_yx")}
    ctx.set_timing(False)
    log(f"step time {dt / args.steps:.3f}s: " + ", ".join(f"{f} {ms / args.steps:.1f} ms" for f, (ms, _) in fam.items()))
    p = X.shape[1]
    w_ms, w_n = fam["wald"]
    r_ms, r_n = fam["rotate"]
    rot_tflops = 2.0 * n * n * m * args.steps / (r_ms * 1e-3) / 1e12 if r_ms > 0 else None
    wald_tflops = 2.0 * n * (p + 2) * tests / (w_ms * 1e-3) / 1e12 if w_ms > 0 else None
    cpu = None
    if not args.no_cpu:
        # the oracle's scan_gene_snps restatement (no permutations) on a few genes
        genes = list(range(0, G, G // 8))[:8]
        t1 = time.perf_counter()
        nt = 0
        for k in genes:
            first = (gch[k] - 1) * per
            lo = first + max(0, (int(cs[k]) - 1 + 1999) // 2000)
            hi = first + min(per - 1, (int(ce[k]) - 1) // 2000)
            rows = ctx.copy_rows(rows_dev, lo, hi - lo + 1, bps)
            dos, _, keep, _ = O.decode_bed(rows, n)
            r = O.cis_gene("lmm", Y[:, k], X, dos[keep], B=0, gene=f"g{k}", Q=Q, lam=lam)
            nt += r["n_tested"]
        dtc = time.perf_counter() - t1
        cpu = {"value": nt / dtc, "unit": "cis LMM tests/s", "cores": 1, "kind": "port",
               "sample": f"{len(genes)} genes ({nt} window SNPs) through the oracle's restatement of "
                         f"scan_gene_snps + test_lmm (REML per gene, Q^T g per test) on one thread, {dtc:.1f} s"}
        log("cpu baseline", cpu)
    out = {
        "metric": "cis LMM tests/s (all pairs reported)",
        "value": tests / dt, "unit": "cis LMM tests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3 contractions, fp64 REML and epilogue", "data": "synthetic (generated in HBM, seeded)",
        "config": {"workload": "cis LMM n=1000 samples x 100k SNPs x 2000 genes, +-1 Mb windows, p=16, 1% missing",
                   "samples": n, "snps": m, "genes": G, "covariates": p, "delta": "reference search"},
        "roofline": {"bound": "mfma", "achieved": rot_tflops, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": rot_tflops / BF16_PEAK_TFLOPS if rot_tflops else None, "traffic": None,
                     "note": "k_rotate: algorithmic 2n^2 FLOP per SNP (Q^T g, once per SNP for all genes) over its "
                             "event-timed launches; the split form issues 3 bf16 products per MAC"},
        "cpu_baseline": cpu,
        "kernel_ms_per_step": {f: ms / args.steps for f, (ms, _) in fam.items()},
        "wald_tflops_algorithmic": wald_tflops,
    }
    print(json.dumps(out))
    ctx.close()


GRM_N, GRM_M, GRM_G = 5000, 1_000_000, 20_000
FP64_PEAK_TFLOPS = 78.6
INT8_PEAK_TOPS = 5000.0


def bench_grm(args):
    """GRM (GCTA form) over n = 5000 samples x 1M SNPs (MAF ~ U(0.01, 0.5), no
    missing calls); one step = decode + classify + the grouped int8 Gram of
    the complete SNPs (held 2-bit rows grouped by minor allele count, expanded
    to int8, exact int32 group Grams added into FP64 with the group weight,
    rank-1 centring in FP64) + the GCTA divisor, K copied to the host. Then
    (untimed by value) the eigendecomposition and the REML adjoint (f,
    df/dlog delta, d2f/dlog delta^2) of 20k intercept-only genes at their
    delta."""
    import ctypes
    import eqtl_amd
    import oracle as O
    from eqtl_amd import synth
    from eqtl_amd._lib import lib, _check
    global BLOCK
    BLOCK = args.block
    n, m = GRM_N, min(args.snps or GRM_M, GRM_M)
    bps = (n + 3) // 4
    ctx = eqtl_amd.Context(0)
    rows_dev = ctx.synth_bed(1005, m, n, 0.01, 0.5, 0.0)
    ctx.set_samples(n)
    ctx.set_policy(True, 0.8, 0.0)

    def grm_pass(mm):
        _check(lib().eqtl_gpu_grm_begin(ctx.handle))
        for j0 in range(0, mm, BLOCK):
            mb = min(BLOCK, mm - j0)
            _check(lib().eqtl_gpu_grm_accumulate(ctx.handle, ctypes.c_void_p(rows_dev + j0 * bps),
                                                 ctypes.c_int32(1), ctypes.c_int64(mb), ctypes.c_int32(bps)))
        return ctx.grm_finish()

    for _ in range(args.warmup):
        grm_pass(min(m, BLOCK))
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        K, Np, used = grm_pass(m)
        log(f"step {s}: {used} SNPs used")
    ctx.synchronize()
    dt = time.perf_counter() - t0
    g_ms, g_n = ctx.timing("grm_i8")
    e_ms, _ = ctx.timing("grm_expand")
    f_ms, f_n = ctx.timing("grm")
    ctx.set_timing(False)
    snps = m * args.steps
    achieved = float(n) * (n + 1) * snps / (g_ms * 1e-3) / 1e12
    log(f"GRM step {dt / args.steps:.3f}s, int8 gram {g_ms / args.steps:.1f} ms in {g_n} launches "
        f"({achieved:.0f} TOP/s algorithmic), expand {e_ms / args.steps:.1f} ms, FP64 gram {f_ms:.1f} ms")
    extra = {"gram_i8_ms_per_step": g_ms / args.steps, "gram_i8_launches_per_step": g_n / args.steps,
             "expand_ms_per_step": e_ms / args.steps, "fp64_gram_ms_per_step": f_ms / args.steps,
             "snps_per_block": BLOCK}
    # eigendecomposition and the REML adjoint at this n
    t1 = time.perf_counter()
    Q, lam = ctx.eigh(K)
    extra["eigh_s"] = time.perf_counter() - t1
    rng = np.random.default_rng(1005)
    causal = np.sort(rng.choice(m, 1000, replace=False))
    Z = np.empty((1000, n))
    for i, j in enumerate(causal):
        Z[i] = synth.standardized(synth.unpack_bed(ctx.copy_rows(rows_dev, int(j), 1, bps), n))[0]
    h2 = rng.uniform(0, 0.8, GRM_G)
    u = Z.T @ rng.standard_normal((1000, GRM_G))
    u = (u - u.mean(0)) / u.std(0)
    Y = rng.standard_normal((n, GRM_G)) + u * np.sqrt(h2 / (1 - h2))
    del u, Z
    X1 = np.ones((n, 1))
    t1 = time.perf_counter()
    delta, _ = ctx.lmm_prep(Y, X1, delta_mode=eqtl_amd.DELTA_NEWTON)
    extra["lmm_prep_s"] = time.perf_counter() - t1
    ld = np.log(delta)
    ctx.reml_grad_hess(ld)
    ctx.synchronize()
    t1 = time.perf_counter()
    for sh in (0.0, -1.0, 1.0):
        ctx.reml_grad_hess(ld + sh)
    t_adj = (time.perf_counter() - t1) / 3
    adj_bytes = 8.0 * n * (1 + 1) * GRM_G
    extra["reml_adjoint"] = {"genes": GRM_G, "ms_per_call": 1e3 * t_adj,
                             "achieved_GBps": adj_bytes / t_adj / 1e9,
                             "note": "host-timed call (delta upload, weights kernel, two DGEMMs, f/g/h read back); "
                                     "algorithmic 8n(1+p) B per gene-evaluation"}
    log(f"eigh {extra['eigh_s']:.2f}s, lmm_prep {extra['lmm_prep_s']:.2f}s, REML adjoint {1e3 * t_adj:.1f} ms/call")
    cpu = None
    if not args.no_cpu:
        ns = 16384
        rows_s = ctx.copy_rows(rows_dev, 0, ns, bps)
        dos, miss, keep, _ = O.decode_bed(rows_s, n)
        threads = O.num_threads()
        t1 = time.perf_counter()
        O.grm(dos, miss, keep)
        dtc = time.perf_counter() - t1
        cpu = {"value": ns / dtc, "unit": "GRM SNPs/s", "cores": threads, "kind": "port",
               "sample": f"{ns} SNPs x {n} samples through the oracle's restatement of compute_grm_from "
                         f"(rank-1 FP64 updates of the lower triangle, {threads} threads, {dtc:.1f} s incl. "
                         f"the GCTA divisor)"}
        log("cpu baseline", cpu)
    out = {
        "metric": "GRM SNPs/s (GCTA form, exact int8 group Grams + FP64 weights)",
        "value": snps / dt, "unit": "GRM SNPs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 (exact int32 group sums), f64 weights and centring", "data": "synthetic (generated in HBM, seeded)",
        "config": {"workload": f"GRM n={n} samples x {m} SNPs, no missing calls; REML adjoint for {GRM_G} "
                               f"intercept-only genes", "samples": n, "snps_per_step": m, "snps_per_block": BLOCK},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                     "frac": achieved / INT8_PEAK_TOPS, "traffic": None,
                     "note": "k_grm_i8: algorithmic n(n+1) int8 ops per SNP (unique lower triangle incl. the "
                             "diagonal) over its event-timed launches, dense int8 matrix peak"},
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out))
    ctx.close()


def main():
    global BLOCK
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--snps", type=int, default=None,
                    help="SNPs per step (default: the workload's full count, 500k for the trans scans)")
    ap.add_argument("--block", type=int, default=BLOCK, help="SNPs per scan block")
    ap.add_argument("--genes", type=int, default=1000, help="cis-perm: genes of the 20k-gene table that are run")
    ap.add_argument("--perm", type=int, default=1000, help="cis-perm: permutations per gene")
    ap.add_argument("--probe-copy", action="store_true",
                    help="first run a 1 GiB device-to-device copy (reference kernel for HBM byte counters)")
    ap.add_argument("--workload", choices=["trans-lmm", "trans-lm", "trans-glm", "cis-lmm", "cis-perm", "grm",
                                           "vcf-gt", "vcf-ds"], default="trans-lmm",
                    help="trans-lmm: the association scan (default); trans-lm: the LM scan; trans-glm: the NB "
                         "score-test scan; cis-lmm: the cis LMM scan, all pairs; cis-perm: cis LM+LMM with gene-level permutations; grm: the GRM at "
                         "n=5000 (+ eigh, REML adjoint); vcf-gt / vcf-ds: the device VCF decode")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.probe_copy:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        log("copy probe: 1 GiB read + 1 GiB written")
        del x, y
    other = {"trans-glm": bench_glm, "trans-lm": bench_trans_lm, "cis-perm": bench_cis_perm, "grm": bench_grm,
             "cis-lmm": bench_cis_lmm,
             "vcf-gt": lambda a: bench_vcf(a, False), "vcf-ds": lambda a: bench_vcf(a, True)}
    if args.workload in other:
        other[args.workload](args)
        return
    import eqtl_amd
    BLOCK = args.block
    m = min(args.snps or M, M)
    ctx = eqtl_amd.Context(0)
    name, ncu, hbm = ctx.device_info()
    log(f"device {name} CUs={ncu} HBM={hbm / 2**30:.0f} GiB")
    t0 = time.perf_counter()
    rows_dev, bps, chrom, pos, X, Y, win, planted = make_inputs(ctx)
    log(f"inputs ready {time.perf_counter() - t0:.1f}s")
    ctx.set_samples(N)
    ctx.set_policy(True, 0.8, 0.0)
    t0 = time.perf_counter()
    ctx.grm_begin_dev = None
    import ctypes
    from eqtl_amd._lib import lib, _check
    _check(lib().eqtl_gpu_grm_begin(ctx.handle))
    for j0 in range(0, M, BLOCK):
        mb = min(BLOCK, M - j0)
        _check(lib().eqtl_gpu_grm_accumulate(ctx.handle, ctypes.c_void_p(rows_dev + j0 * bps), ctypes.c_int32(1),
                                             ctypes.c_int64(mb), ctypes.c_int32(bps)))
    K, _, used = ctx.grm_finish()
    t_grm = time.perf_counter() - t0
    t0 = time.perf_counter()
    Q, lam = ctx.eigh(K)
    t_eigh = time.perf_counter() - t0
    log(f"GRM {t_grm:.2f}s ({used} SNPs)  eigh {t_eigh:.2f}s")
    ctx.set_snp_annot(chrom, pos)
    ctx.set_gene_windows(*win)

    # the phenotypes stay in device memory like the genotype rows (column-major
    # n x G), allocated through the HIP runtime the library runs on
    hip = ctypes.CDLL("libamdhip64.so")
    y_host = np.asfortranarray(Y, dtype=np.float64)
    y_ptr = ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(y_ptr), ctypes.c_size_t(y_host.nbytes)) != 0 or \
            hip.hipMemcpy(y_ptr, y_host.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(y_host.nbytes), 1) != 0:
        raise RuntimeError("device copy of the phenotypes failed")
    Y_dev = (y_ptr.value, Y.shape[1])
    prec = eqtl_amd.PREC_BF16X3
    for w in range(args.warmup):
        trans_pass(ctx, rows_dev, bps, X, Y, min(m, 4 * BLOCK), prec, eqtl_amd.DELTA_NEWTON, Y_dev=Y_dev)
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    tests = 0
    for s in range(args.steps):
        delta, st, nh = trans_pass(ctx, rows_dev, bps, X, Y, m, prec, eqtl_amd.DELTA_NEWTON, Y_dev=Y_dev)
        tests += int(st["n_tested"].sum())
        log(f"step {s}: {int(st['n_tested'].sum())} tests, {nh} hits")
    ctx.synchronize()
    dt = time.perf_counter() - t0
    value = tests / dt
    num_ms, num_n = ctx.timing("num")
    wald_ms, wald_n = ctx.timing("wald")
    tests_ms, tests_n = ctx.timing("tests")
    rot_ms, _ = ctx.timing("rotate")
    dec_ms, _ = ctx.timing("decode")
    reml_ms, _ = ctx.timing("reml")
    basis_ms, _ = ctx.timing("basis")
    log(f"step time {dt / args.steps:.2f}s: num {num_ms / args.steps:.0f} ms + wald {wald_ms / args.steps:.0f} ms + "
        f"tests {tests_ms / args.steps:.0f} ms, "
        f"rotate {rot_ms / args.steps:.0f} ms, decode {dec_ms / args.steps:.0f} ms, reml {reml_ms / args.steps:.0f} ms, "
        f"basis {basis_ms / args.steps:.0f} ms")
    ctx.set_timing(False)
    # Algorithmic work per test of each kernel, over its event-timed launches.
    # Interpolated form, three kernels per block:
    #   k_lr_num   num = (s o r)^T g~ over the samples: 2n FLOP (matrix cores)
    #   k_wald_lr  the p covariate projections and gg interpolated from R
    #              nodes (R averaged over the genes): 2R(p+1) FLOP (matrix cores)
    #   k_lr_tests the FP64 tests over the fold: reads gg, S and num, 12 bytes
    # Eight-wave form (k_wald_lmm, tests included): 2n(p+2) FLOP (num, gg and
    # the p projections contracted over the samples).
    p = X.shape[1]
    lr_on, lr_sg, lr_seg, lr_nodes = ctx.lr_plan()
    dense_flop = 2.0 * N * (p + 2)
    n_launch = max(wald_n, 1)
    tests_per_launch = tests / n_launch
    if lr_on:
        kern = {"k_lr_num": {"ms": num_ms / max(num_n, 1), "bound": "mfma", "work_per_test": 2.0 * N},
                "k_wald_lr": {"ms": wald_ms / n_launch, "bound": "mfma", "work_per_test": 2.0 * lr_nodes * (p + 1)},
                "k_lr_tests": {"ms": tests_ms / max(tests_n, 1), "bound": "hbm", "work_per_test": 12.0}}
    else:
        kern = {"k_wald_lmm": {"ms": wald_ms / n_launch, "bound": "mfma", "work_per_test": dense_flop}}
    for v in kern.values():
        mfma = v["bound"] == "mfma"
        v["unit"] = "TFLOP/s" if mfma else "GB/s"
        v["achieved"] = v["work_per_test"] * tests_per_launch / (v["ms"] * 1e-3) / (1e12 if mfma else 1e9)
        v["peak"] = BF16_PEAK_TFLOPS if mfma else HBM_PEAK_GBS
        v["frac"] = v["achieved"] / v["peak"]
    dom = max(kern, key=lambda k: kern[k]["ms"])
    all_ms = sum(v["ms"] for v in kern.values())
    dense_tflops = dense_flop * tests_per_launch / (all_ms * 1e-3) / 1e12
    extra = {"wald_plan": {"super_groups": lr_sg, "segments": lr_seg, "mean_nodes": lr_nodes,
                           "dense_flop_per_test": dense_flop, "tests_per_launch": tests_per_launch,
                           "kernels": kern}}
    # plain FP32 form on a slice of the same data (REF delta search, as the reference)
    m32 = min(m, 131072)
    ctx.set_timing(True)
    ctx.synchronize()
    t1 = time.perf_counter()
    d32, st32, _ = trans_pass(ctx, rows_dev, bps, X, Y, m32, eqtl_amd.PREC_FP32, eqtl_amd.DELTA_REF, Y_dev=Y_dev)
    ctx.synchronize()
    dt32 = time.perf_counter() - t1
    w32, n32 = ctx.timing("wald")
    ctx.set_timing(False)
    extra["fp32_form"] = {"snps": m32, "tests_per_s": int(st32["n_tested"].sum()) / dt32,
                          "wald_ms_per_launch": w32 / max(n32, 1)}
    if not args.no_accuracy:
        import oracle  # noqa: F401  (checker only)
        extra["accuracy_vs_fp64"] = accuracy_sample(ctx, rows_dev, bps, X, Y, win, delta, planted, prec, Q, lam)
        log("accuracy", extra["accuracy_vs_fp64"])
    cpu = None
    if not args.no_cpu:
        import oracle as O  # noqa: F401  (timed CPU baseline only)
        rows_s = ctx.copy_rows(rows_dev, 0, CPU_SNPS, bps)
        cpu = cpu_baseline(Q, lam, X, Y[:, :CPU_GENES], rows_s, delta[:CPU_GENES])
        log("cpu baseline", cpu)
    out = {
        "metric": "Wald tests/sec trans-LMM (split-BF16 vs FP32 GPU), max p rel err vs FP64",
        "value": value, "unit": "Wald tests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3 contractions, fp64 epilogue", "data": "synthetic (generated in HBM, seeded)",
        "config": {"workload": "trans-LMM n=2000 samples x 500k SNPs x 20k genes, p=21, 1% missing",
                   "snps_per_step": m, "genes": G, "samples": N, "covariates": p, "delta": "batched Newton",
                   "snps_per_block": BLOCK},
        "roofline": {"kernel": dom, "bound": kern[dom]["bound"], "achieved": kern[dom]["achieved"],
                     "peak": kern[dom]["peak"], "unit": kern[dom]["unit"], "frac": kern[dom]["frac"],
                     "traffic": measured_traffic(BLOCK, dom),
                     "note": (f"{dom}, the longest kernel per block; per-test algorithmic work: k_lr_num 2n FLOP "
                              f"({kern['k_lr_num']['frac']:.3f} of bf16 dense peak), k_wald_lr 2R(p+1) FLOP "
                              f"({kern['k_wald_lr']['frac']:.3f}), k_lr_tests 12 B read (gg, S, num; "
                              f"{kern['k_lr_tests']['frac']:.3f} of HBM peak: its FP64 tail, ACAT term and "
                              "near-null refinement bound it by FP64 instruction issue, not by HBM); the "
                              "contraction kernels issue 3 bf16 MFMA products per num MAC and 6 per interpolation "
                              f"MAC; dense-equivalent 2n(p+2) FLOP/test over the three kernels: {dense_tflops:.0f} "
                              "TFLOP/s")
                     if lr_on else "algorithmic 2n(p+2) FLOP/test over the fused Wald kernel's event-timed launches; "
                                   "the split form issues 3 bf16 MFMA products per algorithmic MAC"},
        "cpu_baseline": cpu,
        "setup_s": {"grm": t_grm, "eigh": t_eigh},
        **extra,
    }
    print(json.dumps(out))
    ctx.close()
    hip.hipFree(y_ptr)


if __name__ == "__main__":
    main()
