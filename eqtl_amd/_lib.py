"""ctypes binding of the C ABI in include/eqtl_gpu.h (libeqtl_gpu.so).

The shared library is built in-tree by eqtl_amd.build (hipcc, gfx950). There is
no fallback: if the library is missing or cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeqtl_gpu.so")

EQTL_OK = 0
EQTL_E_CAPACITY = -6
MODEL_LM, MODEL_LMM, MODEL_GLM = 0, 1, 2
PREC_FP32, PREC_BF16X3 = 0, 1
DELTA_REF, DELTA_NEWTON, DELTA_GIVEN = 0, 1, 2
SCOPE_TRANS, SCOPE_CIS = 0, 1
GENO_BED, GENO_DOSE32 = 0, 1
EQTL_E_FORMAT = -7

HIT_DTYPE = np.dtype([("snp", "i8"), ("gene", "i4"), ("pad", "i4"), ("af", "f8"), ("beta", "f8"),
                      ("se", "f8"), ("stat", "f8"), ("p", "f8"), ("r2", "f8")])
VCF_RECORD_DTYPE = np.dtype([("offset", "i8"), ("length", "i4"), ("gt_key", "i2"), ("ds_key", "i2"),
                             ("haploid", "i4"), ("pad", "i4")])
GENE_STATE_DTYPE = np.dtype([("acat_p", "f8"), ("best_p", "f8"), ("best_snp", "i8"), ("best_af", "f8"),
                             ("best_beta", "f8"), ("best_se", "f8"), ("best_stat", "f8"), ("best_r2", "f8"),
                             ("n_tested", "i4"), ("n_sig", "i4")])

# every entry point declared in include/eqtl_gpu.h
EXPORTS = [
    "eqtl_gpu_open", "eqtl_gpu_close", "eqtl_gpu_last_error", "eqtl_gpu_device_info",
    "eqtl_gpu_set_samples", "eqtl_gpu_set_policy", "eqtl_gpu_set_snp_annot", "eqtl_gpu_decode_bed",
    "eqtl_gpu_grm_begin", "eqtl_gpu_grm_accumulate", "eqtl_gpu_grm_finish", "eqtl_gpu_eigh",
    "eqtl_gpu_set_basis", "eqtl_gpu_lmm_prep", "eqtl_gpu_reml_grad_hess", "eqtl_gpu_lm_prep",
    "eqtl_gpu_set_gene_windows", "eqtl_gpu_scan_begin", "eqtl_gpu_scan_block", "eqtl_gpu_scan_end",
    "eqtl_gpu_set_timing", "eqtl_gpu_get_timing", "eqtl_gpu_scan_lr_plan", "eqtl_gpu_synth_bed", "eqtl_gpu_copy_rows",
    "eqtl_gpu_synchronize", "eqtl_gpu_selftest", "eqtl_gpu_upload_bed", "eqtl_gpu_cis_perm",
    "eqtl_gpu_beta_approx", "eqtl_gpu_set_geno_format", "eqtl_gpu_vcf_decode", "eqtl_gpu_glm_prep",
    "eqtl_gpu_set_test_subset", "eqtl_gpu_scan_hits", "eqtl_gpu_grm_accumulate_noalloc", "eqtl_gpu_cis_tuning",
    "eqtl_gpu_grm_set_limits",
]

_lib = None


class EqtlGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"eqtl_gpu error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EqtlGpuError(-99, f"{LIB_PATH} not built (run __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.eqtl_gpu_last_error.restype = ctypes.c_char_p
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != EQTL_OK:
        raise EqtlGpuError(rc, lib().eqtl_gpu_last_error().decode())
    return rc


class Context:
    """One GPU context (one HIP stream)."""

    def __init__(self, device: int = 0):
        self._c = ctypes.c_void_p()
        _check(lib().eqtl_gpu_open(ctypes.c_int32(device), ctypes.byref(self._c)))
        self.n = 0
        self.G = 0

    def close(self):
        if self._c:
            lib().eqtl_gpu_close(self._c)
            self._c = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._c

    def device_info(self):
        name = ctypes.create_string_buffer(128)
        ncu = ctypes.c_int32(0)
        hbm = ctypes.c_int64(0)
        _check(lib().eqtl_gpu_device_info(self._c, name, ctypes.c_int32(128), ctypes.byref(ncu), ctypes.byref(hbm)))
        return name.value.decode(), ncu.value, hbm.value

    def selftest(self):
        m = ctypes.c_int32(-1)
        _check(lib().eqtl_gpu_selftest(self._c, ctypes.byref(m)))
        return m.value

    def synchronize(self):
        _check(lib().eqtl_gpu_synchronize(self._c))

    def set_samples(self, n_file: int, sample_col=None):
        if sample_col is None:
            sample_col = np.arange(n_file, dtype=np.int32)
        sc = np.ascontiguousarray(sample_col, dtype=np.int32)
        _check(lib().eqtl_gpu_set_samples(self._c, ctypes.c_int32(n_file), _p(sc), ctypes.c_int32(sc.size)))
        self.n = sc.size
        self.n_src = sc.size
        self.n_file = n_file
        self.bps = (n_file + 3) // 4

    def set_test_subset(self, keep=None):
        """Restrict tests to analysis samples keep (increasing); None restores all."""
        if keep is None:
            _check(lib().eqtl_gpu_set_test_subset(self._c, None, ctypes.c_int32(0)))
            self.n = self.n_src
            return
        k = np.ascontiguousarray(keep, dtype=np.int32)
        _check(lib().eqtl_gpu_set_test_subset(self._c, _p(k), ctypes.c_int32(k.size)))
        self.n = k.size

    def set_policy(self, impute=True, max_miss=0.8, maf_min=0.0):
        _check(lib().eqtl_gpu_set_policy(self._c, ctypes.c_int32(1 if impute else 0), ctypes.c_double(max_miss),
                                         ctypes.c_double(maf_min)))

    def set_snp_annot(self, chrom, pos):
        ch = np.ascontiguousarray(chrom, dtype=np.int32)
        ps = np.ascontiguousarray(pos, dtype=np.int64)
        _check(lib().eqtl_gpu_set_snp_annot(self._c, _p(ch), _p(ps), ctypes.c_int64(ch.size)))

    def set_geno_format(self, fmt):
        _check(lib().eqtl_gpu_set_geno_format(self._c, ctypes.c_int32(fmt)))
        self.bps = 4 * self.n_file if fmt == GENO_DOSE32 else (self.n_file + 3) // 4

    def vcf_decode(self, text, recs, text_dev=None, text_len=None):
        """Sample columns of VCF records (VCF_RECORD_DTYPE) -> device DOSE32 rows.
        text: host bytes, or text_dev (device pointer, 4-byte aligned, readable 4
        bytes past text_len). Returns the device pointer of the rows (owned by the
        context, valid until the next call)."""
        recs = np.ascontiguousarray(recs, dtype=VCF_RECORD_DTYPE)
        ptr = ctypes.c_void_p()
        if text_dev is not None:
            src, on_dev, n = ctypes.c_void_p(text_dev), 1, text_len
        else:
            src, on_dev, n = ctypes.c_char_p(text), 0, len(text)
        _check(lib().eqtl_gpu_vcf_decode(self._c, src, ctypes.c_int32(on_dev), ctypes.c_int64(n), _p(recs),
                                         ctypes.c_int64(recs.size), ctypes.byref(ptr)))
        return ptr.value

    def decode_bed(self, rows, want_dosage=True):
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        m, bps = rows.shape
        dos = np.empty((m, self.n)) if want_dosage else None
        keep = np.empty(m, dtype=np.uint8)
        af = np.empty(m)
        _check(lib().eqtl_gpu_decode_bed(self._c, _p(rows), ctypes.c_int32(0), ctypes.c_int64(m), ctypes.c_int32(bps),
                                         _p(dos), _p(keep), _p(af)))
        return dos, keep.astype(bool), af

    # ---- GRM ----
    def grm(self, rows, block=4096, noalloc=False):
        """GRM of host bed rows (m, bps) in blocks; noalloc uses the entry point
        that works in the library's fixed device scratch (no allocation)."""
        rows = np.ascontiguousarray(rows, dtype=np.uint8)
        m, bps = rows.shape
        _check(lib().eqtl_gpu_grm_begin(self._c))
        fn = lib().eqtl_gpu_grm_accumulate_noalloc if noalloc else lib().eqtl_gpu_grm_accumulate
        for j0 in range(0, m, block):
            blk = np.ascontiguousarray(rows[j0:j0 + block])
            _check(fn(self._c, _p(blk), ctypes.c_int32(0), ctypes.c_int64(blk.shape[0]), ctypes.c_int32(bps)))
        return self.grm_finish()

    def grm_accumulate_dev(self, rows_dev, m, bps):
        """eqtl_gpu_grm_accumulate on device rows (after grm_begin)."""
        _check(lib().eqtl_gpu_grm_accumulate(self._c, ctypes.c_void_p(rows_dev), ctypes.c_int32(1),
                                             ctypes.c_int64(m), ctypes.c_int32(bps)))

    def grm_accumulate_noalloc_dev(self, rows_dev, m, bps):
        """eqtl_gpu_grm_accumulate_noalloc on device rows (after grm_begin)."""
        _check(lib().eqtl_gpu_grm_accumulate_noalloc(self._c, ctypes.c_void_p(rows_dev), ctypes.c_int32(1),
                                                     ctypes.c_int64(m), ctypes.c_int32(bps)))

    def grm_begin(self):
        _check(lib().eqtl_gpu_grm_begin(self._c))

    def grm_set_limits(self, exact_buffer_bytes=0, chunk_bytes=0):
        """Device memory of the exact-SNP GRM path (0 keeps a value)."""
        _check(lib().eqtl_gpu_grm_set_limits(self._c, ctypes.c_int64(exact_buffer_bytes),
                                             ctypes.c_int64(chunk_bytes)))

    def grm_finish(self):
        n = self.n
        K = np.empty((n, n), order="F")
        N = np.empty(n * (n + 1) // 2, dtype=np.int32)
        used = ctypes.c_int64(0)
        _check(lib().eqtl_gpu_grm_finish(self._c, _p(K), _p(N), ctypes.byref(used)))
        return K, N, used.value

    # ---- basis ----
    def eigh(self, K):
        K = np.asfortranarray(K, dtype=np.float64)
        n = K.shape[0]
        Q = np.empty((n, n), order="F")
        lam = np.empty(n)
        _check(lib().eqtl_gpu_eigh(self._c, _p(K), ctypes.c_int32(n), _p(Q), _p(lam)))
        return Q, lam

    def set_basis(self, Q, lam):
        Q = np.asfortranarray(Q, dtype=np.float64)
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        _check(lib().eqtl_gpu_set_basis(self._c, _p(Q), _p(lam), ctypes.c_int32(lam.size)))

    # ---- genes ----
    def lmm_prep(self, Y, X, delta_mode=DELTA_REF, delta_in=None, Y_dev=None):
        """Y: n x G host array, or None with Y_dev = (device pointer, G) of an
        n x G column-major FP64 matrix already in device memory."""
        X = np.asfortranarray(X, dtype=np.float64)
        if Y_dev is not None:
            yp, G = ctypes.c_void_p(Y_dev[0]), int(Y_dev[1])
        else:
            Y = np.asfortranarray(Y, dtype=np.float64)
            yp, G = _p(Y), Y.shape[1]
        p = X.shape[1]
        din = np.ascontiguousarray(delta_in, dtype=np.float64) if delta_in is not None else None
        dout = np.empty(G)
        rss = np.empty(G)
        _check(lib().eqtl_gpu_lmm_prep(self._c, yp, ctypes.c_int32(G), _p(X), ctypes.c_int32(p),
                                       ctypes.c_int32(delta_mode), _p(din), _p(dout), _p(rss)))
        self.G = G
        return dout, rss

    def reml_grad_hess(self, log_delta):
        ld = np.ascontiguousarray(log_delta, dtype=np.float64)
        f, g, h = np.empty_like(ld), np.empty_like(ld), np.empty_like(ld)
        _check(lib().eqtl_gpu_reml_grad_hess(self._c, _p(ld), ctypes.c_int32(ld.size), _p(f), _p(g), _p(h)))
        return f, g, h

    def lm_prep(self, Y, X):
        Y = np.asfortranarray(Y, dtype=np.float64)
        X = np.asfortranarray(X, dtype=np.float64)
        G, p = Y.shape[1], X.shape[1]
        yty = np.empty(G)
        _check(lib().eqtl_gpu_lm_prep(self._c, _p(Y), ctypes.c_int32(G), _p(X), ctypes.c_int32(p), _p(yty)))
        self.G = G
        return yty

    def glm_prep(self, Y, X):
        """NB GLM null fits of all genes (counts). Returns (phi[G], converged[G])."""
        Y = np.asfortranarray(Y, dtype=np.float64)
        X = np.asfortranarray(X, dtype=np.float64)
        G, p = Y.shape[1], X.shape[1]
        phi = np.empty(G)
        conv = np.empty(G, dtype=np.int32)
        _check(lib().eqtl_gpu_glm_prep(self._c, _p(Y), ctypes.c_int32(G), _p(X), ctypes.c_int32(p), _p(phi),
                                       _p(conv)))
        self.G = G
        return phi, conv.astype(bool)

    def set_gene_windows(self, chrom, cstart, cend):
        ch = np.ascontiguousarray(chrom, dtype=np.int32)
        cs = np.ascontiguousarray(cstart, dtype=np.int64)
        ce = np.ascontiguousarray(cend, dtype=np.int64)
        _check(lib().eqtl_gpu_set_gene_windows(self._c, _p(ch), _p(cs), _p(ce), ctypes.c_int32(ch.size)))

    # ---- scans ----
    def scan_begin(self, model, precision, pthr, scope=SCOPE_TRANS):
        _check(lib().eqtl_gpu_scan_begin(self._c, ctypes.c_int32(model), ctypes.c_int32(precision),
                                         ctypes.c_int32(scope), ctypes.c_double(pthr)))

    def scan_block(self, rows, snp0, hit_cap=1 << 16, rows_dev=None, m=None, bps=None):
        """rows: host (m, bps) uint8 array, or rows_dev (device pointer int) with m, bps."""
        hits = np.empty(max(hit_cap, 1), dtype=HIT_DTYPE)
        nh = ctypes.c_int64(0)
        if rows_dev is not None:
            rc = lib().eqtl_gpu_scan_block(self._c, ctypes.c_void_p(rows_dev), ctypes.c_int32(1), ctypes.c_int64(snp0),
                                           ctypes.c_int64(m), ctypes.c_int32(bps), _p(hits), ctypes.c_int64(hit_cap),
                                           ctypes.byref(nh))
        else:
            rows = np.ascontiguousarray(rows, dtype=np.uint8)
            m, bps = rows.shape
            rc = lib().eqtl_gpu_scan_block(self._c, _p(rows), ctypes.c_int32(0), ctypes.c_int64(snp0), ctypes.c_int64(m),
                                           ctypes.c_int32(bps), _p(hits), ctypes.c_int64(hit_cap), ctypes.byref(nh))
        _check(rc)
        total = nh.value
        if total <= hit_cap:
            return hits[:total]
        # drain the rest of the block's hits
        out = np.empty(total, dtype=HIT_DTYPE)
        out[:hit_cap] = hits[:hit_cap]
        got = ctypes.c_int64(0)
        _check(lib().eqtl_gpu_scan_hits(self._c, ctypes.c_int64(hit_cap), _p(out[hit_cap:]),
                                        ctypes.c_int64(total - hit_cap), ctypes.byref(got)))
        assert got.value == total - hit_cap
        return out

    def scan_end(self):
        st = np.empty(self.G, dtype=GENE_STATE_DTYPE)
        _check(lib().eqtl_gpu_scan_end(self._c, _p(st)))
        return st

    def scan(self, model, precision, rows, pthr=1e-5, block=8192, hit_cap=1 << 18, scope=SCOPE_TRANS):
        """Full scan over host bed rows (m, bps). Returns (hits, gene_state)."""
        self.scan_begin(model, precision, pthr, scope)
        out = []
        for j0 in range(0, rows.shape[0], block):
            out.append(self.scan_block(rows[j0:j0 + block], j0, hit_cap=hit_cap))
        st = self.scan_end()
        hits = np.concatenate(out) if out else np.empty(0, dtype=HIT_DTYPE)
        return hits, st

    # ---- cis permutations ----
    def upload_bed(self, rows=None, rows_dev=None, m=None, bps=None):
        """Resident genotypes: host (m, bps) uint8 rows, or a device pointer with m, bps."""
        if rows_dev is not None:
            _check(lib().eqtl_gpu_upload_bed(self._c, ctypes.c_void_p(rows_dev), ctypes.c_int32(1),
                                             ctypes.c_int64(m), ctypes.c_int32(bps)))
        else:
            rows = np.ascontiguousarray(rows, dtype=np.uint8)
            _check(lib().eqtl_gpu_upload_bed(self._c, _p(rows), ctypes.c_int32(0), ctypes.c_int64(rows.shape[0]),
                                             ctypes.c_int32(rows.shape[1])))

    def cis_perm(self, precision, B, seed, gene_hash, best_p_obs, delta_mode=DELTA_REF):
        """Gene-level permutations after an observed cis scan. Returns (perm_min_p G x B, p_emp G)."""
        G = self.G
        gh = np.ascontiguousarray(gene_hash, dtype=np.uint32)
        bp = np.ascontiguousarray(best_p_obs, dtype=np.float64)
        pm = np.empty((G, B), dtype=np.float64)
        pe = np.empty(G, dtype=np.float64)
        _check(lib().eqtl_gpu_cis_perm(self._c, ctypes.c_int32(precision), ctypes.c_int32(B), ctypes.c_int32(seed),
                                       _p(gh), _p(bp), ctypes.c_int32(delta_mode), _p(pm), _p(pe)))
        return pm, pe

    def cis_tuning(self):
        """(tiles per work item, timed flags) of the cis autotuner for the three window-size classes."""
        t = np.zeros(3, dtype=np.int32)
        d = np.zeros(3, dtype=np.int32)
        _check(lib().eqtl_gpu_cis_tuning(self._c, _p(t), _p(d)))
        return t, d

    def beta_approx(self, perm_min_p, obs_min_p):
        """(p_beta, shape1, shape2) per gene from G x B permutation minima."""
        pm = np.ascontiguousarray(perm_min_p, dtype=np.float64)
        ob = np.ascontiguousarray(obs_min_p, dtype=np.float64)
        G, B = pm.shape
        pb, s1, s2 = np.empty(G), np.empty(G), np.empty(G)
        _check(lib().eqtl_gpu_beta_approx(self._c, _p(pm), _p(ob), ctypes.c_int32(G), ctypes.c_int32(B), _p(pb),
                                          _p(s1), _p(s2)))
        return pb, s1, s2

    # ---- instrumentation ----
    def set_timing(self, on=True):
        _check(lib().eqtl_gpu_set_timing(self._c, ctypes.c_int32(1 if on else 0)))

    def timing(self, family):
        ms = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        _check(lib().eqtl_gpu_get_timing(self._c, family.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def lr_plan(self):
        """Plan of the prepared trans LMM scan: (low_rank, super_groups, segments, mean_nodes)."""
        on, nsg, nseg = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        mn = ctypes.c_double(0)
        _check(lib().eqtl_gpu_scan_lr_plan(self._c, ctypes.byref(on), ctypes.byref(nsg), ctypes.byref(nseg),
                                           ctypes.byref(mn)))
        return bool(on.value), nsg.value, nseg.value, mn.value

    def synth_bed(self, seed, m, n, maf_lo=0.05, maf_hi=0.5, miss=0.01):
        ptr = ctypes.c_void_p()
        _check(lib().eqtl_gpu_synth_bed(self._c, ctypes.c_uint64(seed), ctypes.c_int64(m), ctypes.c_int32(n),
                                        ctypes.c_double(maf_lo), ctypes.c_double(maf_hi), ctypes.c_double(miss),
                                        ctypes.byref(ptr)))
        return ptr.value

    def copy_rows(self, rows_dev, j0, m, bps):
        out = np.empty((m, bps), dtype=np.uint8)
        _check(lib().eqtl_gpu_copy_rows(self._c, ctypes.c_void_p(rows_dev), ctypes.c_int64(j0), ctypes.c_int64(m),
                                        ctypes.c_int32(bps), _p(out)))
        return out
