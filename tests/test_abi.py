"""CPU: the C-ABI library loads and exports every entry point include/eqtl_gpu.h
declares (no compute calls: there is no GPU here); the product fails loudly
without its library; host-side numerics shared with the kernels."""
import ctypes
import os
import re
import subprocess

import pytest

import eqtl_amd
from eqtl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "eqtl_gpu.h")).read()
    return sorted(set(re.findall(r"\b(eqtl_gpu_[a-z_0-9]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libeqtl_gpu.so not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for s in declared_symbols():
        assert re.search(rf"\bT {s}\b", out), s


def test_fails_loudly_without_library(monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libeqtl_gpu.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(eqtl_amd.EqtlGpuError):
        eqtl_amd.Context(0)


def test_acat_x87_emulation_host(tmp_path):
    """The kernels' ACAT term/final (numerics.hpp) against the platform's long double."""
    exe = tmp_path / "acat_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "acat_emul_check.cpp"), "-o",
                           str(exe)])
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr


def test_permutation_stream_matches_libstdcxx(tmp_path):
    """Device MT19937 + shuffle (perm_rng.hpp) against std::mt19937 + std::shuffle."""
    exe = tmp_path / "perm_check"
    subprocess.check_call(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "perm_rng_check.cpp"), "-o",
                           str(exe)])
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
