"""The reference-side binding (eqtl_amd/binding/scan_trans_gpu.cpp): the
reference's scan_lm_snp_outer / scan_lmm_snp_outer templates implemented over
the C ABI, compiled as a real translation unit against the reference's own
headers and vendored Eigen (syntax and semantic analysis, with the explicit
instantiations for PlinkBed and VcfSession). Skipped where the reference tree
is absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "include", "eqtl")), reason="reference tree absent")
def test_binding_compiles_against_reference_headers():
    src = os.path.join(ROOT, "eqtl_amd", "binding", "scan_trans_gpu.cpp")
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-fopenmp", "-Wall", "-Werror=return-type",
           "-I" + os.path.join(REF, "include"), "-I" + os.path.join(REF, "third_party", "eigen"),
           "-I" + os.path.join(ROOT, "include"), src]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
