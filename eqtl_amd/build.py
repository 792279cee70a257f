"""In-tree build of the HIP library (gfx950) — no JIT, no torch extension cache."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "eqtl_gpu.hip")
OUT = os.path.join(HERE, "libeqtl_gpu.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed-fp32 VALU forms stay off (compiler-generated v_pk_* f32 ops)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


HOST_SRC = [os.path.join(HERE, "host", "eqtl_host.cpp"), os.path.join(HERE, "host", "inputs.cpp"),
            os.path.join(HERE, "host", "vcf_source.cpp"), os.path.join(HERE, "host", "main.cpp")]
CLI = os.path.join(HERE, "eqtl")


def build_cli() -> str:
    """The eqtl command-line driver (C++17 host code over libeqtl_gpu.so)."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-pthread", *HOST_SRC, "-L" + HERE, "-leqtl_gpu", "-lz",
           "-Wl,-rpath,$ORIGIN", "-o", CLI + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ failed:\n" + res.stderr[-4000:])
    os.replace(CLI + ".tmp", CLI)
    return CLI


def build_lib(verbose: bool = False) -> str:
    cmd = ["hipcc", *HIPCC_FLAGS, SRC, "-lrocblas", "-lrocsolver", "-o", OUT + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stderr[-4000:])
    if verbose:
        err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
        if err.strip():
            print(err)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build_lib(verbose=True))
    print(build_cli())
