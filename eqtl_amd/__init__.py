"""eqtl_amd — MI355X (gfx950) association engine for eqtl (trans LM/LMM and
NB GLM scans, REML, GRM, cis permutations, VCF decode) behind a C ABI
(include/eqtl_gpu.h).

The compute path is libeqtl_gpu.so (hand-written HIP kernels); this package
holds its ctypes binding, the in-tree build, and the synthetic-input generator.
"""
from ._lib import (Context, EqtlGpuError, HIT_DTYPE, GENE_STATE_DTYPE, MODEL_LM, MODEL_LMM, MODEL_GLM, PREC_FP32,
                   PREC_BF16X3, DELTA_REF, DELTA_NEWTON, DELTA_GIVEN, SCOPE_TRANS, SCOPE_CIS, LIB_PATH)

__all__ = ["Context", "EqtlGpuError", "HIT_DTYPE", "GENE_STATE_DTYPE", "MODEL_LM", "MODEL_LMM", "MODEL_GLM", "PREC_FP32",
           "PREC_BF16X3", "DELTA_REF", "DELTA_NEWTON", "DELTA_GIVEN", "SCOPE_TRANS", "SCOPE_CIS", "LIB_PATH"]
