// Genotype rows (EQTL_GENO_*): the ALT dosage of file column col, or missing.
//   BED    PLINK-1 LUT {00→2, 01→missing, 10→1, 11→0} (plink_bed.cpp:221-224)
//   DOSE32 float dosage, non-finite = missing (rows written by k_vcf_decode)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/eqtl_gpu.h"

namespace eqtl_amd {

template <int FMT>
__device__ inline double geno_at(const uint8_t* row, int col, bool& miss) {
  if constexpr (FMT == EQTL_GENO_BED) {
    const int code = (row[col >> 2] >> ((col & 3) * 2)) & 3;
    miss = code == 1;
    return code == 0 ? 2.0 : (code == 2 ? 1.0 : 0.0);
  } else {
    const float v = reinterpret_cast<const float*>(row)[col];
    miss = !isfinite(v);
    return miss ? 0.0 : static_cast<double>(v);
  }
}

}  // namespace eqtl_amd
