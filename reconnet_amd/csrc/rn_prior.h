// Isotropic TV separable surrogate (PriorFunctionSolver/TVSQS.cu).
#pragma once

#include "rn_common.h"

namespace rn {

// img, s1, s2, var: device [nb][nx][ny][nz][nc] (nz ignored, taken as 1, for dims == 2)
void tv_iso_sqs(const float* img, float* s1, float* s2, float* var, int nb, int nx, int ny, int nz, int nc, int dims,
                float eps, hipStream_t st);

} // namespace rn
