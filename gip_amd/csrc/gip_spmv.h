// SpMV launchers shared by the PCG and batch translation units.
#pragma once

#include "gip_internal.h"

namespace gip {

// y_i (= mode 0 | a*sum + b*y_i mode 1) over non-empty rows; scatter maps row -> output index.
int spmv_launch(gip_ctx *c, const gip_mat *A, const double *x, double *y, double a, double b, int mode,
                const int32_t *scatter);
// FP32 y = A32 x (every row written).
int spmv32_launch(gip_ctx *c, const gip_mat *A, const float *x, float *y);

}  // namespace gip
