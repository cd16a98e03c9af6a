// Triangular recurrences of the DILU and ILU(0) preconditioners.
#pragma once

#include "fvh_device.h"

namespace fvh {

// update_dilu_diagonal!: D (full) and Dinv = 1/D from the values of A
int dilu_update(fvh_mesh *m, const double *val, double *D, double *Dinv);
// ldiv!(y, DILU, x): forward then backward sweep
int dilu_apply(fvh_mesh *m, const double *val, const double *Dinv, const double *x, double *y, const KScal *ks);
// ILU(0) factor on A's pattern and its two sweeps
int ilu0_factor(fvh_mesh *m, const double *val, double *lu);
int ilu0_apply(fvh_mesh *m, const double *lu, const double *x, double *y, const KScal *ks);
// the same sweeps in FP32 (FP32 copies of A / Dinv / LU)
int dilu_apply32(fvh_mesh *m, const float *val, const float *Dinv, const float *x, float *y, const KScal *ks);
int ilu0_apply32(fvh_mesh *m, const float *lu, const float *x, float *y, const KScal *ks);

}  // namespace fvh
