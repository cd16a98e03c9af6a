// Record-based level-scheduled DILU recurrences (unstructured meshes).
#pragma once

#include "fvh_device.h"

struct fvh_eqn;

namespace fvh {

// the record path: at most kLevDeps lower / upper neighbours per row, and not
// switched off with FVH_LEVEL_RECORDS=0
bool lev_path(const fvh_mesh *m);
// update_dilu_diagonal! (D, Dinv) and the sweep records from the values of A
int lev_dilu_update(fvh_eqn *e, const double *val);
int lev_dilu_apply(fvh_eqn *e, const double *x, double *y, const KScal *ks);
// FP32 copies of the sweep records (after each update) and the FP32 sweeps
int lev_to32(fvh_eqn *e);
int lev_dilu_apply32(fvh_eqn *e, const float *x, float *y, const KScal *ks);
// a wait of the sync-free walk ran out since the last check (reported once)
int lev_check(fvh_eqn *e);

}  // namespace fvh
