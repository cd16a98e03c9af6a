// Banded-grid DILU recurrences (anti-diagonal schedule, one thread per grid row).
#pragma once

#include "fvh_device.h"

struct fvh_eqn;

namespace fvh {

int band_setup(fvh_mesh *m);
// the DILU of this mesh runs on the banded (anti-diagonal) kernels
inline bool band_path(const fvh_mesh *m) {
    return m->cases ? m->cases->band : m->sched.band_nx > 0;
}
// per-case geometry of a batch of full-rectangle cases
int band_cases_setup(fvh_mesh *m);
int band_alloc(fvh_eqn *e);
int band_dilu_update(fvh_eqn *e, const double *val);
int band_dilu_apply(fvh_eqn *e, const double *x, double *y, const KScal *ks);
// FP32 records from the FP64 ones (after each update), and the FP32 sweeps
int band_to32(fvh_eqn *e);
int band_dilu_apply32(fvh_eqn *e, const float *x, float *y, const KScal *ks);
int band_dinv_natural(fvh_eqn *e, double *out);
// grid rows per lane the sweeps use (tuned at the first update)
int band_rows_per_lane(const fvh_mesh *m);

// Pipelined band sweeps (skewed layout, waves handing over through LDS and
// device-coherent channels); used for banded meshes unless FVH_BAND_ROWS
// forces one of the single-work-group variants.
struct SkewRect {
    int o, nx, ny;  // cells o + i + j*nx, i < nx, j < ny
};
int skew_setup(fvh_mesh *m, const std::vector<SkewRect> &rects);
bool skew_active(const fvh_mesh *m);
int skew_update(fvh_eqn *e, const double *val);
int skew_copy(fvh_eqn *e, const fvh_eqn *src);
bool skew_tuned32(const fvh_mesh *m);  // the FP32 sweep width is chosen (tuned or forced)
int skew_apply(fvh_eqn *e, const double *x, double *y, const KScal *ks);
int skew_to32(fvh_eqn *e);
int skew_apply32(fvh_eqn *e, const float *x, float *y, const KScal *ks);
// the apply without its last step: z = M^-1 x left in the sweeps' quads
// (*zq), where natural row r's value is word skew_quad_map(m)[r]; x null:
// x was written into the sweep layout by the caller (skew_rhs_target)
int skew_apply_quads(fvh_eqn *e, const double *x, const KScal *ks, const double **zq);
int skew_apply_quads32(fvh_eqn *e, const float *x, const KScal *ks, const float **zq);
const int *skew_quad_map(const fvh_mesh *m);
// What a caller writes in place of the gather (FP64 or FP32 sweeps): natural
// row r's value into word map[r] * stride of dst, and the hand-over arrays
// reset (word q: 0 where q % HS == 0, else the sentinel, whose bits are the
// low sizeof(T) bytes of `sent`)
struct SkewRhs {
    void *dst;
    const int *map;
    int stride;
    void *hand;
    int nhand, HS;
    unsigned long long sent;
};
int skew_rhs_target(fvh_eqn *e, int fp32, SkewRhs *t);
// the reset of the hand-over arrays alone (the gather's other half)
int skew_hand_reset(fvh_eqn *e, int fp32);
int skew_dinv_natural(fvh_eqn *e, double *out);
// reports (once) a sweep whose hand-over wait ran out
int skew_check(fvh_eqn *e);

}  // namespace fvh
