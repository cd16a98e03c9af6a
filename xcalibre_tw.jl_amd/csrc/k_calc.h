// Declarations shared between the kernel translation units.
#pragma once

#include "fvh_device.h"
#include "fvh_mathfn.h"

namespace fvh {

constexpr int kMaxBc = 16;

// Boundary-condition tuple of one field, one component, by value in kernel args.
struct BcVals {
    int nbc;
    int kind[kMaxBc];
    int patch[kMaxBc];
    double value[kMaxBc];
};

int fill_bcvals(const fvh_bc *bcs, int nbc, int comp, BcVals &out);

int launch_interpolate(const fvh_mesh *m, const double *phi, double *phif);
int launch_interpolate_vec(const fvh_mesh *m, const double *ux, const double *uy, double *fx, double *fy);
int launch_interpolate_midpoint_vec(const fvh_mesh *m, const double *ux, const double *uy, const double *uz, double *fx,
                                    double *fy, double *fz);
int launch_correct_boundaries(const fvh_mesh *m, const BcVals &bc, const double *phi, double *phif);
int launch_correct_boundaries_vec(const fvh_mesh *m, const fvh_bc *bcs, int nbc, const double *ux, const double *uy,
                                  const double *uz, double *fx, double *fy, double *fz);
int launch_green_gauss(const fvh_mesh *m, const double *phif, double *dx, double *dy, double *dz);
int launch_correct_interpolation(const fvh_mesh *m, const double *dx, const double *dy, const double *dz,
                                 const double *phi, double *phif);
int launch_grad(const fvh_mesh *m, int scheme, const BcVals &bc, const double *phi, double *phif, double *dx, double *dy,
                double *dz);
int launch_div(const fvh_mesh *m, const double *fx, const double *fy, const double *fz, double *out);
int launch_interpolate_gradient(const fvh_mesh *m, const double *gx, const double *gy, const double *gz,
                                const double *phi, double *fx, double *fy, double *fz);
int launch_nonorthogonal_flux(const fvh_mesh *m, const double *fx, const double *fy, const double *fz, double *phif);
int launch_nonorthogonal_correct(const fvh_mesh *m, int sign, const double *J, const double *corr, double *b);
int launch_flux(const fvh_mesh *m, const double *fx, const double *fy, const double *fz, double *out);

inline dim3 grid1d(int64_t n, int block = 256) {
    int64_t g = (n + block - 1) / block;
    return dim3((unsigned)(g > 0 ? g : 1));
}

// Julia's max(x, 0.0)
__device__ inline double jmax0(double x) { return fvh_jmax(x, 0.0); }

}  // namespace fvh
