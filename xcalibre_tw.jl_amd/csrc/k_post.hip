// Post-processing on a boundary patch: surface flux of a cell field, the
// wall-parallel normal gradient of a velocity field, pressure and viscous
// forces and the wall shear stress (Postprocess/Postprocess_forces.jl,
// Calculate/Calculate_5_surface_normal_gradient.jl).
//
// One thread per patch face fills the per-face arrays; the force sums run
// over the faces in patch order, one after another, as the reference sums
// them, so the totals carry the reference's rounding.
#include "k_calc.h"

namespace fvh {

// surface_flux(snflux, facesID, cellsID, phi): phi[cell] * (area * normal)
__global__ void k_surface_flux(DevMesh m, int b0, int nb, const double *__restrict__ phi, double *__restrict__ ox,
                               double *__restrict__ oy, double *__restrict__ oz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int f = m.bf_face[b0 + i], c = m.bf_cell[b0 + i];
    const double a = m.farea[f], v = phi[c];
    ox[i] = v * (a * m.nx[f]);
    oy[i] = v * (a * m.ny[f]);
    oz[i] = v * (a * m.nz[f]);
}

// surface_normal_gradient(snGrad, facesID, cellsID, U, Uw): the part of
// U[cell] - Uw parallel to the face, over the face's delta
__global__ void k_surface_normal_gradient(DevMesh m, int b0, int nb, const double *__restrict__ ux,
                                          const double *__restrict__ uy, const double *__restrict__ uz, double wx,
                                          double wy, double wz, double *__restrict__ ox, double *__restrict__ oy,
                                          double *__restrict__ oz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int f = m.bf_face[b0 + i], c = m.bf_cell[b0 + i];
    const double dx = ux[c] - wx, dy = uy[c] - wy, dz = uz[c] - wz;
    const double nx = m.nx[f], ny = m.ny[f], nz = m.nz[f];
    const double dn = (dx * nx + dy * ny) + dz * nz;
    const double delta = m.fdelta[f];
    ox[i] = (dx - dn * nx) / delta;
    oy[i] = (dy - dn * ny) / delta;
    oz[i] = (dz - dn * nz) / delta;
}

// Sequential sums over the patch faces (one thread), as the reference's
// loops: kind 0 = sum of the arrays (pressure force), kind 1 = sum of
// array * area * (nu + nut[cell]) (viscous force); out = rho .* sums.
__global__ void k_patch_sum(DevMesh m, int b0, int nb, int kind, const double *__restrict__ x,
                            const double *__restrict__ y, const double *__restrict__ z, double nu,
                            const double *__restrict__ nut, double rho, double *__restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < nb; ++i) {
        if (kind == 0) {
            sx += x[i];
            sy += y[i];
            sz += z[i];
        } else {
            const int f = m.bf_face[b0 + i], c = m.bf_cell[b0 + i];
            const double a = m.farea[f], nue = nu + (nut ? nut[c] : 0.0);
            sx += x[i] * a * nue;
            sy += y[i] * a * nue;
            sz += z[i] * a * nue;
        }
    }
    out[0] = rho * sx;
    out[1] = rho * sy;
    out[2] = rho * sz;
}

// wall_shear_stress: snGrad * (nu + nut[cell]) per face, and face centres
__global__ void k_wall_shear(DevMesh m, int b0, int nb, double nu, const double *__restrict__ nut,
                             double *__restrict__ tx, double *__restrict__ ty, double *__restrict__ tz,
                             double *__restrict__ pos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int f = m.bf_face[b0 + i], c = m.bf_cell[b0 + i];
    const double nue = nu + (nut ? nut[c] : 0.0);
    tx[i] *= nue;
    ty[i] *= nue;
    tz[i] *= nue;
    if (pos) {
        pos[3 * i] = m.fcx[f];
        pos[3 * i + 1] = m.fcy[f];
        pos[3 * i + 2] = m.fcz[f];
    }
}

static int patch_range(const fvh_mesh *m, int32_t patch, int *b0, int *nb) {
    if (!m) return set_error(FVH_ERR_ARG, "null mesh");
    if (patch < 0 || patch >= m->nboundaries)
        return set_error(FVH_ERR_ARG, "patch %d of %d", patch, m->nboundaries);
    *b0 = m->h_bstart[patch];
    *nb = m->h_bstart[patch + 1] - m->h_bstart[patch];
    return FVH_OK;
}

}  // namespace fvh

using namespace fvh;

extern "C" int fvh_patch_size(const fvh_mesh *m, int32_t patch, int64_t *nfaces) {
    int b0 = 0, nb = 0;
    FVH_TRY(patch_range(m, patch, &b0, &nb));
    if (nfaces) *nfaces = nb;
    return FVH_OK;
}

extern "C" int fvh_surface_flux(const fvh_mesh *m, int32_t patch, const double *phi, double *fx, double *fy,
                                double *fz) {
    int b0 = 0, nb = 0;
    FVH_TRY(patch_range(m, patch, &b0, &nb));
    if (!phi || (nb > 0 && (!fx || !fy || !fz))) return set_error(FVH_ERR_ARG, "null argument");
    if (nb > 0) {
        hipLaunchKernelGGL(k_surface_flux, grid1d(nb), dim3(256), 0, stream(), m->d, b0, nb, phi, fx, fy, fz);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

extern "C" int fvh_surface_normal_gradient(const fvh_mesh *m, int32_t patch, const double *ux, const double *uy,
                                           const double *uz, const double *uw, double *gx, double *gy, double *gz) {
    int b0 = 0, nb = 0;
    FVH_TRY(patch_range(m, patch, &b0, &nb));
    if (!ux || !uy || !uz || !uw || (nb > 0 && (!gx || !gy || !gz))) return set_error(FVH_ERR_ARG, "null argument");
    if (nb > 0) {
        hipLaunchKernelGGL(k_surface_normal_gradient, grid1d(nb), dim3(256), 0, stream(), m->d, b0, nb, ux, uy, uz,
                           uw[0], uw[1], uw[2], gx, gy, gz);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

// Per-face scratch of one patch (three arrays) and the three sums.
static int patch_scratch(int nb, DBuf<double> &buf) { return buf.alloc(3 * (size_t)std::max(nb, 1) + 3); }

extern "C" int fvh_pressure_force(const fvh_mesh *m, int32_t patch, const double *p, double rho, double *force) {
    FVH_GUARD_BEGIN
    int b0 = 0, nb = 0;
    FVH_TRY(patch_range(m, patch, &b0, &nb));
    if (!p || !force) return set_error(FVH_ERR_ARG, "null argument");
    DBuf<double> s;
    FVH_TRY(patch_scratch(nb, s));
    const size_t w = std::max(nb, 1);
    double *x = s.p, *y = s.p + w, *z = s.p + 2 * w, *out = s.p + 3 * w;
    if (nb > 0) {
        hipLaunchKernelGGL(k_surface_flux, grid1d(nb), dim3(256), 0, stream(), m->d, b0, nb, p, x, y, z);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_patch_sum, dim3(1), dim3(64), 0, stream(), m->d, b0, nb, 0, x, y, z, 0.0, nullptr, rho, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(force, out, 3 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_viscous_force(const fvh_mesh *m, int32_t patch, const double *ux, const double *uy,
                                 const double *uz, const double *uw, double rho, double nu, const double *nut,
                                 double *force) {
    FVH_GUARD_BEGIN
    int b0 = 0, nb = 0;
    FVH_TRY(patch_range(m, patch, &b0, &nb));
    if (!ux || !uy || !uz || !force) return set_error(FVH_ERR_ARG, "null argument");
    DBuf<double> s;
    FVH_TRY(patch_scratch(nb, s));
    const size_t w = std::max(nb, 1);
    double *x = s.p, *y = s.p + w, *z = s.p + 2 * w, *out = s.p + 3 * w;
    if (nb > 0) {
        if (uw) {
            hipLaunchKernelGGL(k_surface_normal_gradient, grid1d(nb), dim3(256), 0, stream(), m->d, b0, nb, ux, uy,
                               uz, uw[0], uw[1], uw[2], x, y, z);
            HIP_TRY(hipGetLastError());
        } else {  // no velocity BC names the patch: the gradient stays zero
            HIP_TRY(hipMemsetAsync(s.p, 0, 3 * w * sizeof(double), stream()));
        }
    }
    hipLaunchKernelGGL(k_patch_sum, dim3(1), dim3(64), 0, stream(), m->d, b0, nb, 1, x, y, z, nu, nut, rho, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(force, out, 3 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_wall_shear_stress(const fvh_mesh *m, int32_t patch, const double *ux, const double *uy,
                                     const double *uz, const double *uw, double nu, const double *nut, double *tx,
                                     double *ty, double *tz, double *pos) {
    int b0 = 0, nb = 0;
    FVH_TRY(patch_range(m, patch, &b0, &nb));
    if (!ux || !uy || !uz || !uw || (nb > 0 && (!tx || !ty || !tz))) return set_error(FVH_ERR_ARG, "null argument");
    if (nb > 0) {
        hipLaunchKernelGGL(k_surface_normal_gradient, grid1d(nb), dim3(256), 0, stream(), m->d, b0, nb, ux, uy, uz,
                           uw[0], uw[1], uw[2], tx, ty, tz);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_wall_shear, grid1d(nb), dim3(256), 0, stream(), m->d, b0, nb, nu, nut, tx, ty, tz, pos);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}
