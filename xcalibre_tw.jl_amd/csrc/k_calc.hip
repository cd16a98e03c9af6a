// Calculate: face interpolation, boundary-value correction, Green-Gauss
// gradients (Orthogonal and Midpoint), cell divergence and face fluxes.
// Face kernels are face-parallel over the SoA face arrays (coalesced);
// cell kernels gather their faces in the reference's faces_range order and
// add boundary faces afterwards in ascending face id, as the reference does.
#include "k_calc.h"

namespace fvh {

__global__ void k_interpolate(DevMesh m, const double *__restrict__ phi, double *__restrict__ phif) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    double w = m.fw[f];
    double omw = 1.0 - w;
    phif[f] = w * phi[m.own1[f]] + omw * phi[m.own2[f]];
}

__global__ void k_interpolate_vec(DevMesh m, const double *__restrict__ ux, const double *__restrict__ uy,
                                  double *__restrict__ fx, double *__restrict__ fy) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    double w = m.fw[f];
    int c1 = m.own1[f], c2 = m.own2[f];
    double omw = 1.0 - w;
    fx[f] = w * ux[c1] + omw * ux[c2];
    fy[f] = w * uy[c1] + omw * uy[c2];
}

// interpolate_midpoint! for scalars: 0.5*(phi1 + phi2) on all faces
__global__ void k_interpolate_midpoint(DevMesh m, const double *__restrict__ phi, double *__restrict__ phif) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    phif[f] = 0.5 * (phi[m.own1[f]] + phi[m.own2[f]]);
}

// interpolate_midpoint! for vectors: weight*(x1 + x2) (Calculate_0_gradient.jl:140)
__global__ void k_interpolate_midpoint_vec(DevMesh m, const double *__restrict__ ux, const double *__restrict__ uy,
                                           const double *__restrict__ uz, double *__restrict__ fx,
                                           double *__restrict__ fy, double *__restrict__ fz) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    double w = m.fw[f];
    int c1 = m.own1[f], c2 = m.own2[f];
    fx[f] = w * (ux[c1] + ux[c2]);
    fy[f] = w * (uy[c1] + uy[c2]);
    fz[f] = w * (uz[c1] + uz[c2]);
}

// correct_boundaries!: each boundary face takes the last BC of the tuple
// that names its patch; Dirichlet -> value, every other kind -> owner value.
__global__ void k_correct_boundaries(DevMesh m, BcVals bc, const double *__restrict__ phi, double *__restrict__ phif) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.nbf) return;
    int patch = m.bf_patch[i];
    int hit = -1;
    for (int k = 0; k < bc.nbc; ++k)
        if (bc.patch[k] == patch) hit = k;
    if (hit < 0) return;
    phif[m.bf_face[i]] = (bc.kind[hit] == FVH_BC_DIRICHLET) ? bc.value[hit] : phi[m.bf_cell[i]];
}

// The three components of a vector field at once (correct_boundaries! for
// FaceVectorField, Calculate_2_interpolation.jl:75-108).
__global__ void k_correct_boundaries_vec(DevMesh m, BcVals bx, BcVals by, BcVals bz, const double *__restrict__ ux,
                                         const double *__restrict__ uy, const double *__restrict__ uz,
                                         double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.nbf) return;
    int patch = m.bf_patch[i];
    int hit = -1;
    for (int k = 0; k < bx.nbc; ++k)
        if (bx.patch[k] == patch) hit = k;
    if (hit < 0) return;
    int f = m.bf_face[i], c = m.bf_cell[i];
    bool dir = bx.kind[hit] == FVH_BC_DIRICHLET;
    fx[f] = dir ? bx.value[hit] : ux[c];
    fy[f] = dir ? by.value[hit] : uy[c];
    fz[f] = dir ? bz.value[hit] : uz[c];
}

// green_gauss!(dx, dy, dz, phif)
// A cell's faces are taken kFaceGroup at a time: the group's index loads,
// then its data loads, are issued together (not one dependent pair per
// face), and the terms are added in faces_range order as before.
constexpr int kFaceGroup = 4;

__global__ void k_green_gauss(DevMesh m, const double *__restrict__ phif, double *__restrict__ dx,
                              double *__restrict__ dy, double *__restrict__ dz) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    double rx = 0.0, ry = 0.0, rz = 0.0;
    const int i1 = m.cf_ptr[c + 1];
    for (int i = m.cf_ptr[c]; i < i1; i += kFaceGroup) {
        int f[kFaceGroup];
        double ns[kFaceGroup], v[kFaceGroup], sx[kFaceGroup], sy[kFaceGroup], sz[kFaceGroup];
#pragma unroll
        for (int k = 0; k < kFaceGroup; ++k)
            if (i + k < i1) {
                f[k] = m.cf_face[i + k];
                ns[k] = m.cf_sign[i + k];
            }
#pragma unroll
        for (int k = 0; k < kFaceGroup; ++k)
            if (i + k < i1) {
                v[k] = phif[f[k]];
                sx[k] = m.Sfx[f[k]];
                sy[k] = m.Sfy[f[k]];
                sz[k] = m.Sfz[f[k]];
            }
#pragma unroll
        for (int k = 0; k < kFaceGroup; ++k)
            if (i + k < i1) {
                rx += v[k] * (sx[k] * ns[k]);
                ry += v[k] * (sy[k] * ns[k]);
                rz += v[k] * (sz[k] * ns[k]);
            }
    }
    double vol = m.vol[c];
    rx = rx / vol;
    ry = ry / vol;
    rz = rz / vol;
    for (int i = m.cb_ptr[c]; i < m.cb_ptr[c + 1]; ++i) {
        int f = m.cb_face[i];
        double v = phif[f];
        rx += v * m.Sfx[f] / vol;
        ry += v * m.Sfy[f] / vol;
        rz += v * m.Sfz[f] / vol;
    }
    dx[c] = rx;
    dy[c] = ry;
    dz[c] = rz;
}

// correct_interpolation! on internal faces (Calculate_0_gradient.jl:146-173)
__global__ void k_correct_interpolation(DevMesh m, const double *__restrict__ dx, const double *__restrict__ dy,
                                        const double *__restrict__ dz, const double *__restrict__ phi,
                                        double *__restrict__ phif) {
    int f = m.nbf + blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    const double w = 0.5;
    int o1 = m.own1[f], o2 = m.own2[f];
    double phifp = w * (phi[o1] + phi[o2]);
    double gx = w * (dx[o1] + dx[o2]), gy = w * (dy[o1] + dy[o2]), gz = w * (dz[o1] + dz[o2]);
    double rx = m.fcx[f] - w * (m.ccx[o1] + m.ccx[o2]);
    double ry = m.fcy[f] - w * (m.ccy[o1] + m.ccy[o2]);
    double rz = m.fcz[f] - w * (m.ccz[o1] + m.ccz[o2]);
    phif[f] = phifp + ((gx * rx + gy * ry) + gz * rz);
}

// div!(phi, psif): per-term division by the volume (Calculate_0_divergence.jl:94,105)
__global__ void k_div(DevMesh m, const double *__restrict__ fx, const double *__restrict__ fy,
                      const double *__restrict__ fz, double *__restrict__ out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    double vol = m.vol[c];
    double acc = 0.0;
    const int i1 = m.cf_ptr[c + 1];
    for (int i = m.cf_ptr[c]; i < i1; i += kFaceGroup) {
        int f[kFaceGroup];
        double ns[kFaceGroup], d[kFaceGroup];
#pragma unroll
        for (int k = 0; k < kFaceGroup; ++k)
            if (i + k < i1) {
                f[k] = m.cf_face[i + k];
                ns[k] = m.cf_sign[i + k];
            }
#pragma unroll
        for (int k = 0; k < kFaceGroup; ++k)
            if (i + k < i1) d[k] = (fx[f[k]] * m.Sfx[f[k]] + fy[f[k]] * m.Sfy[f[k]]) + fz[f[k]] * m.Sfz[f[k]];
#pragma unroll
        for (int k = 0; k < kFaceGroup; ++k)
            if (i + k < i1) acc += d[k] * ns[k] / vol;
    }
    for (int i = m.cb_ptr[c]; i < m.cb_ptr[c + 1]; ++i) {
        int f = m.cb_face[i];
        double d = (fx[f] * m.Sfx[f] + fy[f] * m.Sfy[f]) + fz[f] * m.Sfz[f];
        acc += d / vol;
    }
    out[c] = acc;
}

// flux!(phif, psif) = psif . (area*normal)
__global__ void k_flux(DevMesh m, const double *__restrict__ fx, const double *__restrict__ fy,
                       const double *__restrict__ fz, double *__restrict__ out) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    out[f] = (fx[f] * m.Sfx[f] + fy[f] * m.Sfy[f]) + fz[f] * m.Sfz[f];
}

int fill_bcvals(const fvh_bc *bcs, int nbc, int comp, BcVals &out) {
    if (nbc > kMaxBc) return set_error(FVH_ERR_SIZE, "at most %d boundary conditions per field (got %d)", kMaxBc, nbc);
    if (comp < 0 || comp > 2) return set_error(FVH_ERR_ARG, "component %d", comp);
    std::memset(&out, 0, sizeof out);
    out.nbc = nbc;
    for (int i = 0; i < nbc; ++i) {
        out.kind[i] = bcs[i].kind;
        out.patch[i] = bcs[i].patch;
        out.value[i] = bcs[i].value[comp];
    }
    return FVH_OK;
}

static inline dim3 grid_of(int64_t n) { return dim3((unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1)); }

int launch_interpolate(const fvh_mesh *m, const double *phi, double *phif) {
    hipLaunchKernelGGL(k_interpolate, grid_of(m->nf), dim3(256), 0, stream(), m->d, phi, phif);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
int launch_interpolate_vec(const fvh_mesh *m, const double *ux, const double *uy, double *fx, double *fy) {
    hipLaunchKernelGGL(k_interpolate_vec, grid_of(m->nf), dim3(256), 0, stream(), m->d, ux, uy, fx, fy);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
int launch_interpolate_midpoint_vec(const fvh_mesh *m, const double *ux, const double *uy, const double *uz, double *fx,
                                    double *fy, double *fz) {
    hipLaunchKernelGGL(k_interpolate_midpoint_vec, grid_of(m->nf), dim3(256), 0, stream(), m->d, ux, uy, uz, fx, fy, fz);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
int launch_correct_boundaries(const fvh_mesh *m, const BcVals &bc, const double *phi, double *phif) {
    if (m->nbf == 0 || bc.nbc == 0) return FVH_OK;
    hipLaunchKernelGGL(k_correct_boundaries, grid_of(m->nbf), dim3(256), 0, stream(), m->d, bc, phi, phif);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
int launch_correct_boundaries_vec(const fvh_mesh *m, const fvh_bc *bcs, int nbc, const double *ux, const double *uy,
                                  const double *uz, double *fx, double *fy, double *fz) {
    if (m->nbf == 0 || nbc == 0) return FVH_OK;
    BcVals bx, by, bz;
    FVH_TRY(fill_bcvals(bcs, nbc, 0, bx));
    FVH_TRY(fill_bcvals(bcs, nbc, 1, by));
    FVH_TRY(fill_bcvals(bcs, nbc, 2, bz));
    hipLaunchKernelGGL(k_correct_boundaries_vec, grid_of(m->nbf), dim3(256), 0, stream(), m->d, bx, by, bz, ux, uy, uz,
                       fx, fy, fz);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
int launch_green_gauss(const fvh_mesh *m, const double *phif, double *dx, double *dy, double *dz) {
    hipLaunchKernelGGL(k_green_gauss, grid_of(m->n), dim3(256), 0, stream(), m->d, phif, dx, dy, dz);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
int launch_correct_interpolation(const fvh_mesh *m, const double *dx, const double *dy, const double *dz,
                                 const double *phi, double *phif) {
    int64_t ni = m->nf - m->nbf;
    if (ni <= 0) return FVH_OK;
    hipLaunchKernelGGL(k_correct_interpolation, grid_of(ni), dim3(256), 0, stream(), m->d, dx, dy, dz, phi, phif);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
int launch_div(const fvh_mesh *m, const double *fx, const double *fy, const double *fz, double *out) {
    hipLaunchKernelGGL(k_div, grid_of(m->n), dim3(256), 0, stream(), m->d, fx, fy, fz, out);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
int launch_flux(const fvh_mesh *m, const double *fx, const double *fy, const double *fz, double *out) {
    hipLaunchKernelGGL(k_flux, grid_of(m->nf), dim3(256), 0, stream(), m->d, fx, fy, fz, out);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// Non-orthogonal correction (Calculate_3_orthogonality_correction.jl,
// Calculate_2_interpolation.jl:152-176).  The reference defines the
// interpolation weight for these operators only for the Linear scheme
// (weight(::Type{Linear}, ...), :81-93): the face centre projected on the
// line between the cell centres.
// ---------------------------------------------------------------------------
__device__ inline double nonorth_weight(const DevMesh &m, int f, int c1, int c2) {
    const double c1x = m.ccx[c1], c1y = m.ccy[c1], c1z = m.ccz[c1];
    const double c2x = m.ccx[c2], c2y = m.ccy[c2], c2z = m.ccz[c2];
    const double ax = m.fcx[f] - c1x, ay = m.fcy[f] - c1y, az = m.fcz[f] - c1z;  // c1_f
    const double bx = c2x - c1x, by = c2y - c1y, bz = c2z - c1z;              // c1_c2
    const double q = ((ax * bx + ay * by) + az * bz) / ((bx * bx + by * by) + bz * bz);
    // f' = c1 - q (c1 - c2)
    const double px = c1x - q * (c1x - c2x), py = c1y - q * (c1y - c2y), pz = c1z - q * (c1z - c2z);
    const double dx = c2x - px, dy = c2y - py, dz = c2z - pz;
    const double ex = c2x - c1x, ey = c2y - c1y, ez = c2z - c1z;
    return sqrt((dx * dx + dy * dy) + dz * dz) / sqrt((ex * ex + ey * ey) + ez * ez);
}

// interpolate!(gradf, grad, phi): internal faces, weighted average of the
// two cell gradients corrected along e by the face-normal difference.
__global__ void k_interpolate_gradient(DevMesh m, const double *__restrict__ gx, const double *__restrict__ gy,
                                       const double *__restrict__ gz, const double *__restrict__ phi,
                                       double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz) {
    const int f = m.nbf + blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    const int c1 = m.own1[f], c2 = m.own2[f];
    const double w = nonorth_weight(m, f, c1, c2);
    const double omw = 1.0 - w;
    const double ax = w * gx[c1] + omw * gx[c2];
    const double ay = w * gy[c1] + omw * gy[c2];
    const double az = w * gz[c1] + omw * gz[c2];
    const double ex = m.ex[f], ey = m.ey[f], ez = m.ez[f];
    const double s = (phi[c2] - phi[c1]) / m.fdelta[f] - ((ax * ex + ay * ey) + az * ez);
    fx[f] = ax + s * ex;
    fy[f] = ay + s * ey;
    fz[f] = az + s * ez;
}

// nonorthogonal_flux!(phif, gradf): internal faces, gradf . (normal - e) area.
__global__ void k_nonorthogonal_flux(DevMesh m, const double *__restrict__ fx, const double *__restrict__ fy,
                                     const double *__restrict__ fz, double *__restrict__ phif) {
    const int f = m.nbf + blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nf) return;
    const double a = m.farea[f];
    const double tx = (m.nx[f] - m.ex[f]) * a, ty = (m.ny[f] - m.ey[f]) * a, tz = (m.nz[f] - m.ez[f]) * a;
    phif[f] = (fx[f] * tx + fy[f] * ty) + fz[f] * tz;
}

// correct!(eqn, term, corr_flux): the reference walks every face in id
// order, adding c = -sign J_f corr_f to its first owner's b and -c to its
// second (a boundary face gives its one cell +c then -c).  Per cell that is
// its boundary faces in ascending id, then its internal faces in ascending
// id (faces_range order), each boundary face first +c then -c.
__global__ void k_nonorthogonal_correct(DevMesh m, double msign, const double *__restrict__ J,
                                        const double *__restrict__ corr, double *__restrict__ b) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    double v = b[c];
    for (int i = m.cb_ptr[c]; i < m.cb_ptr[c + 1]; ++i) {
        const int f = m.cb_face[i];
        const double k = msign * J[f] * corr[f];
        v += k;
        v += k * (-1.0);
    }
    for (int i = m.cf_ptr[c]; i < m.cf_ptr[c + 1]; ++i) {
        const int f = m.cf_face[i];
        const double k = msign * J[f] * corr[f];
        v += (m.own1[f] == c) ? k : k * (-1.0);
    }
    b[c] = v;
}

int launch_interpolate_gradient(const fvh_mesh *m, const double *gx, const double *gy, const double *gz,
                                const double *phi, double *fx, double *fy, double *fz) {
    const int64_t ni = m->nf - m->nbf;
    if (ni <= 0) return FVH_OK;
    hipLaunchKernelGGL(k_interpolate_gradient, grid_of(ni), dim3(256), 0, stream(), m->d, gx, gy, gz, phi, fx, fy, fz);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
int launch_nonorthogonal_flux(const fvh_mesh *m, const double *fx, const double *fy, const double *fz, double *phif) {
    const int64_t ni = m->nf - m->nbf;
    if (ni <= 0) return FVH_OK;
    hipLaunchKernelGGL(k_nonorthogonal_flux, grid_of(ni), dim3(256), 0, stream(), m->d, fx, fy, fz, phif);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
int launch_nonorthogonal_correct(const fvh_mesh *m, int sign, const double *J, const double *corr, double *b) {
    if (!m->cf_ascending)
        return set_error(FVH_ERR_UNSUPPORTED, "correct!: a cell's internal faces are not in ascending face order");
    hipLaunchKernelGGL(k_nonorthogonal_correct, grid_of(m->n), dim3(256), 0, stream(), m->d, (double)(-sign), J, corr,
                       b);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// grad!(grad, phif, phi, BCs) for a scalar field
int launch_grad(const fvh_mesh *m, int scheme, const BcVals &bc, const double *phi, double *phif, double *dx, double *dy,
                double *dz) {
    if (scheme == FVH_MIDPOINT) {
        hipLaunchKernelGGL(k_interpolate_midpoint, grid_of(m->nf), dim3(256), 0, stream(), m->d, phi, phif);
        HIP_TRY(hipGetLastError());
        FVH_TRY(launch_correct_boundaries(m, bc, phi, phif));
        FVH_TRY(launch_green_gauss(m, phif, dx, dy, dz));
        for (int it = 0; it < 2; ++it) {
            FVH_TRY(launch_correct_interpolation(m, dx, dy, dz, phi, phif));
            FVH_TRY(launch_green_gauss(m, phif, dx, dy, dz));
        }
        return FVH_OK;
    }
    if (scheme != FVH_ORTHOGONAL) return set_error(FVH_ERR_UNSUPPORTED, "gradient scheme %d", scheme);
    FVH_TRY(launch_interpolate(m, phi, phif));
    FVH_TRY(launch_correct_boundaries(m, bc, phi, phif));
    return launch_green_gauss(m, phif, dx, dy, dz);
}

}  // namespace fvh

using namespace fvh;

#define SYNC_RETURN()                                   \
    do {                                                \
        HIP_TRY(hipStreamSynchronize(stream()));        \
        return FVH_OK;                                  \
    } while (0)

extern "C" int fvh_interpolate(const fvh_mesh *m, const double *phi, double *phif) {
    if (!m || !phi || !phif) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(launch_interpolate(m, phi, phif));
    SYNC_RETURN();
}
extern "C" int fvh_interpolate_vector(const fvh_mesh *m, const double *ux, const double *uy, double *fx, double *fy) {
    if (!m || !ux || !uy || !fx || !fy) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(launch_interpolate_vec(m, ux, uy, fx, fy));
    SYNC_RETURN();
}
extern "C" int fvh_correct_boundaries(const fvh_mesh *m, const fvh_bc *bcs, int32_t nbc, int32_t comp, const double *phi,
                                      double *phif) {
    if (!m || (!bcs && nbc) || !phi || !phif) return set_error(FVH_ERR_ARG, "null argument");
    BcVals bv;
    FVH_TRY(fill_bcvals(bcs, nbc, comp, bv));
    FVH_TRY(launch_correct_boundaries(m, bv, phi, phif));
    SYNC_RETURN();
}
extern "C" int fvh_green_gauss(const fvh_mesh *m, const double *phif, double *dx, double *dy, double *dz) {
    if (!m || !phif || !dx || !dy || !dz) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(launch_green_gauss(m, phif, dx, dy, dz));
    SYNC_RETURN();
}
extern "C" int fvh_grad(const fvh_mesh *m, int32_t scheme, const fvh_bc *bcs, int32_t nbc, int32_t comp, const double *phi,
                        double *phif, double *dx, double *dy, double *dz) {
    if (!m || (!bcs && nbc) || !phi || !phif || !dx || !dy || !dz) return set_error(FVH_ERR_ARG, "null argument");
    BcVals bv;
    FVH_TRY(fill_bcvals(bcs, nbc, comp, bv));
    FVH_TRY(launch_grad(m, scheme, bv, phi, phif, dx, dy, dz));
    SYNC_RETURN();
}
extern "C" int fvh_div(const fvh_mesh *m, const double *fx, const double *fy, const double *fz, double *out) {
    if (!m || !fx || !fy || !fz || !out) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(launch_div(m, fx, fy, fz, out));
    SYNC_RETURN();
}
extern "C" int fvh_flux(const fvh_mesh *m, const double *fx, const double *fy, const double *fz, double *out) {
    if (!m || !fx || !fy || !fz || !out) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(launch_flux(m, fx, fy, fz, out));
    SYNC_RETURN();
}
extern "C" int fvh_interpolate_gradient(const fvh_mesh *m, const double *gx, const double *gy, const double *gz,
                                        const double *phi, double *fx, double *fy, double *fz) {
    if (!m || !gx || !gy || !gz || !phi || !fx || !fy || !fz) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(launch_interpolate_gradient(m, gx, gy, gz, phi, fx, fy, fz));
    SYNC_RETURN();
}
extern "C" int fvh_nonorthogonal_flux(const fvh_mesh *m, const double *fx, const double *fy, const double *fz,
                                      double *phif) {
    if (!m || !fx || !fy || !fz || !phif) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(launch_nonorthogonal_flux(m, fx, fy, fz, phif));
    SYNC_RETURN();
}
