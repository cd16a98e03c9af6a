// DILU / ILU(0) recurrences, level-scheduled in the reference's row order.
//
// The DILU diagonal recurrence (preconditioners_1_DILU.jl:56-105) and both
// sweeps (:107-162) are evaluated row by row with each row's terms in the
// reference's order, so the preconditioner and its rounding are unchanged.
// Rows of one level have no dependency on each other; one work-group walks
// the levels with a barrier between them (rows of a level in parallel).
#include "k_sweep.h"

namespace fvh {

constexpr int kSweepBlock = 1024;

__global__ void __launch_bounds__(kSweepBlock)
    k_dilu_update_levels(int nlev, const int *__restrict__ lptr, const int *__restrict__ lrows,
                         const int *__restrict__ rowptr, const int *__restrict__ col, const int *__restrict__ diag,
                         const int *__restrict__ tslot, const double *__restrict__ val, double *D, double *Dinv) {
    for (int l = 0; l < nlev; ++l) {
        for (int k = lptr[l] + threadIdx.x; k < lptr[l + 1]; k += kSweepBlock) {
            const int j = lrows[k];
            double d = val[diag[j]];
            for (int p = rowptr[j]; p < diag[j]; ++p) {
                const int i = col[p];
                d = d - val[p] * val[tslot[p]] / D[i];
            }
            D[j] = d;
            Dinv[j] = 1.0 / d;
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(kSweepBlock)
    k_dilu_apply_levels(int nlev, const int *__restrict__ fptr, const int *__restrict__ frows,
                        const int *__restrict__ bptr, const int *__restrict__ brows, const int *__restrict__ rowptr,
                        const int *__restrict__ col, const int *__restrict__ diag, const T *__restrict__ val,
                        const T *__restrict__ Dinv, const T *__restrict__ x, T *y, const KScal *ks) {
    if (ks && ks->done) return;
    // forward: y_i = x_i - sum_{j<i} (A_ij y_j) Dinv_j, j ascending
    // (FP32: val holds A_ij Dinv_j below and A_ij Dinv_i above the diagonal,
    // rounded once, and the terms are fused multiply-adds, j descending)
    for (int l = 0; l < nlev; ++l) {
        for (int k = fptr[l] + threadIdx.x; k < fptr[l + 1]; k += kSweepBlock) {
            const int i = frows[k];
            T s = x[i];
            if constexpr (sizeof(T) == 4) {
                for (int p = diag[i] - 1; p >= rowptr[i]; --p) s = __builtin_fmaf(-val[p], y[col[p]], s);
            } else {
                for (int p = rowptr[i]; p < diag[i]; ++p) s = s - val[p] * y[col[p]] * Dinv[col[p]];
            }
            y[i] = s;
        }
        __syncthreads();
    }
    // backward: z_i = y_i Dinv_i - sum_{j>i} (A_ij z_j) Dinv_i, j ascending
    for (int l = 0; l < nlev; ++l) {
        for (int k = bptr[l] + threadIdx.x; k < bptr[l + 1]; k += kSweepBlock) {
            const int i = brows[k];
            const T di = Dinv[i];
            T s = y[i] * di;
            for (int p = diag[i] + 1; p < rowptr[i + 1]; ++p) {
                if constexpr (sizeof(T) == 4) s = __builtin_fmaf(-val[p], y[col[p]], s);
                else s = s - val[p] * y[col[p]] * di;
            }
            y[i] = s;
        }
        __syncthreads();
    }
}

__global__ void k_copy(int n, const double *__restrict__ a, double *__restrict__ b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// ILU(0), Saad Alg. 10.4 (IKJ) row by row on A's pattern
__global__ void __launch_bounds__(kSweepBlock)
    k_ilu0_factor_levels(int nlev, const int *__restrict__ lptr, const int *__restrict__ lrows,
                         const int *__restrict__ rowptr, const int *__restrict__ col, const int *__restrict__ diag,
                         const int *__restrict__ uptr, const int *__restrict__ pairs, double *lu) {
    for (int l = 0; l < nlev; ++l) {
        for (int k = lptr[l] + threadIdx.x; k < lptr[l + 1]; k += kSweepBlock) {
            const int i = lrows[k];
            for (int pk = rowptr[i]; pk < diag[i]; ++pk) {
                const int kk = col[pk];
                double lik = lu[pk] / lu[diag[kk]];
                lu[pk] = lik;
                for (int q = uptr[pk]; q < uptr[pk + 1]; q += 2) lu[pairs[q]] = lu[pairs[q]] - lik * lu[pairs[q + 1]];
            }
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(kSweepBlock)
    k_ilu0_apply_levels(int nlev, const int *__restrict__ fptr, const int *__restrict__ frows,
                        const int *__restrict__ bptr, const int *__restrict__ brows, const int *__restrict__ rowptr,
                        const int *__restrict__ col, const int *__restrict__ diag, const T *__restrict__ lu,
                        const T *__restrict__ x, T *y, const KScal *ks) {
    if (ks && ks->done) return;
    for (int l = 0; l < nlev; ++l) {
        for (int k = fptr[l] + threadIdx.x; k < fptr[l + 1]; k += kSweepBlock) {
            const int i = frows[k];
            T s = x[i];
            for (int p = rowptr[i]; p < diag[i]; ++p) s = s - lu[p] * y[col[p]];
            y[i] = s;
        }
        __syncthreads();
    }
    for (int l = 0; l < nlev; ++l) {
        for (int k = bptr[l] + threadIdx.x; k < bptr[l + 1]; k += kSweepBlock) {
            const int i = brows[k];
            T s = y[i];
            for (int p = diag[i] + 1; p < rowptr[i + 1]; ++p) s = s - lu[p] * y[col[p]];
            y[i] = s / lu[diag[i]];
        }
        __syncthreads();
    }
}

int dilu_update(fvh_mesh *m, const double *val, double *D, double *Dinv) {
    const Schedule &s = m->sched;
    hipLaunchKernelGGL(k_dilu_update_levels, dim3(1), dim3(kSweepBlock), 0, stream(), s.nlev, s.fptr.p, s.frows.p,
                       m->d.rowptr, m->d.col, m->d.diag, m->d.tslot, val, D, Dinv);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

template <typename T>
static int dilu_apply_t(fvh_mesh *m, const T *val, const T *Dinv, const T *x, T *y, const KScal *ks) {
    const Schedule &s = m->sched;
    hipLaunchKernelGGL((k_dilu_apply_levels<T>), dim3(1), dim3(kSweepBlock), 0, stream(), s.nlev, s.fptr.p, s.frows.p, s.bptr.p,
                       s.brows.p, m->d.rowptr, m->d.col, m->d.diag, val, Dinv, x, y, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int ilu0_factor(fvh_mesh *m, const double *val, double *lu) {
    const Schedule &s = m->sched;
    const int nnz = (int)m->nnz;
    hipLaunchKernelGGL(k_copy, dim3((nnz + 255) / 256), dim3(256), 0, stream(), nnz, val, lu);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_ilu0_factor_levels, dim3(1), dim3(kSweepBlock), 0, stream(), s.nlev, s.fptr.p, s.frows.p,
                       m->d.rowptr, m->d.col, m->d.diag, m->ilu_uptr.p, m->ilu_pairs.p, lu);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

template <typename T>
static int ilu0_apply_t(fvh_mesh *m, const T *lu, const T *x, T *y, const KScal *ks) {
    const Schedule &s = m->sched;
    hipLaunchKernelGGL((k_ilu0_apply_levels<T>), dim3(1), dim3(kSweepBlock), 0, stream(), s.nlev, s.fptr.p, s.frows.p, s.bptr.p,
                       s.brows.p, m->d.rowptr, m->d.col, m->d.diag, lu, x, y, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int dilu_apply(fvh_mesh *m, const double *val, const double *Dinv, const double *x, double *y, const KScal *ks) {
    return dilu_apply_t<double>(m, val, Dinv, x, y, ks);
}
int dilu_apply32(fvh_mesh *m, const float *val, const float *Dinv, const float *x, float *y, const KScal *ks) {
    return dilu_apply_t<float>(m, val, Dinv, x, y, ks);
}
int ilu0_apply(fvh_mesh *m, const double *lu, const double *x, double *y, const KScal *ks) {
    return ilu0_apply_t<double>(m, lu, x, y, ks);
}
int ilu0_apply32(fvh_mesh *m, const float *lu, const float *x, float *y, const KScal *ks) {
    return ilu0_apply_t<float>(m, lu, x, y, ks);
}

}  // namespace fvh
