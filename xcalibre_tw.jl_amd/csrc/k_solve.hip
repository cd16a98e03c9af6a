// Solve: CSR SpMV, deterministic reductions, DILU / ILU(0) / Jacobi /
// NormDiagonal preconditioners and the Krylov.jl 0.9.5 BiCGSTAB, CG and
// GMRES recurrences, all on the device.
//
// Bitwise contract of the FP64 path (compiled with -ffp-contract=off):
//   * SpMV sums each row from 0 in ascending column order, which is the
//     order SparseArrays' column-scatter mul! produces for every row;
//   * every dot / norm / mean uses the canonical blocked order of
//     fvh_device.h (the oracle implements the same order);
//   * the DILU recurrences and sweeps evaluate each row's terms in the
//     reference's order, (A_ij * x_j) * Dinv, level by level.
// Krylov scalars live in a device struct (KScal); every kernel of an
// iteration returns at once when ks->done is set, so the host enqueues
// iterations in batches and reads the flag only between batches.
#include <cmath>
#include <cstdlib>
#include <limits>

#include "fvh_eqn.h"
#include "k_band.h"
#include "k_level.h"
#include "k_sweep.h"
#include "k_spmv.h"
#include "k_util.h"

namespace fvh {

// A bounded wait of the sweeps (pipelined DILU or LDL) ran out since the
// last check: reported once, by the next call that ran them.
static int fault_check(fvh_eqn *e) {
    fvh_eqn *q = e->pshare ? e->pshare : e;
    const int a = skew_check(e);
    const int b = ldl_check(q);
    const int c = lev_check(q);
    return a != FVH_OK ? a : (b != FVH_OK ? b : c);
}

// Banded grids (single case): during a solve the SpMV reads A by band
// position (five values per row, no column indices), bit-identical to CSR.
static bool band_spmv(const fvh_mesh *m) { return !m->cases && m->sched.band_nx > 0; }

static int band_pack(fvh_eqn *e) {
    const fvh_mesh *m = e->m;
    const int n = (int)m->n;
    if (e->kw.bval.n != 5 * (size_t)n) FVH_TRY(e->kw.bval.alloc(5 * (size_t)n));
    hipLaunchKernelGGL((k_band_values<double>), grid1d(n), dim3(256), 0, stream(), n, m->d.rowptr, m->sched.bmask.p,
                       e->nz.p, e->kw.bval.p);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int launch_spmv(fvh_eqn *e, const double *x, double *y, const KScal *ks) {
    const fvh_mesh *m = e->m;
    if (e->kw.band_live)
        hipLaunchKernelGGL((k_spmv_band<double>), grid1d(m->n), dim3(256), 0, stream(), (int)m->n, m->sched.band_nx,
                           m->sched.bmask.p, e->kw.bval.p, x, y, 1.0, ks);
    else
        hipLaunchKernelGGL((k_spmv<double>), grid1d(m->n), dim3(256), 0, stream(), (int)m->n, m->d.rowptr, m->d.col,
                           e->nz.p, x, y, 1.0, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

static int launch_spmv_dot(fvh_eqn *e, const double *x, double *y, double sigma, const KScal *ks, const double *w,
                           double *part) {
    const fvh_mesh *m = e->m;
    int nch = nchunks(m->n);
    if (e->kw.band_live)
        hipLaunchKernelGGL((k_spmv_dot_band<double>), dim3(nch), dim3(kWide), 0, stream(), (int)m->n,
                           m->sched.band_nx, m->sched.bmask.p, e->kw.bval.p, x, y, sigma, ks, w, part, nch);
    else
        hipLaunchKernelGGL((k_spmv_dot<double>), dim3(nch), dim3(kWide), 0, stream(), (int)m->n, m->d.rowptr,
                           m->d.col, e->nz.p, x, y, sigma, ks, w, part, nch);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// Residual: sqrt(mean((Ax-b)^2)) / norm(b)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWide) k_residual_part(int n, const int *__restrict__ rowptr,
                                                         const int *__restrict__ col, const double *__restrict__ val,
                                                         const double *__restrict__ x, const double *__restrict__ b,
                                                         double *part, int nch) {
    __shared__ double p[2][kChunk];
    const int c0 = blockIdx.x * kChunk;
#pragma unroll
    for (int h = 0; h < kChunk / kWide; ++h) {
        const int j = threadIdx.x + h * kWide, i = c0 + j;
        double d2 = 0.0, b2 = 0.0;
        if (i < n) {
            const double d = fabs(row_ax(rowptr, col, val, x, i) - b[i]);
            d2 = d * d;
            b2 = b[i] * b[i];
        }
        p[0][j] = d2;
        p[1][j] = b2;
    }
    chunk_reduce_staged<2>(p, part, nch);
}

__global__ void __launch_bounds__(kBlock) k_residual_final(const double *part, int nch, int n, double *out) {
    double r[2];
    final_reduce<2>(part, nch, r);
    if (threadIdx.x == 0) out[0] = sqrt(r[0] / (double)n) / sqrt(r[1]);
}

int residual_async(fvh_eqn *e, const double *phi, double *res_dev) {
    const fvh_mesh *m = e->m;
    const int nch = nchunks(m->n);
    if (e->red.n < (size_t)(2 * nch)) FVH_TRY(e->red.alloc(2 * nch));
    hipLaunchKernelGGL(k_residual_part, dim3(nch), dim3(kWide), 0, stream(), (int)m->n, m->d.rowptr, m->d.col, e->nz.p,
                       phi, e->b.p, e->red.p, nch);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(kBlock), 0, stream(), e->red.p, nch, (int)m->n, res_dev);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int residual(fvh_eqn *e, const double *phi, double *res_host) {
    DBuf<double> out;
    FVH_TRY(out.alloc(1));
    FVH_TRY(residual_async(e, phi, out.p));
    HIP_TRY(hipMemcpyAsync(res_host, out.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// Preconditioners
// ---------------------------------------------------------------------------
__global__ void k_jacobi_store(int n, const int *__restrict__ diag, const double *__restrict__ val, double *__restrict__ s) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) s[i] = 1.0 / fabs(val[diag[i]]);
}

// 1 / norm of column i (generic_norm2 of a short column: sequential sum of squares)
__global__ void k_normdiag_store(int n, const int *__restrict__ rowptr, const int *__restrict__ tslot,
                                 const double *__restrict__ val, double *__restrict__ s) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double mx = 0.0;
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) mx = fmax(mx, fabs(val[tslot[p]]));
    double nrm;
    if (mx == 0.0 || isinf(mx)) nrm = mx;
    else {
        double acc = 0.0;
        bool first = true;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            double v = val[tslot[p]] * val[tslot[p]];
            acc = first ? v : acc + v;
            first = false;
        }
        nrm = sqrt(acc);
    }
    s[i] = 1.0 / nrm;
}

static int ilu_setup(fvh_mesh *m) {
    if (m->ilu_ready) return FVH_OK;
    const int n = (int)m->n;
    const auto &rp = m->h_rowptr, &col = m->h_col, &dg = m->h_diag;
    std::vector<int> uptr(m->nnz + 1, 0), pairs;
    for (int i = 0; i < n; ++i)
        for (int pk = rp[i]; pk < rp[i + 1]; ++pk) {
            if (pk < dg[i]) {
                int k = col[pk];
                for (int pj = pk + 1; pj < rp[i + 1]; ++pj) {
                    int j = col[pj];
                    auto b = col.begin() + rp[k], e = col.begin() + rp[k + 1];
                    auto it = std::lower_bound(b, e, j);
                    if (it != e && *it == j) {
                        pairs.push_back(pj);
                        pairs.push_back((int)(it - col.begin()));
                    }
                }
            }
            uptr[pk + 1] = (int)pairs.size();
        }
    FVH_TRY(m->ilu_uptr.upload(uptr));
    FVH_TRY(m->ilu_pairs.upload(pairs));
    m->ilu_ready = true;
    return FVH_OK;
}

int precond_update(fvh_eqn *eq) {
    fvh_eqn *e = eq->pshare ? eq->pshare : eq;  // preconditioner object (may be shared)
    fvh_mesh *m = e->m;
    fvh_eqn *src = e->pmat ? e->pmat : e;       // matrix it reads
    const int n = (int)m->n;
    switch (e->pkind) {
        case FVH_PRECOND_NONE: return FVH_OK;
        case FVH_PRECOND_JACOBI:
            hipLaunchKernelGGL(k_jacobi_store, grid1d(n), dim3(256), 0, stream(), n, m->d.diag, src->nz.p, e->pstore.p);
            HIP_TRY(hipGetLastError());
            return FVH_OK;
        case FVH_PRECOND_NORMDIAG:
            hipLaunchKernelGGL(k_normdiag_store, grid1d(n), dim3(256), 0, stream(), n, m->d.rowptr, m->d.tslot, src->nz.p,
                               e->pstore.p);
            HIP_TRY(hipGetLastError());
            return FVH_OK;
        case FVH_PRECOND_DILU:
            if (band_path(m)) return band_dilu_update(e, src->nz.p);
            if (lev_path(m)) return lev_dilu_update(e, src->nz.p);
            return dilu_update(m, src->nz.p, e->D.p, e->Dinv.p);
        case FVH_PRECOND_ILU0: return ilu0_factor(m, src->nz.p, e->lu.p);
        case FVH_PRECOND_LDL: return ldl_setup(e, src->nz.p);  // the first update factors; later ones keep it
        default: return set_error(FVH_ERR_UNSUPPORTED, "preconditioner %d", e->pkind);
    }
}

// y = M^{-1} x.  With independent cases ks is the array of per-case Krylov
// scalars: only the per-case sweeps look at it; the elementwise and
// level-scheduled paths then cover every case.
int precond_apply(fvh_eqn *eq, double *y, const double *x, const KScal *ks_in, double) {
    fvh_eqn *e = eq->pshare ? eq->pshare : eq;
    fvh_mesh *m = e->m;
    fvh_eqn *src = e->pmat ? e->pmat : e;
    const int n = (int)m->n;
    const KScal *ks = m->cases ? nullptr : ks_in;
    switch (e->pkind) {
        case FVH_PRECOND_NONE: {
            return map(n, ks, [=] __device__(int i) { y[i] = x[i]; });
        }
        case FVH_PRECOND_JACOBI:
        case FVH_PRECOND_NORMDIAG: {
            const double *s = e->pstore.p;
            return map(n, ks, [=] __device__(int i) { y[i] = s[i] * x[i]; });
        }
        case FVH_PRECOND_DILU:
            if (band_path(m)) return band_dilu_apply(e, x, y, ks_in);
            if (lev_path(m) && e->lv.fwd.p) return lev_dilu_apply(e, x, y, ks);
            return dilu_apply(m, src->nz.p, e->Dinv.p, x, y, ks);
        case FVH_PRECOND_ILU0: return ilu0_apply(m, e->lu.p, x, y, ks);
        case FVH_PRECOND_LDL: return ldl_apply<double>(e, x, y, ks);
        default: return set_error(FVH_ERR_UNSUPPORTED, "preconditioner %d", e->pkind);
    }
}

// ---------------------------------------------------------------------------
// FP32 operators of the mixed-precision inner solves
// ---------------------------------------------------------------------------
int launch_spmv32(fvh_eqn *e, const float *x, float *y, float sigma, const KScal *ks, const float *w,
                  double *part) {
    const fvh_mesh *m = e->m;
    if (e->kw.band_live && e->kw.bval32.p) {
        if (!w) {
            hipLaunchKernelGGL((k_spmv_band<float>), grid1d(m->n), dim3(256), 0, stream(), (int)m->n,
                               m->sched.band_nx, m->sched.bmask.p, e->kw.bval32.p, x, y, sigma, ks);
        } else {
            const int nch = nchunks(m->n);
            hipLaunchKernelGGL((k_spmv_dot_band<float>), dim3(nch), dim3(kWide), 0, stream(), (int)m->n,
                               m->sched.band_nx, m->sched.bmask.p, e->kw.bval32.p, x, y, sigma, ks, w, part, nch);
        }
        HIP_TRY(hipGetLastError());
        return FVH_OK;
    }
    if (!w) {
        hipLaunchKernelGGL((k_spmv<float>), grid1d(m->n), dim3(256), 0, stream(), (int)m->n, m->d.rowptr, m->d.col,
                           e->kw.valf.p, x, y, sigma, ks);
    } else {
        const int nch = nchunks(m->n);
        hipLaunchKernelGGL((k_spmv_dot<float>), dim3(nch), dim3(kWide), 0, stream(), (int)m->n, m->d.rowptr, m->d.col,
                           e->kw.valf.p, x, y, sigma, ks, w, part, nch);
    }
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// FP32 copies read by the inner solves: A (for the SpMV) and the
// preconditioner's data, rounded once from the FP64 values after each update.
// The DILU reciprocals are computed in FP64 by the update itself.
int prepare32(fvh_eqn *e) {
    const fvh_mesh *m = e->m;
    const int64_t n = m->n, nnz = m->nnz;
    KrylovWork &w = e->kw;
    if (w.valf.n != (size_t)nnz) FVH_TRY(w.valf.alloc(nnz));
    {
        const double *a = e->nz.p;
        float *af = w.valf.p;
        FVH_TRY(map(nnz, nullptr, [=] __device__(int i) { af[i] = (float)a[i]; }));
    }
    if (band_spmv(m)) {  // the same values by band position (each rounded once, as in valf)
        if (w.bval32.n != 5 * (size_t)n) FVH_TRY(w.bval32.alloc(5 * (size_t)n));
        hipLaunchKernelGGL((k_band_values<float>), grid1d(n), dim3(256), 0, stream(), (int)n, m->d.rowptr,
                           m->sched.bmask.p, e->nz.p, w.bval32.p);
        HIP_TRY(hipGetLastError());
    }
    fvh_eqn *P = e->pshare ? e->pshare : e;
    fvh_eqn *src = P->pmat ? P->pmat : P;
    switch (P->pkind) {
        case FVH_PRECOND_NONE: return FVH_OK;
        case FVH_PRECOND_LDL: return FVH_OK;  // the factor keeps its own FP32 copy, made at set-up
        case FVH_PRECOND_JACOBI:
        case FVH_PRECOND_NORMDIAG: {
            if (P->p32a.n != (size_t)n) FVH_TRY(P->p32a.alloc(n));
            const double *s = P->pstore.p;
            float *sf = P->p32a.p;
            return map(n, nullptr, [=] __device__(int i) { sf[i] = (float)s[i]; });
        }
        case FVH_PRECOND_DILU: {
            if (band_path(m)) return band_to32(P);
            if (lev_path(m) && P->lv.fwd.p) FVH_TRY(lev_to32(P));
            if (P->p32a.n != (size_t)n) FVH_TRY(P->p32a.alloc(n));
            if (P->p32b.n != (size_t)nnz) FVH_TRY(P->p32b.alloc(nnz));
            const double *di = P->Dinv.p, *a = src->nz.p;
            float *df = P->p32a.p, *af = P->p32b.p;
            const int *rp = m->d.rowptr, *cl = m->d.col;
            FVH_TRY(map(n, nullptr, [=] __device__(int i) { df[i] = (float)di[i]; }));
            // each off-diagonal premultiplied by the pivot it meets in the
            // sweeps (its column's below the diagonal, its row's above),
            // rounded once
            return map(n, nullptr, [=] __device__(int i) {
                for (int p = rp[i]; p < rp[i + 1]; ++p) {
                    const int j = cl[p];
                    af[p] = (float)(a[p] * (j < i ? di[j] : di[i]));
                }
            });
        }
        case FVH_PRECOND_ILU0: {
            if (P->p32b.n != (size_t)nnz) FVH_TRY(P->p32b.alloc(nnz));
            const double *lu = P->lu.p;
            float *lf = P->p32b.p;
            return map(nnz, nullptr, [=] __device__(int i) { lf[i] = (float)lu[i]; });
        }
        default: return set_error(FVH_ERR_UNSUPPORTED, "preconditioner %d", P->pkind);
    }
}

// y = M^{-1} x in FP32 (same recurrences, FP32 copies of their data)
int precond_apply32(fvh_eqn *eq, float *y, const float *x, const KScal *ks_in) {
    fvh_eqn *e = eq->pshare ? eq->pshare : eq;
    fvh_mesh *m = e->m;
    const int n = (int)m->n;
    const KScal *ks = m->cases ? nullptr : ks_in;
    switch (e->pkind) {
        case FVH_PRECOND_NONE: return map(n, ks, [=] __device__(int i) { y[i] = x[i]; });
        case FVH_PRECOND_JACOBI:
        case FVH_PRECOND_NORMDIAG: {
            const float *s = e->p32a.p;
            return map(n, ks, [=] __device__(int i) { y[i] = s[i] * x[i]; });
        }
        case FVH_PRECOND_DILU:
            if (band_path(m)) return band_dilu_apply32(e, x, y, ks_in);
            if (lev_path(m) && e->lv.fwd32.p) return lev_dilu_apply32(e, x, y, ks);
            return dilu_apply32(m, e->p32b.p, e->p32a.p, x, y, ks);
        case FVH_PRECOND_ILU0: return ilu0_apply32(m, e->p32b.p, x, y, ks);
        case FVH_PRECOND_LDL: return ldl_apply<float>(e, x, y, ks);  // the FP64 factor, FP32 vectors
        default: return set_error(FVH_ERR_UNSUPPORTED, "preconditioner %d", e->pkind);
    }
}

// ---------------------------------------------------------------------------
// Krylov workspace
// ---------------------------------------------------------------------------
int ensure_work(fvh_eqn *e, int basis) {
    KrylovWork &w = e->kw;
    const int64_t n = e->m->n;
    if (w.n != n) {
        w.drop_graphs();  // captured kernels hold the old workspace pointers
        // one set of Krylov scalars per independent case
        const int K = e->m->cases ? e->m->cases->K : 1;
        const int nred = e->m->cases ? std::max(e->m->cases->nblk, nchunks(n)) : nchunks(n);
        for (auto &v : w.v) FVH_TRY(v.alloc(n));
        FVH_TRY(w.part.alloc(4 * (size_t)nred + 4));
        FVH_TRY(w.ks.alloc(K));
        if (w.h_ks) (void)hipHostFree(w.h_ks);
        w.h_ks = nullptr;
        HIP_TRY(hipHostMalloc((void **)&w.h_ks, sizeof(KScal) * K, hipHostMallocDefault));
        w.n = n;
        w.cap_basis = 0;
        w.n32 = 0;
    }
    if (basis > w.cap_basis) {
        w.drop_graphs();
        FVH_TRY(w.basis.alloc((size_t)(basis + 1) * n));
        FVH_TRY(w.small.alloc((size_t)basis * (basis + 1) / 2 + 4 * (size_t)(basis + 2)));
        w.cap_basis = basis;
    }
    return FVH_OK;
}

int ensure_work32(fvh_eqn *e, int basis) {
    FVH_TRY(ensure_work(e, basis));
    KrylovWork &w = e->kw;
    const int64_t n = e->m->n;
    if (w.n32 != n) {
        w.drop_graphs();
        for (auto &v : w.f) FVH_TRY(v.alloc(n));
        w.n32 = n;
        w.cap_basis32 = 0;
    }
    if (basis > w.cap_basis32) {
        FVH_TRY(w.basis32.alloc((size_t)(basis + 1) * n));
        w.cap_basis32 = basis;
    }
    return FVH_OK;
}

static int read_ks(fvh_eqn *e) {
    HIP_TRY(hipMemcpyAsync(e->kw.h_ks, e->kw.ks.p, sizeof(KScal), hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

__device__ inline double dsign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }

// Iterations are enqueued in batches between host checks of ks->done: 2, 2,
// then as many as are done, up to 16 (each check idles the device for a
// host round trip; iterations enqueued past the end skip at once).
static int next_batch(int done_iters) { return done_iters < 4 ? 2 : std::min(16, done_iters); }
// The first batch is enqueued without reading the scalars set up before it:
// a solve that is already done (zero or small residual, itmax <= 0) finds
// ks->done set on the device, and every kernel of the batch returns at once.
// The read after each batch waits for the batch's kernels (and so for every
// sweep the solve ran, before the caller's fault check); a solve that ran no
// batch reads the scalars once at its end.
static void first_batch(fvh_eqn *e) { e->kw.h_ks->done = 0; }

// A linear system as the Krylov recurrences see it.  T is the storage type
// of the vectors: double for the reference path (start value in x, A and M
// in FP64), float for the inner solves of mixed precision (start value 0,
// A and M in FP32).  Scalars and every dot product stay FP64; for T = double
// each expression below is the FP64 one term for term.
template <typename T>
struct Sys {
    fvh_eqn *e;
    int n;
    const T *b;
    bool warm;
    T *v[8];
    T *basis;
};

template <typename T>
__device__ inline double D(T x) {
    return (double)x;
}

static int sys_spmv(const Sys<double> &S, const double *x, double *y, const KScal *ks) {
    return launch_spmv(S.e, x, y, ks);
}
static int sys_spmv(const Sys<float> &S, const float *x, float *y, const KScal *ks) {
    return launch_spmv32(S.e, x, y, 1.0f, ks, nullptr, nullptr);
}
static int sys_spmv_dot(const Sys<double> &S, const double *x, double *y, double sigma, const KScal *ks,
                        const double *w, double *part) {
    return launch_spmv_dot(S.e, x, y, sigma, ks, w, part);
}
static int sys_spmv_dot(const Sys<float> &S, const float *x, float *y, double sigma, const KScal *ks, const float *w,
                        double *part) {
    return launch_spmv32(S.e, x, y, (float)sigma, ks, w, part);
}
static int sys_precond(const Sys<double> &S, double *y, const double *x, const KScal *ks) {
    return precond_apply(S.e, y, x, ks, 1.0);
}
static int sys_precond(const Sys<float> &S, float *y, const float *x, const KScal *ks) {
    return precond_apply32(S.e, y, x, ks);
}
// The map of natural rows into the banded DILU sweeps' result quads when the
// system's preconditioner is that apply (pipelined band sweeps, one case),
// else null: CG then reads z from the quads and skips the scatter.
template <typename T>
static const int *sys_quad_map(const Sys<T> &S) {
    const fvh_eqn *P = S.e->pshare ? S.e->pshare : S.e;
    const fvh_mesh *m = S.e->m;
    if (P->pkind != FVH_PRECOND_DILU || m->cases || !band_path(m) || !skew_active(m)) return nullptr;
    return skew_quad_map(m);
}
static int sys_precond_quads(const Sys<double> &S, const double *x, const KScal *ks, const double **zq) {
    return skew_apply_quads(S.e->pshare ? S.e->pshare : S.e, x, ks, zq);
}
static int sys_precond_quads(const Sys<float> &S, const float *x, const KScal *ks, const float **zq) {
    return skew_apply_quads32(S.e->pshare ? S.e->pshare : S.e, x, ks, zq);
}
template <typename T>
static int sys_rhs_target(const Sys<T> &S, SkewRhs *t) {
    return skew_rhs_target(S.e->pshare ? S.e->pshare : S.e, sizeof(T) == 4, t);
}
template <typename T>
__device__ inline T bits_as(unsigned long long u) {
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, u);
    else return __builtin_bit_cast(T, (unsigned)u);
}

static void fill_stats(fvh_eqn *e, fvh_stats *st) {
    st->niter = e->kw.h_ks->iter;
    st->status = e->kw.h_ks->status;
    st->rnorm = e->kw.h_ks->rnorm;
    st->rnorm0 = e->kw.h_ks->rnorm0;
}

// ---------------------------------------------------------------------------
// BiCGSTAB (Krylov.jl bicgstab!, left preconditioning, c = b)
// ---------------------------------------------------------------------------
template <typename T>
static int bicgstab(const Sys<T> &S, const fvh_solver_cfg &cfg, T *xu, fvh_stats *st) {
    fvh_eqn *e = S.e;
    const int n = S.n;
    KrylovWork &w = e->kw;
    T *xk = S.v[0], *dx = S.v[1], *r = S.v[2], *p = S.v[3], *v = S.v[4], *s = S.v[5], *q = S.v[6], *t = S.v[7];
    const T *b = S.b;
    double *part = w.part.p;
    KScal *ks = w.ks.p;
    const double atol = cfg.atol, rtol = cfg.rtol;
    const int itmax = cfg.itmax;
    FVH_TRY(single(ks, false, [=] __device__(KScal * k) {
        k->done = 0;
        k->iter = 0;
        k->status = FVH_KRYLOV_ITMAX;
        k->alpha = 1.0;
        k->omega = 1.0;
        k->rho = 1.0;
    }));
    if (S.warm) {
        FVH_TRY(map(n, nullptr, [=] __device__(int i) {
            dx[i] = xu[i];
            xk[i] = 0;
            s[i] = 0;
            v[i] = 0;
        }));
        FVH_TRY(sys_spmv(S, dx, q, nullptr));
        FVH_TRY(map(n, nullptr, [=] __device__(int i) { q[i] = b[i] - q[i]; }));
    } else {
        FVH_TRY(map(n, nullptr, [=] __device__(int i) {
            xk[i] = 0;
            s[i] = 0;
            v[i] = 0;
            q[i] = b[i];
        }));
    }
    FVH_TRY(sys_precond(S, r, q, nullptr));
    FVH_TRY(mapred<2>(n, nullptr, part, [=] __device__(int i, double(&a)[2]) {
        p[i] = r[i];
        a[0] = a[0] + D(r[i]) * D(r[i]);
        a[1] = a[1] + D(b[i]) * D(r[i]);
    }));
    FVH_TRY(finalize<2>(n, part, ks, false, [=] __device__(KScal * k, double(&rr)[2]) {
        double rn = sqrt(rr[0]);
        k->rnorm = rn;
        k->rnorm0 = rn;
        k->eps = atol + rtol * rn;
        k->beta = rr[1];  // next_rho
        if (rn == 0.0) { k->done = 1; k->status = FVH_KRYLOV_ZERO_RES; return; }
        if (rr[1] == 0.0) { k->done = 1; k->status = FVH_KRYLOV_BREAKDOWN; return; }
        if (rn <= k->eps) { k->done = 1; k->status = FVH_KRYLOV_SOLVED; return; }
        if (itmax <= 0) { k->done = 1; k->status = FVH_KRYLOV_ITMAX; }
    }));
    int enq = 0;
    first_batch(e);
    while (!w.h_ks->done && enq < itmax) {
        int nb = std::min(next_batch(enq), itmax - enq);
        for (int it = 0; it < nb; ++it) {
            FVH_TRY(sys_spmv(S, p, q, ks));
            FVH_TRY(sys_precond(S, v, q, ks));
            FVH_TRY(mapred<1>(n, ks, part, [=] __device__(int i, double(&a)[1]) { a[0] = a[0] + D(b[i]) * D(v[i]); }));
            FVH_TRY(finalize<1>(n, part, ks, true, [=] __device__(KScal * k, double(&bv)[1]) {
                k->iter += 1;
                k->rho = k->beta;
                k->alpha = k->rho / bv[0];
            }));
            FVH_TRY(map(n, ks, [=] __device__(int i) {
                double al = ks->alpha;
                s[i] = D(r[i]) + (-al) * D(v[i]);
                xk[i] = D(xk[i]) + al * D(p[i]);
            }));
            FVH_TRY(sys_spmv(S, s, q, ks));
            FVH_TRY(sys_precond(S, t, q, ks));
            FVH_TRY(mapred<2>(n, ks, part, [=] __device__(int i, double(&a)[2]) {
                a[0] = a[0] + D(t[i]) * D(s[i]);
                a[1] = a[1] + D(t[i]) * D(t[i]);
            }));
            FVH_TRY(finalize<2>(n, part, ks, true, [=] __device__(KScal * k, double(&tt)[2]) { k->omega = tt[0] / tt[1]; }));
            FVH_TRY(mapred<2>(n, ks, part, [=] __device__(int i, double(&a)[2]) {
                double om = ks->omega;
                xk[i] = D(xk[i]) + om * D(s[i]);
                const T ri = D(s[i]) + (-om) * D(t[i]);
                r[i] = ri;
                a[0] = a[0] + D(b[i]) * D(ri);
                a[1] = a[1] + D(ri) * D(ri);
            }));
            FVH_TRY(finalize<2>(n, part, ks, true, [=] __device__(KScal * k, double(&rr)[2]) {
                double next_rho = rr[0];
                double beta = (next_rho / k->rho) * (k->alpha / k->omega);
                k->beta = next_rho;
                k->ts = beta;
                double rn = sqrt(rr[1]);
                k->rnorm = rn;
                bool solved = (rn <= k->eps) || (rn + 1.0 <= 1.0);
                bool tired = k->iter >= itmax;
                bool brk = (k->alpha == 0.0) || isnan(k->alpha);
                if (solved || tired || brk) {
                    k->done = 1;
                    k->status = solved ? FVH_KRYLOV_SOLVED : (brk ? FVH_KRYLOV_BREAKDOWN : FVH_KRYLOV_ITMAX);
                }
            }));
            FVH_TRY(map(n, ks, [=] __device__(int i) {
                double om = ks->omega, be = ks->ts;
                double pi = D(p[i]) + (-om) * D(v[i]);
                p[i] = D(r[i]) + be * pi;
            }));
        }
        enq += nb;
        FVH_TRY(read_ks(e));
    }
    if (S.warm) FVH_TRY(map(n, nullptr, [=] __device__(int i) { xu[i] = xk[i] + dx[i]; }));
    else FVH_TRY(map(n, nullptr, [=] __device__(int i) { xu[i] = xk[i]; }));
    if (enq == 0) FVH_TRY(read_ks(e));
    fill_stats(e, st);
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// CG's two scalar updates fused into the vector updates that use them.  Every
// block reduces the partial sums itself (the canonical order of k_final, so
// the same bits), computes the scalar and maps its elements with it; block 0
// alone stores the Krylov scalars.  What the blocks read from ks is not
// written in the same launch (gamma and pnorm2 in the alpha stage; gprev and
// eps in the beta stage), and what decides a block's early exit is computed
// from those, so every block takes the same path.  The `done` flag is read
// once per block, before the block-wide reduction.
// ---------------------------------------------------------------------------
__device__ inline bool block_done(const KScal *ks) {
    __shared__ int d;
    if (threadIdx.x == 0) d = ks->done;
    __syncthreads();
    return d != 0;
}
inline dim3 fused_grid(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 1024))); }

// alpha = gamma / pAp (or stop on zero curvature), then x += alpha p,
// r -= alpha Ap: `load(i)` reads an element's inputs, `op(i, v, alpha)`
// writes its results.  Each thread has at most kAlphaPer elements (the grid
// is sized for that), loaded before the reduction so that the two
// latencies overlap.
constexpr int kAlphaPer = 4;
inline dim3 alpha_grid(int64_t n) {
    return dim3((unsigned)std::max<int64_t>(1, (n + kBlock * kAlphaPer - 1) / (kBlock * kAlphaPer)));
}
template <class Ld, class Op>
__global__ void __launch_bounds__(kBlock)
    k_cg_alpha(int n, const double *part, int nch, KScal *ks, double epsT, Ld load, Op op) {
    if (block_done(ks)) return;
    using St = decltype(load(0));
    St v[kAlphaPer];
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
#pragma unroll
    for (int k = 0; k < kAlphaPer; ++k)
        if (i0 + k * stride < n) v[k] = load(i0 + k * stride);
    double pa[1];
    final_reduce<1>(part, nch, pa);
    const double pAp = pa[0], gamma = ks->gamma, pn2 = ks->pnorm2;
    const bool zc = pAp <= epsT * pn2 && fabs(pAp) <= epsT * pn2;
    const double al = gamma / pAp;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ks->gprev = gamma;
        if (zc) {
            ks->done = 1;
            ks->status = FVH_KRYLOV_ZERO_CURV;
        } else {
            ks->alpha = al;
        }
    }
    if (zc) return;
#pragma unroll
    for (int k = 0; k < kAlphaPer; ++k)
        if (i0 + k * stride < n) op(i0 + k * stride, v[k], al);
}

// beta = gamma_new / gamma (or stop: solved, itmax), then p = z + beta p
template <class Op>
__global__ void __launch_bounds__(kBlock) k_cg_beta(int n, const double *part, int nch, KScal *ks, int itmax, Op op) {
    if (block_done(ks)) return;
    double g[1];
    final_reduce<1>(part, nch, g);
    const double gn = g[0], rn = sqrt(gn);
    const bool solved = (rn <= ks->eps) || (rn + 1.0 <= 1.0);
    const double beta = gn / ks->gprev;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ks->rnorm = rn;
        if (!solved) {
            ks->pnorm2 = gn + beta * beta * ks->pnorm2;
            ks->gamma = gn;
            ks->beta = beta;
        }
        ks->iter += 1;
        const bool tired = ks->iter >= itmax;
        if (solved || tired) {
            ks->done = 1;
            ks->status = solved ? FVH_KRYLOV_SOLVED : FVH_KRYLOV_ITMAX;
        }
    }
    if (solved) return;  // at itmax p is updated once more: nothing reads it
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) op(i, beta);
}

// k_cg_beta and the next iteration's k_spmv_dot in one launch, one block of
// kWide threads per chunk: beta as k_cg_beta forms it, then p' = z + beta p
// for the chunk's rows, stored into a second array (the old p of rows in
// other blocks must stay readable), and y = sigma A p' with the partial sums
// of p'.y.  A row's neighbours' p' are formed again from z and p by the same
// expression, so the product reads the bits that the owners store.  At itmax
// the blocks still run the product: nothing reads it.
template <typename V, class Ax>
__global__ void __launch_bounds__(kWide)
    k_cg_beta_spmv(int n, const double *rz, int nch, KScal *ks, int itmax, const V *__restrict__ z,
                   const V *__restrict__ p, V *__restrict__ pn, V *__restrict__ y, V sigma, double *part, Ax ax) {
    if (block_done(ks)) return;
    double g[1];
    final_reduce<1>(rz, nch, g);
    const double gn = g[0], rn = sqrt(gn);
    const bool solved = (rn <= ks->eps) || (rn + 1.0 <= 1.0);
    const double beta = gn / ks->gprev;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ks->rnorm = rn;
        if (!solved) {
            ks->pnorm2 = gn + beta * beta * ks->pnorm2;
            ks->gamma = gn;
            ks->beta = beta;
        }
        ks->iter += 1;
        const bool tired = ks->iter >= itmax;
        if (solved || tired) {
            ks->done = 1;
            ks->status = solved ? FVH_KRYLOV_SOLVED : FVH_KRYLOV_ITMAX;
        }
    }
    if (solved) return;
    const auto pnew = [=](int j) -> V { return (V)((double)z[j] + beta * (double)p[j]); };
    __shared__ double sp[1][kChunk];
    const int c0 = blockIdx.x * kChunk;
#pragma unroll
    for (int h = 0; h < kChunk / kWide; ++h) {
        const int j = threadIdx.x + h * kWide, i = c0 + j;
        double pr = 0.0;
        if (i < n) {
            const V pi = pnew(i);
            pn[i] = pi;
            const V s = sigma * ax(pnew, i);
            y[i] = s;
            pr = (double)pi * (double)s;
        }
        sp[0][j] = pr;
    }
    chunk_reduce_staged<1>(sp, part, nch);
}

template <typename V>
struct BandRows {
    int n, nx;
    const unsigned char *mask;
    const V *C;
    template <class X>
    __device__ V operator()(X xat, int i) const {
        return band_ax_at(n, nx, mask, C, xat, i);
    }
};
template <typename V>
struct CsrRows {
    const int *rowptr, *col;
    const V *val;
    template <class X>
    __device__ V operator()(X xat, int i) const {
        return row_ax_at(rowptr, col, val, xat, i);
    }
};

// The fused beta update and SpMV with the operator the system's own SpMV
// uses (banded or CSR values, FP64 or the FP32 copy).
template <typename T>
static int launch_cg_beta_spmv(fvh_eqn *e, const double *rz, KScal *ks, int itmax, const T *z, const T *p, T *pn,
                               T *y, double sigma, double *part) {
    const fvh_mesh *m = e->m;
    const int n = (int)m->n, nch = nchunks(m->n);
    const T sg = (T)sigma;
    const T *band = sizeof(T) == 8 ? (const T *)e->kw.bval.p : (const T *)e->kw.bval32.p;
    const T *csr = sizeof(T) == 8 ? (const T *)e->nz.p : (const T *)e->kw.valf.p;
    if (e->kw.band_live && band)
        hipLaunchKernelGGL((k_cg_beta_spmv<T, BandRows<T>>), dim3(nch), dim3(kWide), 0, stream(), n, rz, nch, ks,
                           itmax, z, p, pn, y, sg, part, BandRows<T>{n, m->sched.band_nx, m->sched.bmask.p, band});
    else
        hipLaunchKernelGGL((k_cg_beta_spmv<T, CsrRows<T>>), dim3(nch), dim3(kWide), 0, stream(), n, rz, nch, ks,
                           itmax, z, p, pn, y, sg, part, CsrRows<T>{m->d.rowptr, m->d.col, csr});
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// An element's inputs to CG's alpha update (and its slot in the sweeps'
// records when r goes there).
template <typename T>
struct AlphaIn {
    T x, p, r, ap;
    int slot;
};

// ---------------------------------------------------------------------------
// CG (Krylov.jl cg!) on sigma*A x = sigma*b
// ---------------------------------------------------------------------------
template <typename T>
static int cg(const Sys<T> &S, const fvh_solver_cfg &cfg, T *xu, fvh_stats *st, double sigma) {
    fvh_eqn *e = S.e;
    const int n = S.n;
    KrylovWork &w = e->kw;
    T *xk = S.v[0], *dx = S.v[1], *r = S.v[2], *z = S.v[3], *p = S.v[4], *Ap = S.v[5], *p2 = S.v[6];
    const T *b = S.b;
    // r.z partials in part, p.Ap partials after them (the fused beta + SpMV
    // launch reads the one and writes the other)
    double *part = w.part.p, *ppap = w.part.p + nchunks(n);
    KScal *ks = w.ks.p;
    const double atol = cfg.atol, rtol = cfg.rtol;
    const int itmax = cfg.itmax;
    const double epsT = std::numeric_limits<double>::epsilon();
    FVH_TRY(single(ks, false, [=] __device__(KScal * k) {
        k->done = 0;
        k->iter = 0;
        k->status = FVH_KRYLOV_ITMAX;
        k->sigma = sigma;
    }));
    if (S.warm) {
        FVH_TRY(map(n, nullptr, [=] __device__(int i) {
            dx[i] = xu[i];
            xk[i] = 0;
        }));
        FVH_TRY(sys_spmv(S, dx, Ap, nullptr));
        FVH_TRY(map(n, nullptr, [=] __device__(int i) { r[i] = sigma * D(b[i]) - sigma * D(Ap[i]); }));
    } else {
        FVH_TRY(map(n, nullptr, [=] __device__(int i) {
            xk[i] = 0;
            r[i] = sigma * D(b[i]);
        }));
    }
    FVH_TRY(sys_precond(S, z, r, nullptr));
    FVH_TRY(mapred<1>(n, nullptr, part, [=] __device__(int i, double(&a)[1]) {
        const T zi = sigma * D(z[i]);
        z[i] = zi;
        p[i] = zi;
        a[0] = a[0] + D(r[i]) * D(zi);
    }));
    FVH_TRY(finalize<1>(n, part, ks, false, [=] __device__(KScal * k, double(&g)[1]) {
        double gamma = g[0];
        double rn = sqrt(gamma);
        k->gamma = gamma;
        k->rnorm = rn;
        k->rnorm0 = rn;
        k->pnorm2 = gamma;
        k->eps = atol + rtol * rn;
        if (gamma == 0.0) { k->done = 1; k->status = FVH_KRYLOV_ZERO_RES; return; }
        if (rn <= k->eps) { k->done = 1; k->status = FVH_KRYLOV_SOLVED; return; }
        if (itmax <= 0) { k->done = 1; k->status = FVH_KRYLOV_ITMAX; }
    }));
    int enq = 0;
    // with the banded sweeps, r goes into their layout from the alpha update
    // and z comes out of it into the r.z product: no gather or scatter pass
    const int *qmap = sys_quad_map(S);
    SkewRhs rt{};
    if (qmap) FVH_TRY(sys_rhs_target(S, &rt));
    T *bdst = static_cast<T *>(rt.dst), *hand = static_cast<T *>(rt.hand);
    const int *bmap = rt.map;
    const int bstride = rt.stride, nhand = rt.nhand, HS = rt.HS;
    const unsigned long long sent = rt.sent;
    first_batch(e);
    // Each iteration's beta update runs fused with the next iteration's SpMV
    // (k_cg_beta_spmv, p alternating between two arrays); the last one of a
    // batch runs alone, in place, before the host reads the scalars.
    const char *fuse_env = std::getenv("FVH_CG_FUSED");
    const bool fuse = !(fuse_env && fuse_env[0] == '0');
    T *pc = p, *pa = p2;
    while (!w.h_ks->done && enq < itmax) {
        int nb = std::min(next_batch(enq), itmax - enq);
        for (int it = 0; it < nb; ++it) {
            if (it > 0 && fuse) {
                FVH_TRY(launch_cg_beta_spmv<T>(e, part, ks, itmax, z, pc, pa, Ap, sigma, ppap));
                std::swap(pc, pa);
            } else {
                if (it > 0) {
                    hipLaunchKernelGGL(k_cg_beta, fused_grid(n), dim3(kBlock), 0, stream(), (int)n, part, nchunks(n),
                                       ks, itmax, [=] __device__(int i, double be) { pc[i] = D(z[i]) + be * D(pc[i]); });
                    HIP_TRY(hipGetLastError());
                }
                FVH_TRY(sys_spmv_dot(S, pc, Ap, sigma, ks, pc, ppap));
            }
            const T *pk = pc;
            const auto ld = [=] __device__(int i) { return AlphaIn<T>{xk[i], pk[i], r[i], Ap[i], qmap ? bmap[i] : 0}; };
            if (qmap)
                hipLaunchKernelGGL(k_cg_alpha, alpha_grid(n), dim3(kBlock), 0, stream(), (int)n, ppap, nchunks(n), ks,
                                   epsT, ld, [=] __device__(int i, const AlphaIn<T> &v, double al) {
                                       xk[i] = D(v.x) + al * D(v.p);
                                       const T ri = D(v.r) + (-al) * D(v.ap);
                                       r[i] = ri;
                                       bdst[(size_t)v.slot * bstride] = ri;
                                       for (int q = i; q < nhand; q += n) hand[q] = q % HS == 0 ? T(0) : bits_as<T>(sent);
                                   });
            else
                hipLaunchKernelGGL(k_cg_alpha, alpha_grid(n), dim3(kBlock), 0, stream(), (int)n, ppap, nchunks(n), ks,
                                   epsT, ld, [=] __device__(int i, const AlphaIn<T> &v, double al) {
                                       xk[i] = D(v.x) + al * D(v.p);
                                       r[i] = D(v.r) + (-al) * D(v.ap);
                                   });
            HIP_TRY(hipGetLastError());
            if (qmap) {  // r already in the sweep layout, z read straight from the sweeps' quads
                const T *zq = nullptr;
                FVH_TRY(sys_precond_quads(S, nullptr, ks, &zq));
                FVH_TRY(mapred<1>(n, ks, part, [=] __device__(int i, double(&a)[1]) {
                    const T zi = sigma * D(zq[qmap[i]]);
                    z[i] = zi;
                    a[0] = a[0] + D(r[i]) * D(zi);
                }));
            } else {
                FVH_TRY(sys_precond(S, z, r, ks));
                FVH_TRY(mapred<1>(n, ks, part, [=] __device__(int i, double(&a)[1]) {
                    const T zi = sigma * D(z[i]);
                    z[i] = zi;
                    a[0] = a[0] + D(r[i]) * D(zi);
                }));
            }
        }
        hipLaunchKernelGGL(k_cg_beta, fused_grid(n), dim3(kBlock), 0, stream(), (int)n, part, nchunks(n), ks, itmax,
                           [=] __device__(int i, double be) { pc[i] = D(z[i]) + be * D(pc[i]); });
        HIP_TRY(hipGetLastError());
        enq += nb;
        FVH_TRY(read_ks(e));
    }
    if (S.warm) FVH_TRY(map(n, nullptr, [=] __device__(int i) { xu[i] = xk[i] + dx[i]; }));
    else FVH_TRY(map(n, nullptr, [=] __device__(int i) { xu[i] = xk[i]; }));
    if (enq == 0) FVH_TRY(read_ks(e));
    fill_stats(e, st);
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// GMRES (Krylov.jl gmres!, restart = false: one pass of up to itmax MGS steps)
// ---------------------------------------------------------------------------
__device__ inline void sym_givens(double a, double b, double &c, double &s, double &rho) {
    if (b == 0.0) {
        c = (a == 0.0) ? 1.0 : dsign(a);
        s = 0.0;
        rho = fabs(a);
    } else if (a == 0.0) {
        c = 0.0;
        s = dsign(b);
        rho = fabs(b);
    } else if (fabs(b) > fabs(a)) {
        double t = a / b;
        s = dsign(b) / sqrt(1.0 + t * t);
        c = s * t;
        rho = b / s;
    } else {
        double t = b / a;
        c = dsign(a) / sqrt(1.0 + t * t);
        s = c * t;
        rho = a / c;
    }
}

template <typename T>
static int gmres(const Sys<T> &S, const fvh_solver_cfg &cfg, T *xu, fvh_stats *st) {
    fvh_eqn *e = S.e;
    const int itmax = cfg.itmax;
    const int n = S.n;
    KrylovWork &w = e->kw;
    T *xk = S.v[0], *dx = S.v[1], *wv = S.v[2], *r0 = S.v[3], *q = S.v[4];
    T *V = S.basis;
    const size_t ld = (size_t)n;
    const int cap = w.cap_basis;
    double *R = w.small.p;                         // packed upper triangle
    double *cc = R + (size_t)cap * (cap + 1) / 2;  // Givens cosines
    double *ss = cc + (cap + 2);                   // sines
    double *zz = ss + (cap + 2);                   // rhs of the least-squares problem
    double *yy = zz + (cap + 2);
    const T *b = S.b;
    double *part = w.part.p;
    KScal *ks = w.ks.p;
    const double atol = cfg.atol, rtol = cfg.rtol;
    const double btol = std::pow(std::numeric_limits<double>::epsilon(), 0.75);
    FVH_TRY(single(ks, false, [=] __device__(KScal * k) {
        k->done = 0;
        k->iter = 0;
        k->k = 0;
        k->status = FVH_KRYLOV_ITMAX;
    }));
    if (S.warm) {
        FVH_TRY(map(n, nullptr, [=] __device__(int i) {
            dx[i] = xu[i];
            xk[i] = 0;
        }));
        FVH_TRY(sys_spmv(S, dx, wv, nullptr));
        FVH_TRY(map(n, nullptr, [=] __device__(int i) { wv[i] = b[i] - wv[i]; }));
    } else {
        FVH_TRY(map(n, nullptr, [=] __device__(int i) {
            xk[i] = 0;
            wv[i] = b[i];
        }));
    }
    FVH_TRY(sys_precond(S, r0, wv, nullptr));
    FVH_TRY(mapred<1>(n, nullptr, part, [=] __device__(int i, double(&a)[1]) { a[0] = a[0] + D(r0[i]) * D(r0[i]); }));
    FVH_TRY(finalize<1>(n, part, ks, false, [=] __device__(KScal * k, double(&bb)[1]) {
        double beta = sqrt(bb[0]);
        k->rnorm = beta;
        k->rnorm0 = beta;
        k->eps = atol + rtol * beta;
        zz[0] = beta;
        if (beta == 0.0) { k->done = 1; k->status = FVH_KRYLOV_ZERO_RES; return; }
        if (beta <= k->eps) { k->done = 1; k->status = FVH_KRYLOV_SOLVED; return; }
        if (itmax <= 0) { k->done = 1; k->status = FVH_KRYLOV_ITMAX; }
    }));
    FVH_TRY(map(n, ks, [=] __device__(int i) { V[i] = D(r0[i]) / ks->rnorm0; }));
    FVH_TRY(read_ks(e));
    int enq = 0;
    while (!w.h_ks->done && enq < itmax) {
        int nb = std::min(next_batch(enq), itmax - enq);
        for (int it = 0; it < nb; ++it) {
            const int k = enq + it + 1;  // 1-based inner iteration
            const int nr = (k - 1) * k / 2;
            T *Vk1 = V + (size_t)(k - 1) * ld;
            FVH_TRY(sys_spmv(S, Vk1, wv, ks));
            FVH_TRY(sys_precond(S, q, wv, ks));
            // modified Gram-Schmidt: R[nr+i] = V_i . q ; q -= R[nr+i] V_i
            FVH_TRY(mapred<1>(n, ks, part, [=] __device__(int j, double(&a)[1]) { a[0] = a[0] + D(V[j]) * D(q[j]); }));
            for (int i = 0; i < k; ++i) {
                double *Ri = R + nr + i;
                FVH_TRY(finalize<1>(n, part, ks, true, [=] __device__(KScal *, double(&d)[1]) { *Ri = d[0]; }));
                const T *Vi = V + (size_t)i * ld;
                const T *Vn = (i + 1 < k) ? V + (size_t)(i + 1) * ld : nullptr;
                FVH_TRY(mapred<1>(n, ks, part, [=] __device__(int j, double(&a)[1]) {
                    const T qj = D(q[j]) + (-(*Ri)) * D(Vi[j]);
                    q[j] = qj;
                    a[0] = a[0] + (Vn ? D(Vn[j]) : D(qj)) * D(qj);
                }));
            }
            FVH_TRY(finalize<1>(n, part, ks, true, [=] __device__(KScal * kk, double(&h)[1]) {
                double Hbis = sqrt(h[0]);
                for (int i = 0; i < k - 1; ++i) {
                    double Rtmp = cc[i] * R[nr + i] + ss[i] * R[nr + i + 1];
                    R[nr + i + 1] = ss[i] * R[nr + i] - cc[i] * R[nr + i + 1];
                    R[nr + i] = Rtmp;
                }
                double c, s, rho;
                sym_givens(R[nr + k - 1], Hbis, c, s, rho);
                cc[k - 1] = c;
                ss[k - 1] = s;
                R[nr + k - 1] = rho;
                double zeta = s * zz[k - 1];
                zz[k - 1] = c * zz[k - 1];
                double rn = fabs(zeta);
                kk->rnorm = rn;
                kk->k = k;
                kk->hbis = Hbis;
                bool solved = (rn <= kk->eps) || (rn + 1.0 <= 1.0);
                bool brk = Hbis <= btol;
                bool tired = k >= itmax;
                if (solved || tired || brk) {
                    kk->done = 1;
                    kk->status = solved ? FVH_KRYLOV_SOLVED : (brk ? FVH_KRYLOV_BREAKDOWN : FVH_KRYLOV_ITMAX);
                } else {
                    zz[k] = zeta;
                }
            }));
            T *Vk = V + (size_t)k * ld;
            FVH_TRY(map(n, ks, [=] __device__(int i) { Vk[i] = D(q[i]) / ks->hbis; }));
        }
        enq += nb;
        FVH_TRY(read_ks(e));
    }
    // back substitution R y = z (one thread) and x = sum_i y_i V_i
    FVH_TRY(single(ks, false, [=] __device__(KScal * kk) {
        const int k = kk->k;
        const int nr = k * (k + 1) / 2;
        for (int i = 0; i < k; ++i) yy[i] = zz[i];
        for (int i = k; i >= 1; --i) {
            int pos = nr + i - k;
            for (int j = k; j >= i + 1; --j) {
                yy[i - 1] = yy[i - 1] - R[pos - 1] * yy[j - 1];
                pos = pos - j + 1;
            }
            if (fabs(R[pos - 1]) <= btol) yy[i - 1] = 0.0;
            else yy[i - 1] = yy[i - 1] / R[pos - 1];
        }
        kk->iter = k;
    }));
    const bool warm = S.warm;
    FVH_TRY(map(n, nullptr, [=] __device__(int j) {
        const int k = ks->k;
        double xj = D(xk[j]);
        for (int i = 0; i < k; ++i) xj = xj + yy[i] * D(V[(size_t)i * ld + j]);
        xu[j] = warm ? xj + D(dx[j]) : xj;
    }));
    FVH_TRY(read_ks(e));
    fill_stats(e, st);
    return FVH_OK;
}

template <typename T>
static int krylov(const Sys<T> &S, const fvh_solver_cfg &cfg, T *x, fvh_stats *st, double sigma) {
    if (cfg.solver == FVH_BICGSTAB) return bicgstab(S, cfg, x, st);
    if (cfg.solver == FVH_GMRES) return gmres(S, cfg, x, st);
    if (cfg.solver == FVH_CG) return cg(S, cfg, x, st, sigma);
    return set_error(FVH_ERR_UNSUPPORTED, "solver %d", cfg.solver);
}

// CG runs on -A x = -b when A is negative definite and M follows its sign
// (DILU / ILU(0)); the oracle does the same.
int cg_sigma(fvh_eqn *e, double *sigma) {
    *sigma = 1.0;
    const fvh_eqn *P = e->pshare ? e->pshare : e;
    if (P->pkind == FVH_PRECOND_DILU || P->pkind == FVH_PRECOND_ILU0 || P->pkind == FVH_PRECOND_LDL) {
        double a00 = 0.0;
        HIP_TRY(hipMemcpyAsync(&a00, e->nz.p + e->m->h_diag[0], sizeof(double), hipMemcpyDeviceToHost, stream()));
        HIP_TRY(hipStreamSynchronize(stream()));
        if (a00 < 0.0) *sigma = -1.0;
    }
    return FVH_OK;
}

// FP64 defect r = b - A x and its 2-norm (canonical order) on the host
static int defect64(fvh_eqn *e, const double *x, double *r, double *norm) {
    const int n = (int)e->m->n;
    KrylovWork &w = e->kw;
    const double *b = e->b.p;
    FVH_TRY(launch_spmv(e, x, r, nullptr));
    FVH_TRY(mapred<1>(n, nullptr, w.part.p, [=] __device__(int i, double(&a)[1]) {
        const double ri = b[i] - r[i];
        r[i] = ri;
        a[0] = a[0] + ri * ri;
    }));
    KScal *ks = w.ks.p;
    FVH_TRY(finalize<1>(n, w.part.p, ks, false, [=] __device__(KScal * k, double(&s)[1]) { k->tt = sqrt(s[0]); }));
    FVH_TRY(read_ks(e));
    *norm = w.h_ks->tt;
    return FVH_OK;
}

// Mixed precision: FP64 defect correction around FP32 Krylov solves.
//   r = b - A x (FP64);  A d = r solved in FP32 from d = 0 with the same
//   method, preconditioner and stopping rule;  x += d (FP64).
// When rtol >= kInnerRtol the FP32 solve meets the stopping rule by itself
// (its rounding is orders below rtol) and one correction is made, as the
// FP64 solve makes one solve; the pressure-velocity loop around it is the
// outer refinement.  A tighter rtol is beyond FP32: the inner solves then run
// to kInnerRtol and the correction repeats until the FP64 defect meets
// atol + rtol*|r0| (2-norm), stops halving, or the itmax budget is spent.
constexpr double kInnerRtol = 1e-5;
constexpr int kMaxRefine = 8;

static int solve_mixed(fvh_eqn *e, const fvh_solver_cfg &cfg, double *x, fvh_stats *st) {
    const int n = (int)e->m->n;
    const int basis = cfg.solver == FVH_GMRES ? std::max(cfg.itmax, 1) : 0;
    FVH_TRY(ensure_work32(e, basis));
    FVH_TRY(prepare32(e));
    KrylovWork &w = e->kw;
    double sigma = 1.0;
    if (cfg.solver == FVH_CG) FVH_TRY(cg_sigma(e, &sigma));
    double *r = w.v[8].p;
    float *r32 = w.f[8].p, *d32 = w.f[9].p;
    Sys<float> S{e, n, r32, false, {w.f[0].p, w.f[1].p, w.f[2].p, w.f[3].p, w.f[4].p, w.f[5].p, w.f[6].p, w.f[7].p},
                 w.basis32.p};
    const bool refine = cfg.rtol < kInnerRtol;
    double rn0 = 0.0, rn = 0.0;
    FVH_TRY(defect64(e, x, r, &rn0));
    rn = rn0;
    fvh_solver_cfg in = cfg;
    if (refine) in.rtol = kInnerRtol;
    int used = 0;
    for (int ref = 0; ref < (refine ? kMaxRefine : 1) && used < cfg.itmax; ++ref) {
        if (rn == 0.0) break;
        FVH_TRY(map(n, nullptr, [=] __device__(int i) { r32[i] = (float)r[i]; }));
        in.itmax = cfg.itmax - used;
        fvh_stats is{};
        FVH_TRY(krylov(S, in, d32, &is, sigma));
        if (is.status != FVH_KRYLOV_ZERO_RES)
            FVH_TRY(map(n, nullptr, [=] __device__(int i) { x[i] = x[i] + (double)d32[i]; }));
        used += is.niter;
        if (ref == 0) st->rnorm0 = is.rnorm0;
        st->rnorm = is.rnorm;
        st->status = is.status;
        st->refinements = ref + 1;
        if (!refine) break;
        const double rn_prev = rn;
        FVH_TRY(defect64(e, x, r, &rn));
        const bool reached = rn <= cfg.atol + cfg.rtol * rn0;
        const bool progress = rn <= 0.5 * rn_prev;
        if (reached || !progress || is.status == FVH_KRYLOV_ITMAX) break;
    }
    st->niter = used;
    return FVH_OK;
}

static int solve_any(fvh_eqn *e, const fvh_solver_cfg &cfg, double *x, fvh_stats *st) {
    fvh_stats local{};
    if (!st) st = &local;
    std::memset(st, 0, sizeof *st);
    if (cfg.solver != FVH_BICGSTAB && cfg.solver != FVH_CG && cfg.solver != FVH_GMRES)
        return set_error(FVH_ERR_UNSUPPORTED, "solver %d", cfg.solver);
    if (e->m->cases) return solve_cases(e, cfg, x, st);
    if (cfg.precision == FVH_MIXED) return solve_mixed(e, cfg, x, st);
    if (cfg.precision != FVH_FP64) return set_error(FVH_ERR_UNSUPPORTED, "precision %d", cfg.precision);
    const int basis = cfg.solver == FVH_GMRES ? std::max(cfg.itmax, 1) : 0;
    FVH_TRY(ensure_work(e, basis));
    KrylovWork &w = e->kw;
    Sys<double> S{e, (int)e->m->n, e->b.p, true,
                  {w.v[0].p, w.v[1].p, w.v[2].p, w.v[3].p, w.v[4].p, w.v[5].p, w.v[6].p, w.v[7].p}, w.basis.p};
    double sigma = 1.0;
    if (cfg.solver == FVH_CG) FVH_TRY(cg_sigma(e, &sigma));
    return krylov(S, cfg, x, st, sigma);
}

int solve(fvh_eqn *e, const fvh_solver_cfg &cfg, double *x, fvh_stats *st) {
    if (band_spmv(e->m)) {  // A is fixed for the solve: its SpMVs read it by band position
        FVH_TRY(band_pack(e));
        e->kw.band_live = true;
    }
    const int rc = solve_any(e, cfg, x, st);
    e->kw.band_live = false;
    const int fault = fault_check(e);
    return rc != FVH_OK ? rc : fault;
}

}  // namespace fvh

using namespace fvh;

extern "C" int fvh_spmv(fvh_eqn *e, const double *x, double *y) {
    if (!e || !x || !y) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(launch_spmv(e, x, y, nullptr));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

extern "C" int fvh_residual(fvh_eqn *e, const double *phi, double *res) {
    if (!e || !phi || !res) return set_error(FVH_ERR_ARG, "null argument");
    return residual(e, phi, res);
}

// ---------------------------------------------------------------------------
// residual! for callers that cannot allocate: the chunk partials and the
// result live in `scratch`, a file-scope device array of which every GPU
// holds one copy for the life of the process.  Its device address is taken
// once per device (hipGetSymbolAddress, with that device current) and only
// that address reaches kernels and copies: in host code `scratch` names a
// host-side stand-in, not device memory.
// ---------------------------------------------------------------------------
namespace fvh {
constexpr int kNoallocMaxChunks = 8192;                       // n <= 8192 * 2048 = 16,777,216 cells
constexpr int kNoallocResult = 2 * kNoallocMaxChunks;         // result word after both partial arrays
constexpr int kNoallocWords = kNoallocResult + 32;            // 128 KB + one 256-byte line
static __device__ __align__(256) double scratch[kNoallocWords];

constexpr int kNoallocMaxDevices = 64;
static double *g_scratch_ptr[kNoallocMaxDevices];

// Device address of `scratch` on the current device, looked up on the
// device's first call and kept.
static int noalloc_scratch(double **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoallocMaxDevices)
        return set_error(FVH_ERR_SIZE, "device %d beyond the %d devices of the no-allocation entry points", dev,
                         kNoallocMaxDevices);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(FVH_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return FVH_OK;
}
}  // namespace fvh

extern "C" int fvh_residual_noalloc(fvh_eqn *e, const double *phi, double *res) {
    if (!e || !phi || !res) return set_error(FVH_ERR_ARG, "null argument");
    const fvh_mesh *m = e->m;
    const int nch = nchunks(m->n);
    if (nch > kNoallocMaxChunks)
        return set_error(FVH_ERR_SIZE, "fvh_residual_noalloc: %lld cells beyond the maximum of %d",
                         (long long)m->n, kNoallocMaxChunks * kChunk);
    double *s = nullptr;
    FVH_TRY(noalloc_scratch(&s));
    // Every partial this call reads is written by this call (one per chunk,
    // per sum) before the final reduction, so nothing needs clearing.
    hipLaunchKernelGGL(k_residual_part, dim3(nch), dim3(kWide), 0, stream(), (int)m->n, m->d.rowptr, m->d.col, e->nz.p,
                       phi, e->b.p, s, nch);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(kBlock), 0, stream(), s, nch, (int)m->n, s + kNoallocResult);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(res, s + kNoallocResult, sizeof(double), hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

extern "C" int fvh_precond_set(fvh_eqn *e, int32_t kind, fvh_eqn *matrix_of) {
    FVH_GUARD_BEGIN
    if (!e) return set_error(FVH_ERR_ARG, "null equation");
    if (matrix_of && matrix_of->m != e->m) return set_error(FVH_ERR_ARG, "preconditioner matrix lives on another mesh");
    e->pkind = kind;
    e->pmat = matrix_of;
    const int64_t n = e->m->n;
    switch (kind) {
        case FVH_PRECOND_NONE: break;
        case FVH_PRECOND_JACOBI:
        case FVH_PRECOND_NORMDIAG: FVH_TRY(e->pstore.alloc(n)); break;
        case FVH_PRECOND_DILU:
            FVH_TRY(e->D.alloc(n));
            FVH_TRY(e->Dinv.alloc(n));
            break;
        case FVH_PRECOND_ILU0:
            FVH_TRY(ilu_setup(e->m));
            FVH_TRY(e->lu.alloc(e->m->nnz));
            break;
        case FVH_PRECOND_LDL:
            if (e->m->cases) return set_error(FVH_ERR_UNSUPPORTED, "LDL preconditioner: not with independent cases");
            e->ldl.ready = false;  // factored at the next update
            break;
        default: return set_error(FVH_ERR_UNSUPPORTED, "preconditioner %d", kind);
    }
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_precond_update(fvh_eqn *e) {
    if (!e) return set_error(FVH_ERR_ARG, "null equation");
    FVH_TRY(precond_update(e));
    HIP_TRY(hipStreamSynchronize(stream()));
    return fault_check(e);
}

extern "C" int fvh_precond_apply(fvh_eqn *e, double *y, const double *x) {
    if (!e || !x || !y) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(precond_apply(e, y, x, nullptr, 1.0));
    HIP_TRY(hipStreamSynchronize(stream()));
    return fault_check(e);
}

int fvh::precond_dinv_natural(fvh_eqn *e, const double **out) {
    if (e->pkind != FVH_PRECOND_DILU) return set_error(FVH_ERR_ARG, "not a DILU preconditioner");
    if (band_path(e->m)) FVH_TRY(band_dinv_natural(e, e->Dinv.p));
    *out = e->Dinv.p;
    return FVH_OK;
}

extern "C" int fvh_precond_dinv(fvh_eqn *e, double *dinv) {
    if (!e || !dinv) return set_error(FVH_ERR_ARG, "null argument");
    const double *src = nullptr;
    FVH_TRY(precond_dinv_natural(e, &src));
    HIP_TRY(hipStreamSynchronize(stream()));
    HIP_TRY(hipMemcpy(dinv, src, sizeof(double) * e->m->n, hipMemcpyDeviceToHost));
    return FVH_OK;
}

extern "C" int fvh_solve(fvh_eqn *e, const fvh_solver_cfg *cfg, double *x, fvh_stats *st) {
    FVH_GUARD_BEGIN
    if (!e || !cfg || !x) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(solve(e, *cfg, x, st));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_spmv_f32(fvh_eqn *e, float *y, const float *x) {
    FVH_GUARD_BEGIN
    if (!e || !x || !y) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(ensure_work32(e, 0));
    FVH_TRY(prepare32(e));
    FVH_TRY(launch_spmv32(e, x, y, 1.0f, nullptr, nullptr, nullptr));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_precond_apply_f32(fvh_eqn *e, float *y, const float *x) {
    FVH_GUARD_BEGIN
    if (!e || !x || !y) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(ensure_work32(e, 0));
    FVH_TRY(prepare32(e));
    FVH_TRY(precond_apply32(e, y, x, nullptr));
    HIP_TRY(hipStreamSynchronize(stream()));
    return fault_check(e);
    FVH_GUARD_END
}

extern "C" int fvh_time_op(fvh_eqn *e, int32_t op, int32_t reps, const double *x, double *y, double *ms) {
    FVH_GUARD_BEGIN
    if (!e || !ms || reps <= 0) return set_error(FVH_ERR_ARG, "bad argument");
    if ((op == 0 || op == 1) && (!x || !y)) return set_error(FVH_ERR_ARG, "null vector");
    if (op < 0 || op > 7) return set_error(FVH_ERR_UNSUPPORTED, "fvh_time_op: op %d", op);
    if ((op == 5 || op == 6) && (!x || !y)) return set_error(FVH_ERR_ARG, "null vector");
    if (op >= 6 && !band_spmv(e->m)) return set_error(FVH_ERR_UNSUPPORTED, "fvh_time_op: the mesh is not a banded grid");
    float *x32 = nullptr, *y32 = nullptr;
    if (op == 6) FVH_TRY(band_pack(e));
    if (op >= 3 && op != 5 && op != 6) {  // FP32 operators on the inner-solve copies, vectors from x rounded
        FVH_TRY(ensure_work32(e, 0));
        FVH_TRY(prepare32(e));
        x32 = e->kw.f[8].p;
        y32 = e->kw.f[9].p;
        const int n = (int)e->m->n;
        if (x) FVH_TRY(map(n, nullptr, [=] __device__(int i) { x32[i] = (float)x[i]; }));
        else FVH_TRY(map(n, nullptr, [=] __device__(int i) { x32[i] = 1.0f; }));
    }
    auto launch = [&]() -> int {
        switch (op) {
            case 0: return launch_spmv(e, x, y, nullptr);
            case 1: return precond_apply(e, y, x, nullptr, 1.0);
            case 2: return precond_update(e);
            case 3: return launch_spmv32(e, x32, y32, 1.0f, nullptr, nullptr, nullptr);
            case 4: return precond_apply32(e, y32, x32, nullptr);
            case 6:
            case 7: {  // the banded SpMV of the solves (FP64 / FP32)
                e->kw.band_live = true;
                const int rc = op == 6 ? launch_spmv(e, x, y, nullptr)
                                       : launch_spmv32(e, x32, y32, 1.0f, nullptr, nullptr, nullptr);
                e->kw.band_live = false;
                return rc;
            }
            default: {  // plain copy of ncells doubles: the reference point of traffic counters
                const double *xs = x;
                double *ys = y;
                return map(e->m->n, nullptr, [=] __device__(int i) { ys[i] = xs[i]; });
            }
        }
    };
    FVH_TRY(launch());  // warm
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, stream()));
    for (int i = 0; i < reps; ++i) FVH_TRY(launch());
    HIP_TRY(hipEventRecord(b, stream()));
    HIP_TRY(hipEventSynchronize(b));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = (double)t / reps;
    return fault_check(e);
    FVH_GUARD_END
}
