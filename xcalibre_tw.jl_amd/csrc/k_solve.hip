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
#include <limits>

#include "fvh_eqn.h"
#include "k_band.h"
#include "k_sweep.h"
#include "k_util.h"

namespace fvh {

// ---------------------------------------------------------------------------
// SpMV  y = sigma * A x  (optionally with the partial sums of w.y)
// ---------------------------------------------------------------------------
template <bool DOT>
__global__ void __launch_bounds__(kBlock) k_spmv(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                                                 const double *__restrict__ val, const double *__restrict__ x,
                                                 double *__restrict__ y, double sigma, const KScal *ks,
                                                 const double *__restrict__ w, double *part, int nch) {
    if (ks && ks->done) return;
    double acc[1] = {0.0};
    const int base = blockIdx.x * kChunk + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        int i = base + k * kBlock;
        if (i < n) {
            double s = 0.0;
            for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) s = s + val[p] * x[col[p]];
            s = sigma * s;
            y[i] = s;
            if (DOT) acc[0] = acc[0] + w[i] * s;
        }
    }
    if (DOT) chunk_reduce<1>(acc, part, nch);
}

int launch_spmv(fvh_eqn *e, const double *x, double *y, const KScal *ks) {
    const fvh_mesh *m = e->m;
    int nch = nchunks(m->n);
    hipLaunchKernelGGL((k_spmv<false>), dim3(nch), dim3(kBlock), 0, stream(), (int)m->n, m->d.rowptr, m->d.col, e->nz.p, x, y,
                       1.0, ks, (const double *)nullptr, (double *)nullptr, nch);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

static int launch_spmv_dot(fvh_eqn *e, const double *x, double *y, double sigma, const KScal *ks, const double *w,
                           double *part) {
    const fvh_mesh *m = e->m;
    int nch = nchunks(m->n);
    hipLaunchKernelGGL((k_spmv<true>), dim3(nch), dim3(kBlock), 0, stream(), (int)m->n, m->d.rowptr, m->d.col, e->nz.p, x, y,
                       sigma, ks, w, part, nch);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// Residual: sqrt(mean((Ax-b)^2)) / norm(b)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_residual_part(int n, const int *__restrict__ rowptr,
                                                          const int *__restrict__ col, const double *__restrict__ val,
                                                          const double *__restrict__ x, const double *__restrict__ b,
                                                          double *part, int nch) {
    double acc[2] = {0.0, 0.0};
    const int base = blockIdx.x * kChunk + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        int i = base + k * kBlock;
        if (i < n) {
            double s = 0.0;
            for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) s = s + val[p] * x[col[p]];
            double d = fabs(s - b[i]);
            acc[0] = acc[0] + d * d;
            acc[1] = acc[1] + b[i] * b[i];
        }
    }
    chunk_reduce<2>(acc, part, nch);
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
    hipLaunchKernelGGL(k_residual_part, dim3(nch), dim3(kBlock), 0, stream(), (int)m->n, m->d.rowptr, m->d.col, e->nz.p, phi,
                       e->b.p, e->red.p, nch);
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
            if (m->sched.band_nx > 0) return band_dilu_update(e, src->nz.p);
            return dilu_update(m, src->nz.p, e->D.p, e->Dinv.p);
        case FVH_PRECOND_ILU0: return ilu0_factor(m, src->nz.p, e->lu.p);
        default: return set_error(FVH_ERR_UNSUPPORTED, "preconditioner %d", e->pkind);
    }
}

// y = M^{-1} x
int precond_apply(fvh_eqn *eq, double *y, const double *x, const KScal *ks, double) {
    fvh_eqn *e = eq->pshare ? eq->pshare : eq;
    fvh_mesh *m = e->m;
    fvh_eqn *src = e->pmat ? e->pmat : e;
    const int n = (int)m->n;
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
            if (m->sched.band_nx > 0) return band_dilu_apply(e, x, y, ks);
            return dilu_apply(m, src->nz.p, e->Dinv.p, x, y, ks);
        case FVH_PRECOND_ILU0: return ilu0_apply(m, e->lu.p, x, y, ks);
        default: return set_error(FVH_ERR_UNSUPPORTED, "preconditioner %d", e->pkind);
    }
}

// ---------------------------------------------------------------------------
// Krylov workspace
// ---------------------------------------------------------------------------
static int ensure_work(fvh_eqn *e, int basis) {
    KrylovWork &w = e->kw;
    const int64_t n = e->m->n;
    if (w.n != n) {
        for (auto &v : w.v) FVH_TRY(v.alloc(n));
        FVH_TRY(w.part.alloc(4 * (size_t)nchunks(n) + 4));
        FVH_TRY(w.ks.alloc(1));
        if (!w.h_ks) HIP_TRY(hipHostMalloc((void **)&w.h_ks, sizeof(KScal), hipHostMallocDefault));
        w.n = n;
        w.cap_basis = 0;
    }
    if (basis > w.cap_basis) {
        FVH_TRY(w.basis.alloc((size_t)(basis + 1) * n));
        FVH_TRY(w.small.alloc((size_t)basis * (basis + 1) / 2 + 4 * (size_t)(basis + 2)));
        w.cap_basis = basis;
    }
    return FVH_OK;
}

static int read_ks(fvh_eqn *e) {
    HIP_TRY(hipMemcpyAsync(e->kw.h_ks, e->kw.ks.p, sizeof(KScal), hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

__device__ inline double dsign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }

// Iterations are enqueued in batches between host checks of ks->done.
static int next_batch(int done_iters) { return done_iters < 8 ? 2 : (done_iters < 32 ? 4 : 8); }

// ---------------------------------------------------------------------------
// BiCGSTAB (Krylov.jl bicgstab!, left preconditioning, c = b)
// ---------------------------------------------------------------------------
static int bicgstab(fvh_eqn *e, const fvh_solver_cfg &cfg, double *xu, fvh_stats *st) {
    FVH_TRY(ensure_work(e, 0));
    const int n = (int)e->m->n;
    KrylovWork &w = e->kw;
    double *xk = w.v[0].p, *dx = w.v[1].p, *r = w.v[2].p, *p = w.v[3].p, *v = w.v[4].p, *s = w.v[5].p, *q = w.v[6].p,
           *t = w.v[7].p;
    const double *b = e->b.p;
    double *part = w.part.p;
    KScal *ks = w.ks.p;
    const int nch = nchunks(n);
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
    FVH_TRY(map(n, nullptr, [=] __device__(int i) {
        dx[i] = xu[i];
        xk[i] = 0.0;
        s[i] = 0.0;
        v[i] = 0.0;
    }));
    FVH_TRY(launch_spmv(e, dx, q, nullptr));
    FVH_TRY(map(n, nullptr, [=] __device__(int i) { q[i] = b[i] - q[i]; }));
    FVH_TRY(precond_apply(e, r, q, nullptr, 1.0));
    FVH_TRY(mapred<2>(n, nullptr, part, [=] __device__(int i, double(&a)[2]) {
        p[i] = r[i];
        a[0] = a[0] + r[i] * r[i];
        a[1] = a[1] + b[i] * r[i];
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
    FVH_TRY(read_ks(e));
    while (!w.h_ks->done && enq < itmax) {
        int nb = std::min(next_batch(enq), itmax - enq);
        for (int it = 0; it < nb; ++it) {
            FVH_TRY(launch_spmv(e, p, q, ks));
            FVH_TRY(precond_apply(e, v, q, ks, 1.0));
            FVH_TRY(mapred<1>(n, ks, part, [=] __device__(int i, double(&a)[1]) { a[0] = a[0] + b[i] * v[i]; }));
            FVH_TRY(finalize<1>(n, part, ks, true, [=] __device__(KScal * k, double(&bv)[1]) {
                k->iter += 1;
                k->rho = k->beta;
                k->alpha = k->rho / bv[0];
            }));
            FVH_TRY(map(n, ks, [=] __device__(int i) {
                double al = ks->alpha;
                s[i] = r[i] + (-al) * v[i];
                xk[i] = xk[i] + al * p[i];
            }));
            FVH_TRY(launch_spmv(e, s, q, ks));
            FVH_TRY(precond_apply(e, t, q, ks, 1.0));
            FVH_TRY(mapred<2>(n, ks, part, [=] __device__(int i, double(&a)[2]) {
                a[0] = a[0] + t[i] * s[i];
                a[1] = a[1] + t[i] * t[i];
            }));
            FVH_TRY(finalize<2>(n, part, ks, true, [=] __device__(KScal * k, double(&tt)[2]) { k->omega = tt[0] / tt[1]; }));
            FVH_TRY(mapred<2>(n, ks, part, [=] __device__(int i, double(&a)[2]) {
                double om = ks->omega;
                xk[i] = xk[i] + om * s[i];
                double ri = s[i] + (-om) * t[i];
                r[i] = ri;
                a[0] = a[0] + b[i] * ri;
                a[1] = a[1] + ri * ri;
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
                double pi = p[i] + (-om) * v[i];
                p[i] = r[i] + be * pi;
            }));
        }
        enq += nb;
        FVH_TRY(read_ks(e));
    }
    FVH_TRY(map(n, nullptr, [=] __device__(int i) { xu[i] = xk[i] + dx[i]; }));
    FVH_TRY(read_ks(e));
    st->niter = w.h_ks->iter;
    st->status = w.h_ks->status;
    st->rnorm = w.h_ks->rnorm;
    st->rnorm0 = w.h_ks->rnorm0;
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// CG (Krylov.jl cg!) on sigma*A x = sigma*b
// ---------------------------------------------------------------------------
static int cg(fvh_eqn *e, const fvh_solver_cfg &cfg, double *xu, fvh_stats *st, double sigma) {
    FVH_TRY(ensure_work(e, 0));
    const int n = (int)e->m->n;
    KrylovWork &w = e->kw;
    double *xk = w.v[0].p, *dx = w.v[1].p, *r = w.v[2].p, *z = w.v[3].p, *p = w.v[4].p, *Ap = w.v[5].p;
    const double *b = e->b.p;
    double *part = w.part.p;
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
    FVH_TRY(map(n, nullptr, [=] __device__(int i) {
        dx[i] = xu[i];
        xk[i] = 0.0;
    }));
    FVH_TRY(launch_spmv(e, dx, Ap, nullptr));
    FVH_TRY(map(n, nullptr, [=] __device__(int i) { r[i] = sigma * b[i] - sigma * Ap[i]; }));
    FVH_TRY(precond_apply(e, z, r, nullptr, 1.0));
    FVH_TRY(mapred<1>(n, nullptr, part, [=] __device__(int i, double(&a)[1]) {
        double zi = sigma * z[i];
        z[i] = zi;
        p[i] = zi;
        a[0] = a[0] + r[i] * zi;
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
    FVH_TRY(read_ks(e));
    while (!w.h_ks->done && enq < itmax) {
        int nb = std::min(next_batch(enq), itmax - enq);
        for (int it = 0; it < nb; ++it) {
            FVH_TRY(launch_spmv_dot(e, p, Ap, sigma, ks, p, part));
            FVH_TRY(finalize<1>(n, part, ks, true, [=] __device__(KScal * k, double(&pa)[1]) {
                double pAp = pa[0];
                if (pAp <= epsT * k->pnorm2 && fabs(pAp) <= epsT * k->pnorm2) {
                    k->done = 1;
                    k->status = FVH_KRYLOV_ZERO_CURV;
                    return;
                }
                k->alpha = k->gamma / pAp;
            }));
            FVH_TRY(map(n, ks, [=] __device__(int i) {
                double al = ks->alpha;
                xk[i] = xk[i] + al * p[i];
                r[i] = r[i] + (-al) * Ap[i];
            }));
            FVH_TRY(precond_apply(e, z, r, ks, 1.0));
            FVH_TRY(mapred<1>(n, ks, part, [=] __device__(int i, double(&a)[1]) {
                double zi = sigma * z[i];
                z[i] = zi;
                a[0] = a[0] + r[i] * zi;
            }));
            FVH_TRY(finalize<1>(n, part, ks, true, [=] __device__(KScal * k, double(&g)[1]) {
                double gn = g[0];
                double rn = sqrt(gn);
                k->rnorm = rn;
                bool solved = (rn <= k->eps) || (rn + 1.0 <= 1.0);
                if (!solved) {
                    double beta = gn / k->gamma;
                    k->pnorm2 = gn + beta * beta * k->pnorm2;
                    k->gamma = gn;
                    k->beta = beta;
                }
                k->iter += 1;
                bool tired = k->iter >= itmax;
                if (solved || tired) {
                    k->done = 1;
                    k->status = solved ? FVH_KRYLOV_SOLVED : FVH_KRYLOV_ITMAX;
                }
            }));
            FVH_TRY(map(n, ks, [=] __device__(int i) { p[i] = z[i] + ks->beta * p[i]; }));
        }
        enq += nb;
        FVH_TRY(read_ks(e));
    }
    FVH_TRY(map(n, nullptr, [=] __device__(int i) { xu[i] = xk[i] + dx[i]; }));
    FVH_TRY(read_ks(e));
    st->niter = w.h_ks->iter;
    st->status = w.h_ks->status;
    st->rnorm = w.h_ks->rnorm;
    st->rnorm0 = w.h_ks->rnorm0;
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

static int gmres(fvh_eqn *e, const fvh_solver_cfg &cfg, double *xu, fvh_stats *st) {
    const int itmax = cfg.itmax;
    FVH_TRY(ensure_work(e, itmax > 0 ? itmax : 1));
    const int n = (int)e->m->n;
    KrylovWork &w = e->kw;
    double *xk = w.v[0].p, *dx = w.v[1].p, *wv = w.v[2].p, *r0 = w.v[3].p, *q = w.v[4].p;
    double *V = w.basis.p;
    const size_t ld = (size_t)n;
    const int cap = w.cap_basis;
    double *R = w.small.p;                         // packed upper triangle
    double *cc = R + (size_t)cap * (cap + 1) / 2;  // Givens cosines
    double *ss = cc + (cap + 2);                   // sines
    double *zz = ss + (cap + 2);                   // rhs of the least-squares problem
    double *yy = zz + (cap + 2);
    const double *b = e->b.p;
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
    FVH_TRY(map(n, nullptr, [=] __device__(int i) {
        dx[i] = xu[i];
        xk[i] = 0.0;
    }));
    FVH_TRY(launch_spmv(e, dx, wv, nullptr));
    FVH_TRY(map(n, nullptr, [=] __device__(int i) { wv[i] = b[i] - wv[i]; }));
    FVH_TRY(precond_apply(e, r0, wv, nullptr, 1.0));
    FVH_TRY(mapred<1>(n, nullptr, part, [=] __device__(int i, double(&a)[1]) { a[0] = a[0] + r0[i] * r0[i]; }));
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
    FVH_TRY(map(n, ks, [=] __device__(int i) { V[i] = r0[i] / ks->rnorm0; }));
    FVH_TRY(read_ks(e));
    int enq = 0;
    while (!w.h_ks->done && enq < itmax) {
        int nb = std::min(next_batch(enq), itmax - enq);
        for (int it = 0; it < nb; ++it) {
            const int k = enq + it + 1;  // 1-based inner iteration
            const int nr = (k - 1) * k / 2;
            double *Vk1 = V + (size_t)(k - 1) * ld;
            FVH_TRY(launch_spmv(e, Vk1, wv, ks));
            FVH_TRY(precond_apply(e, q, wv, ks, 1.0));
            // modified Gram-Schmidt: R[nr+i] = V_i . q ; q -= R[nr+i] V_i
            FVH_TRY(mapred<1>(n, ks, part, [=] __device__(int j, double(&a)[1]) { a[0] = a[0] + V[j] * q[j]; }));
            for (int i = 0; i < k; ++i) {
                double *Ri = R + nr + i;
                FVH_TRY(finalize<1>(n, part, ks, true, [=] __device__(KScal *, double(&d)[1]) { *Ri = d[0]; }));
                const double *Vi = V + (size_t)i * ld;
                const double *Vn = (i + 1 < k) ? V + (size_t)(i + 1) * ld : nullptr;
                FVH_TRY(mapred<1>(n, ks, part, [=] __device__(int j, double(&a)[1]) {
                    double qj = q[j] + (-(*Ri)) * Vi[j];
                    q[j] = qj;
                    a[0] = a[0] + (Vn ? Vn[j] : qj) * qj;
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
            double *Vk = V + (size_t)k * ld;
            FVH_TRY(map(n, ks, [=] __device__(int i) { Vk[i] = q[i] / ks->hbis; }));
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
    FVH_TRY(map(n, nullptr, [=] __device__(int j) {
        const int k = ks->k;
        double xj = xk[j];
        for (int i = 0; i < k; ++i) xj = xj + yy[i] * V[(size_t)i * ld + j];
        xu[j] = xj + dx[j];
    }));
    FVH_TRY(read_ks(e));
    st->niter = w.h_ks->iter;
    st->status = w.h_ks->status;
    st->rnorm = w.h_ks->rnorm;
    st->rnorm0 = w.h_ks->rnorm0;
    return FVH_OK;
}

int solve(fvh_eqn *e, const fvh_solver_cfg &cfg, double *x, fvh_stats *st) {
    fvh_stats local{};
    if (!st) st = &local;
    std::memset(st, 0, sizeof *st);
    if (cfg.solver == FVH_BICGSTAB) return bicgstab(e, cfg, x, st);
    if (cfg.solver == FVH_GMRES) return gmres(e, cfg, x, st);
    if (cfg.solver == FVH_CG) {
        // CG runs on -A x = -b when A is negative definite and M follows its
        // sign (DILU / ILU(0)); the oracle does the same.
        double sigma = 1.0;
        const fvh_eqn *P = e->pshare ? e->pshare : e;
        if (P->pkind == FVH_PRECOND_DILU || P->pkind == FVH_PRECOND_ILU0) {
            double a00 = 0.0;
            HIP_TRY(hipMemcpyAsync(&a00, e->nz.p + e->m->h_diag[0], sizeof(double), hipMemcpyDeviceToHost, stream()));
            HIP_TRY(hipStreamSynchronize(stream()));
            if (a00 < 0.0) sigma = -1.0;
        }
        return cg(e, cfg, x, st, sigma);
    }
    return set_error(FVH_ERR_UNSUPPORTED, "solver %d", cfg.solver);
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
        default: return set_error(FVH_ERR_UNSUPPORTED, "preconditioner %d (LDL is not on the device path)", kind);
    }
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_precond_update(fvh_eqn *e) {
    if (!e) return set_error(FVH_ERR_ARG, "null equation");
    FVH_TRY(precond_update(e));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

extern "C" int fvh_precond_apply(fvh_eqn *e, double *y, const double *x) {
    if (!e || !x || !y) return set_error(FVH_ERR_ARG, "null argument");
    FVH_TRY(precond_apply(e, y, x, nullptr, 1.0));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

int fvh::precond_dinv_natural(fvh_eqn *e, const double **out) {
    if (e->pkind != FVH_PRECOND_DILU) return set_error(FVH_ERR_ARG, "not a DILU preconditioner");
    if (e->m->sched.band_nx > 0) FVH_TRY(band_dinv_natural(e, e->Dinv.p));
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
    if (cfg->precision != FVH_FP64) return set_error(FVH_ERR_UNSUPPORTED, "fvh_solve: use FVH_FP64 (mixed precision runs inside the loop drivers)");
    FVH_TRY(solve(e, *cfg, x, st));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_time_op(fvh_eqn *e, int32_t op, int32_t reps, const double *x, double *y, double *ms) {
    FVH_GUARD_BEGIN
    if (!e || !ms || reps <= 0) return set_error(FVH_ERR_ARG, "bad argument");
    if ((op == 0 || op == 1) && (!x || !y)) return set_error(FVH_ERR_ARG, "null vector");
    auto launch = [&]() -> int {
        switch (op) {
            case 0: return launch_spmv(e, x, y, nullptr);
            case 1: return precond_apply(e, y, x, nullptr, 1.0);
            case 2: return precond_update(e);
            default: return set_error(FVH_ERR_UNSUPPORTED, "fvh_time_op: op %d", op);
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
    return FVH_OK;
    FVH_GUARD_END
}
