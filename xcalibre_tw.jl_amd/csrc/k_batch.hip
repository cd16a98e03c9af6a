// Independent cases on one mesh (batched mode).
//
// A mesh declared with fvh_mesh_set_cases is the block-diagonal union of K
// cases.  Every cell / face kernel covers all cases in one launch as it is;
// what is per case lives here:
//   * reductions: launch block b takes chunk c of case k (the table built by
//     fvh_mesh_set_cases), so each case's partial sums, and their final
//     reduction, are the ones that case has alone -- a case's FP64 results
//     in the batch are bit-identical to its single-case run;
//   * Krylov scalars: one KScal per case; a finished case's rows are skipped
//     by every later kernel of the solve and its sweep work-group exits;
//   * the DILU sweeps: one work-group per case in a single launch (k_band).
// The recurrences below are those of k_solve.hip, expression for expression.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>

#include "fvh_eqn.h"
#include "k_band.h"
#include "k_spmv.h"
#include "k_util.h"

namespace fvh {

// ---------------------------------------------------------------------------
// Per-case reductions and maps
// ---------------------------------------------------------------------------
template <int NV, class Op>
__global__ void __launch_bounds__(kBlock) k_mapred_cases(CaseView cv, const KScal *ks, double *part, Op op) {
    const int b = blockIdx.x, k = cv.blk_case[b];
    if (ks[k].done) return;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    const int base = cv.blk_row0[b] + threadIdx.x, end = cv.off[k + 1];
    const KScal *K = ks + k;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int i = base + j * kBlock;
        if (i < end) op(i, acc, K);
    }
    chunk_reduce<NV>(acc, part, cv.nblk);
}

template <int NV, class Fin>
__global__ void __launch_bounds__(kBlock) k_final_cases(CaseView cv, const double *part, KScal *ks, bool skip,
                                                        Fin fin) {
    const int k = blockIdx.x;
    if (skip && ks[k].done) return;
    double r[NV];
    final_reduce_at<NV>(part, cv.nblk, cv.cb[k], cv.cb[k + 1] - cv.cb[k], r);
    if (threadIdx.x == 0) fin(ks + k, r);
}

template <class Op>
__global__ void __launch_bounds__(kBlock) k_map_cases(int n, const int *cell_case, const KScal *ks, Op op) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const KScal *K = ks + cell_case[i];
    if (K->done) return;
    op(i, K);
}

template <class Fn>
__global__ void k_single_cases(int K, KScal *ks, bool skip, Fn fn) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K || (skip && ks[k].done)) return;
    fn(ks + k);
}

struct Cases {
    const fvh_mesh *m;
    CaseView cv;
    int n;
    template <int NV, class Op>
    int mapred(const KScal *ks, double *part, Op op) const {
        hipLaunchKernelGGL((k_mapred_cases<NV, Op>), dim3(cv.nblk), dim3(kBlock), 0, stream(), cv, ks, part, op);
        HIP_TRY(hipGetLastError());
        return FVH_OK;
    }
    template <int NV, class Fin>
    int finalize(const double *part, KScal *ks, bool skip, Fin fin) const {
        hipLaunchKernelGGL((k_final_cases<NV, Fin>), dim3(cv.K), dim3(kBlock), 0, stream(), cv, part, ks, skip, fin);
        HIP_TRY(hipGetLastError());
        return FVH_OK;
    }
    template <class Op>
    int map(const KScal *ks, Op op) const {
        hipLaunchKernelGGL((k_map_cases<Op>), grid1d(n), dim3(kBlock), 0, stream(), n, cv.cell_case, ks, op);
        HIP_TRY(hipGetLastError());
        return FVH_OK;
    }
    template <class Fn>
    int single(KScal *ks, bool skip, Fn fn) const {
        hipLaunchKernelGGL((k_single_cases<Fn>), dim3((cv.K + 63) / 64), dim3(64), 0, stream(), cv.K, ks, skip, fn);
        HIP_TRY(hipGetLastError());
        return FVH_OK;
    }
};

// ---------------------------------------------------------------------------
// SpMV and residual per case
// ---------------------------------------------------------------------------
template <typename V>
__global__ void __launch_bounds__(256) k_spmv_cases(int n, const int *__restrict__ cell_case,
                                                    const int *__restrict__ rowptr, const int *__restrict__ col,
                                                    const V *__restrict__ val, const V *__restrict__ x,
                                                    V *__restrict__ y, const KScal *ks) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (ks && ks[cell_case[i]].done)) return;
    y[i] = V(1) * row_ax(rowptr, col, val, x, i);
}

// y = sigma A x with the per-case partial sums of w.y: one kWide block per
// (case, chunk) of the reduction table
template <typename V>
__global__ void __launch_bounds__(kWide) k_spmv_dot_cases(CaseView cv, const int *__restrict__ rowptr,
                                                          const int *__restrict__ col, const V *__restrict__ val,
                                                          const V *__restrict__ x, V *__restrict__ y, V sigma,
                                                          const KScal *ks, const V *__restrict__ w, double *part) {
    const int b = blockIdx.x, k = cv.blk_case[b];
    if (ks[k].done) return;
    __shared__ double p[1][kChunk];
    const int c0 = cv.blk_row0[b], end = cv.off[k + 1];
#pragma unroll
    for (int h = 0; h < kChunk / kWide; ++h) {
        const int j = threadIdx.x + h * kWide, i = c0 + j;
        double pr = 0.0;
        if (i < end) {
            const V s = sigma * row_ax(rowptr, col, val, x, i);
            y[i] = s;
            pr = (double)w[i] * (double)s;
        }
        p[0][j] = pr;
    }
    chunk_reduce_staged<1>(p, part, cv.nblk);
}

__global__ void __launch_bounds__(kWide) k_residual_part_cases(CaseView cv, const int *__restrict__ rowptr,
                                                               const int *__restrict__ col,
                                                               const double *__restrict__ val,
                                                               const double *__restrict__ x,
                                                               const double *__restrict__ b, double *part) {
    const int blk = blockIdx.x, k = cv.blk_case[blk];
    __shared__ double p[2][kChunk];
    const int c0 = cv.blk_row0[blk], end = cv.off[k + 1];
#pragma unroll
    for (int h = 0; h < kChunk / kWide; ++h) {
        const int j = threadIdx.x + h * kWide, i = c0 + j;
        double d2 = 0.0, b2 = 0.0;
        if (i < end) {
            const double d = fabs(row_ax(rowptr, col, val, x, i) - b[i]);
            d2 = d * d;
            b2 = b[i] * b[i];
        }
        p[0][j] = d2;
        p[1][j] = b2;
    }
    chunk_reduce_staged<2>(p, part, cv.nblk);
}

__global__ void __launch_bounds__(kBlock) k_residual_final_cases(CaseView cv, const double *part, double *out,
                                                                 int stride, int comp) {
    const int k = blockIdx.x;
    double r[2];
    final_reduce_at<2>(part, cv.nblk, cv.cb[k], cv.cb[k + 1] - cv.cb[k], r);
    if (threadIdx.x == 0) out[k * stride + comp] = sqrt(r[0] / (double)(cv.off[k + 1] - cv.off[k])) / sqrt(r[1]);
}

int residual_cases_async(fvh_eqn *e, const double *phi, double *out, int stride, int comp) {
    const fvh_mesh *m = e->m;
    const CaseView cv = m->case_view();
    if (e->red.n < (size_t)(2 * cv.nblk)) FVH_TRY(e->red.alloc(2 * cv.nblk));
    hipLaunchKernelGGL(k_residual_part_cases, dim3(cv.nblk), dim3(kWide), 0, stream(), cv, m->d.rowptr, m->d.col,
                       e->nz.p, phi, e->b.p, e->red.p);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_residual_final_cases, dim3(cv.K), dim3(kBlock), 0, stream(), cv, e->red.p, out, stride, comp);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// The system seen by the per-case recurrences (as Sys<T> in k_solve.hip)
// ---------------------------------------------------------------------------
template <typename T>
struct CSys {
    fvh_eqn *e;
    Cases C;
    const T *val;  // A values (FP64 or the FP32 copy)
    const T *b;
    bool warm;
    T *v[8];
};

template <typename T>
__device__ inline double D(T x) {
    return (double)x;
}

template <typename T>
static int c_spmv(const CSys<T> &S, const T *x, T *y, const KScal *ks) {
    const fvh_mesh *m = S.e->m;
    hipLaunchKernelGGL((k_spmv_cases<T>), grid1d(S.C.n), dim3(256), 0, stream(), S.C.n, S.C.cv.cell_case, m->d.rowptr,
                       m->d.col, S.val, x, y, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
template <typename T>
static int c_spmv_dot(const CSys<T> &S, const T *x, T *y, double sigma, const KScal *ks, const T *w, double *part) {
    const fvh_mesh *m = S.e->m;
    hipLaunchKernelGGL((k_spmv_dot_cases<T>), dim3(S.C.cv.nblk), dim3(kWide), 0, stream(), S.C.cv, m->d.rowptr,
                       m->d.col, S.val, x, y, (T)sigma, ks, w, part);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
static int c_precond(const CSys<double> &S, double *y, const double *x, const KScal *ks) {
    return precond_apply(S.e, y, x, ks, 1.0);
}
static int c_precond(const CSys<float> &S, float *y, const float *x, const KScal *ks) {
    return precond_apply32(S.e, y, x, ks);
}
// With the pipelined banded DILU, CG writes r into the sweep layout in its
// alpha update and reads z from the sweeps' result quads (no gather or
// scatter pass; the hand-over arrays are reset by a launch of their own,
// since the per-case maps skip finished cases): the map of rows into the
// quads, or null
template <typename T>
static const int *c_quad_map(const CSys<T> &S) {
    const fvh_eqn *P = S.e->pshare ? S.e->pshare : S.e;
    const fvh_mesh *m = S.e->m;
    if (P->pkind != FVH_PRECOND_DILU || !band_path(m) || !skew_active(m)) return nullptr;
    return skew_quad_map(m);
}
static int c_precond_quads(const CSys<double> &S, const KScal *ks, const double **zq) {
    fvh_eqn *P = S.e->pshare ? S.e->pshare : S.e;
    FVH_TRY(skew_hand_reset(P, 0));
    return skew_apply_quads(P, nullptr, ks, zq);
}
static int c_precond_quads(const CSys<float> &S, const KScal *ks, const float **zq) {
    fvh_eqn *P = S.e->pshare ? S.e->pshare : S.e;
    FVH_TRY(skew_hand_reset(P, 1));
    return skew_apply_quads32(P, nullptr, ks, zq);
}

static int read_all(fvh_eqn *e, bool *all_done) {
    const int K = e->m->cases->K;
    HIP_TRY(hipMemcpyAsync(e->kw.h_ks, e->kw.ks.p, sizeof(KScal) * K, hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    bool d = true;
    for (int k = 0; k < K; ++k) d = d && e->kw.h_ks[k].done;
    *all_done = d;
    return FVH_OK;
}

static int next_batch_c(int done_iters) { return done_iters < 8 ? 2 : (done_iters < 32 ? 4 : 8); }

// Krylov iterations are enqueued by `body(nb)` (nb iterations, kernels only:
// no copies, memsets or synchronisation).  With graphs (the default for
// independent cases; FVH_GRAPH=0 turns them off) the body is captured once
// per solver set-up as a graph of 2 and one of 8 iterations and replayed
// until every case is done; a replay past a case's end is harmless because
// every kernel of an iteration skips finished cases and the iteration count
// is capped on the device (k->iter >= itmax ends a case).
static bool graphs_enabled() {
    const char *env = std::getenv("FVH_GRAPH");
    return !(env && env[0] == '0');
}

template <class Body>
static int replay(fvh_eqn *e, const std::string &key, int nb, Body body) {
    KrylovWork &w = e->kw;
    const std::string k = key + "/" + std::to_string(nb);
    auto it = w.graphs.find(k);
    if (it == w.graphs.end()) {
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
        const int rc = body(nb);
        const hipError_t ce = hipStreamEndCapture(stream(), &graph);
        if (rc != FVH_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HIP_TRY(ce);
        hipGraphExec_t exec = nullptr;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(ie);
        it = w.graphs.emplace(k, exec).first;
    }
    HIP_TRY(hipGraphLaunch(it->second, stream()));
    return FVH_OK;
}

template <class Body>
static int iterate_cases(fvh_eqn *e, const std::string &key, int itmax, Body body) {
    bool all = false;
    FVH_TRY(read_all(e, &all));
    int enq = 0;
    const bool graphs = graphs_enabled();
    while (!all && enq < itmax) {
        if (graphs) {
            const int nb = enq < 8 ? 2 : 8;
            FVH_TRY(replay(e, key, nb, body));
            enq += nb;
        } else {
            const int nb = std::min(next_batch_c(enq), itmax - enq);
            FVH_TRY(body(nb));
            enq += nb;
        }
        FVH_TRY(read_all(e, &all));
    }
    return FVH_OK;
}

// what the captured kernels bake in besides the workspace pointers
static std::string graph_key(const char *solver, const fvh_solver_cfg &cfg, bool fp32, double sigma) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s/%d/%d/%.17g/%.17g/%.17g/%d", solver, fp32 ? 32 : 64, cfg.itmax, cfg.atol,
                  cfg.rtol, sigma, cfg.precond);
    return buf;
}

// ---------------------------------------------------------------------------
// BiCGSTAB per case (the recurrence of bicgstab<T> in k_solve.hip)
// ---------------------------------------------------------------------------
template <typename T>
static int bicgstab_c(const CSys<T> &S, const fvh_solver_cfg &cfg, T *xu) {
    fvh_eqn *e = S.e;
    const Cases &C = S.C;
    KrylovWork &w = e->kw;
    T *xk = S.v[0], *dx = S.v[1], *r = S.v[2], *p = S.v[3], *v = S.v[4], *s = S.v[5], *q = S.v[6], *t = S.v[7];
    const T *b = S.b;
    double *part = w.part.p;
    KScal *ks = w.ks.p;
    const double atol = cfg.atol, rtol = cfg.rtol;
    const int itmax = cfg.itmax;
    FVH_TRY(C.single(ks, false, [=] __device__(KScal * k) {
        k->done = 0;
        k->iter = 0;
        k->status = FVH_KRYLOV_ITMAX;
        k->alpha = 1.0;
        k->omega = 1.0;
        k->rho = 1.0;
    }));
    if (S.warm) {
        FVH_TRY(map(C.n, nullptr, [=] __device__(int i) {
            dx[i] = xu[i];
            xk[i] = 0;
            s[i] = 0;
            v[i] = 0;
        }));
        FVH_TRY(c_spmv(S, dx, q, nullptr));
        FVH_TRY(map(C.n, nullptr, [=] __device__(int i) { q[i] = b[i] - q[i]; }));
    } else {
        FVH_TRY(map(C.n, nullptr, [=] __device__(int i) {
            xk[i] = 0;
            s[i] = 0;
            v[i] = 0;
            q[i] = b[i];
        }));
    }
    FVH_TRY(c_precond(S, r, q, nullptr));
    FVH_TRY(C.mapred<2>(ks, part, [=] __device__(int i, double(&a)[2], const KScal *) {
        p[i] = r[i];
        a[0] = a[0] + D(r[i]) * D(r[i]);
        a[1] = a[1] + D(b[i]) * D(r[i]);
    }));
    FVH_TRY(C.finalize<2>(part, ks, false, [=] __device__(KScal * k, double(&rr)[2]) {
        double rn = sqrt(rr[0]);
        k->rnorm = rn;
        k->rnorm0 = rn;
        k->eps = atol + rtol * rn;
        k->beta = rr[1];
        if (rn == 0.0) { k->done = 1; k->status = FVH_KRYLOV_ZERO_RES; return; }
        if (rr[1] == 0.0) { k->done = 1; k->status = FVH_KRYLOV_BREAKDOWN; return; }
        if (rn <= k->eps) { k->done = 1; k->status = FVH_KRYLOV_SOLVED; return; }
        if (itmax <= 0) { k->done = 1; k->status = FVH_KRYLOV_ITMAX; }
    }));
    auto body = [&](int nb) -> int {
        for (int it = 0; it < nb; ++it) {
            FVH_TRY(c_spmv(S, p, q, ks));
            FVH_TRY(c_precond(S, v, q, ks));
            FVH_TRY(C.mapred<1>(ks, part, [=] __device__(int i, double(&a)[1], const KScal *) {
                a[0] = a[0] + D(b[i]) * D(v[i]);
            }));
            FVH_TRY(C.finalize<1>(part, ks, true, [=] __device__(KScal * k, double(&bv)[1]) {
                k->iter += 1;
                k->rho = k->beta;
                k->alpha = k->rho / bv[0];
            }));
            FVH_TRY(C.map(ks, [=] __device__(int i, const KScal *K) {
                double al = K->alpha;
                s[i] = D(r[i]) + (-al) * D(v[i]);
                xk[i] = D(xk[i]) + al * D(p[i]);
            }));
            FVH_TRY(c_spmv(S, s, q, ks));
            FVH_TRY(c_precond(S, t, q, ks));
            FVH_TRY(C.mapred<2>(ks, part, [=] __device__(int i, double(&a)[2], const KScal *) {
                a[0] = a[0] + D(t[i]) * D(s[i]);
                a[1] = a[1] + D(t[i]) * D(t[i]);
            }));
            FVH_TRY(C.finalize<2>(part, ks, true,
                                  [=] __device__(KScal * k, double(&tt)[2]) { k->omega = tt[0] / tt[1]; }));
            FVH_TRY(C.mapred<2>(ks, part, [=] __device__(int i, double(&a)[2], const KScal *K) {
                double om = K->omega;
                xk[i] = D(xk[i]) + om * D(s[i]);
                const T ri = D(s[i]) + (-om) * D(t[i]);
                r[i] = ri;
                a[0] = a[0] + D(b[i]) * D(ri);
                a[1] = a[1] + D(ri) * D(ri);
            }));
            FVH_TRY(C.finalize<2>(part, ks, true, [=] __device__(KScal * k, double(&rr)[2]) {
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
            FVH_TRY(C.map(ks, [=] __device__(int i, const KScal *K) {
                double om = K->omega, be = K->ts;
                double pi = D(p[i]) + (-om) * D(v[i]);
                p[i] = D(r[i]) + be * pi;
            }));
        }
        return FVH_OK;
    };
    FVH_TRY(iterate_cases(e, graph_key("bicgstab", cfg, sizeof(T) == 4, 1.0), itmax, body));
    if (S.warm) FVH_TRY(map(C.n, nullptr, [=] __device__(int i) { xu[i] = xk[i] + dx[i]; }));
    else FVH_TRY(map(C.n, nullptr, [=] __device__(int i) { xu[i] = xk[i]; }));
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// CG per case (the recurrence of cg<T> in k_solve.hip) on sigma*A x = sigma*b
// ---------------------------------------------------------------------------
template <typename T>
static int cg_c(const CSys<T> &S, const fvh_solver_cfg &cfg, T *xu, double sigma) {
    fvh_eqn *e = S.e;
    const Cases &C = S.C;
    KrylovWork &w = e->kw;
    T *xk = S.v[0], *dx = S.v[1], *r = S.v[2], *z = S.v[3], *p = S.v[4], *Ap = S.v[5];
    const T *b = S.b;
    double *part = w.part.p;
    KScal *ks = w.ks.p;
    const double atol = cfg.atol, rtol = cfg.rtol;
    const int itmax = cfg.itmax;
    const double epsT = std::numeric_limits<double>::epsilon();
    FVH_TRY(C.single(ks, false, [=] __device__(KScal * k) {
        k->done = 0;
        k->iter = 0;
        k->status = FVH_KRYLOV_ITMAX;
        k->sigma = sigma;
    }));
    if (S.warm) {
        FVH_TRY(map(C.n, nullptr, [=] __device__(int i) {
            dx[i] = xu[i];
            xk[i] = 0;
        }));
        FVH_TRY(c_spmv(S, dx, Ap, nullptr));
        FVH_TRY(map(C.n, nullptr, [=] __device__(int i) { r[i] = sigma * D(b[i]) - sigma * D(Ap[i]); }));
    } else {
        FVH_TRY(map(C.n, nullptr, [=] __device__(int i) {
            xk[i] = 0;
            r[i] = sigma * D(b[i]);
        }));
    }
    FVH_TRY(c_precond(S, z, r, nullptr));
    FVH_TRY(C.mapred<1>(ks, part, [=] __device__(int i, double(&a)[1], const KScal *) {
        const T zi = sigma * D(z[i]);
        z[i] = zi;
        p[i] = zi;
        a[0] = a[0] + D(r[i]) * D(zi);
    }));
    FVH_TRY(C.finalize<1>(part, ks, false, [=] __device__(KScal * k, double(&g)[1]) {
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
    const int *qmap = c_quad_map(S);
    SkewRhs rt{};
    if (qmap) FVH_TRY(skew_rhs_target(e->pshare ? e->pshare : e, sizeof(T) == 4, &rt));
    T *bdst = static_cast<T *>(rt.dst);
    const int *bmap = rt.map;
    const int bstride = rt.stride;
    auto body = [&](int nb) -> int {
        for (int it = 0; it < nb; ++it) {
            FVH_TRY(c_spmv_dot(S, p, Ap, sigma, ks, p, part));
            FVH_TRY(C.finalize<1>(part, ks, true, [=] __device__(KScal * k, double(&pa)[1]) {
                double pAp = pa[0];
                if (pAp <= epsT * k->pnorm2 && fabs(pAp) <= epsT * k->pnorm2) {
                    k->done = 1;
                    k->status = FVH_KRYLOV_ZERO_CURV;
                    return;
                }
                k->alpha = k->gamma / pAp;
            }));
            if (qmap) {  // r also into the sweep layout; z from the result quads
                FVH_TRY(C.map(ks, [=] __device__(int i, const KScal *K) {
                    double al = K->alpha;
                    xk[i] = D(xk[i]) + al * D(p[i]);
                    const T ri = D(r[i]) + (-al) * D(Ap[i]);
                    r[i] = ri;
                    bdst[(size_t)bmap[i] * bstride] = ri;
                }));
                const T *zq = nullptr;
                FVH_TRY(c_precond_quads(S, ks, &zq));
                FVH_TRY(C.mapred<1>(ks, part, [=] __device__(int i, double(&a)[1], const KScal *) {
                    const T zi = sigma * D(zq[qmap[i]]);
                    z[i] = zi;
                    a[0] = a[0] + D(r[i]) * D(zi);
                }));
            } else {
                FVH_TRY(C.map(ks, [=] __device__(int i, const KScal *K) {
                    double al = K->alpha;
                    xk[i] = D(xk[i]) + al * D(p[i]);
                    r[i] = D(r[i]) + (-al) * D(Ap[i]);
                }));
                FVH_TRY(c_precond(S, z, r, ks));
                FVH_TRY(C.mapred<1>(ks, part, [=] __device__(int i, double(&a)[1], const KScal *) {
                    const T zi = sigma * D(z[i]);
                    z[i] = zi;
                    a[0] = a[0] + D(r[i]) * D(zi);
                }));
            }
            FVH_TRY(C.finalize<1>(part, ks, true, [=] __device__(KScal * k, double(&g)[1]) {
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
            FVH_TRY(C.map(ks, [=] __device__(int i, const KScal *K) { p[i] = D(z[i]) + K->beta * D(p[i]); }));
        }
        return FVH_OK;
    };
    FVH_TRY(iterate_cases(e, graph_key("cg", cfg, sizeof(T) == 4, sigma), itmax, body));
    if (S.warm) FVH_TRY(map(C.n, nullptr, [=] __device__(int i) { xu[i] = xk[i] + dx[i]; }));
    else FVH_TRY(map(C.n, nullptr, [=] __device__(int i) { xu[i] = xk[i]; }));
    return FVH_OK;
}

template <typename T>
static int krylov_c(const CSys<T> &S, const fvh_solver_cfg &cfg, T *x, double sigma) {
    if (cfg.solver == FVH_BICGSTAB) return bicgstab_c(S, cfg, x);
    if (cfg.solver == FVH_CG) return cg_c(S, cfg, x, sigma);
    return set_error(FVH_ERR_UNSUPPORTED, "independent cases: solver %d (BiCGSTAB and CG are batched)", cfg.solver);
}

static void collect_stats(fvh_eqn *e, fvh_stats *st) {
    const int K = e->m->cases->K;
    auto &cs = e->kw.case_stats;
    cs.assign(K, fvh_stats{});
    std::memset(st, 0, sizeof *st);
    for (int k = 0; k < K; ++k) {
        const KScal &h = e->kw.h_ks[k];
        cs[k].niter = h.iter;
        cs[k].status = h.status;
        cs[k].rnorm = h.rnorm;
        cs[k].rnorm0 = h.rnorm0;
        if (h.iter >= st->niter) {
            st->niter = h.iter;
            st->rnorm = h.rnorm;
            st->rnorm0 = h.rnorm0;
        }
        st->status = std::max(st->status, h.status);
    }
}

int solve_cases(fvh_eqn *e, const fvh_solver_cfg &cfg, double *x, fvh_stats *st) {
    fvh_mesh *m = e->m;
    const Cases C{m, m->case_view(), (int)m->n};
    double sigma = 1.0;
    if (cfg.solver == FVH_CG) FVH_TRY(cg_sigma(e, &sigma));
    if (cfg.precision == FVH_FP64) {
        FVH_TRY(ensure_work(e, 0));
        KrylovWork &w = e->kw;
        CSys<double> S{e, C, e->nz.p, e->b.p, true,
                       {w.v[0].p, w.v[1].p, w.v[2].p, w.v[3].p, w.v[4].p, w.v[5].p, w.v[6].p, w.v[7].p}};
        FVH_TRY(krylov_c(S, cfg, x, sigma));
        bool all = false;
        FVH_TRY(read_all(e, &all));
        collect_stats(e, st);
        return FVH_OK;
    }
    if (cfg.precision != FVH_MIXED) return set_error(FVH_ERR_UNSUPPORTED, "precision %d", cfg.precision);
    // one FP32 correction of the FP64 defect per solve (the regime the
    // single-case path uses at rtol >= 1e-5; tighter tolerances would need
    // per-case refinement counts)
    if (cfg.rtol < 1e-5)
        return set_error(FVH_ERR_UNSUPPORTED, "independent cases, mixed precision: rtol %g below 1e-5", cfg.rtol);
    FVH_TRY(ensure_work32(e, 0));
    FVH_TRY(prepare32(e));
    KrylovWork &w = e->kw;
    double *r = w.v[8].p;
    float *r32 = w.f[8].p, *d32 = w.f[9].p;
    const double *b = e->b.p;
    FVH_TRY(c_spmv(CSys<double>{e, C, e->nz.p, b, true, {}}, x, r, nullptr));
    FVH_TRY(map(C.n, nullptr, [=] __device__(int i) { r32[i] = (float)(b[i] - r[i]); }));
    CSys<float> S{e, C, w.valf.p, r32, false, {w.f[0].p, w.f[1].p, w.f[2].p, w.f[3].p, w.f[4].p, w.f[5].p, w.f[6].p, w.f[7].p}};
    FVH_TRY(krylov_c(S, cfg, d32, sigma));
    {
        // a case with a zero defect keeps x as it is (as the single-case path)
        const KScal *ks = w.ks.p;
        const int *cc = C.cv.cell_case;
        FVH_TRY(map(C.n, nullptr, [=] __device__(int i) {
            if (ks[cc[i]].status != FVH_KRYLOV_ZERO_RES) x[i] = x[i] + (double)d32[i];
        }));
    }
    bool all = false;
    FVH_TRY(read_all(e, &all));
    collect_stats(e, st);
    for (auto &c : e->kw.case_stats) c.refinements = 1;
    st->refinements = 1;
    return FVH_OK;
}

}  // namespace fvh

using namespace fvh;

extern "C" int fvh_mesh_set_cases(fvh_mesh *m, int32_t ncases, const int64_t *cell_offsets) {
    FVH_GUARD_BEGIN
    if (!m || !cell_offsets || ncases < 1) return set_error(FVH_ERR_ARG, "bad argument");
    if (cell_offsets[0] != 0 || cell_offsets[ncases] != m->n)
        return set_error(FVH_ERR_ARG, "case offsets must run from 0 to ncells");
    auto L = std::make_unique<CaseLayout>();
    L->K = ncases;
    L->h_off.assign(ncases + 1, 0);
    std::vector<int> cell_case(m->n), blk_case, blk_row0;
    L->h_cb.assign(ncases + 1, 0);
    for (int k = 0; k < ncases; ++k) {
        const int64_t a = cell_offsets[k], b = cell_offsets[k + 1];
        if (b <= a) return set_error(FVH_ERR_ARG, "case %d is empty or out of order", k);
        L->h_off[k] = (int)a;
        for (int64_t c = a; c < b; ++c) {
            cell_case[c] = k;
            for (int q = m->h_rowptr[c]; q < m->h_rowptr[c + 1]; ++q)
                if (m->h_col[q] < a || m->h_col[q] >= b)
                    return set_error(FVH_ERR_MESH, "cell %lld of case %d has a neighbour in another case", (long long)c, k);
        }
        L->h_cb[k] = (int)blk_case.size();
        for (int64_t r = a; r < b; r += kChunk) {
            blk_case.push_back(k);
            blk_row0.push_back((int)r);
        }
    }
    L->h_off[ncases] = (int)m->n;
    L->h_cb[ncases] = (int)blk_case.size();
    L->nblk = (int)blk_case.size();
    FVH_TRY(L->off.upload(L->h_off));
    FVH_TRY(L->cb.upload(L->h_cb));
    FVH_TRY(L->cell_case.upload(cell_case));
    FVH_TRY(L->blk_case.upload(blk_case));
    FVH_TRY(L->blk_row0.upload(blk_row0));
    m->cases = std::move(L);
    FVH_TRY(band_cases_setup(m));
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_mesh_case_info(const fvh_mesh *m, int32_t *ncases, int32_t *banded) {
    if (!m) return set_error(FVH_ERR_ARG, "null mesh");
    if (ncases) *ncases = m->cases ? m->cases->K : 0;
    if (banded) *banded = m->cases ? (m->cases->band ? 1 : 0) : 0;
    return FVH_OK;
}
