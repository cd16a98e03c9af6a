// Pressure-velocity drivers: the SIMPLE and PISO loop bodies
// (Solvers_1_SIMPLE.jl:69-231, Solvers_2_PISO.jl:69-238) and the k-omega
// transport step (RANSModels_kOmega.jl:114-302), on device-resident fields.
// The host only sequences kernels; it reads three residual scalars per
// iteration for the convergence break and nothing else.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <string>
#include <thread>

#include "fvh_eqn.h"
#include "k_band.h"
#include "k_util.h"

namespace fvh {

// Constants of one wall-function BC (RANSModels_kOmega.jl:200-222 and
// Discretise_0_types.jl:61-77): cmu^0.25 and y+_laminar are evaluated once.
struct WallArgs {
    int nbc;
    double kappa[kMaxBc], beta1[kMaxBc], cmu25[kMaxBc], E[kMaxBc], ylam[kMaxBc];
};

struct WallPlan {
    std::vector<fvh_bc> bcs;  // the BCs of one kind, tuple order
    BcPlan *plan = nullptr;   // per-cell entries in visiting order
    WallArgs args{};
    bool empty() const { return bcs.empty(); }
};

static int wall_plan(fvh_mesh *m, const std::vector<fvh_bc> &all, int kind, WallPlan &out) {
    out.bcs.clear();
    for (const auto &bc : all)
        if (bc.kind == kind) out.bcs.push_back(bc);
    out.plan = nullptr;
    std::memset(&out.args, 0, sizeof out.args);
    if (out.bcs.empty()) return FVH_OK;
    if ((int)out.bcs.size() > kMaxBc) return set_error(FVH_ERR_SIZE, "too many wall-function BCs");
    FVH_TRY(bc_plan(m, out.bcs.data(), (int)out.bcs.size(), &out.plan));
    out.args.nbc = (int)out.bcs.size();
    for (size_t i = 0; i < out.bcs.size(); ++i) {
        const fvh_bc &b = out.bcs[i];
        out.args.kappa[i] = b.wf[0];
        out.args.beta1[i] = b.wf[1];
        out.args.cmu25[i] = std::pow(b.wf[2], 0.25);
        out.args.E[i] = b.wf[4];
        out.args.ylam[i] = fvh_y_plus_laminar(b.wf[4], b.wf[0]);
    }
    return FVH_OK;
}

// Buffers of bound!: the bounded values, then per cell a pending flag, the
// list of cells the chain pass resolves, its length and a failure flag.
struct BoundWork {
    DBuf<double> tmp;
    DBuf<int> idx;
    int last_chained = 0;  // cells the chain pass resolved in the last call
};
int bound_field(const fvh_mesh *m, double *field, double bnd, BoundWork &w);

struct Sim {
    fvh_mesh *m = nullptr;
    fvh_case cs{};
    std::vector<fvh_bc> bU, bp, bk, bw, bnut;
    int iterations = 0, it = 0;
    bool converged = false;
    // fields
    DBuf<double> Ux, Uy, Uz, p, k, omega, nut;
    DBuf<double> mdotf, nueff, rDf, pf, Ufx, Ufy, Ufz, nutf, nueffk, nueffw, nutf_tmp;
    DBuf<double> divHv, rD, Hx, Hy, Hz, prev, gpx, gpy, gpz, Dkf, Dwf, Pk, Pw;
    BoundWork bwork;
    DBuf<double> gfx, gfy, gfz;  // face gradient of p (non-orthogonal correction)
    DBuf<double> gU[9];  // xx yx zx xy yy zy xz yz zz
    DBuf<double> res_dev;
    DBuf<int> flag;
    std::unique_ptr<fvh_eqn> ux, uy, pe, pk_hold, pw_hold;
    std::vector<double> Rux, Ruy, Rp;
    std::map<std::string, fvh_stats> stats;
    // BC tables
    BcVals vUx, vUy, vUz, vp, vk, vw, vnut;
    BcPlan *planU = nullptr, *planP = nullptr, *planK = nullptr, *planW = nullptr;
    WallPlan wK, wW, wNut;
    double ms_asm = NAN, ms_solve = NAN, ms_other = NAN;
    // independent cases (batched mode): K > 1 or a case-split mesh
    bool cases = false;
    int K = 1;
    DBuf<double> nu_face;           // viscosity of the case owning each face
    DBuf<unsigned char> refflag;    // cells taking setReference!
    std::vector<double> Rc;         // residual histories [K][iterations][3]
    // a case that reached the break keeps being computed with the others,
    // but its fields, residual history and statistics stay as they were at
    // its break (the reference's loop would have stopped there)
    std::vector<int> case_break;    // iteration of each case's break, -1 = running
    DBuf<double> snapUx, snapUy, snapP;
    std::map<std::string, std::vector<fvh_stats>> cstats;
    // the second stream of the concurrent momentum solves, and the event it
    // waits on (created at first use)
    struct Side {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        ~Side() {
            if (ev) (void)hipEventDestroy(ev);
            if (s) (void)hipStreamDestroy(s);
        }
    } side;
};

// ---------------------------------------------------------------------------
// Assembly of one transport equation
// ---------------------------------------------------------------------------
static int assemble(Sim &S, fvh_eqn *e, const std::vector<KTerm> &terms, const std::vector<std::pair<double, const double *>> &srcs,
                    const std::vector<fvh_bc> &bcs, int comp, BcPlan *plan, bool relax, double alpha, bool setref) {
    AsmArgs a;
    std::memset(&a, 0, sizeof a);
    a.mode = ASM_FULL;
    a.nterms = (int)terms.size();
    for (size_t i = 0; i < terms.size(); ++i) a.t[i] = terms[i];
    a.nsrc = (int)srcs.size();
    for (size_t i = 0; i < srcs.size(); ++i) {
        a.ssign[i] = srcs[i].first;
        a.sfield[i] = srcs[i].second;
    }
    a.prev = S.prev.p;
    a.rdt = 1.0 / S.cs.dt;
    FVH_TRY(check_bc_terms(a, bcs.data(), (int)bcs.size()));
    FVH_TRY(fill_bcvals(bcs.data(), (int)bcs.size(), comp, a.bc));
    if (plan) {
        a.bptr = plan->ptr.p;
        a.bface = plan->face.p;
        a.bbc = plan->bc.p;
    }
    a.relax = relax ? 1 : 0;
    a.alpha = alpha;
    a.rfield = S.prev.p;
    a.setref = setref ? 1 : 0;
    a.pref = S.cs.pref;
    a.refcell = 0;
    a.refflag = S.cases ? S.refflag.p : nullptr;
    return launch_assemble(e, a);
}

static inline KTerm term(int op, int scheme, double sign, const double *flux, const double *phi) {
    KTerm t;
    t.op = op;
    t.scheme = scheme;
    t.sign = sign;
    t.flux = flux;
    t.phi = phi;
    return t;
}

static int copy_dev(double *dst, const double *src, int64_t n) {
    HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, stream()));
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// Calculate-level helpers
// ---------------------------------------------------------------------------
static int velocity_faces(Sim &S, const double *cx, const double *cy, const double *cz) {
    FVH_TRY(launch_interpolate_vec(S.m, cx, cy, S.Ufx.p, S.Ufy.p));
    return launch_correct_boundaries_vec(S.m, S.bU.data(), (int)S.bU.size(), cx, cy, cz, S.Ufx.p, S.Ufy.p, S.Ufz.p);
}

static int grad_vector(Sim &S) {  // grad!(gradU, Uf, U, U.BCs), TensorField result
    fvh_mesh *m = S.m;
    double *g[9];
    for (int i = 0; i < 9; ++i) g[i] = S.gU[i].p;
    if (S.cs.sU.gradient == FVH_MIDPOINT) {
        FVH_TRY(launch_interpolate_midpoint_vec(m, S.Ux.p, S.Uy.p, S.Uz.p, S.Ufx.p, S.Ufy.p, S.Ufz.p));
        FVH_TRY(launch_correct_boundaries_vec(m, S.bU.data(), (int)S.bU.size(), S.Ux.p, S.Uy.p, S.Uz.p, S.Ufx.p, S.Ufy.p,
                                              S.Ufz.p));
        for (int it = 0; it < 2; ++it) {
            FVH_TRY(launch_correct_interpolation(m, g[0], g[1], g[2], S.Ux.p, S.Ufx.p));
            FVH_TRY(launch_green_gauss(m, S.Ufx.p, g[0], g[1], g[2]));
            FVH_TRY(launch_correct_interpolation(m, g[3], g[4], g[5], S.Uy.p, S.Ufy.p));
            FVH_TRY(launch_green_gauss(m, S.Ufy.p, g[3], g[4], g[5]));
            FVH_TRY(launch_correct_interpolation(m, g[6], g[7], g[8], S.Uz.p, S.Ufz.p));
            FVH_TRY(launch_green_gauss(m, S.Ufz.p, g[6], g[7], g[8]));
        }
        return FVH_OK;
    }
    FVH_TRY(velocity_faces(S, S.Ux.p, S.Uy.p, S.Uz.p));
    FVH_TRY(launch_green_gauss(m, S.Ufx.p, g[0], g[1], g[2]));
    FVH_TRY(launch_green_gauss(m, S.Ufy.p, g[3], g[4], g[5]));
    return launch_green_gauss(m, S.Ufz.p, g[6], g[7], g[8]);
}

static int update_nueff(Sim &S) {
    double nu = S.cs.nu;
    double *nueff = S.nueff.p;
    const double *nutf = S.nutf.p;
    if (S.cases) {  // laminar, one viscosity per case
        const double *nf = S.nu_face.p;
        return map(S.m->nf, nullptr, [=] __device__(int f) { nueff[f] = nf[f]; });
    }
    if (S.cs.model == FVH_KOMEGA) return map(S.m->nf, nullptr, [=] __device__(int f) { nueff[f] = nu + nutf[f]; });
    return map(S.m->nf, nullptr, [=] __device__(int f) { nueff[f] = nu; });
}

// ---------------------------------------------------------------------------
// k-omega  (RANSModels_kOmega.jl:114-302)
// ---------------------------------------------------------------------------
__global__ void k_wall_production(DevMesh m, WallArgs w, const int *bptr, const int *bface, const int *bbc, double nu,
                                  const double *k, const double *ux, const double *uy, const double *uz, double *P) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    for (int e = bptr[c]; e < bptr[c + 1]; ++e) {
        int f = bface[e], b = bbc[e];
        double delta = m.fdelta[f];
        double uStar = w.cmu25[b] * sqrt(k[c]);
        double dUdy = uStar / (w.kappa[b] * delta);
        double yplus = w.cmu25[b] * delta * sqrt(k[c]) / nu;
        double nutw = fvh_nut_wall(nu, yplus, w.kappa[b], w.E[b]);
        double dx = ux[c] - 0.0, dy = uy[c] - 0.0, dz = uz[c] - 0.0;
        double nx = m.nx[f], ny = m.ny[f], nz = m.nz[f];
        double un = (dx * nx + dy * ny) + dz * nz;
        double gx = (dx - un * nx) / delta, gy = (dy - un * ny) / delta, gz = (dz - un * nz) / delta;
        double mag = sqrt((gx * gx + gy * gy) + gz * gz);
        if (yplus > w.ylam[b]) P[c] = (nu + nutw) * mag * dUdy;
    }
}

// constrain_equation!: b += A_cc*wc; A_cc += A_cc once per wall face (H16)
__global__ void k_omega_constrain(DevMesh m, WallArgs w, const int *bptr, const int *bface, const int *bbc, double nu,
                                  const double *k, double *nz, double *bv) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n || bptr[c] == bptr[c + 1]) return;
    int ds = m.diag[c];
    double A = nz[ds], b = bv[c];
    for (int e = bptr[c]; e < bptr[c + 1]; ++e) {
        int f = bface[e], j = bbc[e];
        double y = m.fdelta[f];
        double wvis = 6.0 * nu / (w.beta1[j] * (y * y));
        double wlog = sqrt(k[c]) / (w.cmu25[j] * w.kappa[j] * y);
        double yplus = w.cmu25[j] * y * sqrt(k[c]) / nu;
        double wc = yplus > w.ylam[j] ? wlog : wvis;
        b += A * wc;
        A += A;
    }
    nz[ds] = A;
    bv[c] = b;
}

// set_cell_value!: omega_c = wall value (last wall face of the cell wins)
__global__ void k_omega_set_cell(DevMesh m, WallArgs w, const int *bptr, const int *bface, const int *bbc, double nu,
                                 const double *k, double *omega) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    for (int e = bptr[c]; e < bptr[c + 1]; ++e) {
        int f = bface[e], j = bbc[e];
        double y = m.fdelta[f];
        double wvis = 6.0 * nu / (w.beta1[j] * (y * y));
        double wlog = sqrt(k[c]) / (w.cmu25[j] * w.kappa[j] * y);
        double yplus = w.cmu25[j] * y * sqrt(k[c]) / nu;
        omega[c] = yplus > w.ylam[j] ? wlog : wvis;
    }
}

// correct_nut_wall!: face value of nut on NutWallFunction patches
__global__ void k_nut_wall(DevMesh m, WallArgs w, const int *bptr, const int *bface, const int *bbc, double nu,
                           const double *k, double *nutf) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    for (int e = bptr[c]; e < bptr[c + 1]; ++e) {
        int f = bface[e], j = bbc[e];
        double delta = m.fdelta[f];
        double yplus = w.cmu25[j] * delta * sqrt(k[c]) / nu;
        double nutw = fvh_nut_wall(nu, yplus, w.kappa[j], w.E[j]);
        if (yplus > w.ylam[j]) nutf[f] = nutw;
    }
}

// bound!(field, bound) (RANSModels_functions.jl:62-88) updates in place, in
// cell order, so a cell's neighbour average sees neighbours bounded earlier
// in the same pass.  Only a cell with its sign bit set reads its neighbours
// (a non-negative cell becomes max(phi, bound)), and bounding leaves a
// neighbour's value unchanged when it is non-negative and at least `bound`
// (`moves` is true otherwise, NaN included).  The parallel pass is therefore
// exact for every cell except a negative one with a lower neighbour that
// moves; it lists those cells, which the chain pass then resolves in
// dependency order.
__device__ inline bool bound_moves(double v, double bnd) { return signbit(v) || !(v >= bnd); }

__global__ void k_bound_parallel(DevMesh m, const double *__restrict__ f, double *__restrict__ out, double bnd,
                                 int *__restrict__ pend, int *__restrict__ list, int *__restrict__ count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    const double epsv = 2.220446049250313e-16;
    double sum = 0.0;
    int cnt = 0;
    bool neg = signbit(f[i]);
    bool conflict = false;
    for (int q = m.cf_ptr[i]; q < m.cf_ptr[i + 1]; ++q) {
        int j = m.cf_nbr[q];
        double v = f[j];
        if (neg && j < i && bound_moves(v, bnd)) conflict = true;
        sum += fvh_jmax(v, epsv);
        cnt += 1;
    }
    double avg = sum / (double)cnt;
    double sb = neg ? avg : copysign(0.0, avg);
    out[i] = fvh_jmax(fvh_jmax(f[i], sb), bnd);
    pend[i] = conflict ? 1 : 0;
    if (conflict) list[atomicAdd(count, 1)] = i;
}

// The listed cells, by one work-group in rounds: a cell is evaluated once
// none of its lower neighbours is still pending, reading the new value of
// every lower neighbour that moves and the old value of the others, as the
// sequential pass does.  The lowest pending cell is always ready, so every
// round resolves at least one cell: after count rounds nothing is left, and
// a further round (which cannot happen) sets *fail and ends the loop.
__global__ void __launch_bounds__(1024) k_bound_chains(DevMesh m, const double *__restrict__ f, double *out,
                                                       double bnd, int *pend, const int *__restrict__ list,
                                                       int count, int *fail) {
    __shared__ int left;
    const double epsv = 2.220446049250313e-16;
    for (int round = 0;; ++round) {
        if (round > count) {
            if (threadIdx.x == 0) *fail = 1;
            break;
        }
        if (threadIdx.x == 0) left = 0;
        __syncthreads();
        for (int q = threadIdx.x; q < count; q += blockDim.x) {
            const int i = list[q];
            if (!__hip_atomic_load(&pend[i], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) continue;
            bool ready = true;
            for (int r = m.cf_ptr[i]; r < m.cf_ptr[i + 1]; ++r) {
                const int j = m.cf_nbr[r];
                if (j < i && __hip_atomic_load(&pend[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) ready = false;
            }
            if (!ready) {
                left = 1;
                continue;
            }
            double sum = 0.0;
            int cnt = 0;
            for (int r = m.cf_ptr[i]; r < m.cf_ptr[i + 1]; ++r) {
                const int j = m.cf_nbr[r];
                const double v = (j < i && bound_moves(f[j], bnd))
                                     ? __hip_atomic_load(&out[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                                     : f[j];
                sum += fvh_jmax(v, epsv);
                cnt += 1;
            }
            const double avg = sum / (double)cnt;
            __hip_atomic_store(&out[i], fvh_jmax(fvh_jmax(f[i], avg), bnd), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(&pend[i], 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        __syncthreads();
        const int more = left;
        __syncthreads();
        if (!more) break;
    }
}

// bound!(field, bnd) on the device, in place; w holds n doubles and 2n + 2 ints.
int bound_field(const fvh_mesh *m, double *field, double bnd, BoundWork &w) {
    const int n = (int)m->n;
    if (w.tmp.n < (size_t)n) FVH_TRY(w.tmp.alloc(n));
    if (w.idx.n < 2 * (size_t)n + 2) FVH_TRY(w.idx.alloc(2 * (size_t)n + 2));
    int *pend = w.idx.p, *list = w.idx.p + n, *count = w.idx.p + 2 * (size_t)n, *fail = count + 1;
    HIP_TRY(hipMemsetAsync(count, 0, 2 * sizeof(int), stream()));
    hipLaunchKernelGGL(k_bound_parallel, grid1d(n), dim3(256), 0, stream(), m->d, field, w.tmp.p, bnd, pend, list,
                       count);
    HIP_TRY(hipGetLastError());
    int h = 0;
    HIP_TRY(hipMemcpyAsync(&h, count, sizeof(int), hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    if (h < 0 || h > n) return set_error(FVH_ERR_HIP, "bound!: %d conflicting cells of %d", h, n);
    if (h > 0) {
        hipLaunchKernelGGL(k_bound_chains, dim3(1), dim3(1024), 0, stream(), m->d, field, w.tmp.p, bnd, pend, list, h,
                           fail);
        HIP_TRY(hipGetLastError());
        int bad = 0;
        HIP_TRY(hipMemcpyAsync(&bad, fail, sizeof(int), hipMemcpyDeviceToHost, stream()));
        HIP_TRY(hipStreamSynchronize(stream()));
        if (bad) return set_error(FVH_ERR_HIP, "bound!: the ordered pass did not resolve its %d cells", h);
    }
    w.last_chained = h;
    return copy_dev(field, w.tmp.p, n);
}

static int bound(Sim &S, double *field) {
    return bound_field(S.m, field, std::numeric_limits<double>::epsilon(), S.bwork);
}

static int run_solve(Sim &S, fvh_eqn *e, const fvh_solver_cfg &cfg, double *x, const char *name) {
    fvh_stats st{};
    FVH_TRY(solve(e, cfg, x, &st));
    S.stats[name] = st;
    if (S.cases) {
        auto &cs = S.cstats[name];
        const auto &now = e->kw.case_stats;
        if (cs.size() != now.size()) cs = now;
        for (size_t k = 0; k < now.size() && k < S.case_break.size(); ++k)
            if (S.case_break[k] < 0) cs[k] = now[k];
    }
    return FVH_OK;
}

// residual!(R, eqn, phi, iteration) into slot comp of the device residuals
// (per case: res_dev[3k + comp])
static int residual_slot(Sim &S, fvh_eqn *e, const double *phi, int comp) {
    if (S.cases) return residual_cases_async(e, phi, S.res_dev.p, 3, comp);
    return residual_async(e, phi, S.res_dev.p + comp);
}

static int turbulence(Sim &S) {
    fvh_mesh *m = S.m;
    const int n = (int)m->n, nf = (int)m->nf;
    const double nu = S.cs.nu;
    const double betaP = 0.09, alpha1 = 0.52, beta1 = 0.072, sigk = 0.5, sigw = 0.5;
    {
        const double *g0 = S.gU[0].p, *g1 = S.gU[1].p, *g2 = S.gU[2].p, *g3 = S.gU[3].p, *g4 = S.gU[4].p,
                     *g5 = S.gU[5].p, *g6 = S.gU[6].p, *g7 = S.gU[7].p, *g8 = S.gU[8].p;
        double *Pk = S.Pk.p, *Pw = S.Pw.p, *Dwf = S.Dwf.p, *Dkf = S.Dkf.p;
        const double *omega = S.omega.p, *nut = S.nut.p;
        FVH_TRY(map(n, nullptr, [=] __device__(int c) {
            // M[j][k] = d U_k / d x_j  (TensorField getindex, Calculate_0_gradient.jl:60-74)
            double M[3][3] = {{g0[c], g3[c], g6[c]}, {g1[c], g4[c], g7[c]}, {g2[c], g5[c], g8[c]}};
            double sum = 0.0;
            for (int j = 0; j < 3; ++j)
                for (int kk = 0; kk < 3; ++kk) {
                    double s = 0.5 * (M[j][kk] + M[kk][j]);
                    sum += s * s;
                }
            double pk = sum * 2.0;
            Pw[c] = alpha1 * pk;
            Dwf[c] = beta1 * omega[c];
            Dkf[c] = betaP * omega[c];
            Pk[c] = nut[c] * pk;
        }));
        double *nw = S.nueffw.p, *nk = S.nueffk.p;
        const double *nutf = S.nutf.p;
        FVH_TRY(map(nf, nullptr, [=] __device__(int f) {
            nw[f] = nu + sigw * nutf[f];
            nk[f] = nu + sigk * nutf[f];
        }));
    }
    if (!S.wK.empty()) {
        hipLaunchKernelGGL(k_wall_production, grid1d(n), dim3(256), 0, stream(), m->d, S.wK.args, S.wK.plan->ptr.p,
                           S.wK.plan->face.p, S.wK.plan->bc.p, nu, S.k.p, S.Ux.p, S.Uy.p, S.Uz.p, S.Pk.p);
        HIP_TRY(hipGetLastError());
    }
    // omega equation
    FVH_TRY(copy_dev(S.prev.p, S.omega.p, n));
    {
        std::vector<KTerm> t = {term(FVH_OP_TIME, S.cs.somega.time, 1, nullptr, S.omega.p),
                                term(FVH_OP_DIVERGENCE, S.cs.somega.divergence, 1, S.mdotf.p, S.omega.p),
                                term(FVH_OP_LAPLACIAN, S.cs.somega.laplacian, -1, S.nueffw.p, S.omega.p),
                                term(FVH_OP_SI, 0, 1, S.Dwf.p, S.omega.p)};
        FVH_TRY(assemble(S, S.pe.get(), t, {{1.0, S.Pw.p}}, S.bw, 0, S.planW, true, S.cs.vomega.relax, false));
    }
    if (!S.wW.empty()) {
        hipLaunchKernelGGL(k_omega_constrain, grid1d(n), dim3(256), 0, stream(), m->d, S.wW.args, S.wW.plan->ptr.p,
                           S.wW.plan->face.p, S.wW.plan->bc.p, nu, S.k.p, S.pe->nz.p, S.pe->b.p);
        HIP_TRY(hipGetLastError());
    }
    S.pe->pshare = S.pw_hold.get();
    FVH_TRY(precond_update(S.pe.get()));
    FVH_TRY(run_solve(S, S.pe.get(), S.cs.vomega, S.omega.p, "omega"));
    S.pe->pshare = nullptr;
    if (!S.wW.empty()) {
        hipLaunchKernelGGL(k_omega_set_cell, grid1d(n), dim3(256), 0, stream(), m->d, S.wW.args, S.wW.plan->ptr.p,
                           S.wW.plan->face.p, S.wW.plan->bc.p, nu, S.k.p, S.omega.p);
        HIP_TRY(hipGetLastError());
    }
    FVH_TRY(bound(S, S.omega.p));
    // k equation
    FVH_TRY(copy_dev(S.prev.p, S.k.p, n));
    {
        std::vector<KTerm> t = {term(FVH_OP_TIME, S.cs.sk.time, 1, nullptr, S.k.p),
                                term(FVH_OP_DIVERGENCE, S.cs.sk.divergence, 1, S.mdotf.p, S.k.p),
                                term(FVH_OP_LAPLACIAN, S.cs.sk.laplacian, -1, S.nueffk.p, S.k.p),
                                term(FVH_OP_SI, 0, 1, S.Dkf.p, S.k.p)};
        FVH_TRY(assemble(S, S.pe.get(), t, {{1.0, S.Pk.p}}, S.bk, 0, S.planK, true, S.cs.vk.relax, false));
    }
    S.pe->pshare = S.pk_hold.get();
    FVH_TRY(precond_update(S.pe.get()));
    FVH_TRY(run_solve(S, S.pe.get(), S.cs.vk, S.k.p, "k"));
    S.pe->pshare = nullptr;
    FVH_TRY(bound(S, S.k.p));
    {
        double *nut = S.nut.p;
        const double *k = S.k.p, *w = S.omega.p;
        FVH_TRY(map(n, nullptr, [=] __device__(int c) { nut[c] = k[c] / w[c]; }));
    }
    FVH_TRY(launch_interpolate(m, S.nut.p, S.nutf.p));
    FVH_TRY(launch_correct_boundaries(m, S.vnut, S.nut.p, S.nutf.p));
    if (!S.wNut.empty()) {
        hipLaunchKernelGGL(k_nut_wall, grid1d(n), dim3(256), 0, stream(), m->d, S.wNut.args, S.wNut.plan->ptr.p,
                           S.wNut.plan->face.p, S.wNut.plan->bc.p, nu, S.k.p, S.nutf.p);
        HIP_TRY(hipGetLastError());
    }
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// Loop body
// ---------------------------------------------------------------------------
static int momentum_assemble(Sim &S, fvh_eqn *e, double *U, const double *gp, int comp, bool relax) {
    const int n = (int)S.m->n;
    FVH_TRY(copy_dev(S.prev.p, U, n));
    std::vector<KTerm> t = {term(FVH_OP_TIME, S.cs.sU.time, 1, nullptr, U),
                            term(FVH_OP_DIVERGENCE, S.cs.sU.divergence, 1, S.mdotf.p, U),
                            term(FVH_OP_LAPLACIAN, S.cs.sU.laplacian, -1, S.nueff.p, U)};
    return assemble(S, e, t, {{-1.0, gp}}, S.bU, comp, S.planU, relax, S.cs.vU.relax, false);
}

static int momentum(Sim &S, fvh_eqn *e, double *U, const double *gp, int comp, bool relax, const char *name,
                    bool update_pre) {
    FVH_TRY(momentum_assemble(S, e, U, gp, comp, relax));
    if (update_pre) FVH_TRY(precond_update(e));
    FVH_TRY(run_solve(S, e, S.cs.vU, U, name));
    return residual_slot(S, e, U, comp);
}

static int pressure_corrector(Sim &S) {
    fvh_mesh *m = S.m;
    const int n = (int)m->n;
    {  // H!(Hv, U, ux_eqn, uy_eqn)
        const DevMesh d = m->d;
        const double *ax = S.ux->nz.p, *ay = S.uy->nz.p, *bx = S.ux->b.p, *by = S.uy->b.p, *vx = S.Ux.p, *vy = S.Uy.p;
        double *hx = S.Hx.p, *hy = S.Hy.p, *hz = S.Hz.p;
        FVH_TRY(map(n, nullptr, [=] __device__(int c) {
            double sx = 0.0, sy = 0.0;
            // four neighbours at a time: index loads, then value loads, then
            // the sums in faces_range order
            constexpr int G = 4;
            const int i1 = d.cf_ptr[c + 1];
            for (int i = d.cf_ptr[c]; i < i1; i += G) {
                int nb[G], sl[G];
                double a1[G], a2[G], u1[G], u2[G];
_Pragma("unroll")
                for (int k = 0; k < G; ++k)
                    if (i + k < i1) {
                        nb[k] = d.cf_nbr[i + k];
                        sl[k] = d.cf_slot[i + k];
                    }
_Pragma("unroll")
                for (int k = 0; k < G; ++k)
                    if (i + k < i1) {
                        a1[k] = ax[sl[k]];
                        u1[k] = vx[nb[k]];
                        a2[k] = ay[sl[k]];
                        u2[k] = vy[nb[k]];
                    }
_Pragma("unroll")
                for (int k = 0; k < G; ++k)
                    if (i + k < i1) {
                        sx += a1[k] * u1[k];
                        sy += a2[k] * u2[k];
                    }
            }
            double rD = 1.0 / ax[d.diag[c]];
            hx[c] = (bx[c] - sx) * rD;
            hy[c] = (by[c] - sy) * rD;
            hz[c] = 0.0;
        }));
    }
    FVH_TRY(velocity_faces(S, S.Hx.p, S.Hy.p, S.Hz.p));
    FVH_TRY(launch_div(m, S.Ufx.p, S.Ufy.p, S.Ufz.p, S.divHv.p));
    FVH_TRY(copy_dev(S.prev.p, S.p.p, n));
    {
        std::vector<KTerm> t = {term(FVH_OP_LAPLACIAN, S.cs.sp.laplacian, 1, S.rDf.p, S.p.p)};
        FVH_TRY(assemble(S, S.pe.get(), t, {{1.0, S.divHv.p}}, S.bp, 0, S.planP, false, 1.0, S.cs.has_pref != 0));
    }
    FVH_TRY(precond_update(S.pe.get()));
    FVH_TRY(run_solve(S, S.pe.get(), S.cs.vp, S.p.p, "p"));
    {
        double *p = S.p.p;
        const double *p0 = S.prev.p;
        double a = S.cs.vp.relax;
        FVH_TRY(map(n, nullptr, [=] __device__(int i) { p[i] = p0[i] + a * (p[i] - p0[i]); }));
    }
    FVH_TRY(residual_slot(S, S.pe.get(), S.p.p, 2));
    FVH_TRY(launch_grad(m, S.cs.sp.gradient, S.vp, S.p.p, S.pf.p, S.gpx.p, S.gpy.p, S.gpz.p));
    // the loops' `correct` block (Solvers_1_SIMPLE.jl:165-180, Solvers_2_PISO.jl:170-182),
    // off unless nonorth_correctors > 0; pf carries the correction flux
    for (int i = 0; i < S.cs.nonorth_correctors; ++i) {
        std::vector<KTerm> t = {term(FVH_OP_LAPLACIAN, S.cs.sp.laplacian, 1, S.rDf.p, S.p.p)};
        FVH_TRY(assemble(S, S.pe.get(), t, {{1.0, S.divHv.p}}, S.bp, 0, S.planP, false, 1.0, S.cs.has_pref != 0));
        FVH_TRY(launch_interpolate_gradient(m, S.gpx.p, S.gpy.p, S.gpz.p, S.p.p, S.gfx.p, S.gfy.p, S.gfz.p));
        FVH_TRY(launch_nonorthogonal_flux(m, S.gfx.p, S.gfy.p, S.gfz.p, S.pf.p));
        FVH_TRY(launch_nonorthogonal_correct(m, 1, S.rDf.p, S.pf.p, S.pe->b.p));
        FVH_TRY(run_solve(S, S.pe.get(), S.cs.vp, S.p.p, "p"));
        FVH_TRY(launch_grad(m, S.cs.sp.gradient, S.vp, S.p.p, S.pf.p, S.gpx.p, S.gpy.p, S.gpz.p));
    }
    {
        double *ux = S.Ux.p, *uy = S.Uy.p;
        const double *hx = S.Hx.p, *hy = S.Hy.p, *gx = S.gpx.p, *gy = S.gpy.p, *rD = S.rD.p;
        FVH_TRY(map(n, nullptr, [=] __device__(int i) {
            double r = rD[i];
            ux[i] = hx[i] - gx[i] * r;
            uy[i] = hy[i] - gy[i] * r;
        }));
    }
    FVH_TRY(velocity_faces(S, S.Ux.p, S.Uy.p, S.Uz.p));
    FVH_TRY(launch_flux(m, S.Ufx.p, S.Ufy.p, S.Ufz.p, S.mdotf.p));
    if (S.cs.model == FVH_KOMEGA) {
        FVH_TRY(grad_vector(S));
        FVH_TRY(turbulence(S));
        FVH_TRY(update_nueff(S));
    }
    return FVH_OK;
}

// The two momentum solves run at once when the U preconditioner is DILU on
// the pipelined sweeps (one case; FVH_U_PAIR=0 turns this off).
static bool u_pair(const Sim &S) {
    const char *env = std::getenv("FVH_U_PAIR");
    return !(env && env[0] == '0') && !S.cases && S.cs.vU.precond == FVH_PRECOND_DILU && band_path(S.m) &&
           skew_active(S.m);
}

// Both momentum equations are assembled (neither reads the other's
// solution), Ux's preconditioner is updated and Uy's records are copied from
// it: Uy's preconditioner is Ux's (Solvers_1_SIMPLE.jl:45), and the update
// before the Uy solve (:141) recomputes it from Ux's matrix, unchanged since
// the update before the Ux solve (:132).  The Uy solve and residual then run
// on a second stream from a second thread while this thread runs Ux's: the
// solves share no buffer that either writes, so their results are those of
// the reference's order.  The first mixed-precision solve on a mesh times
// its FP32 sweep widths and runs alone.
// The Uy half of the pair, on the calling thread's stream (exceptions become
// error codes here: none may leave the thread).
static int uy_side(Sim &S, fvh_eqn *ey, fvh_stats *sy) {
    FVH_GUARD_BEGIN
    HIP_TRY(hipStreamWaitEvent(stream(), S.side.ev, 0));
    FVH_TRY(solve(ey, S.cs.vU, S.Uy.p, sy));
    return residual_slot(S, ey, S.Uy.p, 1);
    FVH_GUARD_END
}

static int momentum_pair(Sim &S, bool piso) {
    fvh_eqn *ex = S.ux.get(), *ey = S.uy.get();
    FVH_TRY(momentum_assemble(S, ex, S.Ux.p, S.gpx.p, 0, !piso));
    FVH_TRY(precond_update(ex));
    FVH_TRY(momentum_assemble(S, ey, S.Uy.p, S.gpy.p, 1, !piso));
    FVH_TRY(skew_copy(ey, ex));
    if (S.cs.vU.precision != FVH_FP64 && !skew_tuned32(S.m)) {
        FVH_TRY(run_solve(S, ex, S.cs.vU, S.Ux.p, "Ux"));
        FVH_TRY(residual_slot(S, ex, S.Ux.p, 0));
        FVH_TRY(run_solve(S, ey, S.cs.vU, S.Uy.p, "Uy"));
        return residual_slot(S, ey, S.Uy.p, 1);
    }
    if (!S.side.s) HIP_TRY(hipStreamCreateWithFlags(&S.side.s, hipStreamNonBlocking));
    if (!S.side.ev) HIP_TRY(hipEventCreateWithFlags(&S.side.ev, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(S.side.ev, stream()));
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    fvh_stats sx{}, sy{};
    int rc_y = FVH_OK;
    std::string err_y;
    std::thread th([&] {
        set_thread_stream(S.side.s);
        // the new thread's allocations go to this thread's device
        rc_y = hipSetDevice(dev) == hipSuccess ? uy_side(S, ey, &sy) : set_error(FVH_ERR_HIP, "hipSetDevice failed");
        const hipError_t se = hipStreamSynchronize(S.side.s);  // also after a failure: nothing left in flight
        if (rc_y == FVH_OK && se != hipSuccess) rc_y = set_error(FVH_ERR_HIP, "Uy stream: %s", hipGetErrorString(se));
        if (rc_y != FVH_OK) err_y = last_error_text();
        set_thread_stream(nullptr);
    });
    int rc_x = FVH_OK;
    try {  // the thread is joined on every path
        rc_x = solve(ex, S.cs.vU, S.Ux.p, &sx);
        if (rc_x == FVH_OK) rc_x = residual_slot(S, ex, S.Ux.p, 0);
    } catch (...) {
        th.join();
        throw;
    }
    th.join();
    S.stats["Ux"] = sx;
    S.stats["Uy"] = sy;
    if (rc_x != FVH_OK) return rc_x;
    if (rc_y != FVH_OK) return set_error(rc_y, "%s", err_y.c_str());
    return FVH_OK;
}

static int iterate_once(Sim &S) {
    fvh_mesh *m = S.m;
    const int n = (int)m->n;
    const int iter = S.it;
    const bool piso = S.cs.algorithm == FVH_PISO;
    if (u_pair(S)) {
        S.uy->pshare = nullptr;  // Uy applies its copy of Ux's records
        FVH_TRY(momentum_pair(S, piso));
    } else {
        S.uy->pshare = S.ux.get();
        FVH_TRY(momentum(S, S.ux.get(), S.Ux.p, S.gpx.p, 0, !piso, "Ux", true));
        // Uy's preconditioner is Ux's (Solvers_1_SIMPLE.jl:45), so the update
        // before the Uy solve (:141) recomputes it from Ux's matrix, which has
        // not changed since the update before the Ux solve (:132): the
        // recomputation would store the same values, and is skipped.
        FVH_TRY(momentum(S, S.uy.get(), S.Uy.p, S.gpy.p, 1, !piso, "Uy", false));
    }
    {
        const double *vol = m->d.vol, *ax = S.ux->nz.p, *gx = S.gpx.p, *gy = S.gpy.p;
        const int *diag = m->d.diag;
        double *rD = S.rD.p, *bx = S.ux->b.p, *by = S.uy->b.p;
        FVH_TRY(map(n, nullptr, [=] __device__(int c) {
            double v = vol[c];
            rD[c] = v / ax[diag[c]];                // inverse_diagonal!
            bx[c] -= -1.0 * gx[c] * v;              // remove_pressure_source!
            by[c] -= -1.0 * gy[c] * v;
        }));
    }
    FVH_TRY(launch_interpolate(m, S.rD.p, S.rDf.p));
    const int ncorr = piso ? 2 : 1;
    for (int i = 0; i < ncorr; ++i) FVH_TRY(pressure_corrector(S));
    const double cb = S.cs.conv_break;
    if (S.cases) {
        // every case advances each iteration; the loop stops once all have
        // reached the break
        std::vector<double> r(3 * S.K);
        HIP_TRY(hipMemcpyAsync(r.data(), S.res_dev.p, r.size() * sizeof(double), hipMemcpyDeviceToHost, stream()));
        HIP_TRY(hipStreamSynchronize(stream()));
        bool all = !piso;
        const std::vector<int> &off = m->cases->h_off;
        for (int k = 0; k < S.K; ++k) {
            if (S.case_break[k] >= 0) continue;  // frozen at its break
            for (int q = 0; q < 3; ++q) S.Rc[((size_t)k * S.iterations + iter) * 3 + q] = r[3 * k + q];
            const bool brk = !piso && r[3 * k] <= cb && r[3 * k + 1] <= cb && r[3 * k + 2] <= cb;
            if (brk) {
                S.case_break[k] = iter;
                const size_t o = (size_t)off[k], len = (size_t)(off[k + 1] - off[k]);
                HIP_TRY(hipMemcpyAsync(S.snapUx.p + o, S.Ux.p + o, len * sizeof(double), hipMemcpyDeviceToDevice, stream()));
                HIP_TRY(hipMemcpyAsync(S.snapUy.p + o, S.Uy.p + o, len * sizeof(double), hipMemcpyDeviceToDevice, stream()));
                HIP_TRY(hipMemcpyAsync(S.snapP.p + o, S.p.p + o, len * sizeof(double), hipMemcpyDeviceToDevice, stream()));
            }
        }
        for (int k = 0; k < S.K; ++k) all = all && S.case_break[k] >= 0;
        S.Rux[iter] = r[0];
        S.Ruy[iter] = r[1];
        S.Rp[iter] = r[2];
        S.it++;
        if (all) S.converged = true;
        return FVH_OK;
    }
    double r[3];
    HIP_TRY(hipMemcpyAsync(r, S.res_dev.p, 3 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    HIP_TRY(hipStreamSynchronize(stream()));
    S.Rux[iter] = r[0];
    S.Ruy[iter] = r[1];
    S.Rp[iter] = r[2];
    S.it++;
    if (!piso && r[0] <= cb && r[1] <= cb && r[2] <= cb) S.converged = true;
    return FVH_OK;
}

static int sim_init(Sim &S, const double *Ux, const double *Uy, const double *Uz, const double *p, const double *k,
                    const double *omega, const double *nut) {
    fvh_mesh *m = S.m;
    const int64_t n = m->n, nf = m->nf;
    auto cell = [&](DBuf<double> &b, const double *src, double def) -> int {
        std::vector<double> h(n, def);
        if (src) std::copy(src, src + n, h.begin());
        return b.upload(h);
    };
    auto face = [&](DBuf<double> &b, double def) -> int { return b.upload(std::vector<double>(nf, def)); };
    FVH_TRY(cell(S.Ux, Ux, 0.0));
    FVH_TRY(cell(S.Uy, Uy, 0.0));
    FVH_TRY(cell(S.Uz, Uz, 0.0));
    FVH_TRY(cell(S.p, p, 0.0));
    FVH_TRY(cell(S.k, k, 0.0));
    FVH_TRY(cell(S.omega, omega, 0.0));
    FVH_TRY(cell(S.nut, nut, 0.0));
    for (DBuf<double> *b : {&S.divHv, &S.rD, &S.Hx, &S.Hy, &S.Hz, &S.prev, &S.gpx, &S.gpy, &S.gpz, &S.Dkf, &S.Dwf, &S.Pk,
                            &S.Pw})
        FVH_TRY(cell(*b, nullptr, 0.0));
    for (auto &g : S.gU) FVH_TRY(cell(g, nullptr, 0.0));
    for (DBuf<double> *b : {&S.mdotf, &S.nueff, &S.pf, &S.Ufx, &S.Ufy, &S.Ufz, &S.nutf, &S.nueffk, &S.nueffw, &S.nutf_tmp})
        FVH_TRY(face(*b, 0.0));
    FVH_TRY(face(S.rDf, 1.0));
    FVH_TRY(S.res_dev.alloc(S.cases ? 3 * S.K + 1 : 4));
    FVH_TRY(S.flag.alloc(1));
    if (S.cs.nonorth_correctors < 0 || S.cs.nonorth_correctors > 8)
        return set_error(FVH_ERR_ARG, "nonorth_correctors must be 0..8 (got %d)", S.cs.nonorth_correctors);
    if (S.cs.nonorth_correctors > 0) {
        if (S.cases) return set_error(FVH_ERR_UNSUPPORTED, "independent cases: no non-orthogonal correctors");
        for (auto *b : {&S.gfx, &S.gfy, &S.gfz}) FVH_TRY(face(*b, 0.0));
    }
    // equations and preconditioners (uy shares ux's preconditioner, which
    // reads ux's matrix; k and omega assemble into p's equation)
    fvh_eqn *raw = nullptr;
    FVH_TRY(fvh_eqn_create(m, &raw));
    S.ux.reset(raw);
    FVH_TRY(fvh_eqn_create(m, &raw));
    S.uy.reset(raw);
    FVH_TRY(fvh_eqn_create(m, &raw));
    S.pe.reset(raw);
    FVH_TRY(fvh_precond_set(S.ux.get(), S.cs.vU.precond, nullptr));
    // Uy borrows Ux's preconditioner, or, when the two solves run at once,
    // applies a copy of its records with buffers of its own
    if (S.cs.vU.precond == FVH_PRECOND_DILU) FVH_TRY(fvh_precond_set(S.uy.get(), FVH_PRECOND_DILU, nullptr));
    S.uy->pshare = S.ux.get();
    FVH_TRY(fvh_precond_set(S.pe.get(), S.cs.vp.precond, nullptr));
    if (S.cs.model == FVH_KOMEGA) {
        FVH_TRY(fvh_eqn_create(m, &raw));
        S.pk_hold.reset(raw);
        FVH_TRY(fvh_eqn_create(m, &raw));
        S.pw_hold.reset(raw);
        FVH_TRY(fvh_precond_set(S.pk_hold.get(), S.cs.vk.precond, S.pe.get()));
        FVH_TRY(fvh_precond_set(S.pw_hold.get(), S.cs.vomega.precond, S.pe.get()));
    }
    FVH_TRY(fill_bcvals(S.bU.data(), (int)S.bU.size(), 0, S.vUx));
    FVH_TRY(fill_bcvals(S.bU.data(), (int)S.bU.size(), 1, S.vUy));
    FVH_TRY(fill_bcvals(S.bU.data(), (int)S.bU.size(), 2, S.vUz));
    FVH_TRY(fill_bcvals(S.bp.data(), (int)S.bp.size(), 0, S.vp));
    FVH_TRY(bc_plan(m, S.bU.data(), (int)S.bU.size(), &S.planU));
    FVH_TRY(bc_plan(m, S.bp.data(), (int)S.bp.size(), &S.planP));
    if (S.cs.model == FVH_KOMEGA) {
        FVH_TRY(fill_bcvals(S.bk.data(), (int)S.bk.size(), 0, S.vk));
        FVH_TRY(fill_bcvals(S.bw.data(), (int)S.bw.size(), 0, S.vw));
        FVH_TRY(fill_bcvals(S.bnut.data(), (int)S.bnut.size(), 0, S.vnut));
        FVH_TRY(bc_plan(m, S.bk.data(), (int)S.bk.size(), &S.planK));
        FVH_TRY(bc_plan(m, S.bw.data(), (int)S.bw.size(), &S.planW));
        FVH_TRY(wall_plan(m, S.bk, FVH_BC_KWALL, S.wK));
        FVH_TRY(wall_plan(m, S.bw, FVH_BC_OMEGAWALL, S.wW));
        FVH_TRY(wall_plan(m, S.bnut, FVH_BC_NUTWALL, S.wNut));
    }
    // term / BC compatibility is checked once here (the reference would fail
    // with a MethodError at the first discretise!)
    {
        AsmArgs a;
        std::memset(&a, 0, sizeof a);
        a.nterms = 3;
        a.t[0] = term(FVH_OP_TIME, S.cs.sU.time, 1, nullptr, nullptr);
        a.t[1] = term(FVH_OP_DIVERGENCE, S.cs.sU.divergence, 1, nullptr, nullptr);
        a.t[2] = term(FVH_OP_LAPLACIAN, S.cs.sU.laplacian, -1, nullptr, nullptr);
        FVH_TRY(check_bc_terms(a, S.bU.data(), (int)S.bU.size()));
        if (S.cs.sU.laplacian != FVH_LINEAR || S.cs.sp.laplacian != FVH_LINEAR)
            return set_error(FVH_ERR_UNSUPPORTED, "Laplacian scheme must be Linear");
        if (S.cs.sU.divergence != FVH_LINEAR && S.cs.sU.divergence != FVH_UPWIND)
            return set_error(FVH_ERR_UNSUPPORTED, "Divergence scheme %d", S.cs.sU.divergence);
    }
    // LDL: the reference factors each equation's set-up matrix once
    // (set_preconditioner, preconditioners_2_functions.jl:4-13, before the
    // loop: rDf = 1, and no face flux or viscosity yet, which leaves the
    // momentum and turbulence matrices zero) and never updates the factor.
    if (S.cs.vU.precond == FVH_PRECOND_LDL ||
        (S.cs.model == FVH_KOMEGA && (S.cs.vk.precond == FVH_PRECOND_LDL || S.cs.vomega.precond == FVH_PRECOND_LDL)))
        return set_error(FVH_ERR_UNSUPPORTED,
                         "LDL: the momentum and turbulence equations' set-up matrices are zero (no flux and no "
                         "viscosity before the loop), so their factorization has a zero pivot");
    if (S.cs.vp.precond == FVH_PRECOND_LDL) {
        std::vector<KTerm> t = {term(FVH_OP_LAPLACIAN, S.cs.sp.laplacian, 1, S.rDf.p, S.p.p)};
        FVH_TRY(assemble(S, S.pe.get(), t, {{1.0, S.divHv.p}}, S.bp, 0, S.planP, false, 1.0, false));
        FVH_TRY(precond_update(S.pe.get()));
    }
    S.Rux.assign(S.iterations, 1.0);
    S.Ruy.assign(S.iterations, 1.0);
    S.Rp.assign(S.iterations, 1.0);
    if (S.cases) {
        S.Rc.assign((size_t)S.K * S.iterations * 3, 1.0);
        S.case_break.assign(S.K, -1);
        FVH_TRY(S.snapUx.alloc(m->n));
        FVH_TRY(S.snapUy.alloc(m->n));
        FVH_TRY(S.snapP.alloc(m->n));
    }
    // loop prologue (Solvers_1_SIMPLE.jl:114-119)
    FVH_TRY(velocity_faces(S, S.Ux.p, S.Uy.p, S.Uz.p));
    FVH_TRY(launch_flux(m, S.Ufx.p, S.Ufy.p, S.Ufz.p, S.mdotf.p));
    FVH_TRY(launch_grad(m, S.cs.sp.gradient, S.vp, S.p.p, S.pf.p, S.gpx.p, S.gpy.p, S.gpz.p));
    FVH_TRY(update_nueff(S));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

}  // namespace fvh

using namespace fvh;

struct fvh_sim {
    Sim S;
};

static int sim_create(fvh_mesh *m, const fvh_case *c, const double *nu_case, const int64_t *ref_cells,
                      const double *Ux, const double *Uy, const double *Uz, const double *p, const double *k,
                      const double *omega, const double *nut, fvh_sim **out) {
    if (!m || !c || !out) return set_error(FVH_ERR_ARG, "null argument");
    if (c->iterations < 0) return set_error(FVH_ERR_ARG, "iterations < 0");
    for (const fvh_solver_cfg *v : {&c->vU, &c->vp})
        if (v->precision != FVH_FP64 && v->precision != FVH_MIXED)
            return set_error(v->precision == FVH_MIXED_FP16 ? FVH_ERR_UNSUPPORTED : FVH_ERR_ARG, "precision %d",
                             v->precision);
    auto h = std::make_unique<fvh_sim>();
    Sim &S = h->S;
    S.m = m;
    S.cs = *c;
    S.iterations = c->iterations;
    if (m->cases) {
        if (!nu_case) return set_error(FVH_ERR_ARG, "mesh has independent cases: use fvh_sim_create_cases");
        if (c->model != FVH_LAMINAR)
            return set_error(FVH_ERR_UNSUPPORTED, "independent cases: laminar models only");
        const CaseLayout &L = *m->cases;
        S.cases = true;
        S.K = L.K;
        DBuf<double> nuc;
        FVH_TRY(nuc.upload(std::vector<double>(nu_case, nu_case + L.K)));
        FVH_TRY(S.nu_face.alloc(m->nf));
        {
            const double *nc = nuc.p;
            const int *cc = L.cell_case.p, *own = m->d.own1;
            double *nf = S.nu_face.p;
            FVH_TRY(map(m->nf, nullptr, [=] __device__(int f) { nf[f] = nc[cc[own[f]]]; }));
        }
        std::vector<unsigned char> flag(m->n, 0);
        for (int q = 0; q < L.K; ++q) {
            const int64_t rc = ref_cells ? ref_cells[q] : -1;
            if (rc < 0) continue;
            if (rc >= L.h_off[q + 1] - L.h_off[q]) return set_error(FVH_ERR_ARG, "reference cell of case %d", q);
            flag[L.h_off[q] + rc] = 1;
        }
        FVH_TRY(S.refflag.upload(flag));
        HIP_TRY(hipStreamSynchronize(stream()));
    }
    auto load = [](const fvh_bcset &s, std::vector<fvh_bc> &v) {
        v.assign(s.bcs, s.bcs + s.nbc);
    };
    load(c->U, S.bU);
    load(c->p, S.bp);
    load(c->k, S.bk);
    load(c->omega, S.bw);
    load(c->nut, S.bnut);
    int rc = sim_init(S, Ux, Uy, Uz, p, k, omega, nut);
    if (rc) return rc;
    *out = h.release();
    return FVH_OK;
}

extern "C" int fvh_sim_create(fvh_mesh *m, const fvh_case *c, const double *Ux, const double *Uy, const double *Uz,
                              const double *p, const double *k, const double *omega, const double *nut, fvh_sim **out) {
    FVH_GUARD_BEGIN
    return sim_create(m, c, nullptr, nullptr, Ux, Uy, Uz, p, k, omega, nut, out);
    FVH_GUARD_END
}

extern "C" int fvh_sim_create_cases(fvh_mesh *m, const fvh_case *c, const double *nu_case, const int64_t *ref_cells,
                                    const double *Ux, const double *Uy, const double *Uz, const double *p,
                                    fvh_sim **out) {
    FVH_GUARD_BEGIN
    if (!m || !m->cases) return set_error(FVH_ERR_ARG, "mesh has no independent cases (fvh_mesh_set_cases)");
    if (!nu_case) return set_error(FVH_ERR_ARG, "null viscosities");
    return sim_create(m, c, nu_case, ref_cells, Ux, Uy, Uz, p, nullptr, nullptr, nullptr, out);
    FVH_GUARD_END
}

extern "C" int fvh_sim_case_residuals(fvh_sim *h, int32_t k, double *a, double *b, double *c) {
    if (!h) return set_error(FVH_ERR_ARG, "null sim");
    Sim &S = h->S;
    if (!S.cases || k < 0 || k >= S.K) return set_error(FVH_ERR_ARG, "no case %d", k);
    for (int it = 0; it < S.iterations; ++it) {
        const double *r = &S.Rc[((size_t)k * S.iterations + it) * 3];
        if (a) a[it] = r[0];
        if (b) b[it] = r[1];
        if (c) c[it] = r[2];
    }
    return FVH_OK;
}

extern "C" int fvh_sim_case_stats(fvh_sim *h, int32_t k, const char *field, fvh_stats *out) {
    if (!h || !field || !out) return set_error(FVH_ERR_ARG, "null argument");
    Sim &S = h->S;
    if (!S.cases || k < 0 || k >= S.K) return set_error(FVH_ERR_ARG, "no case %d", k);
    auto it = S.cstats.find(field);
    if (it == S.cstats.end() || (int)it->second.size() != S.K)
        return set_error(FVH_ERR_ARG, "no statistics for %s", field);
    *out = it->second[k];
    return FVH_OK;
}

extern "C" void fvh_sim_destroy(fvh_sim *s) { delete s; }

extern "C" int fvh_sim_iterate(fvh_sim *h, int32_t n, int32_t *done) {
    FVH_GUARD_BEGIN
    if (!h) return set_error(FVH_ERR_ARG, "null sim");
    Sim &S = h->S;
    int32_t k = 0;
    while (k < n && S.it < S.iterations && !S.converged) {
        FVH_TRY(iterate_once(S));
        ++k;
    }
    if (done) *done = k;
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_sim_residuals(fvh_sim *h, double *a, double *b, double *c) {
    if (!h) return set_error(FVH_ERR_ARG, "null sim");
    Sim &S = h->S;
    if (a) std::copy(S.Rux.begin(), S.Rux.end(), a);
    if (b) std::copy(S.Ruy.begin(), S.Ruy.end(), b);
    if (c) std::copy(S.Rp.begin(), S.Rp.end(), c);
    return FVH_OK;
}

extern "C" int fvh_sim_stats(fvh_sim *h, const char *field, fvh_stats *out) {
    if (!h || !field || !out) return set_error(FVH_ERR_ARG, "null argument");
    auto it = h->S.stats.find(field);
    if (it == h->S.stats.end()) return set_error(FVH_ERR_ARG, "no statistics for %s", field);
    *out = it->second;
    return FVH_OK;
}

extern "C" int fvh_sim_timing(fvh_sim *h, double *a, double *s, double *o) {
    if (!h) return set_error(FVH_ERR_ARG, "null sim");
    if (a) *a = h->S.ms_asm;
    if (s) *s = h->S.ms_solve;
    if (o) *o = h->S.ms_other;
    return FVH_OK;
}

// Average device time of one launch of a loop phase on the simulation's
// current fields, by HIP events on the library stream.  op 0: the Ux
// momentum assembly as the loop runs it (terms, source, BCs, relaxation);
// op 1: the pressure equation's assembly; op 2: the Green-Gauss kernel on
// p's face values.  The phase's outputs are recomputed in place.
extern "C" int fvh_sim_time_op(fvh_sim *h, int32_t op, int32_t reps, double *ms) {
    FVH_GUARD_BEGIN
    if (!h || !ms || reps <= 0) return set_error(FVH_ERR_ARG, "bad argument");
    if (op < 0 || op > 2) return set_error(FVH_ERR_UNSUPPORTED, "fvh_sim_time_op: op %d", op);
    Sim &S = h->S;
    fvh_mesh *m = S.m;
    auto launch = [&]() -> int {
        if (op == 0) {
            std::vector<KTerm> t = {term(FVH_OP_TIME, S.cs.sU.time, 1, nullptr, S.Ux.p),
                                    term(FVH_OP_DIVERGENCE, S.cs.sU.divergence, 1, S.mdotf.p, S.Ux.p),
                                    term(FVH_OP_LAPLACIAN, S.cs.sU.laplacian, -1, S.nueff.p, S.Ux.p)};
            return assemble(S, S.ux.get(), t, {{-1.0, S.gpx.p}}, S.bU, 0, S.planU, S.cs.algorithm != FVH_PISO,
                            S.cs.vU.relax, false);
        }
        if (op == 1) {
            std::vector<KTerm> t = {term(FVH_OP_LAPLACIAN, S.cs.sp.laplacian, 1, S.rDf.p, S.p.p)};
            return assemble(S, S.pe.get(), t, {{1.0, S.divHv.p}}, S.bp, 0, S.planP, false, 1.0, S.cs.has_pref != 0);
        }
        return launch_green_gauss(m, S.pf.p, S.gpx.p, S.gpy.p, S.gpz.p);
    };
    FVH_TRY(launch());  // warm
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    int rc = FVH_OK;
    if (hipEventRecord(a, stream()) != hipSuccess) rc = set_error(FVH_ERR_HIP, "hipEventRecord");
    for (int r = 0; r < reps && rc == FVH_OK; ++r) rc = launch();
    if (rc == FVH_OK && hipEventRecord(b, stream()) != hipSuccess) rc = set_error(FVH_ERR_HIP, "hipEventRecord");
    if (rc == FVH_OK && hipEventSynchronize(b) != hipSuccess) rc = set_error(FVH_ERR_HIP, "hipEventSynchronize");
    float t = 0.f;
    if (rc == FVH_OK && hipEventElapsedTime(&t, a, b) != hipSuccess) rc = set_error(FVH_ERR_HIP, "hipEventElapsedTime");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != FVH_OK) return rc;
    *ms = (double)t / reps;
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_bound(const fvh_mesh *m, double *field, double bnd, int32_t *chained) {
    FVH_GUARD_BEGIN
    if (!m || !field) return set_error(FVH_ERR_ARG, "null argument");
    BoundWork w;
    FVH_TRY(bound_field(m, field, bnd, w));
    HIP_TRY(hipStreamSynchronize(stream()));
    if (chained) *chained = w.last_chained;
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_sim_equation(fvh_sim *h, const char *name, fvh_eqn **out) {
    if (!h || !name || !out) return set_error(FVH_ERR_ARG, "null argument");
    std::string nm(name);
    if (nm == "ux") *out = h->S.ux.get();
    else if (nm == "uy") *out = h->S.uy.get();
    else if (nm == "p") *out = h->S.pe.get();
    else return set_error(FVH_ERR_ARG, "no equation %s", name);
    return FVH_OK;
}

extern "C" int fvh_sim_get(fvh_sim *h, const char *name, double *out, int64_t *len) {
    FVH_GUARD_BEGIN
    if (!h || !name) return set_error(FVH_ERR_ARG, "null argument");
    Sim &S = h->S;
    const int64_t n = S.m->n, nf = S.m->nf, nnz = S.m->nnz;
    std::string nm(name);
    const double *src = nullptr;
    int64_t cnt = 0;
    std::map<std::string, std::pair<const double *, int64_t>> tab = {
        {"Ux", {S.Ux.p, n}}, {"Uy", {S.Uy.p, n}}, {"Uz", {S.Uz.p, n}}, {"p", {S.p.p, n}}, {"k", {S.k.p, n}},
        {"omega", {S.omega.p, n}}, {"nut", {S.nut.p, n}}, {"mdotf", {S.mdotf.p, nf}}, {"nueff", {S.nueff.p, nf}},
        {"rDf", {S.rDf.p, nf}}, {"divHv", {S.divHv.p, n}}, {"rD", {S.rD.p, n}}, {"Hv_x", {S.Hx.p, n}},
        {"Hv_y", {S.Hy.p, n}}, {"gradp_x", {S.gpx.p, n}}, {"gradp_y", {S.gpy.p, n}}, {"gradp_z", {S.gpz.p, n}},
        {"pf", {S.pf.p, nf}}, {"Uf_x", {S.Ufx.p, nf}}, {"Uf_y", {S.Ufy.p, nf}}, {"Uf_z", {S.Ufz.p, nf}},
        {"ux_b", {S.ux->b.p, n}}, {"uy_b", {S.uy->b.p, n}}, {"p_b", {S.pe->b.p, n}}, {"ux_A", {S.ux->nz.p, nnz}},
        {"uy_A", {S.uy->nz.p, nnz}}, {"p_A", {S.pe->nz.p, nnz}}, {"nutf", {S.nutf.p, nf}}, {"Pk", {S.Pk.p, n}},
        {"gradU_xx", {S.gU[0].p, n}}, {"gradU_yx", {S.gU[1].p, n}}, {"gradU_xy", {S.gU[3].p, n}},
        {"gradU_yy", {S.gU[4].p, n}}, {"nueffk", {S.nueffk.p, nf}}, {"nueffw", {S.nueffw.p, nf}}};
    auto it = tab.find(nm);
    if (it != tab.end()) {
        src = it->second.first;
        cnt = it->second.second;
    } else if (nm == "ux_dinv" && S.ux->pkind == FVH_PRECOND_DILU) {
        FVH_TRY(precond_dinv_natural(S.ux.get(), &src));
        cnt = n;
    } else if (nm == "p_dinv" && S.pe->pkind == FVH_PRECOND_DILU) {
        FVH_TRY(precond_dinv_natural(S.pe.get(), &src));
        cnt = n;
    } else {
        return set_error(FVH_ERR_ARG, "unknown array %s", name);
    }
    if (len) *len = cnt;
    if (out) {
        HIP_TRY(hipStreamSynchronize(stream()));
        HIP_TRY(hipMemcpy(out, src, sizeof(double) * cnt, hipMemcpyDeviceToHost));
        // independent cases: a case that reached its break reports its fields there
        const double *snap = nm == "Ux" ? S.snapUx.p : nm == "Uy" ? S.snapUy.p : nm == "p" ? S.snapP.p : nullptr;
        if (S.cases && snap) {
            const std::vector<int> &off = S.m->cases->h_off;
            for (int k = 0; k < S.K; ++k)
                if (S.case_break[k] >= 0)
                    HIP_TRY(hipMemcpy(out + off[k], snap + off[k], sizeof(double) * (off[k + 1] - off[k]),
                                      hipMemcpyDeviceToHost));
        }
    }
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_simple(fvh_mesh *m, const fvh_case *c, double *Ux, double *Uy, double *Uz, double *p, double *k,
                          double *omega, double *nut, double *R_ux, double *R_uy, double *R_p, int32_t *done) {
    fvh_sim *h = nullptr;
    FVH_TRY(fvh_sim_create(m, c, Ux, Uy, Uz, p, k, omega, nut, &h));
    std::unique_ptr<fvh_sim> guard(h);
    int32_t dn = 0;
    FVH_TRY(fvh_sim_iterate(h, c->iterations, &dn));
    const int64_t n = m->n;
    auto get = [&](const char *nm, double *dst) -> int {
        if (!dst) return FVH_OK;
        int64_t len = 0;
        return fvh_sim_get(h, nm, dst, &len);
    };
    (void)n;
    FVH_TRY(get("Ux", Ux));
    FVH_TRY(get("Uy", Uy));
    FVH_TRY(get("Uz", Uz));
    FVH_TRY(get("p", p));
    if (c->model == FVH_KOMEGA) {
        FVH_TRY(get("k", k));
        FVH_TRY(get("omega", omega));
        FVH_TRY(get("nut", nut));
    }
    FVH_TRY(fvh_sim_residuals(h, R_ux, R_uy, R_p));
    if (done) *done = dn;
    return FVH_OK;
}
