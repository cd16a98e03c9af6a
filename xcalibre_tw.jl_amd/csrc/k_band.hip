// DILU recurrences on banded grids (cells numbered i-fastest on an nx x ny
// grid, lower neighbours r-1 and r-nx only), in the reference's row order.
//
// Row r = i + j*nx of the forward sweep needs rows r-nx (south) and r-1
// (west); the dependency depth is nx + ny - 1 anti-diagonals.  One
// work-group walks the anti-diagonals T = i + j; thread t owns grid row j = t
// and handles cell (T - t, t) at step T, so
//   * the west value (i-1, j) is its own result of step T-1 (a register),
//   * the south value (i, j-1) is lane t-1's result of step T-1: a DPP
//     wave_shr:1 inside a wave, an LDS slot across wave boundaries.
// One barrier per step.  All coefficients are repacked once per
// preconditioner update in anti-diagonal order (k_band_pack), so the loads of
// one step are contiguous across the threads; they are issued kPD steps ahead
// into a register ring.  The right-hand side is gathered into the packed
// order and the result scattered back by all-CU permutation kernels.
//
// Each row evaluates exactly the reference's expressions in the reference's
// order (preconditioners_1_DILU.jl:56-162):
//   D_r   = (A_rr - (A_rs*A_sr)/D_s) - (A_rw*A_wr)/D_w,    Dinv_r = 1/D_r
//   y_r   = (b_r - (A_rs*y_s)*Dinv_s) - (A_rw*y_w)*Dinv_w
//   z_r   = (y_r*Dinv_r - (A_re*z_e)*Dinv_r) - (A_rn*z_n)*Dinv_r
// so the result is bit-identical to the level-scheduled kernels.
#include "fvh_eqn.h"
#include "k_band.h"

#include <cstdlib>

namespace fvh {

constexpr int kPD = 8;  // prefetch distance (steps)

enum : unsigned char { F_S = 1, F_W = 2, F_E = 4, F_N = 8 };

// Work-group barrier that orders LDS only: global loads issued for later
// steps stay in flight across it (a full __syncthreads waits for them).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ inline double dpp_shr1(double v) {  // lane l <- lane l-1 (lane 0 <- 0)
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double dpp_shl1(double v) {  // lane l <- lane l+1 (lane 63 <- 0)
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// ---------------------------------------------------------------------------
// Packing (all CUs): natural CSR values -> anti-diagonal order
// ---------------------------------------------------------------------------
__global__ void k_band_pack(int n, int nx, const int *__restrict__ iperm, const unsigned char *__restrict__ flags,
                            const int *__restrict__ rowptr, const int *__restrict__ diag,
                            const double *__restrict__ val, BandVals bv) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int r = iperm[p];
    const unsigned char f = flags[p];
    const int d = diag[r];
    bv.ad[p] = val[d];
    bv.ls[p] = (f & F_S) ? val[rowptr[r]] : 0.0;
    bv.lw[p] = (f & F_W) ? val[d - 1] : 0.0;
    bv.ue[p] = (f & F_E) ? val[d + 1] : 0.0;
    bv.un[p] = (f & F_N) ? val[rowptr[r + 1] - 1] : 0.0;
    bv.ts[p] = (f & F_S) ? val[rowptr[r - nx + 1] - 1] : 0.0;  // A[r-nx, r]: last entry of row r-nx
    bv.tw[p] = (f & F_W) ? val[diag[r - 1] + 1] : 0.0;         // A[r-1, r]: after the diagonal of row r-1
}

__global__ void k_band_gather(int n, const int *__restrict__ iperm, const double *__restrict__ x,
                              double *__restrict__ xp, const KScal *ks) {
    if (ks && ks->done) return;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) xp[p] = x[iperm[p]];
}

__global__ void k_band_scatter(int n, const int *__restrict__ iperm, const double *__restrict__ xp,
                               double *__restrict__ x, const KScal *ks) {
    if (ks && ks->done) return;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) x[iperm[p]] = xp[p];
}

// Anti-diagonal walk: offj(T) = (first packed position of diagonal T) - jmin(T),
// advanced in scalar registers (no per-step loads of the offset table).
struct Walk {
    int T, offj, jmin, jmax;
    __device__ void init(const BandGeom &g, int T0) {
        T = T0;
        offj = (T0 >= 0 && T0 < g.nT) ? g.offj[T0] : 0;
        jmin = max(0, T0 - g.nx + 1);
        jmax = min(g.ny - 1, T0);
    }
    __device__ void next(const BandGeom &g) {  // T -> T+1
        int cnt = jmax - jmin + 1;
        offj += cnt;
        T += 1;
        int jmin1 = max(0, T - g.nx + 1);
        offj -= jmin1 - jmin;
        jmin = jmin1;
        jmax = min(g.ny - 1, T);
    }
    __device__ void prev(const BandGeom &g) {  // T -> T-1
        T -= 1;
        int jmin1 = max(0, T - g.nx + 1), jmax1 = min(g.ny - 1, T);
        int cnt = jmax1 - jmin1 + 1;
        offj = offj + jmin - jmin1 - cnt;
        jmin = jmin1;
        jmax = jmax1;
    }
    // packed position of thread t on this diagonal, clamped into [0, n) so the
    // load is always in bounds (the value of an inactive thread is unused)
    __device__ int pos(int t, int n) const {
        int q = offj + t;
        return q < 0 ? 0 : (q >= n ? n - 1 : q);
    }
    __device__ bool active(int t) const { return T >= 0 && t >= jmin && t <= jmax; }
};

// ---------------------------------------------------------------------------
// Diagonal recurrence (update_dilu_diagonal!)
// ---------------------------------------------------------------------------
struct PreU {
    double ad, ls, ts, lw, tw;
    unsigned f;
};

__global__ void __launch_bounds__(1024) k_band_dilu_update(BandGeom g, BandVals bv, double *__restrict__ dinv) {
    __shared__ double sh_d[16][2];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int n = g.nx * g.ny;
    const bool row = t < g.ny;
    PreU buf[kPD];
    Walk cw, pw;  // current and prefetch walks
    cw.init(g, 0);
    pw.init(g, 0);
    auto load = [&](const Walk &w) -> PreU {
        const int p = w.pos(t, n);
        return PreU{bv.ad[p], bv.ls[p], bv.ts[p], bv.lw[p], bv.tw[p], g.flags[p]};
    };
#pragma unroll
    for (int k = 0; k < kPD; ++k) {
        buf[k] = load(pw);
        pw.next(g);
    }
    double dprev = 0.0;  // D of (i-1, j)
    for (int T0 = 0; T0 < g.nT; T0 += kPD) {
#pragma unroll
        for (int k = 0; k < kPD; ++k) {
            const PreU c = buf[k];
            double ds = dpp_shr1(dprev);
            const double dl = sh_d[wave > 0 ? wave - 1 : 0][(cw.T + 1) & 1];
            ds = (lane == 0 && wave > 0) ? dl : ds;
            double d1 = c.ad - c.ls * c.ts / ds;
            double d = (c.f & F_S) ? d1 : c.ad;
            double d2 = d - c.lw * c.tw / dprev;
            d = (c.f & F_W) ? d2 : d;
            const bool act = row && cw.T < g.nT && cw.active(t);
            if (act) dinv[cw.offj + t] = 1.0 / d;
            dprev = act ? d : dprev;
            if (lane == 63) sh_d[wave][cw.T & 1] = dprev;
            lds_barrier();
            buf[k] = load(pw);
            pw.next(g);
            cw.next(g);
        }
    }
}

// ---------------------------------------------------------------------------
// ldiv!: forward then backward sweep, packed in and out
// ---------------------------------------------------------------------------
struct Pre {
    double b, d, c1, c2;
    unsigned f;
};

template <int kPD>
__global__ void __launch_bounds__(1024) k_band_dilu_apply(BandGeom g, BandVals bv, const double *__restrict__ dinv,
                                                          const double *bp, double *yp, const KScal *ks) {
    if (ks && ks->done) return;
    __shared__ double sh_x[16][2], sh_d[16][2];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int n = g.nx * g.ny;
    const bool row = t < g.ny;
    const int nwaves = (blockDim.x + 63) >> 6;
    const int wl = wave > 0 ? wave - 1 : 0, wr = wave + 1 < nwaves ? wave + 1 : wave;
    // ---- forward: y_r = (b_r - (A_rs y_s) Dinv_s) - (A_rw y_w) Dinv_w ----
    {
        Pre buf[kPD];
        Walk cw, pw;
        cw.init(g, 0);
        pw.init(g, 0);
        auto load = [&](const Walk &w) -> Pre {
            const int p = w.pos(t, n);
            return Pre{bp[p], dinv[p], bv.ls[p], bv.lw[p], g.flags[p]};
        };
#pragma unroll
        for (int k = 0; k < kPD; ++k) {
            buf[k] = load(pw);
            pw.next(g);
        }
        double xprev = 0.0, dprev = 0.0;
        for (int T0 = 0; T0 < g.nT; T0 += kPD) {
#pragma unroll
            for (int k = 0; k < kPD; ++k) {
                const Pre c = buf[k];
                double xs = dpp_shr1(xprev), ds = dpp_shr1(dprev);
                const int par = (cw.T + 1) & 1;
                const double xl = sh_x[wl][par], dl = sh_d[wl][par];
                const bool fromlds = lane == 0 && wave > 0;
                xs = fromlds ? xl : xs;
                ds = fromlds ? dl : ds;
                double x1 = c.b - c.c1 * xs * ds;
                double x = (c.f & F_S) ? x1 : c.b;
                double x2 = x - c.c2 * xprev * dprev;
                x = (c.f & F_W) ? x2 : x;
                const bool act = row && cw.T < g.nT && cw.active(t);
                if (act) yp[cw.offj + t] = x;
                xprev = act ? x : xprev;
                dprev = act ? c.d : dprev;
                if (lane == 63) {
                    sh_x[wave][cw.T & 1] = xprev;
                    sh_d[wave][cw.T & 1] = dprev;
                }
                lds_barrier();
                buf[k] = load(pw);
                pw.next(g);
                cw.next(g);
            }
        }
    }
    __syncthreads();  // all forward stores complete before the backward loads
    // ---- backward: z_r = (y_r Dinv_r - (A_re z_e) Dinv_r) - (A_rn z_n) Dinv_r ----
    {
        Pre buf[kPD];
        Walk cw, pw;
        cw.init(g, g.nT - 1);
        pw.init(g, g.nT - 1);
        auto load = [&](const Walk &w) -> Pre {
            const int p = w.pos(t, n);
            return Pre{yp[p], dinv[p], bv.ue[p], bv.un[p], g.flags[p]};
        };
#pragma unroll
        for (int k = 0; k < kPD; ++k) {
            buf[k] = load(pw);
            pw.prev(g);
        }
        double zprev = 0.0;
        for (int T0 = g.nT - 1; T0 >= 0; T0 -= kPD) {
#pragma unroll
            for (int k = 0; k < kPD; ++k) {
                const Pre c = buf[k];
                double zn = dpp_shl1(zprev);
                const double zl = sh_x[wr][(cw.T + 1) & 1];
                zn = (lane == 63 && wave + 1 < nwaves) ? zl : zn;
                double z = c.b * c.d;
                double z1 = z - c.c1 * zprev * c.d;
                z = (c.f & F_E) ? z1 : z;
                double z2 = z - c.c2 * zn * c.d;
                z = (c.f & F_N) ? z2 : z;
                const bool act = row && cw.active(t);
                if (act) yp[cw.offj + t] = z;
                zprev = act ? z : zprev;
                if (lane == 0) sh_x[wave][cw.T & 1] = zprev;
                lds_barrier();
                buf[k] = load(pw);
                pw.prev(g);
                cw.prev(g);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Anti-diagonal geometry of a banded mesh: offsets, packed row order, flags.
int band_setup(fvh_mesh *m) {
    Schedule &s = m->sched;
    const int nx = s.band_nx;
    if (nx <= 0) return FVH_OK;
    const int n = (int)m->n, ny = n / nx;
    if (ny > 1024) {  // one thread per grid row in one work-group
        s.band_nx = 0;
        return FVH_OK;
    }
    const int nT = nx + ny - 1;
    std::vector<int> offj(nT), iperm(n), perm(n);
    std::vector<unsigned char> flags(n);
    const auto &rp = m->h_rowptr, &col = m->h_col, &dg = m->h_diag;
    int p = 0;
    for (int T = 0; T < nT; ++T) {
        int jmin = std::max(0, T - nx + 1), jmax = std::min(ny - 1, T);
        offj[T] = p - jmin;
        for (int j = jmin; j <= jmax; ++j, ++p) {
            int r = (T - j) + j * nx;
            iperm[p] = r;
            perm[r] = p;
            unsigned char f = 0;
            for (int q = rp[r]; q < rp[r + 1]; ++q) {
                int c = col[q];
                if (c == r - nx) f |= F_S;
                else if (c == r - 1) f |= F_W;
                else if (c == r + 1) f |= F_E;
                else if (c == r + nx) f |= F_N;
                else if (c != r) return set_error(FVH_ERR_MESH, "band detection inconsistent at row %d", r);
            }
            // the packing reads the south coefficient as the first entry of
            // the row and the north one as the last: check the layout
            if ((f & F_S) && col[rp[r]] != r - nx) return set_error(FVH_ERR_MESH, "band layout (south) at row %d", r);
            if ((f & F_N) && col[rp[r + 1] - 1] != r + nx) return set_error(FVH_ERR_MESH, "band layout (north) at row %d", r);
            if ((f & F_W) && col[dg[r] - 1] != r - 1) return set_error(FVH_ERR_MESH, "band layout (west) at row %d", r);
            if ((f & F_E) && col[dg[r] + 1] != r + 1) return set_error(FVH_ERR_MESH, "band layout (east) at row %d", r);
            flags[p] = f;
        }
    }
    // symmetric pattern: row r-nx ends with r, row r-1 has r right after its diagonal
    for (int r = 0; r < n; ++r) {
        unsigned char f = flags[perm[r]];
        if ((f & F_S) && col[rp[r - nx + 1] - 1] != r) return set_error(FVH_ERR_MESH, "band transpose (south) at %d", r);
        if ((f & F_W) && col[dg[r - 1] + 1] != r) return set_error(FVH_ERR_MESH, "band transpose (west) at %d", r);
    }
    FVH_TRY(s.b_offj.upload(offj));
    FVH_TRY(s.b_iperm.upload(iperm));
    FVH_TRY(s.b_perm.upload(perm));
    FVH_TRY(s.b_flags.upload(flags));
    s.g.nx = nx;
    s.g.ny = ny;
    s.g.nT = nT;
    s.g.offj = s.b_offj.p;
    s.g.flags = s.b_flags.p;
    return FVH_OK;
}

int band_alloc(fvh_eqn *e) {
    const int64_t n = e->m->n;
    for (DBuf<double> *b : {&e->bd.ad, &e->bd.ls, &e->bd.lw, &e->bd.ue, &e->bd.un, &e->bd.ts, &e->bd.tw, &e->bd.dinv,
                            &e->bd.xp})
        if (b->n != (size_t)n) FVH_TRY(b->alloc(n));
    e->bd.v = BandVals{e->bd.ad.p, e->bd.ls.p, e->bd.lw.p, e->bd.ue.p, e->bd.un.p, e->bd.ts.p, e->bd.tw.p};
    return FVH_OK;
}

static int threads_for(int ny) { return ((ny + 63) / 64) * 64; }

int band_dilu_update(fvh_eqn *e, const double *val) {
    fvh_mesh *m = e->m;
    const Schedule &s = m->sched;
    FVH_TRY(band_alloc(e));
    const int n = (int)m->n;
    hipLaunchKernelGGL(k_band_pack, grid1d(n), dim3(256), 0, stream(), n, s.g.nx, s.b_iperm.p, s.b_flags.p,
                       m->d.rowptr, m->d.diag, val, e->bd.v);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_band_dilu_update, dim3(1), dim3(threads_for(s.g.ny)), 0, stream(), s.g, e->bd.v, e->bd.dinv.p);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int band_dilu_apply(fvh_eqn *e, const double *x, double *y, const KScal *ks) {
    fvh_mesh *m = e->m;
    const Schedule &s = m->sched;
    const int n = (int)m->n;
    double *xp = e->bd.xp.p;
    hipLaunchKernelGGL(k_band_gather, grid1d(n), dim3(256), 0, stream(), n, s.b_iperm.p, x, xp, ks);
    HIP_TRY(hipGetLastError());
    static int pd = -1;
    if (pd < 0) {
        const char *env = std::getenv("FVH_BAND_PD");
        pd = env ? std::atoi(env) : 8;
    }
    const dim3 blk(threads_for(s.g.ny));
    if (pd == 16)
        hipLaunchKernelGGL((k_band_dilu_apply<16>), dim3(1), blk, 0, stream(), s.g, e->bd.v, e->bd.dinv.p, xp, xp, ks);
    else if (pd == 24)
        hipLaunchKernelGGL((k_band_dilu_apply<24>), dim3(1), blk, 0, stream(), s.g, e->bd.v, e->bd.dinv.p, xp, xp, ks);
    else if (pd == 4)
        hipLaunchKernelGGL((k_band_dilu_apply<4>), dim3(1), blk, 0, stream(), s.g, e->bd.v, e->bd.dinv.p, xp, xp, ks);
    else
        hipLaunchKernelGGL((k_band_dilu_apply<8>), dim3(1), blk, 0, stream(), s.g, e->bd.v, e->bd.dinv.p, xp, xp, ks);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_band_scatter, grid1d(n), dim3(256), 0, stream(), n, s.b_iperm.p, xp, y, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// Natural-order copy of the packed reciprocal diagonal.
int band_dinv_natural(fvh_eqn *e, double *out) {
    fvh_mesh *m = e->m;
    const int n = (int)m->n;
    hipLaunchKernelGGL(k_band_scatter, grid1d(n), dim3(256), 0, stream(), n, m->sched.b_iperm.p, e->bd.dinv.p, out,
                       (const KScal *)nullptr);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

}  // namespace fvh
