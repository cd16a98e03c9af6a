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

namespace fvh {

constexpr int kPD = 8;  // prefetch distance (steps)

enum : unsigned char { F_S = 1, F_W = 2, F_E = 4, F_N = 8 };

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

// ---------------------------------------------------------------------------
// Diagonal recurrence (update_dilu_diagonal!)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_band_dilu_update(BandGeom g, BandVals bv, double *__restrict__ dinv) {
    __shared__ double sh_d[16][2];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const bool row = t < g.ny;
    double dprev = 0.0;  // D of (i-1, j)
    for (int T = 0; T < g.nT; ++T) {
        double ds = dpp_shr1(dprev);
        if (lane == 0 && wave > 0) ds = sh_d[wave - 1][(T + 1) & 1];
        const int i = T - t;
        if (row && i >= 0 && i < g.nx) {
            const int p = g.offj[T] + t;
            const unsigned char f = g.flags[p];
            double d = bv.ad[p];
            if (f & F_S) d = d - bv.ls[p] * bv.ts[p] / ds;
            if (f & F_W) d = d - bv.lw[p] * bv.tw[p] / dprev;
            dinv[p] = 1.0 / d;
            dprev = d;
        }
        if (lane == 63) sh_d[wave][T & 1] = dprev;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// ldiv!: forward then backward sweep, packed in and out
// ---------------------------------------------------------------------------
struct Pre {
    double b, d, c1, c2;
    unsigned char f;
};

__global__ void __launch_bounds__(1024) k_band_dilu_apply(BandGeom g, BandVals bv, const double *__restrict__ dinv,
                                                          const double *bp, double *yp,
                                                          const KScal *ks) {
    if (ks && ks->done) return;
    __shared__ double sh_x[16][2], sh_d[16][2];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const bool row = t < g.ny;
    const int nwaves = (blockDim.x + 63) >> 6;
    // ---- forward ----
    {
        Pre buf[kPD];
        auto load = [&](int T) -> Pre {
            Pre q{0.0, 0.0, 0.0, 0.0, 0};
            const int i = T - t;
            if (row && T < g.nT && i >= 0 && i < g.nx) {
                const int p = g.offj[T] + t;
                q.f = g.flags[p];
                q.b = bp[p];
                q.d = dinv[p];
                q.c1 = bv.ls[p];
                q.c2 = bv.lw[p];
            }
            return q;
        };
#pragma unroll
        for (int k = 0; k < kPD; ++k) buf[k] = load(k);
        double xprev = 0.0, dprev = 0.0;
        for (int T0 = 0; T0 < g.nT; T0 += kPD) {
#pragma unroll
            for (int k = 0; k < kPD; ++k) {
                const int T = T0 + k;
                const Pre cur = buf[k];
                buf[k] = load(T + kPD);
                double xs = dpp_shr1(xprev), ds = dpp_shr1(dprev);
                if (lane == 0 && wave > 0) {
                    xs = sh_x[wave - 1][(T + 1) & 1];
                    ds = sh_d[wave - 1][(T + 1) & 1];
                }
                const int i = T - t;
                if (row && T < g.nT && i >= 0 && i < g.nx) {
                    double x = cur.b;
                    if (cur.f & F_S) x = x - cur.c1 * xs * ds;
                    if (cur.f & F_W) x = x - cur.c2 * xprev * dprev;
                    yp[g.offj[T] + t] = x;
                    xprev = x;
                    dprev = cur.d;
                }
                if (lane == 63) {
                    sh_x[wave][T & 1] = xprev;
                    sh_d[wave][T & 1] = dprev;
                }
                __syncthreads();
            }
        }
    }
    // ---- backward (anti-diagonals in reverse) ----
    {
        Pre buf[kPD];
        auto load = [&](int T) -> Pre {
            Pre q{0.0, 0.0, 0.0, 0.0, 0};
            const int i = T - t;
            if (row && T >= 0 && i >= 0 && i < g.nx) {
                const int p = g.offj[T] + t;
                q.f = g.flags[p];
                q.b = yp[p];
                q.d = dinv[p];
                q.c1 = bv.ue[p];
                q.c2 = bv.un[p];
            }
            return q;
        };
        const int Tlast = g.nT - 1;
#pragma unroll
        for (int k = 0; k < kPD; ++k) buf[k] = load(Tlast - k);
        double zprev = 0.0;
        for (int T0 = Tlast; T0 >= 0; T0 -= kPD) {
#pragma unroll
            for (int k = 0; k < kPD; ++k) {
                const int T = T0 - k;
                const Pre cur = buf[k];
                buf[k] = load(T - kPD);
                double zn = dpp_shl1(zprev);
                if (lane == 63 && wave + 1 < nwaves) zn = sh_x[wave + 1][(T + 1) & 1];
                const int i = T - t;
                if (row && T >= 0 && i >= 0 && i < g.nx) {
                    double z = cur.b * cur.d;
                    if (cur.f & F_E) z = z - cur.c1 * zprev * cur.d;
                    if (cur.f & F_N) z = z - cur.c2 * zn * cur.d;
                    yp[g.offj[T] + t] = z;
                    zprev = z;
                }
                if (lane == 0) sh_x[wave][T & 1] = zprev;
                __syncthreads();
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
    hipLaunchKernelGGL(k_band_dilu_apply, dim3(1), dim3(threads_for(s.g.ny)), 0, stream(), s.g, e->bd.v, e->bd.dinv.p, xp,
                       xp, ks);
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
