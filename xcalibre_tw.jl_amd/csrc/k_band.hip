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
#include <cstdlib>
#include <cstring>

#include "fvh_eqn.h"
#include "k_band.h"
#include "k_dpp.h"

namespace fvh {

constexpr int kPD = 8;  // prefetch distance (steps)

// Record access by 32-bit byte offset from a base pointer (one VALU add per
// address; the packed arrays stay below 4 GiB, checked at set-up).
template <typename R>
__device__ inline const char *rec_at(const R *base, unsigned idx) {
    return reinterpret_cast<const char *>(base) + (size_t)(idx * (unsigned)sizeof(R));
}
__device__ inline Rec4 ldrec(const Rec4 *base, unsigned idx) {
    const char *p = rec_at(base, idx);
    const double2 u = *reinterpret_cast<const double2 *>(p), v = *reinterpret_cast<const double2 *>(p + 16);
    return Rec4{u.x, u.y, v.x, v.y};
}
__device__ inline Rec4f ldrec(const Rec4f *base, unsigned idx) {
    const float4 u = *reinterpret_cast<const float4 *>(rec_at(base, idx));
    return Rec4f{u.x, u.y, u.z, u.w};
}
__device__ inline double ldval(const double *base, unsigned idx) { return *reinterpret_cast<const double *>(rec_at(base, idx)); }
// first member of record idx
template <typename T>
__device__ inline void st_a(Rec4T<T> *base, unsigned idx, T v) {
    *reinterpret_cast<T *>(const_cast<char *>(rec_at(base, idx))) = v;
}
template <typename T>
__device__ inline void st_val(T *base, unsigned idx, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (size_t)(idx * (unsigned)sizeof(T))) = v;
}

// ---------------------------------------------------------------------------
// Packing (all CUs): natural CSR values -> anti-diagonal records
// ---------------------------------------------------------------------------
// Row r of a grid with nx columns whose cells are [o, o + nk) (o = 0, nk = n
// for a single mesh; one case of a batch otherwise).
__global__ void k_band_pack(int n, int nx1, const int *__restrict__ cell_case, const int *__restrict__ off,
                            const int *__restrict__ nxs, const int *__restrict__ iperm, const int *__restrict__ rowptr,
                            const int *__restrict__ diag, const double *__restrict__ val, BandRecs br) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int o = 0, nk = n, nx = nx1;
    if (cell_case) {
        const int k = cell_case[p];  // packed positions of a case cover its own cell range
        o = off[k];
        nk = off[k + 1] - o;
        nx = nxs[k];
    }
    const int r = iperm[p], rl = r - o;
    const int i = rl % nx;
    const bool S = rl >= nx, W = i > 0, E = i < nx - 1, N = rl + nx < nk;
    const int d = diag[r];
    const double a_rs = S ? val[rowptr[r]] : 0.0;               // A[r, r-nx]
    const double a_sr = S ? val[rowptr[r - nx + 1] - 1] : 0.0;  // A[r-nx, r]
    const double a_rw = W ? val[d - 1] : 0.0;                   // A[r, r-1]
    const double a_wr = W ? val[diag[r - 1] + 1] : 0.0;         // A[r-1, r]
    br.ud[p] = val[d];
    br.us[p] = a_rs * a_sr;
    br.uw[p] = a_rw * a_wr;
    br.f[p].c = a_rs;
    br.f[p].d = a_rw;
    br.b[p].c = E ? val[d + 1] : 0.0;                  // A[r, r+1]
    br.b[p].d = N ? val[rowptr[r + 1] - 1] : 0.0;      // A[r, r+nx]
}

// FP32 copies of the sweep records (coefficients and Dinv rounded once)
__global__ void k_band_to32(int n, const Rec4 *__restrict__ f, const Rec4 *__restrict__ b, Rec4f *__restrict__ f32,
                            Rec4f *__restrict__ b32) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const Rec4 u = f[p], v = b[p];
    f32[p] = Rec4f{0.f, (float)u.b, (float)u.c, (float)u.d};
    b32[p] = Rec4f{0.f, (float)v.b, (float)v.c, (float)v.d};
}

template <typename T>
__global__ void k_band_gather(int n, const int *__restrict__ iperm, const T *__restrict__ x, Rec4T<T> *__restrict__ f,
                              const KScal *ks) {
    if (ks && ks->done) return;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) f[p].a = x[iperm[p]];
}

template <typename T>
__global__ void k_band_scatter(int n, const int *__restrict__ iperm, const T *__restrict__ xp, T *__restrict__ x,
                               const KScal *ks) {
    if (ks && ks->done) return;
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) x[iperm[p]] = xp[p];
}

__global__ void k_band_dinv_natural(int n, const int *__restrict__ iperm, const Rec4 *__restrict__ f,
                                    double *__restrict__ out) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[iperm[p]] = f[p].b;
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
        offj += jmax - jmin + 1;
        T += 1;
        int jmin1 = max(0, T - g.nx + 1);
        offj -= jmin1 - jmin;
        jmin = jmin1;
        jmax = min(g.ny - 1, T);
    }
    __device__ void prev(const BandGeom &g) {  // T -> T-1
        T -= 1;
        int jmin1 = max(0, T - g.nx + 1), jmax1 = min(g.ny - 1, T);
        offj = offj + jmin - jmin1 - (jmax1 - jmin1 + 1);
        jmin = jmin1;
        jmax = jmax1;
    }
    // packed position of thread t, clamped into [0, n): every load is in
    // bounds; the value read by a thread off the diagonal is never used
    __device__ unsigned pos(int t, int n) const { return min((unsigned)(offj + t), (unsigned)(n - 1)); }
    // first of the R adjacent records of rows t .. t+R-1, clamped so that all
    // R are in bounds (R <= n)
    template <int R>
    __device__ unsigned pos_block(int t, int n) const {
        return min((unsigned)(offj + t), (unsigned)(n - R));
    }
};

// ---------------------------------------------------------------------------
// Diagonal recurrence (update_dilu_diagonal!)
//
// Lane t owns the R consecutive grid rows j = R t + q.  At step T row j
// handles cell (T - j, j); its south value (T - j, j - 1) is row j-1's result
// of step T-1: the lane's own register for q > 0, lane t-1's last row (DPP /
// LDS) for q = 0.  Its west value is its own result of step T-1.
// ---------------------------------------------------------------------------
template <int R>
__device__ inline void band_update_body(const BandGeom &g, const double *__restrict__ ud,
                                        const double *__restrict__ us, const double *__restrict__ uw,
                                        Rec4 *__restrict__ fr, Rec4 *__restrict__ br_) {
    __shared__ double sh_d[16][2];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wl = wave > 0 ? wave - 1 : 0;
    const int n = g.nx * g.ny;
    const int j0 = t * R;
    double bd[kPD][R], bs[kPD][R], bw[kPD][R];  // A_rr, A_rs*A_sr, A_rw*A_wr
    Walk cw, pw;
    cw.init(g, 0);
    pw.init(g, 0);
#pragma unroll
    for (int k = 0; k < kPD; ++k) {
        {
            const unsigned pb = pw.pos_block<R>(j0, n);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                bd[k][q] = ldval(ud, pb + q);
                bs[k][q] = ldval(us, pb + q);
                bw[k][q] = ldval(uw, pb + q);
            }
        }
        __builtin_amdgcn_sched_barrier(0);  // issue in consumption order
        pw.next(g);
    }
    double dprev[R];  // D of (i-1, j) per row
#pragma unroll
    for (int q = 0; q < R; ++q) dprev[q] = 0.0;
    for (int T0 = 0; T0 < g.nT; T0 += kPD) {
#pragma unroll
        for (int k = 0; k < kPD; ++k) {
            const double ds0 = dpp_shr1(dprev[R - 1], sh_d[wl][(cw.T + 1) & 1]);
            double dn[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int j = j0 + q, i = cw.T - j;
                const double ds = q == 0 ? ds0 : dprev[q - 1];
                // D_r = (A_rr - (A_rs A_sr) / D_s) - (A_rw A_wr) / D_w
                double d = bd[k][q];
                const double d1 = d - bs[k][q] / ds;
                d = (j > 0) ? d1 : d;
                const double d2 = d - bw[k][q] / dprev[q];
                d = (i > 0) ? d2 : d;
                if (j < g.ny && (unsigned)i < (unsigned)g.nx) {
                    const double r = 1.0 / d;
                    const unsigned p = cw.offj + j;
                    fr[p].b = r;
                    br_[p].b = r;
                }
                dn[q] = d;
            }
#pragma unroll
            for (int q = 0; q < R; ++q) dprev[q] = dn[q];
            if (lane == 63) sh_d[wave][cw.T & 1] = dprev[R - 1];
            lds_barrier();
            // the slots are consumed above: keep their refill below them
            __builtin_amdgcn_sched_barrier(0);
            {
                const unsigned pb = pw.pos_block<R>(j0, n);
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    bd[k][q] = ldval(ud, pb + q);
                    bs[k][q] = ldval(us, pb + q);
                    bw[k][q] = ldval(uw, pb + q);
                }
            }
            pw.next(g);
            cw.next(g);
        }
    }
}

template <int R>
__global__ void __launch_bounds__(1024 / R) k_band_dilu_update(BandGeom g, BandRecs br) {
    band_update_body<R>(g, br.ud, br.us, br.uw, br.f, br.b);
}

// one work-group per case of a batch
template <int R>
__global__ void __launch_bounds__(1024 / R) k_band_dilu_update_cases(const BandGeom *__restrict__ gs,
                                                                     const int *__restrict__ off, BandRecs br) {
    const int k = blockIdx.x, o = off[k];
    band_update_body<R>(gs[k], br.ud + o, br.us + o, br.uw + o, br.f + o, br.b + o);
}

// ---------------------------------------------------------------------------
// ldiv!: forward then backward sweep over the packed records, in FP64 (the
// reference's arithmetic) or FP32 (the inner solves of mixed precision); the
// row ownership is that of the diagonal recurrence.
// ---------------------------------------------------------------------------
template <typename T, int R, int PD = kPD>
__device__ inline void band_apply_body(const BandGeom &g, const Rec4T<T> *__restrict__ f, Rec4T<T> *bk,
                                       T *__restrict__ zp) {
    // [wave][step parity][value | Dinv] of the wave's edge lane
    __shared__ T sh[16][2][2];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int n = g.nx * g.ny;
    const int j0 = t * R;
    const int nwaves = (blockDim.x + 63) >> 6;
    const int wl = wave > 0 ? wave - 1 : 0, wr = wave + 1 < nwaves ? wave + 1 : wave;
    // ---- forward: y_r = (b_r - (A_rs y_s) Dinv_s) - (A_rw y_w) Dinv_w ----
    {
        Rec4T<T> buf[PD][R];  // b, dinv, A_rs, A_rw
        Walk cw, pw;
        cw.init(g, 0);
        pw.init(g, 0);
#pragma unroll
        for (int k = 0; k < PD; ++k) {
            {
                const unsigned pb = pw.pos_block<R>(j0, n);
#pragma unroll
                for (int q = 0; q < R; ++q) buf[k][q] = ldrec(f, pb + q);
            }
            __builtin_amdgcn_sched_barrier(0);  // issue in consumption order
            pw.next(g);
        }
        T xprev[R], dprev[R];
#pragma unroll
        for (int q = 0; q < R; ++q) xprev[q] = dprev[q] = 0;
        for (int T0 = 0; T0 < g.nT; T0 += PD) {
#pragma unroll
            for (int k = 0; k < PD; ++k) {
                const int par = (cw.T + 1) & 1;
                const T xs0 = dpp_shr1(xprev[R - 1], sh[wl][par][0]);
                const T ds0 = dpp_shr1(dprev[R - 1], sh[wl][par][1]);
                T xn[R];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const Rec4T<T> c = buf[k][q];
                    const int j = j0 + q, i = cw.T - j;
                    const T xs = q == 0 ? xs0 : xprev[q - 1];
                    const T ds = q == 0 ? ds0 : dprev[q - 1];
                    T x = c.a;
                    const T x1 = x - c.c * xs * ds;
                    x = (j > 0) ? x1 : x;
                    const T x2 = x - c.d * xprev[q] * dprev[q];
                    x = (i > 0) ? x2 : x;
                    if (j < g.ny && (unsigned)i < (unsigned)g.nx) st_a(bk, cw.offj + j, x);
                    xn[q] = x;
                }
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    xprev[q] = xn[q];
                    dprev[q] = keep(buf[k][q].b);
                }
                if (lane == 63) {
                    sh[wave][cw.T & 1][0] = xprev[R - 1];
                    sh[wave][cw.T & 1][1] = dprev[R - 1];
                }
                lds_barrier();
                // the slots are consumed above: keep their refill below them
                __builtin_amdgcn_sched_barrier(0);
                {
                    const unsigned pb = pw.pos_block<R>(j0, n);
#pragma unroll
                    for (int q = 0; q < R; ++q) buf[k][q] = ldrec(f, pb + q);
                }
                pw.next(g);
                cw.next(g);
            }
        }
    }
    __syncthreads();  // forward stores complete before the backward loads
    // ---- backward: z_r = (y_r Dinv_r - (A_re z_e) Dinv_r) - (A_rn z_n) Dinv_r ----
    {
        Rec4T<T> buf[PD][R];  // y, dinv, A_re, A_rn
        Walk cw, pw;
        cw.init(g, g.nT - 1);
        pw.init(g, g.nT - 1);
#pragma unroll
        for (int k = 0; k < PD; ++k) {
            {
                const unsigned pb = pw.pos_block<R>(j0, n);
#pragma unroll
                for (int q = 0; q < R; ++q) buf[k][q] = ldrec((const Rec4T<T> *)bk, pb + q);
            }
            __builtin_amdgcn_sched_barrier(0);  // issue in consumption order
            pw.prev(g);
        }
        T zprev[R];  // z of (i+1, j) per row
#pragma unroll
        for (int q = 0; q < R; ++q) zprev[q] = 0;
        for (int T0 = g.nT - 1; T0 >= 0; T0 -= PD) {
#pragma unroll
            for (int k = 0; k < PD; ++k) {
                // north of the lane's last row: lane t+1's first row, step T+1
                const T znl = dpp_shl1(zprev[0], sh[wr][(cw.T + 1) & 1][0]);
                T zq[R];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const Rec4T<T> c = buf[k][q];
                    const int j = j0 + q, i = cw.T - j;
                    const T zn = q == R - 1 ? znl : zprev[q + 1];
                    T z = c.a * c.b;
                    const T z1 = z - c.c * zprev[q] * c.b;
                    z = (i < g.nx - 1) ? z1 : z;
                    const T z2 = z - c.d * zn * c.b;
                    z = (j < g.ny - 1) ? z2 : z;
                    if (j < g.ny && cw.T >= 0 && (unsigned)i < (unsigned)g.nx) st_val(zp, cw.offj + j, z);
                    zq[q] = z;
                }
#pragma unroll
                for (int q = 0; q < R; ++q) zprev[q] = zq[q];
                if (lane == 0) sh[wave][cw.T & 1][0] = zprev[0];
                lds_barrier();
                __builtin_amdgcn_sched_barrier(0);
                {
                    const unsigned pb = pw.pos_block<R>(j0, n);
#pragma unroll
                    for (int q = 0; q < R; ++q) buf[k][q] = ldrec((const Rec4T<T> *)bk, pb + q);
                }
                pw.prev(g);
                cw.prev(g);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pipelined sweeps (one grid row per lane, no work-group barrier per step).
//
// Wave w needs, at step T, one value of wave w-1: its last lane's result of
// step T-1 (forward; backward the roles of w-1 and w+1 swap).  Instead of a
// barrier per step, waves hand these values over through a ring in LDS with
// two step counters per wave: prod[w] (steps wave w has published) and
// cons[w] (steps of wave w its successor has taken).  A consumer takes kHand
// values per wait, so the LDS round trip is paid once per kHand steps; a
// producer may run up to kRing steps ahead of its consumer.  The waves settle
// into a skew of a few steps and none waits for the whole work-group.
// Termination: the first producer of the chain depends on no one, each
// consumer's need is published by a producer that is blocked only by its own
// consumer's lag, and kRing >= 2*kHand, so every wait is eventually met; each
// wait is also capped (kSpinLimit sleeps), so the kernel always ends.
// The arithmetic of every row is that of band_apply_body.
// ---------------------------------------------------------------------------
constexpr int kHand = 4;            // handed-over values taken per wait
constexpr int kRing = 16;           // ring slots per wave
constexpr int kSpinLimit = 1 << 18;
static_assert(kPD % kHand == 0 && kRing >= 2 * kHand, "pipeline sizes");

__device__ inline int lds_acquire(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_release(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void wait_geq(const int *p, int need) {
    for (int s = 0; s < kSpinLimit && lds_acquire(p) < need; ++s) __builtin_amdgcn_s_sleep(1);
}

template <typename T>
__device__ inline void band_apply_pipe_body(const BandGeom &g, const Rec4T<T> *__restrict__ f, Rec4T<T> *bk,
                                            T *__restrict__ zp) {
    __shared__ T ring[16][kRing][2];  // [producer wave][slot][value | Dinv]
    __shared__ int prod[16], cons[16];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int n = g.nx * g.ny;
    const bool row = t < g.ny, S = t > 0, N = t < g.ny - 1;
    const int nwaves = blockDim.x >> 6;
    if (t < 16) {
        prod[t] = 0;
        cons[t] = 0;
    }
    __syncthreads();
    // ---- forward: wave w takes from w-1 ----
    {
        const bool take = wave > 0, give = wave + 1 < nwaves;
        Rec4T<T> buf[kPD];  // b, dinv, A_rs, A_rw
        Walk cw, pw;
        cw.init(g, 0);
        pw.init(g, 0);
#pragma unroll
        for (int k = 0; k < kPD; ++k) {
            buf[k] = ldrec(f, pw.pos(t, n));
            __builtin_amdgcn_sched_barrier(0);  // issue in consumption order
            pw.next(g);
        }
        T xprev = 0, dprev = 0;
        T cx[kHand], cd[kHand];  // handed-over values of steps base .. base+kHand-1
#pragma unroll
        for (int q = 0; q < kHand; ++q) cx[q] = cd[q] = 0;  // step -1: nothing below row 64w yet
        for (int T0 = 0; T0 < g.nT; T0 += kPD) {
#pragma unroll
            for (int k = 0; k < kPD; ++k) {
                const int Tk = T0 + k;
                if (k % kHand == 0 && give) wait_geq(&cons[wave], Tk + kHand - kRing);
                if (k % kHand == 1 && take) {
                    wait_geq(&prod[wave - 1], Tk - 1 + kHand);
#pragma unroll
                    for (int q = 0; q < kHand; ++q) {
                        cx[q] = ring[wave - 1][(Tk - 1 + q) % kRing][0];
                        cd[q] = ring[wave - 1][(Tk - 1 + q) % kRing][1];
                    }
                    if (lane == 0) lds_release(&cons[wave - 1], Tk - 1 + kHand);
                }
                const int qi = (k + kHand - 1) % kHand;  // step Tk-1 in the handed-over block
                const Rec4T<T> c = buf[k];
                const T xs = dpp_shr1(xprev, cx[qi]);
                const T ds = dpp_shr1(dprev, cd[qi]);
                const int i = cw.T - t;
                T x = c.a;
                const T x1 = x - c.c * xs * ds;
                x = S ? x1 : x;
                const T x2 = x - c.d * xprev * dprev;
                x = (i > 0) ? x2 : x;
                if (row && (unsigned)i < (unsigned)g.nx) st_a(bk, cw.offj + t, x);
                xprev = x;
                dprev = keep(c.b);
                if (give && lane == 63) {
                    ring[wave][Tk % kRing][0] = xprev;
                    ring[wave][Tk % kRing][1] = dprev;
                    if (k % kHand == kHand - 1) lds_release(&prod[wave], Tk + 1);
                }
                __builtin_amdgcn_sched_barrier(0);
                buf[k] = ldrec(f, pw.pos(t, n));
                pw.next(g);
                cw.next(g);
            }
        }
    }
    __syncthreads();  // forward stores complete before the backward loads; counters reset below
    if (t < 16) {
        prod[t] = 0;
        cons[t] = 0;
    }
    __syncthreads();
    // ---- backward: wave w takes from w+1 ----
    {
        const bool take = wave + 1 < nwaves, give = wave > 0;
        Rec4T<T> buf[kPD];  // y, dinv, A_re, A_rn
        Walk cw, pw;
        cw.init(g, g.nT - 1);
        pw.init(g, g.nT - 1);
#pragma unroll
        for (int k = 0; k < kPD; ++k) {
            buf[k] = ldrec((const Rec4T<T> *)bk, pw.pos(t, n));
            __builtin_amdgcn_sched_barrier(0);  // issue in consumption order
            pw.prev(g);
        }
        T zprev = 0;
        T cz[kHand];
#pragma unroll
        for (int q = 0; q < kHand; ++q) cz[q] = 0;
        for (int T0 = g.nT - 1, u0 = 0; T0 >= 0; T0 -= kPD, u0 += kPD) {
#pragma unroll
            for (int k = 0; k < kPD; ++k) {
                const int u = u0 + k;  // steps done in walk order
                if (k % kHand == 0 && give) wait_geq(&cons[wave], u + kHand - kRing);
                if (k % kHand == 1 && take) {
                    wait_geq(&prod[wave + 1], u - 1 + kHand);
#pragma unroll
                    for (int q = 0; q < kHand; ++q) cz[q] = ring[wave + 1][(u - 1 + q) % kRing][0];
                    if (lane == 0) lds_release(&cons[wave + 1], u - 1 + kHand);
                }
                const int qi = (k + kHand - 1) % kHand;
                const Rec4T<T> c = buf[k];
                const T zn = dpp_shl1(zprev, cz[qi]);
                const int i = cw.T - t;
                T z = c.a * c.b;
                const T z1 = z - c.c * zprev * c.b;
                z = (i < g.nx - 1) ? z1 : z;
                const T z2 = z - c.d * zn * c.b;
                z = N ? z2 : z;
                if (row && cw.T >= 0 && (unsigned)i < (unsigned)g.nx) st_val(zp, cw.offj + t, z);
                zprev = z;
                if (give && lane == 0) {
                    ring[wave][u % kRing][0] = zprev;
                    if (k % kHand == kHand - 1) lds_release(&prod[wave], u + 1);
                }
                __builtin_amdgcn_sched_barrier(0);
                buf[k] = ldrec((const Rec4T<T> *)bk, pw.pos(t, n));
                pw.prev(g);
                cw.prev(g);
            }
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_band_dilu_apply_pipe(BandGeom g, const Rec4T<T> *__restrict__ f,
                                                               Rec4T<T> *bk, T *__restrict__ zp, const KScal *ks) {
    if (ks && ks->done) return;
    band_apply_pipe_body<T>(g, f, bk, zp);
}

template <typename T>
__global__ void __launch_bounds__(1024) k_band_dilu_apply_pipe_cases(const BandGeom *__restrict__ gs,
                                                                     const int *__restrict__ off,
                                                                     const Rec4T<T> *__restrict__ f, Rec4T<T> *bk,
                                                                     T *__restrict__ zp, const KScal *ks) {
    const int k = blockIdx.x;
    if (ks && ks[k].done) return;
    const int o = off[k];
    band_apply_pipe_body<T>(gs[k], f + o, bk + o, zp + o);
}

// one row per lane, prefetch 16 steps ahead: up to 512 rows (two waves per SIMD
// leave room for the deeper register ring)
constexpr int kPDDeep = 16;
template <typename T>
__global__ void __launch_bounds__(512) k_band_dilu_apply_deep(BandGeom g, const Rec4T<T> *__restrict__ f,
                                                              Rec4T<T> *bk, T *__restrict__ zp, const KScal *ks) {
    if (ks && ks->done) return;
    band_apply_body<T, 1, kPDDeep>(g, f, bk, zp);
}

template <typename T>
__global__ void __launch_bounds__(512) k_band_dilu_apply_deep_cases(const BandGeom *__restrict__ gs,
                                                                    const int *__restrict__ off,
                                                                    const Rec4T<T> *__restrict__ f, Rec4T<T> *bk,
                                                                    T *__restrict__ zp, const KScal *ks) {
    const int k = blockIdx.x;
    if (ks && ks[k].done) return;
    const int o = off[k];
    band_apply_body<T, 1, kPDDeep>(gs[k], f + o, bk + o, zp + o);
}

template <typename T, int R>
__global__ void __launch_bounds__(1024 / R) k_band_dilu_apply(BandGeom g, const Rec4T<T> *__restrict__ f,
                                                              Rec4T<T> *bk, T *__restrict__ zp, const KScal *ks) {
    if (ks && ks->done) return;
    band_apply_body<T, R>(g, f, bk, zp);
}

// one work-group per case of a batch; a case whose solve has finished skips
template <typename T, int R>
__global__ void __launch_bounds__(1024 / R) k_band_dilu_apply_cases(const BandGeom *__restrict__ gs,
                                                                    const int *__restrict__ off,
                                                                    const Rec4T<T> *__restrict__ f, Rec4T<T> *bk,
                                                                    T *__restrict__ zp, const KScal *ks) {
    const int k = blockIdx.x;
    if (ks && ks[k].done) return;
    const int o = off[k];
    band_apply_body<T, R>(gs[k], f + o, bk + o, zp + o);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Anti-diagonal geometry of a banded mesh: offsets and packed row order.
// The fast path needs the full rectangle pattern (every row r = i + j*nx has
// exactly the neighbours r-nx, r-1, r+1, r+nx that exist on the grid), so the
// neighbour flags follow from (i, j).  Rows [c0, c0 + n) of the CSR pattern.
static bool full_rect(const std::vector<int> &rp, const std::vector<int> &col, int c0, int n, int nx) {
    if (nx <= 0 || n % nx != 0) return false;
    const int ny = n / nx;
    if (ny < 1 || ny > 1024) return false;
    int want[5];
    for (int r = 0; r < n; ++r) {
        const int i = r % nx, j = r / nx;
        int w = 0;
        if (j > 0) want[w++] = r - nx;
        if (i > 0) want[w++] = r - 1;
        want[w++] = r;
        if (i < nx - 1) want[w++] = r + 1;
        if (j < ny - 1) want[w++] = r + nx;
        const int a = rp[c0 + r], e = rp[c0 + r + 1];
        if (w != e - a) return false;
        for (int q = 0; q < w; ++q)
            if (col[a + q] - c0 != want[q]) return false;
    }
    return true;
}

// offj[T] = first packed position of diagonal T minus its first row jmin;
// iperm[p] = row (local) at packed position p.
static void anti_diagonals(int nx, int ny, std::vector<int> &offj, std::vector<int> &iperm) {
    const int nT = nx + ny - 1;
    offj.assign(nT, 0);
    iperm.assign((size_t)nx * ny, 0);
    int p = 0;
    for (int T = 0; T < nT; ++T) {
        int jmin = std::max(0, T - nx + 1), jmax = std::min(ny - 1, T);
        offj[T] = p - jmin;
        for (int j = jmin; j <= jmax; ++j, ++p) iperm[p] = (T - j) + j * nx;
    }
}

int band_setup(fvh_mesh *m) {
    Schedule &s = m->sched;
    const int nx = s.band_nx;
    if (nx <= 0) return FVH_OK;
    const int n = (int)m->n, ny = n / nx;
    // records are addressed by 32-bit byte offsets (32-byte records at most)
    if (!full_rect(m->h_rowptr, m->h_col, 0, n, nx) || (int64_t)n * (int64_t)sizeof(Rec4) >= ((int64_t)1 << 32)) {
        s.band_nx = 0;
        return FVH_OK;
    }
    std::vector<int> offj, iperm;
    anti_diagonals(nx, ny, offj, iperm);
    FVH_TRY(s.b_offj.upload(offj));
    FVH_TRY(s.b_iperm.upload(iperm));
    s.g.nx = nx;
    s.g.ny = ny;
    s.g.nT = nx + ny - 1;
    s.g.offj = s.b_offj.p;
    return skew_setup(m, {SkewRect{0, nx, ny}});
}


// Per-case geometry of a batch whose every case is a full rectangle; leaves
// cases->band false otherwise (the level-scheduled sweeps then cover the
// union, one case after another in the same levels).
int band_cases_setup(fvh_mesh *m) {
    CaseLayout &L = *m->cases;
    L.band = false;
    const char *force = std::getenv("FVH_SWEEP");
    if (force && std::strcmp(force, "levels") == 0) return FVH_OK;
    if ((int64_t)m->n * (int64_t)sizeof(Rec4) >= ((int64_t)1 << 32)) return FVH_OK;
    const auto &rp = m->h_rowptr, &col = m->h_col;
    std::vector<BandGeom> geoms(L.K);
    std::vector<int> nxs(L.K), offj_all, iperm_all(m->n), toff(L.K);
    int maxny = 1;
    for (int k = 0; k < L.K; ++k) {
        const int c0 = L.h_off[k], nk = L.h_off[k + 1] - c0;
        int nx = 0;
        for (int r = 0; r < nk && nx == 0; ++r)
            for (int q = rp[c0 + r]; q < m->h_diag[c0 + r]; ++q)
                if (r - (col[q] - c0) != 1) { nx = r - (col[q] - c0); break; }
        if (nx == 0) nx = nk;  // a single grid row
        if (!full_rect(rp, col, c0, nk, nx)) return FVH_OK;
        std::vector<int> offj, iperm;
        anti_diagonals(nx, nk / nx, offj, iperm);
        toff[k] = (int)offj_all.size();
        offj_all.insert(offj_all.end(), offj.begin(), offj.end());
        for (int q = 0; q < nk; ++q) iperm_all[c0 + q] = c0 + iperm[q];
        geoms[k] = BandGeom{nx, nk / nx, nx + nk / nx - 1, nullptr};
        nxs[k] = nx;
        maxny = std::max(maxny, nk / nx);
    }
    FVH_TRY(L.offj_all.upload(offj_all));
    for (int k = 0; k < L.K; ++k) geoms[k].offj = L.offj_all.p + toff[k];
    FVH_TRY(L.geoms.upload(geoms));
    FVH_TRY(L.b_iperm.upload(iperm_all));
    FVH_TRY(L.nxs.upload(nxs));
    L.band_R = 0;
    L.band_maxny = maxny;
    L.band = true;
    std::vector<SkewRect> rects(L.K);
    for (int k = 0; k < L.K; ++k) rects[k] = SkewRect{L.h_off[k], nxs[k], geoms[k].ny};
    return skew_setup(m, rects);
}

int band_alloc(fvh_eqn *e) {
    const int64_t n = e->m->n;
    BandStore &b = e->bd;
    if (b.ud.n != (size_t)n) {
        FVH_TRY(b.ud.alloc(n));
        FVH_TRY(b.us.alloc(n));
        FVH_TRY(b.uw.alloc(n));
        FVH_TRY(b.f.alloc(n));
        FVH_TRY(b.b.alloc(n));
        FVH_TRY(b.zp.alloc(n));
        b.r = BandRecs{b.ud.p, b.us.p, b.uw.p, b.f.p, b.b.p, nullptr, nullptr};
    }
    return FVH_OK;
}

static int threads_for(int ny, int R) { return (((ny + R - 1) / R + 63) / 64) * 64; }

// Launch variant of the sweeps, chosen per mesh by timing all on the mesh's
// own records at its first update (FVH_BAND_ROWS=1|2|3 forces one):
//   1, 2: one / two grid rows per lane, a work-group barrier per step (two
//         rows let four waves, one per SIMD, share the per-step overhead);
//   3:    one row per lane, waves pipelined through LDS without barriers;
//   4:    one row per lane, a barrier per step, loads issued 16 steps ahead
//         (grids up to 512 rows).
// The variant never changes a result: the arithmetic of every row is fixed.
static int forced_rows() {
    const char *env = std::getenv("FVH_BAND_ROWS");
    if (env && env[0] >= '1' && env[0] <= '4' && env[1] == 0) return env[0] - '0';
    return 0;
}
static int lanes_rows(int variant) { return variant == 2 ? 2 : 1; }

struct BandLaunch {
    bool cases;
    int R, threads, blocks;
};

static BandLaunch band_launch(const fvh_mesh *m, int V) {
    const int R = lanes_rows(V);
    if (m->cases && m->cases->band)
        return BandLaunch{true, V, threads_for(m->cases->band_maxny, R), m->cases->K};
    return BandLaunch{false, V, threads_for(m->sched.g.ny, R), 1};
}

static int band_rows(const fvh_mesh *m) {
    const int f = forced_rows();
    if (f) return f;
    const int t = (m->cases && m->cases->band) ? m->cases->band_R : m->sched.band_R;
    if (t) return t;
    return 1;  // before tuning
}

template <typename T>
static int launch_sweep(const fvh_mesh *m, int R, const Rec4T<T> *f, Rec4T<T> *bk, T *zp, const KScal *ks) {
    const BandLaunch b = band_launch(m, R);
    if (b.cases) {
        const CaseLayout &L = *m->cases;
        if (R == 4)
            hipLaunchKernelGGL((k_band_dilu_apply_deep_cases<T>), dim3(b.blocks), dim3(b.threads), 0, stream(),
                               L.geoms.p, L.off.p, f, bk, zp, ks);
        else if (R == 3)
            hipLaunchKernelGGL((k_band_dilu_apply_pipe_cases<T>), dim3(b.blocks), dim3(b.threads), 0, stream(),
                               L.geoms.p, L.off.p, f, bk, zp, ks);
        else if (R == 2)
            hipLaunchKernelGGL((k_band_dilu_apply_cases<T, 2>), dim3(b.blocks), dim3(b.threads), 0, stream(), L.geoms.p,
                               L.off.p, f, bk, zp, ks);
        else
            hipLaunchKernelGGL((k_band_dilu_apply_cases<T, 1>), dim3(b.blocks), dim3(b.threads), 0, stream(), L.geoms.p,
                               L.off.p, f, bk, zp, ks);
    } else {
        const BandGeom &g = m->sched.g;
        if (R == 4)
            hipLaunchKernelGGL((k_band_dilu_apply_deep<T>), dim3(1), dim3(b.threads), 0, stream(), g, f, bk, zp, ks);
        else if (R == 3)
            hipLaunchKernelGGL((k_band_dilu_apply_pipe<T>), dim3(1), dim3(b.threads), 0, stream(), g, f, bk, zp, ks);
        else if (R == 2)
            hipLaunchKernelGGL((k_band_dilu_apply<T, 2>), dim3(1), dim3(b.threads), 0, stream(), g, f, bk, zp, ks);
        else
            hipLaunchKernelGGL((k_band_dilu_apply<T, 1>), dim3(1), dim3(b.threads), 0, stream(), g, f, bk, zp, ks);
    }
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

static int launch_update(const fvh_mesh *m, int V, BandRecs r) {
    const int R = lanes_rows(V);
    const BandLaunch b = band_launch(m, R);
    if (b.cases) {
        const CaseLayout &L = *m->cases;
        if (R == 2)
            hipLaunchKernelGGL(k_band_dilu_update_cases<2>, dim3(b.blocks), dim3(b.threads), 0, stream(), L.geoms.p,
                               L.off.p, r);
        else
            hipLaunchKernelGGL(k_band_dilu_update_cases<1>, dim3(b.blocks), dim3(b.threads), 0, stream(), L.geoms.p,
                               L.off.p, r);
    } else {
        if (R == 2)
            hipLaunchKernelGGL(k_band_dilu_update<2>, dim3(1), dim3(b.threads), 0, stream(), m->sched.g, r);
        else
            hipLaunchKernelGGL(k_band_dilu_update<1>, dim3(1), dim3(b.threads), 0, stream(), m->sched.g, r);
    }
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// Times the FP64 sweep in each launch variant on the freshly packed records
// (the launches only rewrite the sweep's own scratch fields) and keeps the
// fastest.
static int tune_rows(fvh_eqn *e) {
    fvh_mesh *m = e->m;
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    float best = 0.f;
    int bestR = 1;
    const int tallest = (m->cases && m->cases->band) ? m->cases->band_maxny : m->sched.g.ny;
    for (int R = 1; R <= 4; ++R) {
        if (R == 4 && tallest > 512) continue;
        FVH_TRY(launch_sweep<double>(m, R, e->bd.f.p, e->bd.b.p, e->bd.zp.p, nullptr));  // warm
        HIP_TRY(hipEventRecord(a, stream()));
        for (int rep = 0; rep < 3; ++rep)
            FVH_TRY(launch_sweep<double>(m, R, e->bd.f.p, e->bd.b.p, e->bd.zp.p, nullptr));
        HIP_TRY(hipEventRecord(b, stream()));
        HIP_TRY(hipEventSynchronize(b));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, a, b));
        if (R == 1 || t < best) {
            best = t;
            bestR = R;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (m->cases && m->cases->band) m->cases->band_R = bestR;
    else m->sched.band_R = bestR;
    return FVH_OK;
}

static const int *band_iperm(const fvh_mesh *m) {
    return (m->cases && m->cases->band) ? m->cases->b_iperm.p : m->sched.b_iperm.p;
}

int band_dilu_update(fvh_eqn *e, const double *val) {
    fvh_mesh *m = e->m;
    if (skew_active(m)) return skew_update(e, val);
    const Schedule &s = m->sched;
    FVH_TRY(band_alloc(e));
    const int n = (int)m->n;
    if (m->cases && m->cases->band) {
        const CaseLayout &L = *m->cases;
        hipLaunchKernelGGL(k_band_pack, grid1d(n), dim3(256), 0, stream(), n, 0, L.cell_case.p, L.off.p, L.nxs.p,
                           L.b_iperm.p, m->d.rowptr, m->d.diag, val, e->bd.r);
    } else {
        hipLaunchKernelGGL(k_band_pack, grid1d(n), dim3(256), 0, stream(), n, s.g.nx, (const int *)nullptr,
                           (const int *)nullptr, (const int *)nullptr, s.b_iperm.p, m->d.rowptr, m->d.diag, val,
                           e->bd.r);
    }
    HIP_TRY(hipGetLastError());
    FVH_TRY(launch_update(m, band_rows(m), e->bd.r));
    const bool tuned = (m->cases && m->cases->band) ? m->cases->band_R != 0 : m->sched.band_R != 0;
    if (!tuned && !forced_rows()) FVH_TRY(tune_rows(e));
    return FVH_OK;
}

int band_rows_per_lane(const fvh_mesh *m) { return skew_active(m) ? 5 : band_rows(m); }

int band_to32(fvh_eqn *e) {
    if (skew_active(e->m)) return skew_to32(e);
    const int n = (int)e->m->n;
    BandStore &b = e->bd;
    if (b.f32.n != (size_t)n) {
        FVH_TRY(b.f32.alloc(n));
        FVH_TRY(b.b32.alloc(n));
        FVH_TRY(b.zp32.alloc(n));
        b.r.f32 = b.f32.p;
        b.r.b32 = b.b32.p;
    }
    hipLaunchKernelGGL(k_band_to32, grid1d(n), dim3(256), 0, stream(), n, b.f.p, b.b.p, b.f32.p, b.b32.p);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

template <typename T>
static int band_apply_t(fvh_eqn *e, const T *x, T *y, const KScal *ks, Rec4T<T> *f, Rec4T<T> *bk, T *zp) {
    fvh_mesh *m = e->m;
    const int n = (int)m->n;
    const bool cases = m->cases && m->cases->band;
    // per-case solves: the gather / scatter cover every case (a finished
    // case's rows are not read again), the sweep skips finished cases
    const KScal *ks_all = cases ? nullptr : ks;
    hipLaunchKernelGGL((k_band_gather<T>), grid1d(n), dim3(256), 0, stream(), n, band_iperm(m), x, f, ks_all);
    HIP_TRY(hipGetLastError());
    FVH_TRY(launch_sweep<T>(m, band_rows(m), f, bk, zp, ks));
    hipLaunchKernelGGL((k_band_scatter<T>), grid1d(n), dim3(256), 0, stream(), n, band_iperm(m), zp, y, ks_all);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int band_dilu_apply(fvh_eqn *e, const double *x, double *y, const KScal *ks) {
    if (skew_active(e->m)) return skew_apply(e, x, y, ks);
    return band_apply_t<double>(e, x, y, ks, e->bd.f.p, e->bd.b.p, e->bd.zp.p);
}

int band_dilu_apply32(fvh_eqn *e, const float *x, float *y, const KScal *ks) {
    if (skew_active(e->m)) return skew_apply32(e, x, y, ks);
    if (!e->bd.f32.p) return set_error(FVH_ERR_ARG, "FP32 sweep records not prepared");
    return band_apply_t<float>(e, x, y, ks, e->bd.f32.p, e->bd.b32.p, e->bd.zp32.p);
}

// Natural-order copy of the packed reciprocal diagonal.
int band_dinv_natural(fvh_eqn *e, double *out) {
    fvh_mesh *m = e->m;
    if (skew_active(m)) return skew_dinv_natural(e, out);
    const int n = (int)m->n;
    hipLaunchKernelGGL(k_band_dinv_natural, grid1d(n), dim3(256), 0, stream(), n, band_iperm(m), e->bd.r.f, out);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

}  // namespace fvh

extern "C" int fvh_mesh_tuning(const fvh_mesh *m, int32_t *rows_per_lane) {
    if (!m || !rows_per_lane) return fvh::set_error(FVH_ERR_ARG, "null argument");
    *rows_per_lane = fvh::band_path(m) ? fvh::band_rows_per_lane(m) : 0;
    return FVH_OK;
}
