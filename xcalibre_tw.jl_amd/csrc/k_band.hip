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

// Work-group barrier that orders LDS only: global loads issued for later
// steps stay in flight across it (a full __syncthreads waits for them).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Cross-lane shifts; the lane without a source keeps `old`.
__device__ inline double dpp_shr1(double v, double old) {  // lane l <- lane l-1; lane 0 <- old
    int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), 0x138, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), 0x138, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double dpp_shl1(double v, double old) {  // lane l <- lane l+1; lane 63 <- old
    int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), 0x130, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), 0x130, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline float dpp_shr1(float v, float old) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ inline float dpp_shl1(float v, float old) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x130, 0xF, 0xF, false));
}

// A register copy the compiler keeps: a value carried to the next step is
// copied out of its prefetch slot before the slot is refilled, so the refill
// lands in the slot's own registers (otherwise the loop rotates the whole
// ring through copies that wait for every load in flight).
__device__ inline double keep(double v) {
    double r;
    asm volatile("v_mov_b64 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ inline float keep(float v) {
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

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
__device__ inline Rec6 ldrec(const Rec6 *base, unsigned idx) {
    const char *p = rec_at(base, idx);
    const double2 u = *reinterpret_cast<const double2 *>(p), v = *reinterpret_cast<const double2 *>(p + 16),
                  w = *reinterpret_cast<const double2 *>(p + 32);
    return Rec6{u.x, u.y, v.x, v.y, w.x, w.y};
}
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
__global__ void k_band_pack(int n, int nx, const int *__restrict__ iperm, const int *__restrict__ rowptr,
                            const int *__restrict__ diag, const double *__restrict__ val, BandRecs br) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int r = iperm[p];
    const int i = r % nx;
    const bool S = r >= nx, W = i > 0, E = i < nx - 1, N = r + nx < n;
    const int d = diag[r];
    Rec6 u;
    u.a = val[d];
    u.b = S ? val[rowptr[r]] : 0.0;               // A[r, r-nx]
    u.c = S ? val[rowptr[r - nx + 1] - 1] : 0.0;  // A[r-nx, r]
    u.d = W ? val[d - 1] : 0.0;                   // A[r, r-1]
    u.e = W ? val[diag[r - 1] + 1] : 0.0;         // A[r-1, r]
    u.f = 0.0;
    br.u[p] = u;
    br.f[p].c = u.b;
    br.f[p].d = u.d;
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
};

// ---------------------------------------------------------------------------
// Diagonal recurrence (update_dilu_diagonal!)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_band_dilu_update(BandGeom g, BandRecs br) {
    __shared__ double sh_d[16][2];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wl = wave > 0 ? wave - 1 : 0;
    const int n = g.nx * g.ny;
    const bool row = t < g.ny, S = t > 0;
    Rec6 buf[kPD];
    Walk cw, pw;
    cw.init(g, 0);
    pw.init(g, 0);
#pragma unroll
    for (int k = 0; k < kPD; ++k) {
        buf[k] = ldrec(br.u, pw.pos(t, n));
        __builtin_amdgcn_sched_barrier(0);  // issue in consumption order
        pw.next(g);
    }
    double dprev = 0.0;  // D of (i-1, j)
    for (int T0 = 0; T0 < g.nT; T0 += kPD) {
#pragma unroll
        for (int k = 0; k < kPD; ++k) {
            const Rec6 c = buf[k];  // ad, ls, ts, lw, tw
            const double ds = dpp_shr1(dprev, sh_d[wl][(cw.T + 1) & 1]);
            const int i = cw.T - t;
            double d = c.a;
            const double d1 = d - c.b * c.c / ds;
            d = S ? d1 : d;
            const double d2 = d - c.d * c.e / dprev;
            d = (i > 0) ? d2 : d;
            if (row && (unsigned)i < (unsigned)g.nx) {
                const double r = 1.0 / d;
                const unsigned p = cw.offj + t;
                br.f[p].b = r;
                br.b[p].b = r;
            }
            dprev = d;
            if (lane == 63) sh_d[wave][cw.T & 1] = dprev;
            lds_barrier();
            // the slot is consumed above: keep its refill below it
            __builtin_amdgcn_sched_barrier(0);
            buf[k] = ldrec(br.u, pw.pos(t, n));
            pw.next(g);
            cw.next(g);
        }
    }
}

// ---------------------------------------------------------------------------
// ldiv!: forward then backward sweep over the packed records, in FP64 (the
// reference's arithmetic) or FP32 (the inner solves of mixed precision)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(1024) k_band_dilu_apply(BandGeom g, const Rec4T<T> *__restrict__ f, Rec4T<T> *bk,
                                                          T *__restrict__ zp, const KScal *ks) {
    if (ks && ks->done) return;
    // [wave][step parity][value | Dinv] of the wave's edge lane
    __shared__ T sh[16][2][2];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int n = g.nx * g.ny;
    const bool row = t < g.ny, S = t > 0, N = t < g.ny - 1;
    const int nwaves = (blockDim.x + 63) >> 6;
    const int wl = wave > 0 ? wave - 1 : 0, wr = wave + 1 < nwaves ? wave + 1 : wave;
    // ---- forward: y_r = (b_r - (A_rs y_s) Dinv_s) - (A_rw y_w) Dinv_w ----
    {
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
        for (int T0 = 0; T0 < g.nT; T0 += kPD) {
#pragma unroll
            for (int k = 0; k < kPD; ++k) {
                const Rec4T<T> c = buf[k];
                const int par = (cw.T + 1) & 1;
                const T xs = dpp_shr1(xprev, sh[wl][par][0]);
                const T ds = dpp_shr1(dprev, sh[wl][par][1]);
                const int i = cw.T - t;
                T x = c.a;
                const T x1 = x - c.c * xs * ds;
                x = S ? x1 : x;
                const T x2 = x - c.d * xprev * dprev;
                x = (i > 0) ? x2 : x;
                if (row && (unsigned)i < (unsigned)g.nx) st_a(bk, cw.offj + t, x);
                xprev = x;
                dprev = keep(c.b);
                if (lane == 63) {
                    sh[wave][cw.T & 1][0] = xprev;
                    sh[wave][cw.T & 1][1] = dprev;
                }
                lds_barrier();
                // the slot is consumed above: keep its refill below it
                __builtin_amdgcn_sched_barrier(0);
                buf[k] = ldrec(f, pw.pos(t, n));
                pw.next(g);
                cw.next(g);
            }
        }
    }
    __syncthreads();  // forward stores complete before the backward loads
    // ---- backward: z_r = (y_r Dinv_r - (A_re z_e) Dinv_r) - (A_rn z_n) Dinv_r ----
    {
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
        for (int T0 = g.nT - 1; T0 >= 0; T0 -= kPD) {
#pragma unroll
            for (int k = 0; k < kPD; ++k) {
                const Rec4T<T> c = buf[k];
                const T zn = dpp_shl1(zprev, sh[wr][(cw.T + 1) & 1][0]);
                const int i = cw.T - t;
                T z = c.a * c.b;
                const T z1 = z - c.c * zprev * c.b;
                z = (i < g.nx - 1) ? z1 : z;
                const T z2 = z - c.d * zn * c.b;
                z = N ? z2 : z;
                if (row && cw.T >= 0 && (unsigned)i < (unsigned)g.nx) st_val(zp, cw.offj + t, z);
                zprev = z;
                if (lane == 0) sh[wave][cw.T & 1][0] = zprev;
                lds_barrier();
                __builtin_amdgcn_sched_barrier(0);
                buf[k] = ldrec((const Rec4T<T> *)bk, pw.pos(t, n));
                pw.prev(g);
                cw.prev(g);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Anti-diagonal geometry of a banded mesh: offsets and packed row order.
// The fast path needs the full rectangle pattern (every row r = i + j*nx has
// exactly the neighbours r-nx, r-1, r+1, r+nx that exist on the grid), so the
// neighbour flags follow from (i, j).
int band_setup(fvh_mesh *m) {
    Schedule &s = m->sched;
    const int nx = s.band_nx;
    if (nx <= 0) return FVH_OK;
    const int n = (int)m->n, ny = n / nx;
    const auto &rp = m->h_rowptr, &col = m->h_col;
    bool full = ny <= 1024 && ny >= 1;
    for (int r = 0; r < n && full; ++r) {
        const int i = r % nx, j = r / nx;
        std::vector<int> want;
        if (j > 0) want.push_back(r - nx);
        if (i > 0) want.push_back(r - 1);
        want.push_back(r);
        if (i < nx - 1) want.push_back(r + 1);
        if (j < ny - 1) want.push_back(r + nx);
        if ((int)want.size() != rp[r + 1] - rp[r]) { full = false; break; }
        for (size_t q = 0; q < want.size(); ++q)
            if (col[rp[r] + q] != want[q]) { full = false; break; }
    }
    // records are addressed by 32-bit byte offsets (48-byte records at most)
    if (!full || (int64_t)n * (int64_t)sizeof(Rec6) >= ((int64_t)1 << 32)) {
        s.band_nx = 0;
        return FVH_OK;
    }
    const int nT = nx + ny - 1;
    std::vector<int> offj(nT), iperm(n);
    int p = 0;
    for (int T = 0; T < nT; ++T) {
        int jmin = std::max(0, T - nx + 1), jmax = std::min(ny - 1, T);
        offj[T] = p - jmin;
        for (int j = jmin; j <= jmax; ++j, ++p) iperm[p] = (T - j) + j * nx;
    }
    FVH_TRY(s.b_offj.upload(offj));
    FVH_TRY(s.b_iperm.upload(iperm));
    s.g.nx = nx;
    s.g.ny = ny;
    s.g.nT = nT;
    s.g.offj = s.b_offj.p;
    return FVH_OK;
}

int band_alloc(fvh_eqn *e) {
    const int64_t n = e->m->n;
    BandStore &b = e->bd;
    if (b.u.n != (size_t)n) {
        FVH_TRY(b.u.alloc(n));
        FVH_TRY(b.f.alloc(n));
        FVH_TRY(b.b.alloc(n));
        FVH_TRY(b.zp.alloc(n));
        b.r = BandRecs{b.u.p, b.f.p, b.b.p, nullptr, nullptr};
    }
    return FVH_OK;
}

static int threads_for(int ny) { return ((ny + 63) / 64) * 64; }

int band_dilu_update(fvh_eqn *e, const double *val) {
    fvh_mesh *m = e->m;
    const Schedule &s = m->sched;
    FVH_TRY(band_alloc(e));
    const int n = (int)m->n;
    hipLaunchKernelGGL(k_band_pack, grid1d(n), dim3(256), 0, stream(), n, s.g.nx, s.b_iperm.p, m->d.rowptr, m->d.diag,
                       val, e->bd.r);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_band_dilu_update, dim3(1), dim3(threads_for(s.g.ny)), 0, stream(), s.g, e->bd.r);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int band_to32(fvh_eqn *e) {
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
    const Schedule &s = m->sched;
    const int n = (int)m->n;
    hipLaunchKernelGGL((k_band_gather<T>), grid1d(n), dim3(256), 0, stream(), n, s.b_iperm.p, x, f, ks);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_band_dilu_apply<T>), dim3(1), dim3(threads_for(s.g.ny)), 0, stream(), s.g, f, bk, zp, ks);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_band_scatter<T>), grid1d(n), dim3(256), 0, stream(), n, s.b_iperm.p, zp, y, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int band_dilu_apply(fvh_eqn *e, const double *x, double *y, const KScal *ks) {
    return band_apply_t<double>(e, x, y, ks, e->bd.f.p, e->bd.b.p, e->bd.zp.p);
}

int band_dilu_apply32(fvh_eqn *e, const float *x, float *y, const KScal *ks) {
    if (!e->bd.f32.p) return set_error(FVH_ERR_ARG, "FP32 sweep records not prepared");
    return band_apply_t<float>(e, x, y, ks, e->bd.f32.p, e->bd.b32.p, e->bd.zp32.p);
}

// Natural-order copy of the packed reciprocal diagonal.
int band_dinv_natural(fvh_eqn *e, double *out) {
    fvh_mesh *m = e->m;
    const int n = (int)m->n;
    hipLaunchKernelGGL(k_band_dinv_natural, grid1d(n), dim3(256), 0, stream(), n, m->sched.b_iperm.p, e->bd.r.f, out);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

}  // namespace fvh
