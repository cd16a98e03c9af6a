// Level-scheduled DILU recurrences on unstructured meshes, from per-row
// records in level order.
//
// The schedule is the one of k_sweep.hip (rows of a level have no dependency
// on each other; one work-group walks the levels with a barrier between
// them), and every row evaluates the same expression in the same order
// (preconditioners_1_DILU.jl:56-162), so the results are bit-identical.
// What changes is what a level waits for.  The row-by-row kernels chase a
// chain of dependent loads per level (level bounds, row index, row extent,
// column, value, then the neighbour's result).  Here each row's static data
// - its index, its lower (or upper) neighbours, their coefficients and the
// neighbours' reciprocal pivots - sits in one record at its level position,
// and the records (and the row's own input) of the next level are loaded
// while the current level computes.  After each barrier a level waits only
// for its neighbours' fresh results, one memory round trip.
//
// A record holds at most kLevDeps neighbours; meshes with more per row keep
// the row-by-row kernels.  The FP32 sweeps (inner solves of mixed precision)
// use coefficients premultiplied by their pivots in FP64 and rounded once,
// with fused multiply-adds in neighbour order - the arithmetic of every FP32
// DILU sweep of the library.
#include <cstdlib>
#include <cstring>

#include "fvh_eqn.h"
#include "k_level.h"

namespace fvh {

constexpr int kLevBlock = 1024;

// Record of one row at its level position: i, neighbour count m, neighbours
// j[q] (ascending, as the reference visits them), and per mode
//   update:   a[q] = A_ij * A_ji (the recurrence's product), d[0] = A_ii
//   forward:  a[q] = A_ij, d[q] = Dinv_j
//   backward: a[q] = A_ij (upper), d[0] = Dinv_i
template <typename T>
struct alignas(16) LevRec {
    int i, m, j[kLevDeps], pad[2];
    T a[kLevDeps], d[kLevDeps];
};

// ---------------------------------------------------------------------------
// Record packing (all CUs, once per preconditioner update)
// ---------------------------------------------------------------------------
__global__ void k_lev_pack(int n, const int *__restrict__ rows, const int *__restrict__ rowptr,
                           const int *__restrict__ col, const int *__restrict__ diag, const int *__restrict__ tslot,
                           const double *__restrict__ val, const double *__restrict__ Dinv, int mode,
                           LevRec<double> *__restrict__ R) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = rows[k];
    LevRec<double> r{};
    r.i = i;
    const int p0 = mode == 2 ? diag[i] + 1 : rowptr[i];
    const int p1 = mode == 2 ? rowptr[i + 1] : diag[i];
    r.m = p1 - p0;
    for (int q = 0; q < kLevDeps; ++q) r.j[q] = i;  // unused slots point at the row itself (never read)
    for (int q = 0; q < r.m; ++q) {
        const int p = p0 + q, j = col[p];
        r.j[q] = j;
        if (mode == 0) {
            r.a[q] = val[p] * val[tslot[p]];
        } else {
            r.a[q] = val[p];
            if (mode == 1) r.d[q] = Dinv[j];
        }
    }
    if (mode == 0) r.d[0] = val[diag[i]];
    if (mode == 2) r.d[0] = Dinv[i];
    R[k] = r;
}

// FP32 records of the inner solves of mixed precision: coefficients
// premultiplied by their pivot in FP64 and rounded once (forward: A_ij
// Dinv_j; backward: A_ij Dinv_i, with d[0] = Dinv_i); the FP32 sweeps then
// use fused multiply-adds.
__global__ void k_lev_to32(int n, int mode, const LevRec<double> *__restrict__ R, LevRec<float> *__restrict__ R32) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const LevRec<double> r = R[k];
    LevRec<float> s{};
    s.i = r.i;
    s.m = r.m;
    for (int q = 0; q < kLevDeps; ++q) {
        s.j[q] = r.j[q];
        s.a[q] = (float)(r.a[q] * (mode == 1 ? r.d[q] : r.d[0]));
        s.d[q] = (float)r.d[q];
    }
    R32[k] = s;
}

// ---------------------------------------------------------------------------
// The level walk.  MODE 0: diagonal recurrence (y = D, y2 = Dinv); 1: forward
// sweep y = L-solve of x; 2: backward sweep in place on y.
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__device__ inline void lev_row(const LevRec<T> &r, T in, const T *y, T *out, T *out2) {
    if (MODE == 0) {  // D_i = A_ii - sum (A_ij A_ji) / D_j
        T d = r.d[0];
#pragma unroll
        for (int q = 0; q < kLevDeps; ++q)
            if (q < r.m) d = d - r.a[q] / y[r.j[q]];
        out[r.i] = d;
        out2[r.i] = T(1) / d;
    } else if (MODE == 1) {  // y_i = x_i - sum (A_ij y_j) Dinv_j
        T s = in;
#pragma unroll
        for (int q = 0; q < kLevDeps; ++q)
            if (q < r.m) {
                if constexpr (sizeof(T) == 4) s = __builtin_fmaf(-r.a[q], y[r.j[q]], s);  // a = A_ij Dinv_j
                else s = s - r.a[q] * y[r.j[q]] * r.d[q];
            }
        out[r.i] = s;
    } else {  // z_i = y_i Dinv_i - sum (A_ij z_j) Dinv_i
        const T di = r.d[0];
        T s = in * di;
#pragma unroll
        for (int q = 0; q < kLevDeps; ++q)
            if (q < r.m) {
                if constexpr (sizeof(T) == 4) s = __builtin_fmaf(-r.a[q], y[r.j[q]], s);  // a = A_ij Dinv_i
                else s = s - r.a[q] * y[r.j[q]] * di;
            }
        out[r.i] = s;
    }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kLevBlock)
    k_lev_walk(int nlev, const int *__restrict__ lptr, const LevRec<T> *__restrict__ R, const T *__restrict__ x,
               T *y, T *y2, const KScal *ks) {
    if (ks && ks->done) return;
    const int t = threadIdx.x;
    // the row input (x for the forward sweep, the forward result for the
    // backward sweep) is final before the walk starts
    const T *in = MODE == 1 ? x : y;
    int beg = lptr[0], end = lptr[1];
    LevRec<T> r{};
    T v = T(0);
    if (beg + t < end) {
        r = R[beg + t];
        if (MODE != 0) v = in[r.i];
    }
    for (int l = 0; l < nlev; ++l) {
        // next level's first records, loaded while this level computes
        const int nend = l + 1 < nlev ? lptr[l + 2] : end;
        LevRec<T> rn{};
        T vn = T(0);
        if (l + 1 < nlev && end + t < nend) {
            rn = R[end + t];
            if (MODE != 0) vn = in[rn.i];
        }
        if (beg + t < end) lev_row<T, MODE>(r, v, y, y, y2);
        for (int k = beg + t + kLevBlock; k < end; k += kLevBlock) {  // levels wider than the block
            const LevRec<T> rk = R[k];
            lev_row<T, MODE>(rk, MODE != 0 ? in[rk.i] : T(0), y, y, y2);
        }
        __syncthreads();
        r = rn;
        v = vn;
        beg = end;
        end = nend;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static dim3 lev_grid(int n) { return dim3((unsigned)((n + 255) / 256)); }

bool lev_path(const fvh_mesh *m) {
    if (m->sched.nlev <= 0 || m->sched.max_deps > kLevDeps) return false;
    const char *env = std::getenv("FVH_LEVEL_RECORDS");
    return !(env && env[0] == '0' && env[1] == 0);
}

int lev_dilu_update(fvh_eqn *e, const double *val) {
    fvh_mesh *m = e->m;
    const Schedule &s = m->sched;
    const int n = (int)m->n;
    LevelStore &L = e->lv;
    if (L.upd.n != (size_t)n) {
        FVH_TRY(L.upd.alloc(n));
        FVH_TRY(L.fwd.alloc(n));
        FVH_TRY(L.bwd.alloc(n));
    }
    auto *U = reinterpret_cast<LevRec<double> *>(L.upd.p);
    auto *Fw = reinterpret_cast<LevRec<double> *>(L.fwd.p);
    auto *Bw = reinterpret_cast<LevRec<double> *>(L.bwd.p);
    hipLaunchKernelGGL(k_lev_pack, lev_grid(n), dim3(256), 0, stream(), n, s.frows.p, m->d.rowptr, m->d.col, m->d.diag,
                       m->d.tslot, val, (const double *)nullptr, 0, U);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_lev_walk<double, 0>), dim3(1), dim3(kLevBlock), 0, stream(), s.nlev, s.fptr.p, U,
                       (const double *)nullptr, e->D.p, e->Dinv.p, (const KScal *)nullptr);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_lev_pack, lev_grid(n), dim3(256), 0, stream(), n, s.frows.p, m->d.rowptr, m->d.col, m->d.diag,
                       m->d.tslot, val, e->Dinv.p, 1, Fw);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_lev_pack, lev_grid(n), dim3(256), 0, stream(), n, s.brows.p, m->d.rowptr, m->d.col, m->d.diag,
                       m->d.tslot, val, e->Dinv.p, 2, Bw);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

template <typename T>
static int lev_apply_t(fvh_eqn *e, const LevRec<T> *Fw, const LevRec<T> *Bw, const T *x, T *y, const KScal *ks) {
    const Schedule &s = e->m->sched;
    hipLaunchKernelGGL((k_lev_walk<T, 1>), dim3(1), dim3(kLevBlock), 0, stream(), s.nlev, s.fptr.p, Fw, x, y,
                       (T *)nullptr, ks);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_lev_walk<T, 2>), dim3(1), dim3(kLevBlock), 0, stream(), s.nlev, s.bptr.p, Bw, x, y,
                       (T *)nullptr, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int lev_dilu_apply(fvh_eqn *e, const double *x, double *y, const KScal *ks) {
    const LevelStore &L = e->lv;
    if (!L.fwd.p) return set_error(FVH_ERR_ARG, "DILU records not prepared");
    return lev_apply_t<double>(e, reinterpret_cast<const LevRec<double> *>(L.fwd.p),
                               reinterpret_cast<const LevRec<double> *>(L.bwd.p), x, y, ks);
}

int lev_to32(fvh_eqn *e) {
    LevelStore &L = e->lv;
    const int n = (int)e->m->n;
    if (L.fwd32.n != (size_t)n) {
        FVH_TRY(L.fwd32.alloc(n));
        FVH_TRY(L.bwd32.alloc(n));
    }
    hipLaunchKernelGGL(k_lev_to32, lev_grid(n), dim3(256), 0, stream(), n, 1,
                       reinterpret_cast<const LevRec<double> *>(L.fwd.p), reinterpret_cast<LevRec<float> *>(L.fwd32.p));
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_lev_to32, lev_grid(n), dim3(256), 0, stream(), n, 2,
                       reinterpret_cast<const LevRec<double> *>(L.bwd.p), reinterpret_cast<LevRec<float> *>(L.bwd32.p));
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int lev_dilu_apply32(fvh_eqn *e, const float *x, float *y, const KScal *ks) {
    const LevelStore &L = e->lv;
    if (!L.fwd32.p) return set_error(FVH_ERR_ARG, "FP32 DILU records not prepared");
    return lev_apply_t<float>(e, reinterpret_cast<const LevRec<float> *>(L.fwd32.p),
                              reinterpret_cast<const LevRec<float> *>(L.bwd32.p), x, y, ks);
}

static_assert(sizeof(LevRec<double>) == kLevRecBytes64 && sizeof(LevRec<float>) == kLevRecBytes32,
              "record sizes");

}  // namespace fvh
