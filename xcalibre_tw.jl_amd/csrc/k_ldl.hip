// LDL preconditioner on the device: y = P^T L^-T D^-1 L^-1 P x with the
// factor of the set-up matrix (ldl_host.cpp), frozen after the first
// update as the reference freezes opLDL (preconditioners_0_types.jl:36-42,
// preconditioners_2_functions.jl:52-57).
//
// The two triangular solves walk the ranges of the nested-dissection order
// (k_ldl_range, below).  FVH_LDL_RANGES=0 keeps the earlier sync-free
// solves: one wave per row of L (forward, rows in factor order) or of L^T
// (backward, rows in reverse order), the lanes splitting the row's entries.
// In both, a row's inputs from elsewhere are results that other work-groups
// publish with device-coherent stores into an array filled with a
// signalling-NaN sentinel before the solve; a lane polls a value until it is
// no longer the sentinel.  Every work-group waits only for rows of
// work-groups dispatched before it, so the solve always drains; each poll is
// bounded, and a wait that runs out sets a fault word (pinned host memory)
// that the caller reports.  Per row (sync-free solves): lane l sums its
// entries l, l + 64, ... in order from 0, then the lanes are added by a
// fixed tree; a lane issues the loads of up to kLdlBatch inputs before it
// waits for any.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "fvh_eqn.h"
#include "ldl_host.h"

namespace fvh {

namespace {

constexpr unsigned long long kLdlSent = 0x7FF4DEAD5EA11A57ull;
constexpr int kLdlSpin = 1 << 20;  // polls (a few 100 ms) before a wait is declared lost
constexpr int kLdlWaves = 4;       // waves (rows) per work-group
constexpr int kLdlBatch = 16;      // inputs per lane loaded before any is waited for

__device__ inline double ldl_load(const double *p) {
    return __builtin_bit_cast(double, __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p),
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline void ldl_store(double *p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), __builtin_bit_cast(unsigned long long, v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool ldl_sent(double v) { return __builtin_bit_cast(unsigned long long, v) == kLdlSent; }

// Value of row j once published (bounded wait).
__device__ inline double ldl_wait(const double *y, int j, int *fault) {
    double v = ldl_load(y + j);
    for (int s = 0; ldl_sent(v); ++s) {
        if (s >= kLdlSpin) {
            __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            return 0.0;
        }
        __builtin_amdgcn_s_sleep(1);
        v = ldl_load(y + j);
    }
    return v;
}

__device__ inline double wave_sum(double a) {
    for (int off = 32; off > 0; off >>= 1) a = a + __shfl_down(a, off, 64);
    return a;
}

// DPP move of an FP64 value (both halves); lanes that read outside their
// row, and rows the row mask leaves out, get 0
template <int CTRL, int ROWS, bool BOUND>
__device__ inline double dpp_f64(double v) {
    const int2 b = __builtin_bit_cast(int2, v);
    int2 r;
    r.x = __builtin_amdgcn_update_dpp(0, b.x, CTRL, ROWS, 0xF, BOUND);
    r.y = __builtin_amdgcn_update_dpp(0, b.y, CTRL, ROWS, 0xF, BOUND);
    return __builtin_bit_cast(double, r);
}
// Sum of a wave's 64 lanes in a fixed order, without LDS: shifts by 1, 2,
// 4, 8 within each row of 16 lanes (lane 15 of a row then holds the row's
// sum), the sum of rows 0 and 2 broadcast into rows 1 and 3, then lane 31's
// into rows 2 and 3; lane 63 holds (row 2 + row 3) + (row 0 + row 1).
__device__ inline double wave_total(double a) {
    a = a + dpp_f64<0x111, 0xF, true>(a);   // row_shr:1
    a = a + dpp_f64<0x112, 0xF, true>(a);   // row_shr:2
    a = a + dpp_f64<0x114, 0xF, true>(a);   // row_shr:4
    a = a + dpp_f64<0x118, 0xF, true>(a);   // row_shr:8
    a = a + dpp_f64<0x142, 0xA, false>(a);  // row_bcast:15 into rows 1 and 3
    a = a + dpp_f64<0x143, 0xC, false>(a);  // row_bcast:31 into rows 2 and 3
    const int2 b = __builtin_bit_cast(int2, a);
    return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_readlane(b.x, 63), __builtin_amdgcn_readlane(b.y, 63)));
}

}  // namespace

// rhs[k] = x[perm[k]] (in FP64); both result arrays to the sentinel
template <typename T>
__global__ void k_ldl_prep(int n, const int *__restrict__ perm, const T *__restrict__ x, double *__restrict__ rhs,
                           double *__restrict__ y, double *__restrict__ z, const KScal *ks) {
    if (ks && ks->done) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    rhs[k] = (double)x[perm[k]];
    y[k] = __builtin_bit_cast(double, kLdlSent);
    z[k] = __builtin_bit_cast(double, kLdlSent);
}

// BWD = false: y_r = rhs_r - sum_j L_rj y_j (rows of L in order);
// BWD = true:  z_r = y_r / D_r - sum_k L_kr z_k (rows of L^T, last first).
template <bool BWD>
__global__ void __launch_bounds__(64 * kLdlWaves)
    k_ldl_sweep(int n, const int *__restrict__ ptr, const int *__restrict__ idx, const double *__restrict__ val,
                const double *__restrict__ rhs, const double *__restrict__ D, const double *src, double *dst,
                int *fault, const KScal *ks) {
    if (ks && ks->done) return;
    const int w = blockIdx.x * kLdlWaves + (int)(threadIdx.x >> 6);
    if (w >= n) return;
    const int lane = threadIdx.x & 63;
    const int r = BWD ? n - 1 - w : w;
    const int p1 = ptr[r + 1];
    double acc = 0.0;
    // blocks of 64 * kLdlBatch entries: a lane loads all its inputs of a block
    // at once and then polls only those not yet published, so a row waits
    // about one load latency per block rather than one per entry
    for (int pb = ptr[r] + lane; pb < p1; pb += 64 * kLdlBatch) {
        double a[kLdlBatch], v[kLdlBatch];
        int c[kLdlBatch];
#pragma unroll
        for (int k = 0; k < kLdlBatch; ++k) {
            const int p = pb + 64 * k;
            c[k] = p < p1 ? idx[p] : 0;
            a[k] = p < p1 ? val[p] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kLdlBatch; ++k) v[k] = pb + 64 * k < p1 ? ldl_load(src + c[k]) : 0.0;
        for (int s = 0;; ++s) {
            bool missing = false;
#pragma unroll
            for (int k = 0; k < kLdlBatch; ++k) missing |= ldl_sent(v[k]);
            if (!missing) break;
            if (s >= kLdlSpin) {
                __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int k = 0; k < kLdlBatch; ++k)
                if (ldl_sent(v[k])) v[k] = ldl_load(src + c[k]);
        }
#pragma unroll
        for (int k = 0; k < kLdlBatch; ++k)
            if (pb + 64 * k < p1) acc = acc + a[k] * v[k];  // the lane's entries in order
    }
    const double s = wave_sum(acc);
    if (lane == 0) {
        const double base = BWD ? ldl_wait(rhs, r, fault) / D[r] : rhs[r];
        ldl_store(dst + r, base - s);
    }
}

// Range walk (the default): one work-group per range of the factor order
// (a nested-dissection leaf, or a separator, cut at kLdlRangeMax rows),
// ranges in solve order; the range's row pointers are staged in LDS and the
// next row's first batch of entries is loaded while a row is summed.  The work-group takes its rows one after another,
// its threads splitting each row's entries: an input from the range itself
// comes from the range's results in LDS, one from an earlier range from the
// global result array (device-coherent loads polled until published, with
// up to kLdlRangeBatch loads in flight per thread).  Per row: thread t sums
// its entries t, t + 256, ... in order from 0, the lanes of each wave are
// added by a fixed DPP tree (wave_total) and the four waves in order; every thread then
// forms the result and writes it to the row's LDS word (each wave reads only
// words it wrote itself), thread 0 also publishes it.  One barrier per row
// (the wave sums in two alternating slots).  A separator's dependent rows
// thus follow each other at LDS and barrier latency instead of a
// publish-to-poll hop between CUs.
constexpr int kLdlRangeMax = 2048;   // rows per range (its results and row pointers in LDS)
constexpr int kLdlRangeThreads = 256;
constexpr int kLdlRangeBatch = 8;

template <bool BWD>
__global__ void __launch_bounds__(kLdlRangeThreads)
    k_ldl_range(int nranges, const int *__restrict__ bounds, const int *__restrict__ ptr,
                const int *__restrict__ idx, const double *__restrict__ val, int nnz,
                const double *__restrict__ rhs, const double *__restrict__ D, double *out, int *fault,
                const KScal *ks) {
    if (ks && ks->done) return;
    constexpr int B = kLdlRangeBatch, SPAN = kLdlRangeThreads * kLdlRangeBatch;
    __shared__ double loc[kLdlRangeMax];
    __shared__ int rp[kLdlRangeMax + 1];  // ptr of the range's rows (and of the row after)
    __shared__ double red[2][kLdlRangeThreads / 64];
    const int g = BWD ? nranges - 1 - (int)blockIdx.x : (int)blockIdx.x;
    const int r0 = bounds[g], r1 = bounds[g + 1], nrow = r1 - r0;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int i = t; i <= nrow; i += kLdlRangeThreads) rp[i] = ptr[r0 + i];
    __syncthreads();
    auto row_of = [&](int q) { return BWD ? r1 - 1 - q : r0 + q; };
    // the first batch of row q's entries and its right-hand side, loaded
    // whatever the row's length (clamped indices, unused beyond the row), so
    // that the next row's loads are in flight while this row waits
    struct Head {
        int c[B];
        double a[B];
        int p0, p1;
        double base, piv;
    };
    auto fetch = [&](int q, Head &h) {
        const int r = row_of(q);
        h.p0 = rp[r - r0];
        h.p1 = rp[r - r0 + 1];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int p = min(h.p0 + t + kLdlRangeThreads * k, nnz - 1);
            h.c[k] = idx[p];
            h.a[k] = val[p];
        }
        h.base = rhs[r];
        h.piv = BWD ? D[r] : 1.0;
    };
    // inputs of entries pb + 256 k < p1 with columns c: from the range's LDS
    // words or, polled until published, from the global result array
    int lost = 0;
    auto gather = [&](const int (&c)[B], int pb, int p1, double (&v)[B]) {
        bool far[B];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const bool in = pb + kLdlRangeThreads * k < p1;
            far[k] = in && (c[k] < r0 || c[k] >= r1);
            v[k] = far[k] ? ldl_load(out + c[k]) : loc[in ? c[k] - r0 : 0];
        }
        for (int s = 0;; ++s) {
            bool missing = false;
#pragma unroll
            for (int k = 0; k < B; ++k) missing |= far[k] && ldl_sent(v[k]);
            if (!missing) break;
            if (s >= kLdlSpin) {
                lost = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int k = 0; k < B; ++k)
                if (far[k] && ldl_sent(v[k])) v[k] = ldl_load(out + c[k]);
        }
    };
    Head h;
    if (nrow > 0) fetch(0, h);
    for (int q = 0; q < nrow; ++q) {
        const int r = row_of(q);
        Head hn;
        fetch(min(q + 1, nrow - 1), hn);  // the next row's (the last row's again at the end)
        double acc = 0.0;
        {
            double v[B];
            gather(h.c, h.p0 + t, h.p1, v);
#pragma unroll
            for (int k = 0; k < B; ++k)
                if (h.p0 + t + kLdlRangeThreads * k < h.p1) acc = acc + h.a[k] * v[k];  // the thread's entries in order
        }
        for (int pb = h.p0 + t + SPAN; pb < h.p1; pb += SPAN) {  // rows longer than one batch
            int c[B];
            double a[B], v[B];
#pragma unroll
            for (int k = 0; k < B; ++k) {
                const int p = pb + kLdlRangeThreads * k;
                c[k] = p < h.p1 ? idx[p] : r0;
                a[k] = p < h.p1 ? val[p] : 0.0;
            }
            gather(c, pb, h.p1, v);
#pragma unroll
            for (int k = 0; k < B; ++k)
                if (pb + kLdlRangeThreads * k < h.p1) acc = acc + a[k] * v[k];
        }
        acc = wave_total(acc);
        double sum = acc;
        if constexpr (kLdlRangeThreads > 64) {  // the waves' sums in order
            if (lane == 0) red[q & 1][w] = acc;
            __syncthreads();
            sum = red[q & 1][0];
#pragma unroll
            for (int k = 1; k < kLdlRangeThreads / 64; ++k) sum = sum + red[q & 1][k];
        }
        const double y = (BWD ? h.base / h.piv : h.base) - sum;
        loc[r - r0] = y;
        if (t == 0) ldl_store(out + r, y);
        h = hn;
    }
    if (lost && t == 0) __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <typename T>
__global__ void k_ldl_out(int n, const int *__restrict__ perm, const double *__restrict__ z, T *__restrict__ out,
                          const KScal *ks) {
    if (ks && ks->done) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[perm[k]] = (T)z[k];
}

// Factor A (values of the equation the preconditioner reads) on the host,
// once, and upload L by rows (forward) and by columns (backward).
int ldl_setup(fvh_eqn *e, const double *val_dev) {
    fvh_mesh *m = e->m;
    LdlStore &S = e->ldl;
    if (S.ready) return FVH_OK;  // frozen after the set-up factorization
    if (m->cases) return set_error(FVH_ERR_UNSUPPORTED, "LDL preconditioner: not with independent cases");
    const int n = (int)m->n;
    std::vector<double> val(m->nnz), cx(n), cy(n);
    // the values may still be in assembly on the library's stream (a
    // non-blocking stream, which the synchronous copies below do not wait for)
    HIP_TRY(hipStreamSynchronize(stream()));
    HIP_TRY(hipMemcpy(val.data(), val_dev, sizeof(double) * m->nnz, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cx.data(), m->ccx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cy.data(), m->ccy.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    std::vector<int> perm, bounds;
    ldl_order(n, m->h_rowptr.data(), m->h_col.data(), cx.data(), cy.data(), perm, &bounds);
    {  // ranges of at most kLdlRangeMax rows
        std::vector<int> cut{0};
        for (size_t g = 0; g + 1 < bounds.size(); ++g)
            for (int a = bounds[g]; a < bounds[g + 1]; a += kLdlRangeMax) cut.push_back(std::min(a + kLdlRangeMax, bounds[g + 1]));
        if (cut.back() != n) return set_error(FVH_ERR_ARG, "LDL: ordering ranges end at %d of %d rows", cut.back(), n);
        FVH_TRY(S.ranges.upload(cut));
        S.nranges = (int)cut.size() - 1;
    }
    LdlHost F;
    FVH_TRY(ldl_factor(n, m->h_rowptr.data(), m->h_col.data(), val.data(), perm, F));
    // rows of L (strictly lower, columns ascending) from its columns
    const int nnzL = F.Lp[n];
    std::vector<int> rp(n + 1, 0), ri(nnzL);
    std::vector<double> rv(nnzL);
    for (int p = 0; p < nnzL; ++p) rp[F.Li[p] + 1]++;
    for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
    {
        std::vector<int> pos(rp.begin(), rp.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int p = F.Lp[j]; p < F.Lp[j + 1]; ++p) {
                const int q = pos[F.Li[p]]++;
                ri[q] = j;
                rv[q] = F.Lx[p];
            }
    }
    FVH_TRY(S.perm.upload(F.perm));
    FVH_TRY(S.fptr.upload(rp));
    FVH_TRY(S.fidx.upload(ri));
    FVH_TRY(S.fval.upload(rv));
    FVH_TRY(S.bptr.upload(F.Lp));
    FVH_TRY(S.bidx.upload(F.Li));
    FVH_TRY(S.bval.upload(F.Lx));
    FVH_TRY(S.D.upload(F.D));
    FVH_TRY(S.rhs.alloc(n));
    FVH_TRY(S.y.alloc(n));
    FVH_TRY(S.z.alloc(n));
    if (!S.h_fault) {
        HIP_TRY(hipHostMalloc((void **)&S.h_fault, sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
        *S.h_fault = 0;
    }
    S.nnzL = nnzL;
    S.ready = true;
    return FVH_OK;
}

// the range walk unless FVH_LDL_RANGES=0 (one wave per row, sync-free)
static bool ldl_ranges() {
    const char *env = std::getenv("FVH_LDL_RANGES");
    return !(env && env[0] == '0' && env[1] == 0);
}

template <typename T>
int ldl_apply(fvh_eqn *e, const T *x, T *out, const KScal *ks) {
    LdlStore &S = e->ldl;
    if (!S.ready) return set_error(FVH_ERR_ARG, "LDL preconditioner applied before its set-up factorization");
    const int n = (int)e->m->n;
    int *fault = nullptr;
    HIP_TRY(hipHostGetDevicePointer((void **)&fault, S.h_fault, 0));
    hipLaunchKernelGGL((k_ldl_prep<T>), grid1d(n), dim3(256), 0, stream(), n, S.perm.p, x, S.rhs.p, S.y.p, S.z.p, ks);
    HIP_TRY(hipGetLastError());
    if (ldl_ranges() && S.nranges > 0) {
        const dim3 gr((unsigned)S.nranges);
        const int nnz = std::max(S.nnzL, 1);  // the entry arrays hold at least one word
        hipLaunchKernelGGL((k_ldl_range<false>), gr, dim3(kLdlRangeThreads), 0, stream(), S.nranges, S.ranges.p,
                           S.fptr.p, S.fidx.p, S.fval.p, nnz, S.rhs.p, S.D.p, S.y.p, fault, ks);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL((k_ldl_range<true>), gr, dim3(kLdlRangeThreads), 0, stream(), S.nranges, S.ranges.p,
                           S.bptr.p, S.bidx.p, S.bval.p, nnz, S.y.p, S.D.p, S.z.p, fault, ks);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL((k_ldl_out<T>), grid1d(n), dim3(256), 0, stream(), n, S.perm.p, S.z.p, out, ks);
        HIP_TRY(hipGetLastError());
        return FVH_OK;
    }
    const dim3 g((unsigned)((n + kLdlWaves - 1) / kLdlWaves));
    hipLaunchKernelGGL((k_ldl_sweep<false>), g, dim3(64 * kLdlWaves), 0, stream(), n, S.fptr.p, S.fidx.p, S.fval.p,
                       S.rhs.p, S.D.p, S.y.p, S.y.p, fault, ks);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_ldl_sweep<true>), g, dim3(64 * kLdlWaves), 0, stream(), n, S.bptr.p, S.bidx.p, S.bval.p,
                       S.y.p, S.D.p, S.z.p, S.z.p, fault, ks);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_ldl_out<T>), grid1d(n), dim3(256), 0, stream(), n, S.perm.p, S.z.p, out, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
template int ldl_apply<double>(fvh_eqn *, const double *, double *, const KScal *);
template int ldl_apply<float>(fvh_eqn *, const float *, float *, const KScal *);

// A wait of the last solves ran out (reported once).
int ldl_check(fvh_eqn *e) {
    int *f = e->ldl.h_fault;
    if (f && __atomic_load_n(f, __ATOMIC_ACQUIRE)) {
        __atomic_store_n(f, 0, __ATOMIC_RELEASE);
        return set_error(FVH_ERR_HIP, "LDL solve: a row's input did not arrive in time");
    }
    return FVH_OK;
}

}  // namespace fvh
