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
// loaded ahead of use, and each neighbour's result is found where it is
// cheapest (below): after each barrier a level mostly waits for LDS only.
//
// A record holds at most kLevDeps neighbours; meshes with more per row keep
// the row-by-row kernels.  The FP32 sweeps (inner solves of mixed precision)
// use coefficients premultiplied by their pivots in FP64 and rounded once,
// with fused multiply-adds in neighbour order - the arithmetic of every FP32
// DILU sweep of the library.
#include <algorithm>
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
                           const int *__restrict__ lcode, const double *__restrict__ val,
                           const double *__restrict__ Dinv, int mode, LevRec<double> *__restrict__ R) {
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
        r.j[q] = lcode[p];  // where the walk finds neighbour j's result
        if (r.j[q] >= 0 && (r.j[q] & kLevLate)) r.pad[0] = 1;
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
    s.pad[0] = r.pad[0];
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
//
// One work-group walks the levels, thread t taking row t of each level (and
// every kLevBlock-th row after it in wider levels), with a barrier between
// levels.  Each result also goes to an LDS ring at its schedule position
// (mod kLevRing), which still holds most neighbours when their rows come
// (neighbours are mostly a few levels back).  A neighbour the ring has lost
// is two or more levels back, so final when the walk passes the level before
// its row's: its value is loaded then, one level ahead of use, with the row's
// input; the records are loaded two levels ahead.  Three register sets take
// turns (the loop is unrolled by three), so no load is waited for before its
// use.  A level's critical path holds LDS reads and the barrier only, except
// for neighbours one level back that a ring narrower than their level lost
// (read from global memory after the barrier).
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__device__ inline T lev_eval(const LevRec<T> &r, T in, const T (&nb)[kLevDeps]) {
    if (MODE == 0) {  // D_i = A_ii - sum (A_ij A_ji) / D_j
        T d = r.d[0];
#pragma unroll
        for (int q = 0; q < kLevDeps; ++q)
            if (q < r.m) d = d - r.a[q] / nb[q];
        return d;
    } else if (MODE == 1) {  // y_i = x_i - sum (A_ij y_j) Dinv_j
        T s = in;
        if constexpr (sizeof(T) == 4) {  // a = A_ij Dinv_j; fused, the highest column first
#pragma unroll
            for (int q = kLevDeps - 1; q >= 0; --q)
                if (q < r.m) s = __builtin_fmaf(-r.a[q], nb[q], s);
        } else {
#pragma unroll
            for (int q = 0; q < kLevDeps; ++q)
                if (q < r.m) s = s - r.a[q] * nb[q] * r.d[q];
        }
        return s;
    } else {  // z_i = y_i Dinv_i - sum (A_ij z_j) Dinv_i
        const T di = r.d[0];
        T s = in * di;
#pragma unroll
        for (int q = 0; q < kLevDeps; ++q)
            if (q < r.m) {
                if constexpr (sizeof(T) == 4) s = __builtin_fmaf(-r.a[q], nb[q], s);  // a = A_ij Dinv_i
                else s = s - r.a[q] * nb[q] * di;
            }
        return s;
    }
}

template <typename T>
struct LevSet {  // one level's first-pass row of this thread
    LevRec<T> r;
    T p[kLevDeps];  // neighbours the ring has lost (prefetched)
    T v;            // the row input
};

// The walk reads its input and writes its results in walk order (position
// k = row k of the walk), so both are contiguous per level; MODE 0 also
// writes Dinv in row order.  Loads are issued only for present rows and for
// neighbours the ring lost (the walk is bound by the one CU's load
// throughput); the ring is read with a select rather than a branch.  Rows
// with a neighbour to read after the barrier (flag in pad[0]) take a
// per-wave slow path that waits for all.
template <typename T, int MODE>
__global__ void __launch_bounds__(kLevBlock)
    k_lev_walk(int nlev, const int *__restrict__ lptr, const LevRec<T> *__restrict__ R, const T *__restrict__ in,
               T *out, T *dinv, const KScal *ks, int ringmask) {
    if (ks && ks->done) return;
    extern __shared__ unsigned char lev_lds[];
    T *ring = reinterpret_cast<T *>(lev_lds);  // [kLevRing] results by position
    const int t = threadIdx.x;
    auto load_rec = [&](LevSet<T> &S, int l) {
        S.r = LevRec<T>{};
        S.r.m = -1;  // absent row
        if (l < nlev && lptr[l] + t < lptr[l + 1]) S.r = R[lptr[l] + t];
    };
    auto prefetch = [&](LevSet<T> &S, int l) {
#pragma unroll
        for (int q = 0; q < kLevDeps; ++q) {
            const int c = S.r.j[q];
            S.p[q] = (q < S.r.m && c >= 0 && !(c & kLevLate)) ? out[c] : T(0);
        }
        S.v = (MODE != 0 && S.r.m >= 0) ? in[lptr[l] + t] : T(0);
    };
    auto finish = [&](const LevRec<T> &r, T v, int pos) {
        ring[pos & ringmask] = v;
        out[pos] = v;
        if (MODE == 0) dinv[r.i] = T(1) / v;
    };
    auto level = [&](int l, LevSet<T> &S) {
        const int beg = lptr[l], end = lptr[l + 1];
        T nb[kLevDeps];
#pragma unroll
        for (int q = 0; q < kLevDeps; ++q) {
            const int c = S.r.j[q];
            const T rv = ring[c < 0 ? -1 - c : 0];
            nb[q] = c < 0 ? rv : S.p[q];
        }
        if (__builtin_amdgcn_ballot_w64(beg + t < end && S.r.pad[0] != 0)) {  // neighbours one level back, beyond the ring
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#pragma unroll
            for (int q = 0; q < kLevDeps; ++q) {
                const int c = S.r.j[q];
                if (q < S.r.m && c >= 0 && (c & kLevLate)) nb[q] = out[c & ~kLevLate];
            }
        }
        if (beg + t < end) finish(S.r, lev_eval<T, MODE>(S.r, S.v, nb), beg + t);
        for (int k = beg + t + kLevBlock; k < end; k += kLevBlock) {  // levels wider than the block
            const LevRec<T> rk = R[k];
            T nk[kLevDeps];
#pragma unroll
            for (int q = 0; q < kLevDeps; ++q) {
                const int c = rk.j[q];
                nk[q] = q >= rk.m ? T(0) : c < 0 ? ring[-1 - c] : out[c & ~kLevLate];
            }
            finish(rk, lev_eval<T, MODE>(rk, MODE != 0 ? in[k] : T(0), nk), k);
        }
        __syncthreads();
    };
    LevSet<T> A, B, C;  // levels l, l+1, l+2 at the step of level l (roles rotate)
    load_rec(A, 0);
    load_rec(B, 1);
    prefetch(A, 0);  // level 0 has no neighbours: its inputs only
    for (int l = 0; l < nlev; l += 3) {
        load_rec(C, l + 2);
        prefetch(B, l + 1);
        level(l, A);
        if (l + 1 >= nlev) break;
        load_rec(A, l + 3);
        prefetch(C, l + 2);
        level(l + 1, B);
        if (l + 2 >= nlev) break;
        load_rec(B, l + 4);
        prefetch(A, l + 3);
        level(l + 2, C);
    }
}

// Vectors between row order and walk order (all CUs): to[k] = from[idx[k]]
// (gather) or to[idx[k]] = from[k] (scatter).
template <typename T, bool SCATTER>
__global__ void k_lev_perm(int n, const int *__restrict__ idx, const T *__restrict__ from, T *__restrict__ to,
                           const KScal *ks) {
    if (ks && ks->done) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (SCATTER) to[idx[k]] = from[k];
    else to[k] = from[idx[k]];
}

// ---------------------------------------------------------------------------
// Sync-free walk (the default): one lane per row at its walk position, all
// CUs at once.  A row's neighbours are rows of earlier levels, whose lanes
// publish their results with device-coherent stores into an array filled
// with a signalling-NaN sentinel before the sweep; a lane polls its
// neighbours' words until none is the sentinel, then evaluates the row (the
// same lev_eval as the level walk, so the results are bit-identical) and
// publishes it.  Each wave holds rows of one level only (a table of
// (first position, count) per wave, waves in level order), so no lane ever
// waits for a lane of its own wave, whose store the compiler may place
// after the wave's polling loop has ended for every lane; a wave waits only
// for waves of lower numbers, which the in-order dispatch of work-groups
// has started, so the sweep always drains.  The sleep between polls doubles
// while nothing arrives (up to about 2 us), which keeps waves far ahead of
// the front from flooding the memory system; each wait is bounded and one
// that runs out sets the sweep fault word.
// ---------------------------------------------------------------------------
template <typename T>
struct SfWord;
template <>
struct SfWord<double> {
    using U = unsigned long long;
    static constexpr U kSent = 0x7FF4DEAD5EA11A57ull;
};
template <>
struct SfWord<float> {
    using U = unsigned int;
    static constexpr U kSent = 0x7FA11A57u;
};
template <typename T>
__device__ inline T sf_load(const T *p) {
    using U = typename SfWord<T>::U;
    return __builtin_bit_cast(T, __hip_atomic_load(reinterpret_cast<const U *>(p), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT));
}
template <typename T>
__device__ inline void sf_store(T *p, T v) {
    using U = typename SfWord<T>::U;
    __hip_atomic_store(reinterpret_cast<U *>(p), __builtin_bit_cast(U, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline bool sf_sent(T v) {
    return __builtin_bit_cast(typename SfWord<T>::U, v) == SfWord<T>::kSent;
}
template <typename T>
__host__ __device__ inline T sf_sentinel() {
    return __builtin_bit_cast(T, SfWord<T>::kSent);
}
// Dynamic LDS per block of the sync-free walk, unless FVH_SF_LDS says
// otherwise: 90,000 bytes leave room for one 256-lane block per CU (4 waves),
// so that about 1,000 waves poll instead of up to 8,000 (measured on the
// 1M-point Delaunay mesh: FP64 apply 14.4 ms with no cap, 12.8 ms at two
// blocks per CU, 12.4 ms at one)
constexpr size_t kSfLds = 90000;
constexpr int kSfNap = 1;       // longest sleep between polls (s_sleep 2 units), unless FVH_SF_NAP says otherwise
constexpr int kSfSpin = 1 << 19;  // polls (with back-off: about a second) before a wait is declared lost

template <typename T, int MODE>
__global__ void __launch_bounds__(256)
    k_lev_sf(int nw, const int2 *__restrict__ waves, const LevRec<T> *__restrict__ R, const T *__restrict__ in,
             T *out, T *dinv, int *fault, int napmax, const KScal *ks) {
    if (ks && ks->done) return;
    const int w = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (w >= nw) return;
    const int2 seg = waves[w];
    const int lane = threadIdx.x & 63;
    if (lane >= seg.y) return;
    const int k = seg.x + lane;
    const LevRec<T> r = R[k];
    const T v0 = MODE != 0 ? in[k] : T(0);
    T nb[kLevDeps];
#pragma unroll
    for (int q = 0; q < kLevDeps; ++q) nb[q] = T(0);
    unsigned need = r.m > 0 ? (1u << r.m) - 1u : 0u;
    int nap = 1;
    for (int s = 0; need != 0u; ++s) {
        const unsigned before = need;
#pragma unroll
        for (int q = 0; q < kLevDeps; ++q)
            if (need & (1u << q)) {
                const T v = sf_load(out + r.j[q]);
                if (!sf_sent(v)) {
                    nb[q] = v;
                    need &= ~(1u << q);
                }
            }
        if (need == 0u) break;
        if (s >= kSfSpin) {
            __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            break;
        }
        // the front is here when a neighbour has just arrived: poll at once
        if (need != before) nap = 1;
        for (int z = 0; z < nap; ++z) __builtin_amdgcn_s_sleep(2);
        nap = nap < napmax ? 2 * nap : napmax;
    }
    const T res = lev_eval<T, MODE>(r, v0, nb);
    sf_store(out + k, res);
    if (MODE == 0) dinv[r.i] = T(1) / res;
}

// x (row order) into walk order, with the two result arrays of the sweeps
// set to the sentinel
template <typename T>
__global__ void k_lev_sf_prep(int n, const int *__restrict__ rows, const T *__restrict__ x, T *__restrict__ wa,
                              T *__restrict__ wb, T *__restrict__ wc, const KScal *ks) {
    if (ks && ks->done) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (x) wa[k] = x[rows[k]];
    wb[k] = sf_sentinel<T>();
    if (wc) wc[k] = sf_sentinel<T>();
}

// ---------------------------------------------------------------------------
// Tile walk (the default): one work-group per tile of consecutive rows (in
// the sweep's order), in tile order.  Its compute wave takes the tile's rows
// in steps of at most 64 independent rows, one lane per row, and keeps every
// result in LDS at the row's place in the tile, where the later steps of the
// tile read it; a step's critical path is an LDS read, the row's expression
// (tile_eval: the same bits as every other walk) and an LDS write.  Results
// also go to global memory with device-coherent stores, into an array filled
// with the sentinel before the sweep.  Neighbours from other (earlier) tiles
// are staged into LDS words after the tile's results by kTileStages loader
// waves, which take the chunks of steps in turn: a loader packs its chunk's
// records (with the rows' inputs) and step table into an LDS stage buffer
// (its own), polls the chunk's neighbours from other tiles until none is the
// sentinel, writes them to their LDS words and raises the buffer's ready
// word; the compute wave raises the buffer's done word when it has finished
// the chunk, and the loader then fills it with its next chunk.  The ready
// word carries the chunk's step count; the compute wave reads it with the
// step table and the first records during the previous chunk's last step.
// Only loaders wait for global memory, so the compute wave never waits for a
// load or a store.  A tile waits only for tiles before it, whose work-groups
// the in-order dispatch has started, so the sweep always drains; every wait
// is bounded, and one that runs out sets the fault word and ends the wave (a
// loader or compute wave that ends early makes the other waves of its tile
// run out in turn).
// ---------------------------------------------------------------------------
constexpr int kTileStages = 4;  // loader waves and stage buffers per work-group
constexpr int kTileSpin = 1 << 22;  // polls of an LDS word (about half a second) before a wait is declared lost
// ready word of a stage buffer: (chunk + 1) | (steps << kTileStepShift)
constexpr int kTileStepShift = 20, kTileChunkMask = (1 << kTileStepShift) - 1;

typedef unsigned int TileWord __attribute__((ext_vector_type(4)));  // 16 bytes of a record

// The compute wave's record of a row, built by the loaders from the level
// record: the neighbours' LDS words as byte offsets (absent: 0), the row's
// input v, and the coefficients and pivots the mode reads (update: a[q] =
// A_ij A_ji, d[0] = A_ii, the row i; forward: a[q] = A_ij and, in FP64, d[q] =
// Dinv_j; backward: a[q] = A_ij, d[0] = Dinv_i), packed so that a step reads
// as few 16-byte words as the mode needs.
template <int N>
struct TileWordArray {
    TileWord w[N];
};

template <typename T, int MODE, int D>
struct alignas(16) TileRec {
    static constexpr bool has_i = true, has_v = true;
    static constexpr int nd = D;
    int i, m, j[D];
    T v;
    T a[D], d[D];
};
template <>
struct alignas(16) TileRec<double, 0, 3> {  // 64 bytes
    static constexpr bool has_i = true, has_v = false;
    static constexpr int nd = 1;
    int i, m, j[3];
    double a[3], d[1];
};
template <>
struct alignas(16) TileRec<double, 2, 3> {  // 64 bytes
    static constexpr bool has_i = false, has_v = true;
    static constexpr int nd = 1;
    int m, j[3];
    double v, a[3], d[1];
};
template <>
struct alignas(16) TileRec<float, 1, 3> {  // 32 bytes (the FP32 coefficients carry their pivots)
    static constexpr bool has_i = false, has_v = true;
    static constexpr int nd = 0;
    int m, j[3];
    float v, a[3];
};
template <>
struct alignas(16) TileRec<float, 2, 3> {  // 48 bytes
    static constexpr bool has_i = false, has_v = true;
    static constexpr int nd = 1;
    int m, j[3];
    float v, a[3], d[1];
};

// lev_eval for rows of at most D neighbours, with the test for an absent
// neighbour on the term rather than on the running sum (off the chain of
// dependent operations): an absent term is +0, and x - (+0) = x and
// fma(-(+0), +0, x) = x + (-0) = x exactly for every x (signed zeros, NaN
// and infinities included), so the results are lev_eval's bits.
template <typename T, int MODE, int D, typename R>
__device__ inline T tile_eval(const R &r, T in, const T (&nb)[kLevDeps]) {
    if constexpr (MODE == 0) {
        T d = r.d[0];
#pragma unroll
        for (int q = 0; q < D; ++q) d = d - (q < r.m ? r.a[q] / nb[q] : T(0));
        return d;
    } else if constexpr (MODE == 1) {
        T s = in;
        if constexpr (sizeof(T) == 4) {
#pragma unroll
            for (int q = D - 1; q >= 0; --q)
                s = __builtin_fmaf(-(q < r.m ? r.a[q] : 0.0f), q < r.m ? nb[q] : 0.0f, s);
        } else {
#pragma unroll
            for (int q = 0; q < D; ++q) s = s - (q < r.m ? r.a[q] * nb[q] * r.d[q] : T(0));
        }
        return s;
    } else {
        const T di = r.d[0];
        T s = in * di;
#pragma unroll
        for (int q = 0; q < D; ++q) {
            if constexpr (sizeof(T) == 4) s = __builtin_fmaf(-(q < r.m ? r.a[q] : 0.0f), q < r.m ? nb[q] : 0.0f, s);
            else s = s - (q < r.m ? r.a[q] * nb[q] * di : T(0));
        }
        return s;
    }
}

static_assert(sizeof(TileRec<double, 1, 3>) == 80 && sizeof(TileRec<double, 2, 3>) == 64 &&
                  sizeof(TileRec<double, 0, 3>) == 64 && sizeof(TileRec<float, 1, 3>) == 32 &&
                  sizeof(TileRec<float, 2, 3>) == 48 && sizeof(TileRec<double, 1, 4>) == 96,
              "tile records: whole 16-byte words");

template <typename R>
struct alignas(16) TileStage {  // one chunk in LDS
    // records of the chunk's rows, by position from the chunk's first;
    // every slot holds a record (m = 0 past the chunk).  Word-major (16-byte
    // word w of slot k at rec[w][k]), so that a wave's read of one word of 64
    // consecutive slots is free of bank conflicts.
    TileWord rec[sizeof(R) / 16][64];
    int2 step[64];  // the chunk's steps: first position, rows
};
// slot k of a stage into a record in registers
template <typename R>
__device__ inline R stage_rec(const TileStage<R> &S, int k) {
    constexpr int RW = sizeof(R) / 16;
    TileWord w[RW];
#pragma unroll
    for (int q = 0; q < RW; ++q) w[q] = S.rec[q][k];
    return __builtin_bit_cast(R, w);
}

template <typename R>
__host__ __device__ constexpr size_t tile_words_off() {  // byte offset of the tile's result words
    return kTileStages * sizeof(TileStage<R>) + 16 * ((2 * kTileStages * sizeof(int) + 15) / 16);
}

template <typename T, int MODE, int D>
__global__ void __launch_bounds__(64 * (1 + kTileStages))
    k_lev_tile(int tr, const int *__restrict__ tiles, const int *__restrict__ chunks, const int2 *__restrict__ steps,
               const int2 *__restrict__ ents, const LevRec<T> *__restrict__ R, const T *__restrict__ in, T *out,
               T *dinv, int *fault, const KScal *ks) {
    if (ks && ks->done) return;
    using Rec = TileRec<T, MODE, D>;
    extern __shared__ __align__(16) unsigned char tile_mem[];
    auto *stage = reinterpret_cast<TileStage<Rec> *>(tile_mem);
    int *ready = reinterpret_cast<int *>(stage + kTileStages), *done = ready + kTileStages;
    T *res = reinterpret_cast<T *>(tile_mem + tile_words_off<Rec>());  // [tr + staged neighbours]
    const int t = blockIdx.x, wave = (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int c0 = tiles[2 * t], nc = tiles[2 * t + 1];
    if (threadIdx.x < kTileStages) ready[threadIdx.x] = done[threadIdx.x] = 0;  // chunk c is c + 1
    __syncthreads();
    int broken = 0;
    if (wave > 0) {  // loader of chunks wave-1, wave-1 + kTileStages, ...
        const int b = wave - 1;
        TileStage<Rec> &S = stage[b];
        int4 ca = make_int4(0, 0, 0, 0);  // lane k: the loader's k-th chunk record (of the current 64)
        int2 cb = make_int2(0, 0);
        for (int c = b, k = 0; c < nc; c += kTileStages, k = (k + 1) & 63) {
            if (k == 0) {  // the records of this loader's next 64 chunks, one per lane
                const int cl = c + lane * kTileStages;
                if (cl < nc) {
                    const int *C = chunks + (size_t)kTileChunkInts * (c0 + cl);
                    ca = *reinterpret_cast<const int4 *>(C);
                    cb = *reinterpret_cast<const int2 *>(C + 4);
                }
            }
            const int fs = __builtin_amdgcn_readlane(ca.x, k), ns = __builtin_amdgcn_readlane(ca.y, k),
                      fe = __builtin_amdgcn_readlane(ca.z, k), ne = __builtin_amdgcn_readlane(ca.w, k),
                      fp = __builtin_amdgcn_readlane(cb.x, k), nr = __builtin_amdgcn_readlane(cb.y, k);
            constexpr int LW = sizeof(LevRec<T>) / 16, RW = sizeof(Rec) / 16;  // as 16-byte words (no struct copy in memory)
            int2 st = make_int2(0, 0), e = make_int2(0, 0);
            TileWord lw[LW];
            T x = T(0), v = T(0);
#pragma unroll
            for (int w = 0; w < LW; ++w) lw[w] = TileWord{0u, 0u, 0u, 0u};  // m = 0, words 0
            if (lane < ns) st = steps[fs + lane];
            if (lane < nr) {
                const TileWord *src = reinterpret_cast<const TileWord *>(R + fp + lane);
#pragma unroll
                for (int w = 0; w < LW; ++w) lw[w] = src[w];
                if (MODE != 0) v = in[fp + lane];
            }
            // the compute wave's record (neighbour words as byte offsets into
            // the LDS words, absent 0, so the compute wave reads them with no
            // address arithmetic)
            const LevRec<T> lr = __builtin_bit_cast(LevRec<T>, lw);
            Rec rc{};
            rc.m = lr.m;
#pragma unroll
            for (int q = 0; q < D; ++q) {
                rc.j[q] = q < lr.m ? lr.j[q] * (int)sizeof(T) : 0;
                rc.a[q] = lr.a[q];
            }
            if constexpr (Rec::nd > 0) {
#pragma unroll
                for (int q = 0; q < Rec::nd; ++q) rc.d[q] = lr.d[q];
            }
            if constexpr (Rec::has_i) rc.i = lr.i;
            if constexpr (Rec::has_v) rc.v = v;
            TileWord r[RW];
            {
                const auto rb = __builtin_bit_cast(TileWordArray<RW>, rc);
#pragma unroll
                for (int w = 0; w < RW; ++w) r[w] = rb.w[w];
            }
            if (lane < ne) {
                e = ents[fe + lane];
                x = sf_load(out + e.x);
            }
            for (int s = 0; __builtin_amdgcn_ballot_w64(lane < ne && sf_sent(x)); ++s) {
                if (s >= kSfSpin) {
                    broken = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                if (lane < ne && sf_sent(x)) x = sf_load(out + e.x);
            }
            for (int s = 0; !broken && c >= kTileStages &&
                            __hip_atomic_load(done + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) !=
                                c - kTileStages + 1;
                 ++s) {  // the compute wave is still on this buffer's previous chunk
                if (s >= kTileSpin) broken = 1;
                __builtin_amdgcn_s_sleep(1);
            }
            if (broken) break;
            asm volatile("" ::: "memory");
#pragma unroll
            for (int w = 0; w < RW; ++w) S.rec[w][lane] = r[w];
            if (lane < ns) S.step[lane] = st;
            if (lane < ne) res[e.y] = x;
            asm volatile("" ::: "memory");  // the stage is written (LDS keeps a wave's order) before the ready word
            __hip_atomic_store(ready + b, (c + 1) | (ns << kTileStepShift), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    } else {  // compute wave
        const int base = t * tr;
        constexpr int RW = sizeof(Rec) / 16;
        // a chunk's ready word, step table and first step's records (its rows
        // are slots 0 .. rows-1), read in that order: LDS serves the wave's
        // reads in order, so once the ready word shows the chunk, the reads
        // after it see what the loader wrote before it.  They are issued
        // during the previous chunk's last step; a chunk whose loader was late
        // is read again until its ready word shows it.
        int word = 0;
        int2 sv;
        TileWord rw[RW];
        auto take = [&](int c) {
            const TileStage<Rec> &S = stage[c % kTileStages];
            word = __hip_atomic_load(ready + c % kTileStages, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            asm volatile("" ::: "memory");
            sv = S.step[lane];
#pragma unroll
            for (int w = 0; w < RW; ++w) rw[w] = S.rec[w][lane];
        };
        take(0);
        for (int c = 0; c < nc; ++c) {
            const int b = c % kTileStages;
            const TileStage<Rec> &S = stage[b];
            word = __builtin_amdgcn_readfirstlane(word);
            for (int s = 0; (word & kTileChunkMask) != c + 1; ++s) {
                if (s >= kTileSpin) {
                    broken = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                take(c);
                // the values are taken here, so the reads go out together
                asm volatile("" : "+v"(word), "+v"(sv.x), "+v"(sv.y));
#pragma unroll
                for (int w = 0; w < RW; ++w) asm volatile("" : "+v"(rw[w]));
                word = __builtin_amdgcn_readfirstlane(word);
            }
            if (broken) break;  // nothing of this chunk is used
            Rec r = __builtin_bit_cast(Rec, rw);
            const int ns = word >> kTileStepShift, fp = __builtin_amdgcn_readlane(sv.x, 0);
            const int2 svc = sv;
            int sx = fp, sy = __builtin_amdgcn_readlane(sv.y, 0);
            // one step's rows: neighbours from LDS, the row's expression, the
            // result to LDS and global memory; `ahead` issues the reads for the
            // next step (lanes past a step's rows read some record of the chunk:
            // every slot holds one, its words in range; they store nothing)
            auto step = [&](const Rec &cur, auto ahead) {
                T nb[kLevDeps];
#pragma unroll
                for (int q = 0; q < D; ++q)  // byte offsets (absent: word 0)
                    nb[q] = *reinterpret_cast<const T *>(reinterpret_cast<const char *>(res) + cur.j[q]);
                asm volatile("" ::: "memory");  // the neighbour reads go out first (they are on the chain)
                ahead();
                if (lane < sy) {
                    T in_ = T(0);
                    if constexpr (Rec::has_v) in_ = cur.v;
                    const T val = tile_eval<T, MODE, D>(cur, in_, nb);
                    res[sx - base + lane] = val;
                    sf_store(out + sx + lane, val);
                    if constexpr (MODE == 0) dinv[cur.i] = T(1) / val;
                }
            };
            auto within = [&](int s, const Rec &cur, Rec &nxt) {  // step s, not the last
                const int nx = __builtin_amdgcn_readlane(svc.x, s + 1), ny = __builtin_amdgcn_readlane(svc.y, s + 1);
                step(cur, [&]() { nxt = stage_rec(S, (nx - fp + lane) & 63); });
                sx = nx;
                sy = ny;
            };
            auto last = [&](const Rec &cur) {  // the chunk's last step reads the next chunk's stage
                step(cur, [&]() {
                    if (c + 1 < nc) take(c + 1);
                });
            };
            Rec r2;
            int s = 0;
            for (; s + 2 < ns; s += 2) {  // two register sets take turns
                within(s, r, r2);
                within(s + 1, r2, r);
            }
            if (s + 1 < ns) {
                within(s, r, r2);
                last(r2);
            } else {
                last(r);
            }
            asm volatile("" ::: "memory");  // LDS serves the stage reads before this write
            __hip_atomic_store(done + b, c + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    if (broken && lane == 0) __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static dim3 lev_grid(int n) { return dim3((unsigned)((n + 255) / 256)); }
static dim3 sf_grid(int nwaves) { return dim3((unsigned)((nwaves + 3) / 4)); }  // four waves per work-group
// Longest sleep between two polls of the sync-free walk, in units of about
// 50 ns (FVH_SF_NAP, a power of two up to 64)
// Dynamic LDS per block of the sync-free walk (FVH_SF_LDS bytes, at most
// 160 KiB; the kernel uses none): a way to cap how many waves are resident
// and polling ahead of the front
static size_t sf_lds() {
    if (const char *env = std::getenv("FVH_SF_LDS")) {
        const long r = std::atol(env);
        if (r >= 0 && r <= 160 * 1024) return (size_t)r;
    }
    return kSfLds;
}
template <typename T, int MODE>
static int sf_allow(size_t lds) {
    static size_t allowed = 64 * 1024;
    if (lds > allowed) {
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(k_lev_sf<T, MODE>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        allowed = lds;
    }
    return FVH_OK;
}
static int sf_napmax() {
    int v = kSfNap;
    if (const char *env = std::getenv("FVH_SF_NAP")) {
        const int r = std::atoi(env);
        if (r >= 1 && r <= 64 && (r & (r - 1)) == 0) v = r;
    }
    return v;
}

// The walk's LDS ring of results (FP64: 128 KiB) needs the larger dynamic
// LDS allowance, set once per instantiation.
template <typename T, int MODE>
static int launch_walk(int nlev, const int *lptr, const LevRec<T> *R, const T *in, T *out, T *dinv, const KScal *ks,
                       int ring) {
    constexpr size_t lds = kLevRing * sizeof(T);
    static bool allowed = false;
    if (!allowed) {
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(k_lev_walk<T, MODE>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        allowed = true;
    }
    hipLaunchKernelGGL((k_lev_walk<T, MODE>), dim3(1), dim3(kLevBlock), lds, stream(), nlev, lptr, R, in, out, dinv, ks,
                       ring - 1);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// LDS bytes per work-group of the tile walk
template <typename T, int MODE, int D>
static size_t tile_lds(const Schedule &s, const Schedule::TileWalk &w) {
    return tile_words_off<TileRec<T, MODE, D>>() + (size_t)(s.tile_rows + w.maxrem) * sizeof(T);
}
constexpr size_t kLdsMax = 160 * 1024;

// The walk: the tile walk (2) unless FVH_LEVEL_TILES=0 or its tiles need more
// LDS than a CU has (the sync-free walk, 1), or FVH_LEVEL_SYNCFREE=0 (the
// one-work-group level walk, 0); all give the same bits.
static bool env_off(const char *name) {
    const char *env = std::getenv(name);
    return env && env[0] == '0' && env[1] == 0;
}
static int lev_walk(const Schedule &s) {
    if (env_off("FVH_LEVEL_SYNCFREE")) return 0;
    if (env_off("FVH_LEVEL_TILES")) return 1;
    const size_t lds = std::max(tile_lds<double, 1, kLevDeps>(s, s.tf), tile_lds<double, 2, kLevDeps>(s, s.tb));
    return lds <= kLdsMax ? 2 : 1;
}

template <typename T, int MODE, int D>
static int launch_tiles_d(const Schedule &s, const Schedule::TileWalk &w, const LevRec<T> *R, const T *in, T *out,
                          T *dinv, int *fault, const KScal *ks) {
    const size_t lds = tile_lds<T, MODE, D>(s, w);
    if (lds > kLdsMax) return set_error(FVH_ERR_UNSUPPORTED, "tile walk: %zu bytes of LDS per tile", lds);
    static size_t allowed = 64 * 1024;
    if (lds > allowed) {
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(k_lev_tile<T, MODE, D>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        allowed = lds;
    }
    hipLaunchKernelGGL((k_lev_tile<T, MODE, D>), dim3((unsigned)w.ntiles), dim3(64 * (1 + kTileStages)), lds,
                       stream(), s.tile_rows, w.tiles.p, w.chunks.p, reinterpret_cast<const int2 *>(w.steps.p),
                       reinterpret_cast<const int2 *>(w.ents.p), R, in, out, dinv, fault, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}
// rows of at most three neighbours (triangles) or kLevDeps
template <typename T, int MODE>
static int launch_tiles(const Schedule &s, const Schedule::TileWalk &w, const LevRec<T> *R, const T *in, T *out,
                        T *dinv, int *fault, const KScal *ks) {
    if (s.max_deps <= 3) return launch_tiles_d<T, MODE, 3>(s, w, R, in, out, dinv, fault, ks);
    return launch_tiles_d<T, MODE, kLevDeps>(s, w, R, in, out, dinv, fault, ks);
}

static int lev_fault_ptr(fvh_eqn *e, int **dev) {
    LevelStore &L = e->lv;
    if (!L.h_fault) {
        HIP_TRY(hipHostMalloc((void **)&L.h_fault, sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
        *L.h_fault = 0;
    }
    HIP_TRY(hipHostGetDevicePointer((void **)dev, L.h_fault, 0));
    return FVH_OK;
}

int lev_check(fvh_eqn *e) {
    int *f = e->lv.h_fault;
    if (f && __atomic_load_n(f, __ATOMIC_ACQUIRE)) {
        __atomic_store_n(f, 0, __ATOMIC_RELEASE);
        return set_error(FVH_ERR_HIP, "level sweep: a row's neighbour did not arrive in time");
    }
    return FVH_OK;
}

bool lev_path(const fvh_mesh *m) {
    if (m->sched.nlev <= 0 || m->sched.max_deps > kLevDeps || m->n >= kLevLate) return false;
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
        FVH_TRY(L.a.alloc(n));
        FVH_TRY(L.b.alloc(n));
    }
    auto *U = reinterpret_cast<LevRec<double> *>(L.upd.p);
    auto *Fw = reinterpret_cast<LevRec<double> *>(L.fwd.p);
    auto *Bw = reinterpret_cast<LevRec<double> *>(L.bwd.p);
    const int walk = lev_walk(s);
    if (walk > 0 && L.c.n != (size_t)n) FVH_TRY(L.c.alloc(n));
    // where each walk finds a neighbour: the tile walk's LDS words, the
    // sync-free walk's positions, or the level walk's ring codes
    const int *codes = walk == 2 ? s.tcode.p : walk == 1 ? s.lpos.p : s.lcode.p;
    const int *frows = walk == 2 ? s.tf.rows.p : s.frows.p, *brows = walk == 2 ? s.tb.rows.p : s.brows.p;
    hipLaunchKernelGGL(k_lev_pack, lev_grid(n), dim3(256), 0, stream(), n, frows, m->d.rowptr, m->d.col, m->d.diag,
                       m->d.tslot, codes, val, (const double *)nullptr, 0, U);
    HIP_TRY(hipGetLastError());
    if (walk == 2) {
        int *fault = nullptr;
        FVH_TRY(lev_fault_ptr(e, &fault));
        hipLaunchKernelGGL((k_lev_sf_prep<double>), lev_grid(n), dim3(256), 0, stream(), n, frows,
                           (const double *)nullptr, (double *)nullptr, L.b.p, (double *)nullptr,
                           (const KScal *)nullptr);
        HIP_TRY(hipGetLastError());
        FVH_TRY((launch_tiles<double, 0>(s, s.tf, U, (const double *)nullptr, L.b.p, e->Dinv.p, fault,
                                         (const KScal *)nullptr)));
    } else if (walk == 1) {
        int *fault = nullptr;
        FVH_TRY(lev_fault_ptr(e, &fault));
        hipLaunchKernelGGL((k_lev_sf_prep<double>), lev_grid(n), dim3(256), 0, stream(), n, s.frows.p,
                           (const double *)nullptr, (double *)nullptr, L.b.p, (double *)nullptr,
                           (const KScal *)nullptr);
        HIP_TRY(hipGetLastError());
        FVH_TRY((sf_allow<double, 0>(sf_lds())));
        hipLaunchKernelGGL((k_lev_sf<double, 0>), sf_grid(s.nswf), dim3(256), sf_lds(), stream(), s.nswf,
                           reinterpret_cast<const int2 *>(s.swf.p), U, (const double *)nullptr, L.b.p, e->Dinv.p,
                           fault, sf_napmax(), (const KScal *)nullptr);
        HIP_TRY(hipGetLastError());
    } else {
        FVH_TRY((launch_walk<double, 0>(s.nlev, s.fptr.p, U, (const double *)nullptr, L.b.p, e->Dinv.p,
                                        (const KScal *)nullptr, s.ring)));
    }
    hipLaunchKernelGGL(k_lev_pack, lev_grid(n), dim3(256), 0, stream(), n, frows, m->d.rowptr, m->d.col, m->d.diag,
                       m->d.tslot, codes, val, e->Dinv.p, 1, Fw);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_lev_pack, lev_grid(n), dim3(256), 0, stream(), n, brows, m->d.rowptr, m->d.col, m->d.diag,
                       m->d.tslot, codes, val, e->Dinv.p, 2, Bw);
    HIP_TRY(hipGetLastError());
    L.walk = walk;
    return FVH_OK;
}

// x (row order) -> forward walk order -> forward walk -> backward walk order
// -> backward walk -> y (row order); wa / wb: two vectors in walk order.
template <typename T>
static int lev_apply_t(fvh_eqn *e, const LevRec<T> *Fw, const LevRec<T> *Bw, const T *x, T *y, T *wa, T *wb,
                       T *wc, const KScal *ks) {
    const Schedule &s = e->m->sched;
    const int n = (int)e->m->n;
    if (e->lv.walk == 2) {  // records hold LDS words: the tile walk
        int *fault = nullptr;
        FVH_TRY(lev_fault_ptr(e, &fault));
        hipLaunchKernelGGL((k_lev_sf_prep<T>), lev_grid(n), dim3(256), 0, stream(), n, s.tf.rows.p, x, wa, wb, wc, ks);
        HIP_TRY(hipGetLastError());
        FVH_TRY((launch_tiles<T, 1>(s, s.tf, Fw, wa, wb, (T *)nullptr, fault, ks)));
        hipLaunchKernelGGL((k_lev_perm<T, false>), lev_grid(n), dim3(256), 0, stream(), n, s.tf2b.p, wb, wa, ks);
        HIP_TRY(hipGetLastError());
        FVH_TRY((launch_tiles<T, 2>(s, s.tb, Bw, wa, wc, (T *)nullptr, fault, ks)));
        hipLaunchKernelGGL((k_lev_perm<T, true>), lev_grid(n), dim3(256), 0, stream(), n, s.tb.rows.p, wc, y, ks);
        HIP_TRY(hipGetLastError());
        return FVH_OK;
    }
    if (e->lv.walk == 1) {  // records hold walk positions: the sync-free walk
        int *fault = nullptr;
        FVH_TRY(lev_fault_ptr(e, &fault));
        hipLaunchKernelGGL((k_lev_sf_prep<T>), lev_grid(n), dim3(256), 0, stream(), n, s.frows.p, x, wa, wb, wc, ks);
        HIP_TRY(hipGetLastError());
        FVH_TRY((sf_allow<T, 1>(sf_lds())));
        FVH_TRY((sf_allow<T, 2>(sf_lds())));
        hipLaunchKernelGGL((k_lev_sf<T, 1>), sf_grid(s.nswf), dim3(256), sf_lds(), stream(), s.nswf,
                           reinterpret_cast<const int2 *>(s.swf.p), Fw, wa, wb, (T *)nullptr, fault, sf_napmax(), ks);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL((k_lev_perm<T, false>), lev_grid(n), dim3(256), 0, stream(), n, s.f2b.p, wb, wa, ks);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL((k_lev_sf<T, 2>), sf_grid(s.nswb), dim3(256), sf_lds(), stream(), s.nswb,
                           reinterpret_cast<const int2 *>(s.swb.p), Bw, wa, wc, (T *)nullptr, fault, sf_napmax(), ks);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL((k_lev_perm<T, true>), lev_grid(n), dim3(256), 0, stream(), n, s.brows.p, wc, y, ks);
        HIP_TRY(hipGetLastError());
        return FVH_OK;
    }
    hipLaunchKernelGGL((k_lev_perm<T, false>), lev_grid(n), dim3(256), 0, stream(), n, s.frows.p, x, wa, ks);
    HIP_TRY(hipGetLastError());
    FVH_TRY((launch_walk<T, 1>(s.nlev, s.fptr.p, Fw, wa, wb, (T *)nullptr, ks, s.ring)));
    hipLaunchKernelGGL((k_lev_perm<T, false>), lev_grid(n), dim3(256), 0, stream(), n, s.f2b.p, wb, wa, ks);
    HIP_TRY(hipGetLastError());
    FVH_TRY((launch_walk<T, 2>(s.nlev, s.bptr.p, Bw, wa, wb, (T *)nullptr, ks, s.ring)));
    hipLaunchKernelGGL((k_lev_perm<T, true>), lev_grid(n), dim3(256), 0, stream(), n, s.brows.p, wb, y, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int lev_dilu_apply(fvh_eqn *e, const double *x, double *y, const KScal *ks) {
    LevelStore &L = e->lv;
    if (!L.fwd.p) return set_error(FVH_ERR_ARG, "DILU records not prepared");
    return lev_apply_t<double>(e, reinterpret_cast<const LevRec<double> *>(L.fwd.p),
                               reinterpret_cast<const LevRec<double> *>(L.bwd.p), x, y, L.a.p, L.b.p, L.c.p, ks);
}

int lev_to32(fvh_eqn *e) {
    LevelStore &L = e->lv;
    const int n = (int)e->m->n;
    if (L.fwd32.n != (size_t)n) {
        FVH_TRY(L.fwd32.alloc(n));
        FVH_TRY(L.bwd32.alloc(n));
        FVH_TRY(L.a32.alloc(n));
        FVH_TRY(L.b32.alloc(n));
        FVH_TRY(L.c32.alloc(n));
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
    LevelStore &L = e->lv;
    if (!L.fwd32.p) return set_error(FVH_ERR_ARG, "FP32 DILU records not prepared");
    return lev_apply_t<float>(e, reinterpret_cast<const LevRec<float> *>(L.fwd32.p),
                              reinterpret_cast<const LevRec<float> *>(L.bwd32.p), x, y, L.a32.p, L.b32.p, L.c32.p,
                              ks);
}

static_assert(sizeof(LevRec<double>) == kLevRecBytes64 && sizeof(LevRec<float>) == kLevRecBytes32,
              "record sizes");

}  // namespace fvh
