// Pipelined DILU recurrences on banded grids (cells numbered i-fastest on an
// nx x ny rectangle, lower neighbours r-1 and r-nx only).
//
// Lane l of compute wave w of block rb owns grid row j = 64 (W rb + w) + l
// and, at step T, cell (T - j, j) of anti-diagonal T.  Its west value is its
// own result of step T-1 (a register); its south value is lane l-1's result
// of step T-1 (DPP wave_shr:1).  Lane 0's south value is the top lane of the
// wave below: every wave keeps its results of a group of kSkewL steps, and
// its top lane writes them as one batch to an LDS hand-over array; the wave
// above reads each batch into registers a group ahead of use, and step k of
// the group passes word k as the `old` operand of its DPP shift (the value
// lane 0 keeps).  Between blocks (on other CUs), two courier waves per block relay the
// batches: one copies the top wave's LDS batches to a device-coherent global
// array, the other the block below's batches from theirs into its own LDS.
// Only the couriers issue device-coherent loads and stores, so their latency
// never sits in the compute waves' in-order memory counters.  No barrier couples
// the waves: each waits only for the batch it needs, the waves run a few
// batches apart, and every compute wave issues at its own rate (one per
// SIMD).  The backward sweep runs the same pipeline from the top row block
// down (its blocks are numbered top first, so every block waits only on
// blocks dispatched before it).
//
// Hand-over arrays hold, at index s+1, the value of step s (index 0: step
// -1, a padding value); unwritten words hold a signalling-NaN sentinel, which
// no arithmetic produces, so a sentinel means "not yet written".
//
// Layout: each anti-diagonal owns S slots, one per grid row (rows rounded up
// to whole blocks, at least one padding row), so one step's records of a wave
// are contiguous and one step is one uniform buffer offset.  Slots outside the
// rectangle are padding whose values make every term they enter an exact +0:
//   F = (b, Dinv, A_nr, A_rw) = (0, 1, 0, 0),  Bk = (0, Dinv, A_re, A_rn) =
//   (0, 1, 0, 0),  U = (A_rr, A_rs A_sr, A_rw A_wr) = (1, 0, 0), and the
//   sweeps' quad records and right-hand sides built from them;
// a padding cell's pivot evaluates to 1 and its forward / backward values to
// +0.  A missing neighbour of a real cell has the coefficient +0 in F and U
// and copysign(0, Dinv_r) in Bk (set by the diagonal recurrence), so its
// product is +0 and x - (+0) == x bit for bit, also for x = -0.  Each row
// therefore evaluates exactly the reference's expressions over its existing
// neighbours, in the reference's order (preconditioners_1_DILU.jl:56-162):
//   D_r   = (A_rr - (A_rs*A_sr)/D_s) - (A_rw*A_wr)/D_w,    Dinv_r = 1/D_r
//   y_r   = (b_r - (A_rs*y_s)*Dinv_s) - (A_rw*y_w)*Dinv_w
//   z_r   = (y_r*Dinv_r - (A_re*z_e)*Dinv_r) - (A_rn*z_n)*Dinv_r
// and the results are bit-identical to the level-scheduled kernels for finite
// data.  (A zero pivot or an infinite value can carry a NaN through a padding
// slot into an edge cell where the reference has another non-finite value.)
//
// Every wait is bounded: one that runs out records a fault in pinned host
// memory (the solve reports it), and the wave stops waiting and runs to its
// end, so every launch drains.
#include <cstdlib>
#include <cstring>

#include "fvh_eqn.h"
#include "k_band.h"
#include "k_dpp.h"

namespace fvh {

constexpr int kL = kSkewL;
// The record rings hold M hand-over batches of steps (prefetch distance
// M*kL): as deep as the vector-memory counter allows for what the sweeps
// issue (FP64 forward: two 16-byte loads per step and two stores per four
// steps; FP64 backward: ten operations per four steps; FP32: four or five
// per four steps; at most 63 in flight).
template <typename T>
constexpr int ring_batches() {
    return sizeof(T) == 4 ? 4 : 2;
}
static_assert(4 * kL <= kSkewPD, "padding diagonals cover the deepest ring");
constexpr int kSpinLimit = 1 << 19;  // idle polls (tens of ms) before a hand-over is declared lost
constexpr int kWin = 32;             // hand-over words a courier moves per poll (4 batches)

template <typename T>
struct HandWord;
template <>
struct HandWord<double> {
    using U = unsigned long long;
    static constexpr U kSent = 0x7FF4DEAD5EA11A57ull;
};
template <>
struct HandWord<float> {
    using U = unsigned int;
    static constexpr U kSent = 0x7FA11A57u;
};
template <typename T>
__device__ __host__ inline T hand_sentinel() {
    return __builtin_bit_cast(T, HandWord<T>::kSent);
}
template <typename T>
__device__ inline bool is_sentinel(T v) {
    return __builtin_bit_cast(typename HandWord<T>::U, v) == HandWord<T>::kSent;
}
// device-coherent global words (courier only)
template <typename T>
__device__ inline T gload(const T *p) {
    using U = typename HandWord<T>::U;
    return __builtin_bit_cast(
        T, __hip_atomic_load(reinterpret_cast<U *>(const_cast<T *>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
template <typename T>
__device__ inline void gstore(T *p, T v) {
    using U = typename HandWord<T>::U;
    __hip_atomic_store(reinterpret_cast<U *>(p), __builtin_bit_cast(U, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// LDS words shared by the waves of a block
template <typename T>
__device__ inline T lload(const T *p) {
    using U = typename HandWord<T>::U;
    return __builtin_bit_cast(T, __hip_atomic_load(reinterpret_cast<U *>(const_cast<T *>(p)), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP));
}
template <typename T>
__device__ inline void lstore(T *p, T v) {
    using U = typename HandWord<T>::U;
    __hip_atomic_store(reinterpret_cast<U *>(p), __builtin_bit_cast(U, v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ inline void record_fault(int *fault) {
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Re-reads an LDS batch until no lane holds a sentinel; bounded: returns
// with sentinels still in it when the batch never comes (or at once when the
// wave is already `broken`).  Out of line: the sweeps branch here only when a
// batch was not yet written.
template <typename T>
__device__ __attribute__((noinline)) T wait_batch(const T *p, T v, int broken) {
    for (int s = 0; !broken && s < kSpinLimit; ++s) {
        v = lload(p);
        if (!__builtin_amdgcn_ballot_w64(is_sentinel(v))) break;
        __builtin_amdgcn_s_sleep(1);
    }
    return v;
}

struct SkewArgs {
    const SkewGrid *grids;
    const int *blk_grid;
    int nblk, HS;
    int ks_per_grid;  // KScal k belongs to grid k (batched cases)
    unsigned rbytes;  // bytes of one record array (buffer range)
    int *fault;
};

// Buffer access: a per-lane byte offset (fixed for the sweep) plus a uniform
// per-step offset.
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ inline Rsrc rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
// The buffer from the records of step t on (t + kSkewD0 >= 0; `step` bytes
// per step): the sweeps move their buffers once per loop iteration, so every
// step's offset is a constant multiple of the step, held in a scalar register
// for the whole sweep, and a step issues no address arithmetic.
__device__ inline Rsrc rsrc_at(const void *p, unsigned bytes, int t, unsigned step) {
    const unsigned off = (unsigned)(t + kSkewD0) * step;
    return rsrc(static_cast<const char *>(p) + off, bytes - off);
}
__device__ inline Rec4 bload(Rsrc r, unsigned vo, unsigned so, Rec4 *) {
    const auto u = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0));
    const auto v = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, vo + 16, so, 0));
    return Rec4{u.x, u.y, v.x, v.y};
}
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
__device__ inline void bstore(Rsrc r, double v, unsigned vo, unsigned so) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, vo, so, 0);
}

// Quads: the values of four consecutive diagonals of one row, stored
// together (results, right-hand sides) or N coefficients per diagonal
// (the sweeps' quad records, 4N values per row and quad).  The buffer from
// the quad of diagonal t on (t + kSkewQ0 a multiple of 4), `qstride` bytes
// per quad of diagonals: a quad of values takes S*4*sizeof(T) bytes, which
// is the Rec4 records' stride per diagonal (`step`); a quad record N times
// that.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
template <typename T, int N>
struct Vec {
    T v[N];
};
__device__ inline Rsrc rsrc_q(const void *p, unsigned bytes, int t, unsigned qstride) {
    const unsigned off = (unsigned)((t + kSkewQ0) / 4) * qstride;
    return rsrc(static_cast<const char *>(p) + off, bytes - off);
}
__device__ inline void bstore4(Rsrc r, const float (&q)[4], unsigned vo, unsigned so) {
    const u32x4 v = {__float_as_uint(q[0]), __float_as_uint(q[1]), __float_as_uint(q[2]), __float_as_uint(q[3])};
    __builtin_amdgcn_raw_buffer_store_b128(v, r, vo, so, 0);
}
__device__ inline void bstore4(Rsrc r, const double (&q)[4], unsigned vo, unsigned so) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, double2{q[0], q[1]}), r, vo, so, 0);
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, double2{q[2], q[3]}), r, vo + 16, so, 0);
}
// Piece t (16 bytes) of N values into c
template <typename T, int N>
__device__ inline void bload_piece(Vec<T, N> &c, int t, Rsrc r, unsigned vo, unsigned so) {
    const u32x4 u = __builtin_amdgcn_raw_buffer_load_b128(r, vo + 16 * t, so, 0);
    if constexpr (sizeof(T) == 8) {
        const auto d = __builtin_bit_cast(double2, u);
        c.v[2 * t] = d.x;
        c.v[2 * t + 1] = d.y;
    } else {
        const auto f = __builtin_bit_cast(float4, u);
        c.v[4 * t] = f.x;
        c.v[4 * t + 1] = f.y;
        c.v[4 * t + 2] = f.z;
        c.v[4 * t + 3] = f.w;
    }
}
// A whole quad of N values per diagonal (4N values) in 16-byte loads.  The
// sweeps refill a quad after its last step: spreading the pieces over the
// quad's steps (each right after its last reader) measured slower.
template <typename T, int N>
__device__ inline void bload_quad(Vec<T, 4 * N> &c, Rsrc r, unsigned vo, unsigned so) {
#pragma unroll
    for (int t = 0; t < 4 * N * (int)sizeof(T) / 16; ++t) bload_piece(c, t, r, vo, so);
}

// Resets the block's LDS hand-over arrays ([W+1][HS]; array c lies below
// compute wave c): index 0 holds the padding value of step -1, the others
// the sentinel, except array `open`, which no wave or courier writes (the
// source of the grid's first wave): it holds the padding value throughout.
// Words [HS - kDummy, HS) of each array are a dummy area beyond every batch
// index: the lanes that write no batch write there, so that the hand-over
// store needs no lane mask.
template <typename T>
__device__ inline void lds_reset(T *lds, int count, int HS, T first, int open) {
    for (int q = threadIdx.x; q < count; q += blockDim.x)
        lds[q] = (q % HS == 0 || q / HS == open) ? first : hand_sentinel<T>();
}

// The courier waves of a block relay hand-over batches between blocks, in
// whole batches, kWin words per poll.  The incoming courier copies the
// neighbour block's device-coherent global array gin into LDS array lin; it
// keeps kPolls loads of its window in flight, so a batch that lands is seen
// about one load latency later (each sample carries the window it read, and
// the batches it shows complete beyond those already moved are stored).  The
// outgoing courier copies LDS array lout into the block's own global array
// gout as the compute wave writes it (its stores are never waited for).
// Indices 1 .. NG*kL; every loop is bounded and records a fault when a batch
// never comes.
// (measured best: 2 for FP64 batches, 4 for FP32 ones; 8 was slower)
template <typename T>
constexpr int kPolls = sizeof(T) == 8 ? 2 : 4;
constexpr int kPollLimit = 1 << 17;  // samples without progress (tens of ms) before a hand-over is declared lost

template <typename T>
__device__ void courier_in(const T *gin, T *lin, int NG, int *fault) {
    if (!gin) return;
    const int lane = threadIdx.x & 63;
    const int l = lane < kWin ? lane : kWin - 1;
    T val[kPolls<T>];
    int base[kPolls<T>];
#pragma unroll
    for (int p = 0; p < kPolls<T>; ++p) {
        base[p] = 0;
        val[p] = gload(gin + 1 + l);
    }
    int gi = 0, idle = 0;
    while (gi < NG && idle < kPollLimit) {
#pragma unroll
        for (int p = 0; p < kPolls<T>; ++p) {
            const int b0 = base[p];
            const unsigned long long bad = __builtin_amdgcn_ballot_w64(lane < kWin && is_sentinel(val[p]));
            const int upto = min(b0 + (bad ? (int)(__builtin_ctzll(bad) / kL) : kWin / kL), NG);
            if (upto > gi) {  // batches [gi, upto) are new in this sample
                const int idx = b0 * kL + 1 + lane;
                if (lane < kWin && idx > gi * kL && idx <= upto * kL) lstore(lin + idx, val[p]);
                gi = upto;
                idle = 0;
            } else {
                ++idle;
            }
            base[p] = gi;
            val[p] = gload(gin + gi * kL + 1 + l);  // gi <= NG: inside the array
        }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the samples still in flight
    if (gi < NG) record_fault(fault);
}

template <typename T>
__device__ void courier_out(const T *lout, T *gout, int NG, int *fault) {
    if (!gout) return;
    const int lane = threadIdx.x & 63;
    const bool act = lane < kWin;
    int go = 0, idle = 0;
    while (go < NG && idle < kSpinLimit) {
        const int idx = go * kL + 1 + (act ? lane : 0);
        const T v = lload(lout + idx);
        const unsigned long long bad = __builtin_amdgcn_ballot_w64(act && is_sentinel(v));
        const int nb = min(bad ? (int)(__builtin_ctzll(bad) / kL) : kWin / kL, NG - go);
        if (nb > 0) {
            if (lane < nb * kL) gstore(gout + idx, v);
            go += nb;
            idle = 0;
        } else {
            ++idle;
            __builtin_amdgcn_s_sleep(2);  // about 128 cycles (1 and 4 measured slower)
        }
    }
    if (go < NG) record_fault(fault);
}

// ---------------------------------------------------------------------------
// Per-group hand-over, shared by the three sweeps.  A wave's edge lane (the
// top lane forward, the bottom lane backward) keeps its kL results of a group
// and writes them as one batch at the group's end; every other lane writes
// its own kL words of the dummy area, so the store needs no lane mask.  The
// wave on the other side reads a group ahead of use both the batch, into
// registers (the same kL words in every lane: step k of the group takes word
// k as its edge lane's neighbour value, the `old` operand of the step's DPP
// shift, so no lane shift moves the batch along), and one check word per
// lane (lanes q < kL).  At the group's start it tests the check words read a
// group before (the compare's latency hidden behind a group of steps) and
// branches to the out-of-line wait, which reads the batch again, only when
// one was missing.  A wave starts once its source's second batch is there:
// two groups behind its source, every later batch is complete a group ahead
// of use, and the waits stay off the step path.
// ---------------------------------------------------------------------------
// Batch reads reach index (NG + 1) * kL + kL - 1, courier windows
// NG * kL + kWin - kL; the arrays are sized (HS, in skew_setup) to keep
// both below the dummy area [HS - kDummy, HS): kL words per lane at a
// stride of kL + 1 words (odd, so a store's lanes hit distinct banks).
constexpr int kLook = 4;  // groups per unrolled loop iteration
constexpr int kDummy = 64 * (kL + 1);
static_assert(kLook % 2 == 0, "input batches alternate between two register sets with the group's parity");

template <typename T>
struct Hand {
    const T *rd;   // this lane's check word of group 0 (lin + q)
    const T *lin;  // the input array
    T *wr;         // this lane's first write word of group 0 (batch word 1, or its dummy words)
    int wstep;     // words the write words advance per group (kL or 0)
    T chk;         // this lane's check word of the next group
    T in[2][kL];   // the batches of the current and the next group, by G % 2
    T *dummy;      // this lane's dummy words

    __device__ inline void start(const T *lin_, T *lout, int HS, int q, bool writer, int NG) {
        const int lane = threadIdx.x & 63;
        lin = lin_;
        rd = lin + q;
        dummy = lout + HS - kDummy + (kL + 1) * lane;
        wr = writer ? lout + 1 : dummy;
        wstep = writer ? kL : 0;
        // two groups behind the source (bounded: a batch that never comes is
        // found missing, and recorded, by the checks of the walk)
        (void)wait_batch(rd + min(1, NG - 1) * kL, T(0), 0);
        chk = lload(rd);
        read(0, 0);
        wait_lds();
        mirror();
    }
    // the LDS pattern of a group's end (kL words written) into the dummy
    // words, so that the loop's LDS waits count the same operations on entry
    // as on every later iteration
    __device__ inline void mirror() {
#pragma unroll
        for (int k = 0; k < kL; ++k) dummy[k] = T(0);
    }
    // all LDS operations done (ends the paths that re-read a batch, so that
    // the paths agree on what is still in flight)
    __device__ static inline void wait_lds() { __builtin_amdgcn_s_waitcnt(0xC07F); }
    // the whole batch of group G, in every lane (after its check)
    __device__ inline void read(int G, int slot) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);  // not above the check
        const T *p = lin + G * kL;
#pragma unroll
        for (int k = 0; k < kL; ++k) in[slot][k] = p[k];
    }
    // makes the batch of group G (slot h % 2, h = G % kLook) ready, waiting
    // for it and reading it again if its check failed, and reads the next
    // group's check words and batch
    __device__ inline void take(int G, int h, int &broken) {
        if (__builtin_amdgcn_ballot_w64(is_sentinel(chk))) {
            const T v = wait_batch(rd + G * kL, chk, broken);
            if (__builtin_amdgcn_ballot_w64(is_sentinel(v))) broken = 1;  // lost: the fault is recorded after the walk
            read(G, h % 2);
            wait_lds();
        }
        chk = lload(rd + (G + 1) * kL);
        read(G + 1, (h + 1) % 2);
    }
    // DPP shift by one lane whose edge lane takes the neighbour value of
    // step k of group G (h = G % kLook); the result replaces the spent
    // input, so the shift writes the input's own registers
    template <bool kUp>
    __device__ inline T shift_in(T v, int h, int k) {
        T &x = in[h % 2][k];
        x = kUp ? dpp_shr1(v, x) : dpp_shl1(v, x);
        return x;
    }
    // publishes the group's results (the edge lane's are the batch)
    __device__ inline void finish(int G, int h, const T (&res)[kL]) {
        (void)G, (void)h;
#pragma unroll
        for (int k = 0; k < kL; ++k) wr[k] = res[k];
        wr += wstep;
    }
};

// ---------------------------------------------------------------------------
// Diagonal recurrence (update_dilu_diagonal!): D into F.b (the quad-record
// pass, all CUs, turns it into Dinv = 1/D afterwards, off the recurrence's
// serial path)
// ---------------------------------------------------------------------------
template <int W>
__global__ void __launch_bounds__(64 * (W + 2))
    k_skew_update(SkewArgs a, const Rec4 *__restrict__ U, Rec4 *__restrict__ F, Rec4 *__restrict__ Bk,
                  double *__restrict__ hand) {
    extern __shared__ unsigned char skew_lds[];
    double *lds = reinterpret_cast<double *>(skew_lds);
    const int b = blockIdx.x;
    const SkewGrid g = a.grids[a.blk_grid[b]];
    const int HS = a.HS, rb = b - g.blk0[W - 1], nb = (g.nw + W - 1) / W;
    lds_reset(lds, (W + 1) * HS, HS, 1.0, rb == 0 ? 0 : -1);
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (w == W) {  // incoming courier: D of the row below from block rb-1
        courier_in(rb > 0 ? hand + (size_t)(b - 1) * HS : nullptr, lds, g.NG, a.fault);
        return;
    }
    if (w == W + 1) {  // outgoing courier: the top row's D to block rb+1
        courier_out(lds + W * HS, rb < nb - 1 ? hand + (size_t)b * HS : nullptr, g.NG, a.fault);
        return;
    }
    if (rb * W + w >= g.nw) return;  // past the grid's last wave (last block)
    const int lane = threadIdx.x & 63, wg = rb * W + w;  // wave of the grid
    const unsigned S = g.S, j = wg * 64 + lane;
    const unsigned vo = (g.base + j) * (unsigned)sizeof(Rec4), step = S * (unsigned)sizeof(Rec4);

    constexpr int M = ring_batches<double>(), PD = M * kL;
    static_assert(kLook % M == 0, "whole rings per unrolled block");
    Hand<double> hd;
    hd.start(lds + w * HS, lds + (w + 1) * HS, HS, min(lane, kL - 1), lane == 63, g.NG);
    // the prologue issues what a step issues (a store of the pivot 1.0 that
    // the pre-start padding holds already, one record load), so that the
    // loop's memory waits count the same outstanding operations every step
    Rec4 ring[PD];
    {
        const Rsrc ru = rsrc_at(U, a.rbytes, 0, step), rf = rsrc_at(F, a.rbytes, -PD, step);
#pragma unroll
        for (int k = 0; k < PD; ++k) {
            bstore(rf, 1.0, vo + 8, k * step);
            ring[k] = bload(ru, vo, k * step, (Rec4 *)nullptr);
            __builtin_amdgcn_sched_barrier(0);  // issue in consumption order
        }
    }
    double dprev = 1.0, res[kL];
    int broken = 0;
    int G0 = 0;
    do {
        // this iteration's steps G0*kL + i, i < kLook*kL: stores at i*step, refills PD steps on
        const Rsrc ru = rsrc_at(U, a.rbytes, G0 * kL + PD, step), rf = rsrc_at(F, a.rbytes, G0 * kL, step);
#pragma unroll
        for (int h = 0; h < kLook; ++h) {
            const int G = G0 + h;
            hd.take(G, h, broken);
#pragma unroll
            for (int k = 0; k < kL; ++k) {
                const int r = (h % M) * kL + k;
                const unsigned si = (unsigned)(h * kL + k) * step;
                const Rec4 u = ring[r];
                const double ds = hd.template shift_in<true>(dprev, h, k);
                const double d = (u.a - u.b / ds) - u.c / dprev;
                bstore(rf, d, vo + 8, si);
                dprev = res[k] = d;
                __builtin_amdgcn_sched_barrier(0);  // the slot is consumed: refill it below
                ring[r] = bload(ru, vo, si, (Rec4 *)nullptr);
            }
            hd.finish(G, h, res);
        }
        G0 += kLook;
    } while (G0 < g.NG);  // NG >= 1: no guard between the prologue and the loop
    if (broken) record_fault(a.fault);
}

// ---------------------------------------------------------------------------
// ldiv!: forward sweep into quads of y (yq), backward sweep into quads of z
// (zp).  FP64 evaluates the reference's expressions; FP32 (the inner solves
// of mixed precision) works on coefficients premultiplied by their pivots,
// with fused multiply-adds.
//
// FP64 forward: one record per step, F = (b, Dinv, A_nr, A_rw), A_nr the
// north neighbour's coefficient of this row (A[r+nx, r]).  Each lane forms,
// right after its own y_r, the term the row above subtracts, (A_nr y_r)
// Dinv_r, which is that row's (A_rs y_s) Dinv_s with the same operands in
// the same order; the term moves up one lane (DPP) as the step's only
// cross-lane value (the wave's top lane's terms are its hand-over batch),
// and the west pivot is the wave's own previous Dinv.  The other sweeps
// read quad records, N values per diagonal for four diagonals at once (one
// refill per four steps):
//   forward  FP32 (b quads apart)   (A_rs Dinv_s, A_rw Dinv_w)
//   backward FP64 (Dinv, A_re, A_rn),   FP32 (Dinv, A_re Dinv, A_rn Dinv)
// with Dinv_s / Dinv_w the pivots of the south / west neighbours (slots
// (T-1, j-1) and (T-1, j)).  Quad records would carry both pivots in the
// FP64 forward sweep too, but at 48 bytes per step its prefetch depth (63
// memory operations in flight) no longer covers the load latency.
// ---------------------------------------------------------------------------
constexpr int kFwdN = 2;  // FP32 forward quad records
constexpr int kBwdN = 3;

template <typename T, int W>
__global__ void __launch_bounds__(64 * (W + 2))
    k_skew_fwd(SkewArgs a, const void *__restrict__ recs, const T *__restrict__ bq, T *__restrict__ yq,
               T *__restrict__ hand, const KScal *ks) {
    using R = Rec4T<T>;
    extern __shared__ unsigned char skew_lds[];
    T *lds = reinterpret_cast<T *>(skew_lds);
    const int b = blockIdx.x;
    const int gi = a.blk_grid[b];
    if (ks && ks[a.ks_per_grid ? gi : 0].done) return;
    const SkewGrid g = a.grids[gi];
    const int HS = a.HS, rb = b - g.blk0[W - 1], nb = (g.nw + W - 1) / W;
    lds_reset(lds, (W + 1) * HS, HS, T(0), rb == 0 ? 0 : -1);
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (w == W) {  // couriers: the forward channel of block b is hand[2b]
        courier_in(rb > 0 ? hand + (size_t)(b - 1) * 2 * HS : nullptr, lds, g.NG, a.fault);
        return;
    }
    if (w == W + 1) {
        courier_out(lds + W * HS, rb < nb - 1 ? hand + (size_t)b * 2 * HS : nullptr, g.NG, a.fault);
        return;
    }
    if (rb * W + w >= g.nw) return;  // past the grid's last wave (last block)
    const int lane = threadIdx.x & 63, wg = rb * W + w;
    const unsigned S = g.S, j0 = wg * 64;
    const unsigned vo = (g.base + j0 + lane) * (unsigned)sizeof(R), step = S * (unsigned)sizeof(R);
    // lane q < kL of the batch of group G holds step G*kL + q
    const int ql = min(lane, kL - 1);

    constexpr int M = ring_batches<T>(), PD = M * kL;
    static_assert(kLook % M == 0, "whole rings per unrolled block");
    constexpr bool kFused = sizeof(T) == 4;
    constexpr int N = kFwdN, NR = PD / 4;
    const unsigned vq = (g.base + 4 * (j0 + lane)) * (unsigned)sizeof(T), qbytes = a.rbytes / 4;
    const unsigned vc = N * vq, cstride = N * step;
    const R *F = static_cast<const R *>(recs);
    Hand<T> hd;
    hd.start(lds + w * HS, lds + (w + 1) * HS, HS, ql, lane == 63, g.NG);
    R ring[kFused ? 1 : PD];           // FP64 records
    Vec<T, 4 * N> cr[kFused ? NR : 1];  // FP32 coefficient quads
    Vec<T, 4> br[kFused ? NR : 1];      // FP32 right-hand-side quads
    T q[4] = {T(0), T(0), T(0), T(0)};
    // The prologue stores quads of +0 into the wave's own pre-start padding
    // (its value already), so that it issues what the loop issues and the
    // loop's memory waits count the same outstanding operations throughout.
    {
        const Rsrc rq = rsrc_q(yq, qbytes, -PD, step);
#pragma unroll
        for (int k = 0; k < PD; ++k) {
            const int mq = k / 4;
            if (k % 4 == 3) bstore4(rq, q, vq, (unsigned)mq * step);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (kFused) {
                if (k % 4 == 3) {
                    bload_quad<T, N>(cr[mq], rsrc_q(recs, N * qbytes, 0, cstride), vc, (unsigned)mq * cstride);
                    bload_quad<T, 1>(br[mq], rsrc_q(bq, qbytes, 0, step), vq, (unsigned)mq * step);
                }
            } else {
                ring[(k + PD - 1) % PD] = bload(rsrc_at(F, a.rbytes, -1, step), vo, k * step, (R *)nullptr);
            }
        }
    }
    T xprev = 0, mprev = 0, res[kL];
    int broken = 0;
    int G0 = 0;
    do {
        // this iteration's steps G0*kL + i, i < kLook*kL: results in quads
        // from G0*kL on, refills PD steps on
        const Rsrc rq = rsrc_q(yq, qbytes, G0 * kL, step), rf = rsrc_at(F, a.rbytes, G0 * kL + PD - 1, step);
        const Rsrc rc = rsrc_q(recs, N * qbytes, G0 * kL + PD, cstride), rbn = rsrc_q(bq, qbytes, G0 * kL + PD, step);
#pragma unroll
        for (int h = 0; h < kLook; ++h) {
            const int G = G0 + h;
            hd.take(G, h, broken);
#pragma unroll
            for (int k = 0; k < kL; ++k) {
                const int i = h * kL + k, mq = i / 4, wq = i % 4;
                T x;
                if constexpr (kFused) {  // (A_rs Dinv_s, A_rw Dinv_w) and b; the west term first: it
                                         // fills the wait between xprev's FMA and the lane shift
                    const T *c = &cr[mq % NR].v[N * wq];
                    const T t = __builtin_fmaf(-c[1], xprev, br[mq % NR].v[wq]);
                    __builtin_amdgcn_sched_barrier(0);
                    const T xs = hd.template shift_in<true>(xprev, h, k);
                    x = __builtin_fmaf(-c[0], xs, t);
                } else {  // (b, Dinv, A_nr, A_rw): the reference's order of operations, the south
                          // term (A_rs y_s) Dinv_s formed by the lane below and shifted in
                    const T ms = hd.template shift_in<true>(mprev, h, k);
                    const R c = ring[i % PD];
                    const T dprev = ring[(i + PD - 1) % PD].b;  // the previous step's Dinv, still in its slot
                    x = (c.a - ms) - c.d * xprev * dprev;
                    mprev = c.c * x * c.b;  // (A_nr y_r) Dinv_r: the row above's south term
                }
                q[wq] = x;
                if (wq == 3) bstore4(rq, q, vq, (unsigned)mq * step);
                xprev = x;
                res[k] = kFused ? x : mprev;
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (kFused) {
                    if (wq == 3) {
                        bload_quad<T, N>(cr[mq % NR], rc, vc, (unsigned)mq * cstride);
                        bload_quad<T, 1>(br[mq % NR], rbn, vq, (unsigned)mq * step);
                    }
                } else {
                    ring[(i + PD - 1) % PD] = bload(rf, vo, (unsigned)i * step, (R *)nullptr);
                }
            }
            hd.finish(G, h, res);
        }
        G0 += kLook;
    } while (G0 < g.NG);
    if (broken) record_fault(a.fault);
}

template <typename T, int W>
__global__ void __launch_bounds__(64 * (W + 2))
    k_skew_bwd(SkewArgs a, const T *__restrict__ kb, const T *__restrict__ yq, T *__restrict__ zp,
               T *__restrict__ hand, const KScal *ks) {
    using R = Rec4T<T>;
    extern __shared__ unsigned char skew_lds[];
    T *lds = reinterpret_cast<T *>(skew_lds);
    const int b = a.nblk - 1 - (int)blockIdx.x;  // top row blocks first
    const int gi = a.blk_grid[b];
    if (ks && ks[a.ks_per_grid ? gi : 0].done) return;
    const SkewGrid g = a.grids[gi];
    const int HS = a.HS, rb = b - g.blk0[W - 1], nb = (g.nw + W - 1) / W;
    // the grid's top wave reads the array above it, which nobody writes
    lds_reset(lds, (W + 1) * HS, HS, T(0), rb == nb - 1 ? g.nw - rb * W : -1);
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (w == W) {  // couriers: the backward channel of block b is hand[2b+1]
        courier_in(rb < nb - 1 ? hand + ((size_t)(b + 1) * 2 + 1) * HS : nullptr, lds + W * HS, g.NG, a.fault);
        return;
    }
    if (w == W + 1) {
        courier_out(lds, rb > 0 ? hand + ((size_t)b * 2 + 1) * HS : nullptr, g.NG, a.fault);
        return;
    }
    if (rb * W + w >= g.nw) return;  // past the grid's last wave (last block)
    const int lane = threadIdx.x & 63, wg = rb * W + w;
    const unsigned S = g.S, j0 = wg * 64;
    const int Ttop = g.NG * kL - 1;  // step u handles anti-diagonal Ttop - u
    const unsigned step = S * (unsigned)sizeof(R);  // bytes per quad of values
    constexpr int NI = kLook * kL;  // steps per loop iteration
    // lane 63-q of the batch of group G holds step u = G*kL + q (index u)
    const int qh = min(63 - lane, kL - 1);

    constexpr int M = ring_batches<T>(), PD = M * kL;
    static_assert(kLook % M == 0, "whole rings per unrolled block");
    Hand<T> hd;
    hd.start(lds + (w + 1) * HS, lds + w * HS, HS, qh, lane == 0, g.NG);
    // Ttop + 1 + kSkewQ0 is a multiple of 4, so step i of an iteration is
    // word 3 - i%4 of the iteration's quad i/4 (four diagonals, descending)
    constexpr bool kFused = sizeof(T) == 4;
    constexpr int N = kBwdN, NQ = NI / 4, NR = PD / 4;  // quads per loop iteration, quads in the rings
    static_assert(NQ % NR == 0, "whole rings per loop iteration");
    const unsigned vq = (g.base + 4 * (j0 + lane)) * (unsigned)sizeof(T), qbytes = a.rbytes / 4;
    const unsigned vc = N * vq, cstride = N * step;
    Vec<T, 4 * N> cr[NR];  // coefficient quads
    Vec<T, 4> yr[NR];      // y quads
    T q[4] = {T(0), T(0), T(0), T(0)};
    // prologue: the first PD steps' quads, and quads of +0 into the padding
    // diagonals above the top (see the forward sweep), in the loop's order;
    // offsets count down from the top of each range
    {
        const int tl = Ttop - (NI - 1);
        const Rsrc rc = rsrc_q(kb, N * qbytes, tl, cstride), ry = rsrc_q(yq, qbytes, tl, step);
        const Rsrc rzq = rsrc_q(zp, qbytes, Ttop + 1, step);
#pragma unroll
        for (int k = 0; k < PD; ++k) {
            const int mq = k / 4;
            if (k % 4 == 3) bstore4(rzq, q, vq, (unsigned)mq * step);
            __builtin_amdgcn_sched_barrier(0);
            if (k % 4 == 3) {
                bload_quad<T, N>(cr[mq], rc, vc, (unsigned)(NQ - 1 - mq) * cstride);
                bload_quad<T, 1>(yr[mq], ry, vq, (unsigned)(NQ - 1 - mq) * step);
            }
        }
    }
    T zprev = 0, res[kL];
    int broken = 0;
    int G0 = 0;
    do {
        // this iteration's steps u = G0*kL + i: z of diagonal Ttop - u, refills of Ttop - u - PD
        const int tl = Ttop - G0 * kL - (NI - 1);  // lowest diagonal written this iteration
        const Rsrc rc = rsrc_q(kb, N * qbytes, tl - PD, cstride), ry = rsrc_q(yq, qbytes, tl - PD, step);
        const Rsrc rzq = rsrc_q(zp, qbytes, tl, step);
#pragma unroll
        for (int h = 0; h < kLook; ++h) {
            const int G = G0 + h;
            hd.take(G, h, broken);
#pragma unroll
            for (int k = 0; k < kL; ++k) {
                const int i = h * kL + k, mq = i / 4, wq = 3 - i % 4, rq = mq % NR;
                const T *c = &cr[rq].v[N * wq];
                const T y = yr[rq].v[wq];
                T z;
                if constexpr (kFused) {  // the east term first: it fills the wait before the lane shift
                    const T t = __builtin_fmaf(-c[1], zprev, y * c[0]);
                    __builtin_amdgcn_sched_barrier(0);
                    const T zn = hd.template shift_in<false>(zprev, h, k);
                    z = __builtin_fmaf(-c[2], zn, t);
                } else {  // the reference's order of operations
                    const T zn = hd.template shift_in<false>(zprev, h, k);
                    z = ((y * c[0]) - c[1] * zprev * c[0]) - c[2] * zn * c[0];
                }
                q[wq] = z;
                if (wq == 0) bstore4(rzq, q, vq, (unsigned)(NQ - 1 - mq) * step);
                zprev = res[k] = z;
                __builtin_amdgcn_sched_barrier(0);
                if (wq == 0) {
                    bload_quad<T, N>(cr[rq], rc, vc, (unsigned)(NQ - 1 - mq) * cstride);
                    bload_quad<T, 1>(yr[rq], ry, vq, (unsigned)(NQ - 1 - mq) * step);
                }
            }
            hd.finish(G, h, res);
        }
        G0 += kLook;
    } while (G0 < g.NG);
    if (broken) record_fault(a.fault);
}

// ---------------------------------------------------------------------------
// All-CU helpers: padding, packing, gather / scatter, FP32 copies
// ---------------------------------------------------------------------------
__global__ void k_skew_init(size_t ns, Rec4 *F, Rec4 *Bk, Rec4 *U) {
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    F[s] = Rec4{0.0, 1.0, 0.0, 0.0};
    Bk[s] = Rec4{0.0, 1.0, 0.0, 0.0};
    U[s] = Rec4{1.0, 0.0, 0.0, 0.0};
}

template <typename T>
__global__ void k_skew_fill(size_t ns, T *x, T v) {
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < ns) x[s] = v;
}

// Natural CSR values of the full-rectangle pattern -> skewed records; also
// resets the diagonal recurrence's global hand-over arrays.
__global__ void k_skew_pack(int n, const SkewGrid *__restrict__ grids, const int *__restrict__ cell_grid,
                            const int *__restrict__ slot_of, const int *__restrict__ rowptr,
                            const int *__restrict__ diag, const double *__restrict__ val, Rec4 *__restrict__ F,
                            Rec4 *__restrict__ Bk, Rec4 *__restrict__ U, double *__restrict__ hand, int nhand,
                            int HS) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nhand) hand[p] = (p % HS == 0) ? 1.0 : hand_sentinel<double>();
    if (p >= n) return;
    const SkewGrid g = grids[cell_grid ? cell_grid[p] : 0];
    const int rl = p - g.o, nx = g.nx;
    const int i = rl % nx, j = rl / nx;
    const bool S = j > 0, Wn = i > 0, E = i < nx - 1, N = j < g.ny - 1;
    const int d = diag[p];
    const double a_rs = S ? val[rowptr[p]] : 0.0;               // A[r, r-nx]
    const double a_nr = N ? val[rowptr[p + nx]] : 0.0;          // A[r+nx, r]
    const double a_sr = S ? val[rowptr[p - nx + 1] - 1] : 0.0;  // A[r-nx, r]
    const double a_rw = Wn ? val[d - 1] : 0.0;                  // A[r, r-1]
    const double a_wr = Wn ? val[diag[p - 1] + 1] : 0.0;        // A[r-1, r]
    const int s = slot_of[p];
    U[s] = Rec4{val[d], a_rs * a_sr, a_rw * a_wr, 0.0};
    F[s].c = a_nr;
    F[s].d = a_rw;
    Bk[s].c = E ? val[d + 1] : 0.0;                // A[r, r+1]
    Bk[s].d = N ? val[rowptr[p + 1] - 1] : 0.0;    // A[r, r+nx]
}

// Natural order <-> slot order in 64 x 64 tiles of one grid (rows j0..j0+63,
// columns i0..i0+63), through LDS: the natural side is read / written along
// grid rows, the slot side along anti-diagonals (consecutive rows of one
// diagonal are consecutive slots), so both sides are coalesced.
constexpr int kTile = 64;

template <typename T>
struct TileLds {
    T v[kTile][kTile + 1];
};

// x (natural) -> the forward sweep's right-hand side: FP64 into F.a, FP32
// into quads (bq); padding slots keep their +0.  Blocks beyond the tiles
// reset the sweeps' global hand-over arrays.
template <typename T>
__global__ void __launch_bounds__(256)
    k_skew_gather(int ntiles, const SkewTile *__restrict__ tiles, const SkewGrid *__restrict__ grids,
                  const T *__restrict__ x, Rec4T<T> *__restrict__ F, T *__restrict__ bq, T *__restrict__ hand,
                  int nhand, int HS, const KScal *ks) {
    if (ks && ks->done) return;
    if ((int)blockIdx.x >= ntiles) {
        const int p = ((int)blockIdx.x - ntiles) * 256 + threadIdx.x;
        if (p < nhand) hand[p] = (p % HS == 0) ? T(0) : hand_sentinel<T>();
        return;
    }
    __shared__ TileLds<T> t;
    const SkewTile tl = tiles[blockIdx.x];
    const SkewGrid g = grids[tl.grid];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = tl.i0 + lane;
    T v[kTile / 4];  // all loads in flight before the LDS stores
#pragma unroll
    for (int it = 0; it < kTile / 4; ++it) {
        const int j = tl.j0 + wv + 4 * it;
        v[it] = (i < g.nx && j < g.ny) ? x[(size_t)g.o + i + (size_t)j * g.nx] : T(0);
    }
#pragma unroll
    for (int it = 0; it < kTile / 4; ++it) t.v[wv + 4 * it][lane] = v[it];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kTile / 2; ++it) {  // diagonal T = i0 + j0 + d, row j0 + lane
        const int d = wv + 4 * it, j = tl.j0 + lane, il = d - lane;
        const int Td = tl.i0 + tl.j0 + d, Tq = Td + kSkewQ0;
        if (il >= 0 && il < kTile && tl.i0 + il < g.nx && j < g.ny) {
            if constexpr (sizeof(T) == 8)
                F[g.base + (size_t)(Td + kSkewD0) * g.S + j].a = t.v[lane][il];
            else
                bq[g.base + (size_t)(Tq & ~3) * g.S + 4 * (size_t)j + (Tq & 3)] = t.v[lane][il];
        }
    }
}

// The backward sweep's quads (zq) -> y (natural).
template <typename T>
__global__ void __launch_bounds__(256)
    k_skew_scatter_q(const SkewTile *__restrict__ tiles, const SkewGrid *__restrict__ grids,
                     const T *__restrict__ zq, T *__restrict__ y, const KScal *ks) {
    if (ks && ks->done) return;
    __shared__ TileLds<T> t;
    const SkewTile tl = tiles[blockIdx.x];
    const SkewGrid g = grids[tl.grid];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    T v[kTile / 2];  // all loads in flight before the LDS stores
#pragma unroll
    for (int it = 0; it < kTile / 2; ++it) {
        const int d = wv + 4 * it, j = tl.j0 + lane, il = d - lane;
        const bool in = il >= 0 && il < kTile && tl.i0 + il < g.nx && j < g.ny;
        const int Tq = tl.i0 + tl.j0 + d + kSkewQ0;
        v[it] = in ? zq[g.base + (size_t)(Tq & ~3) * g.S + 4 * (size_t)j + (Tq & 3)] : T(0);
    }
#pragma unroll
    for (int it = 0; it < kTile / 2; ++it) {
        const int il = wv + 4 * it - lane;
        if (il >= 0 && il < kTile) t.v[lane][il] = v[it];
    }
    __syncthreads();
    const int i = tl.i0 + lane;
#pragma unroll
    for (int it = 0; it < kTile / 4; ++it) {
        const int r = wv + 4 * it, j = tl.j0 + r;
        if (i < g.nx && j < g.ny) y[(size_t)g.o + i + (size_t)j * g.nx] = t.v[r][lane];
    }
}

// The coefficients a cell does not have in the backward sweep (east of each
// row's last cell, north of the top row) become copysign(0, Dinv), so that
// their products vanish exactly (x - (+0) == x, x == -0 included); Dinv's
// sign is D's (F.b, written by the diagonal recurrence).  One thread per
// edge cell, blockIdx.y: the grid.
__global__ void k_skew_edges(const SkewGrid *__restrict__ grids, const Rec4 *__restrict__ F, Rec4 *__restrict__ Bk) {
    const SkewGrid g = grids[blockIdx.y];
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    auto slot = [&](int T, int j) { return g.base + (size_t)(T + kSkewD0) * g.S + (size_t)j; };
    if (q < g.ny) {  // row q's east-edge cell, step q + nx - 1
        const size_t s = slot(q + g.nx - 1, q);
        Bk[s].c = copysign(0.0, F[s].b);
    } else if (q < g.ny + g.nx) {  // top-row cell of column q - ny
        const int j = g.ny - 1;
        const size_t s = slot(j + q - g.ny, j);
        Bk[s].d = copysign(0.0, F[s].b);
    }
}

// The sweeps' quad records from the records F and Bk (after the diagonal
// recurrence and the edge coefficients): one thread per row and quad of
// diagonals, for the diagonals T in [-kSkewPD, NG*kSkewL + kSkewPD) the
// sweeps read (prefetches reach kSkewPD diagonals beyond either end).
//   forward  FP32 (A_rs Dinv_s, A_rw Dinv_w)   (FP64 reads F itself)
//   backward FP64 (Dinv, A_re, A_rn),  FP32 (Dinv, A_re Dinv, A_rn Dinv)
// FP32 premultiplies in FP64 and rounds once.  Dinv_s / Dinv_w are the
// pivots of slots (T-1, j-1) / (T-1, j); row 0's south slot is the previous
// diagonal's last slot, a padding row (pivot 1).  blockIdx.y: the grid.
static_assert(kSkewPD % 4 == 0 && (kSkewQ0 - kSkewPD) % 4 == 0, "quads start at a padding diagonal");
template <typename T>
__global__ void k_skew_quads(const SkewGrid *__restrict__ grids, Rec4 *__restrict__ F, Rec4 *__restrict__ Bk,
                             T *__restrict__ kf, T *__restrict__ kb) {
    constexpr int NF = kFwdN, NB = kBwdN;
    const SkewGrid g = grids[blockIdx.y];
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ((size_t)g.NG * kSkewL + 2 * kSkewPD) / 4 * g.S) return;
    const int qi = (int)(q / g.S) + (kSkewQ0 - kSkewPD) / 4, j = (int)(q % g.S);
    const size_t s0 = g.base + (size_t)(4 * qi - kSkewQ0 + kSkewD0) * g.S + j;  // slot of (T, j), word 0
    T cf[4 * NF], cb[4 * NB];
#pragma unroll
    for (int w = 0; w < 4; ++w) {  // diagonal T = 4 qi + w - kSkewQ0
        const size_t s = s0 + (size_t)w * g.S;
        if constexpr (sizeof(T) == 8) {  // F.b holds D: Dinv into F.b, Bk.b and the quad record
            const double di = 1.0 / F[s].b;
            F[s].b = di;
            Bk[s].b = di;
            const Rec4 v = Bk[s];
            cb[3 * w] = di;
            cb[3 * w + 1] = v.c;
            cb[3 * w + 2] = v.d;
        } else {
            const Rec4 u = F[s], v = Bk[s];
            const double ds = F[s - g.S - 1].b, dw = F[s - g.S].b;
            cf[2 * w] = (float)(F[s - g.S - 1].c * ds);  // the south slot's A_nr is A_rs
            cf[2 * w + 1] = (float)(u.d * dw);
            cb[3 * w] = (float)v.b;
            cb[3 * w + 1] = (float)(v.c * v.b);
            cb[3 * w + 2] = (float)(v.d * v.b);
        }
    }
    const size_t o = g.base + (size_t)qi * 4 * g.S + 4 * (size_t)j;  // the quad's first value
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int t = 0; t < 4 * NF; ++t) kf[NF * o + t] = cf[t];
    }
#pragma unroll
    for (int t = 0; t < 4 * NB; ++t) kb[NB * o + t] = cb[t];
}

__global__ void k_skew_dinv_natural(int n, const int *__restrict__ slot_of, const Rec4 *__restrict__ F,
                                    double *__restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = F[slot_of[p]].b;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static dim3 grid_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

static int max_lds_bytes() {
    static int v = -1;
    if (v < 0) {
        int dev = 0, b = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&b, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || b <= 0)
            b = 64 * 1024;
        v = b;
    }
    return v;
}

int skew_setup(fvh_mesh *m, const std::vector<SkewRect> &rects) {
    m->skew.reset();
    const char *force = std::getenv("FVH_SWEEP");
    if (force && std::strcmp(force, "levels") == 0) return FVH_OK;
    if (rects.empty()) return FVH_OK;
    int maxNG = 0;
    // step groups: a whole number of rings of every depth (4 batches)
    auto ng_of = [](const SkewRect &r) { return 4 * ((r.nx + r.ny - 1 + 4 * kL - 1) / (4 * kL)); };
    for (const SkewRect &r : rects) maxNG = std::max(maxNG, ng_of(r));
    // hand-over arrays: indices up to NG*kL, read a window (kWin) or a ring of
    // batches (at most 4) past the last one
    const int HS = (maxNG * kL + kWin + kLook * kL + kDummy + 63) / 64 * 64;
    if (((size_t)2 * HS + 64) * sizeof(double) > (size_t)max_lds_bytes()) return FVH_OK;  // too long even for W = 1
    auto L = std::make_unique<SkewLayout>();
    L->HS = HS;
    size_t base = 0;
    std::vector<int> blk_grid[kSkewMaxW];
    for (size_t k = 0; k < rects.size(); ++k) {
        const SkewRect &r = rects[k];
        SkewGrid g{};
        g.nx = r.nx;
        g.ny = r.ny;
        g.o = r.o;
        g.nw = (r.ny + 1 + 63) / 64;  // at least one padding row
        g.NG = ng_of(r);
        g.S = 64 * g.nw;
        g.base = (unsigned)base;
        for (int W = 1; W <= kSkewMaxW; ++W) {
            g.blk0[W - 1] = L->nblk[W - 1];
            const int nb = (g.nw + W - 1) / W;
            for (int q = 0; q < nb; ++q) blk_grid[W - 1].push_back((int)k);
            L->nblk[W - 1] += nb;
        }
        const size_t nd = (size_t)g.NG * kL + kSkewD0 + kSkewTop;
        base += nd * (size_t)g.S;
        // records are addressed by 32-bit byte offsets
        if (base * sizeof(Rec4) >= ((size_t)1 << 31)) return FVH_OK;
        L->h_grids.push_back(g);
    }
    L->NS = base;
    std::vector<int> slot_of(m->n, 0), qoff(m->n, 0);
    for (const SkewGrid &g : L->h_grids)
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) {
                const size_t r = (size_t)g.o + i + (size_t)j * g.nx;
                const int Tq = i + j + kSkewQ0;
                slot_of[r] = (int)(g.base + (size_t)(i + j + kSkewD0) * g.S + j);
                qoff[r] = (int)(g.base + (size_t)(Tq & ~3) * g.S + 4 * (size_t)j + (Tq & 3));
            }
    std::vector<SkewTile> tiles;
    for (size_t k = 0; k < L->h_grids.size(); ++k)
        for (int j0 = 0; j0 < L->h_grids[k].ny; j0 += kTile)
            for (int i0 = 0; i0 < L->h_grids[k].nx; i0 += kTile) tiles.push_back(SkewTile{(int)k, i0, j0});
    FVH_TRY(L->grids.upload(L->h_grids));
    for (int W = 1; W <= kSkewMaxW; ++W) FVH_TRY(L->blk_grid[W - 1].upload(blk_grid[W - 1]));
    FVH_TRY(L->slot_of.upload(slot_of));
    FVH_TRY(L->qoff.upload(qoff));
    FVH_TRY(L->tiles.upload(tiles));
    m->skew = std::move(L);
    return FVH_OK;
}

bool skew_active(const fvh_mesh *m) {
    if (!m->skew) return false;
    const char *env = std::getenv("FVH_BAND_ROWS");
    return !(env && env[0] >= '1' && env[0] <= '4' && env[1] == 0);
}

// Waves per block: FVH_SKEW_WAVES=1..3 forces one, otherwise the tuned value
// (3 before tuning), never more than the LDS holds.
static int forced_waves() {
    const char *env = std::getenv("FVH_SKEW_WAVES");
    if (env && env[0] >= '1' && env[0] <= '0' + kSkewMaxW && env[1] == 0) return env[0] - '0';
    return 0;
}
static int lds_waves_max(const SkewLayout &L) {
    int W = kSkewMaxW;
    while (W > 1 && ((size_t)(W + 1) * L.HS + 64 * W) * sizeof(double) > (size_t)max_lds_bytes()) --W;
    return W;
}
static int skew_waves(const SkewLayout &L, bool fp32 = false) {
    const int f = forced_waves();
    const int tuned = fp32 ? L.W32 : L.W;
    const int w = f ? f : (tuned ? tuned : kSkewMaxW);
    return std::min(w, lds_waves_max(L));
}


static const int *cell_grid(const fvh_mesh *m) { return m->cases ? m->cases->cell_case.p : nullptr; }

static int skew_alloc(fvh_eqn *e) {
    const SkewLayout &L = *e->m->skew;
    SkewStore &s = e->sk;
    if (s.F.n == L.NS) return FVH_OK;
    FVH_TRY(s.F.alloc(L.NS));
    FVH_TRY(s.Bk.alloc(L.NS));
    FVH_TRY(s.U.alloc(L.NS));
    FVH_TRY(s.uhand.alloc((size_t)L.nblk[0] * L.HS));  // W = 1 has the most blocks
    if (!s.h_fault) {
        HIP_TRY(hipHostMalloc((void **)&s.h_fault, sizeof(int), hipHostMallocDefault));
        *s.h_fault = 0;
    }
    hipLaunchKernelGGL(k_skew_init, grid_for(L.NS), dim3(256), 0, stream(), L.NS, s.F.p, s.Bk.p, s.U.p);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// A precision's sweep buffers, all +0 at first (the padding slots of bq keep
// it: the gather writes only cells).
template <typename T>
static int sweep_alloc(SkewSweep<T> &w, const SkewLayout &L) {
    if (w.zp.n == L.NS) return FVH_OK;
    const bool q32 = sizeof(T) == 4;  // FP64's forward sweep reads F
    FVH_TRY(w.kf.alloc(q32 ? kFwdN * L.NS : 0));
    FVH_TRY(w.kb.alloc(kBwdN * L.NS));
    FVH_TRY(w.bq.alloc(q32 ? L.NS : 0));
    FVH_TRY(w.yq.alloc(L.NS));
    FVH_TRY(w.zp.alloc(L.NS));
    FVH_TRY(w.hand.alloc((size_t)L.nblk[0] * 2 * L.HS));  // W = 1 has the most blocks
    for (DBuf<T> *b : {&w.kf, &w.kb, &w.bq, &w.yq, &w.zp})
        hipLaunchKernelGGL(k_skew_fill<T>, grid_for(b->n), dim3(256), 0, stream(), b->n, b->p, T(0));
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// The quad records of one precision from the current F and Bk.
template <typename T>
static int sweep_records(fvh_eqn *e, SkewSweep<T> &w) {
    const SkewLayout &L = *e->m->skew;
    FVH_TRY(sweep_alloc(w, L));
    size_t most = 0;
    for (const SkewGrid &g : L.h_grids) most = std::max(most, ((size_t)g.NG * kSkewL + 2 * kSkewPD) / 4 * g.S);
    hipLaunchKernelGGL((k_skew_quads<T>), dim3((unsigned)((most + 255) / 256), (unsigned)L.h_grids.size()), dim3(256),
                       0, stream(), L.grids.p, e->sk.F.p, e->sk.Bk.p, w.kf.p, w.kb.p);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

template <typename R>
static SkewArgs skew_args(const fvh_eqn *e, int W) {
    const SkewLayout &L = *e->m->skew;
    return SkewArgs{L.grids.p, L.blk_grid[W - 1].p, L.nblk[W - 1], L.HS, e->m->cases ? 1 : 0,
                    (unsigned)(L.NS * sizeof(R)), e->sk.h_fault};
}

// Dynamic LDS beyond 64 KiB has to be allowed per kernel.
template <typename K>
static int allow_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)bytes));
    return FVH_OK;
}

template <int W>
static int launch_update_w(fvh_eqn *e) {
    const SkewLayout &L = *e->m->skew;
    const size_t lds = (size_t)(W + 1) * L.HS * sizeof(double);
    FVH_TRY(allow_lds(k_skew_update<W>, lds));
    hipLaunchKernelGGL((k_skew_update<W>), dim3(L.nblk[W - 1]), dim3(64 * (W + 2)), lds, stream(),
                       skew_args<Rec4>(e, W), e->sk.U.p, e->sk.F.p, e->sk.Bk.p, e->sk.uhand.p);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

static int skew_tune(fvh_eqn *e);

int skew_update(fvh_eqn *e, const double *val) {
    fvh_mesh *m = e->m;
    const SkewLayout &L = *m->skew;
    FVH_TRY(skew_alloc(e));
    const int n = (int)m->n, nhand = L.nblk[0] * L.HS;
    hipLaunchKernelGGL(k_skew_pack, grid_for(std::max(n, nhand)), dim3(256), 0, stream(), n, L.grids.p, cell_grid(m),
                       L.slot_of.p, m->d.rowptr, m->d.diag, val, e->sk.F.p, e->sk.Bk.p, e->sk.U.p, e->sk.uhand.p,
                       nhand, L.HS);
    HIP_TRY(hipGetLastError());
    switch (skew_waves(L)) {
        case 1: FVH_TRY(launch_update_w<1>(e)); break;
        case 2: FVH_TRY(launch_update_w<2>(e)); break;
        default: FVH_TRY(launch_update_w<3>(e)); break;
    }
    int edge = 0;
    for (const SkewGrid &g : L.h_grids) edge = std::max(edge, g.nx + g.ny);
    hipLaunchKernelGGL(k_skew_edges, dim3((unsigned)((edge + 255) / 256), (unsigned)L.h_grids.size()), dim3(256), 0,
                       stream(), L.grids.p, e->sk.F.p, e->sk.Bk.p);
    HIP_TRY(hipGetLastError());
    FVH_TRY(sweep_records(e, e->sk.s64));
    if (!m->skew->W && !forced_waves()) FVH_TRY(skew_tune(e));
    return FVH_OK;
}

// e's records become a copy of src's (both on the same mesh's pipelined
// sweeps): F, Bk and the FP64 backward quads, which are what an update
// writes and the applies read.  e then applies src's preconditioner with
// buffers of its own (right-hand sides, results, hand-over arrays, fault
// flag), so that the two equations' solves can run at the same time.
int skew_copy(fvh_eqn *e, const fvh_eqn *src) {
    const SkewLayout &L = *e->m->skew;
    if (src->m != e->m || src->sk.F.n != L.NS || src->sk.s64.kb.n != (size_t)kBwdN * L.NS)
        return set_error(FVH_ERR_ARG, "DILU records not prepared");
    FVH_TRY(skew_alloc(e));
    FVH_TRY(sweep_alloc(e->sk.s64, L));
    HIP_TRY(hipMemcpyAsync(e->sk.F.p, src->sk.F.p, sizeof(Rec4) * L.NS, hipMemcpyDeviceToDevice, stream()));
    HIP_TRY(hipMemcpyAsync(e->sk.Bk.p, src->sk.Bk.p, sizeof(Rec4) * L.NS, hipMemcpyDeviceToDevice, stream()));
    HIP_TRY(hipMemcpyAsync(e->sk.s64.kb.p, src->sk.s64.kb.p, sizeof(double) * src->sk.s64.kb.n,
                           hipMemcpyDeviceToDevice, stream()));
    return FVH_OK;
}

template <typename T, int W>
static int launch_sweeps_w(fvh_eqn *e, SkewSweep<T> &w, const KScal *ks) {
    const SkewLayout &L = *e->m->skew;
    const SkewArgs a = skew_args<Rec4T<T>>(e, W);
    const size_t lds = (size_t)(W + 1) * L.HS * sizeof(T), lds_f = lds;
    FVH_TRY(allow_lds(k_skew_fwd<T, W>, lds_f));
    FVH_TRY(allow_lds(k_skew_bwd<T, W>, lds));
    const void *recs = sizeof(T) == 8 ? static_cast<const void *>(e->sk.F.p) : w.kf.p;
    hipLaunchKernelGGL((k_skew_fwd<T, W>), dim3(a.nblk), dim3(64 * (W + 2)), lds_f, stream(), a, recs, w.bq.p, w.yq.p,
                       w.hand.p, ks);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((k_skew_bwd<T, W>), dim3(a.nblk), dim3(64 * (W + 2)), lds, stream(), a, w.kb.p, w.yq.p, w.zp.p,
                       w.hand.p, ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// gather with the reset of the hand-over arrays (skipped when the caller has
// written x into the sweep layout and reset the arrays: x null) and the two
// sweeps: z in quads (w.zp)
template <typename T>
static int skew_sweeps_t(fvh_eqn *e, const T *x, const KScal *ks, SkewSweep<T> &w) {
    fvh_mesh *m = e->m;
    const SkewLayout &L = *m->skew;
    const int nhand = 2 * L.nblk[0] * L.HS;
    // per-case solves: gather / scatter cover every case, the sweeps skip
    // finished cases
    const KScal *ks_all = m->cases ? nullptr : ks;
    if (x) {
        const int ntiles = (int)L.tiles.n;
        hipLaunchKernelGGL((k_skew_gather<T>), dim3(ntiles + (nhand + 255) / 256), dim3(256), 0, stream(), ntiles,
                           L.tiles.p, L.grids.p, x,
                           sizeof(T) == 8 ? reinterpret_cast<Rec4T<T> *>(e->sk.F.p) : nullptr, w.bq.p, w.hand.p,
                           nhand, L.HS, ks_all);
        HIP_TRY(hipGetLastError());
    }
    switch (skew_waves(L, sizeof(T) == 4)) {
        case 1: FVH_TRY((launch_sweeps_w<T, 1>(e, w, ks))); break;
        case 2: FVH_TRY((launch_sweeps_w<T, 2>(e, w, ks))); break;
        default: FVH_TRY((launch_sweeps_w<T, 3>(e, w, ks))); break;
    }
    return FVH_OK;
}

template <typename T>
static int skew_apply_t(fvh_eqn *e, const T *x, T *y, const KScal *ks, SkewSweep<T> &w) {
    FVH_TRY(skew_sweeps_t<T>(e, x, ks, w));
    const SkewLayout &L = *e->m->skew;
    hipLaunchKernelGGL((k_skew_scatter_q<T>), dim3((unsigned)L.tiles.n), dim3(256), 0, stream(), L.tiles.p, L.grids.p,
                       w.zp.p, y, e->m->cases ? nullptr : ks);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// Waves per block, chosen per mesh and precision by timing the apply in each
// width on the freshly packed records (the sweeps only rewrite their own
// scratch fields); the width never changes a result.
template <typename T>
static int skew_tune_t(fvh_eqn *e, SkewSweep<T> &w, int &W_out) {
    SkewLayout &L = *e->m->skew;
    const int n = (int)e->m->n;
    DBuf<T> x, y;
    FVH_TRY(x.alloc(n));
    FVH_TRY(y.alloc(n));
    FVH_TRY(x.zero());
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    float best = -1.f;
    int bestW = 1;
    for (int W = 1; W <= lds_waves_max(L); ++W) {
        W_out = W;  // read by skew_waves
        FVH_TRY(skew_apply_t<T>(e, x.p, y.p, nullptr, w));  // warm
        HIP_TRY(hipEventRecord(a, stream()));
        for (int rep = 0; rep < 3; ++rep) FVH_TRY(skew_apply_t<T>(e, x.p, y.p, nullptr, w));
        HIP_TRY(hipEventRecord(b, stream()));
        HIP_TRY(hipEventSynchronize(b));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, a, b));
        if (best < 0.f || t < best) {
            best = t;
            bestW = W;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    W_out = bestW;
    (void)L;
    return FVH_OK;
}

static int skew_tune(fvh_eqn *e) {
    return skew_tune_t<double>(e, e->sk.s64, e->m->skew->W);
}

int skew_apply(fvh_eqn *e, const double *x, double *y, const KScal *ks) {
    SkewStore &s = e->sk;
    if (!s.s64.kf.p) return set_error(FVH_ERR_ARG, "DILU records not prepared");
    return skew_apply_t<double>(e, x, y, ks, s.s64);
}

bool skew_tuned32(const fvh_mesh *m) { return forced_waves() || m->skew->W32 != 0; }

int skew_to32(fvh_eqn *e) {
    SkewStore &s = e->sk;
    FVH_TRY(sweep_records(e, s.s32));
    if (!e->m->skew->W32 && !forced_waves()) FVH_TRY(skew_tune_t<float>(e, s.s32, e->m->skew->W32));
    return FVH_OK;
}

int skew_apply32(fvh_eqn *e, const float *x, float *y, const KScal *ks) {
    SkewStore &s = e->sk;
    if (!s.s32.kf.p) return set_error(FVH_ERR_ARG, "FP32 sweep records not prepared");
    return skew_apply_t<float>(e, x, y, ks, s.s32);
}

int skew_apply_quads(fvh_eqn *e, const double *x, const KScal *ks, const double **zq) {
    SkewStore &s = e->sk;
    if (!s.s64.kf.p) return set_error(FVH_ERR_ARG, "DILU records not prepared");
    FVH_TRY(skew_sweeps_t<double>(e, x, ks, s.s64));
    *zq = s.s64.zp.p;
    return FVH_OK;
}

int skew_apply_quads32(fvh_eqn *e, const float *x, const KScal *ks, const float **zq) {
    SkewStore &s = e->sk;
    if (!s.s32.kf.p) return set_error(FVH_ERR_ARG, "FP32 sweep records not prepared");
    FVH_TRY(skew_sweeps_t<float>(e, x, ks, s.s32));
    *zq = s.s32.zp.p;
    return FVH_OK;
}

const int *skew_quad_map(const fvh_mesh *m) { return m->skew ? m->skew->qoff.p : nullptr; }

int skew_hand_reset(fvh_eqn *e, int fp32) {
    const SkewLayout &L = *e->m->skew;
    const int nhand = 2 * L.nblk[0] * L.HS;
    const dim3 g((unsigned)((nhand + 255) / 256));
    if (fp32)
        hipLaunchKernelGGL((k_skew_gather<float>), g, dim3(256), 0, stream(), 0, L.tiles.p, L.grids.p,
                           (const float *)nullptr, (Rec4T<float> *)nullptr, e->sk.s32.bq.p, e->sk.s32.hand.p, nhand,
                           L.HS, (const KScal *)nullptr);
    else
        hipLaunchKernelGGL((k_skew_gather<double>), g, dim3(256), 0, stream(), 0, L.tiles.p, L.grids.p,
                           (const double *)nullptr, (Rec4T<double> *)nullptr, e->sk.s64.bq.p, e->sk.s64.hand.p, nhand,
                           L.HS, (const KScal *)nullptr);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

int skew_rhs_target(fvh_eqn *e, int fp32, SkewRhs *t) {
    const SkewLayout &L = *e->m->skew;
    t->nhand = 2 * L.nblk[0] * L.HS;
    t->HS = L.HS;
    if (fp32) {
        if (!e->sk.s32.bq.p) return set_error(FVH_ERR_ARG, "FP32 sweep records not prepared");
        t->dst = e->sk.s32.bq.p;
        t->map = L.qoff.p;
        t->stride = 1;
        t->hand = e->sk.s32.hand.p;
        t->sent = HandWord<float>::kSent;
    } else {
        if (!e->sk.F.p) return set_error(FVH_ERR_ARG, "DILU records not prepared");
        t->dst = e->sk.F.p;  // b is word 0 of each 4-word record
        t->map = L.slot_of.p;
        t->stride = 4;
        t->hand = e->sk.s64.hand.p;
        t->sent = HandWord<double>::kSent;
    }
    return FVH_OK;
}

int skew_dinv_natural(fvh_eqn *e, double *out) {
    const int n = (int)e->m->n;
    hipLaunchKernelGGL(k_skew_dinv_natural, grid_for(n), dim3(256), 0, stream(), n, e->m->skew->slot_of.p, e->sk.F.p,
                       out);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// A hand-over wait of a sweep ran out since the last check (reported once).
// The sweeps of an equation that borrows another's preconditioner record
// their faults in the owner's flag, so both are read.
int skew_check(fvh_eqn *e) {
    fvh_eqn *owners[2] = {e, e->pshare};
    for (fvh_eqn *q : owners) {
        int *f = q ? q->sk.h_fault : nullptr;
        if (f && __atomic_load_n(f, __ATOMIC_ACQUIRE)) {
            __atomic_store_n(f, 0, __ATOMIC_RELEASE);
            return set_error(FVH_ERR_HIP, "band sweep: a hand-over between waves did not arrive in time");
        }
    }
    return FVH_OK;
}

}  // namespace fvh

// Reads, and with raise != 0 sets, the pending hand-over fault of the
// equation's pipelined sweeps (pinned host memory; no device work).
extern "C" int fvh_precond_fault(fvh_eqn *e, int32_t raise, int32_t *pending) {
    if (!e) return fvh::set_error(FVH_ERR_ARG, "null equation");
    fvh_eqn *q = e->pshare ? e->pshare : e;
    int *f = q->sk.h_fault;
    if (!f) return fvh::set_error(FVH_ERR_UNSUPPORTED, "the equation's preconditioner has no pipelined sweeps");
    if (raise) __atomic_store_n(f, 1, __ATOMIC_RELEASE);
    if (pending) *pending = __atomic_load_n(f, __ATOMIC_ACQUIRE);
    return FVH_OK;
}
