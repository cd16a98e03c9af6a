// Device-side internals of libfvmhip.so: buffers, the device mesh view that
// kernels receive, the per-equation state, Krylov scalars and the canonical
// reduction order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "fvh_internal.h"
#include "tile_sched.h"

#define HIP_TRY(expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fvh::set_error(FVH_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #expr,          \
                                  hipGetErrorString(e_));                                           \
    } while (0)

#define FVH_TRY(expr)                    \
    do {                                 \
        int rc_ = (expr);                \
        if (rc_ != FVH_OK) return rc_;   \
    } while (0)

namespace fvh {

// Canonical summation order shared with the oracle: chunks of kChunk
// elements, each summed by kBlock lanes (lane t takes t, t+256, ... in
// order), then a pairwise tree (t, t+s) for s = 128 ... 1; the chunk
// partials are reduced the same way by one block.
constexpr int kBlock = 256;
constexpr int kPer = 8;
constexpr int kChunk = kBlock * kPer;

inline int nchunks(int64_t n) { return (int)((n + kChunk - 1) / kChunk); }

hipStream_t stream();
void set_thread_stream(hipStream_t s);  // this thread's launches (null: the library's stream)
int ensure_device();

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return set_error(FVH_ERR_HIP, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
        }
        n = count;
        return FVH_OK;
    }
    int zero() {
        HIP_TRY(hipMemsetAsync(p, 0, n * sizeof(T), stream()));
        return FVH_OK;
    }
    int upload(const std::vector<T> &v) {
        FVH_TRY(alloc(v.size()));
        if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return FVH_OK;
    }
    int download(std::vector<T> &v, size_t count) const {
        v.resize(count);
        HIP_TRY(hipStreamSynchronize(stream()));  // kernels of the library's (non-blocking) stream first
        if (count) HIP_TRY(hipMemcpy(v.data(), p, count * sizeof(T), hipMemcpyDeviceToHost));
        return FVH_OK;
    }
};

// What kernels see of the mesh (plain pointers, int32 indices).
struct DevMesh {
    int n, nf, nbf, nnz, ncf;
    const int *rowptr, *col, *diag, *tslot;         // CSR pattern of A
    const int *cf_ptr, *cf_face, *cf_nbr, *cf_slot;  // cell -> internal faces (faces_range order)
    const double *cf_sign, *cf_w, *cf_omw;           // nsign, divergence weights of that side
    int nif;                                         // internal faces (faces nbf .. nf-1)
    const int2 *fi_slot;                             // internal face: slots of A[o, n] and A[n, o] (~slot: nsign -1)
    const double2 *fi_w;                             // internal face: divergence weights of the o and n sides
    const int *own1, *own2;                          // face owners (boundary: own1 == own2)
    const double *fw, *farea, *fdelta;               // face weight, area, delta
    const double *Sfx, *Sfy, *Sfz;                   // area * normal
    const double *nx, *ny, *nz;                      // unit normals
    const double *ex, *ey, *ez;                      // unit vectors along the centre line (face e)
    const double *vol;                               // cell volumes
    const double *ccx, *ccy, *ccz, *fcx, *fcy, *fcz;  // centres
    const int *cb_ptr, *cb_face;                     // cell -> boundary faces, ascending face id
    const int *bf_face, *bf_cell, *bf_patch;         // boundary faces in boundary (patch) order
};

// Per-cell boundary-condition entries of one BC tuple, in the order
// apply_boundary_conditions! visits them (tuple order, then patch order).
struct BcPlan {
    DBuf<int> ptr, face, bc;
};

// Anti-diagonal geometry of a banded grid, by value in kernel arguments.
struct BandGeom {
    int nx, ny, nT;
    const int *offj;  // packed position of (T - t, t) is offj[T] + t
};

// Per-row records in packed (anti-diagonal) order.
template <typename T>
struct alignas(4 * sizeof(T)) Rec4T {
    T a, b, c, d;
};
using Rec4 = Rec4T<double>;
using Rec4f = Rec4T<float>;
// ud, us, uw: A_rr, A_rs*A_sr and A_rw*A_wr for the diagonal recurrence
// (the products do not depend on D: formed once, in the recurrence's order);
// f: (b_r, Dinv_r, A_rs, A_rw) for the forward sweep;
// b: (y_r, Dinv_r, A_re, A_rn) for the backward sweep;
// f32 / b32: FP32 copies of f / b for the mixed-precision inner solves.
struct BandRecs {
    double *ud, *us, *uw;
    Rec4 *f, *b;
    Rec4f *f32, *b32;
};

// A mesh that is the block-diagonal union of independent cases (cells of
// case k are [off[k], off[k+1]), no face joins two cases).  Reductions run
// per case over the case's own 2048-row chunks, so each case's sums are the
// ones it would have alone; block b of a reduction launch covers chunk
// (blk_case[b], (blk_row0[b] - off[case]) / 2048).
struct CaseLayout {
    int K = 0;
    int nblk = 0;                        // reduction blocks (chunks) over all cases
    std::vector<int> h_off, h_cb;        // cell offsets [K+1], first block of each case [K+1]
    DBuf<int> off, cb, cell_case, blk_case, blk_row0;
    // per-case banded sweeps (all cases full rectangles, i-fastest)
    bool band = false;
    int band_R = 0, band_maxny = 0;  // rows per lane (0: not tuned yet), tallest case
    DBuf<BandGeom> geoms;                // offj of case k points into offj_all
    DBuf<int> offj_all, b_iperm;         // union numbering
    DBuf<int> nxs;
};

// Skewed (fixed-stride anti-diagonal) layout of the pipelined band sweeps.
// Grid g (a whole mesh, or one case of a batch) stores the record of cell
// (i, j) at slot base + (i + j + kSkewD0) * S + j: every anti-diagonal
// T = i + j owns S consecutive slots, one per grid row (S a multiple of 64
// above ny), and the slots of rows j >= ny, of cells with i < 0 or i >= nx
// and of the kSkewD0 diagonals below T = 0 and the prefetch margin above the
// last diagonal are padding.  One compute wave owns 64 rows; a block runs W
// consecutive waves (W is a launch parameter, tuned per mesh).  The FP32
// sweeps keep their results in quads of four consecutive diagonals per row
// (one 16-byte store per four steps): the value of (T, j) at base +
// ((T + kSkewQ0) & ~3) * S + 4 j + ((T + kSkewQ0) & 3), inside the grid's
// slots for every T the sweeps touch.
constexpr int kLevRing = 16384;    // results the level walk keeps in LDS (by schedule position)
constexpr int kLevLate = 1 << 30;  // flag of a neighbour read after the level barrier
constexpr int kTileRows = 1024;    // rows per tile of the tile walk, unless FVH_TILE_ROWS says otherwise
constexpr int kTileLag = 8;        // steps a neighbour from another tile is scheduled behind its producer
constexpr int kSkewL = 8;                    // steps per hand-over batch
constexpr int kSkewPD = 32;                  // deepest record prefetch (steps)
constexpr int kSkewD0 = kSkewPD + 2;         // padding diagonals below T = 0
constexpr int kSkewTop = kSkewPD + 4;        // padding diagonals above the last step (+ the quads' margin)
constexpr int kSkewQ0 = kSkewD0 + 2;         // quad origin: T + kSkewQ0 is a multiple of 4 when T is
constexpr int kSkewMaxW = 3;                 // compute waves per block (plus one courier wave)
struct SkewGrid {
    int nx, ny;     // cell (i, j) is natural row o + i + j*nx
    int o;
    int nw;         // compute waves (S / 64)
    int NG;         // step groups (multiple of 4): steps 0 .. NG*kSkewL-1 cover the nx+ny-1 diagonals
    int S;          // slots per diagonal
    unsigned base;  // first slot of the grid
    int blk0[kSkewMaxW];  // first block of the grid, W = 1 .. kSkewMaxW waves per block
};
struct SkewTile {
    int grid, i0, j0;
};
struct SkewLayout {
    int W = 0, W32 = 0;        // waves per block of the FP64 / FP32 sweeps (0: not tuned yet)
    int HS = 0;                // hand-over array length (values) per block boundary
    size_t NS = 0;             // slots over all grids
    int nblk[kSkewMaxW] = {};  // blocks over all grids, per W
    std::vector<SkewGrid> h_grids;
    DBuf<SkewGrid> grids;
    DBuf<int> blk_grid[kSkewMaxW];  // block -> grid, per W
    DBuf<int> slot_of;              // natural row -> slot
    DBuf<int> qoff;                 // natural row -> word of its value in the sweeps' quads
    DBuf<SkewTile> tiles;           // 64 x 64 tiles of the grids (gather / scatter)
};

// Device view of a CaseLayout, by value in kernel arguments.
struct CaseView {
    int K, nblk;
    const int *off, *cb, *cell_case, *blk_case, *blk_row0;
};

// Sweep schedule of the lower (forward) and upper (backward) triangles.
struct Schedule {
    int nlev = 0;
    int max_deps = 0;  // most lower (or upper) neighbours of a row
    DBuf<int> fptr, frows, bptr, brows;
    std::vector<int> h_fptr, h_bptr;
    // Per matrix entry (lower entries: forward schedule, upper: backward),
    // where the level walk finds the neighbour's result, by the neighbour's
    // position p in the walk's order: -(1 + k) for word k = p mod kLevRing of
    // the LDS ring of results (still held when the row's level ends), p for
    // a result two or more levels back (prefetched a level ahead), p |
    // kLevLate for one a level back that the ring no longer holds (read after
    // the level barrier; only next to levels wider than the ring).  f2b[k]:
    // the forward position of the row at backward position k.
    DBuf<int> lcode, f2b;
    DBuf<int> lpos;  // per entry: the neighbour's position in the walk's order (the sync-free walk)
    // The sync-free walk's waves, forward and backward: (first position,
    // count <= 64) per wave, each wave inside one level, waves in level order
    DBuf<int> swf, swb;
    int nswf = 0, nswb = 0;
    // The tile walk, forward (tf) and backward (tb): tiles of tile_rows
    // consecutive rows in the sweep's order, each walked by one wave in steps
    // of independent rows; a tile's results stay in its LDS, a neighbour from
    // another tile is staged into that LDS by a loader wave before the
    // chunk of steps that first reads it.
    struct TileWalk {
        DBuf<int> rows;    // position -> row; tile t holds positions [t * tile_rows, ...)
        DBuf<int> tiles;   // per tile: first chunk, chunk count
        DBuf<int> chunks;  // per chunk (kTileChunkInts): first step, steps, first entry, entries, first position, rows
        DBuf<int> steps;   // per step: first position, rows (<= 64, no dependency among them)
        DBuf<int> ents;    // per staged neighbour: its position, the LDS word it goes to
        int ntiles = 0, maxrem = 0;  // tiles; most neighbours one tile stages
    };
    TileWalk tf, tb;
    int tile_rows = kTileRows;
    // per entry: the LDS word where the tile walk finds the neighbour (its
    // position in the tile, or tile_rows + its staging slot); the forward
    // position of the row at backward position k
    DBuf<int> tcode, tf2b;
    int ring = kLevRing;  // results the level walk's LDS ring holds (power of two <= kLevRing)
    int band_nx = 0;  // >0: A is a banded grid (deps only r-1, r-nx)
    DBuf<unsigned char> bmask;  // banded grid: bit k of row r = entry r + (-nx, -1, 0, 1, nx)[k] present
    int band_R = 0;   // grid rows per lane of the sweeps (0: not tuned yet)
    BandGeom g{};
    DBuf<int> b_offj, b_iperm;
};

}  // namespace fvh

struct fvh_mesh {
    fvh::DevMesh d{};
    int64_t n = 0, nf = 0, nbf = 0, nnz = 0;
    int nboundaries = 0;
    // owned device storage
    fvh::DBuf<int> rowptr, col, diag, tslot, cf_ptr, cf_face, cf_nbr, cf_slot, own1, own2, cb_ptr, cb_face,
        bf_face, bf_cell, bf_patch, fi_slot;
    fvh::DBuf<double> fi_w, ex, ey, ez;
    bool cf_ascending = true;  // every cell's internal faces in ascending face id
    // per-slot diagonal terms handed from the face pass of the assembly to
    // its cell pass ([slot][face term]); sized by the largest term count seen
    fvh::DBuf<double> asm_terms;
    fvh::DBuf<double> cf_sign, cf_w, cf_omw, fw, farea, fdelta, Sfx, Sfy, Sfz, nx, ny, nz, vol, ccx, ccy, ccz, fcx,
        fcy, fcz;
    // host copies used by set-up code
    std::vector<int> h_rowptr, h_col, h_diag, h_bf_face, h_bf_cell, h_bf_patch;
    std::vector<int> h_bstart;  // boundary b occupies bf[h_bstart[b], h_bstart[b+1])
    std::vector<double> h_fdelta, h_vol, h_nx, h_ny, h_nz;
    fvh::Schedule sched;
    std::map<std::vector<int>, std::unique_ptr<fvh::BcPlan>> bc_plans;
    // ILU(0) update lists: for each lower entry (i,k), pairs (slot(i,j), slot(k,j)), j > k
    bool ilu_ready = false;
    fvh::DBuf<int> ilu_uptr, ilu_pairs;
    // independent cases (batched mode); null for a single case
    std::unique_ptr<fvh::CaseLayout> cases;
    // pipelined band sweeps (null: the mesh is not banded, or too large for them)
    std::unique_ptr<fvh::SkewLayout> skew;
    fvh::CaseView case_view() const {
        const fvh::CaseLayout &c = *cases;
        return fvh::CaseView{c.K, c.nblk, c.off.p, c.cb.p, c.cell_case.p, c.blk_case.p, c.blk_row0.p};
    }
};

namespace fvh {

// Krylov scalars, resident on the device so that an iteration never waits
// for the host.
struct KScal {
    double rho, alpha, omega, beta, rnorm, eps, gamma, pnorm2, rnorm0, ts, tt, sigma;
    double hbis, zeta;
    double gprev;  // CG: gamma at the start of the iteration (read by the fused beta stage)
    int done, iter, status, k;
};

int bc_plan(fvh_mesh *m, const fvh_bc *bcs, int nbc, BcPlan **out);

}  // namespace fvh
