// Library core: error channel, device/stream management and the device mesh
// (CSR pattern, face -> CSR slot maps, boundary tables, sweep schedules).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>

#include "fvh_device.h"
#include "k_band.h"
#include "tile_sched.h"

namespace fvh {

static thread_local std::string g_last_error;
static hipStream_t g_stream = nullptr;
static int g_device_ok = -1;

std::string last_error_text() { return g_last_error; }

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

// The library works on the current HIP device with one non-blocking
// default-priority stream.
int ensure_device() {
    if (g_device_ok == 1) return FVH_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        g_device_ok = 0;
        return set_error(FVH_ERR_NODEV, "no HIP device available");
    }
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, 0));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(FVH_ERR_NODEV, "device is %s; this library is built for gfx950", prop.gcnArchName);
    HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    g_device_ok = 1;
    return FVH_OK;
}

// A thread may route its launches to a stream of its own (the two momentum
// solves of a SIMPLE iteration run at once, one of them on a second thread);
// every other thread uses the library's stream.
static thread_local hipStream_t t_stream = nullptr;
hipStream_t stream() { return t_stream ? t_stream : g_stream; }
void set_thread_stream(hipStream_t s) { t_stream = s; }

// Per-cell boundary-condition lists for one BC tuple, cached on the mesh.
int bc_plan(fvh_mesh *m, const fvh_bc *bcs, int nbc, BcPlan **out) {
    std::vector<int> key(nbc);
    for (int i = 0; i < nbc; ++i) {
        if (bcs[i].patch < 0 || bcs[i].patch >= m->nboundaries)
            return set_error(FVH_ERR_ARG, "boundary condition %d names patch %d of %d", i, bcs[i].patch, m->nboundaries);
        key[i] = bcs[i].patch;
    }
    auto it = m->bc_plans.find(key);
    if (it != m->bc_plans.end()) {
        *out = it->second.get();
        return FVH_OK;
    }
    const int n = (int)m->n;
    std::vector<std::vector<std::pair<int, int>>> per(n);
    for (int k = 0; k < nbc; ++k) {
        int b = bcs[k].patch;
        for (int i = m->h_bstart[b]; i < m->h_bstart[b + 1]; ++i) per[m->h_bf_cell[i]].push_back({m->h_bf_face[i], k});
    }
    std::vector<int> ptr(n + 1, 0), face, bc;
    for (int c = 0; c < n; ++c) {
        for (auto &e : per[c]) {
            face.push_back(e.first);
            bc.push_back(e.second);
        }
        ptr[c + 1] = (int)face.size();
    }
    auto plan = std::make_unique<BcPlan>();
    FVH_TRY(plan->ptr.upload(ptr));
    FVH_TRY(plan->face.upload(face));
    FVH_TRY(plan->bc.upload(bc));
    *out = plan.get();
    m->bc_plans.emplace(key, std::move(plan));
    return FVH_OK;
}

}  // namespace fvh

using namespace fvh;

extern "C" const char *fvh_version(void) { return "fvmhip 0.1.0 (gfx950)"; }

extern "C" const char *fvh_status_string(int s) {
    switch (s) {
        case FVH_OK: return "ok";
        case FVH_ERR_ARG: return "bad argument";
        case FVH_ERR_HIP: return "HIP runtime error";
        case FVH_ERR_NODEV: return "no usable gfx950 device";
        case FVH_ERR_MESH: return "inconsistent mesh";
        case FVH_ERR_UNSUPPORTED: return "unsupported term/boundary combination";
        case FVH_ERR_IO: return "I/O error";
        case FVH_ERR_SIZE: return "size beyond maximum";
        default: return "unknown status";
    }
}

extern "C" int fvh_last_error(char *buf, int64_t len) {
    if (!buf || len <= 0) return FVH_ERR_ARG;
    std::snprintf(buf, (size_t)len, "%s", g_last_error.c_str());
    return FVH_OK;
}

extern "C" int fvh_device_count(int32_t *n) {
    if (!n) return set_error(FVH_ERR_ARG, "null");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return FVH_OK;
}

extern "C" int fvh_malloc(void **ptr, int64_t bytes) {
    if (!ptr || bytes < 0) return set_error(FVH_ERR_ARG, "fvh_malloc: bad arguments");
    FVH_TRY(ensure_device());
    HIP_TRY(hipMalloc(ptr, bytes > 0 ? (size_t)bytes : 1));
    return FVH_OK;
}

extern "C" int fvh_free(void *ptr) {
    if (ptr) HIP_TRY(hipFree(ptr));
    return FVH_OK;
}

extern "C" int fvh_memcpy(void *dst, const void *src, int64_t bytes) {
    if ((!dst || !src) && bytes > 0) return set_error(FVH_ERR_ARG, "fvh_memcpy: null pointer");
    FVH_TRY(ensure_device());
    if (bytes > 0) {
        HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDefault, stream()));
        HIP_TRY(hipStreamSynchronize(stream()));
    }
    return FVH_OK;
}

extern "C" int fvh_synchronize(void) {
    FVH_TRY(ensure_device());
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

namespace {

struct V3 {
    double x, y, z;
};
inline double vnorm(V3 a) { return std::sqrt((a.x * a.x + a.y * a.y) + a.z * a.z); }

// One direction of the tile walk on the device (the schedule: tile_sched.h).
int tile_walk(int n, const std::vector<int> &rp, const std::vector<int> &col, const std::vector<int> &dg, bool upper,
              int tr, int lag, Schedule::TileWalk &tw, std::vector<int> &pos, std::vector<int> &code) {
    TileSchedule ts;
    tile_schedule(n, rp, col, dg, upper, tr, lag, ts, code);
    tw.ntiles = ts.ntiles;
    tw.maxrem = ts.maxrem;
    FVH_TRY(tw.rows.upload(ts.rows));
    FVH_TRY(tw.tiles.upload(ts.tiles));
    FVH_TRY(tw.chunks.upload(ts.chunks));
    FVH_TRY(tw.steps.upload(ts.steps));
    FVH_TRY(tw.ents.upload(ts.ents));
    pos.swap(ts.pos);
    return FVH_OK;
}

int build_schedule(fvh_mesh *m) {
    const int n = (int)m->n;
    const auto &rp = m->h_rowptr, &col = m->h_col, &dg = m->h_diag;
    std::vector<int> lf(n, 0), lb(n, 0);
    int nlev = 0;
    for (int i = 0; i < n; ++i) {
        int l = 0;
        for (int p = rp[i]; p < dg[i]; ++p) l = std::max(l, lf[col[p]] + 1);
        lf[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    int nlevb = 0;
    for (int i = n - 1; i >= 0; --i) {
        int l = 0;
        for (int p = dg[i] + 1; p < rp[i + 1]; ++p) l = std::max(l, lb[col[p]] + 1);
        lb[i] = l;
        nlevb = std::max(nlevb, l + 1);
    }
    auto group = [n](const std::vector<int> &lev, int nl, std::vector<int> &ptr, std::vector<int> &rows, bool desc) {
        ptr.assign(nl + 1, 0);
        for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
        for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
        rows.assign(n, 0);
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        if (!desc)
            for (int i = 0; i < n; ++i) rows[pos[lev[i]]++] = i;
        else
            for (int i = n - 1; i >= 0; --i) rows[pos[lev[i]]++] = i;
    };
    std::vector<int> frows, brows;
    group(lf, nlev, m->sched.h_fptr, frows, false);
    group(lb, nlevb, m->sched.h_bptr, brows, true);
    // a structurally symmetric pattern: the longest dependency chain is the same both ways
    if (nlevb != nlev) return set_error(FVH_ERR_ARG, "matrix pattern is not structurally symmetric");
    m->sched.nlev = nlev;
    // the level walk's LDS ring: kLevRing results, or fewer (a power of two)
    // with FVH_LEVEL_RING, which sends more neighbours through the global
    // and the after-barrier paths (a test knob; the results are the same)
    int ring = kLevRing;
    if (const char *env = std::getenv("FVH_LEVEL_RING")) {
        const int r = std::atoi(env);
        if (r >= 64 && r <= kLevRing && (r & (r - 1)) == 0) ring = r;
    }
    m->sched.ring = ring;
    {  // where the level walk finds each neighbour's result (positions in the walk's order)
        std::vector<int> pos(n), posf(n), code(rp[n]), f2b(n), lpos(rp[n]);
        auto codes = [&](const std::vector<int> &lev, const std::vector<int> &ptr, const std::vector<int> &rows,
                         bool upper) {
            for (int k = 0; k < n; ++k) pos[rows[k]] = k;
            for (int i = 0; i < n; ++i) {
                // the ring holds position q when i's level is walked if no later
                // position up to the end of that level shares its word
                // (q >= end - ring), and no earlier position of q's own level
                // does (their lanes write the word at the same time when a level
                // is wider than the ring)
                const int end = ptr[lev[i] + 1];
                for (int p = upper ? dg[i] + 1 : rp[i]; p < (upper ? rp[i + 1] : dg[i]); ++p) {
                    const int j = col[p];
                    lpos[p] = pos[j];
                    if (pos[j] >= end - ring && pos[j] - ptr[lev[j]] < ring) code[p] = -(1 + (pos[j] & (ring - 1)));
                    else if (lev[i] - lev[j] >= 2) code[p] = pos[j];
                    else code[p] = pos[j] | kLevLate;
                }
            }
        };
        codes(lf, m->sched.h_fptr, frows, false);
        posf = pos;
        codes(lb, m->sched.h_bptr, brows, true);
        for (int i = 0; i < n; ++i) code[dg[i]] = lpos[dg[i]] = i;
        for (int k = 0; k < n; ++k) f2b[k] = posf[brows[k]];
        FVH_TRY(m->sched.lcode.upload(code));
        FVH_TRY(m->sched.lpos.upload(lpos));
        FVH_TRY(m->sched.f2b.upload(f2b));
    }
    {  // the sync-free walk's waves: no wave holds two rows of different levels
        auto waves = [](const std::vector<int> &ptr, std::vector<int> &tab) {
            tab.clear();
            for (size_t l = 0; l + 1 < ptr.size(); ++l)
                for (int q = ptr[l]; q < ptr[l + 1]; q += 64) {
                    tab.push_back(q);
                    tab.push_back(std::min(64, ptr[l + 1] - q));
                }
            return (int)(tab.size() / 2);
        };
        std::vector<int> tab;
        m->sched.nswf = waves(m->sched.h_fptr, tab);
        FVH_TRY(m->sched.swf.upload(tab));
        m->sched.nswb = waves(m->sched.h_bptr, tab);
        FVH_TRY(m->sched.swb.upload(tab));
    }
    {  // the tile walk (FVH_TILE_ROWS: rows per tile, a power of two 64 .. 8192; FVH_TILE_LAG: steps)
        int tr = kTileRows, lag = kTileLag;
        if (const char *env = std::getenv("FVH_TILE_ROWS")) {
            const int r = std::atoi(env);
            if (r >= 64 && r <= 8192 && (r & (r - 1)) == 0) tr = r;
        }
        if (const char *env = std::getenv("FVH_TILE_LAG")) {
            const int r = std::atoi(env);
            if (r >= 1 && r <= 1000) lag = r;
        }
        Schedule &s = m->sched;
        s.tile_rows = tr;
        std::vector<int> code(rp[n], 0), posf, posb, f2b(n);
        FVH_TRY(tile_walk(n, rp, col, dg, false, tr, lag, s.tf, posf, code));
        FVH_TRY(tile_walk(n, rp, col, dg, true, tr, lag, s.tb, posb, code));
        std::vector<int> brows(n);
        for (int i = 0; i < n; ++i) brows[posb[i]] = i;
        for (int k = 0; k < n; ++k) f2b[k] = posf[brows[k]];
        FVH_TRY(s.tcode.upload(code));
        FVH_TRY(s.tf2b.upload(f2b));
    }
    int deg = 0;  // most lower / upper neighbours of a row
    for (int i = 0; i < n; ++i) deg = std::max(deg, std::max(dg[i] - rp[i], rp[i + 1] - dg[i] - 1));
    m->sched.max_deps = deg;
    FVH_TRY(m->sched.fptr.upload(m->sched.h_fptr));
    FVH_TRY(m->sched.frows.upload(frows));
    FVH_TRY(m->sched.bptr.upload(m->sched.h_bptr));
    FVH_TRY(m->sched.brows.upload(brows));
    // banded grid: every lower neighbour of row r is r-1 (r % nx != 0) or r-nx
    int nx = 0;
    bool band = n > 1;
    for (int i = 0; i < n && band; ++i)
        for (int p = rp[i]; p < dg[i]; ++p) {
            int d = i - col[p];
            if (d == 1) continue;
            if (nx == 0) nx = d;
            if (d != nx) { band = false; break; }
        }
    const char *force = std::getenv("FVH_SWEEP");
    if (force && std::strcmp(force, "levels") == 0) band = false;
    if (band && nx > 1 && n % nx == 0) {
        for (int i = 0; i < n && band; ++i)
            for (int p = rp[i]; p < dg[i]; ++p)
                if (i - col[p] == 1 && i % nx == 0) { band = false; break; }
        m->sched.band_nx = band ? nx : 0;
    } else {
        m->sched.band_nx = 0;
    }
    if (m->sched.band_nx > 0) {  // which of r-nx, r-1, r, r+1, r+nx each row holds (the banded SpMV)
        const int bnx = m->sched.band_nx;
        std::vector<unsigned char> mask(n, 0);
        for (int i = 0; i < n; ++i)
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                const int d = col[p] - i;
                const int k = d == -bnx ? 0 : d == -1 ? 1 : d == 0 ? 2 : d == 1 ? 3 : d == bnx ? 4 : -1;
                if (k < 0) return set_error(FVH_ERR_ARG, "banded pattern with an entry off the band");
                mask[i] |= (unsigned char)(1u << k);
            }
        FVH_TRY(m->sched.bmask.upload(mask));
    }
    return FVH_OK;
}

}  // namespace

extern "C" int fvh_mesh_create(const fvh_mesh_desc *d, fvh_mesh **out) {
    FVH_GUARD_BEGIN
    if (!d || !out) return set_error(FVH_ERR_ARG, "fvh_mesh_create: null argument");
    FVH_TRY(ensure_device());
    if (d->ncells <= 0 || d->nfaces <= 0 || d->ncells >= (1LL << 30) || d->nfaces >= (1LL << 30))
        return set_error(FVH_ERR_SIZE, "mesh sizes out of range");
    const int64_t b = d->index_base;
    const int n = (int)d->ncells, nf = (int)d->nfaces, ncf = (int)d->ncellfaces;
    auto m = std::make_unique<fvh_mesh>();
    m->n = n;
    m->nf = nf;
    m->nboundaries = (int)d->nboundaries;
    // cell -> faces
    std::vector<int> cf_ptr(n + 1, 0), cf_face(ncf), cf_nbr(ncf);
    std::vector<double> cf_sign(ncf);
    for (int c = 0; c < n; ++c) {
        int64_t s = d->cell_faces_range[2 * c] - b, e = d->cell_faces_range[2 * c + 1] - b + 1;
        if (s != cf_ptr[c] || e < s || e > ncf) return set_error(FVH_ERR_MESH, "faces_range of cell %d is not contiguous", c);
        cf_ptr[c + 1] = (int)e;
    }
    if (cf_ptr[n] != ncf) return set_error(FVH_ERR_MESH, "faces_range does not cover cell_faces");
    for (int i = 0; i < ncf; ++i) {
        cf_face[i] = (int)(d->cell_faces[i] - b);
        cf_nbr[i] = (int)(d->cell_neighbours[i] - b);
        cf_sign[i] = (double)d->cell_nsign[i];
        if (cf_face[i] < 0 || cf_face[i] >= nf || cf_nbr[i] < 0 || cf_nbr[i] >= n)
            return set_error(FVH_ERR_MESH, "cell_faces entry %d out of range", i);
    }
    std::vector<int> own1(nf), own2(nf);
    for (int f = 0; f < nf; ++f) {
        own1[f] = (int)(d->face_owners[2 * f] - b);
        own2[f] = (int)(d->face_owners[2 * f + 1] - b);
        if (own1[f] < 0 || own1[f] >= n || own2[f] < 0 || own2[f] >= n) return set_error(FVH_ERR_MESH, "face %d owners out of range", f);
    }
    // boundaries
    std::vector<int> bf_face, bf_cell, bf_patch, bstart(d->nboundaries + 1, 0);
    for (int64_t g = 0; g < d->nboundaries; ++g) {
        int64_t s = d->boundary_range[2 * g] - b, e = d->boundary_range[2 * g + 1] - b;
        for (int64_t i = s; i <= e; ++i) {
            bf_face.push_back((int)(d->boundary_faces[i] - b));
            bf_cell.push_back((int)(d->boundary_cells[i] - b));
            bf_patch.push_back((int)g);
        }
        bstart[g + 1] = (int)bf_face.size();
    }
    const int nbf = (int)bf_face.size();
    m->nbf = nbf;
    // the reference loops `for f in 1:nbfaces` over boundary faces and
    // `nbfaces+1:end` over internal ones; the mesh must be laid out so
    for (int f = 0; f < nf; ++f)
        if ((f < nbf) != (own1[f] == own2[f]))
            return set_error(FVH_ERR_MESH, "face %d: boundary faces must be the first %d faces", f, nbf);
    for (int i = 0; i < nbf; ++i)
        if (bf_face[i] < 0 || bf_face[i] >= nbf || own1[bf_face[i]] != bf_cell[i])
            return set_error(FVH_ERR_MESH, "boundary entry %d inconsistent", i);
    // CSR pattern (transpose of the reference CSC; the pattern is symmetric)
    std::vector<int> rowptr(n + 1, 0), col, diag(n);
    col.reserve(n + ncf);
    std::vector<int> tmp;
    for (int c = 0; c < n; ++c) {
        tmp.assign(1, c);
        for (int i = cf_ptr[c]; i < cf_ptr[c + 1]; ++i) tmp.push_back(cf_nbr[i]);
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        for (int j : tmp) {
            if (j == c) diag[c] = (int)col.size();
            col.push_back(j);
        }
        rowptr[c + 1] = (int)col.size();
    }
    const int nnz = (int)col.size();
    m->nnz = nnz;
    auto find = [&](int r, int c) -> int {
        auto bb = col.begin() + rowptr[r], ee = col.begin() + rowptr[r + 1];
        auto it = std::lower_bound(bb, ee, c);
        return (it != ee && *it == c) ? (int)(it - col.begin()) : -1;
    };
    std::vector<int> tslot(nnz);
    for (int r = 0; r < n; ++r)
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            tslot[p] = find(col[p], r);
            if (tslot[p] < 0) return set_error(FVH_ERR_MESH, "connectivity is not symmetric at (%d,%d)", r, col[p]);
        }
    std::vector<int> cf_slot(ncf);
    std::vector<char> used(nnz, 0);
    std::vector<double> cf_w(ncf), cf_omw(ncf);
    for (int c = 0; c < n; ++c)
        for (int i = cf_ptr[c]; i < cf_ptr[c + 1]; ++i) {
            int s = find(c, cf_nbr[i]);
            if (used[s]) return set_error(FVH_ERR_MESH, "cells %d and %d share more than one face", c, cf_nbr[i]);
            used[s] = 1;
            cf_slot[i] = s;
            // Divergence{Linear} weight of this side (Discretise_1_schemes.jl:64)
            int f = cf_face[i], nb = cf_nbr[i];
            V3 xf{d->face_centre[3 * f], d->face_centre[3 * f + 1], d->face_centre[3 * f + 2]};
            V3 xC{d->cell_centre[3 * c], d->cell_centre[3 * c + 1], d->cell_centre[3 * c + 2]};
            V3 xN{d->cell_centre[3 * nb], d->cell_centre[3 * nb + 1], d->cell_centre[3 * nb + 2]};
            double w = vnorm({xf.x - xC.x, xf.y - xC.y, xf.z - xC.z}) / vnorm({xN.x - xC.x, xN.y - xC.y, xN.z - xC.z});
            cf_w[i] = w;
            cf_omw[i] = 1.0 - w;
        }
    // internal faces seen from the face: the slots A[o, n] and A[n, o] and
    // the two sides' divergence weights, for the face-parallel assembly pass
    const int nif = nf - nbf;
    constexpr int kNoSide = INT32_MIN;
    std::vector<int> fi_slot(2 * (size_t)std::max(nif, 1), kNoSide);
    std::vector<double> fi_w(2 * (size_t)std::max(nif, 1), 0.0);
    for (int c = 0; c < n; ++c)
        for (int i = cf_ptr[c]; i < cf_ptr[c + 1]; ++i) {
            const int f = cf_face[i];
            if (f < nbf) return set_error(FVH_ERR_MESH, "cell %d lists boundary face %d among its internal faces", c, f);
            const int side = (own1[f] == c) ? 0 : (own2[f] == c ? 1 : -1);
            if (side < 0) return set_error(FVH_ERR_MESH, "cell %d lists face %d of which it is no owner", c, f);
            const size_t k = 2 * (size_t)(f - nbf) + side;
            if (fi_slot[k] != kNoSide) return set_error(FVH_ERR_MESH, "face %d listed twice by cell %d", f, c);
            // the side's nsign rides in the slot's sign bit: ~slot when -1
            fi_slot[k] = cf_sign[i] > 0.0 ? cf_slot[i] : ~cf_slot[i];
            fi_w[k] = cf_w[i];
        }
    for (int k = 0; k < 2 * nif; ++k)
        if (fi_slot[k] == kNoSide) return set_error(FVH_ERR_MESH, "internal face %d is missing from a cell's face list", nbf + k / 2);
    // cell -> boundary faces in ascending face id
    std::vector<int> cb_ptr(n + 1, 0), cb_face;
    for (int f = 0; f < nbf; ++f) cb_ptr[own1[f] + 1]++;
    for (int c = 0; c < n; ++c) cb_ptr[c + 1] += cb_ptr[c];
    cb_face.resize(nbf);
    {
        std::vector<int> pos(cb_ptr.begin(), cb_ptr.end() - 1);
        for (int f = 0; f < nbf; ++f) cb_face[pos[own1[f]]++] = f;
    }
    // face geometry
    std::vector<double> fw(nf), farea(nf), fdelta(nf), Sfx(nf), Sfy(nf), Sfz(nf), nxv(nf), nyv(nf), nzv(nf), fcx(nf), fcy(nf), fcz(nf);
    for (int f = 0; f < nf; ++f) {
        fw[f] = d->face_weight[f];
        farea[f] = d->face_area[f];
        fdelta[f] = d->face_delta[f];
        nxv[f] = d->face_normal[3 * f];
        nyv[f] = d->face_normal[3 * f + 1];
        nzv[f] = d->face_normal[3 * f + 2];
        Sfx[f] = farea[f] * nxv[f];
        Sfy[f] = farea[f] * nyv[f];
        Sfz[f] = farea[f] * nzv[f];
        fcx[f] = d->face_centre[3 * f];
        fcy[f] = d->face_centre[3 * f + 1];
        fcz[f] = d->face_centre[3 * f + 2];
    }
    std::vector<double> exv(nf), eyv(nf), ezv(nf);
    for (int f = 0; f < nf; ++f) {
        exv[f] = d->face_e[3 * f];
        eyv[f] = d->face_e[3 * f + 1];
        ezv[f] = d->face_e[3 * f + 2];
    }
    for (int c = 0; c < n && m->cf_ascending; ++c)
        for (int i = cf_ptr[c] + 1; i < cf_ptr[c + 1]; ++i)
            if (cf_face[i] <= cf_face[i - 1]) {
                m->cf_ascending = false;
                break;
            }
    std::vector<double> vol(n), ccx(n), ccy(n), ccz(n);
    for (int c = 0; c < n; ++c) {
        vol[c] = d->cell_volume[c];
        ccx[c] = d->cell_centre[3 * c];
        ccy[c] = d->cell_centre[3 * c + 1];
        ccz[c] = d->cell_centre[3 * c + 2];
    }
    FVH_TRY(m->rowptr.upload(rowptr));
    FVH_TRY(m->col.upload(col));
    FVH_TRY(m->diag.upload(diag));
    FVH_TRY(m->tslot.upload(tslot));
    FVH_TRY(m->cf_ptr.upload(cf_ptr));
    FVH_TRY(m->cf_face.upload(cf_face));
    FVH_TRY(m->cf_nbr.upload(cf_nbr));
    FVH_TRY(m->cf_slot.upload(cf_slot));
    FVH_TRY(m->cf_sign.upload(cf_sign));
    FVH_TRY(m->cf_w.upload(cf_w));
    FVH_TRY(m->cf_omw.upload(cf_omw));
    FVH_TRY(m->fi_slot.upload(fi_slot));
    FVH_TRY(m->fi_w.upload(fi_w));
    FVH_TRY(m->own1.upload(own1));
    FVH_TRY(m->own2.upload(own2));
    FVH_TRY(m->cb_ptr.upload(cb_ptr));
    FVH_TRY(m->cb_face.upload(cb_face));
    FVH_TRY(m->bf_face.upload(bf_face));
    FVH_TRY(m->bf_cell.upload(bf_cell));
    FVH_TRY(m->bf_patch.upload(bf_patch));
    FVH_TRY(m->fw.upload(fw));
    FVH_TRY(m->farea.upload(farea));
    FVH_TRY(m->fdelta.upload(fdelta));
    FVH_TRY(m->Sfx.upload(Sfx));
    FVH_TRY(m->Sfy.upload(Sfy));
    FVH_TRY(m->Sfz.upload(Sfz));
    FVH_TRY(m->nx.upload(nxv));
    FVH_TRY(m->ny.upload(nyv));
    FVH_TRY(m->nz.upload(nzv));
    FVH_TRY(m->ex.upload(exv));
    FVH_TRY(m->ey.upload(eyv));
    FVH_TRY(m->ez.upload(ezv));
    FVH_TRY(m->vol.upload(vol));
    FVH_TRY(m->ccx.upload(ccx));
    FVH_TRY(m->ccy.upload(ccy));
    FVH_TRY(m->ccz.upload(ccz));
    FVH_TRY(m->fcx.upload(fcx));
    FVH_TRY(m->fcy.upload(fcy));
    FVH_TRY(m->fcz.upload(fcz));
    m->h_rowptr = rowptr;
    m->h_col = col;
    m->h_diag = diag;
    m->h_bf_face = bf_face;
    m->h_bf_cell = bf_cell;
    m->h_bf_patch = bf_patch;
    m->h_bstart = bstart;
    m->h_fdelta = fdelta;
    m->h_vol = vol;
    m->h_nx = nxv;
    m->h_ny = nyv;
    m->h_nz = nzv;
    FVH_TRY(build_schedule(m.get()));
    FVH_TRY(band_setup(m.get()));
    DevMesh &v = m->d;
    v.n = n; v.nf = nf; v.nbf = nbf; v.nnz = nnz; v.ncf = ncf;
    v.rowptr = m->rowptr.p; v.col = m->col.p; v.diag = m->diag.p; v.tslot = m->tslot.p;
    v.cf_ptr = m->cf_ptr.p; v.cf_face = m->cf_face.p; v.cf_nbr = m->cf_nbr.p; v.cf_slot = m->cf_slot.p;
    v.cf_sign = m->cf_sign.p; v.cf_w = m->cf_w.p; v.cf_omw = m->cf_omw.p;
    v.nif = nif; v.fi_slot = (const int2 *)m->fi_slot.p; v.fi_w = (const double2 *)m->fi_w.p;
    v.own1 = m->own1.p; v.own2 = m->own2.p; v.fw = m->fw.p; v.farea = m->farea.p; v.fdelta = m->fdelta.p;
    v.Sfx = m->Sfx.p; v.Sfy = m->Sfy.p; v.Sfz = m->Sfz.p; v.nx = m->nx.p; v.ny = m->ny.p; v.nz = m->nz.p;
    v.ex = m->ex.p; v.ey = m->ey.p; v.ez = m->ez.p;
    v.vol = m->vol.p; v.ccx = m->ccx.p; v.ccy = m->ccy.p; v.ccz = m->ccz.p; v.fcx = m->fcx.p; v.fcy = m->fcy.p; v.fcz = m->fcz.p;
    v.cb_ptr = m->cb_ptr.p; v.cb_face = m->cb_face.p;
    v.bf_face = m->bf_face.p; v.bf_cell = m->bf_cell.p; v.bf_patch = m->bf_patch.p;
    *out = m.release();
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" void fvh_mesh_destroy(fvh_mesh *m) { delete m; }

extern "C" int fvh_mesh_pattern(const fvh_mesh *m, int64_t *nnz, int32_t *rowptr, int32_t *colidx) {
    if (!m) return set_error(FVH_ERR_ARG, "null mesh");
    if (nnz) *nnz = m->nnz;
    if (rowptr) std::copy(m->h_rowptr.begin(), m->h_rowptr.end(), rowptr);
    if (colidx) std::copy(m->h_col.begin(), m->h_col.end(), colidx);
    return FVH_OK;
}

extern "C" int fvh_mesh_sweep_info(const fvh_mesh *m, int64_t *nlevels, int64_t *band_nx) {
    if (!m) return set_error(FVH_ERR_ARG, "null mesh");
    if (nlevels) *nlevels = m->sched.nlev;
    if (band_nx) *band_nx = m->sched.band_nx;
    return FVH_OK;
}
