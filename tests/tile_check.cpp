// Host check of the tile walk's schedule (csrc/tile_sched.h) on a CSR
// pattern read from a file: the structure (positions a permutation, tiles,
// chunks and steps covering each tile in order, the 64-row and 64-neighbour
// limits) and a sequential replay of what the kernel's waves do - loaders
// copying each chunk's neighbours from other tiles out of the global result
// array into LDS words, the compute wave taking the steps in order and
// reading every neighbour from its LDS word - which must find every word
// written, by an earlier step or a staging, and give the DILU update,
// forward and backward sweeps of a plain row-by-row loop bit for bit.
//
// usage: tile_check pattern.bin tile_rows lag
// pattern.bin: int32 n, int32 nnz, rowptr[n + 1], col[nnz] (ascending per row)
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "tile_sched.h"

using fvh::TileSchedule;

static int fail(const char *what, long a = 0, long b = 0) {
    std::printf("FAIL %s (%ld, %ld)\n", what, a, b);
    return 1;
}

int main(int argc, char **argv) {
    if (argc < 4) return fail("usage");
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return fail("open");
    int n = 0, nnz = 0;
    if (std::fread(&n, 4, 1, f) != 1 || std::fread(&nnz, 4, 1, f) != 1) return fail("header");
    std::vector<int> rp(n + 1), col(nnz), dg(n), tslot(nnz);
    if (std::fread(rp.data(), 4, n + 1, f) != (size_t)n + 1 || std::fread(col.data(), 4, nnz, f) != (size_t)nnz)
        return fail("body");
    std::fclose(f);
    const int tr = std::atoi(argv[2]), lag = std::atoi(argv[3]);
    for (int i = 0; i < n; ++i) {
        dg[i] = -1;
        for (int p = rp[i]; p < rp[i + 1]; ++p)
            if (col[p] == i) dg[i] = p;
        if (dg[i] < 0) return fail("no diagonal", i);
        for (int p = rp[i]; p < rp[i + 1]; ++p) {  // A_ji
            const int j = col[p];
            tslot[p] = -1;
            for (int q = rp[j]; q < rp[j + 1]; ++q)
                if (col[q] == i) tslot[p] = q;
            if (tslot[p] < 0) return fail("pattern not symmetric", i, j);
        }
    }
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0.1, 1.0), X(-1.0, 1.0);
    std::vector<double> val(nnz), x(n);
    for (int i = 0; i < n; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p) val[p] = col[p] == i ? 4.5 : -U(rng);
    for (auto &v : x) v = X(rng);

    // plain row-by-row loops (the reference's order: neighbours by ascending column)
    std::vector<double> D(n), Dinv(n), y(n), z(n);
    for (int i = 0; i < n; ++i) {
        double d = val[dg[i]];
        for (int p = rp[i]; p < dg[i]; ++p) d = d - val[p] * val[tslot[p]] / D[col[p]];
        D[i] = d;
        Dinv[i] = 1.0 / d;
    }
    for (int i = 0; i < n; ++i) {
        double s = x[i];
        for (int p = rp[i]; p < dg[i]; ++p) s = s - val[p] * y[col[p]] * Dinv[col[p]];
        y[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i] * Dinv[i];
        for (int p = dg[i] + 1; p < rp[i + 1]; ++p) s = s - val[p] * z[col[p]] * Dinv[i];
        z[i] = s;
    }

    long nsteps = 0, nchunks = 0, nents = 0;
    int maxrem = 0, ntiles = 0;
    for (int mode = 0; mode < 3; ++mode) {
        const bool upper = mode == 2;
        TileSchedule ts;
        std::vector<int> code(nnz, -1);
        fvh::tile_schedule(n, rp, col, dg, upper, tr, lag, ts, code);
        const int nt = ts.ntiles;
        if (nt != (n + tr - 1) / tr || (int)ts.tiles.size() != 2 * nt) return fail("tile count", nt);
        std::vector<int> seen(n, 0);
        for (int i = 0; i < n; ++i) {
            if (ts.pos[i] < 0 || ts.pos[i] >= n || seen[ts.pos[i]]++) return fail("positions", i);
            if (ts.rows[ts.pos[i]] != i) return fail("rows", i);
            const int k = upper ? n - 1 - i : i;  // the row's place in the sweep's order
            if (ts.pos[i] / tr != k / tr) return fail("row outside its tile", i);
        }
        // replay
        std::vector<double> out(n, 0.0), res;
        std::vector<char> done(n, 0);
        std::vector<long> wstep;  // step after which each LDS word holds its value (-1: not yet)
        long step_id = 0;
        int expect_chunk = 0;
        for (int t = 0; t < nt; ++t) {
            const int c0 = ts.tiles[2 * t], nc = ts.tiles[2 * t + 1];
            if (c0 != expect_chunk || nc < 1) return fail("chunk range", t, c0);
            expect_chunk += nc;
            const int k0 = t * tr, k1 = std::min(n, k0 + tr);
            res.assign(tr + ts.maxrem, 0.0);
            wstep.assign(tr + ts.maxrem, -1);
            int next_pos = k0;
            for (int c = c0; c < c0 + nc; ++c) {
                const int *C = &ts.chunks[(size_t)fvh::kTileChunkInts * c];
                const int fs = C[0], ns = C[1], fe = C[2], ne = C[3], fp = C[4], nr = C[5];
                if (ns < 1 || ns > 64 || nr < 1 || nr > 64 || ne < 0 || ne > 64) return fail("chunk limits", c, nr);
                if (fp != next_pos || ts.steps[2 * fs] != fp) return fail("chunk position", c, fp);
                for (int e = fe; e < fe + ne; ++e) {  // a loader stages the chunk's neighbours
                    const int src = ts.ents[2 * e], word = ts.ents[2 * e + 1];
                    if (src < 0 || src >= k0 || !done[src]) return fail("staged value not ready", c, src);
                    if (word < tr || word >= tr + ts.maxrem || wstep[word] >= 0) return fail("staging word", c, word);
                    res[word] = out[src];
                    wstep[word] = step_id;
                }
                int rows_in_chunk = 0;
                for (int s = fs; s < fs + ns; ++s) {
                    const int sp = ts.steps[2 * s], sn = ts.steps[2 * s + 1];
                    if (sp != next_pos || sn < 1 || sn > 64) return fail("step", s, sp);
                    ++step_id;
                    std::vector<std::pair<int, double>> wr;
                    for (int k = sp; k < sp + sn; ++k) {
                        const int i = ts.rows[k];
                        const int p0 = upper ? dg[i] + 1 : rp[i], p1 = upper ? rp[i + 1] : dg[i];
                        double v;
                        if (mode == 0) v = val[dg[i]];
                        else if (mode == 1) v = x[i];
                        else v = y[i] * Dinv[i];
                        for (int p = p0; p < p1; ++p) {
                            const int w = code[p];
                            if (w < 0 || w >= tr + ts.maxrem || wstep[w] < 0 || wstep[w] >= step_id)
                                return fail("neighbour word not written before the step", i, w);
                            const double nb = res[w];
                            if (mode == 0) v = v - val[p] * val[tslot[p]] / nb;
                            else if (mode == 1) v = v - val[p] * nb * Dinv[col[p]];
                            else v = v - val[p] * nb * Dinv[i];
                        }
                        wr.push_back({k, v});
                    }
                    for (auto &kv : wr) {  // the step's results: LDS word and global array
                        res[kv.first - k0] = kv.second;
                        wstep[kv.first - k0] = step_id;
                        out[kv.first] = kv.second;
                        done[kv.first] = 1;
                    }
                    next_pos += sn;
                    rows_in_chunk += sn;
                    ++nsteps;
                }
                if (rows_in_chunk != nr) return fail("chunk rows", c, nr);
                nents += ne;
                ++nchunks;
            }
            if (next_pos != k1) return fail("tile not covered", t, next_pos);
        }
        const std::vector<double> &want = mode == 0 ? D : mode == 1 ? y : z;
        for (int i = 0; i < n; ++i)
            if (out[ts.pos[i]] != want[i]) return fail("result differs", mode, i);
        maxrem = std::max(maxrem, ts.maxrem);
        ntiles = nt;
    }
    std::printf("ok n=%d tiles=%d chunks=%ld steps=%ld staged=%ld maxrem=%d\n", n, ntiles, nchunks, nsteps, nents,
                maxrem);
    return 0;
}
