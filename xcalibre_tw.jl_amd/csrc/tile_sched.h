// Schedule of the tile walk (host only, plain C++): the level-scheduled DILU
// sweeps of unstructured meshes in tiles of consecutive rows.
#pragma once

#include <algorithm>
#include <vector>

namespace fvh {

constexpr int kTileChunkInts = 8;  // ints per chunk record

struct TileSchedule {
    std::vector<int> rows;    // position -> row; tile t holds positions [t * tr, ...)
    std::vector<int> tiles;   // per tile: first chunk, chunk count
    std::vector<int> chunks;  // per chunk (kTileChunkInts): first step, steps, first entry, entries, first position, rows
    std::vector<int> steps;   // per step: first position, rows
    std::vector<int> ents;    // per staged neighbour: its position, its LDS word
    std::vector<int> pos;     // row -> position
    int ntiles = 0, maxrem = 0;
};

// One direction of the tile walk over the CSR pattern (rp, col, dg: row
// pointers, columns ascending, diagonal entries).  Rows are taken in the
// sweep's order (forward: ascending, neighbours below the diagonal;
// backward: descending, neighbours above), tile t holding the tr rows from
// t * tr on.  Each row gets a step key: the largest key of its neighbours
// plus one for a neighbour of its own tile, plus `lag` for one of another
// tile (whose result travels through global memory and a loader wave).  A
// tile's rows go to its positions by key (then in the sweep's order), and
// its steps are runs of one key, at most 64 rows; chunks are runs of steps
// with at most 64 rows, keys at most `lag` above the first, and at most 64
// neighbours from other tiles not yet staged (each staged once per tile, to
// LDS word tr + slot).  code[p] of each
// entry p of the direction: the LDS word where the walk finds the neighbour
// (its place in the tile, or its staging word).
inline void tile_schedule(int n, const std::vector<int> &rp, const std::vector<int> &col, const std::vector<int> &dg,
                          bool upper, int tr, int lag, TileSchedule &ts, std::vector<int> &code) {
    auto row_at = [&](int k) { return upper ? n - 1 - k : k; };  // k-th row of the sweep
    auto tile_of = [&](int i) { return (upper ? n - 1 - i : i) / tr; };
    auto deps = [&](int i, int &p0, int &p1) {
        p0 = upper ? dg[i] + 1 : rp[i];
        p1 = upper ? rp[i + 1] : dg[i];
    };
    std::vector<int> key(n, 0);
    for (int k = 0; k < n; ++k) {
        const int i = row_at(k);
        int p0, p1, w = 0;
        deps(i, p0, p1);
        for (int p = p0; p < p1; ++p) {
            const int j = col[p];
            w = std::max(w, key[j] + (tile_of(j) == tile_of(i) ? 1 : lag));
        }
        key[i] = w;
    }
    const int nt = (n + tr - 1) / tr;
    std::vector<int> &pos = ts.pos, &rows = ts.rows, &tiles = ts.tiles, &chunks = ts.chunks, &steps = ts.steps,
                     &ents = ts.ents;
    pos.assign(n, 0);
    rows.assign(n, 0);
    tiles.clear();
    chunks.clear();
    steps.clear();
    ents.clear();
    std::vector<int> slot(n, -1), used, list;
    ts.maxrem = 0;
    for (int t = 0; t < nt; ++t) {
        const int k0 = t * tr, k1 = std::min(n, k0 + tr);
        list.clear();
        for (int k = k0; k < k1; ++k) list.push_back(row_at(k));
        std::stable_sort(list.begin(), list.end(), [&](int a, int b) { return key[a] < key[b]; });
        for (int r = 0; r < (int)list.size(); ++r) {
            pos[list[r]] = k0 + r;
            rows[k0 + r] = list[r];
        }
        tiles.push_back((int)chunks.size() / kTileChunkInts);
        const int chunk0 = (int)chunks.size();
        int nrem = 0;
        int c_step = (int)steps.size() / 2, c_ent = (int)ents.size() / 2, c_pos = k0, c_rows = 0, c_ents = 0, c_key = 0;
        int s_pos = k0, s_rows = 0, s_key = 0;
        auto close_step = [&]() {
            if (s_rows == 0) return;
            steps.push_back(s_pos);
            steps.push_back(s_rows);
            s_rows = 0;
        };
        auto close_chunk = [&]() {
            close_step();
            if (c_rows == 0) return;
            const int ns = (int)steps.size() / 2 - c_step;
            const int rec[kTileChunkInts] = {c_step, ns, c_ent, c_ents, c_pos, c_rows, 0, 0};
            chunks.insert(chunks.end(), rec, rec + kTileChunkInts);
            c_step = (int)steps.size() / 2;
            c_ent = (int)ents.size() / 2;
            c_rows = c_ents = 0;
        };
        for (int r = 0; r < (int)list.size(); ++r) {
            const int i = list[r], p = k0 + r;
            int p0, p1, fresh = 0;
            deps(i, p0, p1);
            for (int q = p0; q < p1; ++q)
                if (tile_of(col[q]) != t && slot[col[q]] < 0) ++fresh;
            // a chunk is staged when all its neighbours from other tiles are
            // there, which for a row of key k is about k - lag: rows more than
            // lag behind the chunk's first would hold it back
            if (c_rows + 1 > 64 || c_ents + fresh > 64 || (c_rows > 0 && key[i] > c_key + lag)) close_chunk();
            if (s_rows > 0 && (key[i] != s_key || s_rows == 64)) close_step();
            if (c_rows == 0) {
                c_pos = p;
                c_key = key[i];
            }
            if (s_rows == 0) {
                s_pos = p;
                s_key = key[i];
            }
            for (int q = p0; q < p1; ++q) {
                const int j = col[q];
                if (tile_of(j) == t) {
                    code[q] = pos[j] - k0;
                } else {
                    if (slot[j] < 0) {
                        slot[j] = nrem++;
                        used.push_back(j);
                        ents.push_back(pos[j]);
                        ents.push_back(tr + slot[j]);
                        ++c_ents;
                    }
                    code[q] = tr + slot[j];
                }
            }
            ++s_rows;
            ++c_rows;
        }
        close_chunk();
        tiles.push_back(((int)chunks.size() - chunk0) / kTileChunkInts);
        ts.maxrem = std::max(ts.maxrem, nrem);
        for (int j : used) slot[j] = -1;
        used.clear();
    }
    ts.ntiles = nt;
}

}  // namespace fvh
