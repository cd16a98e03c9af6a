// Host side of the LDL preconditioner: a nested-dissection ordering from
// cell centres and the numeric LDL^T factorization of the permuted matrix.
//
// The reference builds this preconditioner once, when the solver is set up
// (opLDL(A) in Preconditioner{LDL}, preconditioners_0_types.jl:36-42, with
// LinearOperators 2.6.0 over LDLFactorizations 0.10.1), and never updates
// it (update_preconditioner! for LDL is a no-op,
// preconditioners_2_functions.jl:52-57): every later solve is preconditioned
// by the inverse of the set-up matrix.  The factorization is the up-looking
// LDL^T of Davis ("Algorithm 849: a concise sparse Cholesky factorization
// package", ACM TOMS 31(4), 2005), which LDLFactorizations restates; the
// ordering is our own (geometric nested dissection: recursive bisection at
// the median cell centre along the longer side, the cells next to the other
// half forming the separator, numbered last), where LDLFactorizations uses
// AMD.  Ordering and rounding therefore differ from the reference's; the
// preconditioner they approximate, A^-1 of the set-up matrix, does not.
#include "fvh_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "ldl_host.h"

namespace fvh {

namespace {

struct NdState {
    const int *rowptr, *col;
    const double *x, *y;
    std::vector<int> mark;  // per cell: side stamp of the current split
    int stamp = 0;
    std::vector<int> *out;
    std::vector<int> *bounds;  // end of every leaf and separator appended (or null)
};

constexpr int kNdLeaf = 48;

void nd_split(NdState &S, std::vector<int> cells) {
    // explicit stack instead of recursion: (cells, state) frames
    struct Frame {
        std::vector<int> cells;
        int stage;  // 0: split; 1: append separator after children
        std::vector<int> sep;
    };
    std::vector<Frame> st;
    st.push_back({std::move(cells), 0, {}});
    while (!st.empty()) {
        Frame &F = st.back();
        if (F.stage == 1) {
            S.out->insert(S.out->end(), F.sep.begin(), F.sep.end());
            if (S.bounds && !F.sep.empty()) S.bounds->push_back((int)S.out->size());
            st.pop_back();
            continue;
        }
        if ((int)F.cells.size() <= kNdLeaf) {
            std::sort(F.cells.begin(), F.cells.end());
            S.out->insert(S.out->end(), F.cells.begin(), F.cells.end());
            if (S.bounds && !F.cells.empty()) S.bounds->push_back((int)S.out->size());
            st.pop_back();
            continue;
        }
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (int c : F.cells) {
            x0 = std::min(x0, S.x[c]);
            x1 = std::max(x1, S.x[c]);
            y0 = std::min(y0, S.y[c]);
            y1 = std::max(y1, S.y[c]);
        }
        const double *coord = (x1 - x0 >= y1 - y0) ? S.x : S.y;
        std::vector<int> v = F.cells;
        const size_t half = v.size() / 2;
        std::nth_element(v.begin(), v.begin() + half, v.end(), [&](int a, int b) {
            return coord[a] < coord[b] || (coord[a] == coord[b] && a < b);
        });
        const int sl = ++S.stamp, sr = ++S.stamp;
        for (size_t i = 0; i < v.size(); ++i) S.mark[v[i]] = i < half ? sl : sr;
        std::vector<int> left, right, sep;
        for (size_t i = 0; i < half; ++i) {
            const int c = v[i];
            bool touches = false;
            for (int p = S.rowptr[c]; p < S.rowptr[c + 1] && !touches; ++p) touches = S.mark[S.col[p]] == sr;
            (touches ? sep : left).push_back(c);
        }
        right.assign(v.begin() + half, v.end());
        std::sort(sep.begin(), sep.end());
        F.stage = 1;
        F.sep = std::move(sep);
        F.cells.clear();
        F.cells.shrink_to_fit();
        // right is numbered after left: push it first (the stack pops left first)
        st.push_back({std::move(right), 0, {}});
        st.push_back({std::move(left), 0, {}});
    }
}

}  // namespace

void ldl_order(int n, const int *rowptr, const int *col, const double *x, const double *y, std::vector<int> &perm,
               std::vector<int> *bounds) {
    NdState S{rowptr, col, x, y, std::vector<int>(n, 0), 0, &perm, bounds};
    perm.clear();
    perm.reserve(n);
    if (bounds) bounds->assign(1, 0);
    std::vector<int> all(n);
    std::iota(all.begin(), all.end(), 0);
    nd_split(S, std::move(all));
}

int ldl_factor(int n, const int *rowptr, const int *col, const double *val, const std::vector<int> &perm, LdlHost &F) {
    if ((int)perm.size() != n) return set_error(FVH_ERR_ARG, "LDL: ordering of %zu rows for %d", perm.size(), n);
    std::vector<int> iperm(n, -1);
    for (int k = 0; k < n; ++k) {
        if (perm[k] < 0 || perm[k] >= n || iperm[perm[k]] >= 0) return set_error(FVH_ERR_ARG, "LDL: not a permutation");
        iperm[perm[k]] = k;
    }
    // A(r, c) from the upper triangle of A in its own numbering
    auto upper = [&](int r, int c) -> double {
        const int a = std::min(r, c), b = std::max(r, c);
        const int *lo = col + rowptr[a], *hi = col + rowptr[a + 1];
        const int *it = std::lower_bound(lo, hi, b);
        return (it != hi && *it == b) ? val[it - col] : 0.0;
    };
    // symbolic: elimination tree and column counts of L
    std::vector<int> parent(n, -1), flag(n), lnz(n, 0);
    for (int k = 0; k < n; ++k) {
        flag[k] = k;
        const int r = perm[k];
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            int i = iperm[col[p]];
            if (i >= k) continue;
            for (; flag[i] != k; i = parent[i]) {
                if (parent[i] == -1) parent[i] = k;
                lnz[i]++;
                flag[i] = k;
            }
        }
    }
    F.n = n;
    F.perm = perm;
    F.Lp.assign(n + 1, 0);
    for (int k = 0; k < n; ++k) F.Lp[k + 1] = F.Lp[k] + lnz[k];
    const int64_t nnzL = F.Lp[n];
    if (nnzL >= (int64_t)1 << 31) return set_error(FVH_ERR_SIZE, "LDL: %lld entries in L", (long long)nnzL);
    F.Li.assign(nnzL, 0);
    F.Lx.assign(nnzL, 0.0);
    F.D.assign(n, 0.0);
    // numeric: row k of L from a sparse triangular solve along the etree
    std::vector<double> yv(n, 0.0);
    std::vector<int> pattern(n), fill(n, 0);
    for (int k = 0; k < n; ++k) {
        int top = n;
        flag[k] = k;
        const int r = perm[k];
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            int i = iperm[col[p]];
            if (i > k) continue;
            yv[i] += upper(r, col[p]);
            int len = 0;
            for (; flag[i] != k; i = parent[i]) {
                pattern[len++] = i;
                flag[i] = k;
            }
            while (len > 0) pattern[--top] = pattern[--len];
        }
        double d = yv[k];
        yv[k] = 0.0;
        for (; top < n; ++top) {
            const int i = pattern[top];
            const double yi = yv[i];
            yv[i] = 0.0;
            const int p2 = F.Lp[i] + fill[i];
            for (int p = F.Lp[i]; p < p2; ++p) yv[F.Li[p]] -= F.Lx[p] * yi;
            const double lki = yi / F.D[i];
            d -= lki * yi;
            F.Li[p2] = k;
            F.Lx[p2] = lki;
            fill[i]++;
        }
        if (d == 0.0 || !std::isfinite(d))
            return set_error(FVH_ERR_ARG, "LDL: pivot %d (row %d) is %g: the set-up matrix is singular", k, r, d);
        F.D[k] = d;
    }
    return FVH_OK;
}

}  // namespace fvh
