// ============================================================================
// ORACLE — test infrastructure only.  Not part of the product.
//
// A CPU FP64 restatement of XCALibre_TW.jl's per-iteration hot path, used as
// the checker for libfvmhip.so.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it.
//
// It deliberately keeps the reference's own data structures and loop orders
// (citations relative to the XCALibre_TW.jl root):
//   * SparseMatrixCSC with rows sorted in each column ........ ModelFramework_0_types.jl:90-124
//   * cell-by-cell face-loop assembly, slots found by search .. Discretise_2_generated_distretisation.jl:3-81
//   * scheme!/scheme_source! per term ......................... Discretise_1_schemes.jl:11-111
//   * BC functors per patch / face / term ..................... Discretise_3_boundary_conditions.jl,
//                                                               Discretise_4_apply_bcs.jl:3-44
//   * relaxation, pressure reference .......................... Solve_1_api.jl:48-69
//   * column-scatter SpMV (SparseArrays mul!) ................. Solvers_0_functions.jl:18
//   * DILU with Di / Ri / J lists, column-oriented sweeps ..... preconditioners_1_DILU.jl:5-167
//   * LDL of the set-up matrix, never updated ................. preconditioners_0_types.jl:36-42,
//                                                               preconditioners_2_functions.jl:52-57
//   * Jacobi / NormDiagonal ................................... preconditioners_2_functions.jl:15-50
//   * interpolation, boundary correction, Green-Gauss, div .... Calculate_2_interpolation.jl,
//                                                               Calculate_1_green_gauss.jl, Calculate_0_divergence.jl
//   * SIMPLE and PISO loop bodies ............................. Solvers_1_SIMPLE.jl:69-231, Solvers_2_PISO.jl
//   * surface flux / normal gradient, forces, wall shear ...... Calculate_5_surface_normal_gradient.jl,
//                                                               Postprocess/Postprocess_forces.jl:5-96
//   * non-orthogonal correction (the loops' `correct` block) .. Calculate_3_orthogonality_correction.jl,
//                                                               Calculate_2_interpolation.jl:152-176
//   * k-omega transport, wall functions, bound! ............... RANSModels_kOmega.jl, RANSModels_functions.jl
//
// Arithmetic that lives in packages absent from the reference tree is
// restated from the packages' published algorithms:
//   * Krylov.jl 0.9.5 BiCGSTAB / CG / GMRES (left preconditioning, warm start
//     via the Delta-x correction, stopping at atol + rtol*||r0||);
//   * ILUZero.jl 0.2.0 -> textbook ILU(0), Saad "Iterative Methods" Alg. 10.4;
//   * LDLFactorizations.jl 0.10.1 -> Davis' up-looking LDL^T (ACM TOMS 849),
//     on A's upper triangle in A's own order (the package orders with AMD.jl,
//     absent here);
//   * BLAS dot / nrm2 and Statistics.mean -> one fixed blocked summation
//     order (chunks of 2048 = 256 lanes x 8, pairwise tree of 256), which the
//     GPU library reproduces;  Julia's own arithmetic never contracts to FMA,
//     and neither does this file (-ffp-contract=off).
//   * log in the wall functions -> fvh_log (shared with the GPU), a
//     deterministic natural logarithm from the atanh series.
// Parity of those pieces against the Julia packages is unpinned; everything
// else follows the reference operation by operation.
// ============================================================================
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../include/fvmhip.h"
#include "../xcalibre_tw.jl_amd/csrc/fvh_mathfn.h"

namespace ref {

using Vec = std::vector<double>;
using IVec = std::vector<int64_t>;

static thread_local char g_err[512];
static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

// Julia's max(x, y) for Float64: NaN-propagating, max(-0.0, 0.0) = 0.0.
static inline double jmax(double x, double y) {
    bool take_y = (y > x) || (std::signbit(y) < std::signbit(x));
    return take_y ? (std::isnan(x) ? x : y) : (std::isnan(y) ? y : x);
}
static inline double jsign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }

// ---------------------------------------------------------------------------
// Canonical summation order (shared with the GPU reductions).
// ---------------------------------------------------------------------------
template <class F>
static double det_sum(int64_t n, F f) {
    const int64_t nch = (n + 2047) / 2048;
    Vec part((size_t)std::max<int64_t>(nch, 1), 0.0);
    double acc[256];
    for (int64_t c = 0; c < nch; ++c) {
        for (int t = 0; t < 256; ++t) {
            double a = 0.0;
            for (int k = 0; k < 8; ++k) {
                int64_t i = c * 2048 + k * 256 + t;
                if (i < n) a = a + f(i);
            }
            acc[t] = a;
        }
        for (int s = 128; s > 0; s >>= 1)
            for (int t = 0; t < s; ++t) acc[t] = acc[t] + acc[t + s];
        part[c] = acc[0];
    }
    for (int t = 0; t < 256; ++t) {
        double a = 0.0;
        for (int64_t j = t; j < nch; j += 256) a = a + part[j];
        acc[t] = a;
    }
    for (int s = 128; s > 0; s >>= 1)
        for (int t = 0; t < s; ++t) acc[t] = acc[t] + acc[t + s];
    return acc[0];
}
static double dot(const Vec &x, const Vec &y) {
    return det_sum((int64_t)x.size(), [&](int64_t i) { return x[i] * y[i]; });
}
static double nrm2(const Vec &x) { return std::sqrt(dot(x, x)); }

// ---------------------------------------------------------------------------
// Mesh (Mesh2, zero based)
// ---------------------------------------------------------------------------
struct V3 {
    double x, y, z;
};
static inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline double vdot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline double vnorm(V3 a) { return std::sqrt((a.x * a.x + a.y * a.y) + a.z * a.z); }

struct Boundary {
    IVec facesID, cellsID;
};
struct Mesh {
    int64_t n = 0, nf = 0, nbf = 0;
    std::vector<V3> ccentre;
    Vec volume;
    IVec fstart, fend;  // faces_range as [start, end)
    IVec cell_faces, cell_neighbours, cell_nsign;
    IVec own1, own2;
    std::vector<V3> fcentre, fnormal, fe;
    Vec farea, fdelta, fweight;
    std::vector<Boundary> boundaries;
};

static int load_mesh(const fvh_mesh_desc *d, Mesh &m) {
    const int64_t b = d->index_base;
    m.n = d->ncells;
    m.nf = d->nfaces;
    m.ccentre.resize(m.n);
    m.volume.resize(m.n);
    m.fstart.resize(m.n);
    m.fend.resize(m.n);
    for (int64_t c = 0; c < m.n; ++c) {
        m.ccentre[c] = {d->cell_centre[3 * c], d->cell_centre[3 * c + 1], d->cell_centre[3 * c + 2]};
        m.volume[c] = d->cell_volume[c];
        m.fstart[c] = d->cell_faces_range[2 * c] - b;
        m.fend[c] = d->cell_faces_range[2 * c + 1] - b + 1;
    }
    m.cell_faces.resize(d->ncellfaces);
    m.cell_neighbours.resize(d->ncellfaces);
    m.cell_nsign.resize(d->ncellfaces);
    for (int64_t i = 0; i < d->ncellfaces; ++i) {
        m.cell_faces[i] = d->cell_faces[i] - b;
        m.cell_neighbours[i] = d->cell_neighbours[i] - b;
        m.cell_nsign[i] = d->cell_nsign[i];
    }
    m.own1.resize(m.nf);
    m.own2.resize(m.nf);
    m.fcentre.resize(m.nf);
    m.fnormal.resize(m.nf);
    m.fe.resize(m.nf);
    m.farea.resize(m.nf);
    m.fdelta.resize(m.nf);
    m.fweight.resize(m.nf);
    for (int64_t f = 0; f < m.nf; ++f) {
        m.own1[f] = d->face_owners[2 * f] - b;
        m.own2[f] = d->face_owners[2 * f + 1] - b;
        m.fcentre[f] = {d->face_centre[3 * f], d->face_centre[3 * f + 1], d->face_centre[3 * f + 2]};
        m.fnormal[f] = {d->face_normal[3 * f], d->face_normal[3 * f + 1], d->face_normal[3 * f + 2]};
        m.fe[f] = {d->face_e[3 * f], d->face_e[3 * f + 1], d->face_e[3 * f + 2]};
        m.farea[f] = d->face_area[f];
        m.fdelta[f] = d->face_delta[f];
        m.fweight[f] = d->face_weight[f];
    }
    m.boundaries.resize(d->nboundaries);
    m.nbf = 0;
    for (int64_t g = 0; g < d->nboundaries; ++g) {
        int64_t a = d->boundary_range[2 * g] - b, e = d->boundary_range[2 * g + 1] - b;
        for (int64_t i = a; i <= e; ++i) {
            m.boundaries[g].facesID.push_back(d->boundary_faces[i] - b);
            m.boundaries[g].cellsID.push_back(d->boundary_cells[i] - b);
        }
        m.nbf += (int64_t)m.boundaries[g].facesID.size();
    }
    return FVH_OK;
}

// ---------------------------------------------------------------------------
// SparseMatrixCSC (Equation(mesh): sparse(i, j, v) of the cell adjacency)
// ---------------------------------------------------------------------------
struct CSC {
    int64_t n = 0;
    IVec colptr, rowval;
    Vec nzval;
    int64_t slot(int64_t i, int64_t j) const {  // A[i,j] position, -1 if structurally zero
        auto b = rowval.begin() + colptr[j], e = rowval.begin() + colptr[j + 1];
        auto it = std::lower_bound(b, e, i);
        return (it != e && *it == i) ? (int64_t)(it - rowval.begin()) : -1;
    }
    double get(int64_t i, int64_t j) const {
        int64_t s = slot(i, j);
        return s < 0 ? 0.0 : nzval[s];
    }
    double &at(int64_t i, int64_t j) { return nzval[slot(i, j)]; }
};

static CSC make_pattern(const Mesh &m) {
    CSC A;
    A.n = m.n;
    std::vector<IVec> cols(m.n);
    for (int64_t c = 0; c < m.n; ++c) {
        cols[c].push_back(c);  // entry (c, c)
        for (int64_t fi = m.fstart[c]; fi < m.fend[c]; ++fi) cols[m.cell_neighbours[fi]].push_back(c);  // (c, nbr)
    }
    A.colptr.assign(m.n + 1, 0);
    for (int64_t j = 0; j < m.n; ++j) {
        std::sort(cols[j].begin(), cols[j].end());
        cols[j].erase(std::unique(cols[j].begin(), cols[j].end()), cols[j].end());
        A.colptr[j + 1] = A.colptr[j] + (int64_t)cols[j].size();
    }
    A.rowval.reserve(A.colptr[m.n]);
    for (int64_t j = 0; j < m.n; ++j) A.rowval.insert(A.rowval.end(), cols[j].begin(), cols[j].end());
    A.nzval.assign(A.rowval.size(), 0.0);
    return A;
}

// mul!(y, A, x): y = 0; for each column, y[row] += A[row,col]*x[col]
static void spmv(const CSC &A, const Vec &x, Vec &y) {
    std::fill(y.begin(), y.end(), 0.0);
    for (int64_t col = 0; col < A.n; ++col) {
        double xj = x[col];
        for (int64_t p = A.colptr[col]; p < A.colptr[col + 1]; ++p) y[A.rowval[p]] += A.nzval[p] * xj;
    }
}

struct Equation {
    CSC A;
    Vec b, R, Fx;
};

// ---------------------------------------------------------------------------
// Model terms
// ---------------------------------------------------------------------------
struct Term {
    int op, scheme, sign;
    const Vec *flux;  // nullptr = ConstantScalar(1)
    const Vec *phi;
    double fl(int64_t i) const { return flux ? (*flux)[i] : 1.0; }
};
struct Source {
    int sign;
    const Vec *field;
};
struct Model {
    std::vector<Term> terms;
    std::vector<Source> sources;
};

// discretise!(eqn, prev, runtime)
static void discretise(const Mesh &m, const Model &md, Equation &eq, const Vec *prev, double dt) {
    Vec &nz = eq.A.nzval;
    Vec &b = eq.b;
    std::fill(nz.begin(), nz.end(), 0.0);
    for (int64_t c = 0; c < m.n; ++c) {
        int64_t cIndex = -1;
        for (int64_t fi = m.fstart[c]; fi < m.fend[c]; ++fi) {
            int64_t f = m.cell_faces[fi];
            double ns = (double)m.cell_nsign[fi];
            int64_t nID = m.cell_neighbours[fi];
            cIndex = eq.A.slot(c, c);
            int64_t nIndex = eq.A.slot(c, nID);
            for (const Term &t : md.terms) {
                if (t.op == FVH_OP_LAPLACIAN) {
                    double ap = ((double)t.sign * (-t.fl(f) * m.farea[f])) / m.fdelta[f];
                    nz[cIndex] += ap;
                    nz[nIndex] += -ap;
                } else if (t.op == FVH_OP_DIVERGENCE && t.scheme == FVH_LINEAR) {
                    V3 xf = m.fcentre[f], xC = m.ccentre[c], xN = m.ccentre[nID];
                    double weight = vnorm(sub(xf, xC)) / vnorm(sub(xN, xC));
                    double omw = 1.0 - weight;
                    double ap = (double)t.sign * (t.fl(f) * ns);
                    nz[cIndex] += ap * omw;
                    nz[nIndex] += ap * weight;
                } else if (t.op == FVH_OP_DIVERGENCE && t.scheme == FVH_UPWIND) {
                    double ap = (double)t.sign * (t.fl(f) * ns);
                    nz[cIndex] += jmax(ap, 0.0);
                    nz[nIndex] += -jmax(-ap, 0.0);
                }
            }
        }
        b[c] = 0.0;
        double volume = m.volume[c];
        if (cIndex < 0) cIndex = eq.A.slot(c, c);
        for (const Term &t : md.terms) {
            if (t.op == FVH_OP_TIME && t.scheme == FVH_EULER) {
                double rdt = 1.0 / dt;
                double pv = prev ? (*prev)[c] : 0.0;
                nz[cIndex] += volume * rdt;
                b[c] += pv * volume * rdt;
            } else if (t.op == FVH_OP_SI) {
                double flux = (double)t.sign * t.fl(c) * volume;
                nz[cIndex] += flux;
            }
        }
        for (const Source &s : md.sources) b[c] += (double)s.sign * (*s.field)[c] * volume;
    }
}

// apply_boundary_conditions!(eqn, BCs)
static int apply_bcs(const Mesh &m, const Model &md, Equation &eq, const std::vector<fvh_bc> &bcs, int comp) {
    for (const fvh_bc &bc : bcs) {
        const Boundary &B = m.boundaries[bc.patch];
        double value = bc.value[comp];
        for (size_t i = 0; i < B.cellsID.size(); ++i) {
            int64_t f = B.facesID[i], c = B.cellsID[i];
            for (const Term &t : md.terms) {
                double sg = (double)t.sign;
                if (t.op == FVH_OP_TIME) continue;
                if (bc.kind == FVH_BC_NUTWALL) return fail(FVH_ERR_UNSUPPORTED, "NutWallFunction has no boundary method for an equation term");
                if (t.op == FVH_OP_LAPLACIAN) {
                    if (bc.kind == FVH_BC_DIRICHLET) {
                        double J = t.fl(f);
                        double flux = J * m.farea[f] / m.fdelta[f];
                        double ap = sg * (-flux);
                        eq.A.at(c, c) += ap;
                        eq.b[c] += ap * value;
                    }
                } else if (t.op == FVH_OP_DIVERGENCE && t.scheme == FVH_LINEAR) {
                    if (bc.kind == FVH_BC_DIRICHLET) eq.b[c] += sg * (-t.fl(f) * value);
                    else eq.A.at(c, c) += sg * t.fl(f);
                } else if (t.op == FVH_OP_DIVERGENCE && t.scheme == FVH_UPWIND) {
                    double ap = sg * t.fl(f);
                    if (bc.kind == FVH_BC_DIRICHLET) eq.b[c] -= ap * value;
                    else eq.A.at(c, c) += jmax(ap, 0.0);
                } else if (t.op == FVH_OP_SI) {
                    if (bc.kind == FVH_BC_OMEGAWALL) {
                        double phi = (*t.phi)[c];
                        double flux = sg * t.fl(c);
                        eq.b[c] += flux * phi * m.volume[c];
                    }
                }
            }
        }
    }
    return FVH_OK;
}

static void implicit_relaxation(Equation &eq, const Vec &field, double alpha) {
    for (int64_t i = 0; i < eq.A.n; ++i) {
        double &d = eq.A.at(i, i);
        d /= alpha;
        eq.b[i] += (1.0 - alpha) * d * field[i];
    }
}
static void set_reference(Equation &eq, bool has, double pref, int64_t cell) {
    if (!has) return;
    eq.b[cell] += eq.A.at(cell, cell) * pref;
    eq.A.at(cell, cell) += eq.A.at(cell, cell);
}
static void explicit_relaxation(Vec &phi, const Vec &phi0, double alpha) {
    for (size_t i = 0; i < phi.size(); ++i) phi[i] = phi0[i] + alpha * (phi[i] - phi0[i]);
}
// residual!: sqrt(mean((Ax-b)^2))/norm(b)
static double residual(Equation &eq, const Vec &x) {
    spmv(eq.A, x, eq.Fx);
    const int64_t n = eq.A.n;
    for (int64_t i = 0; i < n; ++i) {
        double d = std::fabs(eq.Fx[i] - eq.b[i]);
        eq.R[i] = d * d;
    }
    double mean = det_sum(n, [&](int64_t i) { return eq.R[i]; }) / (double)n;
    return std::sqrt(mean) / nrm2(eq.b);
}

// ---------------------------------------------------------------------------
// Preconditioners
// ---------------------------------------------------------------------------
struct Precond {
    int kind = FVH_PRECOND_NONE;
    CSC *A = nullptr;  // aliased matrix (reference: P.A === eqn.A)
    // DILU
    Vec D;
    IVec Di;
    std::vector<IVec> Ri, J;
    // Jacobi / NormDiagonal
    Vec storage;
    // ILU(0) on a CSR copy
    IVec rp, ci, dg;
    Vec lu;
    std::vector<IVec> upd_pairs;  // per lower entry: (slot_ij, slot_kj) pairs
    // LDL: the set-up matrix's factor (L by columns, strictly lower), kept
    IVec Lp, Li;
    Vec Lx, Dl;
    bool ldl_ready = false;
    int ldl_status = FVH_OK;
};

static void dilu_setup(Precond &P) {
    const CSC &A = *P.A;
    const int64_t n = A.n;
    P.D.assign(n, 0.0);
    P.Di.assign(n, 0);
    for (int64_t i = 0; i < n; ++i) {  // diagonal_indices!
        int64_t idx = 0;
        for (int64_t p = A.colptr[i]; p < A.colptr[i + 1]; ++p)
            if (A.rowval[p] == i) { idx = p; break; }
        P.Di[i] = idx;
    }
    // upper_row_indices in O(nnz): Ri[i], J[i] = slots / columns of A[i,j], j > i ascending
    P.Ri.assign(n, IVec());
    P.J.assign(n, IVec());
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = A.colptr[j]; p <= P.Di[j]; ++p) {
            int64_t row = A.rowval[p];
            if (row < j) {
                P.Ri[row].push_back(p);
                P.J[row].push_back(j);
            }
        }
}
static void dilu_update(Precond &P) {  // update_dilu_diagonal!
    const CSC &A = *P.A;
    const int64_t n = A.n;
    for (int64_t i = 0; i < n; ++i) P.D[i] = A.nzval[P.Di[i]];
    for (int64_t i = 0; i < n; ++i) {
        int64_t c_start = P.Di[i] + 1, c_end = A.colptr[i + 1] - 1;
        int64_t r_count = 0;
        for (int64_t cp = c_start; cp <= c_end; ++cp) {
            int64_t j = A.rowval[cp];
            P.D[j] -= A.nzval[cp] * A.nzval[P.Ri[i][r_count]] / P.D[i];
            ++r_count;
        }
        P.D[i] = 1.0 / P.D[i];
    }
}
static void dilu_apply(const Precond &P, Vec &x, const Vec &b) {
    const CSC &A = *P.A;
    const int64_t n = A.n;
    for (int64_t i = 0; i < n; ++i) x[i] = b[i];
    for (int64_t j = 0; j < n; ++j) {  // forward_substitution!
        int64_t c_start = P.Di[j] + 1, c_end = A.colptr[j + 1] - 1;
        for (int64_t cp = c_start; cp <= c_end; ++cp) {
            int64_t i = A.rowval[cp];
            x[i] -= A.nzval[cp] * x[j] * P.D[j];
        }
    }
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] * P.D[i];  // backward_substitution!
    for (int64_t i = n - 2; i >= 0; --i)
        for (size_t k = 0; k < P.Ri[i].size(); ++k) x[i] -= A.nzval[P.Ri[i][k]] * x[P.J[i][k]] * P.D[i];
}

static void ilu0_setup(Precond &P) {
    const CSC &A = *P.A;
    const int64_t n = A.n;
    std::vector<IVec> rows(n);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t p = A.colptr[j]; p < A.colptr[j + 1]; ++p) rows[A.rowval[p]].push_back(j);
    P.rp.assign(n + 1, 0);
    P.ci.clear();
    P.dg.assign(n, -1);
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t j : rows[i]) {
            if (j == i) P.dg[i] = (int64_t)P.ci.size();
            P.ci.push_back(j);
        }
        P.rp[i + 1] = (int64_t)P.ci.size();
    }
    P.lu.assign(P.ci.size(), 0.0);
    P.upd_pairs.assign(P.ci.size(), IVec());
    for (int64_t i = 0; i < n; ++i)
        for (int64_t pk = P.rp[i]; pk < P.dg[i]; ++pk) {
            int64_t k = P.ci[pk];
            for (int64_t pj = pk + 1; pj < P.rp[i + 1]; ++pj) {
                int64_t j = P.ci[pj];
                auto b = P.ci.begin() + P.rp[k], e = P.ci.begin() + P.rp[k + 1];
                auto it = std::lower_bound(b, e, j);
                if (it != e && *it == j) {
                    P.upd_pairs[pk].push_back(pj);
                    P.upd_pairs[pk].push_back((int64_t)(it - P.ci.begin()));
                }
            }
        }
}
static void ilu0_update(Precond &P) {  // Saad Alg. 10.4 (IKJ), in place on A's pattern
    const CSC &A = *P.A;
    const int64_t n = A.n;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t p = P.rp[i]; p < P.rp[i + 1]; ++p) P.lu[p] = A.get(i, P.ci[p]);
    for (int64_t i = 1; i < n; ++i)
        for (int64_t pk = P.rp[i]; pk < P.dg[i]; ++pk) {
            int64_t k = P.ci[pk];
            P.lu[pk] = P.lu[pk] / P.lu[P.dg[k]];
            const IVec &u = P.upd_pairs[pk];
            for (size_t q = 0; q < u.size(); q += 2) P.lu[u[q]] = P.lu[u[q]] - P.lu[pk] * P.lu[u[q + 1]];
        }
}
static void ilu0_apply(const Precond &P, Vec &x, const Vec &b) {
    const int64_t n = (int64_t)b.size();
    for (int64_t i = 0; i < n; ++i) {
        double y = b[i];
        for (int64_t p = P.rp[i]; p < P.dg[i]; ++p) y = y - P.lu[p] * x[P.ci[p]];
        x[i] = y;
    }
    for (int64_t i = n - 1; i >= 0; --i) {
        double y = x[i];
        for (int64_t p = P.dg[i] + 1; p < P.rp[i + 1]; ++p) y = y - P.lu[p] * x[P.ci[p]];
        x[i] = y / P.lu[P.dg[i]];
    }
}

// LDL (preconditioners_0_types.jl:36-42): opLDL(A) of the set-up matrix,
// never updated (preconditioners_2_functions.jl:52-57).  Restated from
// LDLFactorizations 0.10.1's published algorithm (Davis' up-looking LDL^T,
// ldl_symbolic / ldl_numeric, then ldl_lsolve / ldl_dsolve / ldl_ltsolve),
// on A's upper triangle, in A's own order (LDLFactorizations orders with
// AMD.jl, which is absent here: ordering and rounding are parity unpinned).
static int ldl_setup(Precond &P) {
    const CSC &A = *P.A;
    const int64_t n = A.n;
    IVec parent(n, -1), flag(n), lnz(n, 0);
    for (int64_t k = 0; k < n; ++k) {  // ldl_symbolic: column k of triu(A)
        flag[k] = k;
        for (int64_t p = A.colptr[k]; p < A.colptr[k + 1]; ++p) {
            int64_t i = A.rowval[p];
            if (i >= k) continue;
            for (; flag[i] != k; i = parent[i]) {
                if (parent[i] == -1) parent[i] = k;
                lnz[i]++;
                flag[i] = k;
            }
        }
    }
    P.Lp.assign(n + 1, 0);
    for (int64_t k = 0; k < n; ++k) P.Lp[k + 1] = P.Lp[k] + lnz[k];
    P.Li.assign(P.Lp[n], 0);
    P.Lx.assign(P.Lp[n], 0.0);
    P.Dl.assign(n, 0.0);
    Vec y(n, 0.0);
    IVec pattern(n), fill(n, 0);
    for (int64_t k = 0; k < n; ++k) {  // ldl_numeric
        int64_t top = n;
        flag[k] = k;
        for (int64_t p = A.colptr[k]; p < A.colptr[k + 1]; ++p) {
            int64_t i = A.rowval[p];
            if (i > k) continue;
            y[i] += A.nzval[p];
            int64_t len = 0;
            for (; flag[i] != k; i = parent[i]) {
                pattern[len++] = i;
                flag[i] = k;
            }
            while (len > 0) pattern[--top] = pattern[--len];
        }
        double d = y[k];
        y[k] = 0.0;
        for (; top < n; ++top) {
            int64_t i = pattern[top];
            double yi = y[i];
            y[i] = 0.0;
            int64_t p2 = P.Lp[i] + fill[i];
            for (int64_t p = P.Lp[i]; p < p2; ++p) y[P.Li[p]] -= P.Lx[p] * yi;
            double l_ki = yi / P.Dl[i];
            d -= l_ki * yi;
            P.Li[p2] = k;
            P.Lx[p2] = l_ki;
            fill[i]++;
        }
        if (d == 0.0 || !std::isfinite(d)) return fail(FVH_ERR_ARG, "LDL: pivot %lld is %g", (long long)k, d);
        P.Dl[k] = d;
    }
    return FVH_OK;
}
static void ldl_apply(const Precond &P, Vec &x, const Vec &b) {
    const int64_t n = (int64_t)b.size();
    x = b;
    for (int64_t j = 0; j < n; ++j)  // ldl_lsolve
        for (int64_t p = P.Lp[j]; p < P.Lp[j + 1]; ++p) x[P.Li[p]] -= P.Lx[p] * x[j];
    for (int64_t j = 0; j < n; ++j) x[j] /= P.Dl[j];  // ldl_dsolve
    for (int64_t j = n - 1; j >= 0; --j)  // ldl_ltsolve
        for (int64_t p = P.Lp[j]; p < P.Lp[j + 1]; ++p) x[j] -= P.Lx[p] * x[P.Li[p]];
}

static void precond_setup(Precond &P, int kind, CSC *A) {
    P.kind = kind;
    P.A = A;
    if (kind == FVH_PRECOND_DILU) dilu_setup(P);
    if (kind == FVH_PRECOND_ILU0) ilu0_setup(P);
    if (kind == FVH_PRECOND_JACOBI || kind == FVH_PRECOND_NORMDIAG) P.storage.assign(A->n, 0.0);
}
static void precond_update(Precond &P) {
    const CSC &A = *P.A;
    if (P.kind == FVH_PRECOND_LDL) {
        if (!P.ldl_ready) {
            P.ldl_status = ldl_setup(P);
            P.ldl_ready = true;
        }
        return;
    }
    if (P.kind == FVH_PRECOND_DILU) dilu_update(P);
    else if (P.kind == FVH_PRECOND_ILU0) ilu0_update(P);
    else if (P.kind == FVH_PRECOND_JACOBI) {
        for (int64_t i = 0; i < A.n; ++i) {
            int64_t idx = 0;
            for (int64_t p = A.colptr[i]; p < A.colptr[i + 1]; ++p)
                if (A.rowval[p] == i) { idx = p; break; }
            P.storage[i] = 1.0 / std::fabs(A.nzval[idx]);
        }
    } else if (P.kind == FVH_PRECOND_NORMDIAG) {
        for (int64_t i = 0; i < A.n; ++i) {  // LinearAlgebra.generic_norm2 of a short column
            double mx = 0.0;
            for (int64_t p = A.colptr[i]; p < A.colptr[i + 1]; ++p) mx = std::max(mx, std::fabs(A.nzval[p]));
            double nrm;
            if (mx == 0.0 || std::isinf(mx)) nrm = mx;
            else {
                double s = 0.0;
                bool first = true;
                for (int64_t p = A.colptr[i]; p < A.colptr[i + 1]; ++p) {
                    double v = A.nzval[p] * A.nzval[p];
                    s = first ? v : s + v;
                    first = false;
                }
                nrm = std::sqrt(s);
            }
            P.storage[i] = 1.0 / nrm;
        }
    }
}
static void precond_apply(const Precond &P, Vec &y, const Vec &x) {
    if (P.kind == FVH_PRECOND_LDL) ldl_apply(P, y, x);
    else if (P.kind == FVH_PRECOND_DILU) dilu_apply(P, y, x);
    else if (P.kind == FVH_PRECOND_ILU0) ilu0_apply(P, y, x);
    else if (P.kind == FVH_PRECOND_JACOBI || P.kind == FVH_PRECOND_NORMDIAG)
        for (size_t i = 0; i < x.size(); ++i) y[i] = P.storage[i] * x[i];
    else y = x;
}

// ---------------------------------------------------------------------------
// Krylov.jl 0.9.5 algorithms (restated), warm started from x
// ---------------------------------------------------------------------------
struct Op {
    const CSC *A;
    const Precond *P;
    bool negate_prec = false;  // -M^-1 (CG on -A with a kept LDL factor of A's set-up matrix)
    void mul(const Vec &x, Vec &y) const { spmv(*A, x, y); }
    void prec(const Vec &x, Vec &y) const {
        precond_apply(*P, y, x);
        if (negate_prec)
            for (double &v : y) v = -v;
    }
};

static fvh_stats bicgstab(const Op &op, const Vec &b, Vec &x, int itmax, double atol, double rtol) {
    const int64_t n = (int64_t)b.size();
    fvh_stats st{};
    Vec dx = x, r0(n), r(n), p(n), v(n, 0.0), s(n, 0.0), q(n), t(n);
    std::fill(x.begin(), x.end(), 0.0);
    op.mul(dx, r0);
    for (int64_t i = 0; i < n; ++i) r0[i] = b[i] - r0[i];
    op.prec(r0, r);
    p = r;
    double alpha = 1.0, omega = 1.0, rho = 1.0;
    double rNorm = nrm2(r);
    st.rnorm0 = rNorm;
    st.rnorm = rNorm;
    if (rNorm == 0.0) {
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + dx[i];
        st.status = FVH_KRYLOV_ZERO_RES;
        return st;
    }
    int iter = 0;
    double eps = atol + rtol * rNorm;
    double next_rho = dot(b, r);
    if (next_rho == 0.0) {
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + dx[i];
        st.status = FVH_KRYLOV_BREAKDOWN;
        return st;
    }
    bool solved = rNorm <= eps, tired = iter >= itmax, breakdown = false;
    while (!(solved || tired || breakdown)) {
        ++iter;
        rho = next_rho;
        op.mul(p, q);
        op.prec(q, v);
        alpha = rho / dot(b, v);
        for (int64_t i = 0; i < n; ++i) s[i] = r[i];
        for (int64_t i = 0; i < n; ++i) s[i] = s[i] + (-alpha) * v[i];
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
        op.mul(s, q);  // d aliases q
        op.prec(q, t);
        omega = dot(t, s) / dot(t, t);
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + omega * s[i];
        for (int64_t i = 0; i < n; ++i) r[i] = s[i];
        for (int64_t i = 0; i < n; ++i) r[i] = r[i] + (-omega) * t[i];
        next_rho = dot(b, r);
        double beta = (next_rho / rho) * (alpha / omega);
        for (int64_t i = 0; i < n; ++i) p[i] = p[i] + (-omega) * v[i];
        for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        rNorm = nrm2(r);
        bool mach = rNorm + 1.0 <= 1.0;
        solved = (rNorm <= eps) || mach;
        tired = iter >= itmax;
        breakdown = (alpha == 0.0) || std::isnan(alpha);
    }
    st.status = solved ? FVH_KRYLOV_SOLVED : (breakdown ? FVH_KRYLOV_BREAKDOWN : FVH_KRYLOV_ITMAX);
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + dx[i];
    st.niter = iter;
    st.rnorm = rNorm;
    return st;
}

static fvh_stats cg(const Op &op, const Vec &b, Vec &x, int itmax, double atol, double rtol) {
    const int64_t n = (int64_t)b.size();
    fvh_stats st{};
    const double epsT = std::numeric_limits<double>::epsilon();
    Vec dx = x, r(n), z(n), p(n), Ap(n);
    std::fill(x.begin(), x.end(), 0.0);
    op.mul(dx, r);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    op.prec(r, z);
    p = z;
    double gamma = dot(r, z);
    double rNorm = std::sqrt(gamma);
    st.rnorm0 = rNorm;
    st.rnorm = rNorm;
    if (gamma == 0.0) {
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + dx[i];
        st.status = FVH_KRYLOV_ZERO_RES;
        return st;
    }
    int iter = 0;
    double pNorm2 = gamma;
    double eps = atol + rtol * rNorm;
    bool solved = rNorm <= eps, tired = iter >= itmax, zero_curv = false;
    while (!(solved || tired || zero_curv)) {
        op.mul(p, Ap);
        double pAp = dot(p, Ap);
        if (pAp <= epsT * pNorm2 && std::fabs(pAp) <= epsT * pNorm2) zero_curv = true;
        if (zero_curv) continue;
        double alpha = gamma / pAp;
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
        for (int64_t i = 0; i < n; ++i) r[i] = r[i] + (-alpha) * Ap[i];
        op.prec(r, z);
        double gamma_next = dot(r, z);
        rNorm = std::sqrt(gamma_next);
        bool mach = rNorm + 1.0 <= 1.0;
        solved = (rNorm <= eps) || mach;
        if (!solved) {
            double beta = gamma_next / gamma;
            pNorm2 = gamma_next + beta * beta * pNorm2;
            gamma = gamma_next;
            for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
        ++iter;
        tired = iter >= itmax;
    }
    st.status = solved ? FVH_KRYLOV_SOLVED : (zero_curv ? FVH_KRYLOV_ZERO_CURV : FVH_KRYLOV_ITMAX);
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + dx[i];
    st.niter = iter;
    st.rnorm = rNorm;
    return st;
}

static void sym_givens(double a, double b, double &c, double &s, double &rho) {
    if (b == 0.0) {
        c = (a == 0.0) ? 1.0 : jsign(a);
        s = 0.0;
        rho = std::fabs(a);
    } else if (a == 0.0) {
        c = 0.0;
        s = jsign(b);
        rho = std::fabs(b);
    } else if (std::fabs(b) > std::fabs(a)) {
        double t = a / b;
        s = jsign(b) / std::sqrt(1.0 + t * t);
        c = s * t;
        rho = b / s;
    } else {
        double t = b / a;
        c = jsign(a) / std::sqrt(1.0 + t * t);
        s = c * t;
        rho = a / c;
    }
}

static fvh_stats gmres(const Op &op, const Vec &b, Vec &x, int itmax, double atol, double rtol) {
    const int64_t n = (int64_t)b.size();
    fvh_stats st{};
    const double btol = std::pow(std::numeric_limits<double>::epsilon(), 0.75);
    Vec dx = x, w(n), r0(n), q(n);
    std::fill(x.begin(), x.end(), 0.0);
    op.mul(dx, w);
    for (int64_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
    op.prec(w, r0);
    double beta = nrm2(r0);
    double rNorm = beta;
    st.rnorm0 = rNorm;
    st.rnorm = rNorm;
    double eps = atol + rtol * rNorm;
    if (beta == 0.0) {
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + dx[i];
        st.status = FVH_KRYLOV_ZERO_RES;
        return st;
    }
    int iter = 0, inner_itmax = itmax;
    bool breakdown = false, solved = rNorm <= eps, tired = iter >= itmax;
    std::vector<Vec> V;
    Vec R, c, s, z;
    while (!(solved || tired || breakdown)) {
        V.assign(1, Vec(n, 0.0));
        R.clear();
        c.clear();
        s.clear();
        z.assign(1, 0.0);
        int nr = 0;
        beta = nrm2(r0);
        z[0] = beta;
        for (int64_t i = 0; i < n; ++i) V[0][i] = r0[i] / rNorm;
        int inner_iter = 0;
        bool inner_tired = false;
        while (!(solved || inner_tired || breakdown)) {
            ++inner_iter;
            const int k = inner_iter;
            R.resize((size_t)(nr + k), 0.0);
            c.resize(k, 0.0);
            s.resize(k, 0.0);
            op.mul(V[k - 1], w);
            op.prec(w, q);
            for (int i = 0; i < k; ++i) {
                R[nr + i] = dot(V[i], q);
                double mr = -R[nr + i];
                for (int64_t j = 0; j < n; ++j) q[j] = q[j] + mr * V[i][j];
            }
            double Hbis = nrm2(q);
            for (int i = 0; i < k - 1; ++i) {
                double Rtmp = c[i] * R[nr + i] + s[i] * R[nr + i + 1];
                R[nr + i + 1] = s[i] * R[nr + i] - c[i] * R[nr + i + 1];
                R[nr + i] = Rtmp;
            }
            sym_givens(R[nr + k - 1], Hbis, c[k - 1], s[k - 1], R[nr + k - 1]);
            double zeta = s[k - 1] * z[k - 1];
            z[k - 1] = c[k - 1] * z[k - 1];
            rNorm = std::fabs(zeta);
            nr += k;
            bool mach = rNorm + 1.0 <= 1.0;
            solved = (rNorm <= eps) || mach;
            breakdown = Hbis <= btol;
            inner_tired = inner_iter >= inner_itmax;
            if (!(solved || inner_tired || breakdown)) {
                V.emplace_back(n);
                for (int64_t i = 0; i < n; ++i) V[k][i] = q[i] / Hbis;
                z.push_back(zeta);
            }
        }
        const int k = inner_iter;
        Vec y(z.begin(), z.begin() + k);
        for (int i = k; i >= 1; --i) {
            int pos = nr + i - k;  // 1-based position of r(i,k)
            for (int j = k; j >= i + 1; --j) {
                y[i - 1] = y[i - 1] - R[pos - 1] * y[j - 1];
                pos = pos - j + 1;
            }
            if (std::fabs(R[pos - 1]) <= btol) y[i - 1] = 0.0;
            else y[i - 1] = y[i - 1] / R[pos - 1];
        }
        for (int i = 0; i < k; ++i)
            for (int64_t j = 0; j < n; ++j) x[j] = x[j] + y[i] * V[i][j];
        inner_itmax -= k;
        iter += k;
        tired = iter >= itmax;
    }
    st.status = solved ? FVH_KRYLOV_SOLVED : (breakdown ? FVH_KRYLOV_BREAKDOWN : FVH_KRYLOV_ITMAX);
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + dx[i];
    st.niter = iter;
    st.rnorm = rNorm;
    return st;
}

// run!(eqn, setup): CG runs on -A x = -b (with the preconditioner of -A) when
// the preconditioner is DILU/ILU(0) and A has a negative first diagonal, the
// only form in which a preconditioned CG is defined for the negative-definite
// pressure matrix.
struct Negated {  // -A and its preconditioner, set up once per equation
    CSC An;
    Precond Pn;
    int kind = -1;
};

static fvh_stats run(Equation &eq, Precond &P, const fvh_solver_cfg &cfg, Vec &x) {
    if (cfg.solver == FVH_CG && (P.kind == FVH_PRECOND_DILU || P.kind == FVH_PRECOND_ILU0) &&
        eq.A.nzval[eq.A.slot(0, 0)] < 0.0) {
        static thread_local std::map<const Equation *, std::unique_ptr<Negated>> cache;
        auto &ng = cache[&eq];
        if (!ng || ng->kind != P.kind || ng->An.n != eq.A.n) {
            ng = std::make_unique<Negated>();
            ng->An = eq.A;
            ng->kind = P.kind;
            precond_setup(ng->Pn, P.kind, &ng->An);
        }
        for (size_t i = 0; i < eq.A.nzval.size(); ++i) ng->An.nzval[i] = -eq.A.nzval[i];
        precond_update(ng->Pn);
        Vec bn(eq.b.size());
        for (size_t i = 0; i < bn.size(); ++i) bn[i] = -eq.b[i];
        Op op{&ng->An, &ng->Pn};
        return cg(op, bn, x, cfg.itmax, cfg.atol, cfg.rtol);
    }
    if (cfg.solver == FVH_CG && P.kind == FVH_PRECOND_LDL && eq.A.nzval[eq.A.slot(0, 0)] < 0.0) {
        // CG on -A: the LDL factor is of A's set-up matrix and stays as it
        // is; -A's preconditioner is its negation
        CSC An = eq.A;
        for (double &v : An.nzval) v = -v;
        Vec bn(eq.b.size());
        for (size_t i = 0; i < bn.size(); ++i) bn[i] = -eq.b[i];
        Op op{&An, &P, true};
        return cg(op, bn, x, cfg.itmax, cfg.atol, cfg.rtol);
    }
    Op op{&eq.A, &P};
    if (cfg.solver == FVH_BICGSTAB) return bicgstab(op, eq.b, x, cfg.itmax, cfg.atol, cfg.rtol);
    if (cfg.solver == FVH_CG) return cg(op, eq.b, x, cfg.itmax, cfg.atol, cfg.rtol);
    return gmres(op, eq.b, x, cfg.itmax, cfg.atol, cfg.rtol);
}

// ---------------------------------------------------------------------------
// Calculate
// ---------------------------------------------------------------------------
static void interpolate(const Mesh &m, const Vec &phi, Vec &phif) {
    for (int64_t f = 0; f < m.nf; ++f) {
        double w = m.fweight[f];
        double omw = 1.0 - w;
        phif[f] = w * phi[m.own1[f]] + omw * phi[m.own2[f]];
    }
}
static void interpolate_vec(const Mesh &m, const Vec &ux, const Vec &uy, Vec &fx, Vec &fy) {
    for (int64_t f = 0; f < m.nf; ++f) {
        double w = m.fweight[f];
        int64_t c1 = m.own1[f], c2 = m.own2[f];
        double omw = 1.0 - w;
        fx[f] = w * ux[c1] + omw * ux[c2];
        fy[f] = w * uy[c1] + omw * uy[c2];
    }
}
// correct_boundaries!: Dirichlet -> value, everything else -> owner cell value
static void correct_boundaries(const Mesh &m, const std::vector<fvh_bc> &bcs, int comp, const Vec &phi, Vec &phif) {
    for (const fvh_bc &bc : bcs) {
        const Boundary &B = m.boundaries[bc.patch];
        for (size_t i = 0; i < B.facesID.size(); ++i)
            phif[B.facesID[i]] = (bc.kind == FVH_BC_DIRICHLET) ? bc.value[comp] : phi[B.cellsID[i]];
    }
}
static void green_gauss(const Mesh &m, const Vec &phif, Vec &dx, Vec &dy, Vec &dz) {
    for (int64_t c = 0; c < m.n; ++c) {
        double rx = 0.0, ry = 0.0, rz = 0.0;
        for (int64_t fi = m.fstart[c]; fi < m.fend[c]; ++fi) {
            int64_t f = m.cell_faces[fi];
            double ns = (double)m.cell_nsign[fi];
            double a = m.farea[f];
            V3 nn = m.fnormal[f];
            double v = phif[f];
            rx += v * (a * nn.x * ns);
            ry += v * (a * nn.y * ns);
            rz += v * (a * nn.z * ns);
        }
        double vol = m.volume[c];
        dx[c] = rx / vol;
        dy[c] = ry / vol;
        dz[c] = rz / vol;
    }
    for (int64_t f = 0; f < m.nbf; ++f) {
        int64_t c = m.own1[f];
        double vol = m.volume[c];
        double a = m.farea[f];
        V3 nn = m.fnormal[f];
        double v = phif[f];
        dx[c] += v * (a * nn.x) / vol;
        dy[c] += v * (a * nn.y) / vol;
        dz[c] += v * (a * nn.z) / vol;
    }
}
// Midpoint gradient: interpolate_midpoint! + correct_interpolation! x2
static void interpolate_midpoint(const Mesh &m, const Vec &phi, Vec &phif) {
    for (int64_t f = 0; f < m.nf; ++f) phif[f] = 0.5 * (phi[m.own1[f]] + phi[m.own2[f]]);
}
static void correct_interpolation(const Mesh &m, const Vec &dx, const Vec &dy, const Vec &dz, Vec &phif, const Vec &phi) {
    const double w = 0.5;
    for (int64_t f = m.nbf; f < m.nf; ++f) {
        int64_t o1 = m.own1[f], o2 = m.own2[f];
        V3 rf = m.fcentre[f], rP = m.ccentre[o1], rN = m.ccentre[o2];
        double phifp = w * (phi[o1] + phi[o2]);
        V3 g{w * (dx[o1] + dx[o2]), w * (dy[o1] + dy[o2]), w * (dz[o1] + dz[o2])};
        V3 Ri{rf.x - w * (rP.x + rN.x), rf.y - w * (rP.y + rN.y), rf.z - w * (rP.z + rN.z)};
        phif[f] = phifp + vdot(g, Ri);
    }
}
static void grad_scalar(const Mesh &m, int scheme, const std::vector<fvh_bc> &bcs, int comp, const Vec &phi, Vec &phif,
                        Vec &dx, Vec &dy, Vec &dz) {
    if (scheme == FVH_MIDPOINT) {
        interpolate_midpoint(m, phi, phif);
        correct_boundaries(m, bcs, comp, phi, phif);
        green_gauss(m, phif, dx, dy, dz);
        for (int it = 0; it < 2; ++it) {
            correct_interpolation(m, dx, dy, dz, phif, phi);
            green_gauss(m, phif, dx, dy, dz);
        }
    } else {
        interpolate(m, phi, phif);
        correct_boundaries(m, bcs, comp, phi, phif);
        green_gauss(m, phif, dx, dy, dz);
    }
}
// ---- non-orthogonal correction (Calculate_3_orthogonality_correction.jl) ----
// weight(::Type{Linear}, cells, faces, fi) (:81-93)
static double nonorth_weight(const Mesh &m, int64_t f) {
    V3 c1 = m.ccentre[m.own1[f]], c2 = m.ccentre[m.own2[f]], centre = m.fcentre[f];
    V3 c1_f = sub(centre, c1), c1_c2 = sub(c2, c1);
    double q = vdot(c1_f, c1_c2) / vdot(c1_c2, c1_c2);
    V3 d12 = sub(c1, c2);
    V3 f_prime{c1.x - q * d12.x, c1.y - q * d12.y, c1.z - q * d12.z};
    return vnorm(sub(c2, f_prime)) / vnorm(sub(c2, c1));
}
// interpolate!(gradf, grad, phi) (Calculate_2_interpolation.jl:152-176), internal faces
static void interpolate_gradient(const Mesh &m, const Vec &gx, const Vec &gy, const Vec &gz, const Vec &phi, Vec &fx,
                                 Vec &fy, Vec &fz) {
    for (int64_t f = m.nbf; f < m.nf; ++f) {
        int64_t c1 = m.own1[f], c2 = m.own2[f];
        double w = nonorth_weight(m, f);
        double omw = 1 - w;
        V3 ave{w * gx[c1] + omw * gx[c2], w * gy[c1] + omw * gy[c2], w * gz[c1] + omw * gz[c2]};
        V3 e = m.fe[f];
        double s = (phi[c2] - phi[c1]) / m.fdelta[f] - vdot(ave, e);
        fx[f] = ave.x + s * e.x;
        fy[f] = ave.y + s * e.y;
        fz[f] = ave.z + s * e.z;
    }
}
// nonorthogonal_flux!(phif, gradf) (:66-79), internal faces
static void nonorthogonal_flux(const Mesh &m, const Vec &fx, const Vec &fy, const Vec &fz, Vec &phif) {
    for (int64_t f = m.nbf; f < m.nf; ++f) {
        V3 n = m.fnormal[f], e = m.fe[f];
        double a = m.farea[f];
        V3 T{(n.x - e.x) * a, (n.y - e.y) * a, (n.z - e.z) * a};
        phif[f] = vdot(V3{fx[f], fy[f], fz[f]}, T);
    }
}
// correct!(eqn, term, corr_flux) (:6-24): every face in id order
static void nonorthogonal_correct(const Mesh &m, int sign, const Vec &J, const Vec &corr, Vec &b) {
    for (int64_t f = 0; f < m.nf; ++f) {
        double correction = (double)(-sign) * J[f] * corr[f];
        b[m.own1[f]] += correction;
        b[m.own2[f]] += correction * (-1.0);
    }
}

static void div(const Mesh &m, const Vec &fx, const Vec &fy, const Vec &fz, Vec &out) {
    for (int64_t c = 0; c < m.n; ++c) {
        double vol = m.volume[c];
        out[c] = 0.0;
        for (int64_t fi = m.fstart[c]; fi < m.fend[c]; ++fi) {
            int64_t f = m.cell_faces[fi];
            double ns = (double)m.cell_nsign[fi];
            double a = m.farea[f];
            V3 nn = m.fnormal[f];
            V3 Sf{a * nn.x, a * nn.y, a * nn.z};
            out[c] += vdot(V3{fx[f], fy[f], fz[f]}, Sf) * ns / vol;
        }
    }
    for (int64_t f = 0; f < m.nbf; ++f) {
        int64_t c = m.own1[f];
        double vol = m.volume[c];
        double a = m.farea[f];
        V3 nn = m.fnormal[f];
        V3 Sf{a * nn.x, a * nn.y, a * nn.z};
        out[c] += vdot(V3{fx[f], fy[f], fz[f]}, Sf) / vol;
    }
}
static void flux(const Mesh &m, const Vec &fx, const Vec &fy, const Vec &fz, Vec &mdot) {
    for (int64_t f = 0; f < m.nf; ++f) {
        double a = m.farea[f];
        V3 nn = m.fnormal[f];
        V3 Sf{a * nn.x, a * nn.y, a * nn.z};
        mdot[f] = vdot(V3{fx[f], fy[f], fz[f]}, Sf);
    }
}

// ---------------------------------------------------------------------------
// Wall functions  (RANSModels_kOmega.jl:200-414)
// ---------------------------------------------------------------------------
struct WallConst {
    double kappa, beta1, cmu, B, E, cmu25, ylam;
};
static WallConst wall_const(const fvh_bc &bc) {
    WallConst w{bc.wf[0], bc.wf[1], bc.wf[2], bc.wf[3], bc.wf[4], 0, 0};
    w.cmu25 = std::pow(w.cmu, 0.25);
    w.ylam = fvh_y_plus_laminar(w.E, w.kappa);
    return w;
}

// ---------------------------------------------------------------------------
// SIMPLE / PISO state
// ---------------------------------------------------------------------------
struct Sim {
    Mesh m;
    fvh_case cs{};
    std::vector<fvh_bc> bU, bp, bk, bw, bnut;
    int iterations = 0, it = 0;
    bool converged = false;
    Vec Ux, Uy, Uz, p, k, omega, nut;
    Vec mdotf, nueff, rDf, divHv, rD, Hx, Hy, Hz, prev, pf;
    Vec gpx, gpy, gpz;
    Vec Ufx, Ufy, Ufz;
    Vec gUxx, gUyx, gUzx, gUxy, gUyy, gUzy, gUxz, gUyz, gUzz;  // Grad(U).result (TensorField)
    Vec kf, wf, nutf, nueffk, nueffw, Dkf, Dwf, Pk, Pw;
    Equation ux, uy, pe;  // k and omega share pe
    Precond Pux, Pp, Pk_, Pw_;
    Model mux, muy, mp, mk, mw;
    Vec Rux, Ruy, Rp;
    std::map<std::string, fvh_stats> stats;
};

static void load_bcs(const fvh_bcset &s, std::vector<fvh_bc> &out) { out.assign(s.bcs, s.bcs + s.nbc); }

static void grad_vector(Sim &S) {  // grad!(gradU, Uf, U, U.BCs), tensor form
    const Mesh &m = S.m;
    int scheme = S.cs.sU.gradient;
    if (scheme == FVH_MIDPOINT) {
        for (int64_t f = 0; f < m.nf; ++f) {  // interpolate_midpoint! (vector): weight*(x1+x2)
            double w = m.fweight[f];
            int64_t c1 = m.own1[f], c2 = m.own2[f];
            S.Ufx[f] = w * (S.Ux[c1] + S.Ux[c2]);
            S.Ufy[f] = w * (S.Uy[c1] + S.Uy[c2]);
            S.Ufz[f] = w * (S.Uz[c1] + S.Uz[c2]);
        }
        for (const fvh_bc &bc : S.bU) {
            const Boundary &B = m.boundaries[bc.patch];
            for (size_t i = 0; i < B.facesID.size(); ++i) {
                int64_t f = B.facesID[i], c = B.cellsID[i];
                bool d = bc.kind == FVH_BC_DIRICHLET;
                S.Ufx[f] = d ? bc.value[0] : S.Ux[c];
                S.Ufy[f] = d ? bc.value[1] : S.Uy[c];
                S.Ufz[f] = d ? bc.value[2] : S.Uz[c];
            }
        }
        for (int it = 0; it < 2; ++it) {
            correct_interpolation(m, S.gUxx, S.gUyx, S.gUzx, S.Ufx, S.Ux);
            green_gauss(m, S.Ufx, S.gUxx, S.gUyx, S.gUzx);
            correct_interpolation(m, S.gUxy, S.gUyy, S.gUzy, S.Ufy, S.Uy);
            green_gauss(m, S.Ufy, S.gUxy, S.gUyy, S.gUzy);
            correct_interpolation(m, S.gUxz, S.gUyz, S.gUzz, S.Ufz, S.Uz);
            green_gauss(m, S.Ufz, S.gUxz, S.gUyz, S.gUzz);
        }
    } else {
        interpolate_vec(m, S.Ux, S.Uy, S.Ufx, S.Ufy);
        for (const fvh_bc &bc : S.bU) {
            const Boundary &B = m.boundaries[bc.patch];
            for (size_t i = 0; i < B.facesID.size(); ++i) {
                int64_t f = B.facesID[i], c = B.cellsID[i];
                bool d = bc.kind == FVH_BC_DIRICHLET;
                S.Ufx[f] = d ? bc.value[0] : S.Ux[c];
                S.Ufy[f] = d ? bc.value[1] : S.Uy[c];
                S.Ufz[f] = d ? bc.value[2] : S.Uz[c];
            }
        }
        green_gauss(m, S.Ufx, S.gUxx, S.gUyx, S.gUzx);
        green_gauss(m, S.Ufy, S.gUxy, S.gUyy, S.gUzy);
        green_gauss(m, S.Ufz, S.gUxz, S.gUyz, S.gUzz);
    }
}

static void velocity_boundaries(Sim &S, const Vec &cx, const Vec &cy, const Vec &cz) {
    const Mesh &m = S.m;
    for (const fvh_bc &bc : S.bU) {
        const Boundary &B = m.boundaries[bc.patch];
        for (size_t i = 0; i < B.facesID.size(); ++i) {
            int64_t f = B.facesID[i], c = B.cellsID[i];
            bool d = bc.kind == FVH_BC_DIRICHLET;
            S.Ufx[f] = d ? bc.value[0] : cx[c];
            S.Ufy[f] = d ? bc.value[1] : cy[c];
            S.Ufz[f] = d ? bc.value[2] : cz[c];
        }
    }
}

static void bound(const Mesh &m, Vec &field, double bnd) {  // bound!: sequential, in place
    const double epsv = std::numeric_limits<double>::epsilon();
    for (int64_t i = 0; i < m.n; ++i) {
        double sum_flux = 0.0;
        int64_t sum_area = 0;
        for (int64_t fi = m.fstart[i]; fi < m.fend[i]; ++fi) {
            sum_flux += jmax(field[m.cell_neighbours[fi]], epsv);
            sum_area += 1;
        }
        double average = sum_flux / (double)sum_area;
        double sb = std::signbit(field[i]) ? average : std::copysign(0.0, average);
        field[i] = jmax(jmax(field[i], sb), bnd);
    }
}

// magnitude2!(Pk, S, scale_factor=2): S = 0.5*(gradU + gradU^T), M[j,k] = d U_k / d x_j
// (RANSModels_functions.jl:5-12,46-60)
static void strain_magnitude2(Sim &S) {
    for (int64_t c = 0; c < S.m.n; ++c) {
        double M[3][3] = {{S.gUxx[c], S.gUxy[c], S.gUxz[c]},
                          {S.gUyx[c], S.gUyy[c], S.gUyz[c]},
                          {S.gUzx[c], S.gUzy[c], S.gUzz[c]}};
        double sum = 0.0;
        for (int j = 0; j < 3; ++j)
            for (int kk = 0; kk < 3; ++kk) {
                double sjk = 0.5 * (M[j][kk] + M[kk][j]);
                sum += sjk * sjk;
            }
        S.Pk[c] = sum * 2.0;
    }
}

// constrain_equation! of every OmegaWallFunction face (RANSModels_kOmega.jl:242-277):
// b += A_cc * omega_wall; A_cc *= 2, once per wall face (a corner cell twice).
static void omega_constrain(const Mesh &m, const std::vector<fvh_bc> &bw, const Vec &k, double nu, Equation &pe) {
    for (const fvh_bc &bc : bw) {
        if (bc.kind != FVH_BC_OMEGAWALL) continue;
        WallConst w = wall_const(bc);
        const Boundary &B = m.boundaries[bc.patch];
        for (size_t i = 0; i < B.cellsID.size(); ++i) {
            int64_t c = B.cellsID[i], f = B.facesID[i];
            double y = m.fdelta[f];
            double wvis = 6.0 * nu / (w.beta1 * (y * y));
            double wlog = std::sqrt(k[c]) / (w.cmu25 * w.kappa * y);
            double yplus = w.cmu25 * y * std::sqrt(k[c]) / nu;
            double wc = yplus > w.ylam ? wlog : wvis;
            pe.b[c] += pe.A.at(c, c) * wc;
            pe.A.at(c, c) += pe.A.at(c, c);
        }
    }
}

static void turbulence(Sim &S) {
    const Mesh &m = S.m;
    const double nu = S.cs.nu;
    const double betaP = 0.09, alpha1 = 0.52, beta1 = 0.072, sigk = 0.5, sigw = 0.5;  // get_coeffs
    strain_magnitude2(S);
    for (int64_t c = 0; c < m.n; ++c) S.Pw[c] = alpha1 * S.Pk[c];
    for (int64_t c = 0; c < m.n; ++c) S.Dwf[c] = beta1 * S.omega[c];
    for (int64_t f = 0; f < m.nf; ++f) S.nueffw[f] = nu + sigw * S.nutf[f];
    for (int64_t c = 0; c < m.n; ++c) S.Dkf[c] = betaP * S.omega[c];
    for (int64_t f = 0; f < m.nf; ++f) S.nueffk[f] = nu + sigk * S.nutf[f];
    for (int64_t c = 0; c < m.n; ++c) S.Pk[c] = S.nut[c] * S.Pk[c];
    for (const fvh_bc &bc : S.bk) {  // correct_production!
        if (bc.kind != FVH_BC_KWALL) continue;
        WallConst w = wall_const(bc);
        const Boundary &B = m.boundaries[bc.patch];
        for (size_t i = 0; i < B.cellsID.size(); ++i) {
            int64_t c = B.cellsID[i], f = B.facesID[i];
            double delta = m.fdelta[f];
            V3 nn = m.fnormal[f];
            double uStar = w.cmu25 * std::sqrt(S.k[c]);
            double dUdy = uStar / (w.kappa * delta);
            double yplus = w.cmu25 * delta * std::sqrt(S.k[c]) / nu;
            double nutw = fvh_nut_wall(nu, yplus, w.kappa, w.E);
            V3 Ud{S.Ux[c] - 0.0, S.Uy[c] - 0.0, S.Uz[c] - 0.0};
            double un = vdot(Ud, nn);
            V3 Up{Ud.x - un * nn.x, Ud.y - un * nn.y, Ud.z - un * nn.z};
            V3 g{Up.x / delta, Up.y / delta, Up.z / delta};
            double mag = std::sqrt((g.x * g.x + g.y * g.y) + g.z * g.z);
            if (yplus > w.ylam) S.Pk[c] = (nu + nutw) * mag * dUdy;
        }
    }
    // omega equation
    S.prev = S.omega;
    discretise(m, S.mw, S.pe, &S.prev, S.cs.dt);
    apply_bcs(m, S.mw, S.pe, S.bw, 0);
    implicit_relaxation(S.pe, S.prev, S.cs.vomega.relax);
    omega_constrain(m, S.bw, S.k, nu, S.pe);
    precond_update(S.Pw_);
    S.stats["omega"] = run(S.pe, S.Pw_, S.cs.vomega, S.omega);
    for (const fvh_bc &bc : S.bw) {  // constrain_boundary! -> set_cell_value!
        if (bc.kind != FVH_BC_OMEGAWALL) continue;
        WallConst w = wall_const(bc);
        const Boundary &B = m.boundaries[bc.patch];
        for (size_t i = 0; i < B.cellsID.size(); ++i) {
            int64_t c = B.cellsID[i], f = B.facesID[i];
            double y = m.fdelta[f];
            double wvis = 6.0 * nu / (w.beta1 * (y * y));
            double wlog = std::sqrt(S.k[c]) / (w.cmu25 * w.kappa * y);
            double yplus = w.cmu25 * y * std::sqrt(S.k[c]) / nu;
            S.omega[c] = yplus > w.ylam ? wlog : wvis;
        }
    }
    bound(m, S.omega, std::numeric_limits<double>::epsilon());
    // k equation
    S.prev = S.k;
    discretise(m, S.mk, S.pe, &S.prev, S.cs.dt);
    apply_bcs(m, S.mk, S.pe, S.bk, 0);
    implicit_relaxation(S.pe, S.prev, S.cs.vk.relax);
    precond_update(S.Pk_);
    S.stats["k"] = run(S.pe, S.Pk_, S.cs.vk, S.k);
    bound(m, S.k, std::numeric_limits<double>::epsilon());
    for (int64_t c = 0; c < m.n; ++c) S.nut[c] = S.k[c] / S.omega[c];
    interpolate(m, S.nut, S.nutf);
    correct_boundaries(m, S.bnut, 0, S.nut, S.nutf);
    for (const fvh_bc &bc : S.bnut) {  // correct_eddy_viscosity!
        if (bc.kind != FVH_BC_NUTWALL) continue;
        WallConst w = wall_const(bc);
        const Boundary &B = m.boundaries[bc.patch];
        for (size_t i = 0; i < B.cellsID.size(); ++i) {
            int64_t c = B.cellsID[i], f = B.facesID[i];
            double delta = m.fdelta[f];
            double yplus = w.cmu25 * delta * std::sqrt(S.k[c]) / nu;
            double nutw = fvh_nut_wall(nu, yplus, w.kappa, w.E);
            if (yplus > w.ylam) S.nutf[f] = nutw;
        }
    }
}

static void update_nueff(Sim &S) {
    bool turb = S.cs.model == FVH_KOMEGA;
    for (int64_t f = 0; f < S.m.nf; ++f) S.nueff[f] = turb ? S.cs.nu + S.nutf[f] : S.cs.nu;
}

static int sim_init(Sim &S) {
    Mesh &m = S.m;
    const int64_t n = m.n, nf = m.nf;
    auto cellv = [&](Vec &v) { v.assign(n, 0.0); };
    auto facev = [&](Vec &v) { v.assign(nf, 0.0); };
    for (Vec *v : {&S.rD, &S.Hx, &S.Hy, &S.Hz, &S.prev, &S.divHv, &S.gpx, &S.gpy, &S.gpz, &S.gUxx, &S.gUyx, &S.gUzx,
                   &S.gUxy, &S.gUyy, &S.gUzy, &S.gUxz, &S.gUyz, &S.gUzz, &S.Dkf, &S.Dwf, &S.Pk, &S.Pw})
        cellv(*v);
    for (Vec *v : {&S.mdotf, &S.nueff, &S.rDf, &S.pf, &S.Ufx, &S.Ufy, &S.Ufz, &S.kf, &S.wf, &S.nutf, &S.nueffk, &S.nueffw})
        facev(*v);
    std::fill(S.rDf.begin(), S.rDf.end(), 1.0);
    for (Equation *e : {&S.ux, &S.uy, &S.pe}) {
        e->A = make_pattern(m);
        e->b.assign(n, 0.0);
        e->R.assign(n, 0.0);
        e->Fx.assign(n, 0.0);
    }
    const fvh_case &c = S.cs;
    S.mux.terms = {{FVH_OP_TIME, c.sU.time, 1, nullptr, &S.Ux},
                   {FVH_OP_DIVERGENCE, c.sU.divergence, 1, &S.mdotf, &S.Ux},
                   {FVH_OP_LAPLACIAN, c.sU.laplacian, -1, &S.nueff, &S.Ux}};
    S.mux.sources = {{-1, &S.gpx}};
    S.muy.terms = {{FVH_OP_TIME, c.sU.time, 1, nullptr, &S.Uy},
                   {FVH_OP_DIVERGENCE, c.sU.divergence, 1, &S.mdotf, &S.Uy},
                   {FVH_OP_LAPLACIAN, c.sU.laplacian, -1, &S.nueff, &S.Uy}};
    S.muy.sources = {{-1, &S.gpy}};
    S.mp.terms = {{FVH_OP_LAPLACIAN, c.sp.laplacian, 1, &S.rDf, &S.p}};
    S.mp.sources = {{1, &S.divHv}};
    for (const Term &t : S.mux.terms)
        if (t.op == FVH_OP_LAPLACIAN && t.scheme != FVH_LINEAR) return fail(FVH_ERR_UNSUPPORTED, "Laplacian scheme");
    // set_preconditioner: discretise with prev = 0, apply BCs, build and update
    Vec zero(n, 0.0);
    discretise(m, S.mux, S.ux, &zero, c.dt);
    if (int rc = apply_bcs(m, S.mux, S.ux, S.bU, 0)) return rc;
    precond_setup(S.Pux, c.vU.precond, &S.ux.A);
    precond_update(S.Pux);
    if (S.Pux.ldl_status != FVH_OK) return S.Pux.ldl_status;
    discretise(m, S.mp, S.pe, &zero, c.dt);
    if (int rc = apply_bcs(m, S.mp, S.pe, S.bp, 0)) return rc;
    precond_setup(S.Pp, c.vp.precond, &S.pe.A);
    precond_update(S.Pp);
    if (S.Pp.ldl_status != FVH_OK) return S.Pp.ldl_status;
    if (c.model == FVH_KOMEGA) {
        S.mk.terms = {{FVH_OP_TIME, c.sk.time, 1, nullptr, &S.k},
                      {FVH_OP_DIVERGENCE, c.sk.divergence, 1, &S.mdotf, &S.k},
                      {FVH_OP_LAPLACIAN, c.sk.laplacian, -1, &S.nueffk, &S.k},
                      {FVH_OP_SI, 0, 1, &S.Dkf, &S.k}};
        S.mk.sources = {{1, &S.Pk}};
        S.mw.terms = {{FVH_OP_TIME, c.somega.time, 1, nullptr, &S.omega},
                      {FVH_OP_DIVERGENCE, c.somega.divergence, 1, &S.mdotf, &S.omega},
                      {FVH_OP_LAPLACIAN, c.somega.laplacian, -1, &S.nueffw, &S.omega},
                      {FVH_OP_SI, 0, 1, &S.Dwf, &S.omega}};
        S.mw.sources = {{1, &S.Pw}};
        discretise(m, S.mk, S.pe, &zero, c.dt);
        if (int rc = apply_bcs(m, S.mk, S.pe, S.bk, 0)) return rc;
        precond_setup(S.Pk_, c.vk.precond, &S.pe.A);
        precond_update(S.Pk_);
        if (S.Pk_.ldl_status != FVH_OK) return S.Pk_.ldl_status;
        discretise(m, S.mw, S.pe, &zero, c.dt);
        if (int rc = apply_bcs(m, S.mw, S.pe, S.bw, 0)) return rc;
        precond_setup(S.Pw_, c.vomega.precond, &S.pe.A);
        precond_update(S.Pw_);
        if (S.Pw_.ldl_status != FVH_OK) return S.Pw_.ldl_status;
    }
    S.Rux.assign(S.iterations, 1.0);
    S.Ruy.assign(S.iterations, 1.0);
    S.Rp.assign(S.iterations, 1.0);
    // loop prologue
    interpolate_vec(m, S.Ux, S.Uy, S.Ufx, S.Ufy);
    velocity_boundaries(S, S.Ux, S.Uy, S.Uz);
    flux(m, S.Ufx, S.Ufy, S.Ufz, S.mdotf);
    grad_scalar(m, c.sp.gradient, S.bp, 0, S.p, S.pf, S.gpx, S.gpy, S.gpz);
    update_nueff(S);
    return FVH_OK;
}

static void momentum_solve(Sim &S, Equation &eq, Model &md, Vec &U, int comp, bool relax, double *R, const char *name) {
    S.prev = U;
    discretise(S.m, md, eq, &S.prev, S.cs.dt);
    apply_bcs(S.m, md, eq, S.bU, comp);
    if (relax) implicit_relaxation(eq, S.prev, S.cs.vU.relax);
    precond_update(S.Pux);
    S.stats[name] = run(eq, S.Pux, S.cs.vU, U);
    *R = residual(eq, U);
}

static void pressure_corrector(Sim &S, int iter) {
    Mesh &m = S.m;
    // H!(Hv, U, ux_eqn, uy_eqn)
    for (int64_t c = 0; c < m.n; ++c) {
        double sx = 0.0, sy = 0.0;
        for (int64_t fi = m.fstart[c]; fi < m.fend[c]; ++fi) {
            int64_t nID = m.cell_neighbours[fi];
            sx += S.ux.A.get(c, nID) * S.Ux[nID];
            sy += S.uy.A.get(c, nID) * S.Uy[nID];
        }
        double D = S.ux.A.get(c, c);
        double rD = 1.0 / D;
        S.Hx[c] = (S.ux.b[c] - sx) * rD;
        S.Hy[c] = (S.uy.b[c] - sy) * rD;
        S.Hz[c] = 0.0;
    }
    interpolate_vec(m, S.Hx, S.Hy, S.Ufx, S.Ufy);
    velocity_boundaries(S, S.Hx, S.Hy, S.Hz);
    div(m, S.Ufx, S.Ufy, S.Ufz, S.divHv);
    S.prev = S.p;
    discretise(m, S.mp, S.pe, &S.prev, S.cs.dt);
    apply_bcs(m, S.mp, S.pe, S.bp, 0);
    set_reference(S.pe, S.cs.has_pref != 0, S.cs.pref, 0);
    precond_update(S.Pp);
    S.stats["p"] = run(S.pe, S.Pp, S.cs.vp, S.p);
    explicit_relaxation(S.p, S.prev, S.cs.vp.relax);
    S.Rp[iter] = residual(S.pe, S.p);
    grad_scalar(m, S.cs.sp.gradient, S.bp, 0, S.p, S.pf, S.gpx, S.gpy, S.gpz);
    // the `correct` block (Solvers_1_SIMPLE.jl:165-180, Solvers_2_PISO.jl:170-182)
    for (int i = 0; i < S.cs.nonorth_correctors; ++i) {
        discretise(m, S.mp, S.pe, &S.prev, S.cs.dt);
        apply_bcs(m, S.mp, S.pe, S.bp, 0);
        set_reference(S.pe, S.cs.has_pref != 0, S.cs.pref, 0);
        Vec gfx(m.nf, 0.0), gfy(m.nf, 0.0), gfz(m.nf, 0.0);
        interpolate_gradient(m, S.gpx, S.gpy, S.gpz, S.p, gfx, gfy, gfz);
        nonorthogonal_flux(m, gfx, gfy, gfz, S.pf);
        nonorthogonal_correct(m, 1, S.rDf, S.pf, S.pe.b);
        S.stats["p"] = run(S.pe, S.Pp, S.cs.vp, S.p);
        grad_scalar(m, S.cs.sp.gradient, S.bp, 0, S.p, S.pf, S.gpx, S.gpy, S.gpz);
    }
    for (int64_t c = 0; c < m.n; ++c) {  // correct_velocity!
        double r = S.rD[c];
        S.Ux[c] = S.Hx[c] - S.gpx[c] * r;
        S.Uy[c] = S.Hy[c] - S.gpy[c] * r;
    }
    interpolate_vec(m, S.Ux, S.Uy, S.Ufx, S.Ufy);
    velocity_boundaries(S, S.Ux, S.Uy, S.Uz);
    flux(m, S.Ufx, S.Ufy, S.Ufz, S.mdotf);
    if (S.cs.model == FVH_KOMEGA) {
        grad_vector(S);
        turbulence(S);
        update_nueff(S);
    }
}

static void iterate_once(Sim &S) {
    Mesh &m = S.m;
    const int iter = S.it;
    const bool piso = S.cs.algorithm == FVH_PISO;
    momentum_solve(S, S.ux, S.mux, S.Ux, 0, !piso, &S.Rux[iter], "Ux");
    momentum_solve(S, S.uy, S.muy, S.Uy, 1, !piso, &S.Ruy[iter], "Uy");
    for (int64_t c = 0; c < m.n; ++c) S.rD[c] = m.volume[c] / S.ux.A.get(c, c);  // inverse_diagonal!
    interpolate(m, S.rD, S.rDf);
    for (int64_t c = 0; c < m.n; ++c) {  // remove_pressure_source! (source sign -1)
        double vol = m.volume[c];
        S.ux.b[c] -= -1.0 * S.gpx[c] * vol;
        S.uy.b[c] -= -1.0 * S.gpy[c] * vol;
    }
    int ncorr = piso ? 2 : 1;
    for (int i = 0; i < ncorr; ++i) pressure_corrector(S, iter);
    S.it++;
    double cb = S.cs.conv_break;
    if (!piso && S.Rux[iter] <= cb && S.Ruy[iter] <= cb && S.Rp[iter] <= cb) S.converged = true;
}

static void csr_values(const CSC &A, Vec &out) {  // A values in row-major (CSR) order
    std::vector<std::vector<std::pair<int64_t, double>>> rows(A.n);
    for (int64_t j = 0; j < A.n; ++j)
        for (int64_t p = A.colptr[j]; p < A.colptr[j + 1]; ++p) rows[A.rowval[p]].push_back({j, A.nzval[p]});
    out.clear();
    for (auto &r : rows) {
        std::sort(r.begin(), r.end());
        for (auto &e : r) out.push_back(e.second);
    }
}

}  // namespace ref

using namespace ref;

struct ref_sim {
    Sim S;
};

extern "C" {

const char *ref_last_error(void) { return g_err; }

int ref_sim_create(const fvh_mesh_desc *d, const fvh_case *c, const double *Ux, const double *Uy, const double *Uz,
                   const double *p, const double *k, const double *omega, const double *nut, ref_sim **out) {
    if (!d || !c || !out) return fail(FVH_ERR_ARG, "null");
    auto *h = new ref_sim();
    Sim &S = h->S;
    load_mesh(d, S.m);
    S.cs = *c;
    S.iterations = c->iterations;
    load_bcs(c->U, S.bU);
    load_bcs(c->p, S.bp);
    load_bcs(c->k, S.bk);
    load_bcs(c->omega, S.bw);
    load_bcs(c->nut, S.bnut);
    const int64_t b = d->index_base;
    for (auto *bs : {&S.bU, &S.bp, &S.bk, &S.bw, &S.bnut})
        for (auto &bc : *bs) bc.patch -= (int32_t)b;
    const int64_t n = S.m.n;
    auto cp = [&](Vec &v, const double *src, double def) {
        v.assign(n, def);
        if (src) std::copy(src, src + n, v.begin());
    };
    cp(S.Ux, Ux, 0.0);
    cp(S.Uy, Uy, 0.0);
    cp(S.Uz, Uz, 0.0);
    cp(S.p, p, 0.0);
    cp(S.k, k, 0.0);
    cp(S.omega, omega, 0.0);
    cp(S.nut, nut, 0.0);
    int rc = sim_init(S);
    if (rc) { delete h; return rc; }
    *out = h;
    return FVH_OK;
}

void ref_sim_destroy(ref_sim *h) { delete h; }

int ref_sim_iterate(ref_sim *h, int32_t n, int32_t *done) {
    Sim &S = h->S;
    int32_t k = 0;
    while (k < n && S.it < S.iterations && !S.converged) {
        iterate_once(S);
        ++k;
    }
    if (done) *done = k;
    return FVH_OK;
}

int ref_sim_residuals(ref_sim *h, double *a, double *b, double *c) {
    Sim &S = h->S;
    std::copy(S.Rux.begin(), S.Rux.end(), a);
    std::copy(S.Ruy.begin(), S.Ruy.end(), b);
    std::copy(S.Rp.begin(), S.Rp.end(), c);
    return FVH_OK;
}

int ref_sim_iterations_done(ref_sim *h) { return h->S.it; }

int ref_sim_stats(ref_sim *h, const char *field, fvh_stats *out) {
    auto it = h->S.stats.find(field);
    if (it == h->S.stats.end()) return fail(FVH_ERR_ARG, "no stats for %s", field);
    *out = it->second;
    return FVH_OK;
}

int ref_sim_get(ref_sim *h, const char *name, double *out, int64_t *len) {
    Sim &S = h->S;
    std::string nm(name);
    Vec tmp;
    const Vec *v = nullptr;
    std::map<std::string, const Vec *> tab = {
        {"Ux", &S.Ux}, {"Uy", &S.Uy}, {"Uz", &S.Uz}, {"p", &S.p}, {"k", &S.k}, {"omega", &S.omega}, {"nut", &S.nut},
        {"mdotf", &S.mdotf}, {"nueff", &S.nueff}, {"rDf", &S.rDf}, {"divHv", &S.divHv}, {"rD", &S.rD},
        {"Hv_x", &S.Hx}, {"Hv_y", &S.Hy}, {"gradp_x", &S.gpx}, {"gradp_y", &S.gpy}, {"gradp_z", &S.gpz},
        {"pf", &S.pf}, {"Uf_x", &S.Ufx}, {"Uf_y", &S.Ufy}, {"Uf_z", &S.Ufz}, {"ux_b", &S.ux.b}, {"uy_b", &S.uy.b},
        {"p_b", &S.pe.b}, {"nutf", &S.nutf}, {"Pk", &S.Pk}, {"gradU_xx", &S.gUxx}, {"gradU_yx", &S.gUyx},
        {"gradU_xy", &S.gUxy}, {"gradU_yy", &S.gUyy}, {"nueffk", &S.nueffk}, {"nueffw", &S.nueffw}};
    auto it = tab.find(nm);
    if (it != tab.end()) v = it->second;
    else if (nm == "ux_A") { csr_values(S.ux.A, tmp); v = &tmp; }
    else if (nm == "uy_A") { csr_values(S.uy.A, tmp); v = &tmp; }
    else if (nm == "p_A") { csr_values(S.pe.A, tmp); v = &tmp; }
    else if (nm == "ux_dinv" && S.Pux.kind == FVH_PRECOND_DILU) { v = &S.Pux.D; }
    else if (nm == "p_dinv" && S.Pp.kind == FVH_PRECOND_DILU) { v = &S.Pp.D; }
    else return fail(FVH_ERR_ARG, "unknown array %s", name);
    if (len) *len = (int64_t)v->size();
    if (out) std::copy(v->begin(), v->end(), out);
    return FVH_OK;
}

// ---- operator-level entry points used by the unit tests ----

// The wall-function elementary functions, shared with the GPU build.
double ref_log(double x) { return fvh_log(x); }
double ref_y_plus_laminar(double E, double kappa) { return fvh_y_plus_laminar(E, kappa); }
double ref_nut_wall(double nu, double yplus, double kappa, double E) { return fvh_nut_wall(nu, yplus, kappa, E); }
double ref_pow(double x, double y) { return std::pow(x, y); }

double ref_det_dot(int64_t n, const double *x, const double *y) {
    return det_sum(n, [&](int64_t i) { return x[i] * y[i]; });
}

// A on the mesh pattern: builds the CSC, fills it from CSR-ordered values,
// runs one preconditioned Krylov solve and returns x (in/out) and stats.
int ref_solve_csr(const fvh_mesh_desc *d, const double *csr_vals, const double *b, const fvh_solver_cfg *cfg,
                  double *x, fvh_stats *st, double *dinv_out) {
    Mesh m;
    load_mesh(d, m);
    Equation eq;
    eq.A = make_pattern(m);
    // CSR order -> CSC slots
    std::vector<std::vector<int64_t>> rows(m.n);
    for (int64_t j = 0; j < m.n; ++j)
        for (int64_t p = eq.A.colptr[j]; p < eq.A.colptr[j + 1]; ++p) rows[eq.A.rowval[p]].push_back(p);
    int64_t q = 0;
    for (int64_t i = 0; i < m.n; ++i) {
        std::vector<std::pair<int64_t, int64_t>> r;
        for (int64_t p : rows[i]) {
            int64_t col = std::upper_bound(eq.A.colptr.begin(), eq.A.colptr.end(), p) - eq.A.colptr.begin() - 1;
            r.push_back({col, p});
        }
        std::sort(r.begin(), r.end());
        for (auto &e : r) eq.A.nzval[e.second] = csr_vals[q++];
    }
    eq.b.assign(b, b + m.n);
    eq.R.assign(m.n, 0.0);
    eq.Fx.assign(m.n, 0.0);
    Precond P;
    precond_setup(P, cfg->precond, &eq.A);
    precond_update(P);
    if (dinv_out && P.kind == FVH_PRECOND_DILU) std::copy(P.D.begin(), P.D.end(), dinv_out);
    Vec xv(x, x + m.n);
    fvh_stats s = run(eq, P, *cfg, xv);
    std::copy(xv.begin(), xv.end(), x);
    if (st) *st = s;
    return FVH_OK;
}

// y = M^{-1} x for the preconditioner of the CSR-ordered matrix.
int ref_precond_apply_csr(const fvh_mesh_desc *d, const double *csr_vals, int32_t kind, const double *x, double *y) {
    Mesh m;
    load_mesh(d, m);
    Equation eq;
    eq.A = make_pattern(m);
    std::vector<std::vector<std::pair<int64_t, int64_t>>> rows(m.n);
    for (int64_t j = 0; j < m.n; ++j)
        for (int64_t p = eq.A.colptr[j]; p < eq.A.colptr[j + 1]; ++p) rows[eq.A.rowval[p]].push_back({j, p});
    int64_t q = 0;
    for (int64_t i = 0; i < m.n; ++i) {
        std::sort(rows[i].begin(), rows[i].end());
        for (auto &e : rows[i]) eq.A.nzval[e.second] = csr_vals[q++];
    }
    Precond P;
    precond_setup(P, kind, &eq.A);
    precond_update(P);
    if (P.ldl_status != FVH_OK) return P.ldl_status;
    Vec xv(x, x + m.n), yv(m.n);
    precond_apply(P, yv, xv);
    std::copy(yv.begin(), yv.end(), y);
    return FVH_OK;
}

// y = A x by the column scatter of SparseArrays.mul!
int ref_spmv_csr(const fvh_mesh_desc *d, const double *csr_vals, const double *x, double *y) {
    Mesh m;
    load_mesh(d, m);
    CSC A = make_pattern(m);
    std::vector<std::vector<std::pair<int64_t, int64_t>>> rows(m.n);
    for (int64_t j = 0; j < m.n; ++j)
        for (int64_t p = A.colptr[j]; p < A.colptr[j + 1]; ++p) rows[A.rowval[p]].push_back({j, p});
    int64_t q = 0;
    for (int64_t i = 0; i < m.n; ++i) {
        std::sort(rows[i].begin(), rows[i].end());
        for (auto &e : rows[i]) A.nzval[e.second] = csr_vals[q++];
    }
    Vec xv(x, x + m.n), yv(m.n);
    spmv(A, xv, yv);
    std::copy(yv.begin(), yv.end(), y);
    return FVH_OK;
}

// Scalar gradient grad!(grad, phif, phi, BCs) on its own.
int ref_grad(const fvh_mesh_desc *d, int32_t scheme, const fvh_bc *bcs, int32_t nbc, int32_t comp, const double *phi,
             double *phif, double *dx, double *dy, double *dz) {
    Mesh m;
    load_mesh(d, m);
    std::vector<fvh_bc> b(bcs, bcs + nbc);
    for (auto &bc : b) bc.patch -= (int32_t)d->index_base;
    Vec ph(phi, phi + m.n), pf(m.nf, 0.0), gx(m.n, 0.0), gy(m.n, 0.0), gz(m.n, 0.0);
    grad_scalar(m, scheme, b, comp, ph, pf, gx, gy, gz);
    std::copy(pf.begin(), pf.end(), phif);
    std::copy(gx.begin(), gx.end(), dx);
    std::copy(gy.begin(), gy.end(), dy);
    std::copy(gz.begin(), gz.end(), dz);
    return FVH_OK;
}

// CSC matrix on the mesh pattern from values in CSR order.
static void csc_from_csr(const Mesh &m, CSC &A, const double *csr_vals) {
    A = make_pattern(m);
    std::vector<std::vector<std::pair<int64_t, int64_t>>> rows(m.n);
    for (int64_t j = 0; j < m.n; ++j)
        for (int64_t p = A.colptr[j]; p < A.colptr[j + 1]; ++p) rows[A.rowval[p]].push_back({j, p});
    int64_t q = 0;
    for (int64_t i = 0; i < m.n; ++i) {
        std::sort(rows[i].begin(), rows[i].end());
        for (auto &e : rows[i]) A.nzval[e.second] = csr_vals[q++];
    }
}

// bound!(field, bound): the sequential in-place pass, field updated in place.
int ref_bound(const fvh_mesh_desc *d, double *field, double bnd) {
    Mesh m;
    load_mesh(d, m);
    Vec f(field, field + m.n);
    bound(m, f, bnd);
    std::copy(f.begin(), f.end(), field);
    return FVH_OK;
}

// grad!(gradU, Uf, U, U.BCs) (scheme `gradient`) then magnitude2!(Pk, S, 2):
// the strain-rate production of the k equation before the nut factor.
int ref_strain_production(const fvh_mesh_desc *d, int32_t gradient, const fvh_bc *bcs, int32_t nbc, const double *Ux,
                          const double *Uy, const double *Uz, double *Pk) {
    Sim S;
    load_mesh(d, S.m);
    const int64_t n = S.m.n, nf = S.m.nf;
    S.cs.sU.gradient = gradient;
    S.bU.assign(bcs, bcs + nbc);
    for (auto &bc : S.bU) bc.patch -= (int32_t)d->index_base;
    S.Ux.assign(Ux, Ux + n);
    S.Uy.assign(Uy, Uy + n);
    S.Uz.assign(Uz, Uz + n);
    for (Vec *v : {&S.Ufx, &S.Ufy, &S.Ufz}) v->assign(nf, 0.0);
    for (Vec *v : {&S.gUxx, &S.gUyx, &S.gUzx, &S.gUxy, &S.gUyy, &S.gUzy, &S.gUxz, &S.gUyz, &S.gUzz, &S.Pk})
        v->assign(n, 0.0);
    grad_vector(S);
    strain_magnitude2(S);
    std::copy(S.Pk.begin(), S.Pk.end(), Pk);
    return FVH_OK;
}

// constrain_equation! of the OmegaWallFunction BCs on a matrix given in CSR
// order (values and b updated in place), with the k field and viscosity nu.
int ref_omega_constrain(const fvh_mesh_desc *d, const fvh_bc *bcs, int32_t nbc, const double *k, double nu,
                        double *csr_vals, double *b) {
    Mesh m;
    load_mesh(d, m);
    std::vector<fvh_bc> bw(bcs, bcs + nbc);
    for (auto &bc : bw) bc.patch -= (int32_t)d->index_base;
    Equation eq;
    csc_from_csr(m, eq.A, csr_vals);
    eq.b.assign(b, b + m.n);
    omega_constrain(m, bw, Vec(k, k + m.n), nu, eq);
    Vec vals;
    csr_values(eq.A, vals);
    std::copy(vals.begin(), vals.end(), csr_vals);
    std::copy(eq.b.begin(), eq.b.end(), b);
    return FVH_OK;
}

// ---- per-function entry points (each mirrors one fvmhip.h Calculate /
// Equation call; host arrays, zero-based patches after index_base) ----

int ref_interpolate(const fvh_mesh_desc *d, const double *phi, double *phif) {
    Mesh m;
    load_mesh(d, m);
    Vec ph(phi, phi + m.n), pf(m.nf, 0.0);
    interpolate(m, ph, pf);
    std::copy(pf.begin(), pf.end(), phif);
    return FVH_OK;
}

int ref_interpolate_vector(const fvh_mesh_desc *d, const double *ux, const double *uy, double *fx, double *fy) {
    Mesh m;
    load_mesh(d, m);
    Vec x(ux, ux + m.n), y(uy, uy + m.n), a(fx, fx + m.nf), b(fy, fy + m.nf);
    interpolate_vec(m, x, y, a, b);
    std::copy(a.begin(), a.end(), fx);
    std::copy(b.begin(), b.end(), fy);
    return FVH_OK;
}

// phif is updated in place (faces of the named patches only).
int ref_correct_boundaries(const fvh_mesh_desc *d, const fvh_bc *bcs, int32_t nbc, int32_t comp, const double *phi,
                           double *phif) {
    Mesh m;
    load_mesh(d, m);
    std::vector<fvh_bc> b(bcs, bcs + nbc);
    for (auto &bc : b) bc.patch -= (int32_t)d->index_base;
    Vec ph(phi, phi + m.n), pf(phif, phif + m.nf);
    correct_boundaries(m, b, comp, ph, pf);
    std::copy(pf.begin(), pf.end(), phif);
    return FVH_OK;
}

int ref_green_gauss(const fvh_mesh_desc *d, const double *phif, double *dx, double *dy, double *dz) {
    Mesh m;
    load_mesh(d, m);
    Vec pf(phif, phif + m.nf), gx(m.n), gy(m.n), gz(m.n);
    green_gauss(m, pf, gx, gy, gz);
    std::copy(gx.begin(), gx.end(), dx);
    std::copy(gy.begin(), gy.end(), dy);
    std::copy(gz.begin(), gz.end(), dz);
    return FVH_OK;
}

int ref_interpolate_gradient(const fvh_mesh_desc *d, const double *gx, const double *gy, const double *gz,
                             const double *phi, double *fx, double *fy, double *fz) {
    Mesh m;
    load_mesh(d, m);
    Vec a(gx, gx + m.n), b(gy, gy + m.n), c(gz, gz + m.n), ph(phi, phi + m.n);
    Vec ox(fx, fx + m.nf), oy(fy, fy + m.nf), oz(fz, fz + m.nf);
    interpolate_gradient(m, a, b, c, ph, ox, oy, oz);
    std::copy(ox.begin(), ox.end(), fx);
    std::copy(oy.begin(), oy.end(), fy);
    std::copy(oz.begin(), oz.end(), fz);
    return FVH_OK;
}

int ref_nonorthogonal_flux(const fvh_mesh_desc *d, const double *fx, const double *fy, const double *fz,
                           double *phif) {
    Mesh m;
    load_mesh(d, m);
    Vec a(fx, fx + m.nf), b(fy, fy + m.nf), c(fz, fz + m.nf), o(phif, phif + m.nf);
    nonorthogonal_flux(m, a, b, c, o);
    std::copy(o.begin(), o.end(), phif);
    return FVH_OK;
}

int ref_nonorthogonal_correct(const fvh_mesh_desc *d, int32_t sign, const double *J, const double *corr, double *b) {
    Mesh m;
    load_mesh(d, m);
    Vec j(J, J + m.nf), c(corr, corr + m.nf), bb(b, b + m.n);
    nonorthogonal_correct(m, sign, j, c, bb);
    std::copy(bb.begin(), bb.end(), b);
    return FVH_OK;
}

// ---- post-processing (Postprocess_forces.jl, Calculate_5_surface_normal_gradient.jl) ----
// surface_flux (:3-18): Sf = area*normal; flux = phi[cID]*Sf
static void surface_flux(const Mesh &m, int patch, const Vec &phi, Vec &x, Vec &y, Vec &z) {
    const Boundary &B = m.boundaries[patch];
    const size_t nb = B.facesID.size();
    x.assign(nb, 0.0);
    y.assign(nb, 0.0);
    z.assign(nb, 0.0);
    for (size_t i = 0; i < nb; ++i) {
        int64_t c = B.cellsID[i], f = B.facesID[i];
        double a = m.farea[f];
        V3 n = m.fnormal[f];
        V3 Sf{a * n.x, a * n.y, a * n.z};
        x[i] = phi[c] * Sf.x;
        y[i] = phi[c] * Sf.y;
        z[i] = phi[c] * Sf.z;
    }
}
// surface_normal_gradient (:20-36): Up = Udiff - (Udiff . n) n; grad = Up / delta
static void surface_normal_gradient(const Mesh &m, int patch, const Vec &ux, const Vec &uy, const Vec &uz,
                                    const double *uw, Vec &x, Vec &y, Vec &z) {
    const Boundary &B = m.boundaries[patch];
    const size_t nb = B.facesID.size();
    x.assign(nb, 0.0);
    y.assign(nb, 0.0);
    z.assign(nb, 0.0);
    for (size_t i = 0; i < nb; ++i) {
        int64_t c = B.cellsID[i], f = B.facesID[i];
        V3 n = m.fnormal[f];
        V3 ud{ux[c] - uw[0], uy[c] - uw[1], uz[c] - uw[2]};
        double dn = vdot(ud, n);
        V3 up{ud.x - dn * n.x, ud.y - dn * n.y, ud.z - dn * n.z};
        double delta = m.fdelta[f];
        x[i] = up.x / delta;
        y[i] = up.y / delta;
        z[i] = up.z / delta;
    }
}

int ref_surface_flux(const fvh_mesh_desc *d, int32_t patch, const double *phi, double *fx, double *fy, double *fz) {
    Mesh m;
    load_mesh(d, m);
    if (patch < 0 || patch >= (int)m.boundaries.size()) return fail(FVH_ERR_ARG, "patch %d", patch);
    Vec ph(phi, phi + m.n), x, y, z;
    surface_flux(m, patch, ph, x, y, z);
    std::copy(x.begin(), x.end(), fx);
    std::copy(y.begin(), y.end(), fy);
    std::copy(z.begin(), z.end(), fz);
    return FVH_OK;
}

int ref_surface_normal_gradient(const fvh_mesh_desc *d, int32_t patch, const double *ux, const double *uy,
                                const double *uz, const double *uw, double *gx, double *gy, double *gz) {
    Mesh m;
    load_mesh(d, m);
    if (patch < 0 || patch >= (int)m.boundaries.size()) return fail(FVH_ERR_ARG, "patch %d", patch);
    Vec a(ux, ux + m.n), b(uy, uy + m.n), c(uz, uz + m.n), x, y, z;
    surface_normal_gradient(m, patch, a, b, c, uw, x, y, z);
    std::copy(x.begin(), x.end(), gx);
    std::copy(y.begin(), y.end(), gy);
    std::copy(z.begin(), z.end(), gz);
    return FVH_OK;
}

// pressure_force (Postprocess_forces.jl:5-27): sums in patch face order, then rho .* sums
int ref_pressure_force(const fvh_mesh_desc *d, int32_t patch, const double *p, double rho, double *F) {
    Mesh m;
    load_mesh(d, m);
    if (patch < 0 || patch >= (int)m.boundaries.size()) return fail(FVH_ERR_ARG, "patch %d", patch);
    Vec ph(p, p + m.n), x, y, z;
    surface_flux(m, patch, ph, x, y, z);
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (size_t i = 0; i < x.size(); ++i) {
        sx += x[i];
        sy += y[i];
        sz += z[i];
    }
    F[0] = rho * sx;
    F[1] = rho * sy;
    F[2] = rho * sz;
    return FVH_OK;
}

// viscous_force (:29-60): snGrad * area * (nu + nut[cID]); zero gradient when
// no velocity BC names the patch (uw == NULL); nut NULL = 0
int ref_viscous_force(const fvh_mesh_desc *d, int32_t patch, const double *ux, const double *uy, const double *uz,
                      const double *uw, double rho, double nu, const double *nut, double *F) {
    Mesh m;
    load_mesh(d, m);
    if (patch < 0 || patch >= (int)m.boundaries.size()) return fail(FVH_ERR_ARG, "patch %d", patch);
    const Boundary &B = m.boundaries[patch];
    Vec a(ux, ux + m.n), b(uy, uy + m.n), c(uz, uz + m.n), x(B.facesID.size(), 0.0), y(x), z(x);
    if (uw) surface_normal_gradient(m, patch, a, b, c, uw, x, y, z);
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (size_t i = 0; i < x.size(); ++i) {
        int64_t cID = B.cellsID[i], fID = B.facesID[i];
        double area = m.farea[fID], nue = nu + (nut ? nut[cID] : 0.0);
        sx += x[i] * area * nue;
        sy += y[i] * area * nue;
        sz += z[i] * area * nue;
    }
    F[0] = rho * sx;
    F[1] = rho * sy;
    F[2] = rho * sz;
    return FVH_OK;
}

// wall_shear_stress (:62-96): tauw = snGrad * (nu + nut[cID]), with face centres
int ref_wall_shear_stress(const fvh_mesh_desc *d, int32_t patch, const double *ux, const double *uy,
                          const double *uz, const double *uw, double nu, const double *nut, double *tx, double *ty,
                          double *tz, double *pos) {
    Mesh m;
    load_mesh(d, m);
    if (patch < 0 || patch >= (int)m.boundaries.size()) return fail(FVH_ERR_ARG, "patch %d", patch);
    const Boundary &B = m.boundaries[patch];
    Vec a(ux, ux + m.n), b(uy, uy + m.n), c(uz, uz + m.n), x, y, z;
    surface_normal_gradient(m, patch, a, b, c, uw, x, y, z);
    for (size_t i = 0; i < x.size(); ++i) {
        int64_t cID = B.cellsID[i], fID = B.facesID[i];
        double nueff = nu + (nut ? nut[cID] : 0.0);
        tx[i] = x[i] * nueff;
        ty[i] = y[i] * nueff;
        tz[i] = z[i] * nueff;
        if (pos) {
            pos[3 * i] = m.fcentre[fID].x;
            pos[3 * i + 1] = m.fcentre[fID].y;
            pos[3 * i + 2] = m.fcentre[fID].z;
        }
    }
    return FVH_OK;
}

int ref_div(const fvh_mesh_desc *d, const double *fx, const double *fy, const double *fz, double *out) {
    Mesh m;
    load_mesh(d, m);
    Vec a(fx, fx + m.nf), b(fy, fy + m.nf), c(fz, fz + m.nf), o(m.n);
    div(m, a, b, c, o);
    std::copy(o.begin(), o.end(), out);
    return FVH_OK;
}

int ref_flux(const fvh_mesh_desc *d, const double *fx, const double *fy, const double *fz, double *mdot) {
    Mesh m;
    load_mesh(d, m);
    Vec a(fx, fx + m.nf), b(fy, fy + m.nf), c(fz, fz + m.nf), o(m.nf);
    flux(m, a, b, c, o);
    std::copy(o.begin(), o.end(), mdot);
    return FVH_OK;
}

// One equation assembled as a loop body assembles it: discretise! (terms in
// model order, sources, Time{Euler} from prev with dt), then
// apply_boundary_conditions! (BC-tuple order, component comp), then, when
// alpha > 0, implicit_relaxation! against relax_field, and, when has_ref,
// setReference! at refcell.  Term fluxes and phi are host arrays (faces for
// Laplacian / Divergence, cells for Si; NULL = the constant 1).  Outputs A's
// values in CSR order and b.
int ref_assemble(const fvh_mesh_desc *d, const fvh_term *terms, int32_t nterms, const fvh_source *srcs, int32_t nsrc,
                 const double *prev, double dt, const fvh_bc *bcs, int32_t nbc, int32_t comp,
                 const double *relax_field, double alpha, int32_t has_ref, double pref, int64_t refcell,
                 double *csr_out, double *b_out) {
    Mesh m;
    load_mesh(d, m);
    std::vector<Vec> store;
    store.reserve(2 * (size_t)nterms + (size_t)nsrc + 1);
    Model md;
    for (int k = 0; k < nterms; ++k) {
        const fvh_term &t = terms[k];
        const int64_t len = t.op == FVH_OP_SI ? m.n : m.nf;
        const Vec *fl = nullptr, *ph = nullptr;
        if (t.flux) {
            store.emplace_back(t.flux, t.flux + len);
            fl = &store.back();
        }
        if (t.phi) {
            store.emplace_back(t.phi, t.phi + m.n);
            ph = &store.back();
        }
        md.terms.push_back(Term{t.op, t.scheme, t.sign, fl, ph});
    }
    for (int k = 0; k < nsrc; ++k) {
        store.emplace_back(srcs[k].field, srcs[k].field + m.n);
        md.sources.push_back(Source{srcs[k].sign, &store.back()});
    }
    Equation eq;
    eq.A = make_pattern(m);
    eq.b.assign(m.n, 0.0);
    eq.R.assign(m.n, 0.0);
    eq.Fx.assign(m.n, 0.0);
    Vec pv;
    if (prev) pv.assign(prev, prev + m.n);
    discretise(m, md, eq, prev ? &pv : nullptr, dt);
    std::vector<fvh_bc> b(bcs, bcs + nbc);
    for (auto &bc : b) bc.patch -= (int32_t)d->index_base;
    int rc = apply_bcs(m, md, eq, b, comp);
    if (rc) return rc;
    if (alpha > 0.0 && relax_field) implicit_relaxation(eq, Vec(relax_field, relax_field + m.n), alpha);
    set_reference(eq, has_ref != 0, pref, refcell);
    Vec vals;
    csr_values(eq.A, vals);
    std::copy(vals.begin(), vals.end(), csr_out);
    std::copy(eq.b.begin(), eq.b.end(), b_out);
    return FVH_OK;
}

// Whole run: simple!/piso!(model, config; pref)
int ref_simple(const fvh_mesh_desc *d, const fvh_case *c, double *Ux, double *Uy, double *Uz, double *p, double *k,
               double *omega, double *nut, double *Rux, double *Ruy, double *Rp, int32_t *done) {
    ref_sim *h = nullptr;
    int rc = ref_sim_create(d, c, Ux, Uy, Uz, p, k, omega, nut, &h);
    if (rc) return rc;
    int32_t dn = 0;
    ref_sim_iterate(h, c->iterations, &dn);
    Sim &S = h->S;
    const int64_t n = S.m.n;
    std::copy(S.Ux.begin(), S.Ux.end(), Ux);
    std::copy(S.Uy.begin(), S.Uy.end(), Uy);
    if (Uz) std::copy(S.Uz.begin(), S.Uz.end(), Uz);
    std::copy(S.p.begin(), S.p.end(), p);
    if (k) std::copy(S.k.begin(), S.k.end(), k);
    if (omega) std::copy(S.omega.begin(), S.omega.end(), omega);
    if (nut) std::copy(S.nut.begin(), S.nut.end(), nut);
    (void)n;
    ref_sim_residuals(h, Rux, Ruy, Rp);
    if (done) *done = dn;
    ref_sim_destroy(h);
    return FVH_OK;
}

}  // extern "C"
