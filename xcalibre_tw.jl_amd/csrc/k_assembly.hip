// Discretise: face-loop assembly of the model terms into A (CSR values) and
// b, with boundary conditions, implicit relaxation and the pressure
// reference fused into the cell pass.
//
// Two passes, no atomics.  The face pass (one thread per internal face,
// face arrays read coalesced) evaluates every face term for the owner and
// the neighbour side (Discretise_1_schemes.jl:40-92), writes the two
// off-diagonal slots through the precomputed face->CSR slot map, and leaves
// each side's diagonal terms in a per-slot array.  The cell pass (one
// thread per cell) sums those terms in the reference's order
// (Discretise_2_generated_distretisation.jl:55-73):
//   faces in faces_range order (terms in model order)
//   -> scheme_source! terms -> sources
//   -> boundary conditions (BC tuple order, then patch face order, then terms)
//   -> implicit relaxation -> setReference!.
#include "fvh_eqn.h"

namespace fvh {

__device__ inline void apply_bc_entries(const DevMesh &m, const AsmArgs &a, int c, double &diag, double &b) {
    for (int e = a.bptr[c]; e < a.bptr[c + 1]; ++e) {
        const int f = a.bface[e];
        const int k = a.bbc[e];
        const int kind = a.bc.kind[k];
        const double value = a.bc.value[k];
#pragma unroll
        for (int ti = 0; ti < kMaxTerms; ++ti) {
            if (ti >= a.nterms) break;
            const KTerm &t = a.t[ti];
            if (t.op == FVH_OP_LAPLACIAN) {
                if (kind == FVH_BC_DIRICHLET) {
                    double J = t.flux[f];
                    double flux = J * m.farea[f] / m.fdelta[f];
                    double ap = t.sign * (-flux);
                    diag += ap;
                    b += ap * value;
                }
            } else if (t.op == FVH_OP_DIVERGENCE) {
                if (t.scheme == FVH_LINEAR) {
                    if (kind == FVH_BC_DIRICHLET) b += t.sign * (-t.flux[f] * value);
                    else diag += t.sign * t.flux[f];
                } else {
                    double ap = t.sign * t.flux[f];
                    if (kind == FVH_BC_DIRICHLET) b -= ap * value;
                    else diag += jmax0(ap);
                }
            } else if (t.op == FVH_OP_SI) {
                if (kind == FVH_BC_OMEGAWALL) {
                    double flux = t.sign * t.flux[c];
                    b += flux * t.phi[c] * m.vol[c];
                }
            }
        }
    }
}

// Face pass: one thread per internal face, reading the face arrays
// coalesced.  It evaluates every face term for both sides of the face
// (owner o with its nsign, neighbour n with its own), writes the two
// off-diagonal slots A[o, n] and A[n, o] (each the sum of its terms from 0
// in model order, as the reference accumulates them into the zeroed
// nzval), and hands the per-side diagonal terms to the cell pass through
// td[slot][term], indexed by the off-diagonal slot of that side.
template <int NFT>
__global__ void __launch_bounds__(256) k_asm_faces(DevMesh m, AsmArgs a, double *__restrict__ nz,
                                                   double *__restrict__ td) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.nif) return;
    const int f = m.nbf + k;
    const int2 sv = m.fi_slot[k];
    const int so = sv.x >= 0 ? sv.x : ~sv.x, sn = sv.y >= 0 ? sv.y : ~sv.y;
    const double nso = sv.x >= 0 ? 1.0 : -1.0, nsn = sv.y >= 0 ? 1.0 : -1.0;
    double offo = 0.0, offn = 0.0;
    double dto[NFT > 0 ? NFT : 1], dtn[NFT > 0 ? NFT : 1];
#pragma unroll
    for (int q = 0; q < NFT; ++q) {
        const KTerm &t = a.ft[q];
        const double flux = t.flux[f];
        if (t.op == FVH_OP_LAPLACIAN) {
            // the same coefficient on both sides (no nsign)
            const double ap = (t.sign * (-flux * m.farea[f])) / m.fdelta[f];
            dto[q] = ap;
            dtn[q] = ap;
            offo += -ap;
            offn += -ap;
        } else {
            const double apo = t.sign * (flux * nso), apn = t.sign * (flux * nsn);
            if (t.scheme == FVH_LINEAR) {
                const double2 w = m.fi_w[k];
                dto[q] = apo * (1.0 - w.x);
                dtn[q] = apn * (1.0 - w.y);
                offo += apo * w.x;
                offn += apn * w.y;
            } else {
                dto[q] = jmax0(apo);
                dtn[q] = jmax0(apn);
                offo += -jmax0(-apo);
                offn += -jmax0(-apn);
            }
        }
    }
    nz[so] = offo;
    nz[sn] = offn;
#pragma unroll
    for (int q = 0; q < NFT; ++q) {
        td[(size_t)so * NFT + q] = dto[q];
        td[(size_t)sn * NFT + q] = dtn[q];
    }
}

// Cell pass: one thread per cell.  NFT >= 0: full assembly, the diagonal
// summed from 0 over the cell's internal faces in faces_range order, terms
// in model order within a face (the face pass's per-side terms), then the
// cell terms and sources; NFT < 0: the diagonal and b as they stand.  Then
// boundary conditions, implicit relaxation and the reference cell.
template <int NFT>
__global__ void __launch_bounds__(256) k_asm_cells(DevMesh m, AsmArgs a, const double *__restrict__ td,
                                                   double *__restrict__ nz, double *__restrict__ bv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int dslot = m.diag[c];
    double diag, b;
    if (NFT >= 0) {
        diag = 0.0;
        if (NFT > 0) {
            // four sides at a time: their slot loads, then their terms'
            // loads, issued together; summed in faces_range order
            constexpr int G = 4, NT = NFT > 0 ? NFT : 1;
            const int i1 = m.cf_ptr[c + 1];
            for (int i = m.cf_ptr[c]; i < i1; i += G) {
                int sl[G];
                double v[G][NT];
#pragma unroll
                for (int k = 0; k < G; ++k)
                    if (i + k < i1) sl[k] = m.cf_slot[i + k];
#pragma unroll
                for (int k = 0; k < G; ++k)
                    if (i + k < i1) {
#pragma unroll
                        for (int q = 0; q < NT; ++q) v[k][q] = td[(size_t)sl[k] * NT + q];
                    }
#pragma unroll
                for (int k = 0; k < G; ++k)
                    if (i + k < i1) {
#pragma unroll
                        for (int q = 0; q < NT; ++q) diag += v[k][q];
                    }
            }
        }
        b = 0.0;
        const double volume = m.vol[c];
#pragma unroll
        for (int ti = 0; ti < kMaxTerms; ++ti) {
            if (ti >= a.nterms) break;
            const KTerm &t = a.t[ti];
            if (t.op == FVH_OP_TIME && t.scheme == FVH_EULER) {
                double pv = a.prev ? a.prev[c] : 0.0;
                diag += volume * a.rdt;
                b += pv * volume * a.rdt;
            } else if (t.op == FVH_OP_SI) {
                double flux = t.sign * t.flux[c] * volume;
                diag += flux;
            }
        }
#pragma unroll
        for (int s = 0; s < kMaxSrc; ++s) {
            if (s >= a.nsrc) break;
            b += a.ssign[s] * a.sfield[s][c] * volume;
        }
    } else {
        diag = nz[dslot];
        b = bv[c];
    }
    if (a.bptr) apply_bc_entries(m, a, c, diag, b);
    if (a.relax) {
        diag /= a.alpha;
        b += (1.0 - a.alpha) * diag * a.rfield[c];
    }
    if (a.setref && (a.refflag ? a.refflag[c] != 0 : c == a.refcell)) {
        b += diag * a.pref;
        diag += diag;
    }
    nz[dslot] = diag;
    bv[c] = b;
}

int make_terms(const fvh_term *terms, int nterms, AsmArgs &a) {
    if (nterms < 1 || nterms > kMaxTerms) return set_error(FVH_ERR_SIZE, "1..%d terms per equation (got %d)", kMaxTerms, nterms);
    a.nterms = nterms;
    for (int i = 0; i < nterms; ++i) {
        const fvh_term &t = terms[i];
        KTerm &k = a.t[i];
        k.op = t.op;
        k.scheme = t.scheme;
        k.sign = (double)t.sign;
        k.flux = t.flux;
        k.phi = t.phi;
        if (t.sign != 1 && t.sign != -1) return set_error(FVH_ERR_ARG, "term %d: sign must be +1 or -1", i);
        switch (t.op) {
            case FVH_OP_TIME:
                if (t.scheme != FVH_STEADY && t.scheme != FVH_EULER) return set_error(FVH_ERR_UNSUPPORTED, "Time scheme %d", t.scheme);
                break;
            case FVH_OP_LAPLACIAN:
                if (t.scheme != FVH_LINEAR) return set_error(FVH_ERR_UNSUPPORTED, "Laplacian scheme %d", t.scheme);
                if (!t.flux) return set_error(FVH_ERR_ARG, "Laplacian needs a face flux");
                break;
            case FVH_OP_DIVERGENCE:
                if (t.scheme != FVH_LINEAR && t.scheme != FVH_UPWIND) return set_error(FVH_ERR_UNSUPPORTED, "Divergence scheme %d", t.scheme);
                if (!t.flux) return set_error(FVH_ERR_ARG, "Divergence needs a face flux");
                break;
            case FVH_OP_SI:
                if (!t.flux || !t.phi) return set_error(FVH_ERR_ARG, "Si needs a cell flux and phi");
                break;
            default: return set_error(FVH_ERR_UNSUPPORTED, "operator %d", t.op);
        }
    }
    return FVH_OK;
}

// The reference defines boundary methods for every (BC, term) pair it uses;
// a NutWallFunction on an equation term has none (a MethodError there).
int check_bc_terms(const AsmArgs &a, const fvh_bc *bcs, int nbc) {
    for (int k = 0; k < nbc; ++k) {
        int kind = bcs[k].kind;
        if (kind < FVH_BC_DIRICHLET || kind > FVH_BC_NUTWALL) return set_error(FVH_ERR_ARG, "BC kind %d", kind);
        for (int i = 0; i < a.nterms; ++i)
            if (kind == FVH_BC_NUTWALL && a.t[i].op != FVH_OP_TIME)
                return set_error(FVH_ERR_UNSUPPORTED, "NutWallFunction has no boundary method for operator %d", a.t[i].op);
    }
    return FVH_OK;
}

template <int NFT>
static int launch_full(fvh_eqn *e, const AsmArgs &a) {
    fvh_mesh *m = e->m;
    double *td = nullptr;
    if (NFT > 0) {
        const size_t need = (size_t)m->nnz * NFT;
        if (m->asm_terms.n < need) FVH_TRY(m->asm_terms.alloc(need));
        td = m->asm_terms.p;
    }
    if (m->d.nif > 0) {
        hipLaunchKernelGGL((k_asm_faces<NFT>), grid1d(m->d.nif), dim3(256), 0, stream(), m->d, a, e->nz.p, td);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL((k_asm_cells<NFT>), grid1d(m->n), dim3(256), 0, stream(), m->d, a, td, e->nz.p, e->b.p);
    HIP_TRY(hipGetLastError());
    return FVH_OK;
}

// Full assembly: the face pass then the cell pass (the face terms compacted
// into a.ft in model order); the partial modes run the cell pass alone.
int launch_assemble(fvh_eqn *e, const AsmArgs &a0) {
    if (a0.mode != ASM_FULL) {
        hipLaunchKernelGGL((k_asm_cells<-1>), grid1d(e->m->n), dim3(256), 0, stream(), e->m->d, a0, nullptr, e->nz.p,
                           e->b.p);
        HIP_TRY(hipGetLastError());
        return FVH_OK;
    }
    AsmArgs a = a0;
    a.nft = 0;
    for (int i = 0; i < a.nterms; ++i)
        if (a.t[i].op == FVH_OP_LAPLACIAN || a.t[i].op == FVH_OP_DIVERGENCE) a.ft[a.nft++] = a.t[i];
    switch (a.nft) {
        case 0: return launch_full<0>(e, a);
        case 1: return launch_full<1>(e, a);
        case 2: return launch_full<2>(e, a);
        case 3: return launch_full<3>(e, a);
        case 4: return launch_full<4>(e, a);
        default: return set_error(FVH_ERR_SIZE, "at most %d face terms", kMaxTerms);
    }
}

}  // namespace fvh

using namespace fvh;

extern "C" int fvh_eqn_create(fvh_mesh *m, fvh_eqn **out) {
    FVH_GUARD_BEGIN
    if (!m || !out) return set_error(FVH_ERR_ARG, "null argument");
    auto e = std::make_unique<fvh_eqn>();
    e->m = m;
    FVH_TRY(e->nz.alloc(m->nnz));
    FVH_TRY(e->b.alloc(m->n));
    FVH_TRY(e->R.alloc(m->n));
    FVH_TRY(e->Fx.alloc(m->n));
    FVH_TRY(e->nz.zero());
    FVH_TRY(e->b.zero());
    HIP_TRY(hipStreamSynchronize(stream()));
    *out = e.release();
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" void fvh_eqn_destroy(fvh_eqn *e) { delete e; }

extern "C" int fvh_eqn_arrays(fvh_eqn *e, double **nzval, double **b) {
    if (!e) return set_error(FVH_ERR_ARG, "null equation");
    if (nzval) *nzval = e->nz.p;
    if (b) *b = e->b.p;
    return FVH_OK;
}

extern "C" int fvh_discretise(fvh_eqn *e, const fvh_term *terms, int32_t nterms, const fvh_source *srcs, int32_t nsrc,
                              const double *prev, double dt) {
    FVH_GUARD_BEGIN
    if (!e || !terms || (nsrc && !srcs)) return set_error(FVH_ERR_ARG, "null argument");
    if (nsrc < 0 || nsrc > kMaxSrc) return set_error(FVH_ERR_SIZE, "0..%d sources per equation", kMaxSrc);
    AsmArgs a;
    std::memset(&a, 0, sizeof a);
    a.mode = ASM_FULL;
    FVH_TRY(make_terms(terms, nterms, a));
    a.nsrc = nsrc;
    for (int s = 0; s < nsrc; ++s) {
        if (!srcs[s].field) return set_error(FVH_ERR_ARG, "source %d has no field", s);
        a.ssign[s] = (double)srcs[s].sign;
        a.sfield[s] = srcs[s].field;
    }
    a.prev = prev;
    a.rdt = 1.0 / dt;
    FVH_TRY(launch_assemble(e, a));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_apply_bcs(fvh_eqn *e, const fvh_term *terms, int32_t nterms, const fvh_bc *bcs, int32_t nbc,
                             int32_t comp) {
    FVH_GUARD_BEGIN
    if (!e || !terms || (nbc && !bcs)) return set_error(FVH_ERR_ARG, "null argument");
    AsmArgs a;
    std::memset(&a, 0, sizeof a);
    a.mode = ASM_BC_ONLY;
    FVH_TRY(make_terms(terms, nterms, a));
    FVH_TRY(check_bc_terms(a, bcs, nbc));
    FVH_TRY(fill_bcvals(bcs, nbc, comp, a.bc));
    BcPlan *plan = nullptr;
    FVH_TRY(bc_plan(e->m, bcs, nbc, &plan));
    a.bptr = plan->ptr.p;
    a.bface = plan->face.p;
    a.bbc = plan->bc.p;
    FVH_TRY(launch_assemble(e, a));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_implicit_relax(fvh_eqn *e, const double *phi, double alpha) {
    if (!e || !phi) return set_error(FVH_ERR_ARG, "null argument");
    AsmArgs a;
    std::memset(&a, 0, sizeof a);
    a.mode = ASM_RELAX_ONLY;
    a.relax = 1;
    a.alpha = alpha;
    a.rfield = phi;
    FVH_TRY(launch_assemble(e, a));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

extern "C" int fvh_set_reference(fvh_eqn *e, double pref, int64_t cell) {
    if (!e || cell < 0 || cell >= e->m->n) return set_error(FVH_ERR_ARG, "bad argument");
    AsmArgs a;
    std::memset(&a, 0, sizeof a);
    a.mode = ASM_RELAX_ONLY;
    a.setref = 1;
    a.pref = pref;
    a.refcell = (int)cell;
    FVH_TRY(launch_assemble(e, a));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}

// correct!(eqn, term, corr_flux) [Calculate_3_orthogonality_correction.jl:6-24]
extern "C" int fvh_nonorthogonal_correct(fvh_eqn *e, int32_t sign, const double *J, const double *corr) {
    if (!e || !J || !corr) return set_error(FVH_ERR_ARG, "null argument");
    if (sign != 1 && sign != -1) return set_error(FVH_ERR_ARG, "sign must be +1 or -1");
    FVH_TRY(launch_nonorthogonal_correct(e->m, sign, J, corr, e->b.p));
    HIP_TRY(hipStreamSynchronize(stream()));
    return FVH_OK;
}
