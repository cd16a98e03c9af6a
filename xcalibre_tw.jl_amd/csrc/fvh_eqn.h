// The equation handle: A values on the mesh CSR pattern, b, R, Fx, the
// preconditioner attached to it and the Krylov workspace.
#pragma once

#include "k_calc.h"

namespace fvh {

constexpr int kMaxTerms = 4;
constexpr int kMaxSrc = 2;

struct KTerm {
    int op, scheme;
    double sign;
    const double *flux;  // face array (Laplacian, Divergence) or cell array (Si); null = 1
    const double *phi;   // cell values of the unknown
};

enum { ASM_FULL = 0, ASM_BC_ONLY = 1, ASM_RELAX_ONLY = 2 };

struct AsmArgs {
    int mode;
    int nterms;
    KTerm t[kMaxTerms];
    int nft;              // face terms (Laplacian, Divergence), set by launch_assemble
    KTerm ft[kMaxTerms];  // ... in model order
    int nsrc;
    double ssign[kMaxSrc];
    const double *sfield[kMaxSrc];
    const double *prev;  // Time{Euler} previous values (null = 0)
    double rdt;
    const int *bptr, *bface, *bbc;  // fused boundary conditions (null = none)
    BcVals bc;
    int relax;
    double alpha;
    const double *rfield;
    int setref;
    double pref;
    int refcell;
    const unsigned char *refflag;  // independent cases: cells that take the reference (null: refcell)
};

// Krylov workspace, allocated on first use and kept across solves.
struct KrylovWork {
    int64_t n = 0;
    int cap_basis = 0;
    DBuf<double> v[10];          // FP64 vectors
    DBuf<double> basis;          // GMRES Krylov basis (cap_basis+1) x n
    DBuf<double> part;           // reduction partials
    DBuf<double> small;          // GMRES R, c, s, z, y
    DBuf<KScal> ks;
    KScal *h_ks = nullptr;       // pinned mirror
    // mixed precision (FP32 inner solves)
    int64_t n32 = 0;
    int cap_basis32 = 0;
    DBuf<float> f[10];           // FP32 vectors
    DBuf<float> basis32;         // GMRES basis of the inner solves
    DBuf<float> valf;            // FP32 copy of A
    // banded grids: A's values by band position ([5][n], 0 where a row has
    // no entry), packed for the duration of a solve (band_live)
    DBuf<double> bval;
    DBuf<float> bval32;
    bool band_live = false;
    std::vector<fvh_stats> case_stats;  // last solve, per independent case
    std::map<std::string, hipGraphExec_t> graphs;  // captured iteration batches (independent cases)
    void drop_graphs() {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
    }
    ~KrylovWork() {
        drop_graphs();
        if (h_ks) (void)hipHostFree(h_ks);
    }
};

}  // namespace fvh

namespace fvh {
struct BandStore {
    DBuf<double> ud, us, uw;
    DBuf<Rec4> f, b;
    DBuf<double> zp;
    DBuf<Rec4f> f32, b32;  // FP32 copies for the mixed-precision inner solves
    DBuf<float> zp32;
    BandRecs r{};
};
// The pipelined band sweeps' buffers of one precision, in the mesh's skewed
// slot order: kf / kb the forward / backward quad records (coefficients of
// four diagonals of a row together), bq / yq / zp quads of the right-hand
// side, the forward and the backward results (see kSkewQ0), hand the global
// hand-over arrays between blocks ([block][forward | backward][HS]).
template <typename T>
struct SkewSweep {
    DBuf<T> kf, kb, bq, yq, zp, hand;
};
// The diagonal recurrence's records F (0, Dinv_r, A_rs, A_rw), Bk (0,
// Dinv_r, A_re, A_rn), U (A_rr, A_rs*A_sr, A_rw*A_wr, 0) and its global
// hand-over arrays uhand ([block][HS]); the sweeps' FP64 and FP32 buffers;
// *h_fault (pinned host memory) is set by a sweep whose hand-over wait ran
// out.
struct SkewStore {
    DBuf<Rec4> F, Bk, U;
    DBuf<double> uhand;
    SkewSweep<double> s64;
    SkewSweep<float> s32;
    int *h_fault = nullptr;
    SkewStore() = default;
    SkewStore(const SkewStore &) = delete;
    SkewStore &operator=(const SkewStore &) = delete;
    ~SkewStore() {
        if (h_fault) (void)hipHostFree(h_fault);
    }
};
// Per-row records of the level-scheduled DILU (update, forward, backward;
// FP32 copies of the sweep records), stored as opaque blobs of the record
// size: at most kLevDeps neighbours per row.
constexpr int kLevDeps = 4;
constexpr int kLevRecBytes64 = 96, kLevRecBytes32 = 64;
struct alignas(16) LevBlob64 {
    unsigned char b[kLevRecBytes64];
};
struct alignas(16) LevBlob32 {
    unsigned char b[kLevRecBytes32];
};
// Records of the level walks and their vectors in walk order (a: input,
// b: results).
struct LevelStore {
    DBuf<LevBlob64> upd, fwd, bwd;
    DBuf<LevBlob32> fwd32, bwd32;
    DBuf<double> a, b, c;
    DBuf<float> a32, b32, c32;
    int walk = 0;            // the walk the records are for: 0 level walk, 1 sync-free, 2 tiles
    int *h_fault = nullptr;  // pinned: a wait of the sync-free walk ran out
    LevelStore() = default;
    LevelStore(const LevelStore &) = delete;
    LevelStore &operator=(const LevelStore &) = delete;
    ~LevelStore() {
        if (h_fault) (void)hipHostFree(h_fault);
    }
};
}  // namespace fvh

namespace fvh {
// LDL preconditioner: the factor of the set-up matrix in its ordering (L by
// rows for the forward solve, by columns for the backward one), its pivots,
// the solve's work vectors and a pinned fault word.
struct LdlStore {
    bool ready = false;
    int nnzL = 0;
    DBuf<int> perm, fptr, fidx, bptr, bidx;
    DBuf<int> ranges;  // bounds of the nested-dissection leaves and separators (at most kLdlRangeMax rows each)
    int nranges = 0;
    DBuf<double> fval, bval, D, rhs, y, z;
    int *h_fault = nullptr;
    LdlStore() = default;
    LdlStore(const LdlStore &) = delete;
    LdlStore &operator=(const LdlStore &) = delete;
    ~LdlStore() {
        if (h_fault) (void)hipHostFree(h_fault);
    }
};
}  // namespace fvh

struct fvh_eqn {
    fvh_mesh *m = nullptr;
    fvh::DBuf<double> nz, b, R, Fx;
    // preconditioner
    int pkind = FVH_PRECOND_NONE;
    fvh_eqn *pmat = nullptr;    // equation whose A the preconditioner reads
    fvh_eqn *pshare = nullptr;  // use that equation's preconditioner object instead
    fvh::DBuf<double> D, Dinv, lu, pstore;
    fvh::DBuf<float> p32a, p32b;  // FP32 copies: Dinv / Jacobi scale, and A or LU values
    fvh::BandStore bd;  // DILU on banded grids (packed order)
    fvh::SkewStore sk;  // DILU on banded grids (pipelined sweeps)
    fvh::LevelStore lv; // DILU on unstructured meshes (level-scheduled records)
    fvh::LdlStore ldl;  // LDL: factor of the set-up matrix
    fvh::KrylovWork kw;
    fvh::DBuf<double> red;  // residual reductions
};

namespace fvh {

int launch_assemble(fvh_eqn *e, const AsmArgs &a);
int make_terms(const fvh_term *terms, int nterms, AsmArgs &a);
int check_bc_terms(const AsmArgs &a, const fvh_bc *bcs, int nbc);

// solve-side entry points used by the drivers
int precond_update(fvh_eqn *e);
int precond_apply(fvh_eqn *e, double *y, const double *x, const KScal *ks, double sigma);
int launch_spmv(fvh_eqn *e, const double *x, double *y, const KScal *ks);
int solve(fvh_eqn *e, const fvh_solver_cfg &cfg, double *x, fvh_stats *st);
int residual(fvh_eqn *e, const double *phi, double *res_host);
int residual_async(fvh_eqn *e, const double *phi, double *res_dev);
int precond_dinv_natural(fvh_eqn *e, const double **out);
// LDL preconditioner (k_ldl.hip)
int ldl_setup(fvh_eqn *e, const double *val_dev);
template <typename T>
int ldl_apply(fvh_eqn *e, const T *x, T *out, const KScal *ks);
int ldl_check(fvh_eqn *e);

// shared by the single-case and the per-case (batched) solvers
int ensure_work(fvh_eqn *e, int basis);
int ensure_work32(fvh_eqn *e, int basis);
int prepare32(fvh_eqn *e);
int precond_apply32(fvh_eqn *eq, float *y, const float *x, const KScal *ks);
int launch_spmv32(fvh_eqn *e, const float *x, float *y, float sigma, const KScal *ks, const float *w, double *part);
int cg_sigma(fvh_eqn *e, double *sigma);
// independent cases: every case solved with its own Krylov state; st gets
// the largest iteration count and the worst status, e->kw.case_stats each case
int solve_cases(fvh_eqn *e, const fvh_solver_cfg &cfg, double *x, fvh_stats *st);
int residual_cases_async(fvh_eqn *e, const double *phi, double *out, int stride, int comp);

}  // namespace fvh
