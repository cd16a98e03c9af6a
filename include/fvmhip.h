/*
 * fvmhip.h — C ABI of libfvmhip.so, the MI355X (gfx950) implementation of the
 * per-iteration hot path of the finite-volume solver XCALibre_TW.jl (package
 * FVM_1D v0.2.0): face-loop assembly, Gauss gradients and face interpolation,
 * Krylov solves with DILU/ILU(0)/Jacobi preconditioners, inside the SIMPLE/PISO
 * loops and the k-omega transport equations.
 *
 * Every entry point returns an int status (FVH_OK = 0).  No C++ exception
 * crosses this boundary.  Pointers documented as "device" are HIP device
 * addresses; "host" pointers are ordinary memory.  All calls are ordered on
 * the library's stream and return after completion unless stated otherwise.
 * Handles are not re-entrant: one host thread per handle.
 *
 * Citations in brackets name the reference function each entry replaces
 * (paths relative to the XCALibre_TW.jl repository root).
 */
#ifndef FVMHIP_H
#define FVMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes                                                        */
/* ------------------------------------------------------------------ */
enum {
    FVH_OK = 0,
    FVH_ERR_ARG = 1,          /* bad argument / inconsistent sizes           */
    FVH_ERR_HIP = 2,          /* HIP runtime error                           */
    FVH_ERR_NODEV = 3,        /* no usable gfx950 device                     */
    FVH_ERR_MESH = 4,         /* mesh inconsistent (reference throws)        */
    FVH_ERR_UNSUPPORTED = 5,  /* term/BC combination has no reference method */
    FVH_ERR_IO = 6,           /* file could not be read                      */
    FVH_ERR_SIZE = 7          /* size beyond a documented maximum            */
};

/* Krylov termination (the reference ignores it: src/Solve/Solve_1_api.jl:43) */
enum {
    FVH_KRYLOV_SOLVED = 0,     /* ||r|| <= atol + rtol*||r0||                 */
    FVH_KRYLOV_ITMAX = 1,      /* maximum number of iterations exceeded       */
    FVH_KRYLOV_BREAKDOWN = 2,  /* alpha == 0 / NaN, or b'c == 0 (BiCGSTAB),
                                  H(k+1,k) <= eps^(3/4) (GMRES)               */
    FVH_KRYLOV_ZERO_RES = 3,   /* initial residual is exactly zero            */
    FVH_KRYLOV_ZERO_CURV = 4   /* p'Ap ~ 0 (CG)                               */
};

/* Operators and schemes  [src/ModelFramework/ModelFramework_0_types.jl:24-49,
 *                         src/Discretise/Discretise_0_types.jl:13-21]       */
enum { FVH_OP_TIME = 0, FVH_OP_LAPLACIAN = 1, FVH_OP_DIVERGENCE = 2, FVH_OP_SI = 3 };
enum { FVH_STEADY = 0, FVH_EULER = 1, FVH_LINEAR = 2, FVH_UPWIND = 3,
       FVH_ORTHOGONAL = 4, FVH_MIDPOINT = 5 };

/* Boundary conditions  [src/Discretise/Discretise_0_types.jl:31-78] */
enum { FVH_BC_DIRICHLET = 0, FVH_BC_NEUMANN = 1, FVH_BC_KWALL = 2,
       FVH_BC_OMEGAWALL = 3, FVH_BC_NUTWALL = 4 };

/* Krylov solvers and preconditioners [src/Solve/Solve_1_api.jl:5-26,
 *                          src/Solve/Preconditioners/preconditioners_0_types.jl:6-12] */
enum { FVH_BICGSTAB = 0, FVH_CG = 1, FVH_GMRES = 2 };
enum { FVH_PRECOND_NONE = 0, FVH_PRECOND_DILU = 1, FVH_PRECOND_ILU0 = 2,
       FVH_PRECOND_JACOBI = 3, FVH_PRECOND_NORMDIAG = 4,
       FVH_PRECOND_LDL = 5 };  /* factor of the matrix at the first update, kept
                                  (opLDL, preconditioners_0_types.jl:36-42) */

/* Arithmetic of the Krylov inner solve */
enum { FVH_FP64 = 0,          /* plain FP64, bit-identical to the oracle             */
       FVH_MIXED = 1,         /* FP32 SpMV + preconditioner inside FP64 refinement   */
       FVH_MIXED_FP16 = 2 };  /* reserved (FP16 off-diagonals): FVH_ERR_UNSUPPORTED  */

/* Turbulence models and pressure-velocity algorithms */
enum { FVH_LAMINAR = 0, FVH_KOMEGA = 1 };
enum { FVH_SIMPLE = 0, FVH_PISO = 1 };

/* ------------------------------------------------------------------ */
/* Mesh2 flat layout  [src/Mesh/Mesh_0_types.jl:21-94,                 */
/*                     src/Mesh/Mesh2D/Mesh2D_0_update_mesh.jl:5-142]  */
/* ------------------------------------------------------------------ */
/* Vectors are xyz-interleaved (SVector{3} layout).  `faces_range` and
 * `boundary_range` hold (first, last) pairs, inclusive, in index_base
 * (1 = Julia arrays as they are, 0 = zero based).  All integer arrays
 * use the same base. */
typedef struct fvh_mesh_desc {
    int64_t index_base;
    int64_t ncells, nfaces, nboundaries, ncellfaces, nboundaryfaces;
    const double  *cell_centre;      /* [3*ncells]                     */
    const double  *cell_volume;      /* [ncells]                       */
    const int64_t *cell_faces_range; /* [2*ncells]                     */
    const int64_t *cell_faces;       /* [ncellfaces] internal faces    */
    const int64_t *cell_neighbours;  /* [ncellfaces]                   */
    const int64_t *cell_nsign;       /* [ncellfaces] +1 / -1           */
    const int64_t *face_owners;      /* [2*nfaces] (owner, neighbour); boundary: (c, c) */
    const double  *face_centre;      /* [3*nfaces]                     */
    const double  *face_normal;      /* [3*nfaces]                     */
    const double  *face_e;           /* [3*nfaces]                     */
    const double  *face_area;        /* [nfaces]                       */
    const double  *face_delta;       /* [nfaces]                       */
    const double  *face_weight;      /* [nfaces]                       */
    const int64_t *boundary_range;   /* [2*nboundaries] into boundary_faces/cells */
    const int64_t *boundary_faces;   /* [nboundaryfaces] facesID       */
    const int64_t *boundary_cells;   /* [nboundaryfaces] cellsID       */
} fvh_mesh_desc;

/* One boundary condition of a field.  `patch` is the boundary index
 * (index_base of the mesh it is used with).  Scalar fields use value[0];
 * vector fields use value[component].  Wall functions read wf =
 * (kappa, beta1, cmu, B, E)  [src/Discretise/Discretise_0_types.jl:56-78]. */
typedef struct fvh_bc {
    int32_t kind;
    int32_t patch;
    double  value[3];
    double  wf[5];
} fvh_bc;

/* One term of a model equation  [ModelFramework_0_types.jl:15-49].
 * flux: device face array (Laplacian, Divergence) or cell array (Si);
 * NULL means the constant 1 (Time).  phi: device cell array of the
 * unknown (read by the OmegaWallFunction Si boundary term). */
typedef struct fvh_term {
    int32_t op;
    int32_t scheme;
    int32_t sign;
    int32_t pad_;
    const double *flux;
    const double *phi;
} fvh_term;

/* One explicit source  [ModelFramework_0_types.jl:53-66]: b += sign*field*V */
typedef struct fvh_source {
    int32_t sign;
    int32_t pad_;
    const double *field;   /* device cell array */
} fvh_source;

/* set_schemes(...)  [src/Discretise/Discretise_0_types.jl:171-181] */
typedef struct fvh_schemes {
    int32_t time, divergence, laplacian, gradient;
} fvh_schemes;

/* set_solver(...)  [src/Solve/Solve_1_api.jl:5-26] */
typedef struct fvh_solver_cfg {
    int32_t solver;
    int32_t precond;
    int32_t itmax;        /* reference default 100                 */
    int32_t precision;    /* FVH_FP64 / FVH_MIXED / FVH_MIXED_FP16  */
    double  convergence;  /* stored, never read by the reference   */
    double  relax;
    double  atol;         /* reference default sqrt(eps)           */
    double  rtol;         /* reference default 1e-3                */
} fvh_solver_cfg;

typedef struct fvh_bcset {
    int32_t nbc;
    int32_t pad_;
    const fvh_bc *bcs;    /* host array, BC-tuple order */
} fvh_bcset;

/* Everything simple!/piso! read from (model, config; pref)
 * [src/Solvers/Solvers_1_SIMPLE.jl:3-67, src/Solvers/Solvers_2_PISO.jl:3-67] */
typedef struct fvh_case {
    int32_t model;          /* FVH_LAMINAR / FVH_KOMEGA                     */
    int32_t algorithm;      /* FVH_SIMPLE / FVH_PISO                        */
    int32_t iterations;     /* runtime.iterations                           */
    int32_t has_pref;       /* pref !== nothing                             */
    double  pref;
    double  nu;             /* ConstantScalar viscosity                     */
    double  dt;             /* runtime.dt                                   */
    double  conv_break;     /* loop break threshold; reference hard-codes 1e-7 */
    fvh_bcset U, p, k, omega, nut;
    fvh_schemes sU, sp, sk, somega;
    fvh_solver_cfg vU, vp, vk, vomega;
    int32_t nonorth_correctors;  /* the reference's disabled `correct` block
                                    (Solvers_1_SIMPLE.jl:165-180): 0 = off   */
    int32_t pad_;
} fvh_case;

/* Per-solve Krylov statistics */
typedef struct fvh_stats {
    int32_t niter;
    int32_t status;
    int32_t refinements;  /* outer FP64 refinement steps (mixed precision) */
    int32_t pad_;
    double  rnorm;        /* final preconditioned residual norm          */
    double  rnorm0;
} fvh_stats;

/* ------------------------------------------------------------------ */
/* Library                                                             */
/* ------------------------------------------------------------------ */
const char *fvh_version(void);
const char *fvh_status_string(int status);
/* Copies the message of the last failed call of this thread into buf. */
int fvh_last_error(char *buf, int64_t len);
/* Number of HIP devices (0 on a host without a GPU). */
int fvh_device_count(int32_t *n);
/* Device memory for hosts that have no GPU array type of their own (the
 * Julia caller passes its own device arrays instead).  fvh_memcpy copies in
 * any direction (unified addressing) and returns after completion. */
int fvh_malloc(void **ptr, int64_t bytes);
int fvh_free(void *ptr);
int fvh_memcpy(void *dst, const void *src, int64_t bytes);
int fvh_synchronize(void);

/* ------------------------------------------------------------------ */
/* Host mesh building: the UNV2 builder rules (face order, ownership,   */
/* geometry) applied to points / elements / boundary groups.            */
/* [src/UNV2/UNV2_3_builder.jl:3-287, src/UNV2/UNV2_2_geometry.jl:20-165] */
/* ------------------------------------------------------------------ */
typedef struct fvh_hmesh fvh_hmesh;

/* elem_ptr[nelements+1] (0-based offsets into elem_verts); elem_verts are
 * 1-based point ids; element labels are 1-based positions.  Edge elements
 * (2 vertices) must precede the cells (3 or 4 vertices).  group_elems are
 * 1-based element labels; group_names is a NUL-separated list. */
int fvh_hmesh_build(int64_t npoints, const double *points_xyz,
                    int64_t nelements, const int64_t *elem_ptr,
                    const int64_t *elem_verts,
                    int64_t ngroups, const int64_t *group_ptr,
                    const int64_t *group_elems, const char *group_names,
                    double scale, fvh_hmesh **out);
/* build_mesh(file; scale)  [src/UNV2/UNV2_3_builder.jl:3-23, UNV2_1_loader.jl:3-120] */
int fvh_hmesh_read_unv(const char *path, double scale, fvh_hmesh **out);
/* Fills a zero-based description whose arrays live as long as the hmesh. */
int fvh_hmesh_desc(const fvh_hmesh *m, fvh_mesh_desc *out);
int fvh_hmesh_boundary_name(const fvh_hmesh *m, int64_t i, char *buf, int64_t len);
/* The mesh's nodes (scaled, xyz interleaved) and each cell's nodes in
 * element order, zero based (cell c: cell_nodes[cell_nodes_ptr[c] ..
 * cell_nodes_ptr[c+1]-1]); valid while the hmesh lives
 * [mesh.nodes, mesh.cell_nodes: src/Mesh/Mesh_0_types.jl, UNV2_3_builder.jl:165-175] */
int fvh_hmesh_nodes(const fvh_hmesh *m, int64_t *npoints, const double **xyz, const int64_t **cell_nodes_ptr,
                    const int64_t **cell_nodes);
void fvh_hmesh_free(fvh_hmesh *m);

/* ------------------------------------------------------------------ */
/* Device mesh                                                          */
/* ------------------------------------------------------------------ */
typedef struct fvh_mesh fvh_mesh;

/* Uploads the mesh, builds the CSR pattern [Equation(mesh),
 * src/ModelFramework/ModelFramework_0_types.jl:90-124], the face->CSR slot
 * maps, the triangular-sweep schedules and the structured-band detection. */
int fvh_mesh_create(const fvh_mesh_desc *desc, fvh_mesh **out);
void fvh_mesh_destroy(fvh_mesh *m);
/* nnz and, when the pointers are non-NULL, host copies of the CSR pattern
 * (rows sorted by column: the transpose of the reference's CSC). */
int fvh_mesh_pattern(const fvh_mesh *m, int64_t *nnz, int32_t *rowptr, int32_t *colidx);
/* Sweep schedule summary: number of levels and, for banded grids, the band
 * width nx (0 when the general level-scheduled sweep is used). */
int fvh_mesh_sweep_info(const fvh_mesh *m, int64_t *nlevels, int64_t *band_nx);

/* Independent cases (batched mode; no counterpart in the reference, which
 * runs one case per process).  Declares the mesh as the block-diagonal union
 * of ncases cases, case k owning cells [cell_offsets[k], cell_offsets[k+1])
 * (0-based, no face joining two cases).  Every solve on the mesh then keeps
 * one Krylov state per case and reduces per case, so each case computes
 * exactly what it computes alone; the DILU sweeps run one work-group per
 * case when every case is a full i-fastest rectangle. */
int fvh_mesh_set_cases(fvh_mesh *m, int32_t ncases, const int64_t *cell_offsets);
/* Number of cases (0 = not split) and whether their sweeps are banded. */
int fvh_mesh_case_info(const fvh_mesh *m, int32_t *ncases, int32_t *banded);
/* Autotuned launch shape of the banded DILU sweeps: grid rows per lane
 * (1 or 2, timed on the mesh's records at its first DILU update; 0 = the
 * level-scheduled sweeps are used).  No counterpart in the reference. */
int fvh_mesh_tuning(const fvh_mesh *m, int32_t *rows_per_lane);

/* ------------------------------------------------------------------ */
/* Calculate  (device pointers)                                        */
/* ------------------------------------------------------------------ */
/* interpolate!(phif, phi) scalar  [src/Calculate/Calculate_2_interpolation.jl:112-127] */
int fvh_interpolate(const fvh_mesh *m, const double *phi, double *phif);
/* interpolate!(psif, psi) vector: writes x and y only  [Calculate_2_interpolation.jl:131-148] */
int fvh_interpolate_vector(const fvh_mesh *m, const double *psi_x, const double *psi_y,
                           double *psif_x, double *psif_y);
/* correct_boundaries!(phif, phi, BCs), one component (0,1,2) of the BC values
 * [Calculate_2_interpolation.jl:6-108] */
int fvh_correct_boundaries(const fvh_mesh *m, const fvh_bc *bcs, int32_t nbc,
                           int32_t component, const double *phi, double *phif);
/* green_gauss!(dx, dy, dz, phif)  [src/Calculate/Calculate_1_green_gauss.jl:2-39] */
int fvh_green_gauss(const fvh_mesh *m, const double *phif, double *dx, double *dy, double *dz);
/* grad!(grad, phif, phi, BCs) scalar, Orthogonal or Midpoint
 * [src/Calculate/Calculate_0_gradient.jl:96-104,175-183] */
int fvh_grad(const fvh_mesh *m, int32_t scheme, const fvh_bc *bcs, int32_t nbc,
             int32_t component, const double *phi, double *phif,
             double *dx, double *dy, double *dz);
/* div!(phi, psif)  [src/Calculate/Calculate_0_divergence.jl:75-107] */
int fvh_div(const fvh_mesh *m, const double *psif_x, const double *psif_y,
            const double *psif_z, double *out);
/* flux!(phif, psif)  [src/Solvers/Solvers_0_functions.jl:65-73] */
int fvh_flux(const fvh_mesh *m, const double *psif_x, const double *psif_y,
             const double *psif_z, double *phif);
/* ------------------------------------------------------------------ */
/* Output and post-processing                                          */
/* ------------------------------------------------------------------ */
/* write_vtk(name, mesh, args...): legacy ASCII VTK of cell data in the
 * reference's layout, numbers printed as Julia prints Float64.  Host
 * arrays: xyz [3*npoints], cell nodes zero based; field k is a scalar
 * (ncomp 1: data[k][ncells]) or a vector (ncomp 3: x, y, z blocks of
 * ncells)  [src/VTK/VTK_writer.jl:22-116] */
int fvh_write_vtk(const char *path, int64_t npoints, const double *xyz, int64_t ncells,
                  const int64_t *cell_nodes_ptr, const int64_t *cell_nodes, int32_t nfields,
                  const char *const *names, const int32_t *ncomp, const double *const *data);
/* The number format of fvh_write_vtk: v as Julia's print(::Float64) writes
 * it, into buf (at least 32 bytes). */
int fvh_format_julia(double v, char *buf, int64_t len);
/* Faces of boundary patch `patch` (zero based), in patch order. */
int fvh_patch_size(const fvh_mesh *m, int32_t patch, int64_t *nfaces);
/* surface_flux(snflux, facesID, cellsID, phi): phi[cell] area normal per
 * patch face; device arrays  [src/Calculate/Calculate_5_surface_normal_gradient.jl:3-18] */
int fvh_surface_flux(const fvh_mesh *m, int32_t patch, const double *phi, double *fx, double *fy, double *fz);
/* surface_normal_gradient(snGrad, facesID, cellsID, U, Uw): the part of
 * U[cell] - Uw parallel to the face over delta; uw: host double[3]
 * [src/Calculate/Calculate_5_surface_normal_gradient.jl:20-36] */
int fvh_surface_normal_gradient(const fvh_mesh *m, int32_t patch, const double *ux, const double *uy,
                                const double *uz, const double *uw, double *gx, double *gy, double *gz);
/* pressure_force(patch, p, rho) -> force[3] (host), summed over the patch
 * faces in order  [src/Postprocess/Postprocess_forces.jl:5-27] */
int fvh_pressure_force(const fvh_mesh *m, int32_t patch, const double *p, double rho, double *force);
/* viscous_force(patch, U, rho, nu, nut) -> force[3] (host); uw = the
 * patch's velocity BC value (host double[3]) or NULL when no velocity BC
 * names the patch (the force is then zero); nut: device cell array or NULL
 * (zero)  [src/Postprocess/Postprocess_forces.jl:29-60] */
int fvh_viscous_force(const fvh_mesh *m, int32_t patch, const double *ux, const double *uy, const double *uz,
                      const double *uw, double rho, double nu, const double *nut, double *force);
/* wall_shear_stress(patch, model): tau = snGrad (nu + nut[cell]) per patch
 * face (device arrays) and the face centres (device [3*nfaces], may be
 * NULL)  [src/Postprocess/Postprocess_forces.jl:62-96] */
int fvh_wall_shear_stress(const fvh_mesh *m, int32_t patch, const double *ux, const double *uy, const double *uz,
                          const double *uw, double nu, const double *nut, double *tx, double *ty, double *tz,
                          double *pos);

/* bound!(field, bound): in place, with the reference's in-order semantics
 * (a negative cell takes the mean of max(neighbour, eps) over its
 * neighbours, seeing those bounded earlier in the same pass)
 * [src/RANSModels/RANSModels_functions.jl:62-88].  field: device array of
 * ncells.  *chained (may be NULL) = cells that needed the ordered pass. */
int fvh_bound(const fvh_mesh *m, double *field, double bound, int32_t *chained);

/* Non-orthogonal correction operators.  The reference keeps them behind a
 * switch that is off (`correct = false`, Solvers_1_SIMPLE.jl:165-180) and
 * defines their interpolation weight for the Linear scheme only; these
 * entries use that weight whatever the gradient scheme.
 * interpolate!(gradf, grad, phi): internal faces only, device arrays
 * [src/Calculate/Calculate_2_interpolation.jl:152-176,
 *  Calculate_3_orthogonality_correction.jl:81-93] */
int fvh_interpolate_gradient(const fvh_mesh *m, const double *gx, const double *gy, const double *gz,
                             const double *phi, double *gfx, double *gfy, double *gfz);
/* nonorthogonal_flux!(phif, gradf): phif = gradf . (n - e) area on internal
 * faces (boundary faces untouched)  [Calculate_3_orthogonality_correction.jl:66-79] */
int fvh_nonorthogonal_flux(const fvh_mesh *m, const double *gfx, const double *gfy, const double *gfz,
                           double *phif);

/* ------------------------------------------------------------------ */
/* Equation: A (CSR values on the mesh pattern), b, R, Fx on the device  */
/* [Equation{Ti,Tf}, ModelFramework_0_types.jl:83-101]                  */
/* ------------------------------------------------------------------ */
typedef struct fvh_eqn fvh_eqn;

int fvh_eqn_create(fvh_mesh *m, fvh_eqn **out);
void fvh_eqn_destroy(fvh_eqn *e);
/* Device pointers to nzval (CSR order, nnz) and b (ncells). */
int fvh_eqn_arrays(fvh_eqn *e, double **nzval, double **b);
/* discretise!(eqn, prev, runtime)  [src/Discretise/Discretise_2_generated_distretisation.jl:3-81]
 * prev: device cell array (NULL = 0); dt: runtime.dt */
int fvh_discretise(fvh_eqn *e, const fvh_term *terms, int32_t nterms,
                   const fvh_source *srcs, int32_t nsrc, const double *prev, double dt);
/* apply_boundary_conditions!(eqn, BCs)  [src/Discretise/Discretise_4_apply_bcs.jl:3-44,
 * Discretise_3_boundary_conditions.jl:4-159] */
int fvh_apply_bcs(fvh_eqn *e, const fvh_term *terms, int32_t nterms,
                  const fvh_bc *bcs, int32_t nbc, int32_t component);
/* implicit_relaxation!(eqn, field, alpha)  [src/Solve/Solve_1_api.jl:54-60] */
int fvh_implicit_relax(fvh_eqn *e, const double *phi, double alpha);
/* setReference!(eqn, pref, cellID)  [src/Solve/Solve_1_api.jl:62-69]; cell is 0-based */
int fvh_set_reference(fvh_eqn *e, double pref, int64_t cell);
/* correct!(eqn, term, corr_flux): b[owner] += -sign J_f corr_f and
 * b[neighbour] -= the same, over every face in id order (device arrays J,
 * corr of nfaces)  [src/Calculate/Calculate_3_orthogonality_correction.jl:6-24] */
int fvh_nonorthogonal_correct(fvh_eqn *e, int32_t sign, const double *J, const double *corr);
/* mul!(y, A, x)  [SparseArrays CSC mul!, used at Solvers_0_functions.jl:18] */
int fvh_spmv(fvh_eqn *e, const double *x, double *y);
/* residual!(R, eqn, phi, it): sqrt(mean((Ax-b)^2))/norm(b) -> *res (host)
 * [src/Solvers/Solvers_0_functions.jl:13-63] */
int fvh_residual(fvh_eqn *e, const double *phi, double *res);
/* fvh_residual without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The chunk partials
 * and the result live in a file-scope device array of 128 KB that each GPU
 * holds until the process ends; at most 16,777,216 cells (FVH_ERR_SIZE
 * beyond).  Same bits as fvh_residual.  Calls that overlap on one GPU share
 * that array and race. */
int fvh_residual_noalloc(fvh_eqn *e, const double *phi, double *res);

/* set_preconditioner / update_preconditioner!  [src/Solve/Preconditioners/
 * preconditioners_2_functions.jl:4-72].  `matrix_of` names the equation whose
 * A the preconditioner reads (the reference aliases uy's preconditioner to
 * ux's matrix, Solvers_1_SIMPLE.jl:45); NULL = e itself. */
int fvh_precond_set(fvh_eqn *e, int32_t kind, fvh_eqn *matrix_of);
int fvh_precond_update(fvh_eqn *e);
/* ldiv!(y, P, x)  [preconditioners_1_DILU.jl:164-167]; device pointers */
int fvh_precond_apply(fvh_eqn *e, double *y, const double *x);
/* Host copy of the DILU reciprocal diagonal (ncells). */
int fvh_precond_dinv(fvh_eqn *e, double *dinv_host);
/* The pipelined DILU sweeps bound every wait between work-groups; a wait
 * that runs out leaves a pending fault, which the next call that ran the
 * sweeps (fvh_precond_update / _apply / _apply_f32, fvh_solve, fvh_time_op)
 * returns as FVH_ERR_HIP, once.  This call reads the pending flag and, with
 * raise != 0, sets it, so a caller can exercise its error path.  No
 * counterpart in the reference (whose sweeps are sequential);
 * FVH_ERR_UNSUPPORTED when the preconditioner has no pipelined sweeps. */
int fvh_precond_fault(fvh_eqn *e, int32_t raise, int32_t *pending);

/* run!(eqn, setup): Krylov solve of A x = b warm-started from x (device,
 * overwritten)  [src/Solve/Solve_1_api.jl:32-46; Krylov.jl 0.9.5 solve!] */
int fvh_solve(fvh_eqn *e, const fvh_solver_cfg *cfg, double *x, fvh_stats *stats);

/* FP32 operators of the mixed-precision inner solves (no counterpart in the
 * reference, which runs in FP64 only): y = A x and y = M^{-1} x on FP32
 * copies of A and of the preconditioner's data, taken from the current FP64
 * values (DILU reciprocals computed in FP64 by fvh_precond_update, then
 * rounded).  x, y: device arrays of ncells floats. */
int fvh_spmv_f32(fvh_eqn *e, float *y, const float *x);
int fvh_precond_apply_f32(fvh_eqn *e, float *y, const float *x);

/* ------------------------------------------------------------------ */
/* Pressure-velocity loops                                             */
/* ------------------------------------------------------------------ */
typedef struct fvh_sim fvh_sim;

/* simple!(model, config; pref) / piso!: set-up part.  Field arrays are
 * host, length ncells; k/omega/nut may be NULL for a laminar model. */
int fvh_sim_create(fvh_mesh *m, const fvh_case *c,
                   const double *Ux, const double *Uy, const double *Uz,
                   const double *p, const double *k, const double *omega,
                   const double *nut, fvh_sim **out);
/* simple!/piso! over the independent cases of a split mesh (laminar): one
 * viscosity per case, the setReference! cell of each case (0-based within
 * the case, -1 = none; the value is c->pref).  Fields are host arrays over
 * all cells.  Each iteration advances every case; fvh_sim_iterate stops when
 * all have reached the convergence break. */
int fvh_sim_create_cases(fvh_mesh *m, const fvh_case *c, const double *nu_case,
                         const int64_t *ref_cells, const double *Ux, const double *Uy,
                         const double *Uz, const double *p, fvh_sim **out);
/* Residual histories and Krylov statistics of case k. */
int fvh_sim_case_residuals(fvh_sim *s, int32_t k, double *R_ux, double *R_uy, double *R_p);
int fvh_sim_case_stats(fvh_sim *s, int32_t k, const char *field, fvh_stats *out);
void fvh_sim_destroy(fvh_sim *s);
/* Runs up to `n` more iterations of the loop body
 * [Solvers_1_SIMPLE.jl:125-230 / Solvers_2_PISO.jl:125-237]; stops early
 * at the reference's convergence break.  Returns iterations done in *done. */
int fvh_sim_iterate(fvh_sim *s, int32_t n, int32_t *done);
/* Residual histories R_ux, R_uy, R_p (host, length case.iterations,
 * entries not reached keep 1.0 as in Solvers_1_SIMPLE.jl:110-112). */
int fvh_sim_residuals(fvh_sim *s, double *R_ux, double *R_uy, double *R_p);
/* Host copy of a named state array; *len receives the length.  Names:
 * cell fields Ux Uy Uz p k omega nut nueff rD divHv Hv_x Hv_y gradp_x
 * gradp_y gradp_z gradU_xx gradU_yx gradU_xy gradU_yy Pk nueffk nueffw;
 * face fields mdotf rDf pf Uf_x Uf_y Uf_z nutf; systems ux_b uy_b p_b and
 * ux_A uy_A p_A (CSR order); DILU reciprocals ux_dinv p_dinv. */
int fvh_sim_get(fvh_sim *s, const char *name, double *out, int64_t *len);
/* Krylov statistics of the last solve of field "Ux","Uy","p","k","omega". */
int fvh_sim_stats(fvh_sim *s, const char *field, fvh_stats *out);
/* Device time of the phases of the last iteration (ms), NaN if untimed. */
int fvh_sim_timing(fvh_sim *s, double *ms_assembly, double *ms_solve, double *ms_other);
/* The equation handle of "ux", "uy" or "p" (owned by the sim). */
int fvh_sim_equation(fvh_sim *s, const char *name, fvh_eqn **out);
/* Measurement helper (no counterpart in the reference): average device time
 * of one launch of a loop phase on the simulation's current fields, by HIP
 * events on the library stream.  op 0 = the Ux momentum assembly as the loop
 * runs it (discretise!, sources, BCs, relaxation), 1 = the pressure
 * equation's assembly, 2 = the Green-Gauss kernel on p's face values.  The
 * phase's outputs are recomputed in place from the current fields. */
int fvh_sim_time_op(fvh_sim *s, int32_t op, int32_t reps, double *ms_per_launch);

/* Measurement helper (no counterpart in the reference): average device time
 * of one launch of an operator over `reps` back-to-back launches, by HIP
 * events on the library stream.  op: 0 = SpMV y = A x (FP64), 1 =
 * preconditioner apply y = M^{-1} x, 2 = preconditioner update, 3 = FP32
 * SpMV on the single-precision copy of A, 4 = FP32 preconditioner apply
 * (ops 3/4 round x to FP32 first, all ones when x is NULL; y is unused),
 * 5 = plain copy y = x of ncells doubles (16 bytes per cell moved: the
 * reference point for hardware traffic counters), 6 / 7 = the banded SpMV
 * that solves on a structured grid use (FP64 / FP32, A read as five values
 * per row; FVH_ERR_UNSUPPORTED on other meshes; op 7 rounds x like op 3). */
int fvh_time_op(fvh_eqn *e, int32_t op, int32_t reps, const double *x, double *y, double *ms_per_launch);

/* simple!(model, config; pref): whole run, fields in/out (host). */
int fvh_simple(fvh_mesh *m, const fvh_case *c,
               double *Ux, double *Uy, double *Uz, double *p,
               double *k, double *omega, double *nut,
               double *R_ux, double *R_uy, double *R_p, int32_t *iterations_done);

#ifdef __cplusplus
}
#endif
#endif /* FVMHIP_H */
