# FVMHIP: XCALibre_TW.jl (FVM_1D 0.2.0) methods that run the hot path on an
# MI355X through libfvmhip.so (C ABI: include/fvmhip.h).
#
#   include("FVMHIP.jl"); using .FVMHIP
#   FVMHIP.set_library!("/path/to/libfvmhip.so")
#   Rx, Ry, Rp = FVMHIP.simple_device!(model, config; pref=0.0)   # the whole loop on the GPU
#
# or, step by step, with device-resident fields (AMDGPU.jl ROCArray pointers
# or fvh_malloc'ed buffers): DeviceMesh(mesh), DeviceEquation(dmesh) and the
# discretise! / apply_boundary_conditions! / run! / residual! / grad! ...
# methods below, each a ccall of one entry point.
#
# Julia is not installed where this library was built, so this module has not
# been run; it follows the structs and entry points of include/fvmhip.h field
# by field, and the Python package fvmhip (ctypes over the same ABI) is the
# tested counterpart.
module FVMHIP

using FVM_1D
using Krylov
using SparseArrays

const LIB = Ref{String}("libfvmhip.so")
set_library!(path::AbstractString) = (LIB[] = String(path); nothing)
lib() = LIB[]

# ---- status codes and enums (fvmhip.h) -------------------------------------
const FVH_OK = 0
const OP_TIME, OP_LAPLACIAN, OP_DIVERGENCE, OP_SI = 0, 1, 2, 3
const STEADY, EULER, LINEAR, UPWIND, ORTHOGONAL, MIDPOINT = 0, 1, 2, 3, 4, 5
const BC_DIRICHLET, BC_NEUMANN, BC_KWALL, BC_OMEGAWALL, BC_NUTWALL = 0, 1, 2, 3, 4
const BICGSTAB, CG, GMRES = 0, 1, 2
const PRECOND_NONE, PRECOND_DILU, PRECOND_ILU0, PRECOND_JACOBI, PRECOND_NORMDIAG, PRECOND_LDL = 0, 1, 2, 3, 4, 5
const FP64, MIXED = 0, 1
const LAMINAR, KOMEGA = 0, 1
const SIMPLE, PISO = 0, 1

function check(st, what="fvmhip")
    st == FVH_OK && return nothing
    buf = Vector{UInt8}(undef, 1024)
    ccall((:fvh_last_error, lib()), Cint, (Ptr{UInt8}, Int64), buf, length(buf))
    error("$what: " * unsafe_string(pointer(buf)))
end

# ---- mirrors of the C structs (field order and padding as in fvmhip.h) -----
struct MeshDesc
    index_base::Int64; ncells::Int64; nfaces::Int64; nboundaries::Int64
    ncellfaces::Int64; nboundaryfaces::Int64
    cell_centre::Ptr{Float64}; cell_volume::Ptr{Float64}
    cell_faces_range::Ptr{Int64}; cell_faces::Ptr{Int64}
    cell_neighbours::Ptr{Int64}; cell_nsign::Ptr{Int64}; face_owners::Ptr{Int64}
    face_centre::Ptr{Float64}; face_normal::Ptr{Float64}; face_e::Ptr{Float64}
    face_area::Ptr{Float64}; face_delta::Ptr{Float64}; face_weight::Ptr{Float64}
    boundary_range::Ptr{Int64}; boundary_faces::Ptr{Int64}; boundary_cells::Ptr{Int64}
end
struct Term;   op::Int32; scheme::Int32; sign::Int32; pad::Int32; flux::Ptr{Float64}; phi::Ptr{Float64}; end
struct Source; sign::Int32; pad::Int32; field::Ptr{Float64}; end
struct BC;     kind::Int32; patch::Int32; value::NTuple{3,Float64}; wf::NTuple{5,Float64}; end
struct Schemes; time::Int32; divergence::Int32; laplacian::Int32; gradient::Int32; end
struct SolverCfg
    solver::Int32; precond::Int32; itmax::Int32; precision::Int32
    convergence::Float64; relax::Float64; atol::Float64; rtol::Float64
end
struct BCSet; nbc::Int32; pad::Int32; bcs::Ptr{BC}; end
struct CaseDesc
    model::Int32; algorithm::Int32; iterations::Int32; has_pref::Int32
    pref::Float64; nu::Float64; dt::Float64; conv_break::Float64
    U::BCSet; p::BCSet; k::BCSet; omega::BCSet; nut::BCSet
    sU::Schemes; sp::Schemes; sk::Schemes; somega::Schemes
    vU::SolverCfg; vp::SolverCfg; vk::SolverCfg; vomega::SolverCfg
    nonorth_correctors::Int32; pad::Int32
end
struct Stats; niter::Int32; status::Int32; refinements::Int32; pad::Int32; rnorm::Float64; rnorm0::Float64; end

# ---- Mesh2 -> flat 1-based arrays (src/Mesh/Mesh_0_types.jl:21-94) ----------
function flatten(mesh)
    (; cells, faces, boundaries, cell_faces, cell_neighbours, cell_nsign) = mesh
    n, nf = length(cells), length(faces)
    cc = Vector{Float64}(undef, 3n); vol = Vector{Float64}(undef, n); fr = Vector{Int64}(undef, 2n)
    for (i, c) in enumerate(cells)
        cc[3i-2:3i] .= c.centre
        vol[i] = c.volume
        fr[2i-1], fr[2i] = first(c.faces_range), last(c.faces_range)
    end
    own = Vector{Int64}(undef, 2nf)
    fc, fnrm, fe = (Vector{Float64}(undef, 3nf) for _ in 1:3)
    area, delta, w = (Vector{Float64}(undef, nf) for _ in 1:3)
    for (f, fa) in enumerate(faces)
        own[2f-1], own[2f] = fa.ownerCells[1], fa.ownerCells[2]
        fc[3f-2:3f] .= fa.centre; fnrm[3f-2:3f] .= fa.normal; fe[3f-2:3f] .= fa.e
        area[f], delta[f], w[f] = fa.area, fa.delta, fa.weight
    end
    brange = Int64[]; bfaces = Int64[]; bcells = Int64[]
    for b in boundaries
        push!(brange, length(bfaces) + 1)
        append!(bfaces, b.facesID); append!(bcells, b.cellsID)
        push!(brange, length(bfaces))
    end
    (; cc, vol, fr, cf=Int64.(cell_faces), cn=Int64.(cell_neighbours), cs=Int64.(cell_nsign), own, fc, fnrm, fe,
       area, delta, w, brange, bfaces, bcells, n, nf, nb=length(boundaries))
end

function mesh_desc(a)
    MeshDesc(1, a.n, a.nf, a.nb, length(a.cf), length(a.bfaces),
             pointer(a.cc), pointer(a.vol), pointer(a.fr), pointer(a.cf), pointer(a.cn), pointer(a.cs),
             pointer(a.own), pointer(a.fc), pointer(a.fnrm), pointer(a.fe), pointer(a.area), pointer(a.delta),
             pointer(a.w), pointer(a.brange), pointer(a.bfaces), pointer(a.bcells))
end

# ---- device handles -----------------------------------------------------------
mutable struct DeviceMesh
    h::Ptr{Cvoid}
    host::Any
end
function DeviceMesh(mesh)
    a = flatten(mesh)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve a check(ccall((:fvh_mesh_create, lib()), Cint, (Ref{MeshDesc}, Ref{Ptr{Cvoid}}), mesh_desc(a), h),
                         "fvh_mesh_create")
    m = DeviceMesh(h[], mesh)
    finalizer(x -> ccall((:fvh_mesh_destroy, lib()), Cvoid, (Ptr{Cvoid},), x.h), m)
end

mutable struct DeviceEquation
    h::Ptr{Cvoid}
    mesh::DeviceMesh
end
function DeviceEquation(m::DeviceMesh)       # Equation(mesh), ModelFramework_0_types.jl:90-101
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fvh_eqn_create, lib()), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), m.h, h), "fvh_eqn_create")
    e = DeviceEquation(h[], m)
    finalizer(x -> ccall((:fvh_eqn_destroy, lib()), Cvoid, (Ptr{Cvoid},), x.h), e)
end

# ---- the DSL as descriptors ----------------------------------------------------
scheme_code(::Type{Steady}) = STEADY
scheme_code(::Type{Euler}) = EULER
scheme_code(::Type{Linear}) = LINEAR
scheme_code(::Type{Upwind}) = UPWIND
scheme_code(::Type{Orthogonal}) = ORTHOGONAL
scheme_code(::Type{Midpoint}) = MIDPOINT

op_code(::Time{T}) where T = (OP_TIME, scheme_code(T))
op_code(::Laplacian{T}) where T = (OP_LAPLACIAN, scheme_code(T))
op_code(::Divergence{T}) where T = (OP_DIVERGENCE, scheme_code(T))
op_code(::Si) = (OP_SI, LINEAR)

# `dev(field)` returns the device pointer of a field's values (e.g. the
# ROCArray mirror kept by the caller); constant fluxes map to C_NULL only
# where the ABI reads them as 1 (Time).
function term_descriptors(model::Model, dev)
    terms = Term[]
    for t in model.terms
        op, sc = op_code(t.type)
        flux = t.flux isa ConstantScalar ? Ptr{Float64}(C_NULL) : dev(t.flux)
        push!(terms, Term(op, sc, Int32(t.sign), 0, flux, dev(t.phi)))
    end
    srcs = [Source(Int32(s.sign), 0, dev(s.field)) for s in model.sources]
    terms, srcs
end

wf_tuple(v) = (Float64(v.kappa), Float64(v.beta1), Float64(v.cmu), Float64(v.B), Float64(v.E))
const WF0 = (0.0, 0.0, 0.0, 0.0, 0.0)
vec3(v::Number) = (Float64(v), 0.0, 0.0)
vec3(v) = (Float64(v[1]), Float64(v[2]), Float64(v[3]))

bc_descriptor(b::Dirichlet) = BC(BC_DIRICHLET, Int32(b.ID - 1), vec3(b.value), WF0)
bc_descriptor(b::Neumann) = BC(BC_NEUMANN, Int32(b.ID - 1), vec3(b.value), WF0)
bc_descriptor(b::KWallFunction) = BC(BC_KWALL, Int32(b.ID - 1), (0.0, 0.0, 0.0), wf_tuple(b.value))
bc_descriptor(b::OmegaWallFunction) = BC(BC_OMEGAWALL, Int32(b.ID - 1), (0.0, 0.0, 0.0), wf_tuple(b.value))
bc_descriptor(b::NutWallFunction) = BC(BC_NUTWALL, Int32(b.ID - 1), (0.0, 0.0, 0.0), wf_tuple(b.value))
bc_descriptors(BCs) = BC[bc_descriptor(b) for b in BCs]

solver_kind(::Type{<:Krylov.BicgstabSolver}) = BICGSTAB
solver_kind(::Type{<:Krylov.CgSolver}) = CG
solver_kind(::Type{<:Krylov.GmresSolver}) = GMRES
precond_kind(::DILU) = PRECOND_DILU
precond_kind(::ILU0) = PRECOND_ILU0
precond_kind(::Jacobi) = PRECOND_JACOBI
precond_kind(::NormDiagonal) = PRECOND_NORMDIAG
precond_kind(::LDL) = PRECOND_LDL

solver_cfg(s; precision=FP64) = SolverCfg(solver_kind(s.solver), precond_kind(s.preconditioner), s.itmax, precision,
                                          s.convergence, s.relax, s.atol, s.rtol)
schemes_desc(s) = Schemes(scheme_code(s.time), scheme_code(s.divergence), scheme_code(s.laplacian),
                          scheme_code(s.gradient))

# fvh_case for simple!/piso! (Solvers_1_SIMPLE.jl:3-67); returns the
# descriptor and the arrays it points into (keep both alive for the call)
function case_descriptor(model, config, pref; algorithm=SIMPLE, precision=FP64, conv_break=1e-7)
    (; solvers, schemes, runtime) = config
    keep = Any[]
    function set(BCs)
        v = bc_descriptors(BCs)
        push!(keep, v)
        BCSet(Int32(length(v)), 0, isempty(v) ? Ptr{BC}(C_NULL) : pointer(v))
    end
    turb = model.turbulence !== nothing && hasproperty(model.turbulence, :k)
    empty = BCSet(0, 0, C_NULL)
    z = Schemes(0, 0, 0, 0)
    zc = SolverCfg(0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0)
    c = CaseDesc(turb ? KOMEGA : LAMINAR, algorithm, runtime.iterations, pref === nothing ? 0 : 1,
                 pref === nothing ? 0.0 : Float64(pref), Float64(model.nu.values), Float64(runtime.dt), conv_break,
                 set(model.U.BCs), set(model.p.BCs),
                 turb ? set(model.turbulence.k.BCs) : empty, turb ? set(model.turbulence.omega.BCs) : empty,
                 turb ? set(model.turbulence.nut.BCs) : empty,
                 schemes_desc(schemes.U), schemes_desc(schemes.p),
                 turb ? schemes_desc(schemes.k) : z, turb ? schemes_desc(schemes.omega) : z,
                 solver_cfg(solvers.U; precision), solver_cfg(solvers.p; precision),
                 turb ? solver_cfg(solvers.k) : zc, turb ? solver_cfg(solvers.omega) : zc, 0, 0)
    c, keep
end

# ---- per-function methods (device pointers) ---------------------------------
# discretise!(eqn, prev, runtime)   Discretise_2_generated_distretisation.jl:3-81
function FVM_1D.discretise!(eqn::DeviceEquation, model::Model, dev, prev::Ptr{Float64}, runtime)
    terms, srcs = term_descriptors(model, dev)
    check(ccall((:fvh_discretise, lib()), Cint,
                (Ptr{Cvoid}, Ptr{Term}, Int32, Ptr{Source}, Int32, Ptr{Float64}, Float64),
                eqn.h, terms, length(terms), srcs, length(srcs), prev, runtime.dt), "fvh_discretise")
end

# apply_boundary_conditions!(eqn, BCs)   Discretise_4_apply_bcs.jl:3-44
function FVM_1D.apply_boundary_conditions!(eqn::DeviceEquation, model::Model, dev, BCs, component=0)
    terms, _ = term_descriptors(model, dev)
    bcs = bc_descriptors(BCs)
    check(ccall((:fvh_apply_bcs, lib()), Cint, (Ptr{Cvoid}, Ptr{Term}, Int32, Ptr{BC}, Int32, Int32),
                eqn.h, terms, length(terms), bcs, length(bcs), component), "fvh_apply_bcs")
end

# implicit_relaxation! / setReference!   Solve_1_api.jl:54-69
FVM_1D.implicit_relaxation!(eqn::DeviceEquation, field::Ptr{Float64}, alpha) =
    check(ccall((:fvh_implicit_relax, lib()), Cint, (Ptr{Cvoid}, Ptr{Float64}, Float64), eqn.h, field, alpha))
FVM_1D.setReference!(eqn::DeviceEquation, pref::Nothing, cellID) = nothing
FVM_1D.setReference!(eqn::DeviceEquation, pref, cellID) =
    check(ccall((:fvh_set_reference, lib()), Cint, (Ptr{Cvoid}, Float64, Int64), eqn.h, pref, cellID - 1))

# set_preconditioner / update_preconditioner!   preconditioners_2_functions.jl:4-72
set_preconditioner!(eqn::DeviceEquation, P::PreconditionerType, matrix_of::Union{Nothing,DeviceEquation}=nothing) =
    check(ccall((:fvh_precond_set, lib()), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}),
                eqn.h, precond_kind(P), matrix_of === nothing ? C_NULL : matrix_of.h))
FVM_1D.update_preconditioner!(eqn::DeviceEquation) =
    check(ccall((:fvh_precond_update, lib()), Cint, (Ptr{Cvoid},), eqn.h), "fvh_precond_update")
# ldiv!(y, P, x)   preconditioners_1_DILU.jl:164-167
ldiv_device!(y::Ptr{Float64}, eqn::DeviceEquation, x::Ptr{Float64}) =
    check(ccall((:fvh_precond_apply, lib()), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), eqn.h, y, x))

# run!(eqn, setup)   Solve_1_api.jl:32-46
function FVM_1D.run!(eqn::DeviceEquation, setup, x::Ptr{Float64}; precision=FP64)
    st = Ref{Stats}()
    check(ccall((:fvh_solve, lib()), Cint, (Ptr{Cvoid}, Ref{SolverCfg}, Ptr{Float64}, Ref{Stats}),
                eqn.h, solver_cfg(setup; precision), x, st), "fvh_solve")
    st[]
end

# residual!(R, eqn, phi, iteration)   Solvers_0_functions.jl:13-63
function FVM_1D.residual!(R, eqn::DeviceEquation, phi::Ptr{Float64}, iteration)
    r = Ref{Float64}()
    check(ccall((:fvh_residual, lib()), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ref{Float64}), eqn.h, phi, r))
    R[iteration] = r[]
end

# interpolate! / correct_boundaries! / grad! / div! / flux!   Calculate_*.jl
interpolate_device!(phif::Ptr{Float64}, phi::Ptr{Float64}, m::DeviceMesh) =
    check(ccall((:fvh_interpolate, lib()), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), m.h, phi, phif))
correct_boundaries_device!(phif::Ptr{Float64}, phi::Ptr{Float64}, BCs, m::DeviceMesh, component=0) =
    check(ccall((:fvh_correct_boundaries, lib()), Cint, (Ptr{Cvoid}, Ptr{BC}, Int32, Int32, Ptr{Float64}, Ptr{Float64}),
                m.h, bc_descriptors(BCs), length(BCs), component, phi, phif))
grad_device!(dx, dy, dz, phif, phi, BCs, m::DeviceMesh, scheme::Type=Orthogonal) =
    check(ccall((:fvh_grad, lib()), Cint,
                (Ptr{Cvoid}, Int32, Ptr{BC}, Int32, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}), m.h, scheme_code(scheme), bc_descriptors(BCs), length(BCs), 0, phi, phif, dx, dy, dz))
div_device!(out::Ptr{Float64}, m::DeviceMesh, ux, uy, uz) =
    check(ccall((:fvh_div, lib()), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                m.h, ux, uy, uz, out))
flux_device!(mdotf::Ptr{Float64}, m::DeviceMesh, ux, uy, uz) =
    check(ccall((:fvh_flux, lib()), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                m.h, ux, uy, uz, mdotf))
# bound!(field, bound)   RANSModels_functions.jl:62-88
bound_device!(field::Ptr{Float64}, bnd, m::DeviceMesh) =
    check(ccall((:fvh_bound, lib()), Cint, (Ptr{Cvoid}, Ptr{Float64}, Float64, Ptr{Int32}), m.h, field, bnd, C_NULL))

# ---- the whole loop: simple!(model, config; pref)   Solvers_1_SIMPLE.jl:3-231 ----
function simple_device!(model, config; pref=nothing, algorithm=SIMPLE, precision=FP64)
    dm = DeviceMesh(model.mesh)
    case, keep = case_descriptor(model, config, pref; algorithm, precision)
    turb = case.model == KOMEGA
    ux, uy, uz, p = (copy(v) for v in (model.U.x.values, model.U.y.values, model.U.z.values, model.p.values))
    k, ω, nut = turb ? (copy(model.turbulence.k.values), copy(model.turbulence.omega.values),
                        copy(model.turbulence.nut.values)) : (Float64[], Float64[], Float64[])
    n = config.runtime.iterations
    Rx, Ry, Rp = ones(n), ones(n), ones(n)
    done = Ref{Int32}(0)
    ptr(v) = isempty(v) ? Ptr{Float64}(C_NULL) : pointer(v)
    GC.@preserve keep ux uy uz p k ω nut check(ccall((:fvh_simple, lib()), Cint,
        (Ptr{Cvoid}, Ref{CaseDesc}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{Int32}),
        dm.h, case, ux, uy, uz, p, ptr(k), ptr(ω), ptr(nut), Rx, Ry, Rp, done), "fvh_simple")
    copyto!(model.U.x.values, ux); copyto!(model.U.y.values, uy); copyto!(model.p.values, p)
    if turb
        copyto!(model.turbulence.k.values, k); copyto!(model.turbulence.omega.values, ω)
        copyto!(model.turbulence.nut.values, nut)
    end
    Rx, Ry, Rp
end

end # module
