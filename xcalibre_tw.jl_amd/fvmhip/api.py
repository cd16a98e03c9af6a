"""Host-side mirror of the reference's user API, driving libfvmhip.

Names, argument meanings and defaults follow XCALibre_TW.jl:
  RANS{Laminar}/RANS{KOmega}(mesh, viscosity) ... src/RANSModels/RANSModels_0_types.jl:22,
                                                  src/RANSModels/RANSModels_kOmega.jl:9
  @assign! model U (Dirichlet(:inlet, v), ...) .. src/Discretise/Discretise_0_types.jl:80-169
  set_schemes(time, divergence, laplacian, gradient)  Discretise_0_types.jl:171-181
  set_solver(field; solver, preconditioner, convergence, relax, itmax=100,
             atol=sqrt(eps), rtol=1e-3) ............. src/Solve/Solve_1_api.jl:5-26
  set_runtime(iterations, write_interval, time_step) src/Solve/Solve_1_api.jl:28-30
  Configuration(schemes, solvers, runtime) ......... src/Simulate/Simulate_0_types.jl:3-7
  initialise!(field, value) ........................ src/Fields/Fields_0_types.jl:205-221
  simple!(model, config; pref) -> (R_ux, R_uy, R_p)  src/Solvers/Solvers_1_SIMPLE.jl:3
  piso!(model, config; pref) ....................... src/Solvers/Solvers_2_PISO.jl:3
Julia's `!` and `:symbol` become plain names and strings.  Errors the
reference throws as strings are raised as ValueError.
"""
import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

# ---- schemes -------------------------------------------------------------
Steady, Euler, Linear, Upwind, Orthogonal, Midpoint = L.STEADY, L.EULER, L.LINEAR, L.UPWIND, L.ORTHOGONAL, L.MIDPOINT


def set_schemes(time=Steady, divergence=Linear, laplacian=Linear, gradient=Orthogonal):
    return dict(time=time, divergence=divergence, laplacian=laplacian, gradient=gradient)


# ---- solvers and preconditioners ----------------------------------------
BicgstabSolver, CgSolver, GmresSolver = L.BICGSTAB, L.CG, L.GMRES


class PreconditionerType:
    kind = None


class DILU(PreconditionerType):
    kind = L.PRECOND_DILU


class ILU0(PreconditionerType):
    kind = L.PRECOND_ILU0


class Jacobi(PreconditionerType):
    kind = L.PRECOND_JACOBI


class NormDiagonal(PreconditionerType):
    kind = L.PRECOND_NORMDIAG


class LDL(PreconditionerType):
    """LDL' factor of the equation's set-up matrix (opLDL over
    LDLFactorizations.jl), kept for every later solve."""
    kind = L.PRECOND_LDL


def set_solver(field, *, solver, preconditioner, convergence, relax, itmax=100, atol=None, rtol=1e-3, precision=L.FP64):
    if atol is None:
        atol = math.sqrt(np.finfo(np.float64).eps)
    if isinstance(preconditioner, type):
        preconditioner = preconditioner()
    if preconditioner.kind is None:
        raise ValueError(f"{type(preconditioner).__name__} preconditioner is not supported by the device path")
    return dict(solver=solver, preconditioner=preconditioner, convergence=float(convergence), relax=float(relax),
                itmax=int(itmax), atol=float(atol), rtol=float(rtol), precision=int(precision))


def set_runtime(*, iterations, write_interval, time_step):
    return dict(iterations=int(iterations), dt=float(time_step), write_interval=int(write_interval))


@dataclass
class Configuration:
    schemes: dict
    solvers: dict
    runtime: dict


# ---- fields and boundary conditions --------------------------------------
@dataclass
class ConstantScalar:
    values: float


WALL_DEFAULTS = dict(kappa=0.41, beta1=0.075, cmu=0.09, B=5.2, E=9.8)  # Discretise_0_types.jl:61,69,77


@dataclass
class BoundaryCondition:
    ID: str
    value: object = 0.0
    kind: int = field(default=L.BC_DIRICHLET, repr=False)


def Dirichlet(ID, value):
    if isinstance(value, (list, tuple, np.ndarray)) and len(value) != 3:
        raise ValueError("Only vectors with three components can be used")
    return BoundaryCondition(str(ID).lstrip(":"), value, L.BC_DIRICHLET)


def Neumann(ID, value=0.0):
    return BoundaryCondition(str(ID).lstrip(":"), value, L.BC_NEUMANN)


def KWallFunction(ID, value=None):
    return BoundaryCondition(str(ID).lstrip(":"), dict(value or WALL_DEFAULTS), L.BC_KWALL)


def OmegaWallFunction(ID, value=None):
    return BoundaryCondition(str(ID).lstrip(":"), dict(value or WALL_DEFAULTS), L.BC_OMEGAWALL)


def NutWallFunction(ID, value=None):
    return BoundaryCondition(str(ID).lstrip(":"), dict(value or WALL_DEFAULTS), L.BC_NUTWALL)


class ScalarField:
    def __init__(self, mesh):
        self.mesh = mesh
        self.values = np.zeros(mesh.ncells)
        self.BCs = ()


class VectorField:
    def __init__(self, mesh):
        self.mesh = mesh
        self.x, self.y, self.z = ScalarField(mesh), ScalarField(mesh), ScalarField(mesh)
        self.BCs = ()


def initialise(f, value):
    """initialise!(field, value): scalar or 3-vector."""
    if isinstance(f, VectorField):
        if len(value) != 3:
            raise ValueError("Vectors should have 3 components")
        f.x.values[:] = value[0]
        f.y.values[:] = value[1]
        f.z.values[:] = value[2]
    else:
        f.values[:] = value


class Laminar:
    pass


class KOmega:
    pass


class RANS:
    """RANS{Laminar}(mesh=..., viscosity=...) / RANS{KOmega}(...)."""

    def __init__(self, model, *, mesh, viscosity):
        self.model = model
        self.mesh = mesh
        self.U = VectorField(mesh)
        self.p = ScalarField(mesh)
        self.nu = viscosity
        if model is KOmega:
            self.turbulence = dict(k=ScalarField(mesh), omega=ScalarField(mesh), nut=ScalarField(mesh))
        else:
            self.turbulence = None


def isturbulent(model):
    return model.model is KOmega


def assign(model, name, *bcs):
    """@assign! model U (...) / @assign! model turbulence k (...)."""
    if name in ("U", "p"):
        f = getattr(model, name)
    else:
        f = model.turbulence[name]
    for bc in bcs:
        model.mesh.boundary_index(bc.ID)  # throws for unknown patches, as boundary_index does
    f.BCs = tuple(bcs)
    return model


# ---- translation to the C case description --------------------------------
def _bcs(mesh, bcs, vector):
    arr = (L.BC * max(1, len(bcs)))()
    for i, bc in enumerate(bcs):
        arr[i].kind = bc.kind
        arr[i].patch = mesh.boundary_index(bc.ID)
        if isinstance(bc.value, dict):
            for j, key in enumerate(("kappa", "beta1", "cmu", "B", "E")):
                arr[i].wf[j] = float(bc.value[key])
        elif isinstance(bc.value, (list, tuple, np.ndarray)):
            for j in range(3):
                arr[i].value[j] = float(bc.value[j])
        else:
            for j in range(3):
                arr[i].value[j] = float(bc.value)
    return arr


def _schemes(s):
    out = L.Schemes()
    out.time, out.divergence, out.laplacian, out.gradient = s["time"], s["divergence"], s["laplacian"], s["gradient"]
    return out


def _solver(v):
    out = L.SolverCfg()
    out.solver = v["solver"]
    out.precond = v["preconditioner"].kind
    out.itmax = v["itmax"]
    out.precision = v.get("precision", L.FP64)
    out.convergence = v["convergence"]
    out.relax = v["relax"]
    out.atol = v["atol"]
    out.rtol = v["rtol"]
    return out


class CaseDesc:
    """fvh_case for (model, config; pref) plus the buffers it points into."""

    def __init__(self, model, config, pref=None, algorithm=L.ALG_SIMPLE, conv_break=1e-7, nonorth_correctors=0):
        mesh = model.mesh
        c = L.Case()
        c.nonorth_correctors = int(nonorth_correctors)
        c.model = L.KOMEGA if isturbulent(model) else L.LAMINAR
        c.algorithm = algorithm
        c.iterations = config.runtime["iterations"]
        c.has_pref = 0 if pref is None else 1
        c.pref = 0.0 if pref is None else float(pref)
        nu = model.nu.values if isinstance(model.nu, ConstantScalar) else float(model.nu)
        c.nu = float(nu)
        c.dt = config.runtime["dt"]
        c.conv_break = conv_break
        self._keep = []

        def bcset(bcs):
            arr = _bcs(mesh, bcs, True)
            self._keep.append(arr)
            s = L.BCSet()
            s.nbc = len(bcs)
            s.bcs = C.cast(arr, C.POINTER(L.BC))
            return s

        c.U = bcset(model.U.BCs)
        c.p = bcset(model.p.BCs)
        sc, sv = config.schemes, config.solvers
        c.sU, c.sp = _schemes(sc["U"]), _schemes(sc["p"])
        c.vU, c.vp = _solver(sv["U"]), _solver(sv["p"])
        if c.model == L.KOMEGA:
            t = model.turbulence
            c.k, c.omega, c.nut = bcset(t["k"].BCs), bcset(t["omega"].BCs), bcset(t["nut"].BCs)
            c.sk, c.somega = _schemes(sc["k"]), _schemes(sc["omega"])
            c.vk, c.vomega = _solver(sv["k"]), _solver(sv["omega"])
        else:
            c.k = c.omega = c.nut = bcset(())
        self.c = c


def _fields(model):
    n = model.mesh.ncells
    U, p = model.U, model.p
    out = dict(Ux=U.x.values, Uy=U.y.values, Uz=U.z.values, p=p.values)
    if isturbulent(model):
        t = model.turbulence
        out.update(k=t["k"].values, omega=t["omega"].values, nut=t["nut"].values)
    for k, v in out.items():
        out[k] = np.ascontiguousarray(v, np.float64)
        assert out[k].size == n
    return out


class Simulation:
    """The loop set-up of simple!/piso! on the device, steppable.

    `run(n)` executes up to n loop iterations (stopping at the reference's
    convergence break); `get(name)` copies a state array to the host."""

    def __init__(self, model, config, pref=None, algorithm=L.ALG_SIMPLE, conv_break=1e-7, nonorth_correctors=0):
        self.model, self.config = model, config
        self.case = CaseDesc(model, config, pref, algorithm, conv_break, nonorth_correctors)
        self.dm = model.mesh.device()
        f = _fields(model)
        self._f = f
        h = C.c_void_p()
        g = lambda k: L.dptr(f[k]) if k in f else None  # noqa: E731
        L.check(L.lib().fvh_sim_create(self.dm.h, C.byref(self.case.c), g("Ux"), g("Uy"), g("Uz"), g("p"), g("k"),
                                       g("omega"), g("nut"), C.byref(h)), "fvh_sim_create")
        self.h = h
        self.iterations = self.case.c.iterations

    def run(self, n):
        done = C.c_int32()
        L.check(L.lib().fvh_sim_iterate(self.h, int(n), C.byref(done)), "fvh_sim_iterate")
        return done.value

    def residuals(self):
        R = [np.ones(self.iterations) for _ in range(3)]
        L.check(L.lib().fvh_sim_residuals(self.h, *(L.dptr(r) for r in R)))
        return tuple(R)

    def get(self, name):
        ln = C.c_int64()
        L.check(L.lib().fvh_sim_get(self.h, name.encode(), None, C.byref(ln)), f"fvh_sim_get({name})")
        out = np.zeros(ln.value)
        L.check(L.lib().fvh_sim_get(self.h, name.encode(), L.dptr(out), C.byref(ln)))
        return out

    def stats(self, field_name):
        s = L.Stats()
        L.check(L.lib().fvh_sim_stats(self.h, field_name.encode(), C.byref(s)))
        return dict(niter=s.niter, status=L.KRYLOV_STATUS.get(s.status, s.status), rnorm=s.rnorm, rnorm0=s.rnorm0,
                    refinements=s.refinements)

    def write_back(self):
        m = self.model
        m.U.x.values[:] = self.get("Ux")
        m.U.y.values[:] = self.get("Uy")
        m.p.values[:] = self.get("p")
        if isturbulent(m):
            for k in ("k", "omega", "nut"):
                m.turbulence[k].values[:] = self.get(k)

    def __del__(self):
        try:
            if self.h:
                L.lib().fvh_sim_destroy(self.h)
        except Exception:
            pass


class BatchSimulation:
    """simple!/piso! over many independent laminar cases advanced together:
    one launch per phase covers every case, the Krylov solves keep one state
    per case and the DILU sweeps run one work-group per case.

    `models`: RANS{Laminar} models with their U and p BCs assigned; `prefs`:
    the setReference! value per case (None = none; the non-None values must
    agree).  Cases with the same patch names and the same BCs form a family
    whose patches are merged in the union mesh.  Each case computes exactly
    what it computes in its own Simulation (FP64: bit for bit)."""

    def __init__(self, models, config, prefs=None, algorithm=L.ALG_SIMPLE, conv_break=1e-7):
        from .mesh import union

        prefs = list(prefs) if prefs is not None else [None] * len(models)
        for mo in models:
            if isturbulent(mo):
                raise FVHError_unsupported("independent cases: laminar models only")

        def key(mo):
            def bck(bcs):
                return tuple((b.ID, b.kind, repr(b.value)) for b in bcs)
            return (tuple(mo.mesh.boundary_names), bck(mo.U.BCs), bck(mo.p.BCs))

        fams, fam_of = {}, []
        for mo in models:
            fam_of.append(fams.setdefault(key(mo), f"f{len(fams)}"))
        umesh = union([mo.mesh for mo in models], fam_of)
        first = {}
        for mo, f in zip(models, fam_of):
            first.setdefault(f, mo)
        ubc = dict(U=[], p=[])
        for f, mo in first.items():
            for name in ("U", "p"):
                for b in getattr(mo, name).BCs:
                    ubc[name].append(BoundaryCondition(f"{f}:{b.ID}", b.value, b.kind))
        nus = [float(mo.nu.values if isinstance(mo.nu, ConstantScalar) else mo.nu) for mo in models]
        um = RANS(Laminar, mesh=umesh, viscosity=ConstantScalar(nus[0]))
        for comp in ("x", "y", "z"):
            getattr(um.U, comp).values[:] = np.concatenate([getattr(mo.U, comp).values for mo in models])
        um.p.values[:] = np.concatenate([mo.p.values for mo in models])
        assign(um, "U", *ubc["U"])
        assign(um, "p", *ubc["p"])
        pv = {float(v) for v in prefs if v is not None}
        if len(pv) > 1:
            raise ValueError("independent cases: one pressure reference value for all cases")
        pref = pv.pop() if pv else None
        self.models, self.config, self.union_model = models, config, um
        self.offsets = umesh.case_offsets
        self.case = CaseDesc(um, config, pref, algorithm, conv_break)
        self.dm = umesh.device()
        f = _fields(um)
        self._f = f
        nu = np.array(nus, np.float64)
        refs = np.array([0 if v is not None else -1 for v in prefs], np.int64)
        self._keep = (nu, refs)
        h = C.c_void_p()
        L.check(L.lib().fvh_sim_create_cases(self.dm.h, C.byref(self.case.c), L.dptr(nu),
                                             refs.ctypes.data_as(L.P_I64), L.dptr(f["Ux"]), L.dptr(f["Uy"]),
                                             L.dptr(f["Uz"]), L.dptr(f["p"]), C.byref(h)), "fvh_sim_create_cases")
        self.h = h
        self.iterations = self.case.c.iterations

    @property
    def ncases(self):
        return len(self.models)

    def run(self, n):
        done = C.c_int32()
        L.check(L.lib().fvh_sim_iterate(self.h, int(n), C.byref(done)), "fvh_sim_iterate")
        return done.value

    def get(self, name, case=None):
        ln = C.c_int64()
        L.check(L.lib().fvh_sim_get(self.h, name.encode(), None, C.byref(ln)), f"fvh_sim_get({name})")
        out = np.zeros(ln.value)
        L.check(L.lib().fvh_sim_get(self.h, name.encode(), L.dptr(out), C.byref(ln)))
        if case is None:
            return out
        if out.size != self.offsets[-1]:
            raise ValueError(f"{name} is not a cell array")
        return out[self.offsets[case]:self.offsets[case + 1]]

    def residuals(self, case):
        n = self.iterations
        a, b, c = np.ones(n), np.ones(n), np.ones(n)
        L.check(L.lib().fvh_sim_case_residuals(self.h, int(case), L.dptr(a), L.dptr(b), L.dptr(c)))
        return a, b, c

    def stats(self, field_name, case):
        s = L.Stats()
        L.check(L.lib().fvh_sim_case_stats(self.h, int(case), field_name.encode(), C.byref(s)))
        return dict(niter=s.niter, status=L.KRYLOV_STATUS.get(s.status, s.status), rnorm=s.rnorm, rnorm0=s.rnorm0,
                    refinements=s.refinements)

    def __del__(self):
        try:
            if self.h:
                L.lib().fvh_sim_destroy(self.h)
        except Exception:
            pass


class ScalarEquation:
    """One steady scalar transport equation on the device, driven the way the
    reference drives one (Discretise_2_generated_distretisation.jl:3-81,
    Discretise_4_apply_bcs.jl:3-44, Solve_1_api.jl:32-46):

        div(mdotf phi) - laplacian(gamma phi) == 0

    `mdotf` and `gamma` are face arrays (host, length nfaces); `bcs` are
    Dirichlet / Neumann conditions of phi; `divergence` is Linear or Upwind.
    discretise() assembles A and b with the BCs applied; solve() runs the
    Krylov solve from phi (FP64 or mixed precision)."""

    def __init__(self, mesh, mdotf, gamma, bcs, divergence=Linear):
        self.mesh = mesh
        self.dm = mesh.device()
        nf = mesh.nfaces
        self._mdot = L.DeviceArray.from_host(np.ascontiguousarray(mdotf, np.float64).reshape(nf))
        self._gam = L.DeviceArray.from_host(np.ascontiguousarray(gamma, np.float64).reshape(nf))
        self.phi = L.DeviceArray.from_host(np.zeros(mesh.ncells))
        h = C.c_void_p()
        L.check(L.lib().fvh_eqn_create(self.dm.h, C.byref(h)), "fvh_eqn_create")
        self.h = h
        terms = (L.Term * 2)()
        terms[0].op, terms[0].scheme, terms[0].sign = L.OP_DIVERGENCE, divergence, 1
        terms[0].flux, terms[0].phi = self._mdot.ptr, self.phi.ptr
        terms[1].op, terms[1].scheme, terms[1].sign = L.OP_LAPLACIAN, Linear, -1
        terms[1].flux, terms[1].phi = self._gam.ptr, self.phi.ptr
        self._terms = terms
        self._bcs = _bcs(mesh, bcs, False)
        self._nbc = len(bcs)
        self._pkind = None

    def discretise(self):
        L.check(L.lib().fvh_discretise(self.h, self._terms, 2, None, 0, None, 1.0), "fvh_discretise")
        L.check(L.lib().fvh_apply_bcs(self.h, self._terms, 2, self._bcs, self._nbc, 0), "fvh_apply_bcs")

    def solve(self, solver=BicgstabSolver, preconditioner=None, itmax=1000, atol=0.0, rtol=1e-10,
              precision=L.FP64):
        pre = preconditioner or DILU()
        if self._pkind != pre.kind:
            L.check(L.lib().fvh_precond_set(self.h, pre.kind, None), "fvh_precond_set")
            self._pkind = pre.kind
        L.check(L.lib().fvh_precond_update(self.h), "fvh_precond_update")
        cfg = L.SolverCfg(solver=solver, precond=pre.kind, itmax=int(itmax), precision=int(precision),
                          convergence=0.0, relax=1.0, atol=float(atol), rtol=float(rtol))
        st = L.Stats()
        L.check(L.lib().fvh_solve(self.h, C.byref(cfg), self.phi.vp, C.byref(st)), "fvh_solve")
        return dict(niter=st.niter, status=L.KRYLOV_STATUS.get(st.status, st.status), rnorm=st.rnorm,
                    rnorm0=st.rnorm0, refinements=st.refinements)

    def residual(self):
        r = C.c_double()
        L.check(L.lib().fvh_residual(self.h, self.phi.vp, C.byref(r)))
        return r.value

    def values(self):
        return self.phi.host()

    def matrix(self):
        """Host copies of A (CSR values on the mesh pattern) and b."""
        nz, b = C.c_void_p(), C.c_void_p()
        L.check(L.lib().fvh_eqn_arrays(self.h, C.byref(nz), C.byref(b)))
        rp, ci = self.dm.pattern()
        vals = np.zeros(ci.size)
        rhs = np.zeros(self.mesh.ncells)
        L.check(L.lib().fvh_memcpy(vals.ctypes.data_as(C.c_void_p), nz, vals.nbytes))
        L.check(L.lib().fvh_memcpy(rhs.ctypes.data_as(C.c_void_p), b, rhs.nbytes))
        return vals, rhs

    def __del__(self):
        try:
            if self.h:
                L.lib().fvh_eqn_destroy(self.h)
        except Exception:
            pass


def taylor_green_flux(mesh):
    """Face mass flux of u = (sin(pi x) cos(pi y), -cos(pi x) sin(pi y)) at the
    face centres: mdotf = u_f . S_f."""
    x, y = mesh.face_centre[:, 0], mesh.face_centre[:, 1]
    u = np.sin(np.pi * x) * np.cos(np.pi * y)
    v = -np.cos(np.pi * x) * np.sin(np.pi * y)
    S = mesh.face_area[:, None] * mesh.face_normal
    return u * S[:, 0] + v * S[:, 1]


def FVHError_unsupported(msg):
    return L.FVHError(L.FVH_ERR_UNSUPPORTED, msg)


def _run_writing(sim, model, config, prefix, algorithm):
    """The loop with the reference's output: model2vtk every write_interval
    iterations (`iteration % write_interval + signbit(write_interval) == 0`)
    and, for simple!, at convergence when write_interval >= 0
    (Solvers_1_SIMPLE.jl:200-228, Solvers_2_PISO.jl:228-235)."""
    wi = int(config.runtime.get("write_interval", -1))
    if wi == 0:
        raise ZeroDivisionError("write_interval = 0 (the reference's iteration % 0 throws DivideError)")
    total = sim.iterations
    if wi < 0:
        sim.run(total)
        sim.write_back()
        return
    done = 0
    while done < total:
        want = min(wi - done % wi, total - done)
        k = sim.run(want)
        done += k
        sim.write_back()
        if k < want:  # the loop broke at its convergence test
            if algorithm == L.ALG_SIMPLE:
                model2vtk(model, f"{prefix}_{done:06d}")
            return
        if done % wi == 0:
            model2vtk(model, f"{prefix}_{done:06d}")


def simple(model, config, pref=None, conv_break=1e-7):
    """simple!(model, config; pref): returns (R_ux, R_uy, R_p); fields are
    updated in place on the model, as in the reference; VTK files
    iteration_NNNNNN.vtk as write_interval asks."""
    sim = Simulation(model, config, pref, L.ALG_SIMPLE, conv_break)
    _run_writing(sim, model, config, "iteration", L.ALG_SIMPLE)
    return sim.residuals()


def piso(model, config, pref=None):
    """piso!(model, config; pref); VTK files timestep_NNNNNN.vtk as
    write_interval asks."""
    sim = Simulation(model, config, pref, L.ALG_PISO)
    _run_writing(sim, model, config, "timestep", L.ALG_PISO)
    return sim.residuals()


# ---- output and post-processing (src/VTK/VTK_writer.jl, src/Postprocess/Postprocess_forces.jl) ----
def _field_of(value, n):
    """(ncomp, flat array) of a ScalarField / VectorField / ndarray of n or (n, 3) values."""
    if isinstance(value, VectorField):
        return 3, np.concatenate([value.x.values, value.y.values, value.z.values]).astype(np.float64)
    if isinstance(value, ScalarField):
        return 1, np.ascontiguousarray(value.values, np.float64)
    raise TypeError('Input data should be a ScalarField or VectorField e.g. ("U", U)')


def write_vtk(name, mesh, *args):
    """write_vtk(name, mesh, ("U", U), ("p", p), ...): `name`.vtk, legacy ASCII
    cell data in the reference's layout (src/VTK/VTK_writer.jl:22-116)."""
    if getattr(mesh, "nodes", None) is None:
        raise ValueError("this mesh carries no nodes (a batched union); write each case with its own mesh")
    n = mesh.ncells
    names, comps, datas = [], [], []
    for label, field in args:
        k, d = _field_of(field, n)
        if d.size != k * n:
            raise ValueError(f"{label}: {d.size} values for {n} cells")
        names.append(str(label).encode())
        comps.append(k)
        datas.append(d)
    cn = (C.c_char_p * max(1, len(names)))(*names)
    cc = (C.c_int32 * max(1, len(comps)))(*comps)
    cd = (C.POINTER(C.c_double) * max(1, len(datas)))(*[L.dptr(d) for d in datas])
    L.check(L.lib().fvh_write_vtk((str(name) + ".vtk").encode(), mesh.nodes.shape[0], L.dptr(mesh.nodes.reshape(-1)),
                                  n, L.iptr(mesh.cell_nodes_ptr), L.iptr(mesh.cell_nodes), len(names), cn, cc, cd),
            "fvh_write_vtk")
    return str(name) + ".vtk"


def model2vtk(model, name):
    """model2vtk(model, name): U and p, plus k, omega and nut for k-omega
    (src/VTK/VTK_writer.jl:3-20)."""
    args = [("U", model.U), ("p", model.p)]
    if isturbulent(model):
        t = model.turbulence
        args += [("k", t["k"]), ("omega", t["omega"]), ("nut", t["nut"])]
    return write_vtk(name, model.mesh, *args)


def _dev_cells(values, n):
    a = np.ascontiguousarray(values, np.float64)
    if a.size != n:
        raise ValueError(f"{a.size} values for {n} cells")
    return L.DeviceArray.from_host(a)


def _wall_velocity(U, patch_id):
    """Value of the last velocity BC that names the patch, or None."""
    uw = None
    for bc in U.BCs:
        if bc.ID == patch_id:
            uw = bc.value
    return uw


def surface_flux(patch, phi):
    """surface_flux over a patch: phi[cell] * (area * normal) per face
    (src/Calculate/Calculate_5_surface_normal_gradient.jl:3-18); (x, y, z)."""
    mesh = phi.mesh
    b = mesh.boundary_index(patch)
    nb = C.c_int64()
    L.check(L.lib().fvh_patch_size(mesh.device().h, b, C.byref(nb)))
    out = [L.DeviceArray(max(nb.value, 1)) for _ in range(3)]
    d = _dev_cells(phi.values, mesh.ncells)
    L.check(L.lib().fvh_surface_flux(mesh.device().h, b, d.vp, *(o.vp for o in out)), "fvh_surface_flux")
    return tuple(o.host()[:nb.value] for o in out)


def pressure_force(patch, p, rho):
    """pressure_force(patch, p, rho) (Postprocess_forces.jl:5-27): rho times
    the sum over the patch of p[cell] area normal."""
    mesh = p.mesh
    b = mesh.boundary_index(patch)
    d = _dev_cells(p.values, mesh.ncells)
    F = np.zeros(3)
    L.check(L.lib().fvh_pressure_force(mesh.device().h, b, d.vp, float(rho), L.dptr(F)), "fvh_pressure_force")
    return F


def viscous_force(patch, U, rho, nu, nut):
    """viscous_force(patch, U, rho, nu, nut) (Postprocess_forces.jl:29-60):
    rho times the sum of the wall-parallel normal gradient times area times
    (nu + nut[cell]); zero when no velocity BC names the patch."""
    mesh = U.mesh
    b = mesh.boundary_index(patch)
    name = mesh.boundary_names[b]
    d = [_dev_cells(c.values, mesh.ncells) for c in (U.x, U.y, U.z)]
    uw = _wall_velocity(U, name)
    uwa = None if uw is None else np.ascontiguousarray(uw, np.float64)
    nt = None
    if nut is not None and not isinstance(nut, ConstantScalar):
        nt = _dev_cells(nut.values if isinstance(nut, ScalarField) else nut, mesh.ncells)
    nu0 = float(nu.values if isinstance(nu, ConstantScalar) else nu)
    if isinstance(nut, ConstantScalar):
        nu0 = nu0 + float(nut.values)
    F = np.zeros(3)
    L.check(L.lib().fvh_viscous_force(mesh.device().h, b, *(a.vp for a in d), None if uwa is None else L.dptr(uwa),
                                      float(rho), nu0, None if nt is None else nt.vp, L.dptr(F)), "fvh_viscous_force")
    return F


def wall_shear_stress(patch, model):
    """wall_shear_stress(patch, model) (Postprocess_forces.jl:62-96):
    ((tau_x, tau_y, tau_z), face centres) with tau = snGrad (nu + nut)."""
    mesh = model.mesh
    b = mesh.boundary_index(patch)
    name = mesh.boundary_names[b]
    uw = _wall_velocity(model.U, name)
    if uw is None:
        raise ValueError(f"no velocity boundary condition names patch {name!r}")
    uwa = np.ascontiguousarray(uw, np.float64)
    d = [_dev_cells(c.values, mesh.ncells) for c in (model.U.x, model.U.y, model.U.z)]
    nt = _dev_cells(model.turbulence["nut"].values, mesh.ncells) if isturbulent(model) else None
    nb = C.c_int64()
    L.check(L.lib().fvh_patch_size(mesh.device().h, b, C.byref(nb)))
    out = [L.DeviceArray(max(nb.value, 1)) for _ in range(3)]
    pos = L.DeviceArray(3 * max(nb.value, 1))
    nu = float(model.nu.values if isinstance(model.nu, ConstantScalar) else model.nu)
    L.check(L.lib().fvh_wall_shear_stress(mesh.device().h, b, *(a.vp for a in d), L.dptr(uwa), nu,
                                          None if nt is None else nt.vp, *(o.vp for o in out), pos.vp),
            "fvh_wall_shear_stress")
    k = nb.value
    return tuple(o.host()[:k] for o in out), pos.host()[:3 * k].reshape(k, 3)
