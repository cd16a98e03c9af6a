"""The benchmark cases as reference-API set-ups (shared by tests and bench)."""
import numpy as np

import fvmhip as F


def cavity_model(mesh, nu=0.01, lid=1.0):
    """Lid-driven cavity: patches inlet/outlet/bottom walls, top = lid (CASE_UNV_Cavity.jl:14-28)."""
    model = F.RANS(F.Laminar, mesh=mesh, viscosity=F.ConstantScalar(nu))
    F.assign(model, "U", F.Dirichlet("inlet", [0.0, 0.0, 0.0]), F.Dirichlet("outlet", [0.0, 0.0, 0.0]),
             F.Dirichlet("bottom", [0.0, 0.0, 0.0]), F.Dirichlet("top", [lid, 0.0, 0.0]))
    F.assign(model, "p", F.Neumann("inlet", 0.0), F.Neumann("outlet", 0.0), F.Neumann("bottom", 0.0),
             F.Neumann("top", 0.0))
    F.initialise(model.U, [0.0, 0.0, 0.0])
    F.initialise(model.p, 0.0)
    return model


def config(iterations, U_solver=F.BicgstabSolver, p_solver=F.CgSolver, U_pre=F.DILU, p_pre=F.DILU, relax_U=0.7,
           relax_p=0.3, precision=0, schemes=None, itmax=100, time=F.Steady, dt=1.0, turbulent=False,
           k_solver=F.GmresSolver, k_pre=F.DILU):
    s = schemes or dict(time=time)
    schemes = dict(U=F.set_schemes(**s), p=F.set_schemes(**s))
    solvers = dict(U=F.set_solver(None, solver=U_solver, preconditioner=U_pre(), convergence=1e-7, relax=relax_U,
                                  itmax=itmax, precision=precision),
                   p=F.set_solver(None, solver=p_solver, preconditioner=p_pre(), convergence=1e-7, relax=relax_p,
                                  itmax=itmax, precision=precision))
    if turbulent:
        schemes.update(k=F.set_schemes(**s), omega=F.set_schemes(**s))
        solvers.update(k=F.set_solver(None, solver=k_solver, preconditioner=k_pre(), convergence=1e-7, relax=0.7,
                                      itmax=itmax),
                       omega=F.set_solver(None, solver=k_solver, preconditioner=k_pre(), convergence=1e-7, relax=0.7,
                                          itmax=itmax))
    runtime = F.set_runtime(iterations=iterations, write_interval=-1, time_step=dt)
    return F.Configuration(schemes=schemes, solvers=solvers, runtime=runtime)


def bfs_model(mesh, U=0.5, nu=1e-5):
    """Laminar backward-facing step (inlet 0.5 m/s above the step, walls, zero-gradient outlet)."""
    model = F.RANS(F.Laminar, mesh=mesh, viscosity=F.ConstantScalar(nu))
    z = [0.0, 0.0, 0.0]
    F.assign(model, "U", F.Dirichlet("inlet", [U, 0.0, 0.0]), F.Dirichlet("step", z), F.Neumann("outlet", 0.0),
             F.Dirichlet("bottom", z), F.Dirichlet("top", z))
    F.assign(model, "p", F.Neumann("inlet", 0.0), F.Neumann("step", 0.0), F.Dirichlet("outlet", 0.0),
             F.Neumann("bottom", 0.0), F.Neumann("top", 0.0))
    F.initialise(model.U, [U, 0.0, 0.0])  # as CASE_UNV_BFS.jl:128
    F.initialise(model.p, 0.0)
    return model


def bfs_case_model(mesh, U=0.5, nu=1e-3):
    """CASE_UNV_BFS.jl on the reference's backwardFacingStep meshes (patches
    inlet / outlet / wall / top): uniform inlet, zero-gradient outlet with
    p = 0, no-slip wall and top."""
    model = F.RANS(F.Laminar, mesh=mesh, viscosity=F.ConstantScalar(nu))
    v, z = [U, 0.0, 0.0], [0.0, 0.0, 0.0]
    F.assign(model, "U", F.Dirichlet("inlet", v), F.Neumann("outlet", 0.0), F.Dirichlet("wall", z),
             F.Dirichlet("top", z))
    F.assign(model, "p", F.Neumann("inlet", 0.0), F.Dirichlet("outlet", 0.0), F.Neumann("wall", 0.0),
             F.Neumann("top", 0.0))
    F.initialise(model.U, v)
    F.initialise(model.p, 0.0)
    return model


def channel_model(mesh, nu=0.01, U=1.0):
    """Plane channel: uniform inlet, zero-gradient outlet with p = 0, no-slip walls
    (patches inlet/outlet/bottom/top as the generated channel)."""
    model = F.RANS(F.Laminar, mesh=mesh, viscosity=F.ConstantScalar(nu))
    z = [0.0, 0.0, 0.0]
    F.assign(model, "U", F.Dirichlet("inlet", [U, 0.0, 0.0]), F.Neumann("outlet", 0.0), F.Dirichlet("bottom", z),
             F.Dirichlet("top", z))
    F.assign(model, "p", F.Neumann("inlet", 0.0), F.Dirichlet("outlet", 0.0), F.Neumann("bottom", 0.0),
             F.Neumann("top", 0.0))
    F.initialise(model.U, [U, 0.0, 0.0])
    F.initialise(model.p, 0.0)
    return model


def batch_cases(ncases=64, seed=5, lo=32, hi=512):
    """The batched workload: side lengths exp(U(ln lo, ln hi)) rounded, Re =
    exp(U(ln 100, ln 1000)) on a unit square with unit lid / inlet velocity
    (nu = 1/Re); even cases are cavities (pressure reference 0 at their first
    cell), odd cases channels.  Returns [(model, pref)]."""
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), ncases))).astype(int)
    res = np.exp(rng.uniform(np.log(100.0), np.log(1000.0), ncases))
    out = []
    for k in range(ncases):
        n, nu = int(sides[k]), 1.0 / float(res[k])
        if k % 2 == 0:
            out.append((cavity_model(F.cavity(n), nu=nu), 0.0))
        else:
            out.append((channel_model(F.channel(n, n), nu=nu), None))
    return out


def flatplate_model(mesh, U=10.0, nu=1.5e-5, intensity=0.01, visc_ratio=10.0):
    """k-omega flat plate (CASE_UNV_flatplate_highRe.jl:16-51): k = 1.5 (I U)^2,
    omega = k / (visc_ratio nu)."""
    k_in = 1.5 * (intensity * U) ** 2
    w_in = k_in / (visc_ratio * nu)
    model = F.RANS(F.KOmega, mesh=mesh, viscosity=F.ConstantScalar(nu))
    z = [0.0, 0.0, 0.0]
    F.assign(model, "U", F.Dirichlet("inlet", [U, 0.0, 0.0]), F.Neumann("outlet", 0.0), F.Dirichlet("wall", z),
             F.Neumann("top", 0.0))
    F.assign(model, "p", F.Neumann("inlet", 0.0), F.Dirichlet("outlet", 0.0), F.Neumann("wall", 0.0),
             F.Neumann("top", 0.0))
    F.assign(model, "k", F.Dirichlet("inlet", k_in), F.Neumann("outlet", 0.0), F.KWallFunction("wall"),
             F.Neumann("top", 0.0))
    F.assign(model, "omega", F.Dirichlet("inlet", w_in), F.Neumann("outlet", 0.0), F.OmegaWallFunction("wall"),
             F.Neumann("top", 0.0))
    F.assign(model, "nut", F.Dirichlet("inlet", k_in / w_in), F.Neumann("outlet", 0.0), F.NutWallFunction("wall"),
             F.Neumann("top", 0.0))
    F.initialise(model.U, [U, 0.0, 0.0])
    F.initialise(model.p, 0.0)
    F.initialise(model.turbulence["k"], k_in)
    F.initialise(model.turbulence["omega"], w_in)
    F.initialise(model.turbulence["nut"], k_in / w_in)
    return model


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def same_bits(a, b):
    """Bitwise equality, NaN == NaN."""
    a, b = np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64)
    if a.shape != b.shape:
        return False
    return bool(np.all((a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))))
