"""The LDL preconditioner on the device (preconditioners_0_types.jl:36-42):
the factor of the set-up matrix, kept for every later solve.  The device
factor is ordered by nested dissection and the oracle's by the matrix's own
order (the reference's AMD ordering is absent), so the two agree to rounding,
not bit for bit: 1e-10 relative on an apply (the pressure matrices' condition
numbers reach 1e4-1e5, so a different elimination order moves the last
digits), a residual of 1e-12 relative either way, and 1e-10 relative L2 on
the fields of a whole run (the FP64 bar of this path)."""
import ctypes as C

import numpy as np
import pytest

import fvmhip as F
from fvmhip import _lib as L
from oracle import ref
from tests.cases import bfs_case_model, cavity_model, config, rel_l2
from tests.conftest import REFERENCE_MESHES

pytestmark = pytest.mark.gpu
LIB = L.lib


def _pressure_matrix(mesh, gamma, dm):
    """Laplacian(gamma, p) with the cavity's Neumann walls and a Dirichlet
    outlet, as the p equation's set-up assembles it."""
    eq = C.c_void_p()
    L.check(LIB().fvh_eqn_create(dm.h, C.byref(eq)), "fvh_eqn_create")
    g = L.DeviceArray.from_host(np.ascontiguousarray(gamma, np.float64))
    terms = (L.Term * 1)()
    terms[0].op, terms[0].scheme, terms[0].sign = L.OP_LAPLACIAN, F.Linear, 1
    terms[0].flux, terms[0].phi = g.ptr, None
    L.check(LIB().fvh_discretise(eq, terms, 1, None, 0, None, 0.0), "fvh_discretise")
    bcs = [F.Neumann(nm, 0.0) for nm in mesh.boundary_names]
    bcs[1] = F.Dirichlet(mesh.boundary_names[1], [0.0, 0.0, 0.0])
    barr = F.api._bcs(mesh, bcs, True)
    L.check(LIB().fvh_apply_bcs(eq, terms, 1, barr, len(bcs), 0), "fvh_apply_bcs")
    nz = C.c_void_p()
    L.check(LIB().fvh_eqn_arrays(eq, C.byref(nz), None))
    rp, ci = dm.pattern()
    vals = np.zeros(ci.size)
    L.check(LIB().fvh_memcpy(vals.ctypes.data_as(C.c_void_p), nz, vals.nbytes))
    return eq, g, vals


@pytest.mark.parametrize("solves", ["ranges", "rows"])
@pytest.mark.parametrize("name", REFERENCE_MESHES)
def test_ldl_apply_against_the_oracle(ref_mesh, name, solves, monkeypatch):
    """Both triangular-solve schedules: the range walk over the
    nested-dissection leaves and separators (the default) and one wave per
    row (FVH_LDL_RANGES=0)."""
    monkeypatch.setenv("FVH_LDL_RANGES", "1" if solves == "ranges" else "0")
    mesh = ref_mesh(name)
    dm = mesh.device()
    n = mesh.ncells
    gamma = 1.0 + 0.5 * np.sin(np.arange(mesh.nfaces) * 0.1)
    eq, keep, vals = _pressure_matrix(mesh, gamma, dm)
    try:
        L.check(LIB().fvh_precond_set(eq, L.PRECOND_LDL, None), "fvh_precond_set")
        L.check(LIB().fvh_precond_update(eq), "fvh_precond_update")
        rng = np.random.default_rng(5)
        for t in range(2):
            x = rng.standard_normal(n)
            dx, dy = L.DeviceArray.from_host(x), L.DeviceArray(n)
            L.check(LIB().fvh_precond_apply(eq, dy.vp, dx.vp), "fvh_precond_apply")
            y = dy.host()
            want = ref.precond_apply_csr(mesh, vals, L.PRECOND_LDL, x)
            err = np.linalg.norm(y - want) / np.linalg.norm(want)
            res = np.linalg.norm(ref.spmv_csr(mesh, vals, y) - x) / np.linalg.norm(x)
            print(f"{name} ({solves}): apply {t}: rel diff to the oracle {err:.2e}, residual {res:.2e}", flush=True)
            assert err <= 1e-10 and res <= 1e-12
        # the factor is kept: a changed matrix and another update leave M alone
        L.check(LIB().fvh_precond_update(eq), "fvh_precond_update")
        dx2 = L.DeviceArray.from_host(x)
        L.check(LIB().fvh_precond_apply(eq, dy.vp, dx2.vp), "fvh_precond_apply")
        assert np.array_equal(dy.host(), y)
        # FP32 vectors through the FP64 factor
        xf = x.astype(np.float32)
        fx, fy = L.DeviceArray.from_host(xf), L.DeviceArray(n, np.float32)
        L.check(LIB().fvh_precond_apply_f32(eq, fy.vp, fx.vp), "fvh_precond_apply_f32")
        assert np.linalg.norm(fy.host() - want) / np.linalg.norm(want) <= 1e-6
    finally:
        LIB().fvh_eqn_destroy(eq)


def test_ldl_needs_a_regular_matrix(ref_mesh):
    """All-Neumann pressure (the cavity without its reference cell): the
    set-up matrix is singular and the factorization reports it."""
    mesh = ref_mesh("quad40")
    dm = mesh.device()
    eq = C.c_void_p()
    L.check(LIB().fvh_eqn_create(dm.h, C.byref(eq)))
    try:
        L.check(LIB().fvh_precond_set(eq, L.PRECOND_LDL, None))
        assert LIB().fvh_precond_update(eq) == L.FVH_ERR_ARG  # all-zero matrix: zero pivot
    finally:
        LIB().fvh_eqn_destroy(eq)


@pytest.mark.parametrize("p_solver", [F.GmresSolver, F.BicgstabSolver])
def test_reference_bfs_case_with_ldl(ref_mesh, p_solver):
    """CASE_UNV_BFS.jl as the reference runs it: backwardFacingStep_10mm,
    GMRES + DILU for U, p with LDL, relax 0.8 / 0.2, to the 1e-7 break.  The
    device run and the oracle agree to 1e-10 (relative L2, every field) and
    in their iteration counts."""
    mesh = ref_mesh("backwardFacingStep_10mm")
    model = bfs_case_model(mesh)
    cfg = config(1000, U_solver=F.GmresSolver, p_solver=p_solver, p_pre=F.LDL, relax_U=0.8, relax_p=0.2)
    g = F.Simulation(model, cfg)
    r = ref.RefSimulation(model, F.CaseDesc(model, cfg))
    ng, nr = g.run(1000), r.run(1000)
    errs = {f: rel_l2(g.get(f), r.get(f)) for f in ("Ux", "Uy", "p")}
    print(f"BFS 10 mm, p solver {p_solver} + LDL: {ng} iterations (oracle {nr}); rel L2 {errs}")
    assert ng == nr and ng < 1000
    for f, e in errs.items():
        assert e <= 1e-10, f"{f}: {e:.3e}"


def test_ldl_rejects_the_momentum_equation(ref_mesh):
    """The reference factors the momentum equation's set-up matrix, which has
    no flux and no viscosity yet (all zero): refused at set-up."""
    mesh = ref_mesh("quad40")
    model = cavity_model(mesh)
    cfg = config(10, U_pre=F.LDL)
    with pytest.raises(F.FVHError):
        F.Simulation(model, cfg, 0.0)
