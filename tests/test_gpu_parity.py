"""GPU parity: libfvmhip (HIP, gfx950) against the oracle on the same inputs.

The FP64 path is required to be bit-identical to the oracle (same operation
order, same canonical reductions, no FMA); every comparison below is exact
unless it says otherwise.
"""
import ctypes as C

import numpy as np
import pytest

import fvmhip as F
from fvmhip import _lib as L
from oracle import ref
from tests.cases import bfs_model, cavity_model, channel_model, config, flatplate_model, rel_l2, same_bits

pytestmark = pytest.mark.gpu


def _sims(model, cfg, pref=None, algorithm=L.ALG_SIMPLE, conv_break=1e-7):
    case = F.CaseDesc(model, cfg, pref, algorithm, conv_break)
    return F.Simulation(model, cfg, pref, algorithm, conv_break), ref.RefSimulation(model, case)


PHASE_ARRAYS = ["ux_A", "ux_b", "uy_A", "uy_b", "ux_dinv", "Ux", "Uy", "rD", "rDf", "Hv_x", "Hv_y", "divHv", "p_A",
                "p_b", "p_dinv", "p", "gradp_x", "gradp_y", "pf", "Uf_x", "Uf_y", "mdotf"]


def test_cavity_first_iterations_bitwise(quad40):
    """Every intermediate array of the first two SIMPLE iterations of C1."""
    model = cavity_model(quad40)
    g, r = _sims(model, config(2), pref=0.0)
    for it in range(2):
        assert g.run(1) == 1 and r.run(1) == 1
        for name in PHASE_ARRAYS:
            a, b = g.get(name), r.get(name)
            assert same_bits(a, b), f"iteration {it + 1}: {name} differs (rel {rel_l2(a, b):.3e})"
        for fld in ("Ux", "Uy", "p"):
            assert g.stats(fld)["niter"] == r.stats(fld)["niter"], fld
    Rg, Rr = g.residuals(), r.residuals()
    for a, b in zip(Rg, Rr):
        assert same_bits(a, b)
    assert np.isnan(Rg[1][0])  # R_uy[1] = 0/0 (b == 0 at the first iteration)


@pytest.mark.parametrize("usolver,psolver,upre,ppre", [
    (F.BicgstabSolver, F.CgSolver, F.DILU, F.DILU),
    (F.GmresSolver, F.GmresSolver, F.DILU, F.DILU),
    (F.BicgstabSolver, F.BicgstabSolver, F.ILU0, F.ILU0),
    (F.GmresSolver, F.CgSolver, F.Jacobi, F.Jacobi),
    (F.BicgstabSolver, F.GmresSolver, F.NormDiagonal, F.DILU),
])
def test_cavity_solver_matrix_bitwise(quad40, usolver, psolver, upre, ppre):
    model = cavity_model(quad40)
    g, r = _sims(model, config(12, U_solver=usolver, p_solver=psolver, U_pre=upre, p_pre=ppre), pref=0.0)
    assert g.run(12) == r.run(12)
    for a, b in zip(g.residuals(), r.residuals()):
        assert same_bits(a, b)
    for name in ("Ux", "Uy", "p", "mdotf"):
        assert same_bits(g.get(name), r.get(name)), name


def test_cavity_c1_full_run(quad40):
    """C1 as specified: 1000 iterations (break at 1e-7); histories and fields."""
    model = cavity_model(quad40)
    g, r = _sims(model, config(1000), pref=0.0)
    ng, nr = g.run(1000), r.run(1000)
    assert ng == nr
    Rg, Rr = g.residuals(), r.residuals()
    for a, b in zip(Rg, Rr):
        assert same_bits(a, b)
    for name in ("Ux", "Uy", "p"):
        a, b = g.get(name), r.get(name)
        assert rel_l2(a, b) <= 1e-10
        assert same_bits(a, b), name


@pytest.mark.parametrize("fused", ["1", "0"])
@pytest.mark.parametrize("mesh_name", ["quad40", "trig40"])
def test_cg_beta_fused_with_spmv_bitwise(ref_mesh, mesh_name, fused, monkeypatch):
    """The pressure CG with each beta update fused into the next iteration's
    SpMV (k_cg_beta_spmv, the default) and with the two apart
    (FVH_CG_FUSED=0), on a banded mesh (banded SpMV, DILU sweeps that read r
    in their own layout) and a triangle mesh (CSR SpMV): both bitwise."""
    monkeypatch.setenv("FVH_CG_FUSED", fused)
    model = cavity_model(ref_mesh(mesh_name))
    g, r = _sims(model, config(25), pref=0.0)
    assert g.run(25) == r.run(25)
    for a, b in zip(g.residuals(), r.residuals()):
        assert same_bits(a, b)
    for name in ("Ux", "Uy", "p", "mdotf"):
        assert same_bits(g.get(name), r.get(name)), name
    assert g.stats("p")["niter"] == r.stats("p")["niter"]


@pytest.mark.parametrize("pair", ["1", "0"])
def test_momentum_pair_bitwise(quad40, pair, monkeypatch):
    """The two momentum solves at once (the default on banded meshes: Uy on a
    second stream and thread, applying a copy of Ux's DILU records) and one
    after the other (FVH_U_PAIR=0: Uy borrows Ux's preconditioner): both
    bitwise, with BiCGSTAB and GMRES for U."""
    monkeypatch.setenv("FVH_U_PAIR", pair)
    model = cavity_model(quad40)
    for usolver in (F.BicgstabSolver, F.GmresSolver):
        g, r = _sims(model, config(15, U_solver=usolver), pref=0.0)
        assert g.run(15) == r.run(15)
        for a, b in zip(g.residuals(), r.residuals()):
            assert same_bits(a, b)
        for name in ("Ux", "Uy", "p", "mdotf"):
            assert same_bits(g.get(name), r.get(name)), name
        for fld in ("Ux", "Uy"):
            assert g.stats(fld)["niter"] == r.stats(fld)["niter"], fld


def test_upwind_midpoint_euler_piso(quad40):
    """Upwind divergence, Midpoint gradients, Euler time term, PISO loop."""
    model = cavity_model(quad40)
    cfg = config(4, schemes=dict(time=F.Euler, divergence=F.Upwind, gradient=F.Midpoint), dt=0.05)
    g, r = _sims(model, cfg, pref=0.0, algorithm=L.ALG_PISO)
    assert g.run(4) == r.run(4) == 4
    for name in ("Ux", "Uy", "p", "gradp_x", "divHv", "ux_A", "ux_b"):
        assert same_bits(g.get(name), r.get(name)), name


def test_bfs_small_bitwise():
    """A small BFS (graded mesh, Dirichlet outlet pressure, Neumann outlet U)."""
    mesh = F.backward_facing_step(nx=120, ny=30)
    model = bfs_model(mesh, nu=1e-3)
    g, r = _sims(model, config(20, schemes=dict(divergence=F.Upwind)), conv_break=1e-6)
    assert g.run(20) == r.run(20)
    for a, b in zip(g.residuals(), r.residuals()):
        assert same_bits(a, b)
    for name in ("Ux", "Uy", "p"):
        assert same_bits(g.get(name), r.get(name)), name


@pytest.mark.parametrize("nx,ny", [(3, 4), (5, 4), (6, 4), (9, 64), (9, 65)])
def test_small_grids_bitwise(nx, ny, monkeypatch):
    """Grids at the pipelined sweeps' edges: one group of 8 anti-diagonals
    (3 x 4, 5 x 4), the first with two (6 x 4), a full wave of rows and the
    first with a second wave (9 x 64, 9 x 65)."""
    for env in ("FVH_SWEEP", "FVH_BAND_ROWS", "FVH_SKEW_WAVES", "FVH_LEVEL_RECORDS"):
        monkeypatch.delenv(env, raising=False)
    mesh = F.channel(nx, ny)
    assert mesh.device().sweep_info() == (nx + ny - 1, nx)
    model = channel_model(mesh)
    g, r = _sims(model, config(5), conv_break=1e-12)
    assert g.run(5) == r.run(5)
    for a, b in zip(g.residuals(), r.residuals()):
        assert same_bits(a, b)
    for name in ("ux_dinv", "p_dinv", "Ux", "Uy", "p"):
        assert same_bits(g.get(name), r.get(name)), name


def test_komega_flatplate_bitwise():
    mesh = F.flat_plate(nx=60, ny=30, first=1e-3, ratio=1.1)
    model = flatplate_model(mesh)
    cfg = config(5, schemes=dict(divergence=F.Upwind), turbulent=True, p_solver=F.BicgstabSolver)
    g, r = _sims(model, cfg)
    assert g.run(5) == r.run(5)
    for name in ("Ux", "Uy", "p", "k", "omega", "nut", "nutf", "Pk"):
        a, b = g.get(name), r.get(name)
        assert same_bits(a, b), f"{name} rel {rel_l2(a, b):.3e}"


# ---------------------------------------------------------------------------
# Operator level, at full size
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def bfs_full():
    return F.backward_facing_step()  # 2000 x 500, 1M cells


def _eqn(mesh):
    dm = mesh.device()
    h = C.c_void_p()
    L.check(L.lib().fvh_eqn_create(dm.h, C.byref(h)))
    nz, b = C.c_void_p(), C.c_void_p()
    L.check(L.lib().fvh_eqn_arrays(h, C.byref(nz), C.byref(b)))
    return dm, h, nz.value, b.value


def test_spmv_and_dilu_full_size_bitwise(bfs_full):
    """C2-size (1M cells) SpMV and DILU (update + apply) on random values."""
    mesh = bfs_full
    dm, h, nzp, bp = _eqn(mesh)
    try:
        rp, ci = dm.pattern()
        rng = np.random.default_rng(7)
        nnz = ci.size
        rows = np.repeat(np.arange(mesh.ncells), np.diff(rp))
        vals = np.where(rows == ci, 4.5, 0.0) - rng.uniform(0.1, 1.0, nnz) * (rows != ci)
        L.check(L.lib().fvh_memcpy(C.c_void_p(nzp), vals.ctypes.data_as(C.c_void_p), vals.nbytes))
        x = rng.standard_normal(mesh.ncells)
        xd, yd = L.DeviceArray.from_host(x), L.DeviceArray(mesh.ncells)
        L.check(L.lib().fvh_spmv(h, xd.vp, yd.vp))
        assert same_bits(yd.host(), ref.spmv_csr(mesh, vals, x))
        L.check(L.lib().fvh_precond_set(h, L.PRECOND_DILU, None))
        L.check(L.lib().fvh_precond_update(h))
        L.check(L.lib().fvh_precond_apply(h, yd.vp, xd.vp))
        assert same_bits(yd.host(), ref.precond_apply_csr(mesh, vals, L.PRECOND_DILU, x))
        for kind in (L.PRECOND_ILU0, L.PRECOND_JACOBI):
            L.check(L.lib().fvh_precond_set(h, kind, None))
            L.check(L.lib().fvh_precond_update(h))
            L.check(L.lib().fvh_precond_apply(h, yd.vp, xd.vp))
            assert same_bits(yd.host(), ref.precond_apply_csr(mesh, vals, kind, x)), kind
    finally:
        L.lib().fvh_eqn_destroy(h)


def test_banded_spmv_bitwise(bfs_full):
    """The banded SpMV the solves use on a structured grid (A read as five
    values per row) against the oracle's CSR product, on C2-size random values
    with explicit zeros off the band's edges (the timing entry point returns
    the product of its last launch)."""
    mesh = bfs_full
    dm, h, nzp, bp = _eqn(mesh)
    try:
        assert dm.sweep_info()[1] > 0
        rp, ci = dm.pattern()
        rng = np.random.default_rng(11)
        nnz = ci.size
        rows = np.repeat(np.arange(mesh.ncells), np.diff(rp))
        vals = np.where(rows == ci, 4.5, 0.0) - rng.uniform(0.1, 1.0, nnz) * (rows != ci)
        vals[rng.random(nnz) < 0.01] = 0.0  # stored zeros stay entries of the row
        L.check(L.lib().fvh_memcpy(C.c_void_p(nzp), vals.ctypes.data_as(C.c_void_p), vals.nbytes))
        x = rng.standard_normal(mesh.ncells)
        x[rng.random(mesh.ncells) < 0.01] = -0.0
        xd, yd = L.DeviceArray.from_host(x), L.DeviceArray(mesh.ncells)
        ms = C.c_double()
        L.check(L.lib().fvh_time_op(h, 6, 1, xd.vp, yd.vp, C.byref(ms)))
        assert same_bits(yd.host(), ref.spmv_csr(mesh, vals, x))
    finally:
        L.lib().fvh_eqn_destroy(h)


@pytest.mark.parametrize("sweep,rows", [("levels", None), ("levels", "rowwise"), ("band", None), ("band", "w1"),
                                        ("band", "w2"), ("band", "w3"), ("band", "1"), ("band", "2"), ("band", "3"),
                                        ("band", "4")])
def test_sweep_paths_bitwise(sweep, rows, monkeypatch):
    """Every DILU schedule against the oracle: level sets (per-row records
    prefetched a level ahead, or the row-by-row kernels); the pipelined band
    sweeps (default: skewed slots, waves handing over through LDS within a
    block and through a courier wave per block between blocks; tuned, or one,
    two or three waves per block, so block boundaries fall between different
    waves); and the single-work-group band variants (one or two grid rows per
    lane with a barrier per step, LDS-pipelined waves, deep prefetch).
    60 x 300: 5 waves of grid rows at one row per lane."""
    if sweep == "levels":
        monkeypatch.setenv("FVH_SWEEP", "levels")
    else:
        monkeypatch.delenv("FVH_SWEEP", raising=False)
    monkeypatch.delenv("FVH_BAND_ROWS", raising=False)
    monkeypatch.delenv("FVH_SKEW_WAVES", raising=False)
    monkeypatch.delenv("FVH_LEVEL_RECORDS", raising=False)
    if rows == "rowwise":
        monkeypatch.setenv("FVH_LEVEL_RECORDS", "0")
    elif rows and rows.startswith("w"):  # waves per block
        monkeypatch.setenv("FVH_SKEW_WAVES", rows[1])
    elif rows:
        monkeypatch.setenv("FVH_BAND_ROWS", rows)
    mesh = F.backward_facing_step(nx=60, ny=300)
    nlev, band = mesh.device().sweep_info()
    assert nlev == 60 + 300 - 1
    assert (band == 60) == (sweep == "band")
    if sweep == "band":
        assert mesh.device().tuning() == (int(rows) if rows and rows[0].isdigit() else 5)
    model = bfs_model(mesh, nu=1e-3)
    g, r = _sims(model, config(4, schemes=dict(divergence=F.Upwind)), conv_break=1e-6)
    assert g.run(4) == r.run(4)
    for name in ("ux_dinv", "p_dinv", "Ux", "Uy", "p"):
        assert same_bits(g.get(name), r.get(name)), name
    for a, b in zip(g.residuals(), r.residuals()):
        assert same_bits(a, b)


def test_quad40_is_banded(quad40):
    nlev, band = quad40.device().sweep_info()
    assert band == 40 and nlev == 79
