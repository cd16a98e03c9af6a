"""Steady scalar transport on an unstructured Delaunay mesh (the C3 set-up at
small size): assembly of div(mdotf phi) - laplacian(gamma phi) with Dirichlet
BCs, and BiCGSTAB + DILU on the level-scheduled sweeps (irregular adjacency,
three faces per cell)."""
import numpy as np
import pytest

import fvmhip as F
from fvmhip import _lib as L
from oracle import ref
from tests.cases import rel_l2, same_bits

pytestmark = pytest.mark.gpu


def _eq(mesh, scheme=F.Linear, gamma=1e-3):
    bcs = [F.Dirichlet("left", 1.0), F.Dirichlet("right", 0.0), F.Dirichlet("bottom", 0.0), F.Dirichlet("top", 0.0)]
    return F.ScalarEquation(mesh, F.taylor_green_flux(mesh), np.full(mesh.nfaces, gamma), bcs, divergence=scheme)


def test_delaunay_transport_fp64_matches_oracle_and_mixed_converges():
    # 3000 points: cell size ~0.013, so gamma = 2e-2 keeps the cell Peclet
    # number below 1 (central convection stays well conditioned, as C3's 1M
    # points do at gamma = 1e-3)
    mesh = F.delaunay_unit_square(3000, seed=2)
    nlev, band = mesh.device().sweep_info()
    assert band == 0 and nlev > 10  # level-scheduled
    eq = _eq(mesh, gamma=2e-2)
    eq.discretise()
    vals, b = eq.matrix()
    cfg = L.SolverCfg(solver=L.BICGSTAB, precond=L.PRECOND_DILU, itmax=1000, precision=L.FP64, convergence=0.0,
                      relax=1.0, atol=0.0, rtol=1e-10)
    st = eq.solve(rtol=1e-10, itmax=1000)
    x = eq.values()
    xr, sr, _ = ref.solve_csr(mesh, vals, b, cfg, np.zeros(mesh.ncells))
    assert st["niter"] == sr["niter"] and st["status"] == sr["status"] == "solved", (st, sr)
    assert same_bits(x, xr)
    # physically bounded by the boundary values (Linear convection: small overshoot allowed)
    assert x.min() > -0.05 and x.max() < 1.05
    # mixed precision: FP32 inner solves in FP64 defect correction reach the same solution
    eq2 = _eq(mesh, gamma=2e-2)
    eq2.discretise()
    st2 = eq2.solve(rtol=1e-10, itmax=1000, precision=L.MIXED)
    assert st2["refinements"] >= 2
    assert rel_l2(eq2.values(), x) < 1e-8


def test_upwind_transport_obeys_the_maximum_principle():
    mesh = F.delaunay_unit_square(2000, seed=4)
    eq = _eq(mesh, F.Upwind)
    eq.discretise()
    eq.solve(rtol=1e-12, itmax=2000)
    x = eq.values()
    assert x.min() >= -1e-10 and x.max() <= 1.0 + 1e-10


# ---------------------------------------------------------------------------
# C3 at its real size: the Delaunay triangulation of 1M seeded points (about
# 2M cells and 4,700 levels per sweep)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def delaunay_full():
    return F.delaunay_unit_square(1_000_000, seed=0)


@pytest.mark.parametrize("walk", ["tiles", "tiles-256", "syncfree", "levels", "levels-ring256"])
def test_delaunay_full_size_dilu_bitwise(delaunay_full, walk, monkeypatch):
    """SpMV, the DILU update and the DILU apply (FP64) on random values,
    bit for bit against the oracle, on the tile walk (the default), on the
    tile walk with 256-row tiles (about 8,000 tiles, most rows next to
    another tile's), on the sync-free walk, on the one-work-group level walk,
    and on the level walk with its LDS ring cut to 256 results (so most
    neighbours come through its global-memory and after-barrier paths)."""
    import ctypes as C

    mesh = delaunay_full
    print(f"walk {walk}", flush=True)
    monkeypatch.setenv("FVH_LEVEL_SYNCFREE", "0" if walk.startswith("levels") else "1")
    monkeypatch.setenv("FVH_LEVEL_TILES", "1" if walk.startswith("tiles") else "0")
    if walk == "tiles-256":
        monkeypatch.setenv("FVH_TILE_ROWS", "256")
    else:
        monkeypatch.delenv("FVH_TILE_ROWS", raising=False)
    if walk == "levels-ring256":
        monkeypatch.setenv("FVH_LEVEL_RING", "256")
    else:
        monkeypatch.delenv("FVH_LEVEL_RING", raising=False)
    mesh._dev = None  # the schedule (and its ring) is built with the device mesh
    dm = mesh.device()
    nlev, band = dm.sweep_info()
    assert band == 0 and nlev > 1000
    h = C.c_void_p()
    L.check(L.lib().fvh_eqn_create(dm.h, C.byref(h)))
    try:
        nz = C.c_void_p()
        L.check(L.lib().fvh_eqn_arrays(h, C.byref(nz), None))
        rp, ci = dm.pattern()
        rng = np.random.default_rng(17)
        rows = np.repeat(np.arange(mesh.ncells), np.diff(rp))
        vals = np.where(rows == ci, 4.5, 0.0) - rng.uniform(0.1, 1.0, ci.size) * (rows != ci)
        L.check(L.lib().fvh_memcpy(nz, vals.ctypes.data_as(C.c_void_p), vals.nbytes))
        x = rng.standard_normal(mesh.ncells)
        xd, yd = L.DeviceArray.from_host(x), L.DeviceArray(mesh.ncells)
        L.check(L.lib().fvh_spmv(h, xd.vp, yd.vp))
        assert same_bits(yd.host(), ref.spmv_csr(mesh, vals, x))
        L.check(L.lib().fvh_precond_set(h, L.PRECOND_DILU, None))
        L.check(L.lib().fvh_precond_update(h), "fvh_precond_update")
        dinv = np.zeros(mesh.ncells)
        L.check(L.lib().fvh_precond_dinv(h, L.dptr(dinv)))
        L.check(L.lib().fvh_precond_apply(h, yd.vp, xd.vp), "fvh_precond_apply")
        want = ref.precond_apply_csr(mesh, vals, L.PRECOND_DILU, x)
        got = yd.host()
        print(f"{walk}: {mesh.ncells} cells, {nlev} levels; apply rel diff "
              f"{np.linalg.norm(got - want) / np.linalg.norm(want):.2e}", flush=True)
        assert np.all(np.isfinite(dinv)) and np.all(dinv > 0)
        assert same_bits(got, want)
    finally:
        L.lib().fvh_eqn_destroy(h)
        mesh._dev = None


def test_delaunay_full_size_fp32_walks_agree(delaunay_full, monkeypatch):
    """The FP32 DILU apply (mixed precision's inner solves) at C3's real size:
    the tile walk (1,024- and 256-row tiles), the sync-free walk and the
    one-work-group level walk evaluate the same FP32 expression per row, so
    their results agree bit for bit; and they are the FP64 apply to FP32
    accuracy."""
    import ctypes as C

    mesh = delaunay_full
    rng = np.random.default_rng(23)
    x = rng.standard_normal(mesh.ncells)
    got = {}
    for walk in ("tiles", "tiles-256", "syncfree", "levels"):
        print(f"walk {walk}", flush=True)
        monkeypatch.setenv("FVH_LEVEL_SYNCFREE", "0" if walk == "levels" else "1")
        monkeypatch.setenv("FVH_LEVEL_TILES", "1" if walk.startswith("tiles") else "0")
        if walk == "tiles-256":
            monkeypatch.setenv("FVH_TILE_ROWS", "256")
        else:
            monkeypatch.delenv("FVH_TILE_ROWS", raising=False)
        mesh._dev = None
        dm = mesh.device()
        h = C.c_void_p()
        L.check(L.lib().fvh_eqn_create(dm.h, C.byref(h)))
        try:
            nz = C.c_void_p()
            L.check(L.lib().fvh_eqn_arrays(h, C.byref(nz), None))
            rp, ci = dm.pattern()
            r2 = np.random.default_rng(29)
            rows = np.repeat(np.arange(mesh.ncells), np.diff(rp))
            vals = np.where(rows == ci, 4.5, 0.0) - r2.uniform(0.1, 1.0, ci.size) * (rows != ci)
            L.check(L.lib().fvh_memcpy(nz, vals.ctypes.data_as(C.c_void_p), vals.nbytes))
            L.check(L.lib().fvh_precond_set(h, L.PRECOND_DILU, None))
            L.check(L.lib().fvh_precond_update(h), "fvh_precond_update")
            xf = L.DeviceArray.from_host(x.astype(np.float32))
            yf = L.DeviceArray(mesh.ncells, np.float32)
            L.check(L.lib().fvh_precond_apply_f32(h, yf.vp, xf.vp), "fvh_precond_apply_f32")
            got[walk] = yf.host()
            if walk == "tiles":
                xd, yd = L.DeviceArray.from_host(x), L.DeviceArray(mesh.ncells)
                L.check(L.lib().fvh_precond_apply(h, yd.vp, xd.vp), "fvh_precond_apply")
                y64 = yd.host()
        finally:
            L.lib().fvh_eqn_destroy(h)
            mesh._dev = None
    for walk, y in got.items():
        assert np.array_equal(y.view(np.int32), got["tiles"].view(np.int32)), walk
    err = np.linalg.norm(got["tiles"] - y64) / np.linalg.norm(y64)
    print(f"FP32 apply against FP64: rel diff {err:.2e}", flush=True)
    assert err < 1e-5
