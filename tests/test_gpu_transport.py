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
