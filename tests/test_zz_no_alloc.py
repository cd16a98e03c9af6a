"""fvh_residual_noalloc: residual! with its partial sums and result in the
library's file-scope device scratch array (no allocation on the call path),
against fvh_residual on the same equation, bit for bit, on two calls in a
row, the second on a smaller system.  Also against the oracle's residual."""
import ctypes as C

import numpy as np
import pytest

import fvmhip as F
from fvmhip import _lib as L
from oracle import ref
from tests.cases import same_bits

pytestmark = pytest.mark.gpu
LIB = L.lib


def _dev(a):
    return L.DeviceArray.from_host(np.ascontiguousarray(a, np.float64))


def _system(nx, ny, seed):
    """A convection-diffusion matrix on an nx x ny channel and a field."""
    mesh = F.channel(nx, ny, 1.0, 0.5, grading=1.03)
    dm = mesh.device()
    n, nf = mesh.ncells, mesh.nfaces
    rng = np.random.default_rng(seed)
    ux = np.sin(2.0 * mesh.cell_centre[:, 0]) + 0.01 * rng.standard_normal(n)
    uy = np.cos(3.0 * mesh.cell_centre[:, 1]) + 0.01 * rng.standard_normal(n)
    fx, fy = ref.interpolate_vector(mesh, ux, uy)
    mdotf = ref.flux(mesh, fx, fy, np.zeros(nf))
    gam = np.full(nf, 1e-2)
    keep = [_dev(mdotf), _dev(gam)]
    terms = (L.Term * 2)()
    terms[0].op, terms[0].scheme, terms[0].sign = L.OP_DIVERGENCE, F.Upwind, 1
    terms[0].flux, terms[0].phi = keep[0].ptr, None
    terms[1].op, terms[1].scheme, terms[1].sign = L.OP_LAPLACIAN, F.Linear, -1
    terms[1].flux, terms[1].phi = keep[1].ptr, None
    srcs = (L.Source * 1)()
    src = _dev(np.cos(np.arange(n) * 0.003))
    keep.append(src)
    srcs[0].sign, srcs[0].field = 1, src.ptr
    eq = C.c_void_p()
    L.check(LIB().fvh_eqn_create(dm.h, C.byref(eq)), "fvh_eqn_create")
    L.check(LIB().fvh_discretise(eq, terms, 2, srcs, 1, None, 0.0), "fvh_discretise")
    x = _dev(ux)
    return mesh, dm, eq, x, keep


def _residual(fn, eq, x):
    r = C.c_double(-1.0)
    L.check(fn(eq, x.vp, C.byref(r)), fn.__name__)
    return r.value


def test_residual_noalloc_matches_the_allocating_entry_point():
    # the first call on the larger system, the second on a smaller one:
    # the scratch array keeps the first call's partials, which the second
    # must neither read nor depend on
    sizes = [(700, 300, 5), (90, 40, 6)]
    systems = [_system(*s) for s in sizes]
    try:
        for k, (mesh, dm, eq, x, keep) in enumerate(systems):
            print(f"call {k}: {mesh.ncells} cells", flush=True)
            a = _residual(LIB().fvh_residual_noalloc, eq, x)
            b = _residual(LIB().fvh_residual, eq, x)
            assert np.isfinite(a) and a > 0.0, a
            assert same_bits(np.array([a]), np.array([b])), (k, a, b)
            # the same call again gives the same bits (nothing left over)
            a2 = _residual(LIB().fvh_residual_noalloc, eq, x)
            assert same_bits(np.array([a2]), np.array([a])), (k, a2, a)
    finally:
        for _, _, eq, _, _ in systems:
            LIB().fvh_eqn_destroy(eq)


def test_residual_noalloc_rejects_null_arguments():
    mesh, dm, eq, x, keep = _system(20, 10, 1)
    try:
        r = C.c_double()
        assert LIB().fvh_residual_noalloc(None, x.vp, C.byref(r)) == L.FVH_ERR_ARG
        assert LIB().fvh_residual_noalloc(eq, None, C.byref(r)) == L.FVH_ERR_ARG
    finally:
        LIB().fvh_eqn_destroy(eq)
