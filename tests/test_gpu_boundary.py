"""The C ABI one entry point at a time, as a Julia `ccall` shim binds it:
every per-function call of include/fvmhip.h on the reference's own sample
meshes (unv_sample_meshes, scale 0.001), against the oracle's function of
the same name on the same inputs.  FP64, bit for bit.

Inputs are seeded smooth fields plus noise, so that no term is trivially
zero; boundary conditions mix Dirichlet and Neumann patches."""
import ctypes as C

import numpy as np
import pytest

import fvmhip as F
from fvmhip import _lib as L
from fvmhip.api import _bcs
from oracle import ref
from tests.cases import bfs_model, cavity_model, config, rel_l2, same_bits
from tests.conftest import REFERENCE_MESHES

pytestmark = pytest.mark.gpu
LIB = L.lib


def _dev(a):
    return L.DeviceArray.from_host(np.ascontiguousarray(a, np.float64))


def _fields(mesh, seed):
    rng = np.random.default_rng(seed)
    x, y = mesh.cell_centre[:, 0], mesh.cell_centre[:, 1]
    n = mesh.ncells
    ux = np.sin(3.0 * x) * np.cos(2.0 * y) + 0.01 * rng.standard_normal(n)
    uy = -np.cos(3.0 * x) * np.sin(2.0 * y) + 0.01 * rng.standard_normal(n)
    phi = x * x - 0.5 * y + 0.3 + 0.01 * rng.standard_normal(n)
    return ux, uy, phi


def _bc_tuple(mesh):
    """Dirichlet on even patches (vector values), Neumann on odd ones."""
    out = []
    for k, name in enumerate(mesh.boundary_names):
        if k % 2 == 0:
            out.append(F.Dirichlet(name, [0.25 + k, -0.5 * k, 0.125]))
        else:
            out.append(F.Neumann(name, 0.0))
    return out


@pytest.mark.parametrize("name", REFERENCE_MESHES)
def test_calculate_entry_points_bitwise(ref_mesh, name):
    mesh = ref_mesh(name)
    dm = mesh.device()
    nf, n = mesh.nfaces, mesh.ncells
    ux, uy, phi = _fields(mesh, 7)
    bcs = _bc_tuple(mesh)
    barr = _bcs(mesh, bcs, True)
    print(f"{name}: {n} cells, {nf} faces", flush=True)

    # interpolate! scalar
    d_phi, d_pf = _dev(phi), _dev(np.zeros(nf))
    L.check(LIB().fvh_interpolate(dm.h, d_phi.vp, d_pf.vp), "fvh_interpolate")
    pf_ref = ref.interpolate(mesh, phi)
    assert same_bits(d_pf.host(), pf_ref), "fvh_interpolate"

    # interpolate! vector: x and y only, z untouched
    fz0 = np.full(nf, 3.5)
    d_ux, d_uy, d_fx, d_fy = _dev(ux), _dev(uy), _dev(fz0), _dev(fz0)
    L.check(LIB().fvh_interpolate_vector(dm.h, d_ux.vp, d_uy.vp, d_fx.vp, d_fy.vp), "fvh_interpolate_vector")
    fx_ref, fy_ref = ref.interpolate_vector(mesh, ux, uy, fz0, fz0)
    assert same_bits(d_fx.host(), fx_ref) and same_bits(d_fy.host(), fy_ref), "fvh_interpolate_vector"

    # correct_boundaries! per component, from the interpolated faces
    for comp in (0, 1):
        d_pf.upload(pf_ref)
        L.check(LIB().fvh_correct_boundaries(dm.h, barr, len(bcs), comp, d_phi.vp, d_pf.vp), "fvh_correct_boundaries")
        assert same_bits(d_pf.host(), ref.correct_boundaries(mesh, barr, len(bcs), comp, phi, pf_ref)), comp

    # green_gauss! on a face field
    face_field = ref.correct_boundaries(mesh, barr, len(bcs), 0, phi, pf_ref)
    d_pf.upload(face_field)
    g = [_dev(np.zeros(n)) for _ in range(3)]
    L.check(LIB().fvh_green_gauss(dm.h, d_pf.vp, *(a.vp for a in g)), "fvh_green_gauss")
    for a, b in zip(g, ref.green_gauss(mesh, face_field)):
        assert same_bits(a.host(), b), "fvh_green_gauss"

    # grad! Orthogonal and Midpoint
    for scheme in (F.Orthogonal, F.Midpoint):
        L.check(LIB().fvh_grad(dm.h, scheme, barr, len(bcs), 0, d_phi.vp, d_pf.vp, *(a.vp for a in g)), "fvh_grad")
        pf_r, g_r = ref.grad(mesh, scheme, barr, len(bcs), 0, phi)
        assert same_bits(d_pf.host(), pf_r), f"fvh_grad({scheme}) phif"
        for a, b in zip(g, g_r):
            assert same_bits(a.host(), b), f"fvh_grad({scheme})"

    # div! and flux! of a face vector field (z nonzero)
    d_fz = _dev(np.linspace(-1.0, 1.0, nf))
    out = _dev(np.zeros(n))
    L.check(LIB().fvh_div(dm.h, d_fx.vp, d_fy.vp, d_fz.vp, out.vp), "fvh_div")
    assert same_bits(out.host(), ref.div(mesh, d_fx.host(), d_fy.host(), d_fz.host())), "fvh_div"
    mdot = _dev(np.zeros(nf))
    L.check(LIB().fvh_flux(dm.h, d_fx.vp, d_fy.vp, d_fz.vp, mdot.vp), "fvh_flux")
    assert same_bits(mdot.host(), ref.flux(mesh, d_fx.host(), d_fy.host(), d_fz.host())), "fvh_flux"


class _Terms:
    """The same term list twice: device pointers for libfvmhip, host
    pointers for the oracle."""

    def __init__(self, spec, n):
        self.n = len(spec)
        self.dev = (L.Term * max(1, self.n))()
        self.host = (L.Term * max(1, self.n))()
        self._keep = []
        for k, (op, scheme, sign, flux, phi) in enumerate(spec):
            for arr, use_dev in ((self.dev, True), (self.host, False)):
                arr[k].op, arr[k].scheme, arr[k].sign = op, scheme, sign
                for attr, a in (("flux", flux), ("phi", phi)):
                    if a is None:
                        setattr(arr[k], attr, None)
                    elif use_dev:
                        d = _dev(a)
                        self._keep.append(d)
                        setattr(arr[k], attr, d.ptr)
                    else:
                        h = np.ascontiguousarray(a, np.float64)
                        self._keep.append(h)
                        setattr(arr[k], attr, h.ctypes.data)


def _eq_arrays(eq, dm, n):
    nz, b = C.c_void_p(), C.c_void_p()
    L.check(LIB().fvh_eqn_arrays(eq, C.byref(nz), C.byref(b)))
    rp, ci = dm.pattern()
    vals, rhs = np.zeros(ci.size), np.zeros(n)
    L.check(LIB().fvh_memcpy(vals.ctypes.data_as(C.c_void_p), nz, vals.nbytes))
    L.check(LIB().fvh_memcpy(rhs.ctypes.data_as(C.c_void_p), b, rhs.nbytes))
    return vals, rhs


@pytest.mark.parametrize("name", REFERENCE_MESHES)
@pytest.mark.parametrize("divergence", [F.Linear, F.Upwind])
def test_equation_entry_points_bitwise(ref_mesh, name, divergence):
    """discretise! + apply_boundary_conditions! + implicit_relaxation! +
    setReference!, then mul!, residual!, the DILU update / ldiv! / Dinv and
    one run! solve, each through its own entry point."""
    mesh = ref_mesh(name)
    dm = mesh.device()
    n, nf = mesh.ncells, mesh.nfaces
    ux, uy, phi = _fields(mesh, 11)
    fx, fy = ref.interpolate_vector(mesh, ux, uy)
    mdotf = ref.flux(mesh, fx, fy, np.zeros(nf))
    nueff = 1e-2 * (1.0 + 0.1 * np.cos(np.arange(nf)))
    si = 0.5 + np.abs(phi)
    src = np.sin(np.arange(n) * 0.01)
    spec = [(L.OP_TIME, F.Euler, 1, None, None),
            (L.OP_DIVERGENCE, divergence, 1, mdotf, None),
            (L.OP_LAPLACIAN, F.Linear, -1, nueff, None),
            (L.OP_SI, F.Linear, 1, si, None)]
    T = _Terms(spec, n)
    srcs_d, srcs_h = (L.Source * 1)(), (L.Source * 1)()
    d_src = _dev(src)
    srcs_d[0].sign, srcs_d[0].field = -1, d_src.ptr
    srcs_h[0].sign, srcs_h[0].field = -1, src.ctypes.data
    bcs = _bc_tuple(mesh)
    barr = _bcs(mesh, bcs, True)
    prev, dt, alpha, pref, refcell = ux * 0.9, 0.05, 0.7, 0.25, n // 3
    d_prev, d_ux = _dev(prev), _dev(ux)

    eq = C.c_void_p()
    L.check(LIB().fvh_eqn_create(dm.h, C.byref(eq)), "fvh_eqn_create")
    try:
        L.check(LIB().fvh_discretise(eq, T.dev, T.n, srcs_d, 1, d_prev.vp, dt), "fvh_discretise")
        L.check(LIB().fvh_apply_bcs(eq, T.dev, T.n, barr, len(bcs), 0), "fvh_apply_bcs")
        A0, b0 = _eq_arrays(eq, dm, n)
        A0r, b0r = ref.assemble(mesh, T.host, T.n, srcs_h, 1, prev, dt, barr, len(bcs), 0)
        assert same_bits(A0, A0r) and same_bits(b0, b0r), "fvh_discretise + fvh_apply_bcs"
        L.check(LIB().fvh_implicit_relax(eq, d_ux.vp, alpha), "fvh_implicit_relax")
        L.check(LIB().fvh_set_reference(eq, pref, refcell), "fvh_set_reference")
        A, b = _eq_arrays(eq, dm, n)
        Ar, br = ref.assemble(mesh, T.host, T.n, srcs_h, 1, prev, dt, barr, len(bcs), 0, relax_field=ux,
                              alpha=alpha, pref=pref, refcell=refcell)
        assert same_bits(A, Ar) and same_bits(b, br), "fvh_implicit_relax + fvh_set_reference"

        # mul! and residual!
        y = _dev(np.zeros(n))
        L.check(LIB().fvh_spmv(eq, d_ux.vp, y.vp), "fvh_spmv")
        Ax = ref.spmv_csr(mesh, Ar, ux)
        assert same_bits(y.host(), Ax), "fvh_spmv"
        r = C.c_double()
        L.check(LIB().fvh_residual(eq, d_ux.vp, C.byref(r)), "fvh_residual")
        R = np.abs(Ax - br) ** 2
        r_ref = np.sqrt(ref.det_dot(R, np.ones(n)) / n) / np.sqrt(ref.det_dot(br, br))
        assert same_bits(np.array([r.value]), np.array([r_ref])), ("fvh_residual", r.value, r_ref)

        # DILU: update, reciprocal diagonal, ldiv!
        L.check(LIB().fvh_precond_set(eq, L.PRECOND_DILU, None), "fvh_precond_set")
        L.check(LIB().fvh_precond_update(eq), "fvh_precond_update")
        dinv = np.zeros(n)
        L.check(LIB().fvh_precond_dinv(eq, L.dptr(dinv)), "fvh_precond_dinv")
        cfg = L.SolverCfg(solver=L.BICGSTAB, precond=L.PRECOND_DILU, itmax=100, precision=L.FP64, convergence=1e-7,
                          relax=1.0, atol=float(np.sqrt(np.finfo(float).eps)), rtol=1e-3)
        x_ref, st_ref, dinv_ref = ref.solve_csr(mesh, Ar, br, cfg, ux)
        assert same_bits(dinv, dinv_ref), "fvh_precond_dinv"
        z = _dev(np.zeros(n))
        L.check(LIB().fvh_precond_apply(eq, z.vp, d_ux.vp), "fvh_precond_apply")
        assert same_bits(z.host(), ref.precond_apply_csr(mesh, Ar, L.PRECOND_DILU, ux)), "fvh_precond_apply"

        # run!: BiCGSTAB + DILU warm-started from ux
        x = _dev(ux)
        st = L.Stats()
        L.check(LIB().fvh_solve(eq, C.byref(cfg), x.vp, C.byref(st)), "fvh_solve")
        assert st.niter == st_ref["niter"], (st.niter, st_ref)
        assert same_bits(x.host(), x_ref), f"fvh_solve (rel {rel_l2(x.host(), x_ref):.2e})"
    finally:
        LIB().fvh_eqn_destroy(eq)


def test_sweep_fault_is_reported_by_the_standalone_entry_points(quad40):
    """A pending hand-over fault of the pipelined DILU sweeps is returned by
    the next fvh_precond_apply (once), not swallowed; the flag is set from
    the host, so every kernel runs normally."""
    mesh = quad40
    dm = mesh.device()
    n = mesh.ncells
    g = F.Simulation(cavity_model(mesh), config(1), pref=0.0)
    assert g.run(1) == 1
    eq = C.c_void_p()
    L.check(LIB().fvh_sim_equation(g.h, b"ux", C.byref(eq)))
    pending = C.c_int32(-1)
    L.check(LIB().fvh_precond_fault(eq, 0, C.byref(pending)))
    assert pending.value == 0
    x, y = _dev(np.linspace(0.0, 1.0, n)), _dev(np.zeros(n))
    L.check(LIB().fvh_precond_apply(eq, y.vp, x.vp))
    good = y.host()
    L.check(LIB().fvh_precond_fault(eq, 1, C.byref(pending)))
    assert pending.value == 1
    assert LIB().fvh_precond_apply(eq, y.vp, x.vp) == L.FVH_ERR_HIP
    buf = C.create_string_buffer(512)
    LIB().fvh_last_error(buf, 512)
    assert b"hand-over" in buf.value
    L.check(LIB().fvh_precond_fault(eq, 0, C.byref(pending)))
    assert pending.value == 0  # reported once
    L.check(LIB().fvh_precond_apply(eq, y.vp, x.vp))
    assert same_bits(y.host(), good)
    # uy borrows ux's preconditioner: its calls see ux's pending fault too
    eqy = C.c_void_p()
    L.check(LIB().fvh_sim_equation(g.h, b"uy", C.byref(eqy)))
    L.check(LIB().fvh_precond_fault(eqy, 1, None))
    assert LIB().fvh_precond_apply(eqy, y.vp, x.vp) == L.FVH_ERR_HIP
    L.check(LIB().fvh_precond_apply(eqy, y.vp, x.vp))
    assert dm is not None


@pytest.mark.parametrize("name,kind", [("quad40", "cavity"), ("trig40", "cavity"), ("backwardFacingStep_10mm", "bfs"),
                                       ("quad100", "cavity")])
def test_fvh_simple_whole_run_bitwise(ref_mesh, name, kind):
    """simple!(model, config; pref) in one call (host fields in and out)."""
    mesh = ref_mesh(name)
    if kind == "cavity":
        model, pref = cavity_model(mesh), 0.0
        cfg = config(40)
    else:
        model, pref = bfs_model(mesh), None
        cfg = config(40, U_solver=F.GmresSolver, p_solver=F.GmresSolver, relax_U=0.8, relax_p=0.2)
    case = F.CaseDesc(model, cfg, pref)
    fields = F.api._fields(model)
    f = {k: v.copy() for k, v in fields.items()}
    it = cfg.runtime["iterations"]
    R = [np.ones(it) for _ in range(3)]
    done = C.c_int32()
    g = lambda k: L.dptr(f[k]) if k in f else None  # noqa: E731
    L.check(LIB().fvh_simple(mesh.device().h, C.byref(case.c), g("Ux"), g("Uy"), g("Uz"), g("p"), g("k"),
                             g("omega"), g("nut"), *(L.dptr(r) for r in R), C.byref(done)), "fvh_simple")
    fr, Rr, done_r = ref.simple(mesh, case, fields)
    print(f"{name}: {done.value} iterations, R_p[end] {R[2][done.value - 1]:.3e}", flush=True)
    assert done.value == done_r
    for a, b in zip(R, Rr):
        assert same_bits(a, b)
    for k in ("Ux", "Uy", "p"):
        assert same_bits(f[k], fr[k]), k
