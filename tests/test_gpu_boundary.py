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
from tests.cases import bfs_case_model, bfs_model, cavity_model, config, rel_l2, same_bits
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
            (L.OP_SI, F.Linear, 1, si, ux)]
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
        model, pref = bfs_case_model(mesh), None
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


@pytest.mark.parametrize("bnd", [np.finfo(float).eps, 1e-3])
def test_bound_entry_point_against_the_sequential_oracle(bnd):
    """bound! with chains of adjacent negative cells (a whole grid row, a
    diagonal staircase, a random third of the cells, NaN and -0.0 among
    them), through fvh_bound, bit for bit against the oracle's in-order
    pass; values between 0 and the bound move too when the bound exceeds
    eps."""
    mesh = F.cavity(64)
    dm = mesh.device()
    n = mesh.ncells
    rng = np.random.default_rng(3)
    base = rng.uniform(0.5, 2.0, n)
    fields = []
    f = base.copy()
    f[64 * 10:64 * 11] = -rng.uniform(0.1, 1.0, 64)        # one grid row: a chain of length 64
    fields.append(f)
    f = base.copy()
    for k in range(40):                                     # a staircase (west and south neighbours)
        f[64 * k + k] = -1.0
        f[64 * k + k + 1] = -0.5
    fields.append(f)
    f = np.where(rng.random(n) < 0.33, -rng.uniform(0.0, 1.0, n), base)
    f[5], f[6], f[70] = -0.0, np.nan, 1e-300               # -0.0, a NaN, a positive value below any bound
    f[rng.random(n) < 0.05] = 0.5 * bnd
    fields.append(f)
    for t, f in enumerate(fields):
        d = _dev(f)
        chained = C.c_int32(-1)
        L.check(LIB().fvh_bound(dm.h, d.vp, bnd, C.byref(chained)), "fvh_bound")
        want = ref.bound(mesh, f, bnd)
        got = d.host()
        print(f"field {t}: {chained.value} cells through the ordered pass", flush=True)
        assert chained.value > 0
        assert same_bits(got, want), (t, int(np.sum(~((got == want) | (np.isnan(got) & np.isnan(want))))))
    # no negative cell: nothing is chained, values below the bound are lifted
    d = _dev(base)
    chained = C.c_int32(-1)
    L.check(LIB().fvh_bound(dm.h, d.vp, bnd, C.byref(chained)), "fvh_bound")
    assert chained.value == 0 and same_bits(d.host(), ref.bound(mesh, base, bnd))


@pytest.mark.parametrize("name", REFERENCE_MESHES)
def test_nonorthogonal_entry_points_bitwise(ref_mesh, name):
    """interpolate!(gradf, grad, phi), nonorthogonal_flux! and correct!
    through their entry points, bit for bit against the oracle."""
    mesh = ref_mesh(name)
    dm = mesh.device()
    n, nf = mesh.ncells, mesh.nfaces
    ux, uy, phi = _fields(mesh, 21)
    gz = 0.1 * ux * uy
    start = np.linspace(-3.0, 3.0, nf)
    d_g = [_dev(a) for a in (ux, uy, gz)]
    d_phi = _dev(phi)
    d_f = [_dev(start) for _ in range(3)]
    L.check(LIB().fvh_interpolate_gradient(dm.h, *(a.vp for a in d_g), d_phi.vp, *(a.vp for a in d_f)),
            "fvh_interpolate_gradient")
    want = ref.interpolate_gradient(mesh, ux, uy, gz, phi, start, start, start)
    for a, b in zip(d_f, want):
        assert same_bits(a.host(), b), "fvh_interpolate_gradient"
    d_pf = _dev(start)
    L.check(LIB().fvh_nonorthogonal_flux(dm.h, *(a.vp for a in d_f), d_pf.vp), "fvh_nonorthogonal_flux")
    corr = ref.nonorthogonal_flux(mesh, *want, start)
    assert same_bits(d_pf.host(), corr), "fvh_nonorthogonal_flux"
    # correct! on an assembled equation's b
    J = 0.5 + np.abs(np.cos(np.arange(nf) * 0.37))
    eq = C.c_void_p()
    L.check(LIB().fvh_eqn_create(dm.h, C.byref(eq)), "fvh_eqn_create")
    try:
        T = _Terms([(L.OP_LAPLACIAN, F.Linear, 1, J, None)], n)
        src = np.sin(np.arange(n) * 0.05)
        srcs_d = (L.Source * 1)()
        d_src = _dev(src)
        srcs_d[0].sign, srcs_d[0].field = 1, d_src.ptr
        L.check(LIB().fvh_discretise(eq, T.dev, T.n, srcs_d, 1, None, 0.0), "fvh_discretise")
        _, b0 = _eq_arrays(eq, dm, n)
        d_J = _dev(J)
        L.check(LIB().fvh_nonorthogonal_correct(eq, 1, d_J.vp, d_pf.vp), "fvh_nonorthogonal_correct")
        _, b1 = _eq_arrays(eq, dm, n)
        assert same_bits(b1, ref.nonorthogonal_correct(mesh, 1, J, corr, b0)), "fvh_nonorthogonal_correct"
    finally:
        LIB().fvh_eqn_destroy(eq)


@pytest.mark.parametrize("mesh_name,algorithm", [("trig40", L.ALG_SIMPLE), ("quad40", L.ALG_SIMPLE),
                                                 ("trig40", L.ALG_PISO)])
def test_loop_with_nonorthogonal_correctors_bitwise(ref_mesh, mesh_name, algorithm):
    """The loops' `correct` block switched on (one corrector): a cavity on
    the reference's triangle mesh, bit for bit against the oracle."""
    mesh = ref_mesh(mesh_name)
    model = cavity_model(mesh)
    it = 6
    kw = {} if algorithm == L.ALG_SIMPLE else dict(time=F.Euler, dt=0.01)
    cfg = config(it + 1, **kw)
    g = F.Simulation(model, cfg, 0.0, algorithm, 1e-7, nonorth_correctors=1)
    r = ref.RefSimulation(model, F.CaseDesc(model, cfg, 0.0, algorithm, 1e-7, nonorth_correctors=1))
    assert g.run(it) == r.run(it) == it
    for nm in ("Ux", "Uy", "p"):
        assert same_bits(g.get(nm), r.get(nm)), f"{nm} rel {rel_l2(g.get(nm), r.get(nm)):.3e}"
    for a, b in zip(g.residuals(), r.residuals()):
        assert same_bits(a[:it], b[:it])
    # the correctors change the answer on triangles (not on quads, to rounding)
    g0 = F.Simulation(model, cfg, 0.0, algorithm, 1e-7)
    g0.run(it)
    d = rel_l2(g.get("p"), g0.get("p"))
    print(f"{mesh_name}: correctors move p by {d:.3e} (rel L2)")
    if mesh_name == "trig40":
        assert d > 1e-8


@pytest.mark.parametrize("name", REFERENCE_MESHES)
def test_postprocess_entry_points_bitwise(ref_mesh, name):
    """surface_flux, surface_normal_gradient, pressure_force, viscous_force
    and wall_shear_stress on every patch, bit for bit against the oracle."""
    mesh = ref_mesh(name)
    dm = mesh.device()
    n = mesh.ncells
    ux, uy, phi = _fields(mesh, 31)
    uz = 0.05 * ux
    nut = 1e-3 * (1.0 + np.abs(uy))
    d_u = [_dev(a) for a in (ux, uy, uz)]
    d_p, d_nut = _dev(phi), _dev(nut)
    uw = np.array([0.25, -0.125, 0.0])
    for b, patch in enumerate(mesh.boundary_names):
        nb = C.c_int64()
        L.check(LIB().fvh_patch_size(dm.h, b, C.byref(nb)))
        k = nb.value
        out = [_dev(np.zeros(max(k, 1))) for _ in range(3)]
        L.check(LIB().fvh_surface_flux(dm.h, b, d_p.vp, *(o.vp for o in out)), "fvh_surface_flux")
        for o, w in zip(out, ref.surface_flux(mesh, b, phi)):
            assert same_bits(o.host()[:k], w), ("surface_flux", patch)
        L.check(LIB().fvh_surface_normal_gradient(dm.h, b, *(a.vp for a in d_u), L.dptr(uw), *(o.vp for o in out)),
                "fvh_surface_normal_gradient")
        for o, w in zip(out, ref.surface_normal_gradient(mesh, b, ux, uy, uz, uw)):
            assert same_bits(o.host()[:k], w), ("surface_normal_gradient", patch)
        Fp = np.zeros(3)
        L.check(LIB().fvh_pressure_force(dm.h, b, d_p.vp, 1.2, L.dptr(Fp)), "fvh_pressure_force")
        assert same_bits(Fp, ref.pressure_force(mesh, b, phi, 1.2)), ("pressure_force", patch)
        for with_nut in (False, True):
            for w in (uw, None):
                Fv = np.zeros(3)
                L.check(LIB().fvh_viscous_force(dm.h, b, *(a.vp for a in d_u), None if w is None else L.dptr(w), 1.2,
                                                1.5e-5, d_nut.vp if with_nut else None, L.dptr(Fv)), "fvh_viscous_force")
                want = ref.viscous_force(mesh, b, ux, uy, uz, w, 1.2, 1.5e-5, nut if with_nut else None)
                assert same_bits(Fv, want), ("viscous_force", patch, with_nut, w is None)
        pos = _dev(np.zeros(3 * max(k, 1)))
        L.check(LIB().fvh_wall_shear_stress(dm.h, b, *(a.vp for a in d_u), L.dptr(uw), 1.5e-5, d_nut.vp,
                                            *(o.vp for o in out), pos.vp), "fvh_wall_shear_stress")
        (tx, ty, tz), wp = ref.wall_shear_stress(mesh, b, ux, uy, uz, uw, 1.5e-5, nut)
        for o, w in zip(out, (tx, ty, tz)):
            assert same_bits(o.host()[:k], w), ("wall_shear_stress", patch)
        assert same_bits(pos.host()[:3 * k], wp.reshape(-1)), ("wall_shear_stress positions", patch)


def test_simple_writes_vtk_as_write_interval_asks(ref_mesh, tmp_path, monkeypatch):
    """simple! with write_interval = 4 on the cavity: iteration_000004.vtk,
    _000008 ... and the file at convergence, holding the fields the model
    was updated with (Solvers_1_SIMPLE.jl:200-228)."""
    import os
    mesh = ref_mesh("quad40")
    model = cavity_model(mesh)
    cfg = config(10)
    cfg.runtime["write_interval"] = 4
    monkeypatch.chdir(tmp_path)
    R = F.simple(model, cfg, pref=0.0)
    files = sorted(os.listdir(tmp_path))
    assert files == ["iteration_000004.vtk", "iteration_000008.vtk"], files
    with open(tmp_path / "iteration_000008.vtk") as f:
        text = f.read()
    assert text.startswith("# vtk DataFile Version 3.0\njCFD simulation data\n")
    F.model2vtk(model, str(tmp_path / "again"))
    with open(tmp_path / "again.vtk") as f:
        assert f.read() != text  # fields moved on after iteration 8 (10 iterations ran)
    assert len(R[0]) == 10
    # a run that converges inside an interval writes at the break
    model2 = cavity_model(mesh)
    cfg2 = config(1000)
    cfg2.runtime["write_interval"] = 100
    for fn in files:
        os.remove(tmp_path / fn)
    os.remove(tmp_path / "again.vtk")
    sim_its = F.Simulation(cavity_model(mesh), config(1000), 0.0).run(1000)
    F.simple(model2, cfg2, pref=0.0)
    want = [f"iteration_{k:06d}.vtk" for k in range(100, sim_its, 100)] + [f"iteration_{sim_its:06d}.vtk"]
    assert sorted(os.listdir(tmp_path)) == want
