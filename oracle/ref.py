"""ORACLE — test infrastructure only (the checker, never the product).

ctypes access to libref_cpu.so, the CPU FP64 restatement of the reference
hot path (ref_cpu.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libref_cpu.so")
P_D = C.POINTER(C.c_double)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.ref_last_error.restype = C.c_char_p
        L.ref_sim_create.argtypes = [vp, vp, P_D, P_D, P_D, P_D, P_D, P_D, P_D, C.POINTER(vp)]
        L.ref_sim_destroy.argtypes = [vp]
        L.ref_sim_destroy.restype = None
        L.ref_sim_iterate.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
        L.ref_sim_residuals.argtypes = [vp, P_D, P_D, P_D]
        L.ref_sim_get.argtypes = [vp, C.c_char_p, P_D, C.POINTER(C.c_int64)]
        L.ref_sim_stats.argtypes = [vp, C.c_char_p, vp]
        L.ref_det_dot.restype = C.c_double
        for fn, nargs in (("ref_log", 1), ("ref_y_plus_laminar", 2), ("ref_nut_wall", 4), ("ref_pow", 2)):
            getattr(L, fn).restype = C.c_double
            getattr(L, fn).argtypes = [C.c_double] * nargs
        L.ref_det_dot.argtypes = [C.c_int64, P_D, P_D]
        L.ref_solve_csr.argtypes = [vp, P_D, P_D, vp, P_D, vp, P_D]
        L.ref_precond_apply_csr.argtypes = [vp, P_D, C.c_int32, P_D, P_D]
        L.ref_spmv_csr.argtypes = [vp, P_D, P_D, P_D]
        L.ref_grad.argtypes = [vp, C.c_int32, vp, C.c_int32, C.c_int32, P_D, P_D, P_D, P_D, P_D]
        L.ref_interpolate.argtypes = [vp, P_D, P_D]
        L.ref_interpolate_vector.argtypes = [vp, P_D, P_D, P_D, P_D]
        L.ref_correct_boundaries.argtypes = [vp, vp, C.c_int32, C.c_int32, P_D, P_D]
        L.ref_green_gauss.argtypes = [vp, P_D, P_D, P_D, P_D]
        L.ref_div.argtypes = [vp, P_D, P_D, P_D, P_D]
        L.ref_flux.argtypes = [vp, P_D, P_D, P_D, P_D]
        L.ref_interpolate_gradient.argtypes = [vp, P_D, P_D, P_D, P_D, P_D, P_D, P_D]
        L.ref_surface_flux.argtypes = [vp, C.c_int32, P_D, P_D, P_D, P_D]
        L.ref_surface_normal_gradient.argtypes = [vp, C.c_int32, P_D, P_D, P_D, P_D, P_D, P_D, P_D]
        L.ref_pressure_force.argtypes = [vp, C.c_int32, P_D, C.c_double, P_D]
        L.ref_viscous_force.argtypes = [vp, C.c_int32, P_D, P_D, P_D, P_D, C.c_double, C.c_double, P_D, P_D]
        L.ref_wall_shear_stress.argtypes = [vp, C.c_int32, P_D, P_D, P_D, P_D, C.c_double, P_D, P_D, P_D, P_D, P_D]
        L.ref_nonorthogonal_flux.argtypes = [vp, P_D, P_D, P_D, P_D]
        L.ref_nonorthogonal_correct.argtypes = [vp, C.c_int32, P_D, P_D, P_D]
        L.ref_assemble.argtypes = [vp, vp, C.c_int32, vp, C.c_int32, P_D, C.c_double, vp, C.c_int32, C.c_int32,
                                   P_D, C.c_double, C.c_int32, C.c_double, C.c_int64, P_D, P_D]
        L.ref_bound.argtypes = [vp, P_D, C.c_double]
        L.ref_strain_production.argtypes = [vp, C.c_int32, vp, C.c_int32, P_D, P_D, P_D, P_D]
        L.ref_omega_constrain.argtypes = [vp, vp, C.c_int32, P_D, C.c_double, P_D, P_D]
        L.ref_simple.argtypes = [vp, vp, P_D, P_D, P_D, P_D, P_D, P_D, P_D, P_D, P_D, P_D, C.POINTER(C.c_int32)]
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(P_D)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().ref_last_error().decode())


class RefSimulation:
    """Oracle counterpart of fvmhip.Simulation (same case description)."""

    def __init__(self, model, case):
        from fvmhip.api import _fields  # the field arrays of the model (host data)

        self.mesh = model.mesh
        self.desc = model.mesh.desc()
        self.case = case
        f = _fields(model)
        self._f = {k: v.copy() for k, v in f.items()}
        g = lambda k: _d(self._f[k]) if k in self._f else None  # noqa: E731
        h = C.c_void_p()
        _check(lib().ref_sim_create(C.byref(self.desc), C.byref(case.c), g("Ux"), g("Uy"), g("Uz"), g("p"), g("k"),
                                    g("omega"), g("nut"), C.byref(h)))
        self.h = h
        self.iterations = case.c.iterations

    def run(self, n):
        done = C.c_int32()
        _check(lib().ref_sim_iterate(self.h, int(n), C.byref(done)))
        return done.value

    def residuals(self):
        R = [np.ones(self.iterations) for _ in range(3)]
        _check(lib().ref_sim_residuals(self.h, *(_d(r) for r in R)))
        return tuple(R)

    def get(self, name):
        ln = C.c_int64()
        _check(lib().ref_sim_get(self.h, name.encode(), None, C.byref(ln)))
        out = np.zeros(ln.value)
        _check(lib().ref_sim_get(self.h, name.encode(), _d(out), C.byref(ln)))
        return out

    def stats(self, field):
        from fvmhip import _lib as FL

        s = FL.Stats()
        _check(lib().ref_sim_stats(self.h, field.encode(), C.byref(s)))
        return dict(niter=s.niter, status=FL.KRYLOV_STATUS.get(s.status, s.status), rnorm=s.rnorm, rnorm0=s.rnorm0)

    def __del__(self):
        try:
            lib().ref_sim_destroy(self.h)
        except Exception:
            pass


def det_dot(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().ref_det_dot(x.size, _d(x), _d(y))


def spmv_csr(mesh, vals, x):
    d = mesh.desc()
    vals = np.ascontiguousarray(vals, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(mesh.ncells)
    _check(lib().ref_spmv_csr(C.byref(d), _d(vals), _d(x), _d(y)))
    return y


def precond_apply_csr(mesh, vals, kind, x):
    d = mesh.desc()
    vals = np.ascontiguousarray(vals, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(mesh.ncells)
    _check(lib().ref_precond_apply_csr(C.byref(d), _d(vals), int(kind), _d(x), _d(y)))
    return y


def solve_csr(mesh, vals, b, cfg, x0):
    """One run!-style Krylov solve; cfg is an fvmhip._lib.SolverCfg."""
    from fvmhip import _lib as FL

    d = mesh.desc()
    vals = np.ascontiguousarray(vals, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, np.float64, copy=True)
    st = FL.Stats()
    dinv = np.zeros(mesh.ncells)
    _check(lib().ref_solve_csr(C.byref(d), _d(vals), _d(b), C.byref(cfg), _d(x), C.byref(st), _d(dinv)))
    return x, dict(niter=st.niter, status=FL.KRYLOV_STATUS.get(st.status, st.status), rnorm=st.rnorm), dinv


def grad(mesh, scheme, bcs_arr, nbc, comp, phi):
    d = mesh.desc()
    phi = np.ascontiguousarray(phi, np.float64)
    pf = np.zeros(mesh.nfaces)
    g = [np.zeros(mesh.ncells) for _ in range(3)]
    _check(lib().ref_grad(C.byref(d), int(scheme), C.cast(bcs_arr, C.c_void_p), int(nbc), int(comp), _d(phi), _d(pf),
                          *(_d(a) for a in g)))
    return pf, g


def _arr(a, n=None):
    a = np.ascontiguousarray(a, np.float64)
    return a if n is None else a.reshape(n)


def interpolate(mesh, phi):
    d = mesh.desc()
    out = np.zeros(mesh.nfaces)
    _check(lib().ref_interpolate(C.byref(d), _d(_arr(phi)), _d(out)))
    return out


def interpolate_vector(mesh, ux, uy, fx=None, fy=None):
    """Writes x and y only: fx / fy start from the given arrays (zeros by default)."""
    d = mesh.desc()
    fx = np.zeros(mesh.nfaces) if fx is None else np.array(fx, np.float64)
    fy = np.zeros(mesh.nfaces) if fy is None else np.array(fy, np.float64)
    ux, uy = _arr(ux), _arr(uy)
    _check(lib().ref_interpolate_vector(C.byref(d), _d(ux), _d(uy), _d(fx), _d(fy)))
    return fx, fy


def correct_boundaries(mesh, bcs_arr, nbc, comp, phi, phif):
    d = mesh.desc()
    out = np.array(phif, np.float64)
    _check(lib().ref_correct_boundaries(C.byref(d), C.cast(bcs_arr, C.c_void_p), int(nbc), int(comp),
                                        _d(_arr(phi)), _d(out)))
    return out


def green_gauss(mesh, phif):
    d = mesh.desc()
    g = [np.zeros(mesh.ncells) for _ in range(3)]
    _check(lib().ref_green_gauss(C.byref(d), _d(_arr(phif)), *(_d(a) for a in g)))
    return g


def div(mesh, fx, fy, fz):
    d = mesh.desc()
    out = np.zeros(mesh.ncells)
    fx, fy, fz = _arr(fx), _arr(fy), _arr(fz)
    _check(lib().ref_div(C.byref(d), _d(fx), _d(fy), _d(fz), _d(out)))
    return out


def flux(mesh, fx, fy, fz):
    d = mesh.desc()
    out = np.zeros(mesh.nfaces)
    fx, fy, fz = _arr(fx), _arr(fy), _arr(fz)
    _check(lib().ref_flux(C.byref(d), _d(fx), _d(fy), _d(fz), _d(out)))
    return out


def interpolate_gradient(mesh, gx, gy, gz, phi, fx=None, fy=None, fz=None):
    """interpolate!(gradf, grad, phi): internal faces; the face arrays start
    from the given values (zeros by default) on boundary faces."""
    d = mesh.desc()
    out = [np.zeros(mesh.nfaces) if a is None else np.array(a, np.float64) for a in (fx, fy, fz)]
    gx, gy, gz, phi = _arr(gx), _arr(gy), _arr(gz), _arr(phi)
    _check(lib().ref_interpolate_gradient(C.byref(d), _d(gx), _d(gy), _d(gz), _d(phi), *(_d(a) for a in out)))
    return tuple(out)


def _patch(mesh, patch):
    b = mesh.boundary_index(patch) if isinstance(patch, str) else int(patch)
    s, e = mesh.boundary_range[b]
    return b, int(e - s + 1)


def surface_flux(mesh, patch, phi):
    b, nb = _patch(mesh, patch)
    d = mesh.desc()
    out = [np.zeros(max(nb, 1)) for _ in range(3)]
    _check(lib().ref_surface_flux(C.byref(d), b, _d(_arr(phi)), *(_d(o) for o in out)))
    return tuple(o[:nb] for o in out)


def surface_normal_gradient(mesh, patch, ux, uy, uz, uw):
    b, nb = _patch(mesh, patch)
    d = mesh.desc()
    out = [np.zeros(max(nb, 1)) for _ in range(3)]
    _check(lib().ref_surface_normal_gradient(C.byref(d), b, _d(_arr(ux)), _d(_arr(uy)), _d(_arr(uz)),
                                             _d(_arr(uw)), *(_d(o) for o in out)))
    return tuple(o[:nb] for o in out)


def pressure_force(mesh, patch, p, rho):
    b, _ = _patch(mesh, patch)
    d = mesh.desc()
    F = np.zeros(3)
    _check(lib().ref_pressure_force(C.byref(d), b, _d(_arr(p)), float(rho), _d(F)))
    return F


def viscous_force(mesh, patch, ux, uy, uz, uw, rho, nu, nut=None):
    b, _ = _patch(mesh, patch)
    d = mesh.desc()
    F = np.zeros(3)
    _check(lib().ref_viscous_force(C.byref(d), b, _d(_arr(ux)), _d(_arr(uy)), _d(_arr(uz)),
                                   None if uw is None else _d(_arr(uw)), float(rho), float(nu),
                                   None if nut is None else _d(_arr(nut)), _d(F)))
    return F


def wall_shear_stress(mesh, patch, ux, uy, uz, uw, nu, nut=None):
    b, nb = _patch(mesh, patch)
    d = mesh.desc()
    out = [np.zeros(max(nb, 1)) for _ in range(3)]
    pos = np.zeros(3 * max(nb, 1))
    _check(lib().ref_wall_shear_stress(C.byref(d), b, _d(_arr(ux)), _d(_arr(uy)), _d(_arr(uz)), _d(_arr(uw)),
                                       float(nu), None if nut is None else _d(_arr(nut)), *(_d(o) for o in out),
                                       _d(pos)))
    return tuple(o[:nb] for o in out), pos[:3 * nb].reshape(nb, 3)


def nonorthogonal_flux(mesh, fx, fy, fz, phif):
    """nonorthogonal_flux!: phif (copied) with its internal faces overwritten."""
    d = mesh.desc()
    out = np.array(phif, np.float64)
    fx, fy, fz = _arr(fx), _arr(fy), _arr(fz)
    _check(lib().ref_nonorthogonal_flux(C.byref(d), _d(fx), _d(fy), _d(fz), _d(out)))
    return out


def nonorthogonal_correct(mesh, sign, J, corr, b):
    """correct!(eqn, term, corr_flux) on a copy of b."""
    d = mesh.desc()
    out = np.array(b, np.float64)
    J, corr = _arr(J), _arr(corr)
    _check(lib().ref_nonorthogonal_correct(C.byref(d), int(sign), _d(J), _d(corr), _d(out)))
    return out


def assemble(mesh, terms, nterms, srcs, nsrc, prev, dt, bcs_arr, nbc, comp, relax_field=None, alpha=0.0,
             pref=None, refcell=0):
    """A (CSR order) and b of one equation: discretise!, BCs, then optional
    implicit relaxation and setReference!.  `terms` / `srcs` are fvh_term /
    fvh_source arrays whose pointers are HOST arrays."""
    d = mesh.desc()
    nnz = mesh.ncells + 2 * (mesh.nfaces - mesh.boundary_faces.size)
    vals = np.zeros(nnz)
    b = np.zeros(mesh.ncells)
    prev = None if prev is None else _arr(prev)
    rf = None if relax_field is None else _arr(relax_field)
    _check(lib().ref_assemble(C.byref(d), C.cast(terms, C.c_void_p), int(nterms),
                              None if srcs is None else C.cast(srcs, C.c_void_p), int(nsrc), _d(prev), float(dt),
                              C.cast(bcs_arr, C.c_void_p), int(nbc), int(comp), _d(rf), float(alpha),
                              0 if pref is None else 1, 0.0 if pref is None else float(pref), int(refcell), _d(vals),
                              _d(b)))
    return vals, b


def simple(mesh, case, fields):
    """ref_simple: the whole run on host copies of `fields` (dict of cell
    arrays Ux Uy Uz p [k omega nut]); returns (fields, (R_ux, R_uy, R_p), done)."""
    d = mesh.desc()
    f = {k: np.array(v, np.float64) for k, v in fields.items()}
    it = case.c.iterations
    R = [np.ones(it) for _ in range(3)]
    done = C.c_int32()
    g = lambda k: _d(f[k]) if k in f else None  # noqa: E731
    _check(lib().ref_simple(C.byref(d), C.byref(case.c), g("Ux"), g("Uy"), g("Uz"), g("p"), g("k"), g("omega"),
                            g("nut"), *(_d(r) for r in R), C.byref(done)))
    return f, tuple(R), done.value


def bound(mesh, field, bnd):
    d = mesh.desc()
    f = np.array(field, np.float64)
    _check(lib().ref_bound(C.byref(d), _d(f), float(bnd)))
    return f


def strain_production(mesh, gradient, bcs_arr, nbc, ux, uy, uz):
    d = mesh.desc()
    out = np.zeros(mesh.ncells)
    ux, uy, uz = _arr(ux), _arr(uy), _arr(uz)
    _check(lib().ref_strain_production(C.byref(d), int(gradient), C.cast(bcs_arr, C.c_void_p), int(nbc), _d(ux),
                                       _d(uy), _d(uz), _d(out)))
    return out


def omega_constrain(mesh, bcs_arr, nbc, k, nu, vals, b):
    d = mesh.desc()
    v, bb = np.array(vals, np.float64), np.array(b, np.float64)
    _check(lib().ref_omega_constrain(C.byref(d), C.cast(bcs_arr, C.c_void_p), int(nbc), _d(_arr(k)), float(nu), _d(v),
                                     _d(bb)))
    return v, bb
