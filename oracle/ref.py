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
        L.ref_det_dot.argtypes = [C.c_int64, P_D, P_D]
        L.ref_solve_csr.argtypes = [vp, P_D, P_D, vp, P_D, vp, P_D]
        L.ref_precond_apply_csr.argtypes = [vp, P_D, C.c_int32, P_D, P_D]
        L.ref_spmv_csr.argtypes = [vp, P_D, P_D, P_D]
        L.ref_grad.argtypes = [vp, C.c_int32, vp, C.c_int32, C.c_int32, P_D, P_D, P_D, P_D, P_D]
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
