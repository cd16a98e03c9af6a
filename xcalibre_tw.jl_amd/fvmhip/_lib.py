"""ctypes binding of libfvmhip.so (the C ABI declared in include/fvmhip.h).

The library is loaded from this directory; a missing library is an error,
never a silent fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfvmhip.so")

# ---- enums (fvmhip.h) ----
FVH_OK, FVH_ERR_ARG, FVH_ERR_HIP, FVH_ERR_NODEV, FVH_ERR_MESH, FVH_ERR_UNSUPPORTED, FVH_ERR_IO, FVH_ERR_SIZE = range(8)
KRYLOV_STATUS = {0: "solved", 1: "itmax", 2: "breakdown", 3: "zero residual", 4: "zero curvature"}
OP_TIME, OP_LAPLACIAN, OP_DIVERGENCE, OP_SI = range(4)
STEADY, EULER, LINEAR, UPWIND, ORTHOGONAL, MIDPOINT = range(6)
BC_DIRICHLET, BC_NEUMANN, BC_KWALL, BC_OMEGAWALL, BC_NUTWALL = range(5)
BICGSTAB, CG, GMRES = range(3)
PRECOND_NONE, PRECOND_DILU, PRECOND_ILU0, PRECOND_JACOBI, PRECOND_NORMDIAG, PRECOND_LDL = range(6)
FP64, MIXED, MIXED_FP16 = range(3)
LAMINAR, KOMEGA = range(2)
ALG_SIMPLE, ALG_PISO = range(2)

P_D = C.POINTER(C.c_double)
P_I64 = C.POINTER(C.c_int64)
P_I32 = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [
        ("index_base", C.c_int64), ("ncells", C.c_int64), ("nfaces", C.c_int64), ("nboundaries", C.c_int64),
        ("ncellfaces", C.c_int64), ("nboundaryfaces", C.c_int64),
        ("cell_centre", P_D), ("cell_volume", P_D), ("cell_faces_range", P_I64), ("cell_faces", P_I64),
        ("cell_neighbours", P_I64), ("cell_nsign", P_I64), ("face_owners", P_I64), ("face_centre", P_D),
        ("face_normal", P_D), ("face_e", P_D), ("face_area", P_D), ("face_delta", P_D), ("face_weight", P_D),
        ("boundary_range", P_I64), ("boundary_faces", P_I64), ("boundary_cells", P_I64),
    ]


class BC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("patch", C.c_int32), ("value", C.c_double * 3), ("wf", C.c_double * 5)]


class Term(C.Structure):
    _fields_ = [("op", C.c_int32), ("scheme", C.c_int32), ("sign", C.c_int32), ("pad_", C.c_int32),
                ("flux", C.c_void_p), ("phi", C.c_void_p)]


class Source(C.Structure):
    _fields_ = [("sign", C.c_int32), ("pad_", C.c_int32), ("field", C.c_void_p)]


class Schemes(C.Structure):
    _fields_ = [("time", C.c_int32), ("divergence", C.c_int32), ("laplacian", C.c_int32), ("gradient", C.c_int32)]


class SolverCfg(C.Structure):
    _fields_ = [("solver", C.c_int32), ("precond", C.c_int32), ("itmax", C.c_int32), ("precision", C.c_int32),
                ("convergence", C.c_double), ("relax", C.c_double), ("atol", C.c_double), ("rtol", C.c_double)]


class BCSet(C.Structure):
    _fields_ = [("nbc", C.c_int32), ("pad_", C.c_int32), ("bcs", C.POINTER(BC))]


class Case(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("algorithm", C.c_int32), ("iterations", C.c_int32), ("has_pref", C.c_int32),
        ("pref", C.c_double), ("nu", C.c_double), ("dt", C.c_double), ("conv_break", C.c_double),
        ("U", BCSet), ("p", BCSet), ("k", BCSet), ("omega", BCSet), ("nut", BCSet),
        ("sU", Schemes), ("sp", Schemes), ("sk", Schemes), ("somega", Schemes),
        ("vU", SolverCfg), ("vp", SolverCfg), ("vk", SolverCfg), ("vomega", SolverCfg),
        ("nonorth_correctors", C.c_int32), ("pad_", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [("niter", C.c_int32), ("status", C.c_int32), ("refinements", C.c_int32), ("pad_", C.c_int32),
                ("rnorm", C.c_double), ("rnorm0", C.c_double)]


# name -> (restype, argtypes)
_VP = C.c_void_p
SIGNATURES = {
    "fvh_version": (C.c_char_p, []),
    "fvh_status_string": (C.c_char_p, [C.c_int]),
    "fvh_last_error": (C.c_int, [C.c_char_p, C.c_int64]),
    "fvh_device_count": (C.c_int, [P_I32]),
    "fvh_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64]),
    "fvh_free": (C.c_int, [C.c_void_p]),
    "fvh_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "fvh_synchronize": (C.c_int, []),
    "fvh_hmesh_build": (C.c_int, [C.c_int64, P_D, C.c_int64, P_I64, P_I64, C.c_int64, P_I64, P_I64, C.c_char_p,
                                  C.c_double, C.POINTER(_VP)]),
    "fvh_hmesh_read_unv": (C.c_int, [C.c_char_p, C.c_double, C.POINTER(_VP)]),
    "fvh_hmesh_desc": (C.c_int, [_VP, C.POINTER(MeshDesc)]),
    "fvh_hmesh_boundary_name": (C.c_int, [_VP, C.c_int64, C.c_char_p, C.c_int64]),
    "fvh_hmesh_nodes": (C.c_int, [_VP, C.POINTER(C.c_int64), C.POINTER(P_D), C.POINTER(P_I64), C.POINTER(P_I64)]),
    "fvh_hmesh_free": (None, [_VP]),
    "fvh_mesh_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(_VP)]),
    "fvh_mesh_destroy": (None, [_VP]),
    "fvh_mesh_pattern": (C.c_int, [_VP, P_I64, P_I32, P_I32]),
    "fvh_mesh_sweep_info": (C.c_int, [_VP, P_I64, P_I64]),
    "fvh_interpolate": (C.c_int, [_VP, _VP, _VP]),
    "fvh_interpolate_vector": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fvh_correct_boundaries": (C.c_int, [_VP, C.POINTER(BC), C.c_int32, C.c_int32, _VP, _VP]),
    "fvh_green_gauss": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fvh_grad": (C.c_int, [_VP, C.c_int32, C.POINTER(BC), C.c_int32, C.c_int32, _VP, _VP, _VP, _VP, _VP]),
    "fvh_div": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fvh_flux": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fvh_bound": (C.c_int, [_VP, _VP, C.c_double, P_I32]),
    "fvh_write_vtk": (C.c_int, [C.c_char_p, C.c_int64, P_D, C.c_int64, P_I64, P_I64, C.c_int32,
                                C.POINTER(C.c_char_p), P_I32, C.POINTER(P_D)]),
    "fvh_format_julia": (C.c_int, [C.c_double, C.c_char_p, C.c_int64]),
    "fvh_patch_size": (C.c_int, [_VP, C.c_int32, C.POINTER(C.c_int64)]),
    "fvh_surface_flux": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP, _VP]),
    "fvh_surface_normal_gradient": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP, P_D, _VP, _VP, _VP]),
    "fvh_pressure_force": (C.c_int, [_VP, C.c_int32, _VP, C.c_double, P_D]),
    "fvh_viscous_force": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP, P_D, C.c_double, C.c_double, _VP, P_D]),
    "fvh_wall_shear_stress": (C.c_int, [_VP, C.c_int32, _VP, _VP, _VP, P_D, C.c_double, _VP, _VP, _VP, _VP, _VP]),
    "fvh_interpolate_gradient": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "fvh_nonorthogonal_flux": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "fvh_nonorthogonal_correct": (C.c_int, [_VP, C.c_int32, _VP, _VP]),
    "fvh_eqn_create": (C.c_int, [_VP, C.POINTER(_VP)]),
    "fvh_eqn_destroy": (None, [_VP]),
    "fvh_eqn_arrays": (C.c_int, [_VP, C.POINTER(_VP), C.POINTER(_VP)]),
    "fvh_discretise": (C.c_int, [_VP, C.POINTER(Term), C.c_int32, C.POINTER(Source), C.c_int32, _VP, C.c_double]),
    "fvh_apply_bcs": (C.c_int, [_VP, C.POINTER(Term), C.c_int32, C.POINTER(BC), C.c_int32, C.c_int32]),
    "fvh_implicit_relax": (C.c_int, [_VP, _VP, C.c_double]),
    "fvh_set_reference": (C.c_int, [_VP, C.c_double, C.c_int64]),
    "fvh_spmv": (C.c_int, [_VP, _VP, _VP]),
    "fvh_residual": (C.c_int, [_VP, _VP, P_D]),
    "fvh_residual_noalloc": (C.c_int, [_VP, _VP, P_D]),
    "fvh_precond_set": (C.c_int, [_VP, C.c_int32, _VP]),
    "fvh_precond_update": (C.c_int, [_VP]),
    "fvh_precond_apply": (C.c_int, [_VP, _VP, _VP]),
    "fvh_precond_dinv": (C.c_int, [_VP, P_D]),
    "fvh_precond_fault": (C.c_int, [_VP, C.c_int32, P_I32]),
    "fvh_solve": (C.c_int, [_VP, C.POINTER(SolverCfg), _VP, C.POINTER(Stats)]),
    "fvh_spmv_f32": (C.c_int, [_VP, _VP, _VP]),
    "fvh_mesh_set_cases": (C.c_int, [_VP, C.c_int32, P_I64]),
    "fvh_mesh_case_info": (C.c_int, [_VP, P_I32, P_I32]),
    "fvh_mesh_tuning": (C.c_int, [_VP, P_I32]),
    "fvh_sim_create_cases": (C.c_int, [_VP, C.POINTER(Case), P_D, P_I64, P_D, P_D, P_D, P_D, C.POINTER(_VP)]),
    "fvh_sim_case_residuals": (C.c_int, [_VP, C.c_int32, P_D, P_D, P_D]),
    "fvh_sim_case_stats": (C.c_int, [_VP, C.c_int32, C.c_char_p, C.POINTER(Stats)]),
    "fvh_precond_apply_f32": (C.c_int, [_VP, _VP, _VP]),
    "fvh_sim_create": (C.c_int, [_VP, C.POINTER(Case), P_D, P_D, P_D, P_D, P_D, P_D, P_D, C.POINTER(_VP)]),
    "fvh_sim_destroy": (None, [_VP]),
    "fvh_sim_iterate": (C.c_int, [_VP, C.c_int32, P_I32]),
    "fvh_sim_residuals": (C.c_int, [_VP, P_D, P_D, P_D]),
    "fvh_sim_get": (C.c_int, [_VP, C.c_char_p, P_D, P_I64]),
    "fvh_sim_stats": (C.c_int, [_VP, C.c_char_p, C.POINTER(Stats)]),
    "fvh_sim_timing": (C.c_int, [_VP, P_D, P_D, P_D]),
    "fvh_sim_equation": (C.c_int, [_VP, C.c_char_p, C.POINTER(_VP)]),
    "fvh_sim_time_op": (C.c_int, [_VP, C.c_int32, C.c_int32, P_D]),
    "fvh_time_op": (C.c_int, [_VP, C.c_int32, C.c_int32, _VP, _VP, P_D]),
    "fvh_simple": (C.c_int, [_VP, C.POINTER(Case), P_D, P_D, P_D, P_D, P_D, P_D, P_D, P_D, P_D, P_D, P_I32]),
}

_lib = None


class FVHError(RuntimeError):
    pass


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FVHError(f"{LIB_PATH} is missing: run `python xcalibre_tw.jl_amd/build.py` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != FVH_OK:
        buf = C.create_string_buffer(1024)
        lib().fvh_last_error(buf, 1024)
        raise FVHError(f"{what or 'fvmhip'}: {lib().fvh_status_string(rc).decode()} ({buf.value.decode()})")


def dptr(a):
    """double* of a contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P_D)


def iptr(a):
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P_I64)


def device_count():
    n = C.c_int32(0)
    check(lib().fvh_device_count(C.byref(n)))
    return n.value


class DeviceArray:
    """A float64 (or int) array in device memory owned by libfvmhip."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        check(lib().fvh_malloc(C.byref(p), self.n * self.dtype.itemsize), "fvh_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        check(lib().fvh_memcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes), "fvh_memcpy")

    def host(self):
        out = np.empty(self.n, self.dtype)
        check(lib().fvh_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes), "fvh_memcpy")
        return out

    @property
    def vp(self):
        return C.c_void_p(self.ptr)

    def __del__(self):
        try:
            if self.ptr:
                lib().fvh_free(C.c_void_p(self.ptr))
        except Exception:
            pass
