"""Output and post-processing on the CPU: the VTK writer's layout and number
format (src/VTK/VTK_writer.jl:22-116), and the oracle's surface flux,
surface-normal gradient, forces and wall shear stress
(src/Postprocess/Postprocess_forces.jl, src/Calculate/
Calculate_5_surface_normal_gradient.jl) against closed forms on the
reference's quad40 mesh (the unit square, 40 x 40)."""
import ctypes as C
import os

import numpy as np
import pytest

import fvmhip as F
from fvmhip import _lib as L
from oracle import ref


def julia(v):
    buf = C.create_string_buffer(64)
    L.check(L.lib().fvh_format_julia(float(v), buf, 64))
    return buf.value.decode()


# print(x) for these Float64 values in Julia (shortest round-trip digits,
# decimal when -4 < decimal point <= 6 digits from the first digit)
JULIA_PRINT = [
    (0.1, "0.1"), (1.0, "1.0"), (-7.0, "-7.0"), (0.0, "0.0"), (-0.0, "-0.0"), (12.5, "12.5"),
    (1e-5, "1.0e-5"), (0.0001, "0.0001"), (0.00015, "0.00015"), (1e6, "1.0e6"), (123456.0, "123456.0"),
    (100000.0, "100000.0"), (1234567.0, "1.234567e6"), (1 / 3, "0.3333333333333333"), (2.5e-10, "2.5e-10"),
    (1.5e15, "1.5e15"), (1e16, "1.0e16"), (5e-324, "5.0e-324"), (1.7976931348623157e308, "1.7976931348623157e308"),
    (float("nan"), "NaN"), (float("inf"), "Inf"), (float("-inf"), "-Inf"), (0.001, "0.001"),
    (1e-4 * 1.5, "0.00015000000000000001"), (-2.2250738585072014e-308, "-2.2250738585072014e-308"),
]


@pytest.mark.parametrize("v,text", JULIA_PRINT)
def test_numbers_print_as_julia_prints_them(v, text):
    assert julia(v) == text


def test_numbers_round_trip():
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-12, 12, 2000),
                           rng.integers(-10 ** 6, 10 ** 6, 200).astype(float)])
    for v in vals:
        s = julia(v)
        assert float(s) == v, (v, s)
        mant = s.lstrip("-").split("e")[0]
        assert "." in mant  # Julia always prints a decimal point


def _parse_vtk(path):
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[-1] == ""
    return lines[:-1]


@pytest.mark.parametrize("name", ["quad40", "trig40"])
def test_vtk_layout(ref_mesh, tmp_path, name):
    mesh = ref_mesh(name)
    n = mesh.ncells
    rng = np.random.default_rng(1)
    model = F.RANS(F.Laminar, mesh=mesh, viscosity=F.ConstantScalar(1e-3))
    model.U.x.values[:] = rng.standard_normal(n)
    model.U.y.values[:] = rng.standard_normal(n) * 1e-7
    model.U.z.values[:] = 0.0
    model.p.values[:] = rng.standard_normal(n) * 1e5
    out = F.model2vtk(model, str(tmp_path / "case"))
    assert out.endswith("case.vtk") and os.path.exists(out)
    lines = _parse_vtk(out)
    npt = mesh.nodes.shape[0]
    assert lines[:5] == ["# vtk DataFile Version 3.0", "jCFD simulation data", "ASCII", "DATASET UNSTRUCTURED_GRID",
                         f"POINTS {npt} double"]
    pts = np.array([[float(t) for t in ln.split(" ")] for ln in lines[5:5 + npt]])
    assert np.array_equal(pts, mesh.nodes)
    k = 5 + npt
    ncn = int(mesh.cell_nodes_ptr[-1])
    assert lines[k] == f"CELLS {n} {ncn + n}"
    for c in range(n):
        nodes = mesh.cell_nodes[mesh.cell_nodes_ptr[c]:mesh.cell_nodes_ptr[c + 1]]
        assert lines[k + 1 + c] == f"{nodes.size}  " + " ".join(str(v) for v in nodes)
    k += 1 + n
    assert lines[k] == f"CELL_TYPES {n}"
    types = {3: "5", 4: "9"}
    assert lines[k + 1:k + 1 + n] == [types[int(d)] for d in np.diff(mesh.cell_nodes_ptr)]
    k += 1 + n
    assert lines[k] == f"CELL_DATA {n}"
    assert lines[k + 1] == "VECTORS U double"
    U = np.array([[float(t) for t in ln.split(" ")] for ln in lines[k + 2:k + 2 + n]])
    assert np.array_equal(U[:, 0], model.U.x.values) and np.array_equal(U[:, 1], model.U.y.values)
    k += 2 + n
    assert lines[k:k + 2] == ["SCALARS p double 1", "LOOKUP_TABLE CellColors"]
    assert [julia(v) for v in model.p.values] == lines[k + 2:k + 2 + n]
    assert len(lines) == k + 2 + n


def test_vtk_of_a_komega_model_has_five_fields(tmp_path):
    from tests.cases import flatplate_model

    mesh = F.flat_plate(nx=8, ny=5, first=1e-3, ratio=1.1)
    model = flatplate_model(mesh)
    lines = _parse_vtk(F.model2vtk(model, str(tmp_path / "kw")))
    heads = [ln for ln in lines if ln.startswith(("SCALARS", "VECTORS"))]
    assert heads == ["VECTORS U double", "SCALARS p double 1", "SCALARS k double 1", "SCALARS omega double 1",
                     "SCALARS nut double 1"]


def test_vtk_rejects_what_the_reference_rejects(ref_mesh, tmp_path):
    mesh = ref_mesh("quad40")
    with pytest.raises(TypeError):
        F.write_vtk(str(tmp_path / "bad"), mesh, ("U", np.zeros(mesh.ncells)))


# ---------------------------------------------------------------------------
# Forces and wall shear on the unit square (quad40: h = 1/40, unit depth)
# ---------------------------------------------------------------------------
H = 1.0 / 40


def test_pressure_force_of_a_uniform_pressure(ref_mesh):
    mesh = ref_mesh("quad40")
    p = np.full(mesh.ncells, 2.0)
    total = np.zeros(3)
    outward = {"inlet": (-1, 0), "outlet": (1, 0), "bottom": (0, -1), "top": (0, 1)}
    for name, (nx, ny) in outward.items():
        Fp = ref.pressure_force(mesh, name, p, 1.5)
        assert np.allclose(Fp, [3.0 * nx, 3.0 * ny, 0.0], atol=1e-12), (name, Fp)
        total += Fp
    assert np.allclose(total, 0.0, atol=1e-12)
    fx, fy, fz = ref.surface_flux(mesh, "inlet", p)
    assert np.allclose(fx, -2.0 * H, rtol=1e-12) and np.all(fy == 0.0) and np.all(fz == 0.0)


def test_viscous_force_and_wall_shear_of_a_uniform_slip(ref_mesh):
    """U = (u, v, 0) in every cell against a resting bottom wall: the
    wall-parallel part is (u, 0, 0), over the centre-to-face distance h/2;
    the force is nu u / (h/2) times the wall length (1), the shear nu u /
    (h/2) per face at the face centres."""
    mesh = ref_mesh("quad40")
    n = mesh.ncells
    u, v, nu = 0.3, -0.7, 1e-3
    ux, uy, uz = np.full(n, u), np.full(n, v), np.zeros(n)
    gx, gy, gz = ref.surface_normal_gradient(mesh, "bottom", ux, uy, uz, [0.0, 0.0, 0.0])
    assert np.allclose(gx, u / (H / 2), rtol=1e-12) and np.allclose(gy, 0.0, atol=1e-12) and np.all(gz == 0.0)
    Fv = ref.viscous_force(mesh, "bottom", ux, uy, uz, [0.0, 0.0, 0.0], 1.0, nu)
    assert np.allclose(Fv, [nu * u / (H / 2), 0.0, 0.0], rtol=1e-12, atol=1e-15)
    nut = np.full(n, 4e-3)
    Fv2 = ref.viscous_force(mesh, "bottom", ux, uy, uz, [0.0, 0.0, 0.0], 2.0, nu, nut)
    assert np.allclose(Fv2, [2.0 * (nu + 4e-3) * u / (H / 2), 0.0, 0.0], rtol=1e-12, atol=1e-15)
    # no velocity BC on the patch: zero
    assert np.all(ref.viscous_force(mesh, "bottom", ux, uy, uz, None, 1.0, nu) == 0.0)
    (tx, ty, tz), pos = ref.wall_shear_stress(mesh, "bottom", ux, uy, uz, [0.0, 0.0, 0.0], nu)
    assert np.allclose(tx, nu * u / (H / 2), rtol=1e-12) and np.allclose(ty, 0.0, atol=1e-15)
    f = mesh.boundary("bottom")[0]
    assert np.array_equal(pos, mesh.face_centre[f])
    # a moving lid: only the difference to the wall velocity counts
    (tx2, _, _), _ = ref.wall_shear_stress(mesh, "top", ux, uy, uz, [u, 0.0, 0.0], nu)
    assert np.allclose(tx2, 0.0, atol=1e-15)
