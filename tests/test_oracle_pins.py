"""More pins of the oracle (the CPU restatement), each by something the
restatement's own formula does not supply: exactly rounded values, closed
forms on uniform grids, hand-computed boundary contributions, sign patterns,
invariants of the mathematics, and dense brute-force constructions.  Each is
chosen so that a plausible slip (a dropped term, a wrong sign or index, a
transposed operand, a parallel instead of a sequential update) fails it.
CPU only."""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

import fvmhip as F
from fvmhip import _lib as L
from fvmhip.api import _bcs
from oracle import ref
from tests.test_oracle import csr_of, dense, pattern

EPS = np.finfo(np.float64).eps


def exact_log(x):
    getcontext().prec = 50
    return float(Decimal(x).ln())  # float(Decimal) rounds correctly


# ---------------------------------------------------------------------------
# Wall-function arithmetic (RANSModels_kOmega.jl:200-222, Discretise_0_types.jl:61-77)
# ---------------------------------------------------------------------------
def test_log_is_correctly_rounded():
    """fvh_log (the wall functions' log, shared by the oracle and the GPU)
    against the exactly rounded logarithm over [1e-300, 1e300], near 1 and
    at the edges of the double range."""
    lib = ref.lib()
    rng = np.random.default_rng(42)
    xs = np.concatenate([np.exp(rng.uniform(np.log(1e-300), np.log(1e300), 30000)), rng.uniform(0.5, 2.0, 10000),
                         1.0 + rng.uniform(-1e-7, 1e-7, 5000),
                         [5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 0.5, 2.0, 10.0, 9.8 * 11.0]])
    bad = [float(x) for x in xs if lib.ref_log(float(x)) != exact_log(float(x))]
    assert not bad, f"{len(bad)} inputs not correctly rounded, e.g. {bad[:3]}"
    # the platform log agrees to within 1 ulp (it is not always correctly rounded)
    for x in xs[:2000]:
        g, m = lib.ref_log(float(x)), math.log(float(x))
        assert abs(g - m) <= math.ulp(m)
    assert lib.ref_log(1.0) == 0.0 and math.isinf(lib.ref_log(0.0)) and math.isnan(lib.ref_log(-1.0))


def test_y_plus_laminar_is_the_ten_step_fixed_point():
    """y_plus_laminar(E, kappa): yL = 11; 10 times yL = log(max(yL E, 1))/kappa."""
    lib = ref.lib()
    for E, kappa in ((9.8, 0.41), (9.0, 0.4)):
        y = 11.0
        for _ in range(10):
            y = exact_log(max(y * E, 1.0)) / kappa
        assert lib.ref_y_plus_laminar(E, kappa) == y
    assert abs(lib.ref_y_plus_laminar(9.8, 0.41) - 11.5301) < 1e-3  # the textbook laminar/log-layer crossing


def test_nut_wall_closed_form():
    lib = ref.lib()
    nu, kappa, E = 1.5e-5, 0.41, 9.8
    for yp in (0.5, 5.0, 11.6, 30.0, 300.0, 5000.0):
        want = max(nu * (yp * kappa / exact_log(max(E * yp, 1.0 + 1e-4)) - 1.0), 0.0)
        assert lib.ref_nut_wall(nu, yp, kappa, E) == want


def test_cmu_quarter_power_is_correctly_rounded():
    """cmu^0.25 for the default cmu = 0.09: the constant the wall functions use."""
    getcontext().prec = 50
    exact = float(Decimal(0.09) ** Decimal("0.25"))
    assert ref.lib().ref_pow(0.09, 0.25) == exact
    assert abs(exact - math.sqrt(math.sqrt(0.09))) <= math.ulp(exact)


# ---------------------------------------------------------------------------
# Assembly (Discretise_1_schemes.jl:11-111, Discretise_3_boundary_conditions.jl:4-159)
# ---------------------------------------------------------------------------
class Terms:
    """fvh_term / fvh_source arrays over host arrays, for ref.assemble."""

    def __init__(self, spec, srcs=()):
        self.keep = []
        self.t = (L.Term * max(1, len(spec)))()
        for k, (op, scheme, sign, flux) in enumerate(spec):
            self.t[k].op, self.t[k].scheme, self.t[k].sign = op, scheme, sign
            if flux is not None:
                a = np.ascontiguousarray(flux, np.float64)
                self.keep.append(a)
                self.t[k].flux = a.ctypes.data
        self.n = len(spec)
        self.s = (L.Source * max(1, len(srcs)))()
        for k, (sign, field) in enumerate(srcs):
            a = np.ascontiguousarray(field, np.float64)
            self.keep.append(a)
            self.s[k].sign, self.s[k].field = sign, a.ctypes.data
        self.ns = len(srcs)


def assemble(mesh, spec, bcs=(), srcs=(), prev=None, dt=1.0, comp=0, **kw):
    T = Terms(spec, srcs)
    barr = _bcs(mesh, list(bcs), True)
    vals, b = ref.assemble(mesh, T.t, T.n, T.s, T.ns, prev, dt, barr, len(bcs), comp, **kw)
    return dense(mesh, vals), b


def interior_cells(mesh):
    nb = np.zeros(mesh.ncells, int)
    np.add.at(nb, mesh.boundary_cells, 1)
    return np.flatnonzero(nb == 0)


def test_laplacian_is_the_five_point_stencil_on_quad40(quad40):
    """Gamma = 1, sign -1 (as -Laplacian in the momentum equations): interior
    rows are (4, -1, -1, -1, -1) on the uniform 40 x 40 grid (A/delta = 1),
    the matrix is symmetric, interior rows sum to zero."""
    mesh = quad40
    A, b = assemble(mesh, [(L.OP_LAPLACIAN, F.Linear, -1, np.ones(mesh.nfaces))])
    inner = interior_cells(mesh)
    assert inner.size == 38 * 38
    for c in inner:
        row = A[c]
        nz = np.flatnonzero(row)
        assert list(nz) == [c - 40, c - 1, c, c + 1, c + 40]
        assert np.allclose(row[nz], [-1.0, -1.0, 4.0, -1.0, -1.0], rtol=1e-12, atol=0)
        assert abs(row.sum()) < 1e-12
    assert np.allclose(A, A.T, rtol=0, atol=1e-13)
    assert not b.any()
    # sign +1 (the pressure equation): negative definite (H12)
    Ap, _ = assemble(mesh, [(L.OP_LAPLACIAN, F.Linear, 1, np.ones(mesh.nfaces))])
    assert np.all(np.diag(Ap) < 0) and np.allclose(Ap, -A)


def face_endpoints(mesh):
    c = mesh.face_centre[:, :2]
    t = np.stack([-mesh.face_normal[:, 1], mesh.face_normal[:, 0]], axis=1) * (0.5 * mesh.face_area)[:, None]
    return c - t, c + t


@pytest.mark.parametrize("name", ["quad40", "trig40"])
def test_divergence_linear_rows_sum_to_the_cell_net_flux(ref_mesh, name):
    """m_f = psi(p2) - psi(p1) over each face (p1 -> p2 with the normal on its
    right) is discretely divergence free: every cell's net outflow telescopes
    to zero.  Divergence{Linear} distributes each face's flux between the two
    cells (weights w, 1 - w), so every row of the assembled matrix sums to the
    cell's net internal outflow; with psi = 0 on the boundary that is zero."""
    mesh = ref_mesh(name)
    p1, p2 = face_endpoints(mesh)
    psi = lambda p: np.sin(np.pi * p[:, 0]) * np.sin(np.pi * p[:, 1])  # noqa: E731
    mdot = psi(p2) - psi(p1)
    A, b = assemble(mesh, [(L.OP_DIVERGENCE, F.Linear, 1, mdot)])
    scale = np.abs(A).sum(axis=1)
    assert np.all(np.abs(A.sum(axis=1)) <= 1e-12 * scale + 1e-15)
    assert np.abs(A).max() > 1e-3  # the test is not vacuous
    # an upwind matrix of the same flux: row sums equal too, and an M-matrix pattern
    Au, _ = assemble(mesh, [(L.OP_DIVERGENCE, F.Upwind, 1, mdot)])
    assert np.all(np.abs(Au.sum(axis=1)) <= 1e-12 * np.abs(Au).sum(axis=1) + 1e-15)
    off = Au - np.diag(np.diag(Au))
    assert np.all(np.diag(Au) >= 0) and np.all(off <= 0)


def test_divergence_linear_weights_are_one_half_on_a_uniform_grid(quad40):
    mesh = quad40
    rng = np.random.default_rng(3)
    mdot = rng.standard_normal(mesh.nfaces)
    A, _ = assemble(mesh, [(L.OP_DIVERGENCE, F.Linear, 1, mdot)])
    c = interior_cells(mesh)[100]
    a, bnd = mesh.cell_faces_range[c]
    for fi in range(a, bnd + 1):
        f, nb, ns = mesh.cell_faces[fi], mesh.cell_neighbours[fi], mesh.cell_nsign[fi]
        assert abs(A[c, nb] - 0.5 * mdot[f] * ns) <= 1e-12 * abs(mdot[f])
    assert abs(A[c, c] - 0.5 * sum(mdot[mesh.cell_faces[fi]] * mesh.cell_nsign[fi]
                                   for fi in range(a, bnd + 1))) < 1e-12


def test_upwind_is_an_m_matrix_for_any_flux(quad40):
    mesh = quad40
    mdot = np.random.default_rng(5).standard_normal(mesh.nfaces)
    bcs = [F.Dirichlet(n, [1.0, 0, 0]) for n in mesh.boundary_names]
    A, _ = assemble(mesh, [(L.OP_DIVERGENCE, F.Upwind, 1, mdot), (L.OP_LAPLACIAN, F.Linear, -1,
                                                                   np.full(mesh.nfaces, 1e-3))], bcs)
    off = A - np.diag(np.diag(A))
    assert np.all(off <= 0) and np.all(np.diag(A) > 0)


def test_boundary_functors_on_a_three_by_three_grid():
    """Hand-computed A and b for each BC functor on the 3 x 3 unit square
    (h = 1/3: internal A/delta = 1, boundary A/delta = (1/3)/(1/6) = 2).
    Cell 0 touches inlet (x=0) and bottom (y=0); cell 6 inlet and top."""
    mesh = F.cavity(3)
    nf = mesh.nfaces
    gamma = np.full(nf, 2.0)
    bcs = [F.Dirichlet("inlet", [5.0, 0, 0]), F.Neumann("bottom", 0.0), F.Dirichlet("top", [7.0, 0, 0]),
           F.Neumann("outlet", 0.0)]
    # Laplacian, sign -1: interior ap = +Gamma A/delta = 2 per face; Dirichlet adds 4 to A_cc and 4*value to b
    A, b = assemble(mesh, [(L.OP_LAPLACIAN, F.Linear, -1, gamma)], bcs)
    want_diag = {0: 2 * 2 + 4, 1: 3 * 2, 2: 2 * 2, 3: 3 * 2 + 4, 4: 4 * 2, 6: 2 * 2 + 4 + 4, 8: 2 * 2 + 4}
    want_b = {0: 20.0, 1: 0.0, 2: 0.0, 3: 20.0, 4: 0.0, 6: 20.0 + 28.0, 8: 28.0}
    for c, v in want_diag.items():
        assert abs(A[c, c] - v) < 1e-12, (c, A[c, c], v)
        assert abs(b[c] - want_b[c]) < 1e-12, (c, b[c], want_b[c])
    assert abs(A[0, 1] + 2.0) < 1e-12 and abs(A[0, 3] + 2.0) < 1e-12
    # Divergence: internal fluxes 0, boundary fluxes chosen per patch
    mdot = np.zeros(nf)
    bf = {n: mesh.boundary(n)[0] for n in mesh.boundary_names}
    mdot[bf["inlet"]], mdot[bf["bottom"]], mdot[bf["top"]], mdot[bf["outlet"]] = -0.3, 0.2, -0.1, 0.4
    A, b = assemble(mesh, [(L.OP_DIVERGENCE, F.Linear, 1, mdot)], bcs)
    # Dirichlet: b -= mdot * value; Neumann: A_cc += mdot
    assert abs(b[0] - 1.5) < 1e-15 and abs(A[0, 0] - 0.2) < 1e-15
    assert abs(b[6] - (1.5 + 0.7)) < 1e-15 and A[6, 6] == 0.0
    assert abs(A[2, 2] - (0.2 + 0.4)) < 1e-15 and b[2] == 0.0
    mdot[bf["bottom"]] = -0.2
    A, b = assemble(mesh, [(L.OP_DIVERGENCE, F.Upwind, 1, mdot)], bcs)
    # Upwind Dirichlet: b -= ap * value (no max); Neumann: A_cc += max(ap, 0)
    assert abs(b[0] - 1.5) < 1e-15 and A[0, 0] == 0.0
    assert abs(A[2, 2] - 0.4) < 1e-15 and abs(b[6] - 2.2) < 1e-15
    # Time{Euler}: A_cc = V/dt, b = prev V/dt, no off-diagonals
    prev = np.arange(9.0)
    A, b = assemble(mesh, [(L.OP_TIME, F.Euler, 1, None)], prev=prev, dt=0.25)
    V = mesh.cell_volume
    assert np.allclose(np.diag(A), V / 0.25, rtol=1e-15) and np.allclose(A, np.diag(np.diag(A)))
    assert np.allclose(b, prev * V / 0.25, rtol=1e-15) and abs(V.sum() - 1.0) < 1e-14
    # Si: A_cc += sign * flux_c * V
    si = np.linspace(1.0, 2.0, 9)
    A, b = assemble(mesh, [(L.OP_SI, F.Linear, -1, si)])
    assert np.allclose(np.diag(A), -si * V, rtol=1e-15) and not b.any()
    # explicit source: b += sign * field * V
    A, b = assemble(mesh, [(L.OP_TIME, F.Steady, 1, None)], srcs=[(-1, si)])
    assert np.allclose(b, -si * V, rtol=1e-15)


def test_relaxation_and_reference_cell():
    """implicit_relaxation!: A_ii /= alpha, b_i += (1 - alpha) A_ii(new) phi_i;
    setReference!: b_1 += A_11 pref, A_11 doubled."""
    mesh = F.cavity(3)
    phi = np.linspace(-1.0, 1.0, 9)
    spec = [(L.OP_LAPLACIAN, F.Linear, -1, np.full(mesh.nfaces, 2.0))]
    A0, b0 = assemble(mesh, spec)
    A, b = assemble(mesh, spec, relax_field=phi, alpha=0.5)
    assert np.allclose(np.diag(A), 2.0 * np.diag(A0)) and np.allclose(b, b0 + 0.5 * np.diag(A) * phi)
    A, b = assemble(mesh, spec, pref=3.0, refcell=4)
    assert A[4, 4] == 2.0 * A0[4, 4] and b[4] == b0[4] + A0[4, 4] * 3.0


# ---------------------------------------------------------------------------
# Calculate (Calculate_0_divergence.jl:75-107, Solvers_0_functions.jl:65-73)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["quad40", "trig40"])
def test_div_of_constant_and_linear_fields(ref_mesh, name):
    mesh = ref_mesh(name)
    nf = mesh.nfaces
    d0 = ref.div(mesh, np.full(nf, 1.0), np.full(nf, -2.0), np.zeros(nf))
    assert np.abs(d0).max() < 1e-10
    d1 = ref.div(mesh, mesh.face_centre[:, 0].copy(), np.zeros(nf), np.zeros(nf))
    assert np.allclose(d1, 1.0, atol=1e-10)
    d2 = ref.div(mesh, np.zeros(nf), 3.0 * mesh.face_centre[:, 1], np.zeros(nf))
    assert np.allclose(d2, 3.0, atol=1e-10)


def test_flux_closed_form(quad40):
    mesh = quad40
    nf = mesh.nfaces
    m = ref.flux(mesh, np.full(nf, 1.0), np.full(nf, 2.0), np.full(nf, 5.0))
    vertical = np.abs(mesh.face_normal[:, 0]) > 0.5
    h = 1.0 / 40
    assert np.allclose(np.abs(m[vertical]), h, rtol=1e-12)
    assert np.allclose(np.abs(m[~vertical]), 2 * h, rtol=1e-12)
    assert np.allclose(np.sign(m[vertical]), np.sign(mesh.face_normal[vertical, 0]))


def test_interpolation_weights_and_boundary_correction(quad40):
    """interpolate!: w phi_owner + (1 - w) phi_nbr with w = 1/2 inside a
    uniform grid, w = 1 on boundary faces; correct_boundaries!: Dirichlet
    faces take the value, Neumann faces the owner cell's value."""
    mesh = quad40
    phi = np.random.default_rng(9).standard_normal(mesh.ncells)
    pf = ref.interpolate(mesh, phi)
    o1, o2 = mesh.face_owners[:, 0], mesh.face_owners[:, 1]
    inner = o1 != o2
    mag = np.abs(phi[o1[inner]]) + np.abs(phi[o2[inner]])
    assert np.all(np.abs(pf[inner] - 0.5 * (phi[o1[inner]] + phi[o2[inner]])) <= 1e-13 * mag)  # w = 1/2 (to 1e-14)
    assert np.array_equal(pf[~inner], phi[o1[~inner]])
    bcs = [F.Dirichlet("inlet", [2.5, 0, 0]), F.Neumann("top", 0.0)]
    barr = _bcs(mesh, bcs, True)
    pc = ref.correct_boundaries(mesh, barr, 2, 0, phi, np.full(mesh.nfaces, -9.0))
    fi = mesh.boundary("inlet")[0]
    ft, ct = mesh.boundary("top")
    assert np.all(pc[fi] == 2.5) and np.array_equal(pc[ft], phi[ct])
    assert np.all(pc[mesh.boundary("outlet")[0]] == -9.0)


# ---------------------------------------------------------------------------
# Non-orthogonal correction (Calculate_3_orthogonality_correction.jl,
# Calculate_2_interpolation.jl:152-176)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["quad40", "trig40", "backwardFacingStep_10mm"])
def test_gradient_interpolation_is_exact_for_a_linear_field(ref_mesh, name):
    """phi = a x + b y with its exact gradient in every cell: the weighted
    average is that gradient and the correction along e vanishes, since
    (phi2 - phi1)/delta = g . e for e = (c2 - c1)/delta; boundary faces are
    left as they were."""
    mesh = ref_mesh(name)
    a, b = 2.5, -1.25
    cc = mesh.cell_centre
    phi = a * cc[:, 0] + b * cc[:, 1]
    n = mesh.ncells
    sentinel = np.full(mesh.nfaces, 7.0)
    fx, fy, fz = ref.interpolate_gradient(mesh, np.full(n, a), np.full(n, b), np.zeros(n), phi, sentinel, sentinel,
                                          sentinel)
    inner = mesh.face_owners[:, 0] != mesh.face_owners[:, 1]
    scale = np.abs(cc).max() / mesh.face_delta[inner].min()
    tol = 64 * EPS * scale * max(abs(a), abs(b))
    assert np.abs(fx[inner] - a).max() <= tol and np.abs(fy[inner] - b).max() <= tol
    assert np.abs(fz[inner]).max() <= tol
    assert np.all(fx[~inner] == 7.0) and np.all(fy[~inner] == 7.0) and np.all(fz[~inner] == 7.0)


def test_gradient_interpolation_weight_is_the_projection(ref_mesh):
    """Two cell gradients g1, g2 and phi = 0: the face value is
    w g1 + (1 - w) g2 minus its component along e, with w the distance from
    c2 to the face centre's projection on the centre line over |c2 - c1|
    (1/2 on a uniform grid)."""
    mesh = ref_mesh("trig40")
    n = mesh.ncells
    rng = np.random.default_rng(4)
    gx, gy = rng.standard_normal(n), rng.standard_normal(n)
    fx, fy, _ = ref.interpolate_gradient(mesh, gx, gy, np.zeros(n), np.zeros(n))
    o1, o2 = mesh.face_owners[:, 0], mesh.face_owners[:, 1]
    inner = np.nonzero(o1 != o2)[0]
    c1, c2, cf = mesh.cell_centre[o1[inner]], mesh.cell_centre[o2[inner]], mesh.face_centre[inner]
    d = c2 - c1
    t = np.einsum("ij,ij->i", cf - c1, d) / np.einsum("ij,ij->i", d, d)
    w = 1.0 - t  # |c2 - f'| / |c2 - c1| with f' = c1 + t d, 0 <= t <= 1 here
    assert np.all((t > 0) & (t < 1))
    e = mesh.face_e[inner]
    ax = w * gx[o1[inner]] + (1 - w) * gx[o2[inner]]
    ay = w * gy[o1[inner]] + (1 - w) * gy[o2[inner]]
    s = ax * e[:, 0] + ay * e[:, 1]
    assert np.allclose(fx[inner], ax - s * e[:, 0], atol=1e-12)
    assert np.allclose(fy[inner], ay - s * e[:, 1], atol=1e-12)
    q = ref_mesh("quad40")
    g = np.ones(q.ncells)
    qx, _, _ = ref.interpolate_gradient(q, np.where(np.arange(q.ncells) % 2 == 0, 1.0, 3.0), g * 0, g * 0, g * 0)
    qi = q.face_owners[:, 0] != q.face_owners[:, 1]
    vertical = np.abs(q.face_normal[:, 0]) < 0.5
    # faces between x-neighbours alternate 1 and 3 -> mean 2, then the x
    # component is removed along e = x; faces between y-neighbours keep 2 or
    # the cell value pair mean
    assert np.allclose(qx[qi & vertical], 2.0, atol=1e-12)


def test_nonorthogonal_flux_vanishes_on_orthogonal_faces(ref_mesh):
    """(n - e) A: zero on a uniform quad grid (n = e); on triangles it is
    the face gradient times the tangential offset, here checked in closed
    form for a constant face gradient; boundary faces are not written."""
    q = ref_mesh("quad40")
    nf = q.nfaces
    out = ref.nonorthogonal_flux(q, np.full(nf, 1.5), np.full(nf, -2.0), np.zeros(nf), np.full(nf, 4.0))
    inner = q.face_owners[:, 0] != q.face_owners[:, 1]
    # quad40's node coordinates carry rounding (scale 0.001), so n - e is
    # zero to the last bits only: |g| A 1e-12 at most
    assert np.abs(out[inner]).max() <= 2.5 * q.face_area.max() * 1e-12 and np.all(out[~inner] == 4.0)
    t = ref_mesh("trig40")
    nf = t.nfaces
    out = ref.nonorthogonal_flux(t, np.full(nf, 1.5), np.full(nf, -2.0), np.zeros(nf), np.full(nf, 4.0))
    inner = t.face_owners[:, 0] != t.face_owners[:, 1]
    T = (t.face_normal - t.face_e) * t.face_area[:, None]
    want = 1.5 * T[:, 0] - 2.0 * T[:, 1]
    assert np.allclose(out[inner], want[inner], atol=1e-15)
    assert np.abs(out[inner]).max() > 1e-6  # the triangles are not orthogonal
    assert np.all(out[~inner] == 4.0)


def test_correct_moves_flux_from_owner_to_neighbour(ref_mesh):
    """correct!: each internal face adds c = -sign J corr to its owner and
    -c to its neighbour, so the sum over cells is unchanged (to rounding);
    one face alone changes exactly its two cells; a boundary face adds +c
    then -c to its cell."""
    mesh = ref_mesh("trig40")
    nf, n = mesh.nfaces, mesh.ncells
    rng = np.random.default_rng(2)
    J, corr, b = rng.uniform(0.5, 2.0, nf), rng.standard_normal(nf), rng.standard_normal(n)
    out = ref.nonorthogonal_correct(mesh, 1, J, corr, b)
    assert abs(out.sum() - b.sum()) < 1e-12 * np.abs(b).sum()
    f = mesh.nbfaces + 17
    one = np.zeros(nf)
    one[f] = 0.5
    o1, o2 = mesh.face_owners[f]
    out = ref.nonorthogonal_correct(mesh, -1, J, one, np.zeros(n))
    want = np.zeros(n)
    want[o1], want[o2] = 0.5 * J[f], -0.5 * J[f]
    assert np.array_equal(out, want)
    fb = 3
    one = np.zeros(nf)
    one[fb] = 1e-3
    c = mesh.face_owners[fb, 0]
    b0 = np.zeros(n)
    b0[c] = 1.0
    out = ref.nonorthogonal_correct(mesh, 1, J, one, b0)
    k = -J[fb] * 1e-3
    assert out[c] == (1.0 + k) + (-k) and np.all(np.delete(out, c) == 0.0)


# ---------------------------------------------------------------------------
# k-omega (RANSModels_functions.jl:5-88, RANSModels_kOmega.jl:114-277)
# ---------------------------------------------------------------------------
def test_production_is_twice_the_strain_rate_norm_for_a_linear_field(quad40):
    """U = (a x + b y, c x - a y, 0): grad U is exact on interior cells of a
    uniform grid, S = sym(grad U), and 2 S:S = 4 a^2 + (b + c)^2."""
    mesh = quad40
    a, b, c = 0.7, -1.3, 0.4
    x, y = mesh.cell_centre[:, 0], mesh.cell_centre[:, 1]
    ux, uy, uz = a * x + b * y, c * x - a * y, np.zeros(mesh.ncells)
    bcs = [F.Neumann(n, 0.0) for n in mesh.boundary_names]
    barr = _bcs(mesh, bcs, True)
    pk = ref.strain_production(mesh, F.Orthogonal, barr, len(bcs), ux, uy, uz)
    inner = interior_cells(mesh)
    assert np.allclose(pk[inner], 4 * a * a + (b + c) ** 2, rtol=1e-10)
    # a pure rotation has no strain: b = -c
    pk0 = ref.strain_production(mesh, F.Orthogonal, barr, len(bcs), -0.5 * y, 0.5 * x, uz)
    assert np.abs(pk0[inner]).max() < 1e-20


def test_bound_is_sequential():
    """bound!(phi, eps): a negative cell takes the mean of max(phi_nbr, eps)
    over its neighbours, seeing neighbours already bounded earlier in the
    same pass (cell order).  Cells 5 and 6 (adjacent) and 9 (above 5) are
    negative on a 4 x 4 grid."""
    mesh = F.cavity(4)
    k = np.ones(16)
    k[[5, 6]], k[9] = -1.0, -2.0
    out = ref.bound(mesh, k, EPS)

    def mean_of(c, field):
        a, bnd = mesh.cell_faces_range[c]
        s = 0.0
        for fi in range(a, bnd + 1):
            s += max(field[mesh.cell_neighbours[fi]], EPS)
        return s / (bnd - a + 1)

    seq = k.copy()
    for c in range(16):
        if seq[c] < 0:
            seq[c] = max(seq[c], mean_of(c, seq))
    assert np.array_equal(out, seq)
    assert abs(out[5] - 0.5) < 1e-15 and abs(out[6] - 0.875) < 1e-15 and abs(out[9] - 0.875) < 1e-15
    # a simultaneous (parallel) update would give 0.75 for cell 6
    assert abs(mean_of(6, k) - 0.75) < 1e-15
    # positive cells are untouched; the floor applies
    assert np.all(out[k > 0] == 1.0)
    assert np.all(ref.bound(mesh, np.full(16, 1e-30), 1e-20) == 1e-20)


def test_omega_wall_constraint_doubles_per_wall_face():
    """constrain_equation!: b += A_cc omega_w; A_cc *= 2 per OmegaWallFunction
    face, so the corner cell between two wall patches is hit twice (H16)."""
    mesh = F.cavity(3)
    rp, ci = pattern(mesh)
    rows = np.repeat(np.arange(9), np.diff(rp))
    vals = np.where(rows == ci, 3.0, -0.5)
    kf = np.full(9, 0.01)
    nu = 1e-5
    bcs = [F.OmegaWallFunction("inlet"), F.OmegaWallFunction("bottom")]
    barr = _bcs(mesh, bcs, True)
    v, b = ref.omega_constrain(mesh, barr, 2, kf, nu, vals, np.zeros(9))
    A = dense(mesh, v)
    kappa, beta1, cmu, E = 0.41, 0.075, 0.09, 9.8
    ylam = 11.0
    for _ in range(10):
        ylam = exact_log(max(ylam * E, 1.0)) / kappa

    def omega_w(c, patch):
        faces, cells = mesh.boundary(patch)
        f = faces[list(cells).index(c)]
        y = mesh.face_delta[f]
        yplus = cmu ** 0.25 * y * math.sqrt(kf[c]) / nu
        return math.sqrt(kf[c]) / (cmu ** 0.25 * kappa * y) if yplus > ylam else 6 * nu / (beta1 * y * y)

    assert A[0, 0] == 12.0 and A[3, 3] == 6.0 and A[1, 1] == 6.0 and A[4, 4] == 3.0
    assert math.isclose(b[0], 3.0 * omega_w(0, "inlet") + 6.0 * omega_w(0, "bottom"), rel_tol=1e-14)
    assert math.isclose(b[3], 3.0 * omega_w(3, "inlet"), rel_tol=1e-14)
    assert b[4] == 0.0 and A[0, 1] == -0.5


# ---------------------------------------------------------------------------
# ILU(0) and the Krylov recurrences
# ---------------------------------------------------------------------------
def ilu0_dense(A, mask):
    """Textbook ILU(0) by its definition: (L U)_ij = A_ij on the pattern,
    computed by dense IKJ elimination restricted to the pattern."""
    n = A.shape[0]
    W = A.copy()
    for i in range(1, n):
        for k in range(i):
            if not mask[i, k]:
                continue
            W[i, k] = W[i, k] / W[k, k]
            for j in range(k + 1, n):
                if mask[i, j]:
                    W[i, j] -= W[i, k] * W[k, j]
    return np.tril(W, -1) + np.eye(n), np.triu(W)


@pytest.mark.parametrize("mesh_fn", [lambda: F.channel(6, 5), lambda: F.delaunay_unit_square(30, seed=4)])
def test_ilu0_matches_the_definition_on_a_2d_pattern(mesh_fn):
    mesh = mesh_fn()
    vals = csr_of(mesh, seed=11, diag=5.0)
    A = dense(mesh, vals)
    mask = A != 0
    Ld, Ud = ilu0_dense(A, mask)
    assert np.allclose((Ld @ Ud)[mask], A[mask], rtol=1e-12, atol=1e-13)
    n = mesh.ncells
    Minv = np.stack([ref.precond_apply_csr(mesh, vals, L.PRECOND_ILU0, e) for e in np.eye(n)], axis=1)
    M = np.linalg.inv(Minv)
    assert np.allclose(M, Ld @ Ud, rtol=1e-9, atol=1e-10)
    assert np.allclose(M[mask], A[mask], rtol=1e-9, atol=1e-10)
    assert not np.allclose(M, A, rtol=1e-6)  # fill-in dropped: not the exact LU


def _cfg(solver, itmax, precond=L.PRECOND_DILU):
    return L.SolverCfg(solver=solver, precond=precond, itmax=itmax, precision=0, convergence=0, relax=1,
                       atol=0.0, rtol=0.0)


def test_cg_residuals_are_m_orthogonal():
    """Preconditioned CG: z_i . r_j = 0 for i != j (z = M^-1 r), with the
    symmetric DILU preconditioner of a symmetric positive definite matrix."""
    mesh = F.cavity(12)
    vals = csr_of(mesh, seed=13, diag=2.6, sym=True)
    A = dense(mesh, vals)
    n = mesh.ncells
    b = np.random.default_rng(14).standard_normal(n)
    rs, zs = [], []
    for k in range(0, 7):
        x, st, _ = ref.solve_csr(mesh, vals, b, _cfg(L.CG, max(k, 1)), np.zeros(n)) if k else (np.zeros(n), None,
                                                                                                None)
        r = b - A @ x
        rs.append(r)
        zs.append(ref.precond_apply_csr(mesh, vals, L.PRECOND_DILU, r))
    floor = 1e-13 * np.linalg.norm(zs[0]) * np.linalg.norm(rs[0])  # rounding of the first residuals
    assert np.linalg.norm(rs[-1]) > 1e-6 * np.linalg.norm(rs[0])  # not yet converged: the test has teeth
    for i in range(len(rs)):
        for j in range(len(rs)):
            if i != j:
                assert abs(zs[i] @ rs[j]) <= 1e-9 * np.linalg.norm(zs[i]) * np.linalg.norm(rs[j]) + floor, (i, j)


def test_gmres_tracked_residual_is_the_true_preconditioned_residual():
    """Left-preconditioned GMRES reports |zeta| = ||M^-1 (b - A x_k)||."""
    mesh = F.channel(12, 9)
    vals = csr_of(mesh, seed=15, diag=2.4)
    A = dense(mesh, vals)
    n = mesh.ncells
    b = np.random.default_rng(16).standard_normal(n)
    x0 = np.random.default_rng(17).standard_normal(n) * 0.1
    prev = None
    floor = 1e-13 * np.linalg.norm(ref.precond_apply_csr(mesh, vals, L.PRECOND_DILU, b))
    for k in (1, 2, 4, 7, 10):
        x, st, _ = ref.solve_csr(mesh, vals, b, _cfg(L.GMRES, k), x0)
        true = np.linalg.norm(ref.precond_apply_csr(mesh, vals, L.PRECOND_DILU, b - A @ x))
        assert st["niter"] == k
        assert abs(st["rnorm"] - true) <= 1e-8 * true + floor, (k, st["rnorm"], true)
        if prev is not None:
            assert st["rnorm"] <= prev * (1 + 1e-12)  # minimal residual: non-increasing
        prev = st["rnorm"]


def test_bicgstab_omega_step_minimises_the_residual():
    """One BiCGSTAB iteration: x_1 = x_0 + alpha p + omega s with omega the
    minimiser of ||s - omega t|| in the preconditioned residual, so
    r_1 = M^-1 (b - A x_1) is orthogonal to t = M^-1 A s."""
    mesh = F.channel(9, 7)
    vals = csr_of(mesh, seed=18, diag=3.5)
    A = dense(mesh, vals)
    n = mesh.ncells
    b = np.random.default_rng(19).standard_normal(n)
    Minv = lambda v: ref.precond_apply_csr(mesh, vals, L.PRECOND_DILU, v)  # noqa: E731
    x1, st, _ = ref.solve_csr(mesh, vals, b, _cfg(L.BICGSTAB, 1), np.zeros(n))
    r0 = Minv(b)
    v = Minv(A @ r0)
    alpha = (b @ r0) / (b @ v)
    s = r0 - alpha * v
    t = Minv(A @ s)
    r1 = Minv(b - A @ x1)
    assert abs(r1 @ t) <= 1e-10 * np.linalg.norm(r1) * np.linalg.norm(t)
    assert math.isclose(st["rnorm"], np.linalg.norm(r1), rel_tol=1e-10)
