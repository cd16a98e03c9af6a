"""Mesh ingest and geometry (UNV2 builder rules) against closed forms and
invariants.  CPU only."""
import numpy as np
import pytest

import fvmhip as F


def closed_surface_defect(m):
    """|sum over faces of S_f * nsign| per cell (interior + boundary faces)."""
    S = m.face_area[:, None] * m.face_normal
    acc = np.zeros((m.ncells, 3))
    for c in range(m.ncells):
        a, b = m.cell_faces_range[c]
        for k in range(a, b + 1):
            acc[c] += S[m.cell_faces[k]] * m.cell_nsign[k]
    nb = m.nbfaces
    np.add.at(acc, m.face_owners[:nb, 0], S[:nb])
    return np.abs(acc).max()


def test_quad40_topology(quad40):
    """unv_sample_meshes/quad40.unv: 40x40 quads on the unit square (x0.001)."""
    m = quad40
    assert m.ncells == 1600
    assert m.nfaces == 3280 and m.nbfaces == 160
    assert m.cell_faces.size == 2 * 3120  # each internal face seen from both cells
    assert m.ncells + m.cell_faces.size == 7840  # nnz of Equation(mesh)
    assert m.boundary_names == ["inlet", "outlet", "bottom", "top"]
    assert [len(m.boundary(n)[0]) for n in m.boundary_names] == [40, 40, 40, 40]


def test_quad40_geometry_closed_form(quad40):
    m, h = quad40, 1.0 / 40
    nb = m.nbfaces
    assert np.allclose(m.cell_volume, h * h, rtol=1e-12)
    assert abs(m.cell_volume.sum() - 1.0) < 1e-12
    assert np.allclose(m.face_area, h, rtol=1e-12)
    assert np.allclose(m.face_delta[nb:], h, rtol=1e-12)  # internal: centre-to-centre
    assert np.allclose(m.face_delta[:nb], h / 2, rtol=1e-12)  # boundary: centre-to-face (H3)
    assert np.allclose(m.face_weight[nb:], 0.5, rtol=1e-12)
    assert np.all(m.face_weight[:nb] == 1.0)
    assert np.allclose(np.linalg.norm(m.face_normal, axis=1), 1.0, rtol=1e-14)
    # axis-aligned normals
    assert np.all(np.isclose(np.abs(m.face_normal[:, :2]).max(axis=1), 1.0))


def test_face_order_and_ownership(quad40):
    """H1: boundary faces first (owner == neighbour), owner < neighbour inside;
    normals point from owner to neighbour; nsign +1 for the owner."""
    m = quad40
    nb = m.nbfaces
    own = m.face_owners
    assert np.all(own[:nb, 0] == own[:nb, 1])
    assert np.all(own[nb:, 0] < own[nb:, 1])
    d = m.cell_centre[own[nb:, 1]] - m.cell_centre[own[nb:, 0]]
    assert np.all((d * m.face_normal[nb:]).sum(1) > 0)
    for c in (0, 17, 800, 1599):
        a, b = m.cell_faces_range[c]
        faces = m.cell_faces[a:b + 1]
        assert np.all(np.diff(faces) > 0)  # ascending face id
        for k in range(a, b + 1):
            f = m.cell_faces[k]
            assert m.cell_nsign[k] == (1 if own[f, 0] == c else -1)
            assert m.cell_neighbours[k] == (own[f, 1] if own[f, 0] == c else own[f, 0])


def test_closed_surface_invariant(quad40, trig40_path):
    assert closed_surface_defect(quad40) < 1e-16
    t = F.build_mesh(trig40_path, scale=0.001)
    assert closed_surface_defect(t) < 1e-15
    assert abs(t.cell_volume.sum() - 1.0) < 1e-12
    assert np.all(t.cell_volume > 0)


def test_generated_cavity_matches_unv_geometry(quad40):
    g = F.cavity(40)
    assert (g.ncells, g.nfaces, g.nbfaces) == (quad40.ncells, quad40.nfaces, quad40.nbfaces)
    for a in ("cell_volume", "face_area", "face_delta", "face_weight"):
        assert np.allclose(np.sort(getattr(g, a)), np.sort(getattr(quad40, a)), rtol=1e-12), a
    assert closed_surface_defect(g) < 1e-16


def test_trig40_cells_have_three_faces(trig40_path):
    t = F.build_mesh(trig40_path, scale=0.001)
    nint = np.diff(t.cell_faces_range, axis=1).ravel() + 1
    nbnd = np.bincount(t.face_owners[:t.nbfaces, 0], minlength=t.ncells)
    assert np.all(nint + nbnd == 3)


def test_bfs_grading_and_patches():
    m = F.backward_facing_step(nx=200, ny=50)
    assert m.ncells == 10000 and m.boundary_names == ["inlet", "step", "outlet", "bottom", "top"]
    assert abs(m.cell_volume.sum() - 0.3 * 0.03) < 1e-15
    hy = m.cell_volume[:50 * 200:200] / (0.3 / 200)  # first column, bottom to top
    ratios = hy[1:25] / hy[:24]
    assert np.allclose(ratios, 1.02, rtol=1e-9)  # geometric from the wall
    assert np.allclose(hy, hy[::-1], rtol=1e-9)  # symmetric about mid-height
    yi = m.face_centre[m.boundary("inlet")[0], 1]
    ys = m.face_centre[m.boundary("step")[0], 1]
    assert yi.min() > 0.01 and ys.max() < 0.01
    assert len(m.boundary("outlet")[0]) == 50 and len(m.boundary("top")[0]) == 200


def test_delaunay_mesh():
    m = F.delaunay_unit_square(2000, seed=3)
    assert abs(m.cell_volume.sum() - 1.0) < 1e-12
    assert np.all(m.cell_volume > 0)
    assert closed_surface_defect(m) < 1e-15
    nb = 4 * int(np.sqrt(2000))
    npts = 2000 + nb
    assert m.ncells == 2 * npts - 2 - nb  # Euler: triangles of a triangulation with nb hull points


def test_flat_plate_first_cell():
    m = F.flat_plate(nx=20, ny=40, first=1e-3, ratio=1.1)
    hy = m.cell_volume[::20] / (2.0 / 20)
    assert abs(hy[0] - 1e-3) < 1e-12
    assert np.allclose(hy[1:4] / hy[:3], 1.1)
    assert abs(hy.sum() - 0.5) < 1e-12


def test_builder_rejects_inconsistent_input():
    pts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    eptr = np.array([0, 2, 4, 6, 8, 12])
    ev = np.array([1, 2, 2, 3, 3, 4, 4, 4, 1, 2, 3, 4])  # degenerate edge 4-4
    with pytest.raises(F.FVHError):
        F.from_elements(pts, (eptr, ev), [("all", [1, 2, 3, 4])])


def test_union_keeps_every_case_intact():
    """Batched mode: the union of independent meshes keeps each case's cells,
    geometry and closed surfaces; patches merge by (family, name) with
    boundary faces first."""
    ms = [F.cavity(8), F.channel(6, 4), F.cavity(5)]
    u = F.union(ms, ["a", "b", "a"])
    assert u.ncells == sum(m.ncells for m in ms)
    assert list(u.case_offsets) == [0, 64, 88, 113]
    assert u.boundary_names == ["a:inlet", "a:outlet", "a:bottom", "a:top", "b:inlet", "b:outlet", "b:bottom", "b:top"]
    assert closed_surface_defect(u) < 1e-15
    nb = u.nbfaces
    assert nb == sum(m.nbfaces for m in ms)
    assert np.all(u.face_owners[:nb, 0] == u.face_owners[:nb, 1])
    assert np.all(u.face_owners[nb:, 0] < u.face_owners[nb:, 1])
    for k, m in enumerate(ms):
        a, b = u.case_offsets[k], u.case_offsets[k + 1]
        assert np.array_equal(u.cell_volume[a:b], m.cell_volume)
        assert np.array_equal(u.cell_centre[a:b], m.cell_centre)
        # each cell's faces: same count, ascending ids, neighbours in the case
        rng = u.cell_faces_range[a:b]
        assert np.array_equal(np.diff(rng, axis=1), np.diff(m.cell_faces_range, axis=1))
        for c in range(0, b - a, 7):
            s, e = rng[c]
            assert np.all(np.diff(u.cell_faces[s:e + 1]) > 0)
            nbrs = u.cell_neighbours[s:e + 1]
            assert np.all((nbrs >= a) & (nbrs < b))
    top_a = u.boundary("a:top")[1]
    assert top_a.size == 8 + 5 and np.all(top_a[:8] < 64) and np.all(top_a[8:] >= 88)
