"""Mesh2 on the host: UNV ingest and the synthetic meshes of the benchmark cases.

Every mesh goes through the same C++ builder (`fvh_hmesh_build`), which applies
the UNV2 rules of the reference (src/UNV2/UNV2_3_builder.jl:3-287,
src/UNV2/UNV2_2_geometry.jl:20-165): boundary faces first in group order, then
cell edges in first-appearance order, owner < neighbour, node-mean centres.
A generated grid is therefore numbered exactly as the reference would number
the same grid read from a UNV file with the same element order.
"""
import ctypes as C

import numpy as np

from . import _lib as L

P_D_ = L.P_D
P_I64_ = L.P_I64


class Mesh2:
    """Zero-based Mesh2 arrays (src/Mesh/Mesh_0_types.jl:82-93), numpy."""

    def __init__(self, h):
        d = L.MeshDesc()
        L.check(L.lib().fvh_hmesh_desc(h, C.byref(d)), "fvh_hmesh_desc")
        n, nf, nb = d.ncells, d.nfaces, d.nboundaries
        ncf, nbf = d.ncellfaces, d.nboundaryfaces

        def arr(ptr, count, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

        self.ncells, self.nfaces = n, nf
        self.cell_centre = arr(d.cell_centre, 3 * n, np.float64).reshape(n, 3)
        self.cell_volume = arr(d.cell_volume, n, np.float64)
        self.cell_faces_range = arr(d.cell_faces_range, 2 * n, np.int64).reshape(n, 2)
        self.cell_faces = arr(d.cell_faces, ncf, np.int64)
        self.cell_neighbours = arr(d.cell_neighbours, ncf, np.int64)
        self.cell_nsign = arr(d.cell_nsign, ncf, np.int64)
        self.face_owners = arr(d.face_owners, 2 * nf, np.int64).reshape(nf, 2)
        self.face_centre = arr(d.face_centre, 3 * nf, np.float64).reshape(nf, 3)
        self.face_normal = arr(d.face_normal, 3 * nf, np.float64).reshape(nf, 3)
        self.face_e = arr(d.face_e, 3 * nf, np.float64).reshape(nf, 3)
        self.face_area = arr(d.face_area, nf, np.float64)
        self.face_delta = arr(d.face_delta, nf, np.float64)
        self.face_weight = arr(d.face_weight, nf, np.float64)
        self.boundary_range = arr(d.boundary_range, 2 * nb, np.int64).reshape(nb, 2)
        self.boundary_faces = arr(d.boundary_faces, nbf, np.int64)
        self.boundary_cells = arr(d.boundary_cells, nbf, np.int64)
        npt = C.c_int64()
        xyz, cnp, cn = P_D_(), P_I64_(), P_I64_()
        L.check(L.lib().fvh_hmesh_nodes(h, C.byref(npt), C.byref(xyz), C.byref(cnp), C.byref(cn)), "fvh_hmesh_nodes")
        self.nodes = arr(xyz, 3 * npt.value, np.float64).reshape(-1, 3)  # mesh.nodes coords (scaled)
        self.cell_nodes_ptr = arr(cnp, n + 1, np.int64)
        self.cell_nodes = arr(cn, int(self.cell_nodes_ptr[-1]) if n else 0, np.int64)
        buf = C.create_string_buffer(256)
        self.boundary_names = []
        for i in range(nb):
            L.check(L.lib().fvh_hmesh_boundary_name(h, i, buf, 256))
            self.boundary_names.append(buf.value.decode())
        self._dev = None
        self._keep = None

    # ---- accessors mirroring the reference ----
    @property
    def nboundaries(self):
        return len(self.boundary_names)

    @property
    def nbfaces(self):
        return int(self.boundary_faces.size)

    def boundary_index(self, name):
        """boundary_index(boundaries, name) (src/Mesh/Mesh_1_functions.jl:20-30), zero based."""
        name = str(name).lstrip(":")
        try:
            return self.boundary_names.index(name)
        except ValueError:
            raise KeyError(f"no boundary named {name!r}; have {self.boundary_names}") from None

    def boundary(self, name):
        b = self.boundary_index(name)
        s, e = self.boundary_range[b]
        return self.boundary_faces[s:e + 1], self.boundary_cells[s:e + 1]

    def desc(self):
        """fvh_mesh_desc over these arrays (zero based); keeps them alive."""
        keep = [np.ascontiguousarray(a) for a in (
            self.cell_centre.reshape(-1), self.cell_volume, self.cell_faces_range.reshape(-1), self.cell_faces,
            self.cell_neighbours, self.cell_nsign, self.face_owners.reshape(-1), self.face_centre.reshape(-1),
            self.face_normal.reshape(-1), self.face_e.reshape(-1), self.face_area, self.face_delta,
            self.face_weight, self.boundary_range.reshape(-1), self.boundary_faces, self.boundary_cells)]
        d = L.MeshDesc()
        d.index_base = 0
        d.ncells, d.nfaces, d.nboundaries = self.ncells, self.nfaces, self.nboundaries
        d.ncellfaces, d.nboundaryfaces = self.cell_faces.size, self.boundary_faces.size
        names = ["cell_centre", "cell_volume", "cell_faces_range", "cell_faces", "cell_neighbours", "cell_nsign",
                 "face_owners", "face_centre", "face_normal", "face_e", "face_area", "face_delta", "face_weight",
                 "boundary_range", "boundary_faces", "boundary_cells"]
        for nm, a in zip(names, keep):
            setattr(d, nm, L.dptr(a) if a.dtype == np.float64 else L.iptr(a))
        self._keep = keep
        return d

    def device(self):
        """The device mesh (fvh_mesh), created on first use."""
        if self._dev is None:
            self._dev = DeviceMesh(self)
        return self._dev


class DeviceMesh:
    def __init__(self, mesh):
        self.mesh = mesh
        d = mesh.desc()
        h = C.c_void_p()
        L.check(L.lib().fvh_mesh_create(C.byref(d), C.byref(h)), "fvh_mesh_create")
        self.h = h
        off = getattr(mesh, "case_offsets", None)
        if off is not None:
            off = np.ascontiguousarray(off, np.int64)
            L.check(L.lib().fvh_mesh_set_cases(self.h, len(off) - 1, off.ctypes.data_as(L.P_I64)),
                    "fvh_mesh_set_cases")

    def tuning(self):
        """Grid rows per lane of the banded sweeps (0: level-scheduled)."""
        r = C.c_int32()
        L.check(L.lib().fvh_mesh_tuning(self.h, C.byref(r)))
        return r.value

    def case_info(self):
        k, b = C.c_int32(), C.c_int32()
        L.check(L.lib().fvh_mesh_case_info(self.h, C.byref(k), C.byref(b)))
        return k.value, bool(b.value)

    def pattern(self):
        nnz = C.c_int64()
        L.check(L.lib().fvh_mesh_pattern(self.h, C.byref(nnz), None, None))
        rp = np.zeros(self.mesh.ncells + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        L.check(L.lib().fvh_mesh_pattern(self.h, None, rp.ctypes.data_as(L.P_I32), ci.ctypes.data_as(L.P_I32)))
        return rp, ci

    def sweep_info(self):
        a, b = C.c_int64(), C.c_int64()
        L.check(L.lib().fvh_mesh_sweep_info(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def __del__(self):
        try:
            if self.h:
                L.lib().fvh_mesh_destroy(self.h)
        except Exception:
            pass


def _from_hmesh(h):
    try:
        return Mesh2(h)
    finally:
        L.lib().fvh_hmesh_free(h)


def build_mesh(path, scale=1.0):
    """build_mesh(file; scale) + update_mesh_format (src/UNV2/UNV2_3_builder.jl:3)."""
    h = C.c_void_p()
    L.check(L.lib().fvh_hmesh_read_unv(str(path).encode(), float(scale), C.byref(h)), f"reading {path}")
    return _from_hmesh(h)


def from_elements(points, elements, groups, scale=1.0):
    """Mesh from UNV-like data: points (npt,3); elements = (ptr, verts) with
    1-based node ids, edge elements first; groups = [(name, 1-based labels)]."""
    pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    eptr, everts = (np.ascontiguousarray(a, np.int64) for a in elements)
    gptr = np.zeros(len(groups) + 1, np.int64)
    gel = []
    for i, (_, labels) in enumerate(groups):
        gel.append(np.asarray(labels, np.int64))
        gptr[i + 1] = gptr[i] + len(labels)
    gel = np.ascontiguousarray(np.concatenate(gel) if gel else np.zeros(0, np.int64))
    names = b"".join(n.encode() + b"\0" for n, _ in groups)
    h = C.c_void_p()
    L.check(L.lib().fvh_hmesh_build(pts.shape[0], L.dptr(pts.reshape(-1)), eptr.size - 1, L.iptr(eptr), L.iptr(everts),
                                    len(groups), L.iptr(gptr), L.iptr(gel), names, float(scale), C.byref(h)),
            "fvh_hmesh_build")
    return _from_hmesh(h)


# ---------------------------------------------------------------------------
# Synthetic structured grids
# ---------------------------------------------------------------------------
def graded(n, length, ratio, symmetric=False):
    """n cell widths over `length`, geometric `ratio` from the start (or from both ends)."""
    if ratio == 1.0:
        w = np.full(n, length / n)
    elif symmetric:
        h = n // 2
        half = graded(h, length * h / n if n % 2 else length / 2, ratio)
        if n % 2:
            mid = np.array([half[-1] * ratio])
            w = np.concatenate([half, mid, half[::-1]])
            w *= length / w.sum()
        else:
            w = np.concatenate([half, half[::-1]])
    else:
        w = ratio ** np.arange(n, dtype=np.float64)
        w *= length / w.sum()
    return w


def rect_grid(xs, ys, patches, z=0.0):
    """Quad grid on node coordinates xs (nx+1), ys (ny+1), cells numbered
    i-fastest, counter-clockwise vertices.  `patches` = list of
    (name, side, predicate) with side in {"left","right","bottom","top"} and
    predicate(mid_coord) selecting the edges of that side that belong to it."""
    xs = np.asarray(xs, np.float64)
    ys = np.asarray(ys, np.float64)
    nx, ny = xs.size - 1, ys.size - 1
    X, Y = np.meshgrid(xs, ys)  # (ny+1, nx+1), row j
    pts = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], axis=1)
    nid = lambda i, j: j * (nx + 1) + i + 1  # noqa: E731  (1-based)
    edges = []
    groups = []
    for name, side, pred in patches:
        labels = []
        if side in ("bottom", "top"):
            j = 0 if side == "bottom" else ny
            for i in range(nx):
                if pred(0.5 * (xs[i] + xs[i + 1])):
                    edges.append((nid(i, j), nid(i + 1, j)))
                    labels.append(len(edges))
        else:
            i = 0 if side == "left" else nx
            for j in range(ny):
                if pred(0.5 * (ys[j] + ys[j + 1])):
                    edges.append((nid(i, j), nid(i, j + 1)))
                    labels.append(len(edges))
        groups.append((name, labels))
    nedge = len(edges)
    if nedge != 2 * (nx + ny):
        raise ValueError(f"patches cover {nedge} of {2 * (nx + ny)} boundary edges")
    I, J = np.meshgrid(np.arange(nx), np.arange(ny))
    I, J = I.ravel(), J.ravel()
    base = J * (nx + 1) + I + 1
    quads = np.stack([base, base + 1, base + 1 + (nx + 1), base + (nx + 1)], axis=1)
    everts = np.concatenate([np.asarray(edges, np.int64).ravel(), quads.ravel()])
    eptr = np.concatenate([np.arange(0, 2 * nedge + 1, 2, dtype=np.int64),
                           2 * nedge + 4 * np.arange(1, nx * ny + 1, dtype=np.int64)])
    return pts, (eptr, everts), groups


def cavity(n=40, length=1.0):
    """Unit lid-driven cavity n x n (patches inlet x=0, outlet x=L, bottom, top = lid)."""
    xs = np.linspace(0.0, length, n + 1)
    pts, el, groups = rect_grid(xs, xs, [("inlet", "left", lambda s: True), ("outlet", "right", lambda s: True),
                                         ("bottom", "bottom", lambda s: True), ("top", "top", lambda s: True)])
    return from_elements(pts, el, groups)


def backward_facing_step(nx=2000, ny=500, h=0.01, length=30.0, grading=1.02):
    """Laminar BFS: the channel downstream of a step of height h: x in
    [0, length*h], y in [0, 3h]; inlet = x=0, y>h (2h tall upstream channel),
    step = x=0, y<h (wall); wall-normal grading symmetric about mid-height."""
    xs = np.concatenate([[0.0], np.cumsum(graded(nx, length * h, 1.0))])
    ys = np.concatenate([[0.0], np.cumsum(graded(ny, 3.0 * h, grading, symmetric=True))])
    xs[-1], ys[-1] = length * h, 3.0 * h
    pts, el, groups = rect_grid(xs, ys, [
        ("inlet", "left", lambda s: s > h), ("step", "left", lambda s: s <= h),
        ("outlet", "right", lambda s: True), ("bottom", "bottom", lambda s: True), ("top", "top", lambda s: True)])
    return from_elements(pts, el, groups)


def channel(nx, ny, lx=1.0, ly=1.0, grading=1.0):
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.concatenate([[0.0], np.cumsum(graded(ny, ly, grading, symmetric=True))])
    ys[-1] = ly
    pts, el, groups = rect_grid(xs, ys, [("inlet", "left", lambda s: True), ("outlet", "right", lambda s: True),
                                         ("bottom", "bottom", lambda s: True), ("top", "top", lambda s: True)])
    return from_elements(pts, el, groups)


def flat_plate(nx=1000, ny=500, lx=2.0, ly=0.5, first=4.0e-5, ratio=1.05):
    """Flat plate: wall at y=0 (x>=0), first cell `first`, geometric growth
    `ratio` from the wall until the spacing reaches the uniform remainder."""
    w = [first]
    while True:
        rem_cells = ny - len(w)
        rem = ly - sum(w)
        if rem_cells <= 0:
            break
        if w[-1] * ratio >= rem / rem_cells:
            w += [rem / rem_cells] * rem_cells
            break
        w.append(w[-1] * ratio)
    ys = np.concatenate([[0.0], np.cumsum(w)])
    ys[-1] = ly
    xs = np.linspace(0.0, lx, nx + 1)
    pts, el, groups = rect_grid(xs, ys, [("inlet", "left", lambda s: True), ("outlet", "right", lambda s: True),
                                         ("wall", "bottom", lambda s: True), ("top", "top", lambda s: True)])
    return from_elements(pts, el, groups)


def delaunay_unit_square(npts, seed=0, nboundary=None):
    """Delaunay triangulation of seeded uniform points in [0,1]^2 plus evenly
    spaced boundary points; triangles ordered by the Morton code of their
    centroids; patches left/right/bottom/top."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    nb = nboundary or max(4, int(np.sqrt(npts)))
    t = np.linspace(0.0, 1.0, nb + 1)[:-1]
    bnd = np.concatenate([np.stack([t, np.zeros_like(t)], 1), np.stack([np.ones_like(t), t], 1),
                          np.stack([1.0 - t, np.ones_like(t)], 1), np.stack([np.zeros_like(t), 1.0 - t], 1)])
    inner = rng.uniform(0.0, 1.0, size=(npts, 2)) * (1 - 2e-6) + 1e-6
    P = np.concatenate([bnd, inner])
    tri = Delaunay(P).simplices.astype(np.int64)
    # counter-clockwise
    a, b, c = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]
    cr = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    tri[cr < 0] = tri[cr < 0][:, [0, 2, 1]]
    cen = P[tri].mean(axis=1)
    q = np.clip((cen * (1 << 20)).astype(np.uint64), 0, (1 << 20) - 1)

    def spread(v):
        v = v & np.uint64(0xFFFFF)
        out = np.zeros_like(v)
        for bit in range(20):
            out |= ((v >> np.uint64(bit)) & np.uint64(1)) << np.uint64(2 * bit)
        return out

    order = np.argsort(spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)), kind="stable")
    tri = tri[order]
    # boundary edges from the hull
    edges = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    key = np.sort(edges, axis=1)
    uniq, cnt = np.unique(key, axis=0, return_counts=True)
    bedges = uniq[cnt == 1]
    mid = P[bedges].mean(axis=1)
    side = np.where(mid[:, 0] < 1e-12, 0, np.where(mid[:, 0] > 1 - 1e-12, 1, np.where(mid[:, 1] < 1e-12, 2, 3)))
    groups, edge_list = [], []
    for s, name in enumerate(["left", "right", "bottom", "top"]):
        sel = bedges[side == s]
        coord = mid[side == s][:, 1 if s < 2 else 0]
        sel = sel[np.argsort(coord, kind="stable")]
        labels = list(range(len(edge_list) + 1, len(edge_list) + 1 + len(sel)))
        edge_list.extend(sel.tolist())
        groups.append((name, labels))
    ne = len(edge_list)
    everts = np.concatenate([np.asarray(edge_list, np.int64).ravel() + 1, tri.ravel() + 1])
    eptr = np.concatenate([np.arange(0, 2 * ne + 1, 2, dtype=np.int64), 2 * ne + 3 * np.arange(1, tri.shape[0] + 1)])
    pts = np.concatenate([P, np.zeros((P.shape[0], 1))], axis=1)
    return from_elements(pts, (eptr, everts), groups)


def union(meshes, families):
    """Block-diagonal union of independent meshes, for the batched mode.

    Cells and internal faces of mesh k follow those of mesh k-1.  Boundary
    patches are keyed by (families[k], patch name): meshes of one family must
    have the same patch names in the same order, and their faces of a patch
    are listed mesh after mesh.  The union keeps boundary faces first in patch
    order (as the builder numbers them), so every cell sees its faces, its
    neighbours and its boundary faces in the order it has alone.  The result
    carries `case_offsets` (cell ranges) and `case_face_ranges`."""
    assert len(meshes) == len(families) and meshes
    order = []  # union patches: (family, name)
    fam_names = {}
    for m, fam in zip(meshes, families):
        if fam in fam_names:
            if fam_names[fam] != m.boundary_names:
                raise ValueError(f"meshes of family {fam!r} differ in their patches")
        else:
            fam_names[fam] = list(m.boundary_names)
            order.extend((fam, nm) for nm in m.boundary_names)
    K = len(meshes)
    coff = np.concatenate([[0], np.cumsum([m.ncells for m in meshes])]).astype(np.int64)
    nbt = sum(m.nbfaces for m in meshes)
    nint = [m.nfaces - m.nbfaces for m in meshes]
    ioff = np.concatenate([[0], np.cumsum(nint)]).astype(np.int64)
    # new ids of every mesh's faces
    fmap = [np.full(m.nfaces, -1, np.int64) for m in meshes]
    bfaces, bcells, brange = [], [], []
    nxt = 0
    for fam, nm in order:
        first = nxt
        for k, m in enumerate(meshes):
            if families[k] != fam:
                continue
            f, c = m.boundary(nm)
            fmap[k][f] = np.arange(nxt, nxt + f.size)
            bfaces.append(np.arange(nxt, nxt + f.size))
            bcells.append(c + coff[k])
            nxt += f.size
        brange.append((first, nxt - 1))
    assert nxt == nbt
    for k, m in enumerate(meshes):
        fmap[k][m.nbfaces:] = nbt + ioff[k] + np.arange(nint[k])
    nf = nbt + int(ioff[-1])
    u = Mesh2.__new__(Mesh2)
    u.ncells, u.nfaces = int(coff[-1]), nf
    u.cell_centre = np.concatenate([m.cell_centre for m in meshes])
    u.cell_volume = np.concatenate([m.cell_volume for m in meshes])
    cfo = np.concatenate([[0], np.cumsum([m.cell_faces.size for m in meshes])]).astype(np.int64)
    u.cell_faces_range = np.concatenate([m.cell_faces_range + cfo[k] for k, m in enumerate(meshes)])
    u.cell_faces = np.concatenate([fmap[k][m.cell_faces] for k, m in enumerate(meshes)])
    u.cell_neighbours = np.concatenate([m.cell_neighbours + coff[k] for k, m in enumerate(meshes)])
    u.cell_nsign = np.concatenate([m.cell_nsign for m in meshes])

    def faces(attr, shape_tail):
        out = np.zeros((nf,) + shape_tail, getattr(meshes[0], attr).dtype)
        for k, m in enumerate(meshes):
            out[fmap[k]] = getattr(m, attr)
        return out

    u.face_owners = faces("face_owners", (2,))
    for k, m in enumerate(meshes):
        u.face_owners[fmap[k]] += coff[k]
    u.face_centre = faces("face_centre", (3,))
    u.face_normal = faces("face_normal", (3,))
    u.face_e = faces("face_e", (3,))
    u.face_area = faces("face_area", ())
    u.face_delta = faces("face_delta", ())
    u.face_weight = faces("face_weight", ())
    u.boundary_range = np.array(brange, np.int64).reshape(-1, 2)
    u.boundary_faces = np.concatenate(bfaces).astype(np.int64)
    u.boundary_cells = np.concatenate(bcells).astype(np.int64)
    u.boundary_names = [f"{fam}:{nm}" for fam, nm in order]
    noff = np.concatenate([[0], np.cumsum([m.nodes.shape[0] for m in meshes])]).astype(np.int64)
    u.nodes = np.concatenate([m.nodes for m in meshes])
    u.cell_nodes = np.concatenate([m.cell_nodes + noff[k] for k, m in enumerate(meshes)])
    u.cell_nodes_ptr = np.concatenate([[0], np.cumsum(np.concatenate([np.diff(m.cell_nodes_ptr) for m in meshes]))])
    u._dev = None
    u._keep = None
    u.case_offsets = coff
    return u

