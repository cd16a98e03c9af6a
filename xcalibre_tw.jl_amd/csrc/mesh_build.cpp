// Host-side mesh construction: SALOME UNV reading and the UNV2 builder rules
// that fix face order, ownership and geometry of a 2-D Mesh2.
//
// The rules restated here (all citations relative to the XCALibre_TW.jl root):
//   * UNV datasets 2411/2412/2467 parsing ............ src/UNV2/UNV2_1_loader.jl:3-120
//   * boundary faces first (group order, element order), then cell edges in
//     first-appearance order after unique! ............. src/UNV2/UNV2_3_builder.jl:93-163
//   * owner < neighbour from ascending node->cell lists  src/UNV2/UNV2_3_builder.jl:200-236
//   * boundary face / cell ids by node-pair match ...... src/UNV2/UNV2_3_builder.jl:238-266
//   * centres = node means ............................. src/UNV2/UNV2_3_builder.jl:270-287,
//                                                        src/UNV2/UNV2_2_geometry.jl:11-18
//   * face normals, delta, e, weight; cell volumes ..... src/UNV2/UNV2_2_geometry.jl:20-164
//   * flattening to Mesh2 .............................. src/Mesh/Mesh2D/Mesh2D_0_update_mesh.jl:5-142
// The arithmetic follows the reference operation by operation (StaticArrays
// dot = (a1*b1 + a2*b2) + a3*b3, norm = sqrt of the same sum, cross product
// by its component formula) so that geometry is reproducible.
#include "fvh_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace fvh {

namespace {

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 divs(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 muls(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt((a.x * a.x + a.y * a.y) + a.z * a.z); }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct Element {
    int64_t index;
    int64_t vcount;
    std::vector<int64_t> verts;
};

struct Group {
    std::string name;
    std::vector<int64_t> elems;
};

}  // namespace

// Builds the Mesh2 arrays (zero based) from points, elements and boundary
// groups exactly as build_mesh + update_mesh_format do.
static int build_hmesh(std::vector<V3> &pts, std::vector<Element> &elements,
                       std::vector<Group> &groups, double scale, HMesh &out) {
    if (scale != 1.0)
        for (auto &p : pts) p = {p.x * scale, p.y * scale, p.z * scale};
    const int64_t npts = (int64_t)pts.size();
    const int64_t nel = (int64_t)elements.size();
    int64_t bfaces = 0;
    for (auto &g : groups) bfaces += (int64_t)g.elems.size();
    // first_2d_element
    int64_t first = -1;
    for (int64_t e = 0; e < nel; ++e)
        if (elements[e].vcount > 2) { first = e; break; }
    if (first < 0) return set_error(FVH_ERR_MESH, "mesh has no 2-D elements");
    const int64_t ncells = nel - first;
    for (int64_t e = 0; e < nel; ++e)
        for (int64_t v : elements[e].verts)
            if (v < 1 || v > npts) return set_error(FVH_ERR_MESH, "element %lld references node %lld", (long long)(e + 1), (long long)v);
    for (auto &g : groups)
        for (int64_t el : g.elems)
            if (el < 1 || el > nel || elements[el - 1].verts.size() < 2)
                return set_error(FVH_ERR_MESH, "boundary group %s references element %lld", g.name.c_str(), (long long)el);

    // generate_nodes: node -> cells (ascending cell ids)
    std::vector<std::vector<int64_t>> node_cells(npts);
    for (int64_t e = first; e < nel; ++e)
        for (int64_t v : elements[e].verts) node_cells[v - 1].push_back(e - first);

    // generate_faces: boundary elements first, then cell edges; unique! keeps first
    std::vector<std::pair<int64_t, int64_t>> faces;
    faces.reserve((size_t)(bfaces + 2 * ncells + 16));
    std::unordered_map<uint64_t, int64_t> seen;
    seen.reserve((size_t)(3 * ncells + bfaces));
    auto key = [](int64_t a, int64_t b) { return ((uint64_t)a << 32) ^ (uint64_t)b; };
    auto push_face = [&](int64_t v1, int64_t v2) -> int {
        if (v1 == v2) return set_error(FVH_ERR_MESH, "Boundary elements are inconsistent: possible mesh corruption");
        int64_t a = v1 < v2 ? v1 : v2, b = v1 < v2 ? v2 : v1;
        uint64_t k = key(a, b);
        if (seen.find(k) == seen.end()) {
            seen.emplace(k, (int64_t)faces.size());
            faces.push_back({a, b});
        }
        return FVH_OK;
    };
    for (auto &g : groups)
        for (int64_t el : g.elems) {
            const auto &v = elements[el - 1].verts;
            int rc = push_face(v[0], v[1]);
            if (rc) return rc;
        }
    const int64_t nbface_unique = (int64_t)faces.size();
    if (nbface_unique != bfaces)
        return set_error(FVH_ERR_MESH, "boundary groups share edges (%lld listed, %lld distinct)", (long long)bfaces, (long long)nbface_unique);
    for (int64_t e = first; e < nel; ++e) {
        const auto &v = elements[e].verts;
        const size_t nv = v.size();
        for (size_t vi = 0; vi < nv; ++vi) {
            int rc = push_face(v[vi], v[(vi + 1) % nv]);
            if (rc) return rc;
        }
    }
    const int64_t nfaces = (int64_t)faces.size();

    // face_cell_connectivity!
    std::vector<int64_t> owner(2 * nfaces);
    std::vector<std::vector<int64_t>> cfaces(ncells), cnbrs(ncells);
    for (int64_t f = 0; f < nfaces; ++f) {
        int64_t oc[2] = {-1, -1};
        int cnt = 0;
        for (int64_t c1 : node_cells[faces[f].first - 1])
            for (int64_t c2 : node_cells[faces[f].second - 1])
                if (c1 == c2) {
                    if (cnt >= 2) return set_error(FVH_ERR_MESH, "face %lld shared by more than two cells", (long long)(f + 1));
                    oc[cnt++] = c1;
                }
        if (cnt == 0) return set_error(FVH_ERR_MESH, "face %lld has no owner cell", (long long)(f + 1));
        if (cnt == 2) {
            cfaces[oc[0]].push_back(f);
            cfaces[oc[1]].push_back(f);
            cnbrs[oc[0]].push_back(oc[1]);
            cnbrs[oc[1]].push_back(oc[0]);
            owner[2 * f] = oc[0];
            owner[2 * f + 1] = oc[1];
        } else {
            owner[2 * f] = oc[0];
            owner[2 * f + 1] = oc[0];
        }
    }
    for (int64_t f = bfaces; f < nfaces; ++f)
        if (owner[2 * f] == owner[2 * f + 1])
            return set_error(FVH_ERR_MESH, "edge %lld is on the boundary but in no boundary group", (long long)(f + 1));

    // boundary_connectivity!
    std::vector<int64_t> bptr(groups.size() + 1, 0), bfaces_id, bcells_id;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        for (int64_t el : groups[gi].elems) {
            auto &v = elements[el - 1].verts;  // sort! mutates the element (only its first two are used)
            int64_t a = v[0] < v[1] ? v[0] : v[1], b = v[0] < v[1] ? v[1] : v[0];
            auto it = seen.find(key(a, b));
            if (it == seen.end() || it->second >= bfaces) return set_error(FVH_ERR_MESH, "boundary edge not found");
            bfaces_id.push_back(it->second);
            bcells_id.push_back(owner[2 * it->second]);
        }
        bptr[gi + 1] = (int64_t)bfaces_id.size();
    }

    // geometry: centres
    std::vector<V3> ccen(ncells), fcen(nfaces);
    for (int64_t c = 0; c < ncells; ++c) {
        const auto &v = elements[first + c].verts;
        V3 s{0.0, 0.0, 0.0};
        for (int64_t id : v) s = add(s, pts[id - 1]);
        ccen[c] = divs(s, (double)v.size());
    }
    for (int64_t f = 0; f < nfaces; ++f) {
        V3 s{0.0, 0.0, 0.0};
        s = add(s, pts[faces[f].first - 1]);
        s = add(s, pts[faces[f].second - 1]);
        fcen[f] = divs(s, 2.0);
    }
    std::vector<V3> fnorm(nfaces), fe(nfaces);
    std::vector<double> farea(nfaces), fdelta(nfaces), fweight(nfaces);
    const V3 kz{0.0, 0.0, 1.0};
    const int64_t nb = (int64_t)bfaces_id.size();
    for (int64_t f = nb; f < nfaces; ++f) {  // internal_face_properties!
        V3 p1 = pts[faces[f].first - 1], p2 = pts[faces[f].second - 1];
        V3 t = sub(p2, p1);
        double area = norm(t);
        V3 c1 = ccen[owner[2 * f]], c2 = ccen[owner[2 * f + 1]], cf = fcen[f];
        V3 d1f = sub(cf, c1), df2 = sub(c2, cf), d12 = sub(c2, c1);
        V3 n = cross(divs(t, area), kz);
        if (dot(d12, n) < 0.0) n = muls(-1.0, n);
        double delta = norm(d12);
        fe[f] = divs(d12, delta);
        fweight[f] = std::fabs(dot(d1f, n) / (dot(d1f, n) + dot(df2, n)));
        farea[f] = area;
        fnorm[f] = n;
        fdelta[f] = delta;
    }
    for (size_t gi = 0; gi < groups.size(); ++gi)  // boundary_face_properties!
        for (int64_t i = bptr[gi]; i < bptr[gi + 1]; ++i) {
            int64_t f = bfaces_id[i];
            V3 p1 = pts[faces[f].first - 1], p2 = pts[faces[f].second - 1];
            V3 t = sub(p2, p1);
            double area = norm(t);
            V3 n = cross(divs(t, area), kz);
            V3 dcf = sub(fcen[f], ccen[owner[2 * f]]);
            if (dot(dcf, n) < 0.0) n = muls(-1.0, n);
            double delta = norm(dcf);
            fe[f] = divs(dcf, delta);
            farea[f] = area;
            fnorm[f] = n;
            fdelta[f] = delta;
            fweight[f] = 1.0;
        }
    // cell_properties!: nsign and volume from internal faces
    std::vector<double> vol(ncells, 0.0);
    std::vector<std::vector<int64_t>> cns(ncells);
    for (int64_t c = 0; c < ncells; ++c) {
        double v = 0.0;
        bool any = false;
        for (int64_t f : cfaces[c]) {
            V3 dcf = sub(fcen[f], ccen[c]);
            int64_t s = dot(dcf, fnorm[f]) > 0.0 ? 1 : -1;
            cns[c].push_back(s);
            v += (dot(dcf, fnorm[f]) * (double)s) * farea[f];
            any = true;
        }
        if (any) vol[c] = 0.5 * v;
    }
    // correct_boundary_cell_volumes!
    for (size_t gi = 0; gi < groups.size(); ++gi)
        for (int64_t i = bptr[gi]; i < bptr[gi + 1]; ++i) {
            int64_t c = bcells_id[i], f = bfaces_id[i];
            V3 dcf = sub(fcen[f], ccen[c]);
            vol[c] = vol[c] + 0.5 * dot(dcf, fnorm[f]) * farea[f];
        }

    // flatten (update_mesh_format)
    out = HMesh();
    out.ncells = ncells;
    out.nfaces = nfaces;
    out.cell_centre.resize(3 * ncells);
    for (int64_t c = 0; c < ncells; ++c) {
        out.cell_centre[3 * c] = ccen[c].x;
        out.cell_centre[3 * c + 1] = ccen[c].y;
        out.cell_centre[3 * c + 2] = ccen[c].z;
    }
    out.cell_volume = vol;
    out.cell_faces_range.resize(2 * ncells);
    int64_t fi = 0;
    for (int64_t c = 0; c < ncells; ++c) {
        out.cell_faces_range[2 * c] = fi;
        for (size_t j = 0; j < cfaces[c].size(); ++j, ++fi) {
            out.cell_faces.push_back(cfaces[c][j]);
            out.cell_neighbours.push_back(cnbrs[c][j]);
            out.cell_nsign.push_back(cns[c][j]);
        }
        out.cell_faces_range[2 * c + 1] = fi - 1;
    }
    out.face_owners = owner;
    out.face_centre.resize(3 * nfaces);
    out.face_normal.resize(3 * nfaces);
    out.face_e.resize(3 * nfaces);
    for (int64_t f = 0; f < nfaces; ++f) {
        out.face_centre[3 * f] = fcen[f].x; out.face_centre[3 * f + 1] = fcen[f].y; out.face_centre[3 * f + 2] = fcen[f].z;
        out.face_normal[3 * f] = fnorm[f].x; out.face_normal[3 * f + 1] = fnorm[f].y; out.face_normal[3 * f + 2] = fnorm[f].z;
        out.face_e[3 * f] = fe[f].x; out.face_e[3 * f + 1] = fe[f].y; out.face_e[3 * f + 2] = fe[f].z;
    }
    out.face_area = farea;
    out.face_delta = fdelta;
    out.face_weight = fweight;
    out.boundary_range.resize(2 * groups.size());
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        out.boundary_range[2 * gi] = bptr[gi];
        out.boundary_range[2 * gi + 1] = bptr[gi + 1] - 1;
        out.names.push_back(groups[gi].name);
    }
    out.boundary_faces = bfaces_id;
    out.boundary_cells = bcells_id;
    // nodes (scaled) and each cell's nodes in element order, for output
    out.node_xyz.resize(3 * (size_t)npts);
    for (int64_t i = 0; i < npts; ++i) {
        out.node_xyz[3 * i] = pts[i].x;
        out.node_xyz[3 * i + 1] = pts[i].y;
        out.node_xyz[3 * i + 2] = pts[i].z;
    }
    out.cell_nodes_ptr.assign(1, 0);
    for (int64_t e = first; e < nel; ++e) {
        for (int64_t v : elements[e].verts) out.cell_nodes.push_back(v - 1);
        out.cell_nodes_ptr.push_back((int64_t)out.cell_nodes.size());
    }
    return FVH_OK;
}

static bool is_int(const std::string &s, long long *v) {
    if (s.empty()) return false;
    char *end = nullptr;
    long long x = std::strtoll(s.c_str(), &end, 10);
    if (*end != '\0') return false;
    if (v) *v = x;
    return true;
}

// UNV2.load: the loader's own line classification rules.
static int read_unv(const char *path, std::vector<V3> &pts, std::vector<Element> &els,
                    std::vector<Group> &groups) {
    std::ifstream in(path);
    if (!in) return set_error(FVH_ERR_IO, "cannot open %s", path);
    bool block = false, d2411 = false, d2412 = false, d2467 = false;
    int64_t index = 0, vcount = 0;
    std::string line;
    std::vector<std::string> tok;
    while (std::getline(in, line)) {
        tok.clear();
        {
            std::istringstream ss(line);
            std::string t;
            while (ss >> t) tok.push_back(t);
        }
        if (tok.empty()) continue;
        if (!block && tok[0] == "-1") { block = true; continue; }
        if (block && tok[0] == "-1") { block = d2411 = d2412 = d2467 = false; continue; }
        if (block && !d2411 && !d2412 && !d2467) {
            if (tok[0] == "2411") { d2411 = true; continue; }
            if (tok[0] == "2412") { d2412 = true; continue; }
            if (tok[0] == "2467") { d2467 = true; continue; }
        }
        if (d2411) {
            if (tok[0].find('.') != std::string::npos && tok.size() >= 3) {
                pts.push_back({std::strtod(tok[0].c_str(), nullptr), std::strtod(tok[1].c_str(), nullptr),
                               std::strtod(tok[2].c_str(), nullptr)});
                continue;
            }
        }
        if (d2412) {
            long long v0 = 0;
            if (is_int(tok[0], &v0) && v0 == 0) continue;
            long long last = 0;
            bool lastint = is_int(tok.back(), &last);
            if (tok.size() == 6 && lastint && (last == 2 || last == 3 || last == 4)) {
                vcount = last;
                index = v0;
                continue;
            }
            if (tok.size() == 2 || tok.size() == 3 || tok.size() == 4) {
                Element e;
                e.index = index;
                e.vcount = vcount;
                for (auto &t : tok) e.verts.push_back(std::strtoll(t.c_str(), nullptr, 10));
                els.push_back(e);
                continue;
            }
        }
        if (d2467) {
            long long a = 0, b = 1;
            if (is_int(tok[0], &a) && tok.size() > 1 && is_int(tok[1], &b) && b == 0) {
                groups.push_back(Group{"default", {}});
                continue;
            }
            if (!is_int(tok[0], nullptr)) {
                if (groups.empty()) return set_error(FVH_ERR_IO, "group name before group header");
                groups.back().name = tok[0];
                continue;
            }
            if (groups.empty() || tok.size() < 2) return set_error(FVH_ERR_IO, "malformed dataset 2467");
            groups.back().elems.push_back(std::strtoll(tok[1].c_str(), nullptr, 10));
            if (tok.size() >= 6) groups.back().elems.push_back(std::strtoll(tok[5].c_str(), nullptr, 10));
        }
    }
    return FVH_OK;
}

}  // namespace fvh

using namespace fvh;

extern "C" int fvh_hmesh_build(int64_t npoints, const double *points_xyz, int64_t nelements,
                               const int64_t *elem_ptr, const int64_t *elem_verts, int64_t ngroups,
                               const int64_t *group_ptr, const int64_t *group_elems,
                               const char *group_names, double scale, fvh_hmesh **out) {
    FVH_GUARD_BEGIN
    if (!out || npoints <= 0 || nelements <= 0 || !points_xyz || !elem_ptr || !elem_verts || ngroups < 0)
        return set_error(FVH_ERR_ARG, "fvh_hmesh_build: bad arguments");
    std::vector<V3> pts((size_t)npoints);
    for (int64_t i = 0; i < npoints; ++i) pts[i] = {points_xyz[3 * i], points_xyz[3 * i + 1], points_xyz[3 * i + 2]};
    std::vector<Element> els((size_t)nelements);
    for (int64_t e = 0; e < nelements; ++e) {
        int64_t a = elem_ptr[e], b = elem_ptr[e + 1];
        if (b - a < 2 || b - a > 4) return set_error(FVH_ERR_MESH, "element %lld has %lld vertices", (long long)(e + 1), (long long)(b - a));
        els[e].index = e + 1;
        els[e].vcount = b - a;
        els[e].verts.assign(elem_verts + a, elem_verts + b);
    }
    std::vector<Group> groups((size_t)ngroups);
    const char *nm = group_names;
    for (int64_t g = 0; g < ngroups; ++g) {
        groups[g].name = nm ? std::string(nm) : ("group" + std::to_string(g + 1));
        if (nm) nm += std::strlen(nm) + 1;
        groups[g].elems.assign(group_elems + group_ptr[g], group_elems + group_ptr[g + 1]);
    }
    auto *h = new fvh_hmesh();
    int rc = build_hmesh(pts, els, groups, scale, h->m);
    if (rc) { delete h; return rc; }
    *out = h;
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_hmesh_read_unv(const char *path, double scale, fvh_hmesh **out) {
    FVH_GUARD_BEGIN
    if (!path || !out) return set_error(FVH_ERR_ARG, "fvh_hmesh_read_unv: bad arguments");
    std::vector<V3> pts;
    std::vector<Element> els;
    std::vector<Group> groups;
    int rc = read_unv(path, pts, els, groups);
    if (rc) return rc;
    // element labels are used as positions by the reference: check they agree
    for (size_t i = 0; i < els.size(); ++i)
        if (els[i].index != (int64_t)i + 1)
            return set_error(FVH_ERR_MESH, "element label %lld at position %zu", (long long)els[i].index, i + 1);
    auto *h = new fvh_hmesh();
    rc = build_hmesh(pts, els, groups, scale, h->m);
    if (rc) { delete h; return rc; }
    *out = h;
    return FVH_OK;
    FVH_GUARD_END
}

extern "C" int fvh_hmesh_desc(const fvh_hmesh *h, fvh_mesh_desc *d) {
    if (!h || !d) return set_error(FVH_ERR_ARG, "fvh_hmesh_desc: null");
    const HMesh &m = h->m;
    std::memset(d, 0, sizeof(*d));
    d->index_base = 0;
    d->ncells = m.ncells;
    d->nfaces = m.nfaces;
    d->nboundaries = (int64_t)m.names.size();
    d->ncellfaces = (int64_t)m.cell_faces.size();
    d->nboundaryfaces = (int64_t)m.boundary_faces.size();
    d->cell_centre = m.cell_centre.data();
    d->cell_volume = m.cell_volume.data();
    d->cell_faces_range = m.cell_faces_range.data();
    d->cell_faces = m.cell_faces.data();
    d->cell_neighbours = m.cell_neighbours.data();
    d->cell_nsign = m.cell_nsign.data();
    d->face_owners = m.face_owners.data();
    d->face_centre = m.face_centre.data();
    d->face_normal = m.face_normal.data();
    d->face_e = m.face_e.data();
    d->face_area = m.face_area.data();
    d->face_delta = m.face_delta.data();
    d->face_weight = m.face_weight.data();
    d->boundary_range = m.boundary_range.data();
    d->boundary_faces = m.boundary_faces.data();
    d->boundary_cells = m.boundary_cells.data();
    return FVH_OK;
}

extern "C" int fvh_hmesh_boundary_name(const fvh_hmesh *h, int64_t i, char *buf, int64_t len) {
    if (!h || !buf || len <= 0 || i < 0 || i >= (int64_t)h->m.names.size())
        return set_error(FVH_ERR_ARG, "fvh_hmesh_boundary_name: bad arguments");
    std::snprintf(buf, (size_t)len, "%s", h->m.names[i].c_str());
    return FVH_OK;
}

extern "C" int fvh_hmesh_nodes(const fvh_hmesh *h, int64_t *npoints, const double **xyz, const int64_t **cell_nodes_ptr,
                               const int64_t **cell_nodes) {
    if (!h || !npoints || !xyz || !cell_nodes_ptr || !cell_nodes) return set_error(FVH_ERR_ARG, "fvh_hmesh_nodes: null");
    *npoints = (int64_t)(h->m.node_xyz.size() / 3);
    *xyz = h->m.node_xyz.data();
    *cell_nodes_ptr = h->m.cell_nodes_ptr.data();
    *cell_nodes = h->m.cell_nodes.data();
    return FVH_OK;
}

extern "C" void fvh_hmesh_free(fvh_hmesh *h) { delete h; }
