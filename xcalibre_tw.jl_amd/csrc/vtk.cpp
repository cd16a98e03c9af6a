// VTK legacy ASCII writer of cell data, in the reference's file layout
// (src/VTK/VTK_writer.jl:22-116): header, POINTS, CELLS (a count, then two
// spaces and the zero-based node ids), CELL_TYPES (5 triangle, 9 quad,
// 7 polygon), CELL_DATA with SCALARS (LOOKUP_TABLE CellColors) and VECTORS.
//
// Numbers are printed as Julia's print(::Float64) prints them: the shortest
// digits that read back to the same double, in decimal notation when the
// decimal point falls within -4 < point <= 6 digits of the first digit
// ("0.0001", "123456.0"), otherwise as d.ddde<exp> ("1.0e-5", "1.0e6").
#include "fvh_internal.h"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

namespace fvh {

// Writes v into buf (at least 32 bytes) as Julia prints it; returns length.
int julia_double(double v, char *buf) {
    char *p = buf;
    if (std::isnan(v)) return std::snprintf(buf, 32, "NaN");
    if (std::signbit(v)) *p++ = '-';
    const double a = std::fabs(v);
    if (std::isinf(a)) return (int)(p - buf) + std::snprintf(p, 8, "Inf");
    if (a == 0.0) return (int)(p - buf) + std::snprintf(p, 8, "0.0");
    // shortest round-trip digits: d[.ddd]e<E>
    char sci[40];
    auto r = std::to_chars(sci, sci + sizeof sci, a, std::chars_format::scientific);
    *r.ptr = '\0';
    char digits[24];
    int nd = 0;
    const char *q = sci;
    for (; *q && *q != 'e'; ++q)
        if (*q != '.') digits[nd++] = *q;
    const int E = std::atoi(q + 1);
    const int pt = E + 1;  // position of the decimal point after the first digit group
    if (pt > -4 && pt <= 6) {
        if (pt <= 0) {
            *p++ = '0';
            *p++ = '.';
            for (int i = 0; i < -pt; ++i) *p++ = '0';
            for (int i = 0; i < nd; ++i) *p++ = digits[i];
        } else if (pt < nd) {
            for (int i = 0; i < pt; ++i) *p++ = digits[i];
            *p++ = '.';
            for (int i = pt; i < nd; ++i) *p++ = digits[i];
        } else {
            for (int i = 0; i < nd; ++i) *p++ = digits[i];
            for (int i = nd; i < pt; ++i) *p++ = '0';
            *p++ = '.';
            *p++ = '0';
        }
        *p = '\0';
        return (int)(p - buf);
    }
    *p++ = digits[0];
    *p++ = '.';
    if (nd > 1)
        for (int i = 1; i < nd; ++i) *p++ = digits[i];
    else
        *p++ = '0';
    return (int)(p - buf) + std::snprintf(p, 12, "e%d", E);
}

}  // namespace fvh

using namespace fvh;

extern "C" int fvh_format_julia(double v, char *buf, int64_t len) {
    if (!buf || len < 32) return set_error(FVH_ERR_ARG, "fvh_format_julia: buffer of at least 32 bytes");
    julia_double(v, buf);
    return FVH_OK;
}

extern "C" int fvh_write_vtk(const char *path, int64_t npoints, const double *xyz, int64_t ncells,
                             const int64_t *cell_nodes_ptr, const int64_t *cell_nodes, int32_t nfields,
                             const char *const *names, const int32_t *ncomp, const double *const *data) {
    FVH_GUARD_BEGIN
    if (!path || npoints < 0 || ncells < 0 || (npoints && !xyz) || (ncells && (!cell_nodes_ptr || !cell_nodes)) ||
        nfields < 0 || (nfields && (!names || !ncomp || !data)))
        return set_error(FVH_ERR_ARG, "fvh_write_vtk: bad arguments");
    for (int32_t k = 0; k < nfields; ++k)
        if (!names[k] || (ncomp[k] != 1 && ncomp[k] != 3) || !data[k])
            return set_error(FVH_ERR_ARG, "fvh_write_vtk: field %d must be a scalar (1) or vector (3) with data", k);
    for (int64_t c = 0; c < ncells; ++c) {
        const int64_t nn = cell_nodes_ptr[c + 1] - cell_nodes_ptr[c];
        if (nn < 3) return set_error(FVH_ERR_MESH, "fvh_write_vtk: cell %lld has %lld nodes", (long long)c, (long long)nn);
        for (int64_t i = cell_nodes_ptr[c]; i < cell_nodes_ptr[c + 1]; ++i)
            if (cell_nodes[i] < 0 || cell_nodes[i] >= npoints)
                return set_error(FVH_ERR_MESH, "fvh_write_vtk: cell %lld names node %lld", (long long)c,
                                 (long long)cell_nodes[i]);
    }
    std::FILE *f = std::fopen(path, "w");
    if (!f) return set_error(FVH_ERR_IO, "fvh_write_vtk: cannot open %s", path);
    std::string out;
    out.reserve(1 << 20);
    char num[40];
    auto flush = [&]() {
        std::fwrite(out.data(), 1, out.size(), f);
        out.clear();
    };
    auto put = [&](double v) {
        julia_double(v, num);
        out += num;
    };
    out += "# vtk DataFile Version 3.0\njCFD simulation data\nASCII\nDATASET UNSTRUCTURED_GRID\n";
    out += "POINTS " + std::to_string(npoints) + " double\n";
    for (int64_t i = 0; i < npoints; ++i) {
        put(xyz[3 * i]);
        out += ' ';
        put(xyz[3 * i + 1]);
        out += ' ';
        put(xyz[3 * i + 2]);
        out += '\n';
        if (out.size() > (1 << 20)) flush();
    }
    const int64_t listsize = cell_nodes_ptr[ncells] - cell_nodes_ptr[0] + ncells;
    out += "CELLS " + std::to_string(ncells) + " " + std::to_string(listsize) + "\n";
    for (int64_t c = 0; c < ncells; ++c) {
        out += std::to_string(cell_nodes_ptr[c + 1] - cell_nodes_ptr[c]);
        out += ' ';
        for (int64_t i = cell_nodes_ptr[c]; i < cell_nodes_ptr[c + 1]; ++i) {
            out += ' ';
            out += std::to_string(cell_nodes[i]);
        }
        out += '\n';
        if (out.size() > (1 << 20)) flush();
    }
    out += "CELL_TYPES " + std::to_string(ncells) + "\n";
    for (int64_t c = 0; c < ncells; ++c) {
        const int64_t nn = cell_nodes_ptr[c + 1] - cell_nodes_ptr[c];
        out += nn == 3 ? "5\n" : (nn == 4 ? "9\n" : "7\n");
    }
    out += "CELL_DATA " + std::to_string(ncells) + "\n";
    for (int32_t k = 0; k < nfields; ++k) {
        const double *d = data[k];
        if (ncomp[k] == 1) {
            out += std::string("SCALARS ") + names[k] + " double 1\nLOOKUP_TABLE CellColors\n";
            for (int64_t c = 0; c < ncells; ++c) {
                put(d[c]);
                out += '\n';
                if (out.size() > (1 << 20)) flush();
            }
        } else {
            out += std::string("VECTORS ") + names[k] + " double\n";
            for (int64_t c = 0; c < ncells; ++c) {
                put(d[c]);
                out += ' ';
                put(d[ncells + c]);
                out += ' ';
                put(d[2 * ncells + c]);
                out += '\n';
                if (out.size() > (1 << 20)) flush();
            }
        }
    }
    flush();
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) return set_error(FVH_ERR_IO, "fvh_write_vtk: writing %s failed", path);
    return FVH_OK;
    FVH_GUARD_END
}
