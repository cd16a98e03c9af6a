// Host-side internals shared by every translation unit of libfvmhip.so:
// the host mesh container and the thread-local error channel.
#pragma once

#include <cstdint>
#include <new>
#include <string>
#include <vector>

#include "../../include/fvmhip.h"

namespace fvh {

// Zero-based Mesh2 arrays (layout of fvh_mesh_desc).
struct HMesh {
    int64_t ncells = 0, nfaces = 0;
    std::vector<double> cell_centre, cell_volume;
    std::vector<int64_t> cell_faces_range, cell_faces, cell_neighbours, cell_nsign;
    std::vector<int64_t> face_owners;
    std::vector<double> face_centre, face_normal, face_e, face_area, face_delta, face_weight;
    std::vector<int64_t> boundary_range, boundary_faces, boundary_cells;
    std::vector<std::string> names;
    std::vector<double> node_xyz;                         // [3 * npoints]
    std::vector<int64_t> cell_nodes_ptr, cell_nodes;      // zero-based node ids per cell
};

// Records a formatted message for fvh_last_error and returns `code`.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// The calling thread's last message (each thread records its own).
std::string last_error_text();

}  // namespace fvh

struct fvh_hmesh {
    fvh::HMesh m;
};

#define FVH_GUARD_BEGIN try {
#define FVH_GUARD_END                                                              \
    }                                                                              \
    catch (const std::bad_alloc &) {                                               \
        return fvh::set_error(FVH_ERR_SIZE, "%s: out of host memory", __func__);   \
    }                                                                              \
    catch (const std::exception &ex) {                                             \
        return fvh::set_error(FVH_ERR_ARG, "%s: %s", __func__, ex.what());         \
    }                                                                              \
    catch (...) {                                                                  \
        return fvh::set_error(FVH_ERR_ARG, "%s: unknown failure", __func__);       \
    }
