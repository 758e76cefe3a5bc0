// mesh.h -- a loaded model (src/common/model.h:32-80) kept indexed: the
// vertex positions of every file read, three vertex indices and a material
// id per triangle, and the reference model's per-corner lists, materials and
// bounding box.  Triangle i's record is v1 = pos[idx[3i]],
// e1 = pos[idx[3i+1]] - v1, e2 = pos[idx[3i+2]] - v1, prim_id = i,
// geom_id = geom[i].
#pragma once

#include "vr_internal.h"

#include <string>
#include <vector>

namespace vr {

// the triangles of one file: its geometric normal rule
struct mesh_span {
    uint64_t first, count;
    bool normalized;  // OBJ: normalize(cross(e1, e2)); PLY: cross(e1, e2)
};

struct mesh_model {
    std::vector<float> pos;        // 3 per vertex
    std::vector<uint32_t> idx;     // 3 per triangle, into pos
    std::vector<uint32_t> geom;    // material id per triangle
    std::vector<float> shading;    // model::shading_normals, 3 per entry
    std::vector<float> tex;        // model::tex_coords, 2 per entry
    std::vector<vr_obj_material> mats;
    std::vector<std::string> names;
    std::vector<mesh_span> spans;
    f3 lo{FLT_MAX, FLT_MAX, FLT_MAX}, hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};  // aabb::invalidate
    uint64_t degenerate = 0, missing = 0;
    // device copy of pos / idx / geom, made by the first vr_mesh_upload
    int device = -1;
    float* d_pos = nullptr;
    uint32_t* d_idx = nullptr;
    uint32_t* d_geom = nullptr;
};

// expand the indexed triangles on the device into d_prims (and the geometric
// normals into d_normals when it is not null)
int assemble_triangles(mesh_model& m, vr_triangle* d_prims, float* d_normals, hipStream_t st);

}  // namespace vr
