// mesh.h -- a loaded model (src/common/model.h:32-80) kept indexed: the
// vertex positions of every file read, three vertex indices and a material
// id per triangle, and the reference model's per-corner lists, materials and
// bounding box.  Triangle i's record is v1 = pos[idx[3i]],
// e1 = pos[idx[3i+1]] - v1, e2 = pos[idx[3i+2]] - v1, prim_id = i,
// geom_id = geom[i].
#pragma once

#include "vr_internal.h"

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace vr {

// std::vector whose resize leaves new elements default-initialized (no
// zero fill of arrays that are overwritten at once; their pages are first
// touched by the threads that fill them).  Arrays of 8 MiB and more sit on
// 2 MiB-aligned memory advised for transparent huge pages: a model of
// millions of triangles then costs hundreds, not hundreds of thousands, of
// page faults.
void* huge_alloc(size_t bytes);
void huge_free(void* p);
constexpr size_t HUGE_ALLOC_MIN = (size_t)8 << 20;

template <typename T>
struct uninit_alloc : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = uninit_alloc<U>;
    };
    uninit_alloc() = default;
    template <typename U>
    uninit_alloc(const uninit_alloc<U>&) noexcept
    {
    }
    T* allocate(size_t n)
    {
        if (n * sizeof(T) >= HUGE_ALLOC_MIN) return static_cast<T*>(huge_alloc(n * sizeof(T)));
        return std::allocator<T>::allocate(n);
    }
    void deallocate(T* p, size_t n)
    {
        if (n * sizeof(T) >= HUGE_ALLOC_MIN) huge_free(p);
        else std::allocator<T>::deallocate(p, n);
    }
    template <typename U, typename... A>
    void construct(U* p, A&&... a)
    {
        if constexpr (sizeof...(A) == 0)
            ::new ((void*)p) U;
        else
            ::new ((void*)p) U(std::forward<A>(a)...);
    }
};
template <typename T>
using pod_vector = std::vector<T, uninit_alloc<T>>;

// the triangles of one file: its geometric normal rule
struct mesh_span {
    uint64_t first, count;
    bool normalized;  // OBJ: normalize(cross(e1, e2)); PLY: cross(e1, e2)
};

struct mesh_model {
    pod_vector<float> pos;        // 3 per vertex
    pod_vector<uint32_t> idx;     // 3 per triangle, into pos
    pod_vector<uint32_t> geom;    // material id per triangle
    pod_vector<float> shading;    // model::shading_normals, 3 per entry
    pod_vector<float> tex;        // model::tex_coords, 2 per entry
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
