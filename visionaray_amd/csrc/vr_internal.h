// vr_internal.h -- host-side state and helpers shared by the library's
// sources: error handling, the state behind a vr_bvh handle, launch helpers.
// The device layouts, exact arithmetic and the traversal are in the public
// visionaray_amd_core.h (user kernels compile the same code).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <string>

#include "../../include/visionaray_amd.h"
#include "../../include/visionaray_amd_core.h"

namespace vr {

// ---------------------------------------------------------------------------
// error handling: every C entry point returns a status and records a message
// ---------------------------------------------------------------------------

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define VR_HIP_CHECK(expr)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::vr::fail(VR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define VR_CHECK(cond, code, msg)                                                          \
    do {                                                                                   \
        if (!(cond)) return ::vr::fail((code), (msg));                                     \
    } while (0)

// C entry points: host exceptions become status codes
#define VR_TRY_BEGIN try {
#define VR_TRY_END                                                                                      \
    }                                                                                                   \
    catch (const std::bad_alloc&) { return ::vr::fail(VR_ERR_OUT_OF_MEMORY, "host allocation failed"); } \
    catch (const std::exception& e) { return ::vr::fail(VR_ERR_INVALID_ARGUMENT, e.what()); }

// create the current device's library state on first use
int ensure_device();

struct bvh_state;
// adjoint limits: path vertices kept per sample (num_bounces), materials in
// the per-block LDS partials
constexpr uint32_t ADJ_MAX_BOUNCES = 16;
constexpr size_t ADJ_MAX_MATERIALS_TOTAL = 2048;
// sah_cost of a BVH over its binary nodes (FP64 sums); synchronizes st
int sah_cost_device(const bvh_state& s, float ci, float cl, float cp, double* out, hipStream_t st);
// reference-layout node arrays on the host: sah_cost and tree rotations
int sah_cost_host(const vr_bvh_node* nodes, size_t n, float ci, float cl, float cp, float* out);
int optimize_rotations_host(vr_bvh_node* nodes, size_t n, int* count);
// SAH-driven tree rotations of a freshly built LBVH (before its collapse)
int lbvh_rotate(bvh_state& s, uint32_t leaf_max, uint32_t* rotations, hipStream_t st);


// Host-side state behind a vr_bvh handle.
struct bvh_state {
    int device = 0;
    size_t num_prims = 0;
    size_t num_pairs = 0;
    uint32_t root = 0;
    uint32_t depth = 0;
    uint32_t max_leaf = 1;
    bool is_lbvh = false;
    pair_node* pairs = nullptr;      // num_pairs
    quad_node* quads = nullptr;      // traversal nodes, one per pair (only those below quad 0 are reached)
    uint32_t collapse_leaf_max = 0;  // leaf size the quads were collapsed with
    tri_node* tris = nullptr;        // num_prims, traversal order
    vr_triangle* prims = nullptr;    // num_prims, caller order (tree.primitives())
    uint32_t* sorted_codes = nullptr;
    int32_t* sorted_ids = nullptr;   // traversal position -> original index
    int32_t* parents = nullptr;      // LBVH: parent pair of every node (2n-1)
    uint32_t* flags = nullptr;       // LBVH: bottom-up split slots (num_pairs)
    uint32_t* heights = nullptr;     // bottom-up subtree heights (num_pairs)
    vr_bvh_node* ref_nodes = nullptr;   // optional reference layout
    uint32_t* indices = nullptr;        // tree.indices()
    size_t num_ref_nodes = 0;
    float* root_box = nullptr;          // 6 floats
    // 8-wide nodes (vr_bvh_set_width): NODE_OCT / NODE_OCTQ per pair index,
    // remade after every build, rebuild and refit of the handle
    int wide_kind = NODE_QUAD;
    bool wide_binary_leaves = false;    // collapse over single-triangle leaves (the reference's rules)
    void* wide = nullptr;
    size_t wide_bytes = 0;
    uint32_t wide_depth = 0;            // levels of 8-wide nodes
    bool rotated = false;               // LBVH after tree rotations (no refit, no reference layout)
    bool want_rotate = false;           // LBVH builds of this handle run the rotation pass
    bool pooled = false;                // arrays from the stream-ordered pool (hipMallocAsync)
    void* arena = nullptr;              // LBVH: one pooled block holding all the arrays above
    void* scratch = nullptr;            // build scratch
    size_t scratch_bytes = 0;

    bvh_view view() const
    {
        return bvh_view{quads, tris, pairs, (uint32_t)num_prims, root, (uint32_t)num_pairs, wide};
    }
};


// ---------------------------------------------------------------------------
// shared host helpers
// ---------------------------------------------------------------------------
int current_device();
inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// the traversal stack holds the deferred children of a tree of this depth
inline bool stack_fits(const bvh_state& s) { return 3u * s.depth + 1u <= (uint32_t)MAX_STACK; }
// the node kind the traversal kernels use: the 8-wide nodes when the handle
// has them and its tree fits the wide stack and 32-bit node offsets, else
// the 4-wide ones
inline int traversal_kind(const bvh_state& s)
{
    if (s.wide_kind == NODE_QUAD || !s.wide || s.num_prims == 0 || s.num_pairs == 0) return NODE_QUAD;
    if (s.wide_depth == 0 || 7u * s.wide_depth + 1u > (uint32_t)MAX_STACK_WIDE) return NODE_QUAD;
    if (s.num_pairs > (s.wide_kind == NODE_OCT ? OCT_NEAR_FAR_MAX : QN_NEAR_FAR_MAX)) return NODE_QUAD;
    return s.wide_kind;
}
// (re)make the 8-wide nodes of s.wide_kind from the pair records
int collapse_wide(bvh_state& s, hipStream_t st);

// Collapse the binary pair records of `s` into its 4-wide traversal nodes
// (s.quads, s.root): quad p is the collapse rooted at pair p.  leaf_max > 0
// (LBVH only: pair e[3] holds the range end) turns inner subtrees of at most
// leaf_max triangles into one leaf over their contiguous range.
int collapse_quads(bvh_state& s, uint32_t leaf_max, hipStream_t st);
// binned SAH build on the device into s (the reference's tree node for node)
int sah_build_device(const vr_triangle* d_prims, size_t n, int max_leaf_size, bvh_state& s, hipStream_t st);
// an LBVH's index list (the identity) into d[n]
int lbvh_indices(uint32_t* d, size_t n, hipStream_t st);

// the work counter of the next persistent launch on the current device (one
// of a ring of counters; the launch resets it on its stream first)
uint32_t* next_work_counter();

// host side of the lbvh over boxes (host_build.cpp): the CPU lbvh_builder
// tree over 6-float boxes (lo xyz, hi xyz)
int lbvh_build_host_boxes(const float* boxes, size_t n, int max_leaf_size, vr_bvh_node* nodes, size_t* num_nodes,
                          uint32_t* indices);

}  // namespace vr

// the state behind a vr_bvh handle
struct vr_bvh_t {
    vr::bvh_state s;
    hipEvent_t built = nullptr;  // recorded after the last build / rebuild / refit on its stream
};
