/*
 * visionaray_amd.h -- C ABI of the MI355X (gfx950) ray-tracing hot path.
 *
 * The reference exposes this path only as C++ templates (no ABI).  Each entry
 * point below replaces one reference interface, cited as file:line relative to
 * the reference repository root (headers under include/visionaray/).  The C++
 * facade in visionaray_amd.hpp re-presents the reference's names on top.
 *
 * Conventions
 *   - Every function returns an int status: VR_OK (0) or a VR_ERR_* code; the
 *     message of the last failure on the calling thread is vr_last_error().
 *     (The reference prints through *_SAFE_CALL and continues,
 *     hip/safe_call.h:17-39; errors here are returned, never printed.)
 *   - Pointers named d_* are device memory (vr_malloc or any hipMalloc);
 *     h_* are host memory.  Sizes are element counts unless named *_bytes.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Work is enqueued
 *     asynchronously unless the function says it synchronizes.
 *   - Handles (vr_bvh) own their device memory; destroy them with
 *     vr_bvh_destroy.  A handle may be used from one host thread at a time.
 */
#ifndef VISIONARAY_AMD_H
#define VISIONARAY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VR_OK 0
#define VR_ERR_INVALID_ARGUMENT 1
#define VR_ERR_OUT_OF_MEMORY 2
#define VR_ERR_HIP 3
#define VR_ERR_NOT_INITIALIZED 4
#define VR_ERR_UNSUPPORTED 5

/* basic_triangle<3, float, unsigned> (math/primitive.h:14-20,
 * math/triangle.h:34-36): 44 bytes, e1 = v2 - v1, e2 = v3 - v1. */
typedef struct vr_triangle {
    uint32_t geom_id;
    uint32_t prim_id;
    float v1[3];
    float e1[3];
    float e2[3];
} vr_triangle;

/* bvh_node (bvh.h:70-141): 32 bytes; num_prims == 0 marks an inner node whose
 * children are nodes first, first + 1; a leaf covers indices
 * [first, first + num_prims). */
typedef struct vr_bvh_node {
    float lo[3];
    float hi[3];
    uint32_t first;
    uint16_t num_prims;
    uint8_t axis;
    uint8_t sign;
} vr_bvh_node;

/* basic_ray<float> (math/ray.h:16-40): 32 bytes. */
typedef struct vr_ray {
    float ori[3];
    float dir[3];
    float tmin;
    float tmax;
} vr_ray;

/* hit_record<ray, primitive<unsigned>> reduced to what callers read
 * (math/intersect.h:151-169): prim_id (-1 on a miss), t, u = b1, v = b2. */
typedef struct vr_hit {
    int32_t prim_id;
    float t;
    float u;
    float v;
} vr_hit;

/* pinhole_camera after begin_frame (detail/pinhole_camera.inl:87-97). */
typedef struct vr_camera {
    float eye[3];
    float U[3];
    float V[3];
    float W[3];
} vr_camera;

/* generic_material<emissive, matte>: type 0 = matte (cd, kd;
 * detail/material/matte.inl:43-56), 1 = emissive (ce, ls;
 * detail/material/emissive.inl:31-40).  36 bytes. */
typedef struct vr_material {
    int32_t type;
    float cd[3];
    float kd;
    float ce[3];
    float ls;
} vr_material;

/* area_light<float, basic_triangle<3, float>> (area_light.h:16-50,
 * detail/area_light.inl): intensity cl * kl over the triangle.  60 bytes. */
typedef struct vr_area_light {
    float cl[3];
    float kl;
    vr_triangle geometry;
} vr_area_light;

/* kernel_params subset used without a light list (kernels.h:38-132):
 * bounces, epsilon, background and ambient intensity (cl * kl). */
typedef struct vr_kernel_params {
    uint32_t num_bounces;
    float epsilon;
    float bg[3];
    float ambient[3];
} vr_kernel_params;

/* jittered_blend pixel sampler plus the frame counter of the scheduler
 * (pixel_sampler_types.h:51-69, detail/cuda_sched.inl:226-246). */
typedef struct vr_sched_params {
    uint32_t width;
    uint32_t height;
    uint32_t spp;
    uint32_t frame_id;
    float sfactor;
    float dfactor;
} vr_sched_params;

typedef struct vr_bvh_info {
    uint64_t num_prims;
    uint64_t num_nodes;      /* reference layout: 2n - 1 for an LBVH */
    uint32_t depth;          /* levels of the deepest leaf (root = 1) */
    uint32_t max_leaf_size;
} vr_bvh_info;

typedef struct vr_bvh_t* vr_bvh;

/* vr_lbvh_build flags */
#define VR_BVH_KEEP_NODES 1u   /* also write the reference bvh_node[2n-1] layout */
/* also run one top-down pass of SAH-driven tree rotations (Kensler, the
 * exchanges of detail/bvh/optimize.h:22-262 with up-to-date costs) on the
 * device before the traversal nodes are made: a lower sah_cost, cheaper
 * traversal, a longer build (two small device read-backs).  Not with
 * VR_BVH_KEEP_NODES; such a BVH is rebuilt (vr_lbvh_rebuild, which rotates
 * again), not refitted; vr_bvh_get_info's depth is then an upper bound. */
#define VR_BVH_ROTATE 2u

/* ---- runtime ---------------------------------------------------------- */
/* Select device `device` for this thread and create its state. */
int vr_init(int device);
const char* vr_last_error(void);
int vr_version(int* major, int* minor);
int vr_malloc(void** d_ptr, size_t bytes);
int vr_free(void* d_ptr);
int vr_memcpy_htod(void* d_dst, const void* h_src, size_t bytes);
int vr_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes);
int vr_memset(void* d_dst, int value, size_t bytes);
int vr_synchronize(void);
int vr_stream_create(void** stream);
int vr_stream_destroy(void* stream);
int vr_stream_synchronize(void* stream);
int vr_event_create(void** event);
int vr_event_destroy(void* event);
int vr_event_record(void* event, void* stream);
int vr_event_elapsed_ms(void* start, void* stop, float* ms);
/* Device-to-device copy by a kernel (16-byte vectors; bytes must be a
 * multiple of 16): graph-capturable, and a kernel of known HBM traffic for
 * calibrating hardware counters. */
int vr_device_copy(void* d_dst, const void* d_src, size_t bytes, void* stream);
/* Reads the 128-byte lines d_lines[0..n) of d_src (src_bytes long, a
 * multiple of 128; the caller keeps every index below src_bytes / 128), one
 * line per thread in eight 16-byte loads -- the access pattern of a traversal
 * lane reading a 4-wide node -- and xors what it read into *d_out.  With
 * distinct lines its HBM traffic is known (n * 128 + n * 4 bytes): it
 * calibrates hardware traffic counters for gathers.  No reference counterpart. */
int vr_device_gather_lines(const void* d_src, size_t src_bytes, const uint32_t* d_lines, size_t n, uint32_t* d_out,
                           void* stream);

/* ---- camera ------------------------------------------------------------- */
/* pinhole_camera::perspective + look_at + begin_frame
 * (detail/pinhole_camera.inl:12-29, 31-64, 87-97). Host only. */
int vr_camera_look_at(const float eye[3], const float center[3], const float up[3],
                      float fovy, float aspect, vr_camera* out);

/* ---- BVH construction ------------------------------------------------- */
/* lbvh_builder::build(cuda_index_bvh<P>{}, P* prims, n)
 * (detail/bvh/lbvh.h:681-844): Morton codes, stable radix sort, Karras
 * hierarchy, bottom-up boxes.  d_prims stays owned by the caller; the BVH
 * keeps its triangles in its own traversal order, read from d_prims by the
 * build's kernels, so d_prims must stay valid until the build's work on
 * `stream` has run (the reference copies the primitives first, from the same
 * stream-ordered buffer).  n == 0 yields an empty BVH. */
int vr_lbvh_build(const vr_triangle* d_prims, size_t n, uint32_t flags, void* stream, vr_bvh* out);
/* The same build into an existing LBVH handle (a new frame of moving
 * geometry): its device arrays are reused when n is unchanged, so no device
 * memory is allocated; the node layout is kept if the handle kept it.  The
 * reference rebuilds with a new lbvh_builder::build call per frame. */
int vr_lbvh_rebuild(vr_bvh bvh, const vr_triangle* d_prims, size_t n, void* stream);

/* assign_morton_codes (detail/bvh/lbvh.h:242-269) alone: codes[i] for
 * primitive i, unsorted. */
int vr_lbvh_morton(const vr_triangle* d_prims, size_t n, uint32_t* d_codes, void* stream);

/* The sorted prim_refs of the last build (lbvh.h:740-755 output):
 * codes and original ids in sorted order, n each. */
int vr_bvh_sorted_refs(vr_bvh bvh, uint32_t* d_codes, int32_t* d_ids, void* stream);

/* index_bvh_t built elsewhere (e.g. binned_sah_builder on the host,
 * detail/bvh/sah.h:176-807): adopt a reference-layout node array and index
 * list (bvh.h:870-1010).  h_nodes/h_indices are host arrays. */
int vr_bvh_from_nodes(const vr_triangle* h_prims, size_t num_prims, const vr_bvh_node* h_nodes,
                      size_t num_nodes, const uint32_t* h_indices, vr_bvh* out);

/* bvh_refitter::refit (detail/bvh/refit.h:28-114): recompute every box for
 * moved vertices (same count and order as at build time); d_prims as for
 * vr_lbvh_build. */
int vr_bvh_refit(vr_bvh bvh, const vr_triangle* d_prims, void* stream);

/* vr_bvh_set_width flags */
#define VR_WIDE_COMPRESSED 1u     /* 8-bit child planes, one 128-byte line per node */
#define VR_WIDE_BINARY_LEAVES 2u  /* an LBVH's single-triangle leaves (the reference's collapse exactly) */
/* Traversal width of a BVH: the reference's wide trees (bvh_collapser,
 * detail/bvh/collapse.h:19-169, into index_bvh8 of bvh_multi_node<8>,
 * bvh.h:167-312; with VR_WIDE_COMPRESSED the 8-bit form of
 * bvh_compressed_node<8>, bvh.h:316-426, quantized outward so that a
 * compressed tree finds exactly the hits of the uncompressed one).  Width 8
 * collapses the binary tree into nodes of up to eight children, top down by
 * the reference's rules; the handle keeps them through later rebuilds and
 * refits, and vr_trace*, vr_render_pt and vr_render_pt_counted traverse them
 * (a tree of more than 48 8-wide levels keeps the 4-wide traversal).  By
 * default an LBVH's 8-wide leaves are the leaf ranges of its 4-wide nodes.
 * Width 4 drops them.  Waits for `stream` (the 8-wide depth is read back). */
int vr_bvh_set_width(vr_bvh bvh, int width, uint32_t flags, void* stream);
/* The 8-wide nodes to the host: *bytes = their size (256-byte float nodes or
 * 128-byte compressed ones, indexed by the pair index of their root binary
 * node, node 0 the root; only nodes reachable from the root are written),
 * *depth = 8-wide levels.  h_out may be null to ask for the size. */
int vr_bvh_download_wide(vr_bvh bvh, void* h_out, size_t capacity, size_t* bytes, uint32_t* depth);

/* Counts and depth.  The depth of an LBVH is read back from the device after
 * the handle's last build / rebuild / refit (a wait on that work's stream up
 * to it, not on the whole device). */
int vr_bvh_get_info(vr_bvh bvh, vr_bvh_info* info);

/* Device view of a BVH for kernels written against the device API
 * (visionaray_amd_device.hpp): tree.ref() (bvh.h:919-931), the POD
 * index_bvh_ref_t (bvh.h:495-569) kernels take by value.  Non-owning: valid
 * while the handle lives and is not rebuilt. */
typedef struct vr_bvh_ref {
    const void* nodes;   /* 128-byte 4-wide traversal nodes */
    const void* prims;   /* 48-byte triangles in traversal order */
    const void* pairs;   /* 64-byte binary node pairs */
    uint32_t num_prims;
    uint32_t root;       /* entry of the root node */
    uint32_t num_nodes;  /* traversal nodes (= pair records) */
    uint32_t reserved;
} vr_bvh_ref;
/* VR_ERR_UNSUPPORTED for a tree deeper than the traversal stack. */
int vr_bvh_get_ref(vr_bvh bvh, vr_bvh_ref* out);
/* Reference-layout nodes (tree.nodes(), bvh.h:910) into d_nodes[num_nodes]. */
int vr_bvh_nodes(vr_bvh bvh, vr_bvh_node* d_nodes, void* stream);
/* tree.indices() (bvh.h:913) into d_indices[num_prims]. */
int vr_bvh_indices(vr_bvh bvh, uint32_t* d_indices, void* stream);
/* Releases the handle.  Waits for the device first, as the reference's
 * device_vector destructor does (cudaFree, cuda/device_vector.h): work on any
 * stream may still read the tree. */
int vr_bvh_destroy(vr_bvh bvh);
/* Stream-ordered release with no host wait (per-frame rebuild loops, several
 * streams): the arrays go back to the device's pool once the work enqueued on
 * `stream` before this call is done; the caller orders any use of the tree on
 * other streams before it.  No counterpart in the reference. */
int vr_bvh_destroy_async(vr_bvh bvh, void* stream);

/* Host binned SAH build (binned_sah_builder::build, detail/bvh/sah.h:176-807,
 * 16 bins, no spatial splits).  h_nodes must hold 2n - 1 nodes; the node count
 * used is returned in *num_nodes. */
int vr_sah_build(const vr_triangle* h_prims, size_t n, int max_leaf_size, vr_bvh_node* h_nodes,
                 size_t* num_nodes, uint32_t* h_indices);

/* The same binned SAH build on the device, into a BVH handle: the tree is
 * the reference's node for node (the node array and index list of
 * vr_bvh_nodes / vr_bvh_indices equal binned_sah_builder's, boxes up to the
 * sign of a zero), built one level at a time.  d_prims as for
 * vr_lbvh_build; max_leaf_size <= 0 means 4.  VR_ERR_UNSUPPORTED for a tree
 * with a leaf of more than 32 primitives (the traversal's limit).  Waits for
 * `stream` once per tree level (the level's node counts are read back). */
int vr_sah_build_device(const vr_triangle* d_prims, size_t n, int max_leaf_size, void* stream, vr_bvh* out);

/* Host LBVH build in the reference's CPU layout (lbvh_builder::build on an
 * index_bvh, detail/bvh/lbvh.h:500-656 driven by build_top_down): Morton
 * order, splits at the highest differing code bit, equal codes split at the
 * midpoint, leaves of up to max_leaf_size (default 4).  Arrays as
 * vr_sah_build. */
int vr_lbvh_build_host(const vr_triangle* h_prims, size_t n, int max_leaf_size, vr_bvh_node* h_nodes,
                       size_t* num_nodes, uint32_t* h_indices);

/* ---- traversal -------------------------------------------------------- */
/* ---- tree quality ------------------------------------------------------
 * sah_cost(tree, ci, cl, cp) (detail/bvh/statistics.h:38-68) over a
 * reference-layout node array on the host: the float sums in node order, as
 * the reference computes them. */
int vr_bvh_sah_cost_host(const vr_bvh_node* h_nodes, size_t num_nodes, float ci, float cl, float cp, float* cost);
/* The same SAH cost of a BVH handle, evaluated on the device over its binary
 * nodes (FP64 sums; agrees with the float sums to their rounding).
 * Synchronizes the stream. */
int vr_bvh_sah_cost(vr_bvh bvh, float ci, float cl, float cp, double* cost, void* stream);
/* bvh_optimizer::optimize_tree_rotations (detail/bvh/optimize.h:22-262,
 * Kensler tree rotations) on a reference-layout node array on the host, in
 * place; *num_rotations = rotations performed.  The reference's result, node
 * for node. */
int vr_bvh_optimize_rotations_host(vr_bvh_node* h_nodes, size_t num_nodes, int* num_rotations);

/* closest_hit / any_hit over one BVH (detail/traverse_linear.inl:149-214 ->
 * intersect<ClosestHit|AnyHit>, detail/bvh/intersect.inl:41-357). */
int vr_trace(vr_bvh bvh, int any_hit, const vr_ray* d_rays, size_t n, vr_hit* d_hits, void* stream);

/* vr_trace that allocates no device memory (no hipMalloc, hipFree or
 * allocating library call in any call, the first included).  No counterpart
 * in the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  Its work counter
 * lives in a 256-byte file-scope device array resolved per device at vr_init
 * (or at the first call on a device) instead of the library's counter ring.
 * Limit: at most 2^31 rays per call (VR_ERR_INVALID_ARGUMENT beyond).  Every
 * call on one GPU shares that array, so calls that overlap on one GPU (other
 * streams, other host threads) race; serialize them. */
int vr_trace_noalloc(vr_bvh bvh, int any_hit, const vr_ray* d_rays, size_t n, vr_hit* d_hits, void* stream);

/* Same, also counting per ray the traversal-node visits and triangle tests
 * (d_counts[2 * i] = nodes, d_counts[2 * i + 1] = triangles); for the
 * algorithmic-bytes model. */
int vr_trace_counted(vr_bvh bvh, int any_hit, const vr_ray* d_rays, size_t n, vr_hit* d_hits,
                     uint32_t* d_counts, void* stream);

/* ---- instanced two-level BVH ------------------------------------------- */
/* One instance of a bottom-level BVH (index_bvh_t::inst(mat4x3),
 * bvh.h:679-782): transform is the mat4x3 as its 12 floats, column by column
 * (three columns of the 3x3 part, then the translation; mat4x3(T const[12]),
 * math/detail/matrix4x3.inl:55-61).  As in the reference, the instance maps
 * object space to world space by the TRANSPOSE of the 3x3 part
 * (top_left(mat4x3) transposes, matrix4x3.inl:206-218) and then the
 * translation; inst_id is what set_inst_id stores and what hits report. */
typedef struct vr_instance {
    uint32_t blas;    /* index into the BVH list given to vr_tlas_build */
    int32_t inst_id;
    float transform[12];
} vr_instance;

/* hit_record_bvh_inst reduced to what callers read
 * (detail/bvh/hit_record.h:54-76, math/intersect.h:151-169): prim_id and
 * geom_id of the triangle hit (-1 on a miss), t, u, v in the instance's
 * space (t is the same parameter in both spaces), the instance's inst_id and
 * its position in the instance list. */
typedef struct vr_inst_hit {
    int32_t prim_id;
    float t;
    float u;
    float v;
    int32_t inst_id;
    int32_t instance;
    int32_t geom_id;
    int32_t pad;
} vr_inst_hit;

typedef struct vr_tlas_t* vr_tlas;

/* The top level over instances of built BVHs: each instance's box is its BVH
 * root box's eight corners through the instance transform (get_bounds,
 * detail/bvh/get_bounds.inl:30-56), the tree over the boxes is the
 * reference's CPU lbvh_builder tree (lbvh.h:500-656, leaves of up to 4
 * instances), built on the host.  The BVHs must outlive the handle and keep
 * their arrays (no rebuild to a different size while it lives). */
int vr_tlas_build(const vr_bvh* blas, uint32_t num_blas, const vr_instance* h_instances, uint32_t num_instances,
                  void* stream, vr_tlas* out);
/* closest_hit over the top level (detail/bvh/intersect.inl:362-394: the
 * ray transformed into each instance's space, the instance's BVH traversed,
 * the closest hit kept). */
int vr_tlas_closest_hit(vr_tlas tlas, const vr_ray* d_rays, size_t n, vr_inst_hit* d_hits, void* stream);
/* any_hit over the top level (the same overload with the AnyHit traversal):
 * a ray stops at its first hit inside [tmin, tmax], in any instance. */
int vr_tlas_any_hit(vr_tlas tlas, const vr_ray* d_rays, size_t n, vr_inst_hit* d_hits, void* stream);
/* the top level's reference-layout nodes and instance order (host copies;
 * num_nodes gets the node count, nodes/indices may be null to query it) */
int vr_tlas_nodes(vr_tlas tlas, vr_bvh_node* h_nodes, uint32_t* h_indices, size_t* num_nodes);
int vr_tlas_destroy(vr_tlas tlas);

/* ---- kernels under the scheduler -------------------------------------- */
/* Ambient occlusion kernel of src/examples/ao/main.cpp:192-260 under the
 * jittered_blend sampler; d_accum/d_color are RGBA32F width*height. */
int vr_render_ao(vr_bvh bvh, const vr_camera* cam, const vr_sched_params* sp, uint32_t ao_samples,
                 float ao_radius, const float bg[3], float* d_accum, float* d_color, void* stream);

/* pathtracing::kernel (detail/pathtracing.inl:43-233) under the
 * jittered_blend sampler and per-pixel seeds make_random_seed(y*w+x, frame)
 * (detail/cuda_sched.inl:32-69), on a persistent wavefront that regenerates
 * paths.  d_materials is indexed by the triangle's geom_id.  d_ray_counts
 * (optional, 2 x uint64, accumulated) receives primary and secondary
 * closest-hit queries. */
int vr_render_pt(vr_bvh bvh, const vr_material* d_materials, size_t num_materials,
                 const vr_kernel_params* kp, const vr_camera* cam, const vr_sched_params* sp,
                 float* d_accum, float* d_color, uint64_t* d_ray_counts, void* stream);

/* vr_render_pt with per-frame traversal statistics for the algorithmic-bytes
 * model: d_stats (4 x uint64, accumulated) = primary queries, secondary
 * queries, traversal-node visits, triangle tests.  Same image as vr_render_pt. */
int vr_render_pt_counted(vr_bvh bvh, const vr_material* d_materials, size_t num_materials,
                         const vr_kernel_params* kp, const vr_camera* cam, const vr_sched_params* sp,
                         float* d_accum, float* d_color, uint64_t* d_stats, void* stream);

/* pathtracing::kernel with a light list (make_kernel_params w/ lights,
 * kernels.h:137-187; next-event estimation with multiple importance sampling,
 * detail/pathtracing.inl:116-196): every hit samples one of the num_lights
 * area lights uniformly (sample_random_light, sampling.h:194-210) and traces
 * a shadow ray (any hit); emitter hits after the first bounce are MIS
 * weighted.  d_lights is device memory.  num_lights == 0 is vr_render_pt.
 * d_stats (optional, 2 x uint64, or 4 with `counted`) as in vr_render_pt /
 * vr_render_pt_counted; shadow rays count as secondary queries.  The
 * reference-scheme launch shape (VR_SCHEDULE_PER_PIXEL) does not apply: the
 * persistent kernel runs. */
int vr_render_pt_lights(vr_bvh bvh, const vr_material* d_materials, size_t num_materials,
                        const vr_area_light* d_lights, size_t num_lights, const vr_kernel_params* kp,
                        const vr_camera* cam, const vr_sched_params* sp, float* d_accum, float* d_color,
                        uint64_t* d_stats, int counted, void* stream);

/* ---- differentiation (no reference counterpart: the reference renders
 * only; these two calls are the forward/backward pair of a differentiable
 * frame) --------------------------------------------------------------- */
/* L2 image loss sum over pixels and RGB of (I - T)^2 into *d_loss (FP64,
 * accumulated) and, when d_dl_dimage is not NULL, its image gradient
 * 2 (I - T) (RGBA32F, alpha 0).  Images are RGBA32F, num_pixels each. */
int vr_image_l2_loss(const float* d_image, const float* d_target, size_t num_pixels, float* d_dl_dimage,
                     double* d_loss, void* stream);

/* Path-replay adjoint of vr_render_pt with the same arguments: replays every
 * path of that frame (same seeds, hits recomputed) and accumulates into
 * d_grad[6 * m + 0..2] d loss / d cd (matte albedo) and d_grad[6 * m + 3..5]
 * d loss / d ce (emission) of material m, FP64, given d loss / d image
 * (RGBA32F).  Russian-roulette probabilities are held constant.  At most 2048
 * materials and 16 bounces.  d_ray_counts as in vr_render_pt. */
int vr_render_pt_backward(vr_bvh bvh, const vr_material* d_materials, size_t num_materials,
                          const vr_kernel_params* kp, const vr_camera* cam, const vr_sched_params* sp,
                          const float* d_dl_dimage, double* d_grad, uint64_t* d_ray_counts, void* stream);

/* vr_render_pt with the Russian-roulette probabilities taken from the path
 * throughput under d_rr_materials (same count and layout) while shading uses
 * d_materials.  With both tables equal it renders exactly vr_render_pt's
 * frame.  Holding d_rr_materials at a parameter point p0 gives the estimator
 * whose derivative at p0 vr_render_pt_backward returns: central differences
 * of it with common random numbers keep every roulette decision, so they
 * check the adjoint without the noise of decisions that flip. */
int vr_render_pt_frozen_rr(vr_bvh bvh, const vr_material* d_materials, const vr_material* d_rr_materials,
                           size_t num_materials, const vr_kernel_params* kp, const vr_camera* cam,
                           const vr_sched_params* sp, float* d_accum, float* d_color, uint64_t* d_ray_counts,
                           void* stream);

/* ---- launch shape and autotuning (no reference counterpart: cuda_sched
 * takes a fixed 16x16 block, detail/cuda_sched.h:43) -------------------- */
/* Launch shape of the path-tracing frame kernel: threads per block (128 or
 * 256), traversal-stack entries per lane kept in LDS (8, 16 or 32; deeper
 * entries spill to private memory), the pixel-queue tile (pixels are
 * handed out in tile x tile squares, 1 = scanlines; 1..64) and the schedule:
 * VR_SCHEDULE_PERSISTENT (a resident grid whose lanes take the next pixel
 * when theirs is done, 4-wide traversal) or VR_SCHEDULE_PER_PIXEL (the
 * reference's GPU scheme, for comparisons: one thread per pixel over the
 * whole image as cuda_sched launches it, detail/cuda_sched.inl:27-70, with
 * the binary-tree traversal of detail/bvh/intersect.inl:166-356; stack_lds 16
 * only; same image up to ties).  Batches always run persistent. */
#define VR_SCHEDULE_PERSISTENT 0u
#define VR_SCHEDULE_PER_PIXEL 1u
typedef struct vr_launch_shape {
    uint32_t block_size;
    uint32_t stack_lds;
    uint32_t tile;
    uint32_t schedule;
} vr_launch_shape;

int vr_set_launch_shape(const vr_launch_shape* shape);
int vr_get_launch_shape(vr_launch_shape* shape);
/* Time every compiled shape with tiles {1, 8, 16} on this frame (arguments
 * as vr_render_pt; one warm-up and `reps` timed frames each), keep the
 * fastest as the current shape and return it with its mean frame time.
 * Synchronizes the stream. */
int vr_autotune_pt(vr_bvh bvh, const vr_material* d_materials, size_t num_materials, const vr_kernel_params* kp,
                   const vr_camera* cam, const vr_sched_params* sp, float* d_accum, float* d_color, int reps,
                   vr_launch_shape* best, float* best_ms, void* stream);

/* ---- batched scenes (no reference counterpart: the reference launches one
 * cuda_sched::frame per scene, detail/cuda_sched.inl:226-246) ------------- */
/* One scene of a batch: what one reference frame() call renders. */
typedef struct vr_scene {
    vr_bvh bvh;
    const vr_material* d_materials;
    size_t num_materials;
    vr_kernel_params kp;
    vr_camera cam;
    vr_sched_params sp;          /* frame_id: the scene's first frame */
    float* d_accum;
    float* d_color;
    const float* d_dl_dimage;    /* for vr_batch_backward only; may be NULL */
} vr_scene;

typedef struct vr_batch_t* vr_batch;

/* Prepare a batch: scene descriptors and the pixel prefix go to device
 * memory once.  The scenes' BVHs and buffers must outlive the batch. */
int vr_batch_create(const vr_scene* h_scenes, size_t num_scenes, vr_batch* out);
/* One frame of every scene in a single persistent launch (pixels of all
 * scenes back to back; paths regenerate across scene boundaries).  Frame k
 * of the batch (k = 0, 1, ...) seeds scene s with sp.frame_id + k.  Enqueues
 * kernels only, so it can be captured in a HIP graph. */
int vr_batch_render(vr_batch batch, uint64_t* d_ray_counts, void* stream);
/* Adjoint of the batch's last frame: d_grad holds 6 slots per material, the
 * scenes' material tables concatenated in scene order (at most 2048 in
 * total); each scene's d_dl_dimage must be set. */
int vr_batch_backward(vr_batch batch, double* d_grad, uint64_t* d_ray_counts, void* stream);
/* Capture vr_batch_render on `stream` (not the NULL stream) into a graph
 * owned by the batch; vr_batch_launch_graph replays it (one frame each). */
int vr_batch_capture_graph(vr_batch batch, uint64_t* d_ray_counts, void* stream);
int vr_batch_launch_graph(vr_batch batch, void* stream);
int vr_batch_destroy(vr_batch batch);

/* ---- the minimum export set named for the drop-in boundary -------------
 * The same operations under the names of the boundary a binding expects
 * (a context per host thread, scene upload, node download, closest / any
 * queries, forward / backward, autotune, destroy, error string); each is a
 * thin equivalent of the entry point named in its comment. */
typedef struct vr_context_t* vr_context;
/* vr_init(device) and a handle that records the device; the library keeps
 * its per-device state itself, so other calls do not take the context. */
int vr_create_context(int device, vr_context* out);
int vr_destroy_context(vr_context ctx);
/* Device copy of n host triangles (44-B reference records) in *d_out,
 * released with vr_free. */
int vr_scene_upload(const vr_triangle* h_prims, size_t n, vr_triangle** d_out, void* stream);
/* The reference node layout (vr_bvh_nodes) into host memory: at most
 * *num_nodes records; *num_nodes returns the count (call with h_nodes NULL
 * to query it). */
int vr_bvh_download_nodes(vr_bvh bvh, vr_bvh_node* h_nodes, size_t* num_nodes);
/* vr_trace with any_hit 0 / 1. */
int vr_trace_closest(vr_bvh bvh, const vr_ray* d_rays, size_t n, vr_hit* d_hits, void* stream);
int vr_trace_any(vr_bvh bvh, const vr_ray* d_rays, size_t n, vr_hit* d_hits, void* stream);
/* vr_render_pt, the forward half of vr_render_pt_backward. */
int vr_render_pt_forward(vr_bvh bvh, const vr_material* d_materials, size_t num_materials, const vr_kernel_params* kp,
                         const vr_camera* cam, const vr_sched_params* sp, float* d_accum, float* d_color,
                         uint64_t* d_ray_counts, void* stream);
/* vr_autotune_pt. */
int vr_autotune(vr_bvh bvh, const vr_material* d_materials, size_t num_materials, const vr_kernel_params* kp,
                const vr_camera* cam, const vr_sched_params* sp, float* d_accum, float* d_color, int reps,
                vr_launch_shape* best, float* best_ms, void* stream);
/* vr_bvh_destroy, vr_batch_destroy, vr_last_error. */
int vr_destroy_bvh(vr_bvh bvh);
int vr_destroy_batch(vr_batch batch);
const char* vr_last_error_string(void);

/* ---- scene ingestion -----------------------------------------------------
 * model::load (src/common/model.cpp:109-200) for Wavefront OBJ files with
 * their MTL libraries (load_obj, src/common/obj_loader.cpp:306-520) and PLY
 * files (load_ply, src/common/ply_loader.cpp:64-211).  A mesh keeps the
 * triangles indexed (vertex positions, three indices and a material id per
 * triangle); vr_mesh_triangles expands them on the host, vr_mesh_upload on
 * the device.  Triangles, ids, per-corner normals and texture coordinates,
 * materials and the bounding box are the reference model's. */

/* sg::obj_material (src/common/sg/material.h:41-72) as filled by
 * add_material (src/common/obj_loader.cpp:248-263). */
typedef struct vr_obj_material {
    float ca[3];
    float cd[3];
    float cs[3];
    float ce[3];
    float ior[3];
    float transmission;
    float specular_exp;
    int32_t illum;
} vr_obj_material;

typedef struct vr_mesh_info {
    uint64_t num_prims;           /* model::primitives */
    uint64_t num_vertices;        /* positions read (every OBJ `v`, every PLY vertex) */
    uint64_t num_shading_normals; /* model::shading_normals (3 per triangle that has them) */
    uint64_t num_tex_coords;      /* model::tex_coords */
    uint64_t num_materials;       /* model::materials */
    uint64_t num_degenerate;      /* OBJ triangles of zero area, rejected (obj_loader.cpp:51-74) */
    uint64_t num_missing;         /* mtllib files and usemtl names that were not found */
    float bbox_lo[3];             /* model::bbox */
    float bbox_hi[3];
} vr_mesh_info;

typedef struct vr_mesh_t* vr_mesh;

/* Load num_paths files into one model, in order (model::load).  The format
 * follows each file's extension (.obj/.OBJ, .ply/.PLY); other formats are
 * VR_ERR_UNSUPPORTED.  An OBJ file is parsed by all host threads
 * (VR_IO_THREADS caps them). */
int vr_mesh_load(const char* const* paths, size_t num_paths, vr_mesh* out);
int vr_mesh_get_info(vr_mesh mesh, vr_mesh_info* info);
/* model::primitives: num_prims 44-B triangles into host memory. */
int vr_mesh_triangles(vr_mesh mesh, vr_triangle* h_prims);
/* model::geometric_normals, 3 floats per triangle: normalize(cross(e1, e2))
 * for OBJ triangles (obj_loader.cpp:506-511), cross(e1, e2) for PLY ones
 * (ply_loader.cpp:169). */
int vr_mesh_geometric_normals(vr_mesh mesh, float* h_normals);
/* model::shading_normals (3 floats each) and model::tex_coords (2 each). */
int vr_mesh_shading_normals(vr_mesh mesh, float* h_normals);
int vr_mesh_tex_coords(vr_mesh mesh, float* h_tex_coords);
/* model::materials (num_materials records) and the name of material i
 * (the MTL newmtl name, empty for defaults) into buf[cap]. */
int vr_mesh_materials(vr_mesh mesh, vr_obj_material* h_materials);
int vr_mesh_material_name(vr_mesh mesh, size_t i, char* buf, size_t cap);
/* Path-tracing materials from model::materials by map_material
 * (src/viewer/viewer.cpp:745-795): emissive when length(ce) > 0 (ce,
 * ls = 1), matte when illum == 1 (cd, kd = 1).  The other kinds it makes
 * (mirror, glass, plastic) are VR_ERR_UNSUPPORTED, unless diffuse_fallback is
 * set: then they become matte with cd. */
int vr_mesh_pt_materials(vr_mesh mesh, int diffuse_fallback, vr_material* h_materials);
/* The triangles (and, when d_normals is not NULL, the geometric normals)
 * built on the device from the indexed mesh: vertex positions, indices and
 * material ids are copied over (12 B per vertex + 16 B per triangle rather
 * than 44 B per triangle) and expanded by one kernel into d_prims[num_prims]
 * (and d_normals[3 num_prims]), identical to vr_mesh_triangles and
 * vr_mesh_geometric_normals.  The mesh keeps its device copy until
 * vr_mesh_destroy. */
int vr_mesh_upload(vr_mesh mesh, vr_triangle* d_prims, float* d_normals, void* stream);
int vr_mesh_destroy(vr_mesh mesh);

#ifdef __cplusplus
}
#endif

#endif
