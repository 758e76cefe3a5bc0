/*
 * vr_oracle.h -- CPU restatement of Visionaray's ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in visionaray_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links or calls it.
 *
 * Every function restates one piece of the reference (file:line cited,
 * relative to the reference repository root) with the same IEEE float
 * operations in the same order.  It is compiled with -ffp-contract=off so
 * that no multiply-add is fused; the GPU kernels are built the same way.
 *
 * Pinning (what makes this restatement trustworthy, see tests/):
 *   morton_encode3d     KATs of test/unittests/morton.cpp:49-67, brute-force
 *                       bit interleave on random inputs, reference golden codes
 *   morton_codes        golden codes from the reference's own CPU
 *                       lbvh_builder::init (same formula), special cases
 *   determine_range     the reference's host-callable determine_range run on the
 *                       same sorted codes (oracle/_ref), brute-force radix tree
 *   lbvh_build          structural invariants (every leaf once, parent box is
 *                       the exact union) + reference determine_range splits
 *   trace               KATs test/unittests/math/intersect.cpp:60-121,
 *                       test/unittests/get_normal.cpp:17-75, brute force over
 *                       all triangles, reference intersect() golden hits
 *   make_random_seed    golden table from the reference header
 *   rng                 golden next() streams from the reference CPU engine
 *   render_ao/render_pt golden images from the reference tiled_sched path
 *   collapse_wide       the reference's bvh_collapser run on the same binary
 *                       trees (oracle/_ref), node for node, W = 4 and 8
 *   compress_wide       the reference's bvh_compressor (bvh_compressed_node::init) on the same
 *                       wide nodes (oracle/_ref), byte for byte on used slots
 */
#ifndef VR_ORACLE_H
#define VR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* basic_triangle<3, float, unsigned>: primitive<unsigned>{geom_id, prim_id}
 * followed by v1, e1, e2 (math/primitive.h:14-20, math/triangle.h:34-36). 44 B. */
typedef struct {
    uint32_t geom_id;
    uint32_t prim_id;
    float v1[3];
    float e1[3];
    float e2[3];
} vro_tri;

/* bvh_node (bvh.h:70-141): bbox, first_child|first_prim, num_prims, axis, sign. 32 B. */
typedef struct {
    float lo[3];
    float hi[3];
    uint32_t first;
    uint16_t num_prims;
    uint8_t axis;
    uint8_t sign;
} vro_node;

/* basic_ray<float>: ori, dir, tmin, tmax. 32 B. */
typedef struct {
    float ori[3];
    float dir[3];
    float tmin;
    float tmax;
} vro_ray;

/* Hit record: prim_id (-1 = miss), t, u (b1), v (b2). 16 B. */
typedef struct {
    int32_t prim_id;
    float t;
    float u;
    float v;
} vro_hit;

/* Pinhole camera after begin_frame(): eye and the U, V, W basis
 * (detail/pinhole_camera.inl:87-97). */
typedef struct {
    float eye[3];
    float U[3];
    float V[3];
    float W[3];
} vro_camera;

/* generic_material<emissive, matte>: type 0 = matte (cd, kd), 1 = emissive (ce, ls). */
typedef struct {
    int32_t type;
    float cd[3];
    float kd;
    float ce[3];
    float ls;
} vro_material;

/* area_light<float, basic_triangle<3, float>> (area_light.h:16-50): cl, kl,
 * then the triangle.  60 B. */
typedef struct {
    float cl[3];
    float kl;
    vro_tri geometry;
} vro_light;

/* Path-tracer / AO parameters (kernels.h:38-74, pixel_sampler_types.h:51-69). */
typedef struct {
    uint32_t width;
    uint32_t height;
    uint32_t spp;
    uint32_t frame_id;
    uint32_t num_bounces;
    float epsilon;
    float bg[3];       /* background intensity, cl * kl */
    float ambient[3];  /* ambient (escape) intensity, cl * kl */
    float sfactor;
    float dfactor;
    uint32_t ao_samples;
    float ao_radius;
    /* order in which two gen.next() calls in one argument list are evaluated:
     * 0 = left argument first, 1 = right argument first */
    uint32_t arg_order;
} vro_render_params;

/* ---- LBVH construction (detail/bvh/lbvh.h) ---- */
uint32_t vro_morton_encode3d(uint32_t x, uint32_t y, uint32_t z);
void vro_tri_bounds(const vro_tri* t, float lo[3], float hi[3]);
void vro_centroid_bounds(const vro_tri* tris, size_t n, float lo[3], float hi[3]);
void vro_morton_codes(const vro_tri* tris, size_t n, uint32_t* codes);
void vro_sort_prim_refs(const uint32_t* codes, size_t n, uint32_t* sorted_codes, int32_t* sorted_ids);
void vro_determine_range(const uint32_t* sorted_codes, int n, int i, int* first, int* last, int* split);
int vro_lbvh_build(const vro_tri* tris, size_t n, vro_node* nodes, uint32_t* indices,
                   uint32_t* sorted_codes, int32_t* sorted_ids);
void vro_refit(vro_node* nodes, size_t num_nodes, const vro_tri* tris, const uint32_t* indices);

/* ---- traversal (detail/bvh/intersect.inl, math/intersect.h) ---- */
void vro_intersect_tri(const vro_ray* r, const vro_tri* t, int* hit, float* tt, float* u, float* v);
void vro_trace(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
               const vro_ray* rays, size_t num_rays, int any_hit, vro_hit* hits);
void vro_trace_brute(const vro_tri* tris, size_t num_prims, const vro_ray* rays, size_t num_rays,
                     int any_hit, vro_hit* hits);

/* ---- sampling (make_random_seed.h, random_generator.h) ---- */
uint32_t vro_make_random_seed(int32_t v0, int32_t v1);
void vro_rng_stream(uint32_t seed, size_t count, float* out);

/* ---- kernels (examples/ao, detail/pathtracing.inl) over a whole frame ---- */
void vro_render_ao(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                   const vro_camera* cam, const vro_render_params* p, float* accum, float* color,
                   int y_begin, int y_end);
void vro_render_pt(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                   const vro_material* mats, const vro_camera* cam, const vro_render_params* p,
                   float* accum, float* color, int y_begin, int y_end, uint64_t* ray_counts);

/* render_pt with a light list (make_kernel_params w/ lights, kernels.h:137-187):
 * next-event estimation with multiple importance sampling. */
void vro_render_pt_lights(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                          const vro_material* mats, const vro_light* lights, size_t num_lights,
                          const vro_camera* cam, const vro_render_params* p, float* accum, float* color,
                          int y_begin, int y_end, uint64_t* ray_counts);

/* Diagnostics: log every query the path tracer traces (closest hits and
 * shadow rays) into buf[cap] until vro_query_log(NULL, 0). */
typedef struct {
    vro_ray ray;
    int32_t any;
    int32_t hit;
    int32_t prim_id;
    float t;
    uint32_t bounce;
    uint32_t pixel;
} vro_query;
void vro_query_log(vro_query* buf, size_t cap);
size_t vro_query_log_count(void);

/* Adjoint of render_pt (no reference counterpart): FP64 gradients of a loss
 * with image gradient dl_dimg (RGBA per pixel) w.r.t. matte albedo cd and
 * emission ce, 3 per material each, accumulated into grad_cd / grad_ce. */
void vro_render_pt_adjoint(const vro_node* nodes, const vro_tri* tris, const uint32_t* indices, size_t num_prims,
                           const vro_material* mats, const vro_camera* cam, const vro_render_params* p,
                           const float* dl_dimg, double* grad_cd, double* grad_ce, int y_begin, int y_end);

/* ---- wide trees (bvh.h:167-426, detail/bvh/collapse.h:19-169) ---- */
/* bvh_collapser::collapse of a binary tree (reference bvh_node layout) into
 * bvh_multi_node<W>, W = 4 or 8, after the reference's removal of empty
 * nodes; child bounds per node as minx[W] miny[W] minz[W] maxx[W] maxy[W]
 * maxz[W] (out_bounds, 6 W floats) and children[W] (out_children, int64:
 * >= 0 inner node, < 0 an encoded leaf ~(count << 48 | first), INT64_MAX
 * unused).  Returns the number of nodes; max_stack the deepest the
 * collapse's stack got (the reference's holds 64). */
size_t vro_collapse_wide(const vro_node* nodes, size_t num_nodes, int W, float* out_bounds, int64_t* out_children,
                         size_t* max_stack);
/* bvh_compressed_node<W>::init of each node, in the reference's byte layout
 * (W = 4: 64 bytes per node, W = 8: 112) */
void vro_compress_wide(const float* bounds, const int64_t* children, size_t num, int W, uint8_t* out);

#ifdef __cplusplus
}
#endif

#endif
