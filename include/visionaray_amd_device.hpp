// visionaray_amd_device.hpp -- the reference's kernel-level operator surface
// for user kernels: HIP code for gfx950, compiled with hipcc
// (--offload-arch=gfx950 -ffp-contract=off, as the library is).
//
//   reference (include/visionaray/...)                  here (namespace visionaray_amd)
//   basic_ray<float>                math/ray.h           basic_ray
//   tree.ref() / index_bvh_ref_t    bvh.h:495-569,919    hip_index_bvh::ref() -> bvh_ref
//   intersect(ray, aabb, inv_dir)   math/intersect.h:44  intersect(basic_ray, aabb, vec3)
//   intersect(ray, triangle)        math/intersect.h:176 intersect(basic_ray, triangle)
//   closest_hit / any_hit (one ray) traverse_linear.inl:149-214
//                                                        closest_hit / any_hit (device)
//   basic_intersector<Derived>      intersector.h:24-82  basic_intersector<Derived>
//   random_generator<float>         random_generator.h   random_generator
//   cosine_sample_hemisphere        sampling.h:117-127   cosine_sample_hemisphere
//   make_orthonormal_basis          vector3.inl:356-366  make_orthonormal_basis
//   get_normal(normals, hit_rec)    get_normal.h         get_normal(bvh_ref, hit_rec)
//   result_record<S>                result_record.h      result_record
//   sched.frame(kernel, sparams)    cuda_sched.inl:32-246
//                                                        hip_sched<R>::frame(kernel, sparams, stream)
//
// A kernel functor is called once per sample as the reference's
// invoke_kernel does (detail/sched_common.h:35-90): kernel(ray, gen, x, y),
// kernel(ray, x, y), kernel(ray, gen) or kernel(ray), and returns a
// result_record.  The scheduler runs it under the jittered-blend sampler
// (sched_common.h:427-526): per-pixel seed make_random_seed(y * w + x,
// frame_id), jittered primary rays, the mean of spp samples blended into
// accum (color * sfactor + accum * dfactor) and copied to color.
//
// Traversal here is the library's own code (visionaray_amd_core.h) in an
// order that depends on the ray and the tree alone (TRAV_LANE), so a user
// kernel's hits, and the counts a user intersector takes, do not depend on
// which rays share its wave.  Its stack keeps 16 entries per thread in
// dynamic LDS: hip_sched::frame provides it; a kernel the caller launches
// itself must use at most traversal_block threads per block and pass
// traversal_lds_bytes of dynamic shared memory (and use none of its own).
//
// Random numbers are the reference host build's stream (minstd_rand0 through
// generate_canonical<float, 24>).  The reference writes
// f(gen.next(), gen.next()) in several places; C++ leaves the order of the two
// draws unspecified, and the reference's host compiler (GCC) draws the right
// argument first while hipcc draws the left one first.  Kernels that must
// reproduce the reference draw into named variables (right argument first).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "visionaray_amd.hpp"
#include "visionaray_amd_core.h"

namespace visionaray_amd {

// ---------------------------------------------------------------------------
// vectors: component-wise arithmetic in the reference's operation order
// (math/detail/vector3.inl)
// ---------------------------------------------------------------------------

__host__ __device__ inline vr::f3 to_f3(vec3 v) { return vr::mk3(v.x, v.y, v.z); }
__host__ __device__ inline vec3 to_vec3(vr::f3 v) { return vec3{v.x, v.y, v.z}; }
__host__ __device__ inline vec3 operator+(vec3 a, vec3 b) { return to_vec3(vr::add3(to_f3(a), to_f3(b))); }
__host__ __device__ inline vec3 operator-(vec3 a, vec3 b) { return to_vec3(vr::sub3(to_f3(a), to_f3(b))); }
__host__ __device__ inline vec3 operator*(vec3 a, vec3 b) { return to_vec3(vr::mul3(to_f3(a), to_f3(b))); }
__host__ __device__ inline vec3 operator*(vec3 a, float s) { return to_vec3(vr::muls3(to_f3(a), s)); }
__host__ __device__ inline vec3 operator*(float s, vec3 a) { return to_vec3(vr::smul3(s, to_f3(a))); }
__host__ __device__ inline vec3 operator-(vec3 a) { return to_vec3(vr::neg3(to_f3(a))); }
__host__ __device__ inline float dot(vec3 a, vec3 b) { return vr::dot3(to_f3(a), to_f3(b)); }
__host__ __device__ inline vec3 cross(vec3 a, vec3 b) { return to_vec3(vr::cross3(to_f3(a), to_f3(b))); }
__host__ __device__ inline vec3 normalize(vec3 a) { return to_vec3(vr::normalize3(to_f3(a))); }
__host__ __device__ inline vec3 faceforward(vec3 n, vec3 i, vec3 nref)
{
    return to_vec3(vr::faceforward3(to_f3(n), to_f3(i), to_f3(nref)));
}

// ---------------------------------------------------------------------------
// rays, boxes, triangles and their hit records
// ---------------------------------------------------------------------------

// basic_ray<float> (math/ray.h); the same 32 bytes as vr_ray
struct basic_ray {
    vec3 ori;
    vec3 dir;
    float tmin;
    float tmax;
};
static_assert(sizeof(basic_ray) == sizeof(vr_ray), "basic_ray is a vr_ray");

struct aabb {
    vec3 min;
    vec3 max;
};

// hit_record<basic_ray<float>, basic_aabb<float>> (math/intersect.h:30-41)
struct hit_record_aabb {
    bool hit = false;
    float tnear = FLT_MAX;
    float tfar = -FLT_MAX;
};

// basic_triangle<3, float> as the traversal holds it (math/triangle.h)
struct triangle {
    vec3 v1, e1, e2;
    int prim_id;
    int geom_id;
};

// hit_record<basic_ray<float>, primitive<unsigned>> (math/intersect.h:151-169)
struct hit_record_tri {
    bool hit = false;
    float t = FLT_MAX;
    float u = 0.0f;
    float v = 0.0f;
    int prim_id = -1;
    int geom_id = -1;
};

// hit_record_bvh (detail/bvh/hit_record.h:26-35): the triangle's position in
// the tree's primitive list (tree.indices()[primitive_list_index] is its
// index in the caller's array)
struct hit_record_bvh : hit_record_tri {
    unsigned primitive_list_index = 0;
    vec3 isect_pos{0.0f, 0.0f, 0.0f};
};

// intersect(ray, aabb, inv_dir) (math/intersect.h:44-62): slab distances,
// tnear = min_max(...), tfar = max_min(...), hit iff tfar >= tnear
__device__ inline hit_record_aabb intersect(const basic_ray& r, const aabb& b, vec3 inv_dir)
{
    const vec3 t1 = (b.min - r.ori) * inv_dir;
    const vec3 t2 = (b.max - r.ori) * inv_dir;
    hit_record_aabb h;
    h.tnear = vr::tmax_f(vr::tmin_f(t1.x, t2.x), vr::tmax_f(vr::tmin_f(t1.y, t2.y), vr::tmin_f(t1.z, t2.z)));
    h.tfar = vr::tmin_f(vr::tmax_f(t1.x, t2.x), vr::tmin_f(vr::tmax_f(t1.y, t2.y), vr::tmax_f(t1.z, t2.z)));
    h.hit = h.tfar >= h.tnear;
    return h;
}

// intersect(ray, basic_triangle) (math/intersect.h:176-229): Moller-Trumbore,
// two-sided, inclusive edges
__device__ inline hit_record_tri intersect(const basic_ray& r, const triangle& tri)
{
    hit_record_tri h;
    float t, u, v;
    if (vr::intersect_tri(to_f3(r.ori), to_f3(r.dir), to_f3(tri.v1), to_f3(tri.e1), to_f3(tri.e2), t, u, v)) {
        h.hit = true;
        h.t = t;
        h.u = u;
        h.v = v;
        h.prim_id = tri.prim_id;
        h.geom_id = tri.geom_id;
    }
    return h;
}

// ---------------------------------------------------------------------------
// custom intersectors (intersector.h:24-82): a type derived from
// basic_intersector<Derived> overrides the box and / or the triangle test
// (e.g. to count them, src/viewer/bvh_costs.h:20-45) and is passed to
// closest_hit / any_hit.  The traversal calls it once per child box of every
// node it visits (four per 4-wide node; an unused slot is the empty box
// lo = hi = +inf, which no ray enters) and once per triangle it tests.
// ---------------------------------------------------------------------------

template <typename Derived>
struct basic_intersector {
    __device__ hit_record_aabb operator()(const basic_ray& r, const aabb& box, vec3 inv_dir)
    {
        return intersect(r, box, inv_dir);
    }
    __device__ hit_record_tri operator()(const basic_ray& r, const triangle& tri) { return intersect(r, tri); }
};

struct default_intersector : basic_intersector<default_intersector> {
};

namespace detail {

// the traversal's box / triangle policy over a user intersector: a child is
// entered when the intersector reports a hit, max(tnear, tmin) <= min(tfar,
// tmax), and max(tnear, tmin) is below the closest hit so far -- with the
// default intersector exactly the library's own slab decision
template <typename I>
struct isect_adapter {
    static constexpr bool near_far = false;  // the user sees whole boxes
    I* user;
    __device__ float box(float lx, float ly, float lz, float hx, float hy, float hz, const vr::trav_state& T,
                         float tclosest) const
    {
        const basic_ray r{to_vec3(T.ori), to_vec3(T.dir), T.tmin, T.tmax};
        const hit_record_aabb h = (*user)(r, aabb{vec3{lx, ly, lz}, vec3{hx, hy, hz}}, to_vec3(T.inv));
        const float tn = vr::tmax_f(h.tnear, T.tmin);
        const float tf = vr::tmin_f(h.tfar, T.tmax);
        return (h.hit && tf >= tn && tn < tclosest) ? tn : __builtin_inff();
    }
    __device__ bool tri(const vr::trav_state& T, float4 a, float4 b, float4 c, float& t, float& u, float& v) const
    {
        const basic_ray r{to_vec3(T.ori), to_vec3(T.dir), T.tmin, T.tmax};
        const triangle tr{vec3{a.x, a.y, a.z}, vec3{b.x, b.y, b.z}, vec3{c.x, c.y, c.z}, __float_as_int(a.w),
                          __float_as_int(b.w)};
        const hit_record_tri h = (*user)(r, tr);
        if (!h.hit) return false;
        t = h.t;
        u = h.u;
        v = h.v;
        return true;
    }
};

}  // namespace detail

// ---------------------------------------------------------------------------
// BVH references and traversal from device code
// ---------------------------------------------------------------------------

// a BVH as device code sees it (index_bvh_ref_t, bvh.h:495-569): non-owning,
// valid while its hip_index_bvh lives and is not rebuilt
struct bvh_ref {
    vr::bvh_view view;
    bvh_ref() = default;
    bvh_ref(const vr_bvh_ref& r)  // NOLINT: implicit, as tree.ref() converts
    {
        view.quads = static_cast<const vr::quad_node*>(r.nodes);
        view.tris = static_cast<const vr::tri_node*>(r.prims);
        view.pairs = static_cast<const vr::pair_node*>(r.pairs);
        view.num_prims = r.num_prims;
        view.root = r.root;
        view.num_quads = r.num_nodes;
        view.wide = nullptr;  // user kernels traverse the 4-wide nodes
    }
    // tree.primitive(i) for a hit's primitive_list_index
    __device__ triangle primitive(unsigned i) const
    {
        const vr::tri_node& t = view.tris[i];
        return triangle{vec3{t.a.x, t.a.y, t.a.z}, vec3{t.b.x, t.b.y, t.b.z}, vec3{t.c.x, t.c.y, t.c.z},
                        __float_as_int(t.a.w), __float_as_int(t.b.w)};
    }
};

// threads per block and dynamic LDS bytes a kernel calling closest_hit /
// any_hit needs (the traversal stack's LDS part: vr::LDS_STACK entries per thread)
constexpr int traversal_block = 256;
constexpr size_t traversal_lds_bytes = (size_t)vr::LDS_STACK * traversal_block * sizeof(uint32_t);

namespace detail {

__device__ inline uint32_t* lane_stack()
{
    extern __shared__ uint32_t vr_traversal_lds[];
    return vr_traversal_lds + (threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
}

__device__ inline hit_record_bvh make_hit(const bvh_ref& b, const vr::hit_t& h)
{
    hit_record_bvh r;
    if (h.tri < 0) return r;
    const vr::tri_node& t = b.view.tris[h.tri];
    r.hit = true;
    r.t = h.t;
    r.u = h.u;
    r.v = h.v;
    r.prim_id = __float_as_int(t.a.w);
    r.geom_id = __float_as_int(t.b.w);
    r.primitive_list_index = (unsigned)h.tri;
    return r;
}

template <bool ANY, typename I>
__device__ inline hit_record_bvh trace_one(const basic_ray& r, const bvh_ref& b, I& isect)
{
    const isect_adapter<I> ad{&isect};
    const vr::hit_t h = vr::traverse<ANY, false, vr::LDS_STACK, traversal_block, vr::TRAV_LANE>(
        b.view, to_f3(r.ori), to_f3(r.dir), r.tmin, r.tmax, lane_stack(), nullptr, ad);
    return make_hit(b, h);
}

}  // namespace detail

// closest_hit(ray, bvh[, intersector]) (traverse_linear.inl:149-214, one BVH)
template <typename I>
__device__ inline hit_record_bvh closest_hit(const basic_ray& r, const bvh_ref& b, I& isect)
{
    return detail::trace_one<false>(r, b, isect);
}
__device__ inline hit_record_bvh closest_hit(const basic_ray& r, const bvh_ref& b)
{
    const vr::hit_t h = vr::traverse<false, false, vr::LDS_STACK, traversal_block, vr::TRAV_LANE>(
        b.view, to_f3(r.ori), to_f3(r.dir), r.tmin, r.tmax, detail::lane_stack(), nullptr);
    return detail::make_hit(b, h);
}

// any_hit(ray, bvh[, intersector]): the first accepted hit in [tmin, tmax]
template <typename I>
__device__ inline hit_record_bvh any_hit(const basic_ray& r, const bvh_ref& b, I& isect)
{
    return detail::trace_one<true>(r, b, isect);
}
__device__ inline hit_record_bvh any_hit(const basic_ray& r, const bvh_ref& b)
{
    const vr::hit_t h = vr::traverse<true, false, vr::LDS_STACK, traversal_block, vr::TRAV_LANE>(
        b.view, to_f3(r.ori), to_f3(r.dir), r.tmin, r.tmax, detail::lane_stack(), nullptr);
    return detail::make_hit(b, h);
}

// get_normal(geometric normals, hit_rec) of the BVH's triangles
// (get_normal.h): normalize(cross(e1, e2)) of the hit triangle
__device__ inline vec3 get_normal(const bvh_ref& b, const hit_record_bvh& h)
{
    const triangle t = b.primitive(h.primitive_list_index);
    return normalize(cross(t.e1, t.e2));
}

// ---------------------------------------------------------------------------
// sampling
// ---------------------------------------------------------------------------

// random_generator<float> of the reference's host build (random_generator.h:29-62)
struct random_generator {
    vr::rng_t g;
    __device__ explicit random_generator(unsigned seed) { vr::rng_seed(g, seed); }
    __device__ float next() { return vr::rng_next(g); }
};

__device__ inline unsigned make_random_seed(int v0, int v1) { return vr::make_random_seed(v0, v1); }

__device__ inline vec3 cosine_sample_hemisphere(float u1, float u2) { return to_vec3(vr::cosine_hemisphere(u1, u2)); }

// make_orthonormal_basis(u, v, w) (math/detail/vector3.inl:356-366)
__device__ inline void make_orthonormal_basis(vec3& u, vec3& v, vec3 w)
{
    vr::f3 uu, vv;
    vr::onb(to_f3(w), uu, vv);
    u = to_vec3(uu);
    v = to_vec3(vv);
}

// result_record<float> (result_record.h)
struct result_record {
    bool hit = false;
    vec4 color{0.0f, 0.0f, 0.0f, 0.0f};
    float depth = FLT_MAX;
};

// ---------------------------------------------------------------------------
// hip_sched<R>::frame for any kernel functor (cuda_sched.inl:32-246)
// ---------------------------------------------------------------------------

namespace detail {

// invoke_kernel (detail/sched_common.h:35-90): the first signature the
// functor accepts, found by overload resolution on device declarations
// (ranked by the tag's derivation depth)
struct pick0 {};
struct pick1 : pick0 {};
struct pick2 : pick1 {};
struct pick3 : pick2 {};
template <typename K>
__device__ inline auto invoke_as(K& k, const basic_ray& r, random_generator& gen, int x, int y, pick3)
    -> decltype(k(r, gen, x, y))
{
    return k(r, gen, x, y);
}
template <typename K>
__device__ inline auto invoke_as(K& k, const basic_ray& r, random_generator&, int x, int y, pick2)
    -> decltype(k(r, x, y))
{
    return k(r, x, y);
}
template <typename K>
__device__ inline auto invoke_as(K& k, const basic_ray& r, random_generator& gen, int, int, pick1)
    -> decltype(k(r, gen))
{
    return k(r, gen);
}
template <typename K>
__device__ inline auto invoke_as(K& k, const basic_ray& r, random_generator&, int, int, pick0) -> decltype(k(r))
{
    return k(r);
}
template <typename K>
__device__ inline result_record invoke_kernel(K& k, const basic_ray& r, random_generator& gen, int x, int y)
{
    return invoke_as(k, r, gen, x, y, pick3{});
}

// one thread per pixel (scanline order), spp samples, jittered-blend store
template <typename K>
__global__ __launch_bounds__(traversal_block) void k_user_frame(K kernel, vr_camera cam, vr_sched_params sp,
                                                                float* accum, float* color)
{
    const uint32_t w = sp.width, h = sp.height;
    const uint32_t px = blockIdx.x * (uint32_t)traversal_block + threadIdx.x;
    if (px >= w * h) return;
    const int x = (int)(px % w), y = (int)(px / w);
    random_generator gen(vr::make_random_seed(y * (int)w + x, (int32_t)sp.frame_id));
    float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t s = 0; s < sp.spp; ++s) {
        vr::f3 o, d;
        vr::primary_ray(cam, gen.g, vr::REFERENCE_ARG_ORDER, x, y, (int)w, (int)h, o, d);
        const basic_ray r{to_vec3(o), to_vec3(d), 0.0f, FLT_MAX};
        const result_record res = invoke_kernel(kernel, r, gen, x, y);
        sum.x += res.color.x;
        sum.y += res.color.y;
        sum.z += res.color.z;
        sum.w += res.color.w;
    }
    const float spp = (float)sp.spp;
    sum.x = sum.x / spp;
    sum.y = sum.y / spp;
    sum.z = sum.z / spp;
    sum.w = sum.w / spp;
    float4* acc = reinterpret_cast<float4*>(accum) + px;
    const float4 old = *acc;
    const float sf = sp.sfactor, df = sp.dfactor;
    const float4 out = make_float4(sum.x * sf + old.x * df, sum.y * sf + old.y * df, sum.z * sf + old.z * df,
                                   sum.w * sf + old.w * df);
    *acc = out;
    reinterpret_cast<float4*>(color)[px] = out;
}

}  // namespace detail

template <typename R>
template <typename K>
void hip_sched<R>::frame(K kernel, sched_params sp, void* stream)
{
    sp.cam.begin_frame();
    const vr_sched_params s = params(sp);
    const size_t npix = (size_t)s.width * s.height;
    if (npix >= (1ull << 31)) throw error(VR_ERR_INVALID_ARGUMENT, "hip_sched::frame: image too large");
    if (npix > 0) {
        const unsigned blocks = (unsigned)((npix + traversal_block - 1) / traversal_block);
        hipLaunchKernelGGL(detail::k_user_frame<K>, dim3(blocks), dim3(traversal_block), traversal_lds_bytes,
                           (hipStream_t)stream, kernel, sp.cam.basis(), s, sp.rt.accum, sp.rt.color);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw error(VR_ERR_HIP, std::string("hip_sched::frame: ") + hipGetErrorString(e));
    }
    ++frame_id_;
}

}  // namespace visionaray_amd
