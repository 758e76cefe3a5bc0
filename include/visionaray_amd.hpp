// visionaray_amd.hpp -- header-only C++ facade over the C ABI
// (visionaray_amd.h) that keeps the reference's names for this path, so a
// caller of the reference's GPU path changes its namespace and include:
//
//   reference (include/visionaray/...)            here (namespace visionaray_amd)
//   hip::device_vector<T>   hip/device_vector.h    device_vector<T>
//   hip_index_bvh<P>        bvh.h:1100             hip_index_bvh
//   lbvh_builder::build     detail/bvh/lbvh.h:682  lbvh_builder::build
//   binned_sah_builder      detail/bvh/sah.h:176   binned_sah_builder (host build + upload)
//   bvh_refitter::refit     detail/bvh/refit.h:28  bvh_refitter::refit
//   closest_hit / any_hit   traverse_linear.inl    closest_hit / any_hit (ray batches)
//   tree.inst(mat4x3) + lbvh_builder over index_bvh<bvh_inst>  bvh.h:679-782
//                                                  bvh_inst + instanced_bvh (top level built on the host)
//   pinhole_camera          pinhole_camera.h       pinhole_camera
//   make_kernel_params      kernels.h:87           make_kernel_params
//   pathtracing::kernel     detail/pathtracing.inl pathtracing::kernel
//   pixel_sampler::jittered_blend_type             pixel_sampler::jittered_blend_type
//   make_sched_params       scheduler.h:103        make_sched_params
//   cuda_sched<R>::frame    detail/cuda_sched.inl  hip_sched<R>::frame
//
// Errors throw visionaray_amd::error (the C ABI returns codes; the reference
// prints and continues).
#pragma once

#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "visionaray_amd.h"

namespace visionaray_amd {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc)
{
    if (rc != VR_OK) throw error(rc, vr_last_error());
}

using basic_triangle = vr_triangle;  // basic_triangle<3, float>
using bvh_node = vr_bvh_node;
using ray = vr_ray;                  // basic_ray<float>
using hit_record = vr_hit;

struct vec3 {
    float x, y, z;
};
struct vec4 {
    float x, y, z, w;
};

// Owning device buffer (hip::device_vector, hip/device_vector.h:22-76): move
// only, freed on destruction.
template <typename T>
class device_vector {
public:
    device_vector() = default;
    explicit device_vector(size_t n) { resize(n); }
    device_vector(const T* host, size_t n)
    {
        resize(n);
        check(vr_memcpy_htod(data_, host, n * sizeof(T)));
    }
    explicit device_vector(const std::vector<T>& host) : device_vector(host.data(), host.size()) {}
    device_vector(const device_vector&) = delete;
    device_vector& operator=(const device_vector&) = delete;
    device_vector(device_vector&& o) noexcept : data_(o.data_), size_(o.size_)
    {
        o.data_ = nullptr;
        o.size_ = 0;
    }
    device_vector& operator=(device_vector&& o) noexcept
    {
        std::swap(data_, o.data_);
        std::swap(size_, o.size_);
        return *this;
    }
    ~device_vector() { vr_free(data_); }

    void resize(size_t n)
    {
        vr_free(data_);
        data_ = nullptr;
        size_ = 0;
        void* p = nullptr;
        check(vr_malloc(&p, n * sizeof(T)));
        data_ = static_cast<T*>(p);
        size_ = n;
    }
    T* data() { return data_; }
    const T* data() const { return data_; }
    size_t size() const { return size_; }
    std::vector<T> to_host() const
    {
        std::vector<T> h(size_);
        check(vr_memcpy_dtoh(h.data(), data_, size_ * sizeof(T)));
        return h;
    }

private:
    T* data_ = nullptr;
    size_t size_ = 0;
};

// A BVH in device memory (hip_index_bvh<P>, bvh.h:1100).
class hip_index_bvh {
public:
    hip_index_bvh() = default;
    explicit hip_index_bvh(vr_bvh h) : h_(h) {}
    hip_index_bvh(const hip_index_bvh&) = delete;
    hip_index_bvh& operator=(const hip_index_bvh&) = delete;
    hip_index_bvh(hip_index_bvh&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    hip_index_bvh& operator=(hip_index_bvh&& o) noexcept
    {
        std::swap(h_, o.h_);
        return *this;
    }
    ~hip_index_bvh() { vr_bvh_destroy(h_); }

    vr_bvh handle() const { return h_; }
    vr_bvh_info info() const
    {
        vr_bvh_info i{};
        check(vr_bvh_get_info(h_, &i));
        return i;
    }
    // tree.ref(): the device view kernels take (visionaray_amd_device.hpp)
    vr_bvh_ref ref() const
    {
        vr_bvh_ref r{};
        check(vr_bvh_get_ref(h_, &r));
        return r;
    }
    size_t num_primitives() const { return info().num_prims; }
    size_t num_nodes() const { return info().num_nodes; }
    // tree.nodes() / tree.indices() copied to the host
    std::vector<bvh_node> nodes() const
    {
        device_vector<bvh_node> d(num_nodes());
        check(vr_bvh_nodes(h_, d.data(), nullptr));
        return d.to_host();
    }
    std::vector<unsigned> indices() const
    {
        device_vector<unsigned> d(num_primitives());
        check(vr_bvh_indices(h_, d.data(), nullptr));
        return d.to_host();
    }

private:
    vr_bvh h_ = nullptr;
};

// GPU LBVH builder (lbvh_builder::build, detail/bvh/lbvh.h:681-844).
struct lbvh_builder {
    hip_index_bvh build(const hip_index_bvh& /* tree type tag */, const basic_triangle* d_prims, size_t num_prims,
                        void* stream = nullptr, bool keep_nodes = true)
    {
        vr_bvh h = nullptr;
        check(vr_lbvh_build(d_prims, num_prims, keep_nodes ? VR_BVH_KEEP_NODES : 0u, stream, &h));
        return hip_index_bvh(h);
    }
};

// Host binned SAH build uploaded for device traversal (binned_sah_builder,
// detail/bvh/sah.h:176-807; spatial splits are not provided).
struct binned_sah_builder {
    void enable_spatial_splits(bool enable)
    {
        if (enable) throw error(VR_ERR_UNSUPPORTED, "spatial splits are not provided");
    }
    hip_index_bvh build(const hip_index_bvh& /* tag */, const basic_triangle* h_prims, size_t num_prims,
                        int max_leaf_size = -1)
    {
        std::vector<bvh_node> nodes(num_prims ? 2 * num_prims - 1 : 1);
        std::vector<uint32_t> idx(num_prims ? num_prims : 1);
        size_t nn = 0;
        check(vr_sah_build(h_prims, num_prims, max_leaf_size, nodes.data(), &nn, idx.data()));
        vr_bvh h = nullptr;
        check(vr_bvh_from_nodes(h_prims, num_prims, nodes.data(), nn, idx.data(), &h));
        return hip_index_bvh(h);
    }
};

// The reference's CPU lbvh_builder tree (top-down, leaf size 4) built on the
// host and uploaded (lbvh_builder::build on an index_bvh, detail/bvh/lbvh.h:534-656).
struct host_lbvh_builder {
    hip_index_bvh build(const hip_index_bvh& /* tag */, const basic_triangle* h_prims, size_t num_prims,
                        int max_leaf_size = -1)
    {
        std::vector<bvh_node> nodes(num_prims ? 2 * num_prims - 1 : 1);
        std::vector<uint32_t> idx(num_prims ? num_prims : 1);
        size_t nn = 0;
        check(vr_lbvh_build_host(h_prims, num_prims, max_leaf_size, nodes.data(), &nn, idx.data()));
        vr_bvh h = nullptr;
        check(vr_bvh_from_nodes(h_prims, num_prims, nodes.data(), nn, idx.data(), &h));
        return hip_index_bvh(h);
    }
};

// bvh_refitter::refit (detail/bvh/refit.h:28-114) for LBVHs on the device.
struct bvh_refitter {
    void refit(hip_index_bvh& tree, const basic_triangle* d_prims, void* stream = nullptr)
    {
        check(vr_bvh_refit(tree.handle(), d_prims, stream));
    }
};

// closest_hit / any_hit over ray batches (detail/traverse_linear.inl:149-214)
inline void closest_hit(const hip_index_bvh& b, const ray* d_rays, size_t n, hit_record* d_hits,
                        void* stream = nullptr)
{
    check(vr_trace(b.handle(), 0, d_rays, n, d_hits, stream));
}
inline void any_hit(const hip_index_bvh& b, const ray* d_rays, size_t n, hit_record* d_hits, void* stream = nullptr)
{
    check(vr_trace(b.handle(), 1, d_rays, n, d_hits, stream));
}

// One instance of a built BVH (index_bvh_inst_t, bvh.h:679-782): the
// mat4x3 as its 12 floats column by column (mat4x3(T const[12])); set_inst_id
// as in the reference.  The top level over instances (the viewer's
// lbvh_builder over index_bvh<bvh_inst>, src/viewer/viewer.cpp:1534-1560) is
// instanced_bvh; the instanced BVHs must outlive it.
using inst_hit_record = vr_inst_hit;

struct bvh_inst {
    vr_instance rec{};
    bvh_inst(uint32_t blas_index, const float mat4x3[12])
    {
        rec.blas = blas_index;
        rec.inst_id = -1;  // ~0u, as index_bvh_inst_t starts
        for (int k = 0; k < 12; ++k) rec.transform[k] = mat4x3[k];
    }
    void set_inst_id(int id) { rec.inst_id = id; }
    int get_inst_id() const { return rec.inst_id; }
};

class instanced_bvh {
public:
    instanced_bvh(const std::vector<const hip_index_bvh*>& blas, const std::vector<bvh_inst>& instances,
                  void* stream = nullptr)
    {
        std::vector<vr_bvh> hs;
        for (const hip_index_bvh* b : blas) hs.push_back(b->handle());
        std::vector<vr_instance> recs;
        for (const bvh_inst& i : instances) recs.push_back(i.rec);
        check(vr_tlas_build(hs.data(), (uint32_t)hs.size(), recs.data(), (uint32_t)recs.size(), stream, &h_));
    }
    instanced_bvh(const instanced_bvh&) = delete;
    instanced_bvh& operator=(const instanced_bvh&) = delete;
    ~instanced_bvh() { vr_tlas_destroy(h_); }
    vr_tlas handle() const { return h_; }
    // the top level's reference-layout nodes and instance order
    std::pair<std::vector<bvh_node>, std::vector<uint32_t>> nodes() const
    {
        size_t n = 0;
        check(vr_tlas_nodes(h_, nullptr, nullptr, &n));
        std::vector<bvh_node> nodes(n);
        std::vector<uint32_t> idx(n ? (n + 1) / 2 : 0);
        check(vr_tlas_nodes(h_, nodes.data(), idx.data(), &n));
        return {nodes, idx};
    }

private:
    vr_tlas h_ = nullptr;
};

// closest_hit / any_hit over the top level (detail/bvh/intersect.inl:362-394)
inline void closest_hit(const instanced_bvh& b, const ray* d_rays, size_t n, inst_hit_record* d_hits,
                        void* stream = nullptr)
{
    check(vr_tlas_closest_hit(b.handle(), d_rays, n, d_hits, stream));
}
inline void any_hit(const instanced_bvh& b, const ray* d_rays, size_t n, inst_hit_record* d_hits,
                    void* stream = nullptr)
{
    check(vr_tlas_any_hit(b.handle(), d_rays, n, d_hits, stream));
}

// pinhole_camera (pinhole_camera.h:20-116): perspective + look_at, then
// begin_frame computes the U, V, W basis the kernels use.
class pinhole_camera {
public:
    void perspective(float fovy, float aspect, float /*z_near*/, float /*z_far*/)
    {
        fovy_ = fovy;
        aspect_ = aspect;
    }
    void look_at(vec3 eye, vec3 center, vec3 up = vec3{0.0f, 1.0f, 0.0f})
    {
        eye_ = eye;
        center_ = center;
        up_ = up;
    }
    void begin_frame()
    {
        const float e[3] = {eye_.x, eye_.y, eye_.z}, c[3] = {center_.x, center_.y, center_.z},
                    u[3] = {up_.x, up_.y, up_.z};
        check(vr_camera_look_at(e, c, u, fovy_, aspect_, &basis_));
    }
    void end_frame() {}
    const vr_camera& basis() const { return basis_; }

private:
    vec3 eye_{0, 0, 0}, center_{0, 0, -1}, up_{0, 1, 0};
    float fovy_ = 0.785398f, aspect_ = 1.0f;
    vr_camera basis_{};
};

namespace pixel_sampler {
// jittered_blend_type (pixel_sampler_types.h:51-69)
struct jittered_blend_type {
    unsigned spp = 1;
    float sfactor = 1.0f;
    float dfactor = 0.0f;
};
}  // namespace pixel_sampler

// kernel_params without lights (kernels.h:38-132): the BVH, materials
// indexed by geom_id, bounces, epsilon, background and ambient (rgb, kl).
struct kernel_params {
    const hip_index_bvh* bvh;
    const vr_material* materials;
    size_t num_materials;
    vr_kernel_params p;
};

inline kernel_params make_kernel_params(const hip_index_bvh& bvh, const vr_material* d_materials,
                                        size_t num_materials, unsigned num_bounces = 5,
                                        float epsilon = FLT_EPSILON, vec4 bg_color = vec4{0, 0, 0, 0},
                                        vec4 ambient_color = vec4{0, 0, 0, 0})
{
    kernel_params k{&bvh, d_materials, num_materials, {}};
    k.p.num_bounces = num_bounces;
    k.p.epsilon = epsilon;
    k.p.bg[0] = bg_color.x * bg_color.w;
    k.p.bg[1] = bg_color.y * bg_color.w;
    k.p.bg[2] = bg_color.z * bg_color.w;
    k.p.ambient[0] = ambient_color.x * ambient_color.w;
    k.p.ambient[1] = ambient_color.y * ambient_color.w;
    k.p.ambient[2] = ambient_color.z * ambient_color.w;
    return k;
}

namespace pathtracing {
// pathtracing::kernel (detail/pathtracing.inl:33-244)
template <typename Params = kernel_params>
struct kernel {
    Params params;
};
}  // namespace pathtracing

// The AO kernel of src/examples/ao/main.cpp:192-260.
struct ao_kernel {
    const hip_index_bvh* bvh;
    unsigned samples = 8;
    float radius = 0.1f;
    vec3 bg{0, 0, 0};
};

// Device RGBA32F color + accum render target (render_target_ref,
// render_target.h:44-113).
struct render_target_ref {
    float* color;
    float* accum;
    int width;
    int height;
};

struct sched_params {
    pixel_sampler::jittered_blend_type sample_params;
    pinhole_camera cam;
    render_target_ref rt;
};

inline sched_params make_sched_params(pixel_sampler::jittered_blend_type sp, const pinhole_camera& cam,
                                      render_target_ref rt)
{
    return sched_params{sp, cam, rt};
}

// hip_sched<R>::frame (cuda_sched<R>::frame, detail/cuda_sched.inl:226-246):
// begin_frame on the camera, one asynchronous launch on `stream`, then the
// frame counter advances (it seeds make_random_seed(y * w + x, frame_id)).
template <typename R = ray>
class hip_sched {
public:
    void frame(const pathtracing::kernel<kernel_params>& k, sched_params sp, void* stream = nullptr,
               uint64_t* d_ray_counts = nullptr)
    {
        sp.cam.begin_frame();
        const vr_sched_params s = params(sp);
        check(vr_render_pt(k.params.bvh->handle(), k.params.materials, k.params.num_materials, &k.params.p,
                           &sp.cam.basis(), &s, sp.rt.accum, sp.rt.color, d_ray_counts, stream));
        ++frame_id_;
    }
    void frame(const ao_kernel& k, sched_params sp, void* stream = nullptr)
    {
        sp.cam.begin_frame();
        const vr_sched_params s = params(sp);
        const float bg[3] = {k.bg.x, k.bg.y, k.bg.z};
        check(vr_render_ao(k.bvh->handle(), &sp.cam.basis(), &s, k.samples, k.radius, bg, sp.rt.accum, sp.rt.color,
                           stream));
        ++frame_id_;
    }
    // Any kernel functor (a user kernel): one thread per pixel under the
    // jittered-blend sampler.  Defined in visionaray_amd_device.hpp (HIP code).
    template <typename K>
    void frame(K kernel, sched_params sp, void* stream = nullptr);

    unsigned frame_id() const { return frame_id_; }

    // Path-replay adjoint of the last frame (no reference counterpart):
    // d_grad[6 m + 0..2] += d loss / d cd, d_grad[6 m + 3..5] += d loss / d ce.
    void backward(const pathtracing::kernel<kernel_params>& k, sched_params sp, const float* d_dl_dimage,
                  double* d_grad, void* stream = nullptr)
    {
        sp.cam.begin_frame();
        vr_sched_params s = params(sp);
        s.frame_id = frame_id_ ? frame_id_ - 1 : 0;
        check(vr_render_pt_backward(k.params.bvh->handle(), k.params.materials, k.params.num_materials, &k.params.p,
                                    &sp.cam.basis(), &s, d_dl_dimage, d_grad, nullptr, stream));
    }

    // Pick the fastest launch shape for this frame (vr_autotune_pt).
    vr_launch_shape autotune(const pathtracing::kernel<kernel_params>& k, sched_params sp, int reps = 2,
                             void* stream = nullptr)
    {
        sp.cam.begin_frame();
        const vr_sched_params s = params(sp);
        vr_launch_shape best{};
        float ms = 0.0f;
        check(vr_autotune_pt(k.params.bvh->handle(), k.params.materials, k.params.num_materials, &k.params.p,
                             &sp.cam.basis(), &s, sp.rt.accum, sp.rt.color, reps, &best, &ms, stream));
        return best;
    }

private:
    vr_sched_params params(const sched_params& sp) const
    {
        vr_sched_params s;
        s.width = (uint32_t)sp.rt.width;
        s.height = (uint32_t)sp.rt.height;
        s.spp = sp.sample_params.spp;
        s.frame_id = frame_id_;
        s.sfactor = sp.sample_params.sfactor;
        s.dfactor = sp.sample_params.dfactor;
        return s;
    }
    unsigned frame_id_ = 0;
};

// Many scenes in one persistent launch (vr_batch_*), optionally replayed as a
// captured HIP graph; no reference counterpart (one cuda_sched::frame per scene).
class scene_batch {
public:
    scene_batch(const std::vector<vr_scene>& scenes) { check(vr_batch_create(scenes.data(), scenes.size(), &b_)); }
    scene_batch(const scene_batch&) = delete;
    scene_batch& operator=(const scene_batch&) = delete;
    ~scene_batch() { vr_batch_destroy(b_); }
    void render(uint64_t* d_ray_counts = nullptr, void* stream = nullptr)
    {
        check(vr_batch_render(b_, d_ray_counts, stream));
    }
    void backward(double* d_grad, void* stream = nullptr) { check(vr_batch_backward(b_, d_grad, nullptr, stream)); }
    void capture(void* stream, uint64_t* d_ray_counts = nullptr)
    {
        check(vr_batch_capture_graph(b_, d_ray_counts, stream));
    }
    void replay(void* stream) { check(vr_batch_launch_graph(b_, stream)); }

private:
    vr_batch b_ = nullptr;
};

}  // namespace visionaray_amd
