// ref_driver.cpp -- runs the reference library's own CPU code on raw binary
// inputs, for golden vectors and the CPU baseline timing.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference headers where they lie (-I <reference>/include); the binary goes to
// oracle/_ref/ and is never part of the product.  Nothing from the reference is
// copied here: this file only instantiates the reference's templates.
//
// Usage: vr_ref <mode> <args...>; every array is a raw little-endian file.
//   morton  TRIS N OUT                  lbvh_builder::init -> sorted {code,id}
//   karras  CODES N OUT                 detail::lbvh::determine_range per inner node
//   trace   NODES NN TRIS NP IDX RAYS NR ANY OUT   intersect over index_bvh
//   brute   TRIS NP RAYS NR ANY OUT     closest_hit/any_hit over the plain list
//   seed    PAIRS N OUT                 make_random_seed
//   rng     SEED COUNT OUT              random_generator<float>::next stream
//   probe   SEED OUT                    argument evaluation order probes
//   camera  CAMIN OUT                   pinhole_camera basis after begin_frame
//   ao      NODES NN TRIS NP IDX CAMIN PARAMS OUT   AO example kernel, tiled_sched
//   pt      NODES NN TRIS NP IDX MATS NM CAMIN PARAMS OUT  pathtracing::kernel
//   ptl     NODES NN TRIS NP IDX MATS NM LIGHTS NL CAMIN PARAMS OUT
//                                       pathtracing::kernel with area lights (NEE + MIS)
//   sah     TRIS N OUTNODES OUTIDX      binned_sah_builder (no spatial splits)
//   sahcost NODES NN TRIS NP IDX OUT     sah_cost(tree) (statistics.h:38-68), default costs
//   rotate  NODES NN TRIS NP IDX OUTNODES OUTCOUNT   bvh_optimizer::optimize_tree_rotations
//   inst    TRIS NP BLASOF NB INST NI RAYS NR OUT OUTNODES OUTIDX
//                                       two-level BVH: a CPU lbvh_builder tree per bottom
//                                       level (triangles BLASOF[b]..BLASOF[b+1]), index_bvh_inst_t per
//                                       instance ({blas, inst_id, mat4x3 as 12 floats}), lbvh_builder over
//                                       the instances, closest_hit -> {prim_id, t, u, v, inst_id, geom_id}
//   wide    NODES NN TRIS NP IDX W OUTNODES OUTCOUNT OUTCOMP
//                                       bvh_collapser into index_bvh4/8 (W), the wide nodes
//                                       as stored, and bvh_compressor::compress of them
//   tracew  NODES NN TRIS NP IDX W COMP RAYS NR OUT
//                                       intersect_ray1_bvhN (COMP 0) or
//                                       intersect_ray1_bvhN_compressed (COMP 1) closest hits
//                                       over the collapsed tree
//   bench   ...                         CPU baseline timing (see below)

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <visionaray/math/math.h>
#include <visionaray/bvh.h>
#include <visionaray/area_light.h>
#include <visionaray/generic_material.h>
#include <visionaray/kernels.h>
#include <visionaray/material.h>
#include <visionaray/morton.h>
#include <visionaray/pinhole_camera.h>
#include <visionaray/random_generator.h>
#include <visionaray/make_random_seed.h>
#include <visionaray/render_target.h>
#include <visionaray/sampling.h>
#include <visionaray/scheduler.h>
#include <visionaray/traverse.h>
#include <visionaray/get_normal.h>
#include <visionaray/detail/bvh/collapse.h>
#include <visionaray/detail/bvh/compress.h>

// the reference's 8-wide traversal needs its 8-lane SIMD types (AVX)
#if defined(__AVX__)
#define WIDE8_TRAVERSAL 1
#else
#define WIDE8_TRAVERSAL 0
#endif
#include <visionaray/detail/bvh/optimize.h>
#include <visionaray/detail/bvh/statistics.h>
#include <visionaray/detail/thread_pool.h>

using namespace visionaray;

using tri_t = basic_triangle<3, float>;
static_assert(sizeof(tri_t) == 44, "triangle layout");

template <typename T>
static std::vector<T> load(const char* path, size_t count)
{
    std::vector<T> v(count);
    FILE* f = std::fopen(path, "rb");
    if (!f) { std::perror(path); std::exit(2); }
    if (count && std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fprintf(stderr, "short read %s\n", path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

template <typename T>
static void save(const char* path, const T* data, size_t count)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) { std::perror(path); std::exit(2); }
    if (count) std::fwrite(data, sizeof(T), count, f);
    std::fclose(f);
}

// camera input: eye[3] center[3] up[3] fovy aspect (11 floats)
struct cam_probe : pinhole_camera
{
    vec3 u() const { return U; }
    vec3 v() const { return V; }
    vec3 w() const { return W; }
};

static cam_probe load_camera(const char* path)
{
    auto c = load<float>(path, 11);
    cam_probe cam;
    cam.perspective(c[9], c[10], 0.001f, 1000.0f);
    cam.look_at(vec3(c[0], c[1], c[2]), vec3(c[3], c[4], c[5]), vec3(c[6], c[7], c[8]));
    return cam;
}

// render params: width height spp frame_id num_bounces epsilon bg[3] amb[3]
//                sfactor dfactor ao_samples ao_radius threads   (all as float)
struct rparams
{
    int w, h, spp, frame_id, bounces;
    float eps, bg[3], amb[3], sf, df;
    int ao_samples;
    float ao_radius;
    int threads;
};

static rparams load_params(const char* path)
{
    auto p = load<float>(path, 17);
    rparams r;
    r.w = (int)p[0]; r.h = (int)p[1]; r.spp = (int)p[2]; r.frame_id = (int)p[3]; r.bounces = (int)p[4];
    r.eps = p[5];
    for (int k = 0; k < 3; ++k) { r.bg[k] = p[6 + k]; r.amb[k] = p[9 + k]; }
    r.sf = p[12]; r.df = p[13]; r.ao_samples = (int)p[14]; r.ao_radius = p[15]; r.threads = (int)p[16];
    if (r.threads <= 0) r.threads = (int)std::thread::hardware_concurrency();
    return r;
}

// RGBA32F color + RGBA32F accum render target (simple_buffer_rt cannot carry
// an accum format, simple_buffer_rt.h:34,59).
struct accum_rt
{
    using ref_type = render_target_ref<PF_RGBA32F, PF_UNSPECIFIED, PF_RGBA32F>;
    int w, h;
    std::vector<vec4> color, accum;
    accum_rt(int w_, int h_) : w(w_), h(h_), color(size_t(w_) * h_, vec4(0.0f)), accum(size_t(w_) * h_, vec4(0.0f)) {}
    ref_type ref() { return ref_type{color.data(), nullptr, accum.data(), w, h}; }
    int width() const { return w; }
    int height() const { return h; }
    void begin_frame() {}
    void end_frame() {}
};

struct tree_in
{
    index_bvh<tri_t> tree;
};

static index_bvh<tri_t> load_tree(char** a)
{
    size_t nn = std::strtoull(a[1], nullptr, 10);
    size_t np = std::strtoull(a[3], nullptr, 10);
    auto tris = load<tri_t>(a[2], np);
    index_bvh<tri_t> t(tris.data(), np);
    auto nodes = load<bvh_node>(a[0], nn);
    t.nodes().resize(nn);
    std::copy(nodes.begin(), nodes.end(), t.nodes().begin());
    auto idx = load<unsigned>(a[4], np);
    std::copy(idx.begin(), idx.end(), t.indices().begin());
    return t;
}

struct hit_out { int32_t prim_id; float t, u, v; };

template <typename HR>
static hit_out to_out(HR const& h)
{
    hit_out o;
    o.prim_id = h.hit ? (int32_t)h.prim_id : -1;
    o.t = h.hit ? h.t : FLT_MAX;
    o.u = h.hit ? h.u : 0.0f;
    o.v = h.hit ? h.v : 0.0f;
    return o;
}

struct probe_cam
{
    template <typename R, typename T = typename R::scalar_type>
    R primary_ray(R /* */, T const& x, T const& y, T const& w, T const& h) const
    {
        R r;
        r.ori = vector<3, T>(x, y, w);
        r.dir = vector<3, T>(h, T(0.0), T(0.0));
        r.tmin = T(0.0);
        r.tmax = T(1.0);
        return r;
    }
};

// AO kernel: the lambda of src/examples/ao/main.cpp:192-260, with geometric
// normals normalize(cross(e1, e2)) indexed by prim_id.
template <typename R, typename Sched>
static void render_ao(Sched& sched, index_bvh<tri_t> const& tree, std::vector<vec3> const& normals,
                      cam_probe const& cam, rparams const& p, accum_rt& rt)
{
    using S = typename R::scalar_type;
    using V = vector<3, S>;
    using C = vector<4, S>;
    std::vector<index_bvh<tri_t>::bvh_ref> bvhs{tree.ref()};
    auto prims_begin = bvhs.data();
    auto prims_end = bvhs.data() + bvhs.size();
    vec3 bgcolor(p.bg[0], p.bg[1], p.bg[2]);
    int AO_Samples = p.ao_samples;
    float AO_Radius = p.ao_radius;
    auto blend_params = pixel_sampler::jittered_blend_type{};
    blend_params.spp = p.spp;
    blend_params.sfactor = p.sf;
    blend_params.dfactor = p.df;
    auto sparams = make_sched_params(blend_params, cam, rt);
    sched.frame([&](R ray, random_generator<S>& gen) -> result_record<S>
    {
        result_record<S> result;
        result.color = C(bgcolor, 1.0f);
        auto hit_rec = closest_hit(ray, prims_begin, prims_end);
        result.hit = hit_rec.hit;
        if (any(hit_rec.hit))
        {
            hit_rec.isect_pos = ray.ori + ray.dir * hit_rec.t;
            result.depth = hit_rec.t;
            auto n = get_normal(normals.data(), hit_rec, tri_t{});
            n = faceforward(n, -ray.dir, n);
            V u;
            V v;
            V w = n;
            make_orthonormal_basis(u, v, w);
            S radius = AO_Radius;
            S ao(0.0);
            for (int i = 0; i < AO_Samples; ++i)
            {
                auto sp = cosine_sample_hemisphere(gen.next(), gen.next());
                auto dir = normalize(sp.x * u + sp.y * v + sp.z * w);
                R ao_ray;
                ao_ray.ori = hit_rec.isect_pos + dir * S(1E-3f);
                ao_ray.dir = dir;
                ao_ray.tmin = S(0.0);
                ao_ray.tmax = radius;
                auto ao_rec = any_hit(ao_ray, prims_begin, prims_end);
                ao = select(ao_rec.hit, ao + S(1.0f / AO_Samples), ao);
            }
            ao = S(1.0) - ao;
            result.color = select(hit_rec.hit, C(ao, ao, ao, S(1.0)), result.color);
        }
        return result;
    }, sparams);
}

// material input: type cd[3] kd ce[3] ls (9 x 4 B)
using mat_t = generic_material<emissive<float>, matte<float>>;

static std::vector<mat_t> load_mats(const char* path, size_t nm)
{
    auto raw = load<float>(path, nm * 9);
    std::vector<mat_t> mats;
    mats.reserve(nm);
    for (size_t i = 0; i < nm; ++i) {
        const float* r = &raw[i * 9];
        int32_t type;
        std::memcpy(&type, r, 4);
        if (type == 1) {
            emissive<float> e;
            e.ce() = from_rgb(vec3(r[5], r[6], r[7]));
            e.ls() = r[8];
            mats.emplace_back(e);
        } else {
            matte<float> m;
            m.ca() = from_rgb(vec3(0.0f));
            m.ka() = 0.0f;
            m.cd() = from_rgb(vec3(r[1], r[2], r[3]));
            m.kd() = r[4];
            mats.emplace_back(m);
        }
    }
    return mats;
}

template <typename R, typename Sched>
static void render_pt(Sched& sched, index_bvh<tri_t> const& tree, std::vector<mat_t> const& mats,
                      cam_probe const& cam, rparams const& p, accum_rt& rt)
{
    std::vector<index_bvh<tri_t>::bvh_ref> bvhs{tree.ref()};
    auto blend_params = pixel_sampler::jittered_blend_type{};
    blend_params.spp = p.spp;
    blend_params.sfactor = p.sf;
    blend_params.dfactor = p.df;
    auto sparams = make_sched_params(blend_params, cam, rt);
    auto kparams = make_kernel_params(
            bvhs.data(), bvhs.data() + bvhs.size(), mats.data(), (unsigned)p.bounces, p.eps,
            vec4(p.bg[0], p.bg[1], p.bg[2], 1.0f), vec4(p.amb[0], p.amb[1], p.amb[2], 1.0f));
    pathtracing::kernel<decltype(kparams)> kernel;
    kernel.params = kparams;
    sched.frame(kernel, sparams);
}

// area light input: cl[3] kl, then the triangle (geom_id prim_id v1 e1 e2):
// 15 x 4 B, the member order of area_light<float, basic_triangle<3, float>>
using light_t = area_light<float, tri_t>;

static std::vector<light_t> load_lights(const char* path, size_t nl)
{
    auto raw = load<float>(path, nl * 15);
    std::vector<light_t> lights;
    lights.reserve(nl);
    for (size_t i = 0; i < nl; ++i) {
        const float* r = &raw[i * 15];
        tri_t t;
        std::memcpy(&t, r + 4, sizeof(tri_t));
        light_t l(t);
        l.set_cl(vec3(r[0], r[1], r[2]));
        l.set_kl(r[3]);
        lights.push_back(l);
    }
    return lights;
}

template <typename R, typename Sched>
static void render_ptl(Sched& sched, index_bvh<tri_t> const& tree, std::vector<mat_t> const& mats,
                       std::vector<light_t> const& lights, cam_probe const& cam, rparams const& p, accum_rt& rt)
{
    std::vector<index_bvh<tri_t>::bvh_ref> bvhs{tree.ref()};
    auto blend_params = pixel_sampler::jittered_blend_type{};
    blend_params.spp = p.spp;
    blend_params.sfactor = p.sf;
    blend_params.dfactor = p.df;
    auto sparams = make_sched_params(blend_params, cam, rt);
    auto kparams = make_kernel_params(
            bvhs.data(), bvhs.data() + bvhs.size(), mats.data(), lights.data(), lights.data() + lights.size(),
            (unsigned)p.bounces, p.eps, vec4(p.bg[0], p.bg[1], p.bg[2], 1.0f),
            vec4(p.amb[0], p.amb[1], p.amb[2], 1.0f));
    pathtracing::kernel<decltype(kparams)> kernel;
    kernel.params = kparams;
    sched.frame(kernel, sparams);
}

static void save_rt(const char* path, accum_rt const& rt)
{
    save(path, reinterpret_cast<const float*>(rt.color.data()), rt.color.size() * 4);
}

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv)
{
    if (argc < 2) { std::fprintf(stderr, "usage: vr_ref <mode> ...\n"); return 2; }
    std::string mode = argv[1];
    char** a = argv + 2;
    int na = argc - 2;

    if (mode == "morton" && na == 3) {
        size_t n = std::strtoull(a[1], nullptr, 10);
        auto tris = load<tri_t>(a[0], n);
        lbvh_builder b;
        b.init(tris.data(), tris.data() + n);
        std::vector<uint32_t> out(2 * n);
        for (size_t i = 0; i < n; ++i) { out[2 * i] = b.prim_refs[i].morton_code; out[2 * i + 1] = (uint32_t)b.prim_refs[i].id; }
        save(a[2], out.data(), out.size());
        return 0;
    }
    if (mode == "karras" && na == 3) {
        int n = std::atoi(a[1]);
        auto codes = load<uint32_t>(a[0], n);
        std::vector<detail::lbvh::prim_ref> refs(n);
        for (int i = 0; i < n; ++i) { refs[i].id = i; refs[i].morton_code = codes[i]; }
        std::vector<int32_t> out;
        for (int i = 0; i + 1 < n; ++i) {
            int split = -1;
            vec2i r = detail::lbvh::determine_range(refs.data(), n, i, split);
            out.push_back(r.x); out.push_back(r.y); out.push_back(split);
        }
        save(a[2], out.data(), out.size());
        return 0;
    }
    if (mode == "trace" && na == 9) {
        auto tree = load_tree(a);
        size_t nr = std::strtoull(a[6], nullptr, 10);
        auto rays = load<basic_ray<float>>(a[5], nr);
        int anyh = std::atoi(a[7]);
        std::vector<index_bvh<tri_t>::bvh_ref> bvhs{tree.ref()};
        std::vector<hit_out> out(nr);
        for (size_t i = 0; i < nr; ++i) {
            if (anyh) out[i] = to_out(any_hit(rays[i], bvhs.data(), bvhs.data() + 1));
            else out[i] = to_out(closest_hit(rays[i], bvhs.data(), bvhs.data() + 1));
        }
        save(a[8], out.data(), out.size());
        return 0;
    }
    if ((mode == "wide" && na == 9) || (mode == "tracew" && na == 10)) {
        auto tree = load_tree(a);
        const int W = std::atoi(a[5]);
        auto run = [&](auto wide, auto comp) {
            using WT = decltype(wide);
            using CT = decltype(comp);
            thread_pool pool(1);
            bvh_collapser().collapse(tree, wide, pool);
            bvh_compressor().compress(wide, comp);
            if (mode == "wide") {
                save(a[6], wide.nodes().data(), wide.nodes().size());
                uint64_t count = wide.nodes().size();
                save(a[7], &count, 1);
                save(a[8], comp.nodes().data(), comp.nodes().size());
                return;
            }
            const int use_comp = std::atoi(a[6]);
            size_t nr = std::strtoull(a[8], nullptr, 10);
            auto rays = load<basic_ray<float>>(a[7], nr);
            std::vector<hit_out> out(nr);
            auto wref = wide.ref();
            auto cref = comp.ref();
            if constexpr (WT::Width == 4) {
                for (size_t i = 0; i < nr; ++i)
                    out[i] = use_comp ? to_out(intersect_ray1_bvhN_compressed(rays[i], cref))
                                      : to_out(intersect_ray1_bvhN(rays[i], wref));
            } else if constexpr (WIDE8_TRAVERSAL) {
                // the reference's compressed 8-wide traversal has no x86
                // build (its sign_extend for 8 bytes exists for NEON only)
                if (use_comp) {
                    std::fprintf(stderr, "tracew: no compressed 8-wide traversal on this ISA\n");
                    std::exit(3);
                }
                for (size_t i = 0; i < nr; ++i) out[i] = to_out(intersect_ray1_bvhN(rays[i], wref));
            } else {
                std::fprintf(stderr, "tracew: 8-wide traversal needs the AVX build\n");
                std::exit(3);
            }
            save(a[9], out.data(), out.size());
            (void)sizeof(CT);
        };
        if (W == 4) run(index_bvh4<tri_t>{}, compressed_index_bvh4<tri_t>{});
        else if (W == 8) run(index_bvh8<tri_t>{}, compressed_index_bvh8<tri_t>{});
        else return 2;
        return 0;
    }
    if (mode == "brute" && na == 6) {
        size_t np = std::strtoull(a[1], nullptr, 10);
        size_t nr = std::strtoull(a[3], nullptr, 10);
        auto tris = load<tri_t>(a[0], np);
        auto rays = load<basic_ray<float>>(a[2], nr);
        int anyh = std::atoi(a[4]);
        std::vector<hit_out> out(nr);
        for (size_t i = 0; i < nr; ++i) {
            if (anyh) out[i] = to_out(any_hit(rays[i], tris.data(), tris.data() + np));
            else out[i] = to_out(closest_hit(rays[i], tris.data(), tris.data() + np));
        }
        save(a[5], out.data(), out.size());
        return 0;
    }
    if (mode == "seed" && na == 3) {
        size_t n = std::strtoull(a[1], nullptr, 10);
        auto pairs = load<int32_t>(a[0], 2 * n);
        std::vector<uint32_t> out(n);
        for (size_t i = 0; i < n; ++i) out[i] = make_random_seed(pairs[2 * i], pairs[2 * i + 1]);
        save(a[2], out.data(), n);
        return 0;
    }
    if (mode == "rng" && na == 3) {
        unsigned seed = (unsigned)std::strtoul(a[0], nullptr, 10);
        size_t count = std::strtoull(a[1], nullptr, 10);
        random_generator<float> gen(seed);
        std::vector<float> out(count);
        for (auto& x : out) x = gen.next();
        save(a[2], out.data(), count);
        return 0;
    }
    if (mode == "probe" && na == 2) {
        unsigned seed = (unsigned)std::strtoul(a[0], nullptr, 10);
        std::vector<float> out;
        {   // jitter of make_primary_ray (detail/sched_common.h:222)
            random_generator<float> gen(seed);
            probe_cam pc;
            auto r = detail::make_primary_ray(basic_ray<float>{}, pixel_sampler::jittered_type{}, gen, 0, 0, 1, 1, pc);
            out.push_back(r.ori.x);
            out.push_back(r.ori.y);
        }
        {   // cosine sample of lambertian::sample_f (brdf.h:67)
            random_generator<float> gen(seed);
            lambertian<float> l;
            l.cd = spectrum<float>(1.0f);
            l.kd = 1.0f;
            vec3 wi;
            float pdf;
            int inter;
            l.sample_f(vec3(0.0f, 0.0f, 1.0f), vec3(0.0f, 0.0f, 1.0f), wi, pdf, inter, gen);
            out.push_back(wi.x); out.push_back(wi.y); out.push_back(wi.z);
        }
        {   // cosine sample in the AO kernel expression (examples/ao/main.cpp:231)
            random_generator<float> gen(seed);
            auto sp = cosine_sample_hemisphere(gen.next(), gen.next());
            out.push_back(sp.x); out.push_back(sp.y); out.push_back(sp.z);
        }
        save(a[1], out.data(), out.size());
        return 0;
    }
    if (mode == "camera" && na == 2) {
        auto cam = load_camera(a[0]);
        cam.begin_frame();
        vec3 u = cam.u(), v = cam.v(), w = cam.w(), e = cam.eye();
        float out[12] = {e.x, e.y, e.z, u.x, u.y, u.z, v.x, v.y, v.z, w.x, w.y, w.z};
        save(a[1], out, 12);
        return 0;
    }
    if (mode == "ao" && na == 8) {
        auto tree = load_tree(a);
        size_t np = std::strtoull(a[3], nullptr, 10);
        std::vector<vec3> normals(np);
        for (size_t i = 0; i < np; ++i) {
            auto const& t = tree.primitives()[i];
            normals[t.prim_id] = normalize(cross(t.e1, t.e2));
        }
        auto cam = load_camera(a[5]);
        auto p = load_params(a[6]);
        accum_rt rt(p.w, p.h);
        tiled_sched<basic_ray<float>> sched(p.threads);
        cam.set_viewport(0, 0, p.w, p.h);
        for (int f = 0; f <= p.frame_id; ++f) render_ao<basic_ray<float>>(sched, tree, normals, cam, p, rt);
        save_rt(a[7], rt);
        return 0;
    }
    if (mode == "pt" && na == 10) {
        auto tree = load_tree(a);
        size_t nm = std::strtoull(a[6], nullptr, 10);
        auto mats = load_mats(a[5], nm);
        auto cam = load_camera(a[7]);
        auto p = load_params(a[8]);
        accum_rt rt(p.w, p.h);
        tiled_sched<basic_ray<float>> sched(p.threads);
        for (int f = 0; f <= p.frame_id; ++f) render_pt<basic_ray<float>>(sched, tree, mats, cam, p, rt);
        save_rt(a[9], rt);
        return 0;
    }
    if (mode == "ptl" && na == 12) {
        auto tree = load_tree(a);
        size_t nm = std::strtoull(a[6], nullptr, 10);
        auto mats = load_mats(a[5], nm);
        size_t nl = std::strtoull(a[8], nullptr, 10);
        auto lights = load_lights(a[7], nl);
        auto cam = load_camera(a[9]);
        auto p = load_params(a[10]);
        accum_rt rt(p.w, p.h);
        tiled_sched<basic_ray<float>> sched(p.threads);
        for (int f = 0; f <= p.frame_id; ++f) render_ptl<basic_ray<float>>(sched, tree, mats, lights, cam, p, rt);
        save_rt(a[11], rt);
        return 0;
    }
    if (mode == "sahcost" && na == 6) {
        auto tree = load_tree(a);
        float c = sah_cost(tree);
        save(a[5], &c, 1);
        return 0;
    }
    if (mode == "rotate" && na == 7) {
        auto tree = load_tree(a);
        thread_pool pool(1);
        bvh_optimizer opt;
        int count = opt.optimize_tree_rotations(tree, pool);
        save(a[5], tree.nodes().data(), tree.nodes().size());
        save(a[6], &count, 1);
        return 0;
    }
    if (mode == "inst" && na == 11) {
        struct inst_in { uint32_t blas; int32_t inst_id; float m[12]; };
        struct inst_out { int32_t prim_id; float t, u, v; int32_t inst_id, geom_id; };
        size_t np = std::strtoull(a[1], nullptr, 10);
        size_t nb = std::strtoull(a[3], nullptr, 10);
        size_t ni = std::strtoull(a[5], nullptr, 10);
        size_t nr = std::strtoull(a[7], nullptr, 10);
        auto tris = load<tri_t>(a[0], np);
        auto of = load<uint32_t>(a[2], nb + 1);
        auto ins = load<inst_in>(a[4], ni);
        auto rays = load<basic_ray<float>>(a[6], nr);
        lbvh_builder b;
        std::vector<index_bvh<tri_t>> blas;
        blas.reserve(nb);
        for (size_t k = 0; k < nb; ++k) blas.push_back(b.build(index_bvh<tri_t>{}, tris.data() + of[k], of[k + 1] - of[k]));
        using inst_t = index_bvh<tri_t>::bvh_inst;
        std::vector<inst_t> insts;
        for (size_t i = 0; i < ni; ++i) {
            auto x = blas[ins[i].blas].inst(mat4x3(ins[i].m));
            x.set_inst_id(ins[i].inst_id);
            insts.push_back(x);
        }
        auto top = b.build(index_bvh<inst_t>{}, insts.data(), ni);
        std::vector<index_bvh<inst_t>::bvh_ref> tops{top.ref()};
        std::vector<inst_out> out(nr);
        for (size_t i = 0; i < nr; ++i) {
            auto hr = closest_hit(rays[i], tops.data(), tops.data() + 1);
            inst_out& o = out[i];
            o.prim_id = hr.hit ? (int32_t)hr.prim_id : -1;
            o.t = hr.hit ? hr.t : FLT_MAX;
            o.u = hr.hit ? hr.u : 0.0f;
            o.v = hr.hit ? hr.v : 0.0f;
            o.inst_id = hr.hit ? (int32_t)hr.inst_id : -1;
            o.geom_id = hr.hit ? (int32_t)hr.geom_id : -1;
        }
        save(a[8], out.data(), out.size());
        save(a[9], top.nodes().data(), top.nodes().size());
        save(a[10], top.indices().data(), top.indices().size());
        return 0;
    }
    if (mode == "lbvhcpu" && na == 4) {
        size_t n = std::strtoull(a[1], nullptr, 10);
        auto tris = load<tri_t>(a[0], n);
        lbvh_builder b;
        auto tree = b.build(index_bvh<tri_t>{}, tris.data(), n, 4);
        save(a[2], tree.nodes().data(), tree.nodes().size());
        save(a[3], tree.indices().data(), tree.indices().size());
        std::printf("%zu %zu\n", tree.nodes().size(), tree.indices().size());
        return 0;
    }
    if (mode == "sah" && na == 4) {
        size_t n = std::strtoull(a[1], nullptr, 10);
        auto tris = load<tri_t>(a[0], n);
        binned_sah_builder b;
        b.enable_spatial_splits(false);
        auto tree = b.build(index_bvh<tri_t>{}, tris.data(), n);
        save(a[2], tree.nodes().data(), tree.nodes().size());
        save(a[3], tree.indices().data(), tree.indices().size());
        std::printf("%zu %zu\n", tree.nodes().size(), tree.indices().size());
        return 0;
    }
    // bench TRIS NP MATS NM CAMIN PARAMS PACKET(1|4) BUILD(lbvh|sah) FRAMES OUTJSON
    //   CPU baseline: the reference's own CPU builder, then tiled_sched +
    //   pathtracing::kernel frames timed with steady_clock around frame() only.
    if (mode == "bench" && na == 10) {
        size_t np = std::strtoull(a[1], nullptr, 10);
        size_t nm = std::strtoull(a[3], nullptr, 10);
        auto tris = load<tri_t>(a[0], np);
        auto mats = load_mats(a[2], nm);
        auto cam = load_camera(a[4]);
        auto p = load_params(a[5]);
        int packet = std::atoi(a[6]);
        std::string build = a[7];
        int frames = std::atoi(a[8]);
        double t0 = now_s();
        index_bvh<tri_t> tree;
        if (build == "sah") {
            binned_sah_builder b;
            b.enable_spatial_splits(false);
            tree = b.build(index_bvh<tri_t>{}, tris.data(), np);
        } else {
            lbvh_builder b;
            tree = b.build(index_bvh<tri_t>{}, tris.data(), np, 4);
        }
        double build_s = now_s() - t0;
        std::fprintf(stderr, "[cpu-ref] %s build of %zu triangles: %.3f s\n", build.c_str(), np, build_s);
        std::fflush(stderr);
        accum_rt rt(p.w, p.h);
        std::vector<double> ft;
        // one scheduler for all frames, as the viewer keeps one: its thread
        // pool notifies its workers without holding their mutex, so a pool
        // torn down per frame can miss a wake-up at join and never return;
        // the process leaves with _Exit below for the same reason
        tiled_sched<basic_ray<simd::float4>> sched4(packet == 4 ? p.threads : 1);
        tiled_sched<basic_ray<float>> sched1(packet == 4 ? 1 : p.threads);
        for (int f = 0; f < frames + 1; ++f) {
            double s0 = now_s();
            if (packet == 4) render_pt<basic_ray<simd::float4>>(sched4, tree, mats, cam, p, rt);
            else render_pt<basic_ray<float>>(sched1, tree, mats, cam, p, rt);
            double s1 = now_s();
            if (f > 0) ft.push_back(s1 - s0);
            std::fprintf(stderr, "[cpu-ref] frame %d of %d (%s): %.3f s\n", f, frames, f ? "timed" : "warm-up",
                         s1 - s0);
            std::fflush(stderr);
        }
        std::sort(ft.begin(), ft.end());
        double med = ft.empty() ? 0.0 : ft[ft.size() / 2];
        FILE* f = std::fopen(a[9], "w");
        std::fprintf(f, "{\"build_s\": %.6f, \"frame_s_median\": %.6f, \"frames\": %d, \"threads\": %d, \"packet\": %d, \"build\": \"%s\"}\n",
                     build_s, med, (int)ft.size(), p.threads, packet, build.c_str());
        std::fclose(f);
        save_rt((std::string(a[9]) + ".img").c_str(), rt);
        std::fflush(stdout);
        std::fflush(stderr);
        std::_Exit(0);
    }
    std::fprintf(stderr, "bad mode/args: %s (%d args)\n", mode.c_str(), na);
    return 2;
}
