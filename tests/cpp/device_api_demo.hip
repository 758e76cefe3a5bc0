// device_api_demo.hip -- user kernels written against the device API
// (include/visionaray_amd_device.hpp) the way the reference's callers write
// them; built into tests/cpp/libdevice_api_demo.so and driven by
// tests/test_gpu_device_api.py.
//
// * demo_ao_frame: the AO kernel of src/examples/ao/main.cpp:192-260 as a
//   lambda, rendered by hip_sched<>::frame(kernel, sched_params).
// * demo_bvh_costs: bvh_cost_intersector of src/viewer/bvh_costs.h:20-45 (a
//   basic_intersector that counts box and triangle tests) passed to
//   closest_hit from a kernel the caller launches itself.
#include "visionaray_amd_device.hpp"

using namespace visionaray_amd;

namespace {

struct bvh_cost_intersector : basic_intersector<bvh_cost_intersector> {
    using basic_intersector<bvh_cost_intersector>::operator();

    __device__ hit_record_aabb operator()(const basic_ray& ray, const aabb& box, vec3 inv_dir)
    {
        ++num_boxes;
        return intersect(ray, box, inv_dir);
    }

    __device__ hit_record_tri operator()(const basic_ray& ray, const triangle& tri)
    {
        ++num_tris;
        return intersect(ray, tri);
    }

    unsigned num_boxes = 0;
    unsigned num_tris = 0;
};

// per ray: boxes tested, triangles tested, prim_id, t (bits)
__global__ __launch_bounds__(traversal_block) void k_bvh_costs(bvh_ref b, const basic_ray* rays, uint32_t n,
                                                               uint32_t* out)
{
    const uint32_t i = blockIdx.x * (uint32_t)traversal_block + threadIdx.x;
    if (i >= n) return;
    bvh_cost_intersector isect;
    const hit_record_bvh hr = closest_hit(rays[i], b, isect);
    out[4 * (size_t)i + 0] = isect.num_boxes;
    out[4 * (size_t)i + 1] = isect.num_tris;
    out[4 * (size_t)i + 2] = (uint32_t)(hr.hit ? hr.prim_id : -1);
    out[4 * (size_t)i + 3] = __float_as_uint(hr.hit ? hr.t : FLT_MAX);
}

}  // namespace

extern "C" {

// The AO example's kernel over the BVH `tree` (not owned here), camera from
// eye / center / up / fovy / aspect, frame 0 of a fresh scheduler.
int demo_ao_frame(vr_bvh tree, const float eye[3], const float center[3], const float up[3], float fovy,
                  float aspect, int width, int height, unsigned spp, unsigned ao_samples, float ao_radius,
                  const float bg[3], float* d_accum, float* d_color)
{
    try {
        vr_bvh_ref r{};
        check(vr_bvh_get_ref(tree, &r));
        const bvh_ref bvh(r);
        const unsigned samples = ao_samples;
        const float radius = ao_radius;
        const vec4 bgcolor{bg[0], bg[1], bg[2], 1.0f};

        pinhole_camera cam;
        cam.perspective(fovy, aspect, 0.001f, 1000.0f);
        cam.look_at(vec3{eye[0], eye[1], eye[2]}, vec3{center[0], center[1], center[2]}, vec3{up[0], up[1], up[2]});
        pixel_sampler::jittered_blend_type blend;
        blend.spp = spp;
        hip_sched<> sched;

        sched.frame(
            [=] __device__(basic_ray ray, random_generator& gen) -> result_record {
                result_record result;
                result.color = bgcolor;

                auto hit_rec = closest_hit(ray, bvh);

                result.hit = hit_rec.hit;

                if (hit_rec.hit) {
                    hit_rec.isect_pos = ray.ori + ray.dir * hit_rec.t;
                    result.depth = hit_rec.t;

                    vec3 n = get_normal(bvh, hit_rec);
                    n = faceforward(n, -ray.dir, n);

                    vec3 u, v;
                    vec3 w = n;
                    make_orthonormal_basis(u, v, w);

                    float ao = 0.0f;
                    for (unsigned i = 0; i < samples; ++i) {
                        // cosine_sample_hemisphere(gen.next(), gen.next()) as
                        // the reference's host build evaluates it: right first
                        const float u2 = gen.next();
                        const float u1 = gen.next();
                        auto sp = cosine_sample_hemisphere(u1, u2);

                        auto dir = normalize(sp.x * u + sp.y * v + sp.z * w);

                        basic_ray ao_ray;
                        ao_ray.ori = hit_rec.isect_pos + dir * 1e-3f;
                        ao_ray.dir = dir;
                        ao_ray.tmin = 0.0f;
                        ao_ray.tmax = radius;

                        auto ao_rec = any_hit(ao_ray, bvh);

                        ao = ao_rec.hit ? ao + 1.0f / samples : ao;
                    }

                    ao = 1.0f - ao;
                    result.color = vec4{ao, ao, ao, 1.0f};
                }

                return result;
            },
            make_sched_params(blend, cam, render_target_ref{d_color, d_accum, width, height}));
        return hipDeviceSynchronize() == hipSuccess ? 0 : VR_ERR_HIP;
    } catch (const error& e) {
        return e.code;
    }
}

// bvh_cost_intersector through closest_hit for n device rays; d_out holds
// 4 words per ray (boxes, triangles, prim_id, t bits).
int demo_bvh_costs(vr_bvh tree, const vr_ray* d_rays, uint32_t n, uint32_t* d_out)
{
    vr_bvh_ref r{};
    const int rc = vr_bvh_get_ref(tree, &r);
    if (rc) return rc;
    if (n == 0) return 0;
    const unsigned blocks = (n + traversal_block - 1) / traversal_block;
    hipLaunchKernelGGL(k_bvh_costs, dim3(blocks), dim3(traversal_block), traversal_lds_bytes, nullptr, bvh_ref(r),
                       reinterpret_cast<const basic_ray*>(d_rays), n, d_out);
    if (hipGetLastError() != hipSuccess) return VR_ERR_HIP;
    return hipDeviceSynchronize() == hipSuccess ? 0 : VR_ERR_HIP;
}

}  // extern "C"
