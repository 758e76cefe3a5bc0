// facade_demo.cpp -- a C++ caller written against the reference-named facade
// (include/visionaray_amd.hpp): LBVH and SAH builds, closest/any hit, refit,
// and a path-traced + AO frame, printed as one JSON object for the GPU test.
#include <cstdio>
#include <vector>

#include "visionaray_amd.hpp"

using namespace visionaray_amd;

static basic_triangle tri(vec3 a, vec3 b, vec3 c, unsigned id, unsigned geom = 0)
{
    basic_triangle t{};
    t.geom_id = geom;
    t.prim_id = id;
    const float v[9] = {a.x, a.y, a.z, b.x - a.x, b.y - a.y, b.z - a.z, c.x - a.x, c.y - a.y, c.z - a.z};
    for (int k = 0; k < 3; ++k) {
        t.v1[k] = v[k];
        t.e1[k] = v[3 + k];
        t.e2[k] = v[6 + k];
    }
    return t;
}

int main()
{
    check(vr_init(0));
    // the two triangles of test/unittests/get_normal.cpp:17-75
    std::vector<basic_triangle> prims = {
        tri({-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, 0, 0),
        tri({1, -1, -1}, {-1, -1, -1}, {-1, 1, -1}, 1, 1),
    };
    device_vector<basic_triangle> d_prims(prims);

    lbvh_builder builder;
    hip_index_bvh lbvh = builder.build(hip_index_bvh{}, d_prims.data(), d_prims.size());
    binned_sah_builder sah;
    hip_index_bvh sbvh = sah.build(hip_index_bvh{}, prims.data(), prims.size());

    ray r{};
    r.ori[0] = 0.5f; r.ori[1] = -0.5f; r.ori[2] = 2.0f;
    r.dir[2] = -1.0f;
    r.tmax = FLT_MAX;
    std::vector<ray> rays(4, r);
    rays[1].ori[2] = 0.0f;   // between the triangles, hits prim 1 at t = 1
    rays[2].tmax = 0.5f;     // too short
    rays[3].dir[2] = 1.0f;   // away from both
    device_vector<ray> d_rays(rays);
    device_vector<hit_record> d_hits(rays.size());

    std::printf("{\"lbvh_nodes\": %zu, \"sah_nodes\": %zu", lbvh.num_nodes(), sbvh.num_nodes());
    for (int which = 0; which < 3; ++which) {
        const hip_index_bvh& b = which == 1 ? sbvh : lbvh;
        if (which == 2) any_hit(b, d_rays.data(), rays.size(), d_hits.data());
        else closest_hit(b, d_rays.data(), rays.size(), d_hits.data());
        check(vr_synchronize());
        auto h = d_hits.to_host();
        std::printf(", \"%s\": [", which == 0 ? "lbvh" : which == 1 ? "sah" : "any");
        for (size_t i = 0; i < h.size(); ++i)
            std::printf("%s[%d, %.9g]", i ? ", " : "", h[i].prim_id, (double)h[i].t);
        std::printf("]");
    }

    // two instances of the SAH tree: in place (inst_id 7) and 10 units down
    // z (inst_id 8); the same rays see the first one only
    {
        const float eye3[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
        const float down[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, -10};
        bvh_inst a(0, eye3), b(0, down);
        a.set_inst_id(7);
        b.set_inst_id(8);
        instanced_bvh top({&sbvh}, {a, b});
        device_vector<inst_hit_record> d_ih(rays.size());
        closest_hit(top, d_rays.data(), rays.size(), d_ih.data());
        auto ih = d_ih.to_host();
        std::printf(", \"inst\": [");
        for (size_t i = 0; i < ih.size(); ++i)
            std::printf("%s[%d, %.9g, %d]", i ? ", " : "", ih[i].prim_id, (double)ih[i].t, ih[i].inst_id);
        std::printf("], \"inst_top_nodes\": %zu", top.nodes().first.size());
    }

    // refit after moving the far triangle from z = -1 to z = -3
    for (int k = 0; k < 1; ++k) prims[1].v1[2] = -3.0f;
    device_vector<basic_triangle> d_moved(prims);
    bvh_refitter refitter;
    refitter.refit(lbvh, d_moved.data());
    closest_hit(lbvh, d_rays.data(), rays.size(), d_hits.data());
    auto h = d_hits.to_host();
    std::printf(", \"refit_t1\": %.9g", (double)h[1].t);

    // one path-traced and one AO frame of a tiny image
    std::vector<vr_material> mats(2);
    mats[0].type = 0; mats[0].cd[0] = mats[0].cd[1] = mats[0].cd[2] = 0.5f; mats[0].kd = 1.0f;
    mats[1].type = 1; mats[1].ce[0] = mats[1].ce[1] = mats[1].ce[2] = 1.0f; mats[1].ls = 2.0f;
    device_vector<vr_material> d_mats(mats);
    const int w = 16, hgt = 16;
    device_vector<float> color(w * hgt * 4), accum(w * hgt * 4);
    check(vr_memset(accum.data(), 0, accum.size() * sizeof(float)));
    pinhole_camera cam;
    cam.perspective(0.8f, 1.0f, 0.001f, 1000.0f);
    cam.look_at({0.0f, 0.0f, 4.0f}, {0.0f, 0.0f, 0.0f});
    hip_sched<ray> sched;
    auto kparams = make_kernel_params(sbvh, d_mats.data(), mats.size(), 4, 1e-4f, vec4{0, 0, 0, 1},
                                      vec4{0.5f, 0.5f, 0.5f, 1.0f});
    pathtracing::kernel<> kernel{kparams};
    render_target_ref rt{color.data(), accum.data(), w, hgt};
    pixel_sampler::jittered_blend_type blend;
    blend.spp = 2;
    sched.frame(kernel, make_sched_params(blend, cam, rt));
    auto img = color.to_host();
    double sum = 0.0;
    for (float v : img) sum += v;
    ao_kernel ao{&sbvh, 8, 0.5f, {0, 0, 0}};
    sched.frame(ao, make_sched_params(blend, cam, rt));
    auto aoimg = color.to_host();
    double aosum = 0.0;
    for (float v : aoimg) aosum += v;
    std::printf(", \"pt_sum\": %.9g, \"ao_sum\": %.9g, \"frames\": %u}\n", sum, aosum, sched.frame_id());
    return 0;
}
