"""The reference's kernel-level operator surface from user code
(include/visionaray_amd_device.hpp), exercised by tests/cpp/device_api_demo.hip:

* the AO kernel of src/examples/ao/main.cpp:192-260, written as the
  reference writes it (a lambda with closest_hit / any_hit on the BVH ref)
  and run by hip_sched<>::frame(kernel, sched_params), renders the library's
  own AO frame (vr_render_ao, k_ao) bit for bit;
* bvh_cost_intersector of src/viewer/bvh_costs.h:20-45, a basic_intersector
  counting box and triangle tests, passed to closest_hit in a kernel the
  caller launches: per ray it counts exactly the triangle tests and 4x the
  4-wide node visits vr_trace_counted reports, and finds the same hits."""
import ctypes as C

import numpy as np
import pytest

from conftest import ROOT
from visionaray_amd import scenes as S

pytestmark = pytest.mark.gpu
DEMO = ROOT / "tests" / "cpp" / "libdevice_api_demo.so"


@pytest.fixture(scope="module")
def demo(V):
    assert DEMO.exists(), "tests/cpp/libdevice_api_demo.so is not built (run __graft_entry__.build())"
    lib = C.CDLL(str(DEMO))
    F3 = C.POINTER(C.c_float)
    lib.demo_ao_frame.argtypes = [C.c_void_p, F3, F3, F3, C.c_float, C.c_float, C.c_int, C.c_int, C.c_uint,
                                  C.c_uint, C.c_float, F3, C.c_void_p, C.c_void_p]
    lib.demo_ao_frame.restype = C.c_int
    lib.demo_bvh_costs.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    lib.demo_bvh_costs.restype = C.c_int
    return lib


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


@pytest.mark.parametrize("sah", [False, True])
def test_user_ao_kernel_equals_library_ao(V, demo, sah):
    tris = S.config0_scene(level=4)
    if sah:
        nodes, idx = V.sah_build(tris, 4)
        bvh = V.bvh_from_nodes(tris, nodes, idx)
    else:
        bvh = V.lbvh_builder().build(tris)
    w, h, spp, samples, radius = 96, 72, 3, 8, 0.5
    bg = (0.1, 0.2, 0.3)
    eye, center, up, fovy, _ = S.config0_camera()
    aspect = w / h
    cam = V.pinhole_camera(eye, center, up, fovy, aspect)
    rt = V.render_target(w, h)
    V.hip_sched().frame(V.ao_kernel(bvh, samples=samples, radius=radius, bg=bg),
                        V.make_sched_params(V.jittered_blend_type(spp=spp), cam, rt))
    want = rt.color_host()
    urt = V.render_target(w, h)
    rc = demo.demo_ao_frame(bvh.handle, _f3(eye), _f3(center), _f3(up), float(fovy), aspect, w, h, spp, samples,
                            radius, _f3(bg), urt.accum.ptr, urt.color.ptr)
    assert rc == 0
    got = urt.color_host()
    occl = np.mean(want[..., 0] < 1.0)
    print(f"[device api] user AO kernel ({'SAH' if sah else 'LBVH'}): {np.mean(np.all(got == want, axis=-1)):.4f} "
          f"of pixels identical, {occl:.3f} occluded")
    assert np.array_equal(got, want)
    assert 0.05 < occl < 1.0


def test_cost_intersector_counts_equal_vr_trace_counted(V, demo):
    rng = np.random.default_rng(3)
    tris = S.random_clusters(60_000, seed=31, clusters=6, sigma=0.4)
    bvh = V.lbvh_builder().build(tris, keep_nodes=False)
    n = 40_000
    rays = np.zeros(n, V.RAY_DTYPE)
    rays["ori"] = rng.uniform(-1.5, 1.5, size=(n, 3))
    d = rng.normal(size=(n, 3))
    rays["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmax"] = np.finfo(np.float32).max
    d_rays = V.DeviceArray.from_host(rays)
    out = V.DeviceArray(4 * n, np.uint32)
    assert demo.demo_bvh_costs(bvh.handle, d_rays.ptr, n, out.ptr) == 0
    o = out.to_host().reshape(n, 4)
    hits = V.DeviceArray(4 * n, np.float32)
    counts = V.DeviceArray(2 * n, np.uint32)
    V.check(V.lib().vr_trace_counted(bvh.handle, 0, d_rays.ptr, n, hits.ptr, counts.ptr, None))
    V.synchronize()
    c = counts.to_host().reshape(n, 2)
    hh = hits.to_host().view(np.int32).reshape(n, 4)
    assert np.array_equal(o[:, 1], c[:, 1]), "triangle tests differ"
    assert np.array_equal(o[:, 0], 4 * c[:, 0]), "box tests differ from 4 x node visits"
    assert np.array_equal(o[:, 2].view(np.int32), hh[:, 0])
    assert np.array_equal(o[:, 3], hh[:, 1].view(np.uint32))
    print(f"[device api] cost intersector over {n} rays: {o[:, 0].mean():.2f} boxes, {o[:, 1].mean():.2f} "
          f"triangles per ray, equal to vr_trace_counted; {np.mean(o[:, 2].view(np.int32) >= 0):.3f} hit")
