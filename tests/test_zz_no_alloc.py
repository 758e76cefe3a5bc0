"""vr_trace_noalloc: the closest-hit / any-hit batch that allocates no device
memory (its work counter lives in a file-scope device array) must return
exactly what vr_trace returns, call after call: two calls in a row, the
second with a smaller input, for closest and any hits, through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from visionaray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, S.RAY_DTYPE)
    r["ori"] = rng.uniform(-1.2, 1.2, size=(n, 3))
    d = rng.normal(size=(n, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmax"] = np.finfo(np.float32).max
    return r


def test_trace_noalloc_matches_trace(V):
    lib = V.lib()
    tris = S.random_clusters(120_000, seed=21, clusters=8, sigma=0.4)
    bvh = V.lbvh_builder().build(tris, keep_nodes=False)
    rays = _rays(200_000, seed=22)
    print("[no-alloc] 120000 triangles, calls of 200000 then 7001 rays", flush=True)
    for any_hit in (0, 1):
        for n in (200_000, 7_001):
            d_rays = V.DeviceArray.from_host(rays[:n])
            want = V.DeviceArray(4 * n, np.float32)
            got = V.DeviceArray(4 * n, np.float32)
            # stale output must be overwritten: fill it with a pattern first
            got.upload(np.full(4 * n, -7.0, np.float32))
            V.check(lib.vr_trace(bvh.handle, any_hit, d_rays.ptr, n, want.ptr, None))
            V.check(lib.vr_trace_noalloc(bvh.handle, any_hit, d_rays.ptr, n, got.ptr, None))
            V.synchronize()
            w = want.to_host().view(np.int32)
            g = got.to_host().view(np.int32)
            assert np.array_equal(g, w), f"any_hit={any_hit} n={n}: {int(np.sum(g != w))} words differ"
            hit = w.reshape(-1, 4)[:, 0] >= 0
            print(f"[no-alloc] any_hit={any_hit} n={n}: identical to vr_trace ({hit.mean():.3f} hit)", flush=True)


def test_trace_noalloc_rejects_oversize(V):
    """More than 2^31 rays is refused on the host, before any launch."""
    lib = V.lib()
    tris = S.random_clusters(1000, seed=23)
    bvh = V.lbvh_builder().build(tris, keep_nodes=False)
    dummy = V.DeviceArray(8, np.float32)
    rc = lib.vr_trace_noalloc(bvh.handle, 0, dummy.ptr, (1 << 31) + 1, dummy.ptr, None)
    assert rc == 1 and b"2^31" in lib.vr_last_error()
    assert lib.vr_trace_noalloc(bvh.handle, 0, dummy.ptr, 0, dummy.ptr, None) == 0
    assert lib.vr_trace_noalloc(None, 0, dummy.ptr, 1, dummy.ptr, None) == 1
