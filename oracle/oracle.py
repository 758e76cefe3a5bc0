"""Python access to the oracle (test infrastructure only).

* `liboracle.so` -- our C restatement of the reference algorithms
  (vr_oracle.c), loaded with ctypes.
* `_ref/vr_ref` -- the reference's own CPU code compiled from its headers
  (ref_driver.cpp), run as a subprocess on raw binary files.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_BIN = HERE / "_ref" / "vr_ref"
REF_BIN_V3 = HERE / "_ref" / "vr_ref_v3"

TRI_DTYPE = np.dtype([("geom_id", "<u4"), ("prim_id", "<u4"), ("v1", "<f4", (3,)), ("e1", "<f4", (3,)),
                      ("e2", "<f4", (3,))])
NODE_DTYPE = np.dtype([("lo", "<f4", (3,)), ("hi", "<f4", (3,)), ("first", "<u4"), ("num_prims", "<u2"),
                       ("axis", "u1"), ("sign", "u1")])
RAY_DTYPE = np.dtype([("ori", "<f4", (3,)), ("dir", "<f4", (3,)), ("tmin", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim_id", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
CAMERA_DTYPE = np.dtype([("eye", "<f4", (3,)), ("U", "<f4", (3,)), ("V", "<f4", (3,)), ("W", "<f4", (3,))])
MAT_DTYPE = np.dtype([("type", "<i4"), ("cd", "<f4", (3,)), ("kd", "<f4"), ("ce", "<f4", (3,)), ("ls", "<f4")])
# area_light<float, basic_triangle<3, float>>: cl, kl, triangle (60 B)
LIGHT_DTYPE = np.dtype([("cl", "<f4", (3,)), ("kl", "<f4"), ("geometry", TRI_DTYPE)])
RPARAMS_DTYPE = np.dtype([
    ("width", "<u4"), ("height", "<u4"), ("spp", "<u4"), ("frame_id", "<u4"), ("num_bounces", "<u4"),
    ("epsilon", "<f4"), ("bg", "<f4", (3,)), ("ambient", "<f4", (3,)), ("sfactor", "<f4"), ("dfactor", "<f4"),
    ("ao_samples", "<u4"), ("ao_radius", "<f4"), ("arg_order", "<u4"),
])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-C", str(HERE), "liboracle.so"], check=True, capture_output=True)
        l = C.CDLL(str(LIB))
        P, SZ = C.c_void_p, C.c_size_t
        l.vro_morton_encode3d.restype = C.c_uint32
        l.vro_morton_encode3d.argtypes = [C.c_uint32] * 3
        l.vro_make_random_seed.restype = C.c_uint32
        l.vro_make_random_seed.argtypes = [C.c_int32, C.c_int32]
        l.vro_lbvh_build.restype = C.c_int
        l.vro_lbvh_build.argtypes = [P, SZ, P, P, P, P]
        l.vro_morton_codes.argtypes = [P, SZ, P]
        l.vro_centroid_bounds.argtypes = [P, SZ, P, P]
        l.vro_sort_prim_refs.argtypes = [P, SZ, P, P]
        l.vro_determine_range.argtypes = [P, C.c_int, C.c_int, P, P, P]
        l.vro_trace.argtypes = [P, P, P, SZ, P, SZ, C.c_int, P]
        l.vro_trace_brute.argtypes = [P, SZ, P, SZ, C.c_int, P]
        l.vro_intersect_tri.argtypes = [P, P, P, P, P, P]
        l.vro_rng_stream.argtypes = [C.c_uint32, SZ, P]
        l.vro_refit.argtypes = [P, SZ, P, P]
        l.vro_render_ao.argtypes = [P, P, P, SZ, P, P, P, P, C.c_int, C.c_int]
        l.vro_render_pt.argtypes = [P, P, P, SZ, P, P, P, P, P, C.c_int, C.c_int, P]
        l.vro_render_pt_lights.argtypes = [P, P, P, SZ, P, P, SZ, P, P, P, P, C.c_int, C.c_int, P]
        l.vro_query_log.argtypes = [P, SZ]
        l.vro_query_log_count.restype = SZ
        l.vro_render_pt_adjoint.argtypes = [P, P, P, SZ, P, P, P, P, P, P, C.c_int, C.c_int]
        l.vro_collapse_wide.restype = SZ
        l.vro_collapse_wide.argtypes = [P, SZ, C.c_int, P, P, P]
        l.vro_compress_wide.argtypes = [P, P, SZ, C.c_int, P]
        _lib = l
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------------------
# restatement
# ---------------------------------------------------------------------------

def morton_encode3d(x: int, y: int, z: int) -> int:
    return int(lib().vro_morton_encode3d(x, y, z))


def morton_codes(tris: np.ndarray) -> np.ndarray:
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    out = np.zeros(tris.size, np.uint32)
    lib().vro_morton_codes(_p(tris), tris.size, _p(out))
    return out


def centroid_bounds(tris: np.ndarray):
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lib().vro_centroid_bounds(_p(tris), tris.size, _p(lo), _p(hi))
    return lo, hi


def sort_prim_refs(codes: np.ndarray):
    codes = np.ascontiguousarray(codes, np.uint32)
    sc = np.zeros(codes.size, np.uint32)
    si = np.zeros(codes.size, np.int32)
    lib().vro_sort_prim_refs(_p(codes), codes.size, _p(sc), _p(si))
    return sc, si


def determine_range(sorted_codes: np.ndarray, i: int):
    sorted_codes = np.ascontiguousarray(sorted_codes, np.uint32)
    f = np.zeros(1, np.int32)
    l = np.zeros(1, np.int32)
    s = np.zeros(1, np.int32)
    lib().vro_determine_range(_p(sorted_codes), sorted_codes.size, i, _p(f), _p(l), _p(s))
    return int(f[0]), int(l[0]), int(s[0])


def lbvh_build(tris: np.ndarray):
    """-> (nodes[2n-1], indices[n], sorted_codes[n], sorted_ids[n])"""
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    n = tris.size
    nodes = np.zeros(max(2 * n - 1, 1), NODE_DTYPE)
    idx = np.zeros(max(n, 1), np.uint32)
    sc = np.zeros(max(n, 1), np.uint32)
    si = np.zeros(max(n, 1), np.int32)
    rc = lib().vro_lbvh_build(_p(tris), n, _p(nodes), _p(idx), _p(sc), _p(si))
    assert rc == 0
    return nodes[: max(2 * n - 1, 0)], idx[:n], sc[:n], si[:n]


def refit(nodes: np.ndarray, tris: np.ndarray, indices: np.ndarray) -> np.ndarray:
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE).copy()
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    indices = np.ascontiguousarray(indices, np.uint32)
    lib().vro_refit(_p(nodes), nodes.size, _p(tris), _p(indices))
    return nodes


def trace(nodes, tris, indices, rays, any_hit=False) -> np.ndarray:
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    indices = np.ascontiguousarray(indices, np.uint32)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(rays.size, HIT_DTYPE)
    lib().vro_trace(_p(nodes), _p(tris), _p(indices), tris.size, _p(rays), rays.size, int(any_hit), _p(hits))
    return hits


def collapse_wide(nodes, width: int):
    """bvh_collapser into bvh_multi_node<width> -> (bounds (num, 6, width)
    float32: minx miny minz maxx maxy maxz per slot, children (num, width)
    int64, deepest collapse stack)"""
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    n = max(nodes.size, 1)
    b = np.zeros((n, 6, width), np.float32)
    c = np.zeros((n, width), np.int64)
    ms = np.zeros(1, np.uint64)
    num = lib().vro_collapse_wide(_p(nodes), nodes.size, width, _p(b), _p(c), _p(ms))
    return b[:num].copy(), c[:num].copy(), int(ms[0])


def compress_wide(bounds, children, width: int) -> np.ndarray:
    """bvh_compressed_node<width>::init of each node, reference bytes
    (num, 64 | 112) uint8"""
    bounds = np.ascontiguousarray(bounds, np.float32)
    children = np.ascontiguousarray(children, np.int64)
    stride = 64 if width == 4 else 112
    out = np.zeros((max(len(children), 1), stride), np.uint8)
    lib().vro_compress_wide(_p(bounds), _p(children), len(children), width, _p(out))
    return out[: len(children)]


def wide_leaf(child: int):
    """decode_leaf of an encoded leaf child -> (first, count)"""
    a = ~int(child) & 0xFFFFFFFFFFFFFFFF
    return a & 0xFFFFFFFFFFFF, a >> 48


def trace_brute(tris, rays, any_hit=False) -> np.ndarray:
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.zeros(rays.size, HIT_DTYPE)
    lib().vro_trace_brute(_p(tris), tris.size, _p(rays), rays.size, int(any_hit), _p(hits))
    return hits


def intersect(tris, rays, which) -> np.ndarray:
    """Ray i against triangle tris[which[i]] alone (math/intersect.h:176-229);
    prim_id is which[i] on a hit, -1 otherwise."""
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    out = np.zeros(rays.size, HIT_DTYPE)
    out["prim_id"] = -1
    hit, t, u, v = C.c_int(), C.c_float(), C.c_float(), C.c_float()
    for i, k in enumerate(np.asarray(which)):
        lib().vro_intersect_tri(C.c_void_p(rays.ctypes.data + i * RAY_DTYPE.itemsize),
                                C.c_void_p(tris.ctypes.data + int(k) * TRI_DTYPE.itemsize),
                                C.byref(hit), C.byref(t), C.byref(u), C.byref(v))
        if hit.value:
            out[i] = (int(k), t.value, u.value, v.value)
    return out


def check_hits(got, want, tris, rays, any_hit=False) -> int:
    """Compare hit records of another traversal with the oracle's under the
    closest-hit contract: the same rays hit, at bit-identical t; where the
    reported triangle differs, it must be hit at exactly that t (a tie between
    triangles, resolved by visit order).  Any-hit: the same rays hit, each at
    a real intersection inside [tmin, tmax].  Returns the number of ties;
    raises AssertionError otherwise."""
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    index = {int(p): i for i, p in enumerate(tris["prim_id"])}
    gh, wh = got["prim_id"] >= 0, want["prim_id"] >= 0
    assert np.array_equal(gh, wh), f"hit/miss differs on {int(np.sum(gh != wh))} rays"
    if any_hit:
        sel = np.nonzero(gh)[0]
    else:
        assert np.array_equal(got["t"][gh], want["t"][gh]), "closest t differs"
        same = gh & (got["prim_id"] == want["prim_id"])
        assert np.array_equal(got["u"][same], want["u"][same]) and np.array_equal(got["v"][same], want["v"][same])
        sel = np.nonzero(gh & ~same)[0]
    if sel.size == 0:
        return 0
    which = np.array([index[int(p)] for p in got["prim_id"][sel]])
    re = intersect(tris, rays[sel], which)
    ok = (re["prim_id"] >= 0) & (re["t"] == got["t"][sel]) & (re["u"] == got["u"][sel]) & (re["v"] == got["v"][sel])
    ok &= (re["t"] >= rays["tmin"][sel]) & (re["t"] <= rays["tmax"][sel])
    assert np.all(ok), f"{int(np.sum(~ok))} reported hits are not intersections at the reported t"
    return 0 if any_hit else int(sel.size)


def make_random_seed(v0: int, v1: int) -> int:
    return int(lib().vro_make_random_seed(v0, v1))


def rng_stream(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, np.float32)
    lib().vro_rng_stream(seed, count, _p(out))
    return out


def render_params(width, height, spp=1, frame_id=0, num_bounces=5, epsilon=float(np.finfo(np.float32).eps),
                  bg=(0, 0, 0), ambient=(0, 0, 0), sfactor=1.0, dfactor=0.0, ao_samples=8, ao_radius=0.1,
                  arg_order=1) -> np.ndarray:
    p = np.zeros(1, RPARAMS_DTYPE)
    p["width"], p["height"], p["spp"], p["frame_id"] = width, height, spp, frame_id
    p["num_bounces"], p["epsilon"] = num_bounces, epsilon
    p["bg"], p["ambient"] = bg, ambient
    p["sfactor"], p["dfactor"] = sfactor, dfactor
    p["ao_samples"], p["ao_radius"], p["arg_order"] = ao_samples, ao_radius, arg_order
    return p


def render_ao(nodes, tris, indices, cam, params, rows=None) -> np.ndarray:
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    indices = np.ascontiguousarray(indices, np.uint32)
    cam = np.ascontiguousarray(cam, CAMERA_DTYPE)
    w, h = int(params["width"][0]), int(params["height"][0])
    acc = np.zeros((h, w, 4), np.float32)
    col = np.zeros((h, w, 4), np.float32)
    y0, y1 = rows if rows else (0, h)
    lib().vro_render_ao(_p(nodes), _p(tris), _p(indices), tris.size, _p(cam), _p(params), _p(acc), _p(col), y0, y1)
    return col


def render_pt(nodes, tris, indices, mats, cam, params, rows=None, lights=None, accum=None):
    """-> (color image, [primary rays, secondary rays]); `lights` (LIGHT_DTYPE)
    turns on next-event estimation (shadow rays are counted as secondary);
    `accum` ((h, w, 4) float32, updated in place) is the render target's
    accumulation buffer the frame blends into (zeros when None)."""
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    indices = np.ascontiguousarray(indices, np.uint32)
    mats = np.ascontiguousarray(mats, MAT_DTYPE)
    cam = np.ascontiguousarray(cam, CAMERA_DTYPE)
    w, h = int(params["width"][0]), int(params["height"][0])
    if accum is None:
        acc = np.zeros((h, w, 4), np.float32)
    else:
        acc = accum
        assert acc.dtype == np.float32 and acc.shape == (h, w, 4) and acc.flags["C_CONTIGUOUS"]
    col = np.zeros((h, w, 4), np.float32)
    counts = np.zeros(2, np.uint64)
    y0, y1 = rows if rows else (0, h)
    if lights is None or len(lights) == 0:
        lib().vro_render_pt(_p(nodes), _p(tris), _p(indices), tris.size, _p(mats), _p(cam), _p(params), _p(acc),
                            _p(col), y0, y1, _p(counts))
    else:
        lights = np.ascontiguousarray(lights, LIGHT_DTYPE)
        lib().vro_render_pt_lights(_p(nodes), _p(tris), _p(indices), tris.size, _p(mats), _p(lights), lights.size,
                                   _p(cam), _p(params), _p(acc), _p(col), y0, y1, _p(counts))
    return col, counts


QUERY_DTYPE = np.dtype([("ray", RAY_DTYPE), ("any", "<i4"), ("hit", "<i4"), ("prim_id", "<i4"), ("t", "<f4"),
                        ("bounce", "<u4"), ("pixel", "<u4")])


def render_pt_queries(nodes, tris, indices, mats, cam, params, rows, lights=None, cap=1 << 20):
    """render_pt over `rows` that also returns every query it traced
    (QUERY_DTYPE: ray, any-hit flag, hit, prim_id, t, bounce, pixel)."""
    buf = np.zeros(cap, QUERY_DTYPE)
    lib().vro_query_log(_p(buf), cap)
    try:
        img, counts = render_pt(nodes, tris, indices, mats, cam, params, rows=rows, lights=lights)
        n = lib().vro_query_log_count()
    finally:
        lib().vro_query_log(None, 0)
    return img, counts, buf[:n].copy()


def render_pt_adjoint(nodes, tris, indices, mats, cam, params, dl_dimg, rows=None):
    """FP64 adjoint: -> (grad_cd (nm, 3), grad_ce (nm, 3))"""
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    tris = np.ascontiguousarray(tris, TRI_DTYPE)
    indices = np.ascontiguousarray(indices, np.uint32)
    mats = np.ascontiguousarray(mats, MAT_DTYPE)
    cam = np.ascontiguousarray(cam, CAMERA_DTYPE)
    dl = np.ascontiguousarray(dl_dimg, np.float32)
    h = int(params["height"][0])
    gcd = np.zeros((mats.size, 3), np.float64)
    gce = np.zeros((mats.size, 3), np.float64)
    y0, y1 = rows if rows else (0, h)
    lib().vro_render_pt_adjoint(_p(nodes), _p(tris), _p(indices), tris.size, _p(mats), _p(cam), _p(params), _p(dl),
                                _p(gcd), _p(gce), y0, y1)
    return gcd, gce


# ---------------------------------------------------------------------------
# the reference's own code (oracle/_ref/vr_ref), where it was built
# ---------------------------------------------------------------------------

def ref_available() -> bool:
    return REF_BIN.exists()


def ref_binary() -> Path:
    """The baseline build the host CPU supports (AVX2 when present)."""
    try:
        flags = Path("/proc/cpuinfo").read_text()
        if " avx2 " in flags and " fma " in flags and REF_BIN_V3.exists():
            return REF_BIN_V3
    except OSError:
        pass
    return REF_BIN


class _Tmp:
    def __init__(self):
        self.dir = tempfile.TemporaryDirectory(prefix="vr_ref_")

    def put(self, name: str, a: np.ndarray) -> str:
        p = os.path.join(self.dir.name, name)
        np.ascontiguousarray(a).tofile(p)
        return p

    def path(self, name: str) -> str:
        return os.path.join(self.dir.name, name)

    def close(self):
        self.dir.cleanup()


def _run(args, binary=None):
    b = str(binary or REF_BIN)
    res = subprocess.run([b, *map(str, args)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"vr_ref {args[0]} failed: {res.stderr.strip()}")
    return res.stdout


def camera_input(eye, center, up, fovy, aspect) -> np.ndarray:
    return np.array([*eye, *center, *up, fovy, aspect], np.float32)


def ref_params(width, height, spp=1, frame_id=0, num_bounces=5, epsilon=float(np.finfo(np.float32).eps),
               bg=(0, 0, 0), ambient=(0, 0, 0), sfactor=1.0, dfactor=0.0, ao_samples=8, ao_radius=0.1,
               threads=0) -> np.ndarray:
    return np.array([width, height, spp, frame_id, num_bounces, epsilon, *bg, *ambient, sfactor, dfactor,
                     ao_samples, ao_radius, threads], np.float32)


def ref_morton(tris):
    t = _Tmp()
    try:
        _run(["morton", t.put("tris.bin", tris), tris.size, t.path("out.bin")])
        out = np.fromfile(t.path("out.bin"), np.uint32).reshape(-1, 2)
        return out[:, 0].copy(), out[:, 1].astype(np.int32)
    finally:
        t.close()


def ref_karras(sorted_codes):
    t = _Tmp()
    try:
        _run(["karras", t.put("codes.bin", np.asarray(sorted_codes, np.uint32)), len(sorted_codes),
              t.path("out.bin")])
        return np.fromfile(t.path("out.bin"), np.int32).reshape(-1, 3)
    finally:
        t.close()


def ref_trace(nodes, tris, indices, rays, any_hit=False):
    t = _Tmp()
    try:
        _run(["trace", t.put("nodes.bin", nodes), nodes.size, t.put("tris.bin", tris), tris.size,
              t.put("idx.bin", np.asarray(indices, np.uint32)), t.put("rays.bin", rays), rays.size,
              int(any_hit), t.path("out.bin")])
        return np.fromfile(t.path("out.bin"), HIT_DTYPE)
    finally:
        t.close()


def ref_brute(tris, rays, any_hit=False):
    t = _Tmp()
    try:
        _run(["brute", t.put("tris.bin", tris), tris.size, t.put("rays.bin", rays), rays.size, int(any_hit),
              t.path("out.bin")])
        return np.fromfile(t.path("out.bin"), HIT_DTYPE)
    finally:
        t.close()


def ref_seed(pairs):
    pairs = np.ascontiguousarray(pairs, np.int32)
    t = _Tmp()
    try:
        _run(["seed", t.put("pairs.bin", pairs), pairs.shape[0], t.path("out.bin")])
        return np.fromfile(t.path("out.bin"), np.uint32)
    finally:
        t.close()


def ref_rng(seed, count):
    t = _Tmp()
    try:
        _run(["rng", seed, count, t.path("out.bin")])
        return np.fromfile(t.path("out.bin"), np.float32)
    finally:
        t.close()


def ref_probe(seed):
    t = _Tmp()
    try:
        _run(["probe", seed, t.path("out.bin")])
        return np.fromfile(t.path("out.bin"), np.float32)
    finally:
        t.close()


def ref_camera(cam_in):
    t = _Tmp()
    try:
        _run(["camera", t.put("cam.bin", np.asarray(cam_in, np.float32)), t.path("out.bin")])
        return np.fromfile(t.path("out.bin"), np.float32)
    finally:
        t.close()


def ref_ao(nodes, tris, indices, cam_in, params):
    t = _Tmp()
    try:
        _run(["ao", t.put("nodes.bin", nodes), nodes.size, t.put("tris.bin", tris), tris.size,
              t.put("idx.bin", np.asarray(indices, np.uint32)), t.put("cam.bin", cam_in),
              t.put("params.bin", params), t.path("out.bin")])
        w, h = int(params[0]), int(params[1])
        return np.fromfile(t.path("out.bin"), np.float32).reshape(h, w, 4)
    finally:
        t.close()


def ref_pt(nodes, tris, indices, mats, cam_in, params):
    t = _Tmp()
    try:
        _run(["pt", t.put("nodes.bin", nodes), nodes.size, t.put("tris.bin", tris), tris.size,
              t.put("idx.bin", np.asarray(indices, np.uint32)), t.put("mats.bin", mats), mats.size,
              t.put("cam.bin", cam_in), t.put("params.bin", params), t.path("out.bin")])
        w, h = int(params[0]), int(params[1])
        return np.fromfile(t.path("out.bin"), np.float32).reshape(h, w, 4)
    finally:
        t.close()


def ref_ptl(nodes, tris, indices, mats, lights, cam_in, params):
    """The reference's pathtracing::kernel with an area-light list (NEE + MIS)."""
    t = _Tmp()
    try:
        lights = np.ascontiguousarray(lights, LIGHT_DTYPE)
        _run(["ptl", t.put("nodes.bin", nodes), nodes.size, t.put("tris.bin", tris), tris.size,
              t.put("idx.bin", np.asarray(indices, np.uint32)), t.put("mats.bin", mats), mats.size,
              t.put("lights.bin", lights), lights.size, t.put("cam.bin", cam_in), t.put("params.bin", params),
              t.path("out.bin")])
        w, h = int(params[0]), int(params[1])
        return np.fromfile(t.path("out.bin"), np.float32).reshape(h, w, 4)
    finally:
        t.close()


def ref_sah(tris):
    t = _Tmp()
    try:
        _run(["sah", t.put("tris.bin", tris), tris.size, t.path("nodes.bin"), t.path("idx.bin")])
        return np.fromfile(t.path("nodes.bin"), NODE_DTYPE), np.fromfile(t.path("idx.bin"), np.uint32)
    finally:
        t.close()


def ref_sah_cost(nodes, tris, indices) -> float:
    """The reference's sah_cost(tree) (detail/bvh/statistics.h:38-68) with its
    default costs (ci 1.2, cl 0, cp 1)."""
    t = _Tmp()
    try:
        _run(["sahcost", t.put("nodes.bin", np.ascontiguousarray(nodes, NODE_DTYPE)), len(nodes),
              t.put("tris.bin", tris), tris.size, t.put("idx.bin", np.asarray(indices, np.uint32)),
              t.path("out.bin")])
        return float(np.fromfile(t.path("out.bin"), np.float32)[0])
    finally:
        t.close()


def ref_rotate(nodes, tris, indices):
    """The reference's bvh_optimizer::optimize_tree_rotations
    (detail/bvh/optimize.h:22-269): (rotated nodes, rotation count)."""
    t = _Tmp()
    try:
        _run(["rotate", t.put("nodes.bin", np.ascontiguousarray(nodes, NODE_DTYPE)), len(nodes),
              t.put("tris.bin", tris), tris.size, t.put("idx.bin", np.asarray(indices, np.uint32)),
              t.path("out.bin"), t.path("count.bin")])
        return np.fromfile(t.path("out.bin"), NODE_DTYPE), int(np.fromfile(t.path("count.bin"), np.int32)[0])
    finally:
        t.close()


def ref_wide(nodes, tris, indices, width: int):
    """The reference's bvh_collapser (detail/bvh/collapse.h:19-169) into
    index_bvh<width> and bvh_compressed_node<width>::init (bvh.h:352-400) of
    every wide node -> (bounds, children, compressed bytes) as collapse_wide
    and compress_wide give them."""
    t = _Tmp()
    try:
        _run(["wide", t.put("nodes.bin", np.ascontiguousarray(nodes, NODE_DTYPE)), len(nodes),
              t.put("tris.bin", tris), tris.size, t.put("idx.bin", np.asarray(indices, np.uint32)), width,
              t.path("wide.bin"), t.path("count.bin"), t.path("comp.bin")])
        num = int(np.fromfile(t.path("count.bin"), np.uint64)[0])
        raw = np.fromfile(t.path("wide.bin"), np.uint8).reshape(num, 32 * width)
        bounds = raw[:, : 24 * width].copy().view(np.float32).reshape(num, 6, width)
        children = raw[:, 24 * width:].copy().view(np.int64).reshape(num, width)
        comp = np.fromfile(t.path("comp.bin"), np.uint8).reshape(num, 64 if width == 4 else 112)
        return bounds, children, comp
    finally:
        t.close()


def ref_trace_wide(nodes, tris, indices, width: int, rays, compressed=False, binary=None):
    """Closest hits of the reference's wide traversal over its collapsed tree:
    intersect_ray1_bvhN (detail/bvh/intersect_ray1_bvhN.inl:166-410) or, for
    compressed nodes, intersect_ray1_bvhN_compressed.  width 8 runs the AVX
    build of the driver (the reference's 8-lane SIMD types)."""
    t = _Tmp()
    try:
        _run(["tracew", t.put("nodes.bin", np.ascontiguousarray(nodes, NODE_DTYPE)), len(nodes),
              t.put("tris.bin", tris), tris.size, t.put("idx.bin", np.asarray(indices, np.uint32)), width,
              int(compressed), t.put("rays.bin", rays), rays.size, t.path("out.bin")],
             binary=binary or (REF_BIN_V3 if width == 8 else REF_BIN))
        return np.fromfile(t.path("out.bin"), HIT_DTYPE)
    finally:
        t.close()


def ref_lbvh_cpu(tris):
    """The reference's CPU lbvh_builder tree (leaf size 4)."""
    t = _Tmp()
    try:
        _run(["lbvhcpu", t.put("tris.bin", tris), tris.size, t.path("nodes.bin"), t.path("idx.bin")])
        return np.fromfile(t.path("nodes.bin"), NODE_DTYPE), np.fromfile(t.path("idx.bin"), np.uint32)
    finally:
        t.close()


def ref_bench(tris, mats, cam_in, params, packet=4, build="lbvh", frames=3, binary=None, timeout=600,
              stall=60.0):
    """Time the reference CPU path (builder + tiled_sched frames).  The
    driver's per-frame progress lines are echoed as they come; a run that
    prints nothing for `stall` seconds (or outlives `timeout`) is stopped and
    the timed frames it reported are used."""
    import re
    import sys
    import threading
    import time as _time
    t = _Tmp()
    try:
        out = t.path("bench.json")
        args = ["bench", t.put("tris.bin", tris), tris.size, t.put("mats.bin", mats), mats.size,
                t.put("cam.bin", cam_in), t.put("params.bin", params), packet, build, frames, out]
        b = str(binary or ref_binary())
        proc = subprocess.Popen([b, *map(str, args)], stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
        seen = {"last": _time.time(), "frames": [], "build": None}

        def pump():
            for line in proc.stderr:
                sys.stderr.write(line)
                sys.stderr.flush()
                seen["last"] = _time.time()
                m = re.search(r"frame \d+ of \d+ \(timed\): ([0-9.]+) s", line)
                if m:
                    seen["frames"].append(float(m.group(1)))
                m = re.search(r"build of \d+ triangles: ([0-9.]+) s", line)
                if m:
                    seen["build"] = float(m.group(1))

        th = threading.Thread(target=pump, daemon=True)
        th.start()
        t0 = _time.time()
        while proc.poll() is None:
            now = _time.time()
            if now - seen["last"] > stall or now - t0 > timeout:
                proc.kill()
                proc.wait()
                th.join(timeout=5)
                if not seen["frames"]:
                    raise RuntimeError("vr_ref bench stalled before a timed frame")
                fr = sorted(seen["frames"])
                return {"build_s": seen["build"] or 0.0, "frame_s_median": fr[len(fr) // 2],
                        "frames": len(fr), "stopped": f"no progress for {stall:.0f} s"}
            _time.sleep(0.2)
        th.join(timeout=5)
        if proc.returncode != 0:
            raise RuntimeError(f"vr_ref bench failed with status {proc.returncode}")
        with open(out) as f:
            return json.load(f)
    finally:
        t.close()


# ---------------------------------------------------------------------------
# instanced two-level BVH (bvh.h:679-782, detail/bvh/intersect.inl:362-394)
# ---------------------------------------------------------------------------

INSTANCE_DTYPE = np.dtype([("blas", "<u4"), ("inst_id", "<i4"), ("transform", "<f4", (12,))])
INST_HIT_DTYPE = np.dtype([("prim_id", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("inst_id", "<i4"),
                           ("geom_id", "<i4")])


def inverse3(m):
    """inverse(matrix<3, 3>) (math/detail/matrix3.inl:300-321) in float32, each
    product and sum rounded as written; m[r][c], returns the same layout."""
    f = np.float32
    m = [[f(m[r][c]) for c in range(3)] for r in range(3)]

    def det2(m00, m01, m10, m11):
        return f(f(m00 * m11) - f(m10 * m01))

    a00 = det2(m[1][1], m[1][2], m[2][1], m[2][2])
    a01 = det2(m[1][0], m[1][2], m[2][0], m[2][2])
    a02 = det2(m[1][0], m[1][1], m[2][0], m[2][1])
    a10 = det2(m[0][1], m[0][2], m[2][1], m[2][2])
    a11 = det2(m[0][0], m[0][2], m[2][0], m[2][2])
    a12 = det2(m[0][0], m[0][1], m[2][0], m[2][1])
    a20 = det2(m[0][1], m[0][2], m[1][1], m[1][2])
    a21 = det2(m[0][0], m[0][2], m[1][0], m[1][2])
    a22 = det2(m[0][0], m[0][1], m[1][0], m[1][1])
    det = f(f(f(m[0][0] * a00) - f(m[0][1] * a01)) + f(m[0][2] * a02))
    # matrix<3, 3>(m00, m10, m20, m01, ...): entries column by column
    return [[f(a00 / det), f(-a10 / det), f(a20 / det)],
            [f(-a01 / det), f(a11 / det), f(-a21 / det)],
            [f(a02 / det), f(-a12 / det), f(a22 / det)]]


def instance_maps(inst):
    """(affine_inv, trans_inv) of index_bvh_inst_t(ref, mat4x3) (bvh.h:696-702):
    inverse(top_left(M)) with top_left(mat4x3)(r, c) = M(c, r), the transpose
    (math/detail/matrix4x3.inl:206-218), and -column 3."""
    t = np.asarray(inst["transform"], np.float32)
    lin = [[t[3 * r + c] for c in range(3)] for r in range(3)]
    return inverse3(lin), [np.float32(-t[9 + k]) for k in range(3)]


def _mat_rows(m, x, y, z):
    """matrix<3, 3> * vector<3> (matrix3.inl:226-235) over float32 arrays"""
    return [(m[r][0] * x + m[r][1] * y) + m[r][2] * z for r in range(3)]


def instance_box(inst, root_box):
    """get_bounds(index_bvh_inst_t) (detail/bvh/get_bounds.inl:30-56): the
    BVH box's corners in compute_vertices order (math/detail/aabb.inl:305-324)
    through inverse(affine_inv), plus -trans_inv; aabb::insert min / max."""
    ainv, ti = instance_maps(inst)
    aff = inverse3(ainv)
    f = np.float32
    lo, hi = [f(v) for v in root_box[:3]], [f(v) for v in root_box[3:]]
    corners = [(hi[0], hi[1], hi[2]), (lo[0], hi[1], hi[2]), (lo[0], lo[1], hi[2]), (hi[0], lo[1], hi[2]),
               (lo[0], hi[1], lo[2]), (hi[0], hi[1], lo[2]), (hi[0], lo[1], lo[2]), (lo[0], lo[1], lo[2])]
    blo = [f(np.finfo(np.float32).max)] * 3
    bhi = [f(-np.finfo(np.float32).max)] * 3
    for v in corners:
        w = _mat_rows(aff, *v)
        for k in range(3):
            x = f(w[k] + f(-ti[k]))
            blo[k] = x if x < blo[k] else blo[k]
            bhi[k] = x if bhi[k] < x else bhi[k]
    return np.array(blo + bhi, np.float32)


def transform_rays(inst, rays):
    """index_bvh_inst_t::transform_ray (bvh.h:769-776):
    ori' = affine_inv * (ori + trans_inv), dir' = affine_inv * dir."""
    ainv, ti = instance_maps(inst)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    out = rays.copy()
    o = rays["ori"].astype(np.float32)
    d = rays["dir"].astype(np.float32)
    ox, oy, oz = o[:, 0] + ti[0], o[:, 1] + ti[1], o[:, 2] + ti[2]
    oo = _mat_rows(ainv, ox, oy, oz)
    dd = _mat_rows(ainv, d[:, 0], d[:, 1], d[:, 2])
    out["ori"] = np.stack(oo, axis=1)
    out["dir"] = np.stack(dd, axis=1)
    return out


def trace_instanced(blas, instances, rays):
    """Closest hit over instances: for each instance the rays moved into its
    space and traced through its bottom level (trace: the reference's
    intersect over index_bvh), the nearest hit kept (is_closer: strictly
    smaller t).  blas = [(nodes, tris, indices)].  Instances are visited in
    list order here; the reference visits them in its top-level order, so
    only rays that hit two instances at exactly the same t can differ (the
    callers count them)."""
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    out = np.zeros(rays.size, INST_HIT_DTYPE)
    out["prim_id"] = -1
    out["t"] = np.finfo(np.float32).max
    out["inst_id"] = -1
    out["geom_id"] = -1
    ties = np.zeros(rays.size, bool)
    for inst in np.ascontiguousarray(instances, INSTANCE_DTYPE):
        nodes, tris, idx = blas[int(inst["blas"])]
        h = trace(nodes, tris, idx, transform_rays(inst, rays))
        hit = h["prim_id"] >= 0
        ties |= hit & (h["t"] == out["t"]) & (out["prim_id"] >= 0)
        better = hit & (h["t"] < out["t"])
        out["prim_id"][better] = h["prim_id"][better]
        out["t"][better] = h["t"][better]
        out["u"][better] = h["u"][better]
        out["v"][better] = h["v"][better]
        out["inst_id"][better] = inst["inst_id"]
        gid = {int(p): int(g) for p, g in zip(tris["prim_id"], tris["geom_id"])}
        out["geom_id"][better] = [gid[int(p)] for p in h["prim_id"][better]]
    return out, ties


def ref_inst(tris, offsets, instances, rays):
    """The reference's two-level BVH (vr_ref inst): CPU lbvh_builder per
    bottom level and over the instances -> (closest hits, top-level nodes,
    top-level instance order)."""
    t = _Tmp()
    try:
        _run(["inst", t.put("tris.bin", tris), tris.size, t.put("of.bin", np.asarray(offsets, np.uint32)),
              len(offsets) - 1, t.put("inst.bin", np.ascontiguousarray(instances, INSTANCE_DTYPE)), len(instances),
              t.put("rays.bin", rays), rays.size, t.path("out.bin"), t.path("nodes.bin"), t.path("idx.bin")])
        return (np.fromfile(t.path("out.bin"), INST_HIT_DTYPE), np.fromfile(t.path("nodes.bin"), NODE_DTYPE),
                np.fromfile(t.path("idx.bin"), np.uint32))
    finally:
        t.close()
