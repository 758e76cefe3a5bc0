"""visionaray_amd -- MI355X (gfx950) implementation of Visionaray's ray-tracing
hot path behind a C ABI, with a host mirror of the reference's interface.

Reference names kept (reference files cited relative to include/visionaray/):
    lbvh_builder().build(tree_type, prims, n)     detail/bvh/lbvh.h:681-844
    bvh.refit(prims)                              detail/bvh/refit.h:28-114
    closest_hit / any_hit                         detail/traverse_linear.inl:149-214
    pinhole_camera                                pinhole_camera.h, detail/pinhole_camera.inl
    make_kernel_params                            kernels.h:87-132
    make_sched_params + jittered_blend_type       scheduler.h:103-194, pixel_sampler_types.h:51-69
    pathtracing.kernel                            detail/pathtracing.inl:33-244
    hip_sched(...).frame(kernel, sparams)         detail/cuda_sched.inl:226-246 (CUDA analogue)

Device memory is handled through `DeviceArray` (vr_malloc / vr_memcpy); every
compute call runs the HIP kernels of libvisionaray_amd.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (BVH_INFO_DTYPE, CAMERA_DTYPE, HIT_DTYPE, KERNEL_PARAMS_DTYPE, NODE_DTYPE, RAY_DTYPE,
                   SCHED_PARAMS_DTYPE, VR_BVH_KEEP_NODES, VR_WIDE_BINARY_LEAVES, VR_WIDE_COMPRESSED, VRError, check,
                   ptr)
from .scenes import LIGHT_DTYPE, MAT_DTYPE, TRI_DTYPE

__all__ = [
    "init", "DeviceArray", "Stream", "Event", "lbvh_builder", "BVH", "bvh_from_nodes", "sah_build", "sah_build_device",
    "closest_hit", "any_hit", "trace", "sah_cost_host", "optimize_tree_rotations", "pinhole_camera", "make_kernel_params", "make_sched_params",
    "jittered_blend_type", "pathtracing", "ao_kernel", "hip_sched", "VRError", "TRI_DTYPE", "MAT_DTYPE",
    "RAY_DTYPE", "HIT_DTYPE", "NODE_DTYPE", "CAMERA_DTYPE", "LIGHT_DTYPE", "synchronize", "model",
    "tlas", "make_instances", "INSTANCE_DTYPE", "INST_HIT_DTYPE",
]

_initialized = False


def lib():
    return _lib.load()


def init(device: int = 0):
    """Select the HIP device (vr_init)."""
    global _initialized
    check(lib().vr_init(device))
    _initialized = True


def _ensure():
    if not _initialized:
        init(0)


def synchronize():
    check(lib().vr_synchronize())


class DeviceArray:
    """Owning device allocation holding `count` records of `dtype`."""

    def __init__(self, count: int, dtype, zero: bool = False):
        _ensure()
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        self._p = C.c_void_p()
        check(lib().vr_malloc(C.byref(self._p), self.nbytes))
        if zero and self.nbytes:
            check(lib().vr_memset(self._p, 0, self.nbytes))

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if d.nbytes:
            check(lib().vr_memcpy_htod(d._p, ptr(a), d.nbytes))
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.count
        if self.nbytes:
            check(lib().vr_memcpy_htod(self._p, ptr(a), self.nbytes))

    def to_host(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        if self.nbytes:
            check(lib().vr_memcpy_dtoh(ptr(out), self._p, self.nbytes))
        return out

    @property
    def ptr(self) -> C.c_void_p:
        return self._p

    def free(self):
        if self._p and self._p.value:
            lib().vr_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        _ensure()
        self.handle = C.c_void_p()
        check(lib().vr_stream_create(C.byref(self.handle)))

    def synchronize(self):
        check(lib().vr_stream_synchronize(self.handle))

    def __del__(self):
        try:
            lib().vr_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        _ensure()
        self.handle = C.c_void_p()
        check(lib().vr_event_create(C.byref(self.handle)))

    def record(self, stream: Stream | None = None):
        check(lib().vr_event_record(self.handle, stream.handle if stream else None))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().vr_event_elapsed_ms(self.handle, end.handle, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            lib().vr_event_destroy(self.handle)
        except Exception:
            pass


def _stream(s):
    return s.handle if isinstance(s, Stream) else s


def _as_device(a, dtype) -> DeviceArray:
    if isinstance(a, DeviceArray):
        return a
    return DeviceArray.from_host(np.ascontiguousarray(a, dtype=dtype))


class BVH:
    """Handle of a device BVH (the analogue of hip_index_bvh<P>, bvh.h:1100)."""

    def __init__(self, handle: C.c_void_p, prims: DeviceArray | None = None):
        self.handle = handle
        self.prims = prims

    def info(self) -> dict:
        a = np.zeros(1, BVH_INFO_DTYPE)
        check(lib().vr_bvh_get_info(self.handle, ptr(a)))
        return {k: int(a[k][0]) for k in a.dtype.names}

    def nodes(self) -> np.ndarray:
        """tree.nodes() in the reference bvh_node layout."""
        n = self.info()["num_nodes"]
        d = DeviceArray(n, NODE_DTYPE)
        check(lib().vr_bvh_nodes(self.handle, d.ptr, None))
        return d.to_host()

    def indices(self) -> np.ndarray:
        n = self.info()["num_prims"]
        d = DeviceArray(n, np.uint32)
        check(lib().vr_bvh_indices(self.handle, d.ptr, None))
        return d.to_host()

    def sorted_refs(self):
        """(sorted Morton codes, original ids) of an LBVH build."""
        n = self.info()["num_prims"]
        c = DeviceArray(n, np.uint32)
        i = DeviceArray(n, np.int32)
        check(lib().vr_bvh_sorted_refs(self.handle, c.ptr, i.ptr, None))
        return c.to_host(), i.to_host()

    def sah_cost(self, ci: float = 1.2, cl: float = 0.0, cp: float = 1.0, stream=None) -> float:
        """sah_cost(tree, ci, cl, cp) (detail/bvh/statistics.h:38-68), on the
        device over the binary nodes (FP64 sums)."""
        c = C.c_double()
        check(lib().vr_bvh_sah_cost(self.handle, ci, cl, cp, C.byref(c), _stream(stream)))
        return float(c.value)

    def rebuild(self, prims, stream=None):
        """vr_lbvh_rebuild: a new LBVH of `prims` into this handle, reusing its
        device arrays when the primitive count is unchanged."""
        p = _as_device(prims, TRI_DTYPE)
        check(lib().vr_lbvh_rebuild(self.handle, p.ptr, p.count, _stream(stream)))
        self.prims = p

    def refit(self, prims, stream=None):
        """bvh_refitter::refit (detail/bvh/refit.h:28-114) for moved vertices."""
        p = _as_device(prims, TRI_DTYPE)
        check(lib().vr_bvh_refit(self.handle, p.ptr, _stream(stream)))
        self.prims = p

    def set_width(self, width: int, compressed: bool = False, binary_leaves: bool = False, stream=None):
        """vr_bvh_set_width: 8-wide traversal nodes (bvh_collapser into
        bvh_multi_node<8>, detail/bvh/collapse.h:19-169; compressed: 8-bit
        child planes as bvh_compressed_node<8>, bvh.h:316-426), or back to 4."""
        flags = (VR_WIDE_COMPRESSED if compressed else 0) | (VR_WIDE_BINARY_LEAVES if binary_leaves else 0)
        check(lib().vr_bvh_set_width(self.handle, int(width), flags, _stream(stream)))

    def wide_nodes(self):
        """(raw 8-wide node bytes (num_pairs, 256 | 128) uint8, 8-wide depth)"""
        n, d = C.c_size_t(), C.c_uint32()
        check(lib().vr_bvh_download_wide(self.handle, None, 0, C.byref(n), C.byref(d)))
        out = np.zeros(n.value, np.uint8)
        if n.value:
            check(lib().vr_bvh_download_wide(self.handle, C.c_void_p(out.ctypes.data), n.value, C.byref(n),
                                             C.byref(d)))
        return out, int(d.value)

    def destroy(self):
        if self.handle and self.handle.value:
            lib().vr_bvh_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class lbvh_builder:
    """GPU LBVH builder (lbvh_builder::build, detail/bvh/lbvh.h:681-844)."""

    def build(self, prims, n: int | None = None, keep_nodes: bool = True, stream=None, rotate: bool = False) -> BVH:
        """rotate: SAH-driven tree rotations before the traversal nodes are
        made (VR_BVH_ROTATE; implies keep_nodes=False)."""
        p = _as_device(prims, TRI_DTYPE)
        n = p.count if n is None else int(n)
        h = C.c_void_p()
        flags = (VR_BVH_KEEP_NODES if keep_nodes and not rotate else 0) | (_lib.VR_BVH_ROTATE if rotate else 0)
        check(lib().vr_lbvh_build(p.ptr, n, flags, _stream(stream), C.byref(h)))
        return BVH(h, p)

    @staticmethod
    def morton_codes(prims, stream=None) -> np.ndarray:
        """assign_morton_codes alone (detail/bvh/lbvh.h:242-269), unsorted."""
        p = _as_device(prims, TRI_DTYPE)
        d = DeviceArray(p.count, np.uint32)
        check(lib().vr_lbvh_morton(p.ptr, p.count, d.ptr, _stream(stream)))
        return d.to_host()


def bvh_from_nodes(prims: np.ndarray, nodes: np.ndarray, indices: np.ndarray) -> BVH:
    """Adopt a reference-layout tree built on the host (index_bvh_t, bvh.h:870-1010)."""
    _ensure()
    prims = np.ascontiguousarray(prims, TRI_DTYPE)
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    indices = np.ascontiguousarray(indices, np.uint32)
    h = C.c_void_p()
    check(lib().vr_bvh_from_nodes(ptr(prims), prims.size, ptr(nodes), nodes.size, ptr(indices), C.byref(h)))
    return BVH(h)


def sah_build(prims: np.ndarray, max_leaf_size: int = 4):
    """Host binned SAH build (binned_sah_builder, detail/bvh/sah.h:176-807):
    returns (nodes, indices) in the reference layout."""
    prims = np.ascontiguousarray(prims, TRI_DTYPE)
    n = prims.size
    nodes = np.zeros(max(2 * n - 1, 1), NODE_DTYPE)
    idx = np.zeros(max(n, 1), np.uint32)
    nn = C.c_size_t()
    check(lib().vr_sah_build(ptr(prims), n, max_leaf_size, ptr(nodes), C.byref(nn), ptr(idx)))
    return nodes[: nn.value].copy(), idx[:n].copy()


def sah_build_device(prims, max_leaf_size: int = 4, stream=None) -> BVH:
    """Binned SAH build on the device (vr_sah_build_device): the reference's
    binned_sah_builder tree node for node, as a BVH handle (nodes() and
    indices() give the reference layout)."""
    p = _as_device(prims, TRI_DTYPE)
    h = C.c_void_p()
    check(lib().vr_sah_build_device(p.ptr, p.count, int(max_leaf_size), _stream(stream), C.byref(h)))
    return BVH(h, p)


def sah_cost_host(nodes: np.ndarray, ci: float = 1.2, cl: float = 0.0, cp: float = 1.0) -> float:
    """sah_cost over a reference-layout node array, summed as the reference does."""
    nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
    c = C.c_float()
    check(lib().vr_bvh_sah_cost_host(ptr(nodes), nodes.size, ci, cl, cp, C.byref(c)))
    return float(c.value)


def optimize_tree_rotations(nodes: np.ndarray):
    """bvh_optimizer::optimize_tree_rotations (detail/bvh/optimize.h:22-262) on a
    copy of a reference-layout node array: (rotated nodes, rotations)."""
    out = np.array(nodes, NODE_DTYPE, copy=True)
    n = C.c_int()
    check(lib().vr_bvh_optimize_rotations_host(ptr(out), out.size, C.byref(n)))
    return out, int(n.value)


def lbvh_build_host(prims: np.ndarray, max_leaf_size: int = 4):
    """Host LBVH in the reference's CPU layout (lbvh_builder on an index_bvh,
    detail/bvh/lbvh.h:500-656): returns (nodes, indices)."""
    prims = np.ascontiguousarray(prims, TRI_DTYPE)
    n = prims.size
    nodes = np.zeros(max(2 * n - 1, 1), NODE_DTYPE)
    idx = np.zeros(max(n, 1), np.uint32)
    nn = C.c_size_t()
    check(lib().vr_lbvh_build_host(ptr(prims), n, max_leaf_size, ptr(nodes), C.byref(nn), ptr(idx)))
    return nodes[: nn.value].copy(), idx[:n].copy()


def trace(bvh: BVH, rays, any_hit_: bool = False, counted: bool = False, stream=None):
    """Trace rays (numpy RAY_DTYPE or DeviceArray); returns host hits (and
    per-ray [node visits, triangle tests] when counted)."""
    r = _as_device(rays, RAY_DTYPE)
    hits = DeviceArray(r.count, HIT_DTYPE)
    if counted:
        cnt = DeviceArray(2 * r.count, np.uint32)
        check(lib().vr_trace_counted(bvh.handle, int(any_hit_), r.ptr, r.count, hits.ptr, cnt.ptr, _stream(stream)))
        return hits.to_host(), cnt.to_host().reshape(-1, 2)
    check(lib().vr_trace(bvh.handle, int(any_hit_), r.ptr, r.count, hits.ptr, _stream(stream)))
    return hits.to_host()


def closest_hit(bvh: BVH, rays, stream=None) -> np.ndarray:
    return trace(bvh, rays, False, stream=stream)


def any_hit(bvh: BVH, rays, stream=None) -> np.ndarray:
    return trace(bvh, rays, True, stream=stream)


INSTANCE_DTYPE = _lib.INSTANCE_DTYPE
INST_HIT_DTYPE = _lib.INST_HIT_DTYPE


def make_instances(blas_ids, transforms, inst_ids=None) -> np.ndarray:
    """INSTANCE_DTYPE records: transforms are (n, 12) mat4x3 floats, column by
    column (three columns of the 3x3 part, then the translation)."""
    blas_ids = np.asarray(blas_ids, np.uint32).reshape(-1)
    t = np.asarray(transforms, np.float32).reshape(-1, 12)
    out = np.zeros(blas_ids.size, INSTANCE_DTYPE)
    out["blas"] = blas_ids
    out["inst_id"] = np.arange(blas_ids.size, dtype=np.int32) if inst_ids is None else inst_ids
    out["transform"] = t
    return out


class tlas:
    """Instanced two-level BVH: index_bvh<index_bvh_inst_t> over instances of
    built BVHs (bvh.h:679-782), the top level built like the reference's CPU
    lbvh_builder; closest_hit follows detail/bvh/intersect.inl:362-394."""

    def __init__(self, blas: list, instances: np.ndarray, stream=None):
        self.blas = list(blas)  # the bottom levels must outlive the top level
        inst = np.ascontiguousarray(instances, INSTANCE_DTYPE)
        handles = (C.c_void_p * max(len(self.blas), 1))(*[b.handle.value for b in self.blas])
        h = C.c_void_p()
        check(lib().vr_tlas_build(C.cast(handles, C.c_void_p), len(self.blas), ptr(inst), inst.size,
                                  _stream(stream), C.byref(h)))
        self.handle = h
        self.num_instances = int(inst.size)

    def closest_hit(self, rays, stream=None) -> np.ndarray:
        r = _as_device(rays, RAY_DTYPE)
        hits = DeviceArray(r.count, INST_HIT_DTYPE)
        check(lib().vr_tlas_closest_hit(self.handle, r.ptr, r.count, hits.ptr, _stream(stream)))
        return hits.to_host()

    def any_hit(self, rays, stream=None) -> np.ndarray:
        r = _as_device(rays, RAY_DTYPE)
        hits = DeviceArray(r.count, INST_HIT_DTYPE)
        check(lib().vr_tlas_any_hit(self.handle, r.ptr, r.count, hits.ptr, _stream(stream)))
        return hits.to_host()

    def closest_hit_device(self, rays: DeviceArray, hits: DeviceArray, stream=None):
        check(lib().vr_tlas_closest_hit(self.handle, rays.ptr, rays.count, hits.ptr, _stream(stream)))

    def nodes(self):
        """(reference-layout top-level nodes, instance order)"""
        n = C.c_size_t()
        check(lib().vr_tlas_nodes(self.handle, None, None, C.byref(n)))
        nodes = np.zeros(max(n.value, 1), NODE_DTYPE)
        idx = np.zeros(max(self.num_instances, 1), np.uint32)
        check(lib().vr_tlas_nodes(self.handle, ptr(nodes), ptr(idx), C.byref(n)))
        return nodes[: n.value].copy(), idx[: self.num_instances].copy()

    def destroy(self):
        if self.handle and self.handle.value:
            lib().vr_tlas_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def pinhole_camera(eye, center, up, fovy: float, aspect: float) -> np.ndarray:
    """perspective + look_at + begin_frame (detail/pinhole_camera.inl); returns
    the CAMERA_DTYPE record {eye, U, V, W}."""
    cam = np.zeros(1, CAMERA_DTYPE)
    e = np.asarray(eye, np.float32)
    c = np.asarray(center, np.float32)
    u = np.asarray(up, np.float32)
    check(lib().vr_camera_look_at(ptr(e), ptr(c), ptr(u), float(fovy), float(aspect), ptr(cam)))
    return cam


def make_kernel_params(num_bounces: int = 5, epsilon: float = float(np.finfo(np.float32).eps),
                       bg_color=(0.0, 0.0, 0.0, 0.0), ambient_color=(0.0, 0.0, 0.0, 0.0)) -> np.ndarray:
    """kernels.h:87-132 without lights; colors are (r, g, b, kl) and the
    intensity passed on is rgb * kl (ambient_light::intensity)."""
    kp = np.zeros(1, KERNEL_PARAMS_DTYPE)
    kp["num_bounces"] = num_bounces
    kp["epsilon"] = np.float32(epsilon)
    bg = np.asarray(bg_color, np.float32)
    am = np.asarray(ambient_color, np.float32)
    kp["bg"] = bg[:3] * bg[3]
    kp["ambient"] = am[:3] * am[3]
    return kp


@dataclass
class jittered_blend_type:
    """pixel_sampler::jittered_blend_type (pixel_sampler_types.h:51-69)."""
    spp: int = 1
    sfactor: float = 1.0
    dfactor: float = 0.0


@dataclass
class render_target:
    """RGBA32F color + accum render target on the device."""
    width: int
    height: int
    color: DeviceArray = field(init=False)
    accum: DeviceArray = field(init=False)

    def __post_init__(self):
        self.color = DeviceArray(self.width * self.height * 4, np.float32, zero=True)
        self.accum = DeviceArray(self.width * self.height * 4, np.float32, zero=True)

    def color_host(self) -> np.ndarray:
        return self.color.to_host().reshape(self.height, self.width, 4)


@dataclass
class sched_params:
    sample_params: jittered_blend_type
    cam: np.ndarray
    rt: render_target


def make_sched_params(sample_params: jittered_blend_type, cam: np.ndarray, rt: render_target) -> sched_params:
    """scheduler.h:103-194 (pixel sampler, camera, render target)."""
    return sched_params(sample_params, cam, rt)


class pathtracing:
    @dataclass
    class kernel:
        """pathtracing::kernel (detail/pathtracing.inl:33-244): BVH, materials
        indexed by geom_id, kernel params and, optionally, the area-light list
        of make_kernel_params w/ lights (kernels.h:137-187; LIGHT_DTYPE on the
        device), which turns on next-event estimation with MIS."""
        bvh: BVH
        materials: DeviceArray
        params: np.ndarray
        lights: "DeviceArray | None" = None


@dataclass
class ao_kernel:
    """The AO kernel of src/examples/ao/main.cpp:192-260."""
    bvh: BVH
    samples: int = 8
    radius: float = 0.1
    bg: tuple = (0.0, 0.0, 0.0)


LAUNCH_SHAPE_DTYPE = np.dtype([("block_size", "<u4"), ("stack_lds", "<u4"), ("tile", "<u4"), ("schedule", "<u4")])
SCHEDULE_PERSISTENT = 0
SCHEDULE_PER_PIXEL = 1


def set_launch_shape(block_size: int = 256, stack_lds: int = 16, tile: int = 1, schedule: int = SCHEDULE_PERSISTENT):
    """Launch shape of the path-tracing frame kernel (vr_set_launch_shape);
    schedule SCHEDULE_PER_PIXEL launches one thread per pixel as cuda_sched does."""
    a = np.array([(block_size, stack_lds, tile, schedule)], LAUNCH_SHAPE_DTYPE)
    check(lib().vr_set_launch_shape(ptr(a)))


def get_launch_shape() -> dict:
    a = np.zeros(1, LAUNCH_SHAPE_DTYPE)
    check(lib().vr_get_launch_shape(ptr(a)))
    return {k: int(a[k][0]) for k in a.dtype.names}


def autotune(kernel, sp, reps: int = 2, stream=None) -> tuple[dict, float]:
    """vr_autotune_pt: time every launch shape on this frame and keep the
    fastest; returns (shape, mean frame ms)."""
    s = np.zeros(1, SCHED_PARAMS_DTYPE)
    s["width"], s["height"], s["spp"] = sp.rt.width, sp.rt.height, sp.sample_params.spp
    s["sfactor"], s["dfactor"] = sp.sample_params.sfactor, sp.sample_params.dfactor
    best = np.zeros(1, LAUNCH_SHAPE_DTYPE)
    ms = C.c_float()
    check(lib().vr_autotune_pt(kernel.bvh.handle, kernel.materials.ptr, kernel.materials.count, ptr(kernel.params),
                               ptr(sp.cam), ptr(s), sp.rt.accum.ptr, sp.rt.color.ptr, int(reps), ptr(best),
                               C.byref(ms), _stream(stream)))
    return {k: int(best[k][0]) for k in best.dtype.names}, float(ms.value)


class batch:
    """Many scenes rendered by one persistent launch (vr_batch_*).  Each entry
    is (pathtracing.kernel, sched_params[, d loss / d image DeviceArray]); the
    scenes' BVHs, materials and render targets must stay alive."""

    def __init__(self, entries, frame_id: int = 0):
        _ensure()
        self._keep = list(entries)
        sc = np.zeros(len(self._keep), _lib.SCENE_DTYPE)
        for i, e in enumerate(self._keep):
            k, sp = e[0], e[1]
            dl = e[2] if len(e) > 2 else None
            sc[i]["bvh"] = k.bvh.handle.value or 0
            sc[i]["materials"] = k.materials.ptr.value or 0
            sc[i]["num_materials"] = k.materials.count
            sc[i]["kp"] = k.params[0]
            sc[i]["cam"] = sp.cam[0]
            sc[i]["sp"]["width"], sc[i]["sp"]["height"] = sp.rt.width, sp.rt.height
            sc[i]["sp"]["spp"], sc[i]["sp"]["frame_id"] = sp.sample_params.spp, frame_id
            sc[i]["sp"]["sfactor"], sc[i]["sp"]["dfactor"] = sp.sample_params.sfactor, sp.sample_params.dfactor
            sc[i]["accum"] = sp.rt.accum.ptr.value or 0
            sc[i]["color"] = sp.rt.color.ptr.value or 0
            sc[i]["dl"] = (dl.ptr.value or 0) if dl is not None else 0
        self.num_materials = int(sum(int(e[0].materials.count) for e in self._keep))
        self.handle = C.c_void_p()
        check(lib().vr_batch_create(ptr(sc), sc.size, C.byref(self.handle)))
        self.ray_counts = DeviceArray(2, np.uint64, zero=True)

    def render(self, stream=None):
        check(lib().vr_batch_render(self.handle, self.ray_counts.ptr, _stream(stream)))

    def backward(self, grad: DeviceArray, stream=None):
        check(lib().vr_batch_backward(self.handle, grad.ptr, self.ray_counts.ptr, _stream(stream)))

    def capture_graph(self, stream: Stream):
        check(lib().vr_batch_capture_graph(self.handle, self.ray_counts.ptr, _stream(stream)))

    def launch_graph(self, stream: Stream):
        check(lib().vr_batch_launch_graph(self.handle, _stream(stream)))

    def rays(self) -> tuple[int, int]:
        c = self.ray_counts.to_host()
        return int(c[0]), int(c[1])

    def reset_counters(self):
        check(lib().vr_memset(self.ray_counts.ptr, 0, self.ray_counts.nbytes))

    def __del__(self):
        try:
            lib().vr_batch_destroy(self.handle)
        except Exception:
            pass


def image_l2_loss(image: DeviceArray, target: DeviceArray, dl_dimage: DeviceArray | None = None,
                  stream=None) -> float:
    """sum (I - T)^2 over pixels and RGB (FP64); writes 2 (I - T) into dl_dimage."""
    npix = image.count // 4
    loss = DeviceArray(1, np.float64, zero=True)
    check(lib().vr_image_l2_loss(image.ptr, target.ptr, npix, dl_dimage.ptr if dl_dimage is not None else None,
                                 loss.ptr, _stream(stream)))
    return float(loss.to_host()[0])


class hip_sched:
    """Scheduler launching one frame of a kernel (cuda_sched<R>::frame,
    detail/cuda_sched.inl:226-246): per-pixel seeds make_random_seed(y*w+x,
    frame_id), frame_id incremented after every frame."""

    def __init__(self):
        _ensure()
        self.frame_id = 0
        self.ray_counts = DeviceArray(2, np.uint64, zero=True)

    def _sp(self, sp: sched_params) -> np.ndarray:
        a = np.zeros(1, SCHED_PARAMS_DTYPE)
        a["width"] = sp.rt.width
        a["height"] = sp.rt.height
        a["spp"] = sp.sample_params.spp
        a["frame_id"] = self.frame_id
        a["sfactor"] = sp.sample_params.sfactor
        a["dfactor"] = sp.sample_params.dfactor
        return a

    def frame(self, kernel, sp: sched_params, stream=None, stats: "DeviceArray | None" = None):
        """One frame; with `stats` (4 x uint64 device array) the counting
        variant also accumulates traversal-node visits and triangle tests."""
        s = self._sp(sp)
        if isinstance(kernel, pathtracing.kernel) and kernel.lights is not None and kernel.lights.count > 0:
            counted = stats is not None
            check(lib().vr_render_pt_lights(kernel.bvh.handle, kernel.materials.ptr, kernel.materials.count,
                                            kernel.lights.ptr, kernel.lights.count, ptr(kernel.params),
                                            ptr(sp.cam), ptr(s), sp.rt.accum.ptr, sp.rt.color.ptr,
                                            (stats if counted else self.ray_counts).ptr, int(counted),
                                            _stream(stream)))
        elif isinstance(kernel, pathtracing.kernel) and stats is not None:
            check(lib().vr_render_pt_counted(kernel.bvh.handle, kernel.materials.ptr, kernel.materials.count,
                                             ptr(kernel.params), ptr(sp.cam), ptr(s), sp.rt.accum.ptr,
                                             sp.rt.color.ptr, stats.ptr, _stream(stream)))
        elif isinstance(kernel, pathtracing.kernel):
            check(lib().vr_render_pt(kernel.bvh.handle, kernel.materials.ptr, kernel.materials.count,
                                     ptr(kernel.params), ptr(sp.cam), ptr(s), sp.rt.accum.ptr, sp.rt.color.ptr,
                                     self.ray_counts.ptr, _stream(stream)))
        elif isinstance(kernel, ao_kernel):
            bg = np.asarray(kernel.bg, np.float32)
            check(lib().vr_render_ao(kernel.bvh.handle, ptr(sp.cam), ptr(s), kernel.samples, float(kernel.radius),
                                     ptr(bg), sp.rt.accum.ptr, sp.rt.color.ptr, _stream(stream)))
        else:
            raise TypeError(f"unsupported kernel {type(kernel).__name__}")
        self.frame_id += 1

    def frame_frozen_rr(self, kernel, rr_materials: "DeviceArray", sp: sched_params, stream=None):
        """vr_render_pt_frozen_rr: a frame whose roulette probabilities follow
        the throughput under `rr_materials` (the estimator the adjoint
        differentiates); frame_id is incremented like frame()."""
        s = self._sp(sp)
        check(lib().vr_render_pt_frozen_rr(kernel.bvh.handle, kernel.materials.ptr, rr_materials.ptr,
                                           kernel.materials.count, ptr(kernel.params), ptr(sp.cam), ptr(s),
                                           sp.rt.accum.ptr, sp.rt.color.ptr, self.ray_counts.ptr, _stream(stream)))
        self.frame_id += 1

    def backward(self, kernel, sp: sched_params, dl_dimage: "DeviceArray", grad: "DeviceArray", stream=None,
                 frame_id: int | None = None):
        """Path-replay adjoint of the frame rendered with `frame_id` (default:
        the last one): accumulates d loss / d (cd rgb, ce rgb) per material
        into `grad` (6 x num_materials float64) given d loss / d image."""
        s = self._sp(sp)
        s["frame_id"] = self.frame_id - 1 if frame_id is None else frame_id
        check(lib().vr_render_pt_backward(kernel.bvh.handle, kernel.materials.ptr, kernel.materials.count,
                                          ptr(kernel.params), ptr(sp.cam), ptr(s), dl_dimage.ptr, grad.ptr,
                                          self.ray_counts.ptr, _stream(stream)))

    def rays(self) -> tuple[int, int]:
        """(primary, secondary) closest-hit queries accumulated so far."""
        c = self.ray_counts.to_host()
        return int(c[0]), int(c[1])

    def reset_counters(self):
        check(lib().vr_memset(self.ray_counts.ptr, 0, self.ray_counts.nbytes))


from .model import model  # noqa: E402  (scene ingestion, src/common/model.h)
