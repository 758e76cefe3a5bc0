"""ctypes binding of libvisionaray_amd.so (include/visionaray_amd.h).

Loading fails loudly when the HIP library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# VR_LIBRARY selects another build of the same library (e.g. one with the host
# code under AddressSanitizer); default: the in-tree build
LIB_PATH = Path(os.environ.get("VR_LIBRARY") or Path(__file__).resolve().parent / "libvisionaray_amd.so")

VR_OK = 0
VR_ERR_INVALID_ARGUMENT = 1
VR_ERR_OUT_OF_MEMORY = 2
VR_ERR_HIP = 3
VR_ERR_NOT_INITIALIZED = 4
VR_ERR_UNSUPPORTED = 5
VR_BVH_KEEP_NODES = 1
VR_BVH_ROTATE = 2
VR_WIDE_COMPRESSED = 1
VR_WIDE_BINARY_LEAVES = 2

# every symbol the C header declares: name -> (restype, argtypes)
P = C.c_void_p
SZ = C.c_size_t
I = C.c_int
U32 = C.c_uint32
F = C.c_float
SIGNATURES = {
    "vr_init": (I, [I]),
    "vr_last_error": (C.c_char_p, []),
    "vr_version": (I, [C.POINTER(I), C.POINTER(I)]),
    "vr_malloc": (I, [C.POINTER(P), SZ]),
    "vr_free": (I, [P]),
    "vr_memcpy_htod": (I, [P, P, SZ]),
    "vr_memcpy_dtoh": (I, [P, P, SZ]),
    "vr_memset": (I, [P, I, SZ]),
    "vr_synchronize": (I, []),
    "vr_stream_create": (I, [C.POINTER(P)]),
    "vr_stream_destroy": (I, [P]),
    "vr_stream_synchronize": (I, [P]),
    "vr_event_create": (I, [C.POINTER(P)]),
    "vr_event_destroy": (I, [P]),
    "vr_event_record": (I, [P, P]),
    "vr_event_elapsed_ms": (I, [P, P, C.POINTER(F)]),
    "vr_device_copy": (I, [P, P, SZ, P]),
    "vr_device_gather_lines": (I, [P, SZ, P, SZ, P, P]),
    "vr_camera_look_at": (I, [P, P, P, F, F, P]),
    "vr_lbvh_build": (I, [P, SZ, U32, P, C.POINTER(P)]),
    "vr_lbvh_rebuild": (I, [P, P, SZ, P]),
    "vr_lbvh_morton": (I, [P, SZ, P, P]),
    "vr_bvh_sorted_refs": (I, [P, P, P, P]),
    "vr_bvh_from_nodes": (I, [P, SZ, P, SZ, P, C.POINTER(P)]),
    "vr_bvh_refit": (I, [P, P, P]),
    "vr_bvh_set_width": (I, [P, I, U32, P]),
    "vr_bvh_download_wide": (I, [P, P, SZ, C.POINTER(SZ), C.POINTER(U32)]),
    "vr_bvh_get_info": (I, [P, P]),
    "vr_bvh_get_ref": (I, [P, P]),
    "vr_bvh_nodes": (I, [P, P, P]),
    "vr_bvh_indices": (I, [P, P, P]),
    "vr_bvh_destroy": (I, [P]),
    "vr_bvh_destroy_async": (I, [P, P]),
    "vr_sah_build": (I, [P, SZ, I, P, C.POINTER(SZ), P]),
    "vr_sah_build_device": (I, [P, SZ, I, P, C.POINTER(P)]),
    "vr_lbvh_build_host": (I, [P, SZ, I, P, C.POINTER(SZ), P]),
    "vr_trace": (I, [P, I, P, SZ, P, P]),
    "vr_trace_noalloc": (I, [P, I, P, SZ, P, P]),
    "vr_tlas_build": (I, [P, U32, P, U32, P, C.POINTER(P)]),
    "vr_tlas_closest_hit": (I, [P, P, SZ, P, P]),
    "vr_tlas_any_hit": (I, [P, P, SZ, P, P]),
    "vr_tlas_nodes": (I, [P, P, P, C.POINTER(SZ)]),
    "vr_tlas_destroy": (I, [P]),
    "vr_trace_counted": (I, [P, I, P, SZ, P, P, P]),
    "vr_render_ao": (I, [P, P, P, U32, F, P, P, P, P]),
    "vr_render_pt": (I, [P, P, SZ, P, P, P, P, P, P, P]),
    "vr_render_pt_counted": (I, [P, P, SZ, P, P, P, P, P, P, P]),
    "vr_image_l2_loss": (I, [P, P, SZ, P, P, P]),
    "vr_render_pt_backward": (I, [P, P, SZ, P, P, P, P, P, P, P]),
    "vr_render_pt_frozen_rr": (I, [P, P, P, SZ, P, P, P, P, P, P, P]),
    "vr_set_launch_shape": (I, [P]),
    "vr_get_launch_shape": (I, [P]),
    "vr_autotune_pt": (I, [P, P, SZ, P, P, P, P, P, I, P, P, P]),
    "vr_batch_create": (I, [P, SZ, C.POINTER(P)]),
    "vr_batch_render": (I, [P, P, P]),
    "vr_batch_backward": (I, [P, P, P, P]),
    "vr_batch_capture_graph": (I, [P, P, P]),
    "vr_batch_launch_graph": (I, [P, P]),
    "vr_batch_destroy": (I, [P]),
    "vr_create_context": (I, [I, C.POINTER(P)]),
    "vr_destroy_context": (I, [P]),
    "vr_scene_upload": (I, [P, SZ, C.POINTER(P), P]),
    "vr_bvh_download_nodes": (I, [P, P, C.POINTER(SZ)]),
    "vr_trace_closest": (I, [P, P, SZ, P, P]),
    "vr_trace_any": (I, [P, P, SZ, P, P]),
    "vr_render_pt_forward": (I, [P, P, SZ, P, P, P, P, P, P, P]),
    "vr_autotune": (I, [P, P, SZ, P, P, P, P, P, I, P, P, P]),
    "vr_destroy_bvh": (I, [P]),
    "vr_destroy_batch": (I, [P]),
    "vr_last_error_string": (C.c_char_p, []),
    "vr_render_pt_lights": (I, [P, P, SZ, P, SZ, P, P, P, P, P, P, I, P]),
    "vr_bvh_sah_cost_host": (I, [P, SZ, F, F, F, C.POINTER(F)]),
    "vr_bvh_sah_cost": (I, [P, F, F, F, C.POINTER(C.c_double), P]),
    "vr_bvh_optimize_rotations_host": (I, [P, SZ, C.POINTER(I)]),
    "vr_mesh_load": (I, [C.POINTER(C.c_char_p), SZ, C.POINTER(P)]),
    "vr_mesh_get_info": (I, [P, P]),
    "vr_mesh_triangles": (I, [P, P]),
    "vr_mesh_geometric_normals": (I, [P, P]),
    "vr_mesh_shading_normals": (I, [P, P]),
    "vr_mesh_tex_coords": (I, [P, P]),
    "vr_mesh_materials": (I, [P, P]),
    "vr_mesh_material_name": (I, [P, SZ, C.c_char_p, SZ]),
    "vr_mesh_pt_materials": (I, [P, I, P]),
    "vr_mesh_upload": (I, [P, P, P, P]),
    "vr_mesh_destroy": (I, [P]),
}

# record layouts shared with the C header
RAY_DTYPE = np.dtype([("ori", "<f4", (3,)), ("dir", "<f4", (3,)), ("tmin", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim_id", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
NODE_DTYPE = np.dtype([("lo", "<f4", (3,)), ("hi", "<f4", (3,)), ("first", "<u4"), ("num_prims", "<u2"),
                       ("axis", "u1"), ("sign", "u1")])
CAMERA_DTYPE = np.dtype([("eye", "<f4", (3,)), ("U", "<f4", (3,)), ("V", "<f4", (3,)), ("W", "<f4", (3,))])
KERNEL_PARAMS_DTYPE = np.dtype([("num_bounces", "<u4"), ("epsilon", "<f4"), ("bg", "<f4", (3,)),
                                ("ambient", "<f4", (3,))])
SCHED_PARAMS_DTYPE = np.dtype([("width", "<u4"), ("height", "<u4"), ("spp", "<u4"), ("frame_id", "<u4"),
                               ("sfactor", "<f4"), ("dfactor", "<f4")])
SCENE_DTYPE = np.dtype([
    ("bvh", "<u8"), ("materials", "<u8"), ("num_materials", "<u8"),
    ("kp", KERNEL_PARAMS_DTYPE), ("cam", CAMERA_DTYPE), ("sp", SCHED_PARAMS_DTYPE),
    ("accum", "<u8"), ("color", "<u8"), ("dl", "<u8"),
])
BVH_INFO_DTYPE = np.dtype([("num_prims", "<u8"), ("num_nodes", "<u8"), ("depth", "<u4"),
                           ("max_leaf_size", "<u4")])
OBJ_MATERIAL_DTYPE = np.dtype([("ca", "<f4", (3,)), ("cd", "<f4", (3,)), ("cs", "<f4", (3,)), ("ce", "<f4", (3,)),
                               ("ior", "<f4", (3,)), ("transmission", "<f4"), ("specular_exp", "<f4"),
                               ("illum", "<i4")])
MESH_INFO_DTYPE = np.dtype([("num_prims", "<u8"), ("num_vertices", "<u8"), ("num_shading_normals", "<u8"),
                            ("num_tex_coords", "<u8"), ("num_materials", "<u8"), ("num_degenerate", "<u8"),
                            ("num_missing", "<u8"), ("bbox_lo", "<f4", (3,)), ("bbox_hi", "<f4", (3,))])
INSTANCE_DTYPE = np.dtype([("blas", "<u4"), ("inst_id", "<i4"), ("transform", "<f4", (12,))])
INST_HIT_DTYPE = np.dtype([("prim_id", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("inst_id", "<i4"),
                           ("instance", "<i4"), ("geom_id", "<i4"), ("pad", "<i4")])
assert INSTANCE_DTYPE.itemsize == 56 and INST_HIT_DTYPE.itemsize == 32
assert OBJ_MATERIAL_DTYPE.itemsize == 72 and MESH_INFO_DTYPE.itemsize == 80
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16 and NODE_DTYPE.itemsize == 32
assert CAMERA_DTYPE.itemsize == 48 and KERNEL_PARAMS_DTYPE.itemsize == 32 and SCHED_PARAMS_DTYPE.itemsize == 24
assert SCENE_DTYPE.itemsize == 152


class VRError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load(path: Path | str = LIB_PATH):
    """Load the shared library (once) and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path)
    if not p.exists():
        raise RuntimeError(f"{p} is not built: run `python visionaray_amd/build.py` "
                           "(there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int):
    if rc != VR_OK:
        msg = _lib.vr_last_error().decode(errors="replace") if _lib is not None else ""
        raise VRError(rc, msg)


def ptr(a: np.ndarray) -> C.c_void_p:
    """Host pointer of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)
