"""Scene ingestion: the reference's `model` (src/common/model.h:32-80,
src/common/model.cpp:109-200) over the library's OBJ / PLY loaders.

    mod = model()
    mod.load("scene.obj")            # or a list of files, model::load(filenames)
    mod.primitives                   # (n,) TRI_DTYPE, model::primitives
    mod.geometric_normals            # (n, 3) float32
    mod.shading_normals              # (m, 3) float32, 3 per triangle that has them
    mod.tex_coords                   # (k, 2) float32
    mod.materials                    # (j,) OBJ_MATERIAL_DTYPE, model::materials
    mod.bbox                         # (lo, hi) float32[3] each
    tris, normals = mod.upload()     # device triangles built from the indexed mesh

Parsing and triangle assembly run in libvisionaray_amd.so (host threads for
parsing, a HIP kernel for vr_mesh_upload); failures raise VRError with the
library's message, where the reference prints a warning and returns.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import MESH_INFO_DTYPE, OBJ_MATERIAL_DTYPE, check, ptr
from .scenes import MAT_DTYPE, TRI_DTYPE

__all__ = ["model", "OBJ_MATERIAL_DTYPE", "MESH_INFO_DTYPE"]


def _lib():
    from . import lib
    return lib()


class model:
    """A loaded model (the reference class of the same name)."""

    def __init__(self):
        self._h = C.c_void_p()

    # -- loading ---------------------------------------------------------
    def load(self, filenames) -> bool:
        """model::load: one file name or a list; the format follows the
        extension (.obj/.OBJ, .ply/.PLY).  Replaces what was loaded before."""
        if isinstance(filenames, (str, os.PathLike)):
            filenames = [filenames]
        names = [os.fsencode(f) for f in filenames]
        arr = (C.c_char_p * len(names))(*names)
        h = C.c_void_p()
        check(_lib().vr_mesh_load(arr, len(names), C.byref(h)))
        self._destroy()
        self._h = h
        return True

    def _destroy(self):
        if self._h and self._h.value:
            _lib().vr_mesh_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass

    # -- contents ----------------------------------------------------------
    @property
    def info(self) -> np.void:
        out = np.zeros(1, MESH_INFO_DTYPE)
        check(_lib().vr_mesh_get_info(self._h, ptr(out)))
        return out[0]

    def _fetch(self, fn, count: int, dtype, shape=()):
        out = np.empty((count, *shape), dtype)
        check(fn(self._h, ptr(out) if out.size else None))
        return out

    @property
    def primitives(self) -> np.ndarray:
        return self._fetch(_lib().vr_mesh_triangles, int(self.info["num_prims"]), TRI_DTYPE)

    @property
    def geometric_normals(self) -> np.ndarray:
        return self._fetch(_lib().vr_mesh_geometric_normals, int(self.info["num_prims"]), np.float32, (3,))

    @property
    def shading_normals(self) -> np.ndarray:
        return self._fetch(_lib().vr_mesh_shading_normals, int(self.info["num_shading_normals"]), np.float32, (3,))

    @property
    def tex_coords(self) -> np.ndarray:
        return self._fetch(_lib().vr_mesh_tex_coords, int(self.info["num_tex_coords"]), np.float32, (2,))

    @property
    def materials(self) -> np.ndarray:
        return self._fetch(_lib().vr_mesh_materials, int(self.info["num_materials"]), OBJ_MATERIAL_DTYPE)

    def material_names(self) -> list[str]:
        buf = C.create_string_buffer(4096)
        names = []
        for i in range(int(self.info["num_materials"])):
            check(_lib().vr_mesh_material_name(self._h, i, buf, len(buf)))
            names.append(buf.value.decode(errors="replace"))
        return names

    @property
    def bbox(self) -> tuple[np.ndarray, np.ndarray]:
        i = self.info
        return np.array(i["bbox_lo"], np.float32), np.array(i["bbox_hi"], np.float32)

    def pt_materials(self, diffuse_fallback: bool = False) -> np.ndarray:
        """Path-tracing materials by the viewer's map_material rule
        (emissive / matte; mirror, glass and plastic only as matte with
        diffuse_fallback)."""
        out = np.zeros(int(self.info["num_materials"]), MAT_DTYPE)
        check(_lib().vr_mesh_pt_materials(self._h, int(bool(diffuse_fallback)), ptr(out) if out.size else None))
        return out

    # -- device ------------------------------------------------------------
    def upload(self, stream=None, normals: bool = True):
        """Triangles (and geometric normals) built on the device from the
        indexed mesh: (DeviceArray of TRI_DTYPE, DeviceArray of float32 x 3n
        or None)."""
        from . import DeviceArray, _stream
        n = int(self.info["num_prims"])
        tris = DeviceArray(n, TRI_DTYPE)
        nrm = DeviceArray(3 * n, np.float32) if normals else None
        check(_lib().vr_mesh_upload(self._h, tris.ptr, nrm.ptr if nrm is not None else None, _stream(stream)))
        return tris, nrm
