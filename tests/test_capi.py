"""The C-ABI library loads and exports every symbol include/visionaray_amd.h
declares; host-only entry points and argument checks behave as documented.
No compute call is made here (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, load_golden

HEADER = ROOT / "include" / "visionaray_amd.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(vr_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    from visionaray_amd import _lib
    return _lib.load()


def test_library_exports_every_declared_symbol(lib):
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s


def test_python_binding_covers_header():
    from visionaray_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_version(lib):
    from visionaray_amd._lib import check
    a, b = C.c_int(), C.c_int()
    check(lib.vr_version(C.byref(a), C.byref(b)))
    assert (a.value, b.value) >= (0, 1)


def test_camera_basis_matches_reference(lib):
    """vr_camera_look_at is host arithmetic; the reference's pinhole_camera
    basis after begin_frame must match bit for bit."""
    import visionaray_amd as V
    g = load_golden("camera")
    for inp, want in zip(g["inputs"], g["basis"]):
        cam = V.pinhole_camera(inp[0:3], inp[3:6], inp[6:9], inp[9], inp[10])
        assert np.array_equal(cam.view(np.float32).reshape(-1), want)


def test_null_arguments_are_rejected(lib):
    from visionaray_amd._lib import VR_ERR_INVALID_ARGUMENT
    assert lib.vr_lbvh_build(None, 10, 0, None, None) == VR_ERR_INVALID_ARGUMENT
    assert b"null" in lib.vr_last_error()
    assert lib.vr_trace(None, 0, None, 1, None, None) == VR_ERR_INVALID_ARGUMENT
    assert lib.vr_render_pt(None, None, 0, None, None, None, None, None, None, None) == VR_ERR_INVALID_ARGUMENT
    assert lib.vr_camera_look_at(None, None, None, 1.0, 1.0, None) == VR_ERR_INVALID_ARGUMENT
    # the wide-node and device SAH entry points check before touching a device
    assert lib.vr_bvh_set_width(None, 8, 0, None) == VR_ERR_INVALID_ARGUMENT
    assert lib.vr_bvh_download_wide(None, None, 0, None, None) == VR_ERR_INVALID_ARGUMENT
    assert lib.vr_sah_build_device(None, 10, 4, None, None) == VR_ERR_INVALID_ARGUMENT


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode != 0 and "one GPU" in r.stdout


def test_sah_build_matches_reference_tree():
    """vr_sah_build (host) reproduces the reference's binned_sah_builder tree
    node for node (detail/bvh/sah.h, spatial splits off)."""
    import visionaray_amd as V
    g = load_golden("c0_small")
    nodes, idx = V.sah_build(g["tris"], 4)
    want_n, want_i = g["sah_nodes"], g["sah_indices"]
    assert nodes.size == want_n.size
    assert np.array_equal(idx, want_i)
    for k in ("lo", "hi", "first", "num_prims"):
        assert np.array_equal(nodes[k], want_n[k]), k
    inner = want_n["num_prims"] == 0
    assert np.array_equal(nodes["axis"][inner], want_n["axis"][inner])
    assert np.array_equal(nodes["sign"][inner], want_n["sign"][inner])


def test_host_lbvh_matches_reference_cpu_tree():
    """vr_lbvh_build_host reproduces the reference's CPU lbvh_builder tree
    (top-down Morton splits, leaf size 4) node for node."""
    import visionaray_amd as V
    g = load_golden("c0_small")
    nodes, idx = V.lbvh_build_host(g["tris"], 4)
    want_n, want_i = g["cpu_lbvh_nodes"], g["cpu_lbvh_indices"]
    assert nodes.size == want_n.size and np.array_equal(idx, want_i)
    for k in ("lo", "hi", "first", "num_prims"):
        assert np.array_equal(nodes[k], want_n[k]), k


def test_bvh_from_nodes_rejects_bad_trees(lib):
    """vr_bvh_from_nodes checks the host tree before any device call: a child
    index that wraps 32-bit arithmetic, and shared subtrees (a DAG), are
    invalid arguments."""
    from visionaray_amd import scenes as S
    from visionaray_amd._lib import NODE_DTYPE, VR_ERR_INVALID_ARGUMENT, ptr
    tris = np.zeros(2, S.TRI_DTYPE)
    tris["e1"] = (1, 0, 0)
    tris["e2"] = (0, 1, 0)
    idx = np.arange(2, dtype=np.uint32)
    out = C.c_void_p()

    def call(nodes):
        return lib.vr_bvh_from_nodes(ptr(tris), 2, ptr(nodes), nodes.size, ptr(idx), C.byref(out))

    # first = 0xFFFFFFFF: first + 1 wraps to 0 in 32 bits
    nodes = np.zeros(3, NODE_DTYPE)
    nodes["first"][0] = 0xFFFFFFFF
    nodes["first"][1:] = (0, 1)
    nodes["num_prims"][1:] = 1
    assert call(nodes) == VR_ERR_INVALID_ARGUMENT
    assert b"out of range" in lib.vr_last_error()
    # a DAG: inner nodes 1 and 2 share their children (leaves 3 and 4)
    dag = np.zeros(5, NODE_DTYPE)
    dag["first"] = (1, 3, 3, 0, 1)
    dag["num_prims"] = (0, 0, 0, 1, 1)
    assert call(dag) == VR_ERR_INVALID_ARGUMENT
    assert b"reached twice" in lib.vr_last_error()
