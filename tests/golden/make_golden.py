"""Writes the golden fixtures of tests/golden/ by running the reference's own
CPU code (oracle/_ref/vr_ref, compiled from the reference headers by
oracle/Makefile) on small seeded inputs.  Run in the container that has the
reference checkout:  python tests/golden/make_golden.py

Each .npz holds inputs and the reference's outputs only (data, no code)."""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
import oracle as O  # noqa: E402
from visionaray_amd import scenes as S  # noqa: E402

OUT = Path(__file__).resolve().parent
FMAX = np.finfo(np.float32).max


def rays_for(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, O.RAY_DTYPE)
    r["ori"] = rng.uniform(-0.9, 0.9, (n, 3))
    d = rng.normal(size=(n, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmax"] = FMAX
    return r


def main():
    assert O.ref_available(), "oracle/_ref/vr_ref is not built (make -C oracle ref)"
    tris = S.config0_scene(level=3)

    # LBVH pieces: the reference's CPU lbvh_builder::init (same Morton formula,
    # std::stable_sort) and its determine_range on those sorted codes
    sc, si = O.ref_morton(tris)
    karras = O.ref_karras(sc)
    # traversal: the reference's intersect() over a GPU-layout LBVH node array
    # (input data here) and over its own binned-SAH tree
    nodes, idx, _, _ = O.lbvh_build(tris)
    rays = rays_for(4096, seed=1)
    sah_nodes, sah_idx = O.ref_sah(tris)
    cpu_nodes, cpu_idx = O.ref_lbvh_cpu(tris)
    np.savez_compressed(
        OUT / "c0_small.npz", tris=tris, ref_sorted_codes=sc, ref_sorted_ids=si, ref_karras=karras,
        lbvh_nodes=nodes, lbvh_indices=idx, rays=rays,
        ref_closest=O.ref_trace(nodes, tris, idx, rays, False),
        ref_any=O.ref_trace(nodes, tris, idx, rays, True),
        ref_brute=O.ref_brute(tris, rays, False),
        sah_nodes=sah_nodes, sah_indices=sah_idx,
        ref_sah_closest=O.ref_trace(sah_nodes, tris, sah_idx, rays, False),
        cpu_lbvh_nodes=cpu_nodes, cpu_lbvh_indices=cpu_idx)

    # seeds and random streams (make_random_seed.h, random_generator.h)
    rng = np.random.default_rng(2)
    pairs = np.concatenate([
        np.array([[0, 0], [1, 0], [0, 1], [-1, -1], [2147483647, 7], [-2147483648, 3]], np.int32),
        rng.integers(-2**31, 2**31 - 1, size=(250, 2)).astype(np.int32)])
    seeds = np.array([0, 1, 2, 12345, 2147483646, 2147483647, 2147483648, 4294967295], np.uint32)
    streams = np.stack([O.ref_rng(int(s), 64) for s in seeds])
    np.savez_compressed(OUT / "rng.npz", pairs=pairs, seeds=O.ref_seed(pairs), stream_seeds=seeds,
                        streams=streams, probe=O.ref_probe(777), probe_seed=np.uint32(777))

    # camera basis after begin_frame
    cams = np.array([
        O.camera_input(*S.config0_camera(512, 512)),
        O.camera_input(*S.config1_camera()),
        O.camera_input((1.0, 2.0, 3.0), (0.5, -0.25, 0.0), (0.0, 0.0, 1.0), 0.7, 1.5),
    ], np.float32)
    np.savez_compressed(OUT / "camera.npz", inputs=cams, basis=np.stack([O.ref_camera(c) for c in cams]))

    # AO (examples/ao) and path tracing (pathtracing::kernel) images under
    # tiled_sched<basic_ray<float>> with the jittered-blend sampler
    w = h = 48
    cam_in = O.camera_input(*S.config0_camera(w, h))
    ao_p = O.ref_params(w, h, spp=2, frame_id=0, ao_samples=8, ao_radius=0.5, bg=(0.1, 0.2, 0.3), threads=4)
    ao = O.ref_ao(nodes, tris, idx, cam_in, ao_p)
    mats = S.lambert_materials(8, seed=3)
    mats["type"][5] = 1
    mats["ce"][5] = (1.0, 0.9, 0.8)
    mats["ls"][5] = 4.0
    tris_m = tris.copy()
    tris_m["geom_id"] = np.arange(tris.size) % 8
    n_m, i_m, _, _ = O.lbvh_build(tris_m)
    pt_p = O.ref_params(w, h, spp=4, frame_id=0, num_bounces=6, epsilon=1e-4, bg=(0.05, 0.05, 0.05),
                        ambient=(0.6, 0.7, 0.9), threads=4)
    pt = O.ref_pt(n_m, tris_m, i_m, mats, cam_in, pt_p)
    np.savez_compressed(OUT / "images.npz", cam_in=cam_in, ao_params=ao_p, ao=ao, mats=mats, tris_m=tris_m,
                        pt_params=pt_p, pt=pt)
    nee()
    progressive()
    quality()
    instances()


def progressive():
    """Progressive accumulation (the viewer's per-frame blend): three frames
    of pathtracing::kernel on one render target under the jittered-blend
    sampler with sfactor = dfactor = 0.5, frames 0, 1, 2 seeded by the
    scheduler's frame counter (sched_common.h:427-526, pixel_access.h:1040-1131)."""
    g = np.load(OUT / "images.npz")
    tris_m, mats, cam_in = g["tris_m"], g["mats"], g["cam_in"]
    n_m, i_m, _, _ = O.lbvh_build(tris_m)
    w = h = 48
    p = O.ref_params(w, h, spp=2, frame_id=2, num_bounces=6, epsilon=1e-4, bg=(0.05, 0.05, 0.05),
                     ambient=(0.6, 0.7, 0.9), sfactor=0.5, dfactor=0.5, threads=4)
    img = O.ref_pt(n_m, tris_m, i_m, mats, cam_in, p)
    np.savez_compressed(OUT / "progressive.npz", params=p, frames=np.int32(3), sfactor=np.float32(0.5),
                        dfactor=np.float32(0.5), pt=img)


def quality():
    """sah_cost and bvh_optimizer::optimize_tree_rotations of the reference
    over the GPU-layout LBVH of the small config-0 scene, and sah_cost of the
    reference's binned-SAH tree."""
    tris = S.config0_scene(level=3)
    nodes, idx, _, _ = O.lbvh_build(tris)
    rot, count = O.ref_rotate(nodes, tris, idx)
    sah_nodes, sah_idx = O.ref_sah(tris)
    np.savez_compressed(OUT / "quality.npz", tris=tris, lbvh_nodes=nodes, lbvh_indices=idx,
                        lbvh_sah=np.float32(O.ref_sah_cost(nodes, tris, idx)), rotated=rot,
                        rotations=np.int32(count), rotated_sah=np.float32(O.ref_sah_cost(rot, tris, idx)),
                        sah_nodes=sah_nodes, sah_indices=sah_idx,
                        sah_tree_sah=np.float32(O.ref_sah_cost(sah_nodes, tris, sah_idx)))


def nee():
    """pathtracing::kernel with an area-light list (make_kernel_params w/
    lights): the lit room (two emissive ceiling triangles as the lights)."""
    tris, mats, lights = S.lit_room_scene(level=3)
    nodes, idx, _, _ = O.lbvh_build(tris)
    w = h = 48
    cam_in = O.camera_input(*S.config0_camera(w, h))
    p = O.ref_params(w, h, spp=4, frame_id=0, num_bounces=6, epsilon=1e-4, bg=(0.05, 0.05, 0.05),
                     ambient=(0.0, 0.0, 0.0), threads=4)
    img = O.ref_ptl(nodes, tris, idx, mats, lights, cam_in, p)
    np.savez_compressed(OUT / "nee.npz", tris=tris, mats=mats, lights=lights, cam_in=cam_in,
                        cam_basis=O.ref_camera(cam_in), params=p, pt=img)


def instances():
    """The reference's two-level BVH (index_bvh_inst_t instances of CPU
    lbvh_builder trees, lbvh_builder over the instances): closest hits of
    random rays, the top-level nodes and order, and each instance's box (the
    top-level root box of that instance alone)."""
    blas, inst = S.instanced_scene()
    tris = np.concatenate(blas)
    offsets = np.cumsum([0] + [b.size for b in blas]).astype(np.uint32)
    rays = S.instanced_rays(4096, seed=12)
    hits, top_nodes, top_idx = O.ref_inst(tris, offsets, inst, rays)
    boxes = np.zeros((inst.size, 6), np.float32)
    for i in range(inst.size):
        _, nodes1, _ = O.ref_inst(tris, offsets, inst[i:i + 1], rays[:1])
        boxes[i, :3], boxes[i, 3:] = nodes1[0]["lo"], nodes1[0]["hi"]
    np.savez_compressed(OUT / "instances.npz", tris=tris, offsets=offsets, instances=inst, rays=rays, hits=hits,
                        top_nodes=top_nodes, top_indices=top_idx, boxes=boxes)


def wide():
    """bvh_collapser into 4- and 8-wide trees (detail/bvh/collapse.h:19-169),
    bvh_compressor::compress (detail/bvh/compress.h, bvh_compressed_node::init,
    bvh.h:352-400) of the wide nodes, and the
    closest hits of the reference's wide traversals (intersect_ray1_bvhN,
    4-wide compressed too) over the GPU-layout LBVH and the reference's own
    binned-SAH tree of the 344-triangle config-0 scene."""
    tris = S.config0_scene(level=2)
    nodes, idx, _, _ = O.lbvh_build(tris)
    sah_nodes, sah_idx = O.ref_sah(tris)
    rays = rays_for(2048, seed=21)
    # rays from outside the scene too, some of them missing it
    far = rays_for(1024, seed=22)
    far["ori"] *= 4.0
    rays = np.concatenate([rays, far])
    out = dict(tris=tris, rays=rays, lbvh_nodes=nodes, lbvh_indices=idx, sah_nodes=sah_nodes, sah_indices=sah_idx)
    for name, nd, ix in (("lbvh", nodes, idx), ("sah", sah_nodes, sah_idx)):
        for w in (4, 8):
            b, c, comp = O.ref_wide(nd, tris, ix, w)
            out[f"{name}{w}_bounds"], out[f"{name}{w}_children"], out[f"{name}{w}_comp"] = b, c, comp
            out[f"{name}{w}_hits"] = O.ref_trace_wide(nd, tris, ix, w, rays)
        out[f"{name}4c_hits"] = O.ref_trace_wide(nd, tris, ix, 4, rays, compressed=True)
    np.savez_compressed(OUT / "wide.npz", **out)


if __name__ == "__main__":
    assert O.ref_available(), "oracle/_ref/vr_ref is not built (make -C oracle ref)"
    if len(sys.argv) > 1:
        for name in sys.argv[1:]:
            globals()[name]()
    else:
        main()
    for f in sorted(OUT.glob("*.npz")):
        print(f.name, f.stat().st_size)
