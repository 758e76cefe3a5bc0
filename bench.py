"""Benchmark of the ray-tracing hot path on one MI355X.

Workload (the large single-scene configuration): a 1448 x 1448-cell fBm
heightfield (4,193,408 triangles) plus 2,000 level-3 icospheres (2,560,000
triangles), 6,753,408 fp32 triangles; GPU LBVH build; 1920 x 1080 path tracing,
16 spp, depth 8, Lambertian albedo U(0.2, 0.8) per triangle, 1% emissive.

A step is one path-traced frame (all 16 spp of every pixel).  `value` is the
incoherent (bounce >= 1) closest-hit rays per second of the timed frames, in
Mrays/s.  LBVH build time is measured separately on the same scene.  The
roofline object is for the path-tracing kernel (HIP events on its stream);
the CPU baseline is the reference's own tiled_sched path (oracle/_ref) on a
bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "incoherent Mrays/s @4M tris; LBVH build ms; adjoint vs FD speedup"
HBM_PEAK_GBS = 8000.0
# MI355X vector ALU: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz, one 32-bit lane
# operation per lane per cycle (a wave64 instruction issues over 4 cycles):
# 39.3 T lane-operations/s, i.e. one wave64 VALU instruction per CU per cycle
CUS, CLOCK_HZ = 256, 2.4e9
VALU_PEAK_LANE_OPS = CUS * 4 * 16 * CLOCK_HZ
# hardware-counter summary of the path-tracing kernel (rocprofv3 --pmc over
# `bench.py --profile-pass`) and the sources whose hash it was collected with
COUNTERS_JSON = ROOT / "profiles" / "pathtrace_c1_counters.json"
KERNEL_SOURCES = ("visionaray_amd/csrc/render.hip", "include/visionaray_amd_core.h")
# algorithmic bytes of one query: ray in, hit out, one 128-byte line per
# traversal-node visit, 48 bytes per triangle test
RAY_IN, HIT_OUT, NODE_BYTES, TRI_BYTES = 32, 16, 128, 48


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_source_sha() -> str:
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update((ROOT / f).read_bytes())
    return h.hexdigest()[:16]


def roofline(kernel_ms: float, alg_bytes: float, per_ray: dict) -> dict:
    """Roofline of the frame kernel.  Bound: VALU issue (the counters show the
    vector ALU busy most of the kernel's time, with a fraction of its lanes
    active, and HBM far below its peak).  achieved = active-lane VALU
    operations per second = SQ_INSTS_VALU x 64 x lane utilisation / kernel
    time, against 39.3 T lane-ops/s; traffic = HBM bytes per launch from
    FETCH_SIZE / WRITE_SIZE scaled by what they read for kernels of known
    traffic (a 128-byte line gather for reads, a copy for writes).  Counts come
    from the committed counter summary, the time from this run's HIP events."""
    out = {"bound": "valu", "achieved": None, "peak": round(VALU_PEAK_LANE_OPS / 1e12, 3), "unit": "Tlane-op/s",
           "frac": None, "traffic": None, "kernel": "k_pathtrace", "kernel_ms": round(kernel_ms, 4)}
    alg_gbs = alg_bytes / (kernel_ms * 1e-3) / 1e9
    hbm = {"algorithmic_bytes_per_launch": int(alg_bytes), "algorithmic_gbs": round(alg_gbs, 1),
           "algorithmic_frac_of_peak": round(alg_gbs / HBM_PEAK_GBS, 4), "per_ray": per_ray,
           "note": "algorithmic bytes count every 128-byte node and 48-byte triangle fetch; L2 / MALL serve "
                   "most of them, so they exceed what HBM moves"}
    try:
        prof = json.loads(COUNTERS_JSON.read_text())
    except (OSError, ValueError):
        out["hbm"] = hbm
        out["error"] = f"no counter summary at {COUNTERS_JSON.relative_to(ROOT)}"
        return out
    cur = kernel_source_sha()
    stale = prof.get("kernel_source_sha") != cur
    c = prof["counters_per_launch"]
    t = kernel_ms * 1e-3
    util = c["SQ_THREAD_CYCLES_VALU"] / (c["SQ_ACTIVE_INST_VALU"] * 64.0)
    lane_ops = c["SQ_INSTS_VALU"] * 64.0 * util
    issue = c["SQ_INSTS_VALU"] / (CUS * t * CLOCK_HZ)
    traffic = prof["hbm_bytes_per_launch"]
    out["achieved"] = round(lane_ops / t / 1e12, 3)
    out["frac"] = round(lane_ops / t / VALU_PEAK_LANE_OPS, 4)
    out["traffic"] = int(traffic)
    out["valu"] = {"issue_frac": round(issue, 4), "lane_utilization": round(util, 4),
                   "insts_valu_per_launch": c["SQ_INSTS_VALU"], "insts_salu_per_launch": c.get("SQ_INSTS_SALU"),
                   "note": "issue_frac = wave64 VALU instructions per CU per cycle (peak 1); frac = issue_frac x "
                           "lane_utilization"}
    hbm.update({"measured_bytes_per_launch": int(traffic), "measured_gbs": round(traffic / t / 1e9, 1),
                "measured_frac_of_peak": round(traffic / t / 1e9 / HBM_PEAK_GBS, 4),
                "calibration": prof.get("calibration")})
    out["hbm"] = hbm
    out["profile"] = {"file": str(COUNTERS_JSON.relative_to(ROOT)), "kernel": prof.get("kernel"),
                      "kernel_source_sha": prof.get("kernel_source_sha"), "current_source_sha": cur,
                      "stale": stale}
    if stale:
        out["profile"]["note"] = "the kernel sources changed since the counters were collected"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1448)
    ap.add_argument("--spheres", type=int, default=2000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--builds", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-spp", type=int, default=4)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--out", type=str, default="")
    ap.add_argument("--config", type=str, default="c1", choices=["c1", "c2", "c3", "c4", "ingest"],
                    help="c1 (default): the headline workload; c2 batched scenes; c3 adjoint vs finite "
                         "differences; c4 dynamic geometry (refit vs rebuild) with batched adjoint; ingest: "
                         "OBJ / PLY scene loading and device triangle assembly")
    ap.add_argument("--grid", type=int, default=2000,
                    help="ingest: the mesh is a grid of grid x grid quads (2 grid^2 triangles)")
    ap.add_argument("--scenes", type=int, default=0, help="c2/c4: number of scenes (default 64 / 32)")
    ap.add_argument("--fd-params", type=int, default=0, help="c3: finite-difference parameters (0 = all)")
    ap.add_argument("--profile-pass", action="store_true",
                    help="for hardware-counter runs: one build, one frame, and one device copy of "
                         "--copy-mb MiB (the calibration kernel of known traffic); prints a short JSON line")
    ap.add_argument("--copy-mb", type=int, default=1024)
    ap.add_argument("--no-extras", action="store_true",
                    help="c1: skip the reference-scheme frame and the heightfield-only run")
    ap.add_argument("--build-only", action="store_true", help="c1: time the LBVH builds alone (--builds of them)")
    ap.add_argument("--no-c2", action="store_true", help="c1: skip the batched-scenes measurement (C2)")
    ap.add_argument("--no-c3", action="store_true", help="c1: skip the adjoint-vs-FD measurement (C3)")
    ap.add_argument("--c2-steps", type=int, default=2, help="c1: timed frames of the C2 measurement")
    ap.add_argument("--trace-spp", type=int, default=4,
                    help="c1: bounce rays per pixel of the incoherent closest-hit batch (vr_trace); 0 skips it")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(the path shards by image tiles or scenes, see DESIGN.md)"}))
        sys.exit(2)

    import visionaray_amd as V
    from visionaray_amd import scenes as S

    if args.config == "ingest":
        line = bench_ingest(args, V)
        print(json.dumps(line), flush=True)
        if args.out:
            Path(args.out).write_text(json.dumps(line) + "\n")
        return
    import bench_workloads as BW
    if args.config != "c1":
        line = {"c2": BW.bench_batched, "c3": BW.bench_adjoint, "c4": BW.bench_dynamic}[args.config](args, V, S, log)
        print(json.dumps(line), flush=True)
        if args.out:
            Path(args.out).write_text(json.dumps(line) + "\n")
        return

    t0 = time.time()
    tris = S.config1_scene(cells=args.cells, spheres=args.spheres)
    mats = S.config1_materials(tris.size)
    log(f"scene: {tris.size} triangles generated in {time.time() - t0:.1f}s")

    V.init(0)
    stream = V.Stream()
    d_tris = V.DeviceArray.from_host(tris)
    d_mats = V.DeviceArray.from_host(mats)

    if args.profile_pass:
        profile_pass(args, V, S, tris, d_tris, d_mats, stream)
        return

    # ---- LBVH build timing (same scene) ----
    builder = V.lbvh_builder()
    build_ms = []
    bvh = None
    for i in range(args.builds + 1):
        if bvh is not None:
            bvh.destroy()
        e0, e1 = V.Event(), V.Event()
        e0.record(stream)
        bvh = builder.build(d_tris, keep_nodes=False, stream=stream)
        e1.record(stream)
        ms = e0.elapsed_ms(e1)
        if i > 0:
            build_ms.append(ms)
    info = bvh.info()
    lbvh_ms = float(np.median(build_ms))
    log(f"LBVH: {lbvh_ms:.3f} ms median of {len(build_ms)} (depth {info['depth']})")
    if args.build_only:
        print(json.dumps({"metric": "LBVH build ms", "value": round(lbvh_ms, 4), "unit": "ms",
                          "higher_is_better": False, "builds": len(build_ms), "depth": info["depth"],
                          "config": {"workload": f"C1 large scene, {tris.size} fp32 triangles, GPU LBVH"}}),
              flush=True)
        return

    # ---- path tracing ----
    eye, center, up, fovy, aspect = S.config1_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, args.width / args.height)
    kp = V.make_kernel_params(num_bounces=args.depth, epsilon=1e-4, bg_color=(0.0, 0.0, 0.0, 1.0),
                              ambient_color=(0.6, 0.7, 0.9, 1.0))
    kernel = V.pathtracing.kernel(bvh, d_mats, kp)
    rt = V.render_target(args.width, args.height)
    sched = V.hip_sched()
    sp = V.make_sched_params(V.jittered_blend_type(spp=args.spp, sfactor=1.0, dfactor=0.0), cam, rt)

    for _ in range(args.warmup):
        sched.frame(kernel, sp, stream=stream)
    stream.synchronize()
    sched.reset_counters()

    frame_ms = []
    V.synchronize()
    wall0 = time.perf_counter()
    for _ in range(args.steps):
        e0, e1 = V.Event(), V.Event()
        e0.record(stream)
        sched.frame(kernel, sp, stream=stream)
        e1.record(stream)
        frame_ms.append(e0.elapsed_ms(e1))
    V.synchronize()
    wall = time.perf_counter() - wall0
    n_primary, n_secondary = sched.rays()
    prim_per, sec_per = n_primary / args.steps, n_secondary / args.steps
    ms_step = wall * 1e3 / args.steps
    kernel_ms = float(np.mean(frame_ms))
    incoherent_mrays = n_secondary / (wall * 1e6)
    total_mrays = (n_primary + n_secondary) / (wall * 1e6)
    log(f"frames: {ms_step:.2f} ms/step wall, {kernel_ms:.2f} ms kernel; {prim_per:.0f} primary + "
        f"{sec_per:.0f} secondary rays per frame; {incoherent_mrays:.1f} incoherent Mrays/s")

    # ---- algorithmic bytes of one frame (counting variant, untimed) ----
    stats = V.DeviceArray(4, np.uint64, zero=True)
    sched.frame_id -= 1  # replay the last timed frame's seeds
    sched.frame(kernel, sp, stream=stream, stats=stats)
    stream.synchronize()
    st = stats.to_host().astype(np.float64)
    rays_frame = st[0] + st[1]
    alg_bytes = rays_frame * (RAY_IN + HIT_OUT) + st[2] * NODE_BYTES + st[3] * TRI_BYTES
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    log(f"per ray: {st[2] / rays_frame:.2f} node visits, {st[3] / rays_frame:.2f} triangles; "
        f"{alg_bytes / 1e9:.2f} GB algorithmic per frame -> {achieved:.0f} GB/s")

    # ---- a batch of incoherent closest-hit queries (vr_trace) ----
    closest, d_rays = trace_batch_bench(args, V, tris, bvh, cam, stream) if args.trace_spp > 0 else (None, None)


    # ---- CPU baseline: the reference's own tiled_sched path ----
    # ---- the same frame on the reference's GPU scheme, and the @4M point ----
    extras = {}
    if not args.no_extras:
        extras["reference_scheme_frame"] = reference_scheme_frame(V, kernel, sp, stream, kernel_ms)
        extras["nee_frame"] = nee_frame(V, S, tris, mats, kernel, sp, stream, kernel_ms)
        extras["lbvh_rotated"] = rotated_lbvh(V, d_tris, bvh, kernel, sp, stream, kernel_ms)
        extras["wide_nodes"] = wide_nodes(V, bvh, kernel, sp, stream, kernel_ms, d_rays, closest)
        extras["sah_device"] = sah_device(V, d_tris, bvh, kernel, sp, stream, kernel_ms)
        extras["heightfield_only"] = heightfield_only(args, V, S, stream)

    cpu, parity = None, None
    if not args.no_cpu_baseline:
        cpu, parity = cpu_baseline(args, tris, mats, V, S, bvh, d_mats, kp, stream)

    # ---- the metric's third part (C3) and the batched launch (C2) ----
    adjoint = None
    if not args.no_c3:
        c3 = BW.bench_adjoint(args, V, S, log, quick=True)
        adjoint = {k: c3[k] for k in ("value", "unit", "adjoint_ms", "forward_loss_ms", "fd_params_timed",
                                      "fd_ms_timed", "fd_ms_all_params", "cosine_vs_plain_fd",
                                      "cosine_at_config_spp")}
        adjoint["workload"] = c3["config"]["workload"]
        adjoint["params"] = c3["config"]["params"]
    batched = None
    if not args.no_c2:
        c2 = BW.bench_batched(args, V, S, log, steps=args.c2_steps)
        batched = {k: c2[k] for k in ("speedup", "speedup_vs_persistent_per_scene", "per_scene_plain_launches",
                                      "per_scene_persistent_launches", "batched_graph", "rays_per_frame")}
        batched["workload"] = c2["config"]["workload"]
        batched["batch_shape"] = c2["config"]["batch_shape"]

    result = {
        "metric": METRIC,
        "value": round(incoherent_mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "C1 large scene: fBm heightfield 1448x1448 cells + 2000 level-3 icospheres, "
                        f"{tris.size} fp32 triangles; GPU LBVH; {args.width}x{args.height} path tracing, "
                        f"{args.spp} spp, depth {args.depth}",
            "triangles": int(tris.size), "width": args.width, "height": args.height, "spp": args.spp,
            "depth": args.depth,
        },
        "lbvh_build_ms": round(lbvh_ms, 4),
        "lbvh_depth": info["depth"],
        "total_mrays_per_s": round(total_mrays, 3),
        "rays_per_frame": {"primary": int(prim_per), "secondary": int(sec_per)},
        "adjoint_vs_fd_speedup": adjoint,
        "batched": batched,
        "parity": parity,
        "closest_hit_batch": closest,
        **extras,
        "roofline": roofline(kernel_ms, alg_bytes, {"nodes": round(st[2] / rays_frame, 3),
                                                    "triangles": round(st[3] / rays_frame, 3)}),
        "cpu_baseline": cpu,
    }
    line = json.dumps(result)
    print(line, flush=True)
    if args.out:
        Path(args.out).write_text(line + "\n")


def reference_scheme_frame(V, kernel, sp, stream, kernel_ms):
    """One frame of the same workload as the reference's GPU path runs it:
    one thread per pixel (16 x 16 pixel blocks), binary-tree traversal, no
    work queue (vr_launch_shape schedule VR_SCHEDULE_PER_PIXEL)."""
    V.set_launch_shape(256, 16, 16, V.SCHEDULE_PER_PIXEL)
    try:
        sched = V.hip_sched()
        sched.frame(kernel, sp, stream=stream)
        stream.synchronize()
        sched.reset_counters()
        e0, e1 = V.Event(), V.Event()
        e0.record(stream)
        sched.frame(kernel, sp, stream=stream)
        e1.record(stream)
        ms = e0.elapsed_ms(e1)
        sec = sched.rays()[1]
    finally:
        V.set_launch_shape(256, 16, 1)
    out = {"ms": round(ms, 3), "incoherent_mrays_per_s": round(sec / (ms * 1e-3) / 1e6, 3),
           "speedup_of_this_build": round(ms / kernel_ms, 3),
           "scheme": "one thread per pixel in 16x16 blocks, binary-tree traversal (cuda_sched + intersect.inl)"}
    log(f"reference-scheme frame: {ms:.2f} ms ({out['incoherent_mrays_per_s']:.1f} incoherent Mrays/s); "
        f"this build x{out['speedup_of_this_build']:.2f}")
    return out


def rotated_lbvh(V, d_tris, bvh, kernel, sp, stream, kernel_ms, builds=3):
    """The LBVH with the optional rotation pass (VR_BVH_ROTATE): build time,
    sah_cost against the plain LBVH, and the same frame on it."""
    import time as _t
    times, rb = [], None
    for _ in range(builds):
        if rb is not None:
            rb.destroy()
        stream.synchronize()
        t0 = _t.perf_counter()
        rb = V.lbvh_builder().build(d_tris, keep_nodes=False, rotate=True, stream=stream)
        stream.synchronize()
        times.append((_t.perf_counter() - t0) * 1e3)
    sah_plain, sah_rot = bvh.sah_cost(stream=stream), rb.sah_cost(stream=stream)
    k = V.pathtracing.kernel(rb, kernel.materials, kernel.params)
    sched = V.hip_sched()
    sched.frame(k, sp, stream=stream)
    stream.synchronize()
    sched.reset_counters()
    e0, e1 = V.Event(), V.Event()
    e0.record(stream)
    sched.frame(k, sp, stream=stream)
    e1.record(stream)
    ms = e0.elapsed_ms(e1)
    sec = sched.rays()[1]
    rb.destroy()
    out = {"build_ms_wall": round(float(np.median(times)), 3), "sah_cost_plain": round(sah_plain, 3),
           "sah_cost_rotated": round(sah_rot, 3), "frame_ms": round(ms, 3),
           "incoherent_mrays_per_s": round(sec / (ms * 1e-3) / 1e6, 3),
           "frame_speedup": round(kernel_ms / ms, 4)}
    log(f"rotated LBVH: build {out['build_ms_wall']:.2f} ms wall, SAH {sah_plain:.2f} -> {sah_rot:.2f}, frame "
        f"{ms:.2f} ms ({out['incoherent_mrays_per_s']:.1f} Mrays/s, x{out['frame_speedup']:.3f})")
    return out


def wide_nodes(V, bvh, kernel, sp, stream, kernel_ms, d_rays, closest, frames=3):
    """The same BVH traversed through 8-wide nodes (vr_bvh_set_width: the
    reference's bvh_collapser into bvh_multi_node<8>, and the 8-bit
    compressed form): collapse time, the frame, node visits per ray and the
    closest-hit batch, against the 4-wide traversal of the headline."""
    import time as _t
    out = {}
    lib = V.lib()
    for name, comp in (("oct_float", False), ("oct_compressed", True)):
        stream.synchronize()
        t0 = _t.perf_counter()
        bvh.set_width(8, compressed=comp, stream=stream)
        collapse_ms = (_t.perf_counter() - t0) * 1e3
        sched = V.hip_sched()
        sched.frame(kernel, sp, stream=stream)
        stream.synchronize()
        sched.reset_counters()
        ms = []
        for _ in range(frames):
            e0, e1 = V.Event(), V.Event()
            e0.record(stream)
            sched.frame(kernel, sp, stream=stream)
            e1.record(stream)
            ms.append(e0.elapsed_ms(e1))
        stream.synchronize()
        sec = sched.rays()[1] / frames
        stats = V.DeviceArray(4, np.uint64, zero=True)
        sched.frame(kernel, sp, stream=stream, stats=stats)
        stream.synchronize()
        st = stats.to_host().astype(np.float64)
        f_ms = float(np.median(ms))
        rec = {"collapse_ms_wall": round(collapse_ms, 3), "frame_ms": round(f_ms, 3),
               "incoherent_mrays_per_s": round(sec / (f_ms * 1e-3) / 1e6, 3),
               "frame_speedup_vs_4wide": round(kernel_ms / f_ms, 4),
               "per_ray": {"nodes": round(st[2] / (st[0] + st[1]), 3), "triangles": round(st[3] / (st[0] + st[1]), 3)}}
        if d_rays is not None and closest is not None:
            n = closest["rays"]
            d_hits = V.DeviceArray(n * 4, np.float32)
            V.check(lib.vr_trace(bvh.handle, 0, d_rays.ptr, n, d_hits.ptr, stream.handle))
            tms = []
            for _ in range(3):
                e0, e1 = V.Event(), V.Event()
                e0.record(stream)
                V.check(lib.vr_trace(bvh.handle, 0, d_rays.ptr, n, d_hits.ptr, stream.handle))
                e1.record(stream)
                tms.append(e0.elapsed_ms(e1))
            stream.synchronize()
            t = float(np.median(tms))
            rec["closest_hit_grays_per_s"] = round(n / (t * 1e-3) / 1e9, 4)
        out[name] = rec
        log(f"8-wide {name}: collapse {collapse_ms:.2f} ms wall, frame {f_ms:.2f} ms "
            f"({rec['incoherent_mrays_per_s']:.1f} Mrays/s, x{rec['frame_speedup_vs_4wide']:.3f}), "
            f"{rec['per_ray']['nodes']:.2f} node visits per ray, closest-hit "
            f"{rec.get('closest_hit_grays_per_s', 0):.3f} Grays/s")
    bvh.set_width(4, stream=stream)
    stream.synchronize()
    return out


def sah_device(V, d_tris, bvh, kernel, sp, stream, kernel_ms, builds=3):
    """The reference's binned SAH tree built on the device
    (vr_sah_build_device, node for node binned_sah_builder's): build time,
    sah_cost against the LBVH's, and the same frame on it."""
    import time as _t
    times, sb = [], None
    for _ in range(builds):
        if sb is not None:
            sb.destroy()
        stream.synchronize()
        t0 = _t.perf_counter()
        sb = V.sah_build_device(d_tris, stream=stream)
        stream.synchronize()
        times.append((_t.perf_counter() - t0) * 1e3)
    cost_lbvh, cost_sah = bvh.sah_cost(stream=stream), sb.sah_cost(stream=stream)
    k = V.pathtracing.kernel(sb, kernel.materials, kernel.params)
    sched = V.hip_sched()
    sched.frame(k, sp, stream=stream)
    stream.synchronize()
    sched.reset_counters()
    ms = []
    for _ in range(3):
        e0, e1 = V.Event(), V.Event()
        e0.record(stream)
        sched.frame(k, sp, stream=stream)
        e1.record(stream)
        ms.append(e0.elapsed_ms(e1))
    stream.synchronize()
    f_ms = float(np.median(ms))
    sec = sched.rays()[1] / len(ms)
    info = sb.info()
    sb.destroy()
    out = {"build_ms_wall": round(float(np.median(times)), 3), "nodes": info["num_nodes"], "depth": info["depth"],
           "sah_cost_lbvh": round(cost_lbvh, 3), "sah_cost_sah": round(cost_sah, 3), "frame_ms": round(f_ms, 3),
           "incoherent_mrays_per_s": round(sec / (f_ms * 1e-3) / 1e6, 3),
           "frame_speedup_vs_lbvh": round(kernel_ms / f_ms, 4)}
    log(f"device binned SAH: build {out['build_ms_wall']:.1f} ms wall, SAH {cost_lbvh:.2f} -> {cost_sah:.2f}, "
        f"frame {f_ms:.2f} ms ({out['incoherent_mrays_per_s']:.1f} Mrays/s, x{out['frame_speedup_vs_lbvh']:.3f})")
    return out


def nee_frame(V, S, tris, mats, kernel, sp, stream, kernel_ms):
    """The same frame with next-event estimation: every emissive triangle of
    the scene is an area light (vr_render_pt_lights); shadow rays count as
    incoherent rays."""
    lights = S.lights_from_emissive(tris, mats)
    d_lights = V.DeviceArray.from_host(lights)
    k = V.pathtracing.kernel(kernel.bvh, kernel.materials, kernel.params, d_lights)
    sched = V.hip_sched()
    sched.frame(k, sp, stream=stream)
    stream.synchronize()
    sched.reset_counters()
    e0, e1 = V.Event(), V.Event()
    e0.record(stream)
    sched.frame(k, sp, stream=stream)
    e1.record(stream)
    ms = e0.elapsed_ms(e1)
    prim, sec = sched.rays()
    out = {"lights": int(lights.size), "ms": round(ms, 3), "rays_per_frame": {"primary": prim, "secondary": sec},
           "incoherent_mrays_per_s": round(sec / (ms * 1e-3) / 1e6, 3),
           "relative_to_plain_frame": round(ms / kernel_ms, 3)}
    log(f"NEE frame ({lights.size} area lights): {ms:.2f} ms, {out['incoherent_mrays_per_s']:.1f} incoherent "
        f"Mrays/s (closest + shadow)")
    return out


def heightfield_only(args, V, S, stream, steps=2):
    """The literal @4M point: the 1448 x 1448-cell heightfield alone
    (4,193,408 triangles), LBVH build and the same frame."""
    tris = S.config1_scene(cells=args.cells, spheres=0)
    mats = S.config1_materials(tris.size)
    d_tris = V.DeviceArray.from_host(tris)
    d_mats = V.DeviceArray.from_host(mats)
    builder = V.lbvh_builder()
    ms_b = []
    bvh = None
    for i in range(4):
        if bvh is not None:
            bvh.destroy()
        e0, e1 = V.Event(), V.Event()
        e0.record(stream)
        bvh = builder.build(d_tris, keep_nodes=False, stream=stream)
        e1.record(stream)
        if i > 0:
            ms_b.append(e0.elapsed_ms(e1))
    eye, center, up, fovy, _ = S.config1_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, args.width / args.height)
    kp = V.make_kernel_params(num_bounces=args.depth, epsilon=1e-4, bg_color=(0.0, 0.0, 0.0, 1.0),
                              ambient_color=(0.6, 0.7, 0.9, 1.0))
    kernel = V.pathtracing.kernel(bvh, d_mats, kp)
    rt = V.render_target(args.width, args.height)
    sched = V.hip_sched()
    sp = V.make_sched_params(V.jittered_blend_type(spp=args.spp, sfactor=1.0, dfactor=0.0), cam, rt)
    sched.frame(kernel, sp, stream=stream)
    stream.synchronize()
    sched.reset_counters()
    e0, e1 = V.Event(), V.Event()
    e0.record(stream)
    for _ in range(steps):
        sched.frame(kernel, sp, stream=stream)
    e1.record(stream)
    ms = e0.elapsed_ms(e1) / steps
    _, sec = sched.rays()
    out = {"triangles": int(tris.size), "lbvh_build_ms": round(float(np.median(ms_b)), 4),
           "frame_ms": round(ms, 3), "incoherent_mrays_per_s": round(sec / steps / (ms * 1e-3) / 1e6, 3)}
    log(f"heightfield only ({tris.size} tris): LBVH {out['lbvh_build_ms']:.3f} ms, frame {ms:.2f} ms, "
        f"{out['incoherent_mrays_per_s']:.1f} incoherent Mrays/s")
    bvh.destroy()
    return out


def cosine_about(nrm, rng):
    """cosine-distributed unit directions about the unit normals nrm (n x 3)"""
    n = nrm.shape[0]
    t = np.where(np.abs(nrm[:, 0:1]) > 0.9, np.array([[0, 1, 0]], np.float32), np.array([[1, 0, 0]], np.float32))
    u = np.cross(t, nrm)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w = np.cross(nrm, u)
    r1, r2 = rng.random(n, dtype=np.float32), rng.random(n, dtype=np.float32)
    r, phi = np.sqrt(r1), 2 * np.pi * r2
    d = u * (r * np.cos(phi))[:, None] + w * (r * np.sin(phi))[:, None] + nrm * np.sqrt(1 - r1)[:, None]
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def secondary_rays(V, tris, bvh, cam, width, height, spp, seed=5):
    """The first diffuse bounce of a frame: jittered primary rays through
    every pixel (spp per pixel, pixel order), traced with vr_trace; at each
    hit a cosine-distributed direction about the face normal turned toward
    the viewer, origin offset 1e-4 along it.  Returns the bounce rays."""
    rng = np.random.default_rng(seed)
    n = width * height * spp
    pix = np.repeat(np.arange(width * height), spp)
    x = (pix % width).astype(np.float32) + rng.random(n, dtype=np.float32)
    y = (pix // width).astype(np.float32) + rng.random(n, dtype=np.float32)
    c = cam[0]
    d = (c["U"][None, :] * (2 * x / width - 1)[:, None] + c["V"][None, :] * (2 * y / height - 1)[:, None]
         + c["W"][None, :])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    prim = np.zeros(n, V.RAY_DTYPE)
    prim["ori"] = c["eye"]
    prim["dir"] = d
    prim["tmax"] = np.finfo(np.float32).max
    hits = V.closest_hit(bvh, prim)
    hit = hits["prim_id"] >= 0
    k = hits["prim_id"][hit]  # prim_id is the triangle index in these scenes
    pos = prim["ori"][hit] + prim["dir"][hit] * hits["t"][hit][:, None]
    nrm = np.cross(tris["e1"][k], tris["e2"][k])
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    nrm *= np.where(np.sum(nrm * prim["dir"][hit], axis=1) > 0, -1.0, 1.0).astype(np.float32)[:, None]
    dirs = cosine_about(nrm, rng)
    rays = np.zeros(int(hit.sum()), V.RAY_DTYPE)
    rays["ori"] = pos + 1e-4 * dirs
    rays["dir"] = dirs
    rays["tmax"] = np.finfo(np.float32).max
    return rays


def trace_batch_bench(args, V, tris, bvh, cam, stream, reps=3):
    """vr_trace closest-hit over the first diffuse bounce of a 1920x1080 frame
    at 4 rays per pixel (rays in device memory): queries per second (HIP
    events) and node / triangle visits per query."""
    t0 = time.time()
    rays = secondary_rays(V, tris, bvh, cam, args.width, args.height, args.trace_spp)
    d_rays = V.DeviceArray.from_host(rays)
    d_hits = V.DeviceArray(rays.size * 4, np.float32)
    lib = V.lib()
    log(f"closest-hit batch: {rays.size} bounce rays generated in {time.time() - t0:.1f}s")
    V.check(lib.vr_trace(bvh.handle, 0, d_rays.ptr, rays.size, d_hits.ptr, stream.handle))
    ms = []
    for _ in range(reps):
        e0, e1 = V.Event(), V.Event()
        e0.record(stream)
        V.check(lib.vr_trace(bvh.handle, 0, d_rays.ptr, rays.size, d_hits.ptr, stream.handle))
        e1.record(stream)
        ms.append(e0.elapsed_ms(e1))
    counts = V.DeviceArray(2 * rays.size, np.uint32)
    V.check(lib.vr_trace_counted(bvh.handle, 0, d_rays.ptr, rays.size, d_hits.ptr, counts.ptr, stream.handle))
    stream.synchronize()
    c = counts.to_host().reshape(-1, 2).astype(np.float64)
    hits = d_hits.to_host().view(np.int32).reshape(-1, 4)[:, 0]
    t = float(np.median(ms))
    out = {"rays": int(rays.size), "ms": round(t, 4), "grays_per_s": round(rays.size / (t * 1e-3) / 1e9, 4),
           "hit_fraction": round(float(np.mean(hits >= 0)), 4),
           "per_ray": {"nodes": round(float(c[:, 0].mean()), 3), "triangles": round(float(c[:, 1].mean()), 3)},
           "rays_kind": f"first diffuse bounce of a {args.width}x{args.height} frame at {args.trace_spp} rays "
                        "per pixel, pixel order"}
    log(f"closest-hit batch: {t:.2f} ms -> {out['grays_per_s']:.3f} Grays/s; {out['per_ray']['nodes']:.2f} node "
        f"visits, {out['per_ray']['triangles']:.2f} triangles per ray; {out['hit_fraction']:.3f} hit")
    if not args.no_extras:
        out["instanced"] = instanced_batch(args, V, d_rays, rays.size, hits, t, stream)
    return out, d_rays


def instanced_batch(args, V, d_rays, n, flat_prim, flat_ms, stream, reps=3):
    """The same bounce rays through the C1 scene as a two-level BVH (the
    height field as one instance, 2000 instances of one 1280-triangle
    icosphere): vr_tlas_closest_hit queries per second, and how many rays
    agree with the flat scene on hit / miss."""
    from visionaray_amd import scenes as S
    t0 = time.time()
    blas_tris, inst = S.config1_instanced(cells=args.cells, spheres=args.spheres)
    blas = [V.lbvh_builder().build(b, keep_nodes=False, stream=stream) for b in blas_tris]
    e0, e1 = V.Event(), V.Event()
    tl = V.tlas(blas, inst, stream=stream)
    stream.synchronize()
    build_s = time.time() - t0
    d_hits = V.DeviceArray(n, V.INST_HIT_DTYPE)
    tl.closest_hit_device(d_rays, d_hits, stream)
    ms = []
    for _ in range(reps):
        e0, e1 = V.Event(), V.Event()
        e0.record(stream)
        tl.closest_hit_device(d_rays, d_hits, stream)
        e1.record(stream)
        ms.append(e0.elapsed_ms(e1))
    h = d_hits.to_host()
    t = float(np.median(ms))
    tris = sum(int(b.size) for b in blas_tris)
    out = {"instances": int(inst.size), "bottom_level_triangles": tris, "ms": round(t, 4),
           "grays_per_s": round(n / (t * 1e-3) / 1e9, 4), "relative_to_flat": round(flat_ms / t, 4),
           "hit_miss_agreement": round(float(np.mean((h["prim_id"] >= 0) == (flat_prim >= 0))), 5),
           "scene": "C1 as a two-level BVH: the height field as one identity instance plus 2000 instances of one "
                    "unit level-3 icosphere (the flat scene's radii and centres)"}
    log(f"instanced: {inst.size} instances over {tris} bottom-level triangles (built in {build_s:.1f}s): "
        f"{t:.2f} ms -> {out['grays_per_s']:.3f} Grays/s (x{out['relative_to_flat']:.2f} of the flat scene), "
        f"hit/miss agreement {out['hit_miss_agreement']:.4f}")
    tl.destroy()
    for b in blas:
        b.destroy()
    return out


def profile_pass(args, V, S, tris, d_tris, d_mats, stream):
    """One LBVH build, one counted and one plain path-traced frame, a copy of
    known size and a gather of known size (every 128-byte line of 4 GiB once,
    random order): the dispatches a hardware-counter run reads.  FETCH_SIZE /
    WRITE_SIZE of the copy and the gather calibrate the counters against bytes
    moved, for streaming and for line-gather access."""
    bvh = V.lbvh_builder().build(d_tris, keep_nodes=False, stream=stream)
    eye, center, up, fovy, _ = S.config1_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, args.width / args.height)
    kp = V.make_kernel_params(num_bounces=args.depth, epsilon=1e-4, bg_color=(0.0, 0.0, 0.0, 1.0),
                              ambient_color=(0.6, 0.7, 0.9, 1.0))
    rt = V.render_target(args.width, args.height)
    sched = V.hip_sched()
    sp = V.make_sched_params(V.jittered_blend_type(spp=args.spp), cam, rt)
    stats = V.DeviceArray(4, np.uint64, zero=True)
    sched.frame(V.pathtracing.kernel(bvh, d_mats, kp), sp, stream=stream, stats=stats)
    sched.frame(V.pathtracing.kernel(bvh, d_mats, kp), sp, stream=stream)
    nbytes = args.copy_mb * (1 << 20)
    src = V.DeviceArray(nbytes // 4, np.float32, zero=True)
    dst = V.DeviceArray(nbytes // 4, np.float32)
    V.check(V.lib().vr_device_copy(dst.ptr, src.ptr, nbytes, stream.handle))
    stream.synchronize()
    del src, dst
    # gather calibration: every 128-byte line of a 4 GiB buffer exactly once,
    # in random order (16x the L2 + MALL capacity, so no line is reused)
    gbytes = 4 << 30
    nlines = gbytes // 128
    gsrc = V.DeviceArray(gbytes // 4, np.uint32)
    V.check(V.lib().vr_memset(gsrc.ptr, 0xA5, gbytes))
    lines = V.DeviceArray.from_host(np.random.default_rng(0).permutation(nlines).astype(np.uint32))
    gout = V.DeviceArray(1, np.uint32, zero=True)
    V.check(V.lib().vr_device_gather_lines(gsrc.ptr, gbytes, lines.ptr, nlines, gout.ptr, stream.handle))
    stream.synchronize()
    st = stats.to_host().astype(np.float64)
    rays = st[0] + st[1]
    alg = rays * (RAY_IN + HIT_OUT) + st[2] * NODE_BYTES + st[3] * TRI_BYTES
    print(json.dumps({"profile_pass": True, "copy_bytes": nbytes, "gather_bytes": gbytes + 4 * nlines,
                      "gather_lines": nlines, "pathtrace_algorithmic_bytes": int(alg),
                      "rays": int(rays), "nodes": int(st[2]), "triangles": int(st[3])}), flush=True)


def cpu_baseline(args, tris, mats, V, S, bvh, d_mats, kp, stream):
    """The leg that runs the oracle/ code, as baseline and checker only.

    Baseline: the reference's tiled_sched + pathtracing::kernel (compiled from
    the reference headers into oracle/_ref) timed on a bounded sample, the same
    scene, camera and depth at --cpu-spp samples per pixel; rays/s uses the
    query count of the GPU run of the identical sample (same per-pixel seeds).

    Parity: the GPU frame (16 spp, depth 8, frame 0) against the oracle's
    restatement of the same frame on four bands of four rows, per pixel within
    4 sigma (sigma from four more GPU frames), with the global relative RMSE.
    Returns (cpu_baseline, parity)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    try:
        import oracle as O
    except Exception as e:  # pragma: no cover
        return {"value": None, "error": f"oracle unavailable: {e}"}, None
    parity = None
    try:
        parity = frame_parity(args, tris, mats, V, S, bvh, d_mats, kp, stream, O)
    except Exception as e:  # the checker must not take the measurement down
        parity = {"error": str(e)}
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    w, h = args.width, args.height
    eye, center, up, fovy, _ = S.config1_camera()
    cam_in = O.camera_input(eye, center, up, fovy, w / h)
    kpv = kp[0]
    # GPU query count for the same sample
    cam = V.pinhole_camera(eye, center, up, fovy, w / h)
    rt = V.render_target(w, h)
    sched = V.hip_sched()
    sp = V.make_sched_params(V.jittered_blend_type(spp=args.cpu_spp), cam, rt)
    sched.frame(V.pathtracing.kernel(bvh, d_mats, kp), sp, stream=stream)
    stream.synchronize()
    n_primary, n_secondary = sched.rays()
    if O.ref_available():
        params = O.ref_params(w, h, spp=args.cpu_spp, num_bounces=int(kpv["num_bounces"]),
                              epsilon=float(kpv["epsilon"]), bg=tuple(kpv["bg"]), ambient=tuple(kpv["ambient"]),
                              threads=threads)
        log(f"cpu baseline: reference tiled_sched, {threads} threads, {w}x{h} at {args.cpu_spp} spp")
        t0 = time.time()
        r = O.ref_bench(tris, mats, cam_in, params, packet=4, build="lbvh", frames=5, timeout=300, stall=60.0)
        log(f"cpu baseline done in {time.time() - t0:.1f}s: frame {r['frame_s_median']:.3f}s, "
            f"build {r['build_s']:.2f}s")
        frame_s = r["frame_s_median"]
        extra = cpu_baseline_variants(args, tris, mats, V, S, bvh, d_mats, kp, stream, O, cam_in, threads)
        return {
            "value": round(n_secondary / frame_s / 1e6, 4),
            "unit": "Mrays/s",
            "cores": threads,
            "kind": "reference",
            "sample": f"same scene and camera, {w}x{h} at {args.cpu_spp} spp, depth {int(kpv['num_bounces'])}; "
                      f"tiled_sched<basic_ray<simd::float4>> + pathtracing::kernel over the reference's CPU "
                      f"lbvh_builder tree (leaf size 4); median of 5 frames after 1 warm-up; rays counted "
                      f"on the GPU run of the identical sample",
            "frame_s": round(frame_s, 4),
            "total_mrays_per_s": round((n_primary + n_secondary) / frame_s / 1e6, 4),
            "cpu_lbvh_build_s": round(r["build_s"], 3),
            "binary": O.ref_binary().name,
            "host": host_info(),
            **extra,
            **({"stopped": r["stopped"], "timed_frames": r["frames"]} if "stopped" in r else {}),
        }, parity
    return {"value": None, "error": "oracle/_ref not built"}, parity


def cpu_baseline_variants(args, tris, mats, V, S, bvh, d_mats, kp, stream, O, cam_in, threads):
    """The reference's CPU path with scalar rays (tiled_sched<basic_ray<float>>)
    on the same scene at 1 spp, and its binned_sah_builder on the config-0
    scene (single-threaded, as written)."""
    out = {}
    w, h = args.width, args.height
    kpv = kp[0]
    try:
        eye, center, up, fovy, _ = S.config1_camera()
        cam = V.pinhole_camera(eye, center, up, fovy, w / h)
        rt = V.render_target(w, h)
        sched = V.hip_sched()
        sched.frame(V.pathtracing.kernel(bvh, d_mats, kp), V.make_sched_params(V.jittered_blend_type(spp=1), cam, rt),
                    stream=stream)
        stream.synchronize()
        _, sec = sched.rays()
        params = O.ref_params(w, h, spp=1, num_bounces=int(kpv["num_bounces"]), epsilon=float(kpv["epsilon"]),
                              bg=tuple(kpv["bg"]), ambient=tuple(kpv["ambient"]), threads=threads)
        r = O.ref_bench(tris, mats, cam_in, params, packet=1, build="lbvh", frames=5, timeout=300, stall=60.0)
        out["scalar_rays"] = {"value": round(sec / r["frame_s_median"] / 1e6, 4), "unit": "Mrays/s",
                              "sample": f"same scene and camera, {w}x{h} at 1 spp; tiled_sched<basic_ray<float>>; "
                                        "median of 5 frames after 1 warm-up",
                              "frame_s": round(r["frame_s_median"], 4)}
        log(f"cpu baseline, scalar rays: frame {r['frame_s_median']:.3f}s at 1 spp "
            f"({out['scalar_rays']['value']:.2f} incoherent Mrays/s)")
    except Exception as e:
        out["scalar_rays"] = {"error": str(e)}
    try:
        c0 = S.config0_scene(level=5)
        c0m = S.lambert_materials(1, seed=3)
        c0["geom_id"] = 0
        cin = O.camera_input(*S.config0_camera(64, 64))
        p0 = O.ref_params(64, 64, spp=1, num_bounces=2, epsilon=1e-4, threads=threads)
        r0 = O.ref_bench(c0, c0m, cin, p0, packet=1, build="sah", frames=1, timeout=120, stall=60.0)
        t0 = time.perf_counter()
        V.sah_build(c0, 4)
        ours = time.perf_counter() - t0
        out["c0_sah_build"] = {"reference_s": round(r0["build_s"], 4), "this_build_host_s": round(ours, 4),
                               "triangles": int(c0.size),
                               "note": "binned_sah_builder, 16 bins, no spatial splits, leaf size 4; the "
                                       "reference's single-threaded build against vr_sah_build (same tree)"}
        log(f"cpu baseline, C0 SAH build of {c0.size} triangles: reference {r0['build_s'] * 1e3:.1f} ms, "
            f"this build's host SAH {ours * 1e3:.1f} ms")
    except Exception as e:
        out["c0_sah_build"] = {"error": str(e)}
    return out


def host_info():
    """CPU model, core count and compiler of the host running the baseline."""
    info = {"nproc": os.cpu_count()}
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        import subprocess
        out = subprocess.run(["g++", "--version"], capture_output=True, text=True, timeout=10).stdout
        info["compiler"] = out.splitlines()[0] if out else None
    except Exception:
        info["compiler"] = None
    return info


PARITY_BANDS = [(300, 304), (540, 544), (760, 764), (1060, 1064)]


def frame_parity(args, tris, mats, V, S, bvh, d_mats, kp, stream, O):
    """GPU frame 0 against the oracle's frame 0 on PARITY_BANDS rows."""
    t0 = time.time()
    w, h = args.width, args.height
    eye, center, up, fovy, _ = S.config1_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, w / h)
    rt = V.render_target(w, h)
    sched = V.hip_sched()
    sp = V.make_sched_params(V.jittered_blend_type(spp=args.spp), cam, rt)
    kern = V.pathtracing.kernel(bvh, d_mats, kp)
    rows = [r for a, b in PARITY_BANDS if b <= h for r in range(a, b)]
    sched.frame(kern, sp, stream=stream)
    stream.synchronize()
    got = rt.color_host()[rows]
    frames = []
    for _ in range(4):
        sched.frame(kern, sp, stream=stream)
        stream.synchronize()
        frames.append(rt.color_host()[rows])
    sigma = np.stack(frames).std(axis=0)
    nodes, idx, _, _ = O.lbvh_build(tris)
    kpv = kp[0]
    p = O.render_params(w, h, spp=args.spp, num_bounces=int(kpv["num_bounces"]), epsilon=float(kpv["epsilon"]),
                        bg=tuple(kpv["bg"]), ambient=tuple(kpv["ambient"]))
    cam_o = cam.astype(O.CAMERA_DTYPE)
    want = []
    for a, b in PARITY_BANDS:
        if b <= h:
            img, _ = O.render_pt(nodes, tris, idx, mats, cam_o, p, rows=(a, b))
            want.append(img[a:b])
    want = np.concatenate(want)
    exact = float(np.mean(np.all(got == want, axis=-1)))
    diff = np.abs(got[..., :3].astype(np.float64) - want[..., :3])
    within = float(np.mean(np.all(diff <= 4.0 * sigma[..., :3] + 1e-6, axis=-1)))
    rel = float(np.sqrt(np.mean(diff ** 2)) / max(np.sqrt(np.mean(want[..., :3].astype(np.float64) ** 2)), 1e-300))
    out = {"pixels": int(got.shape[0] * got.shape[1]), "rows": [list(b) for b in PARITY_BANDS],
           "exact_pixel_fraction": round(exact, 5), "within_4_sigma": round(within, 5),
           "global_rel_rmse": float(f"{rel:.4e}"),
           "against": "the oracle's restatement of the reference's path tracer, frame 0, same seeds",
           "check_s": round(time.time() - t0, 2)}
    log(f"parity: {out['pixels']} pixels of frame 0: exact {exact:.4f}, within 4 sigma {within:.4f}, "
        f"global rel RMSE {rel:.3e}")
    return out


# ---------------------------------------------------------------------------
# ingest: OBJ / PLY loading (host threads) and triangle assembly (k_assemble)
# ---------------------------------------------------------------------------

def grid_quads(n: int, seed: int = 3):
    """An n x n-quad height field: (n+1)^2 vertices, n^2 quads (a, b, c, d)."""
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(n + 1, dtype=np.float32), np.arange(n + 1, dtype=np.float32), indexing="ij")
    zs = (rng.normal(size=xs.shape) * 0.25).astype(np.float32)
    pos = np.stack([xs, ys, zs], -1).reshape(-1, 3) * np.float32(1.0 / n)
    i = (np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]).reshape(-1)
    return pos, np.stack([i, i + n + 1, i + n + 2, i + 1], -1)


def write_obj(path: Path, pos, quads):
    """v / vt / vn per vertex and one 4-corner face v/t/n per quad (1-based)."""
    with open(path, "w") as f:
        f.write("# grid\n")
        f.write("".join(f"v {x:.7g} {y:.7g} {z:.7g}\n" for x, y, z in pos.tolist()))
        f.write("".join(f"vt {x:.7g} {y:.7g}\n" for x, y in pos[:, :2].tolist()))
        f.write("vn 0 0 1\n" * len(pos))
        f.write("".join(f"f {a}/{a}/{a} {b}/{b}/{b} {c}/{c}/{c} {d}/{d}/{d}\n"
                        for a, b, c, d in (quads + 1).tolist()))


def write_ply(path: Path, pos, quads):
    tris = np.concatenate([quads[:, [0, 1, 2]], quads[:, [0, 2, 3]]])
    head = (f"ply\nformat binary_little_endian 1.0\nelement vertex {len(pos)}\nproperty float x\n"
            f"property float y\nproperty float z\nelement face {len(tris)}\n"
            "property list uchar int vertex_indices\nend_header\n")
    rec = np.zeros(len(tris), np.dtype([("n", "u1"), ("i", "<i4", (3,))]))
    rec["n"] = 3
    rec["i"] = tris
    with open(path, "wb") as f:
        f.write(head.encode())
        f.write(np.ascontiguousarray(pos, "<f4").tobytes())
        f.write(rec.tobytes())


def bench_ingest(args, V):
    """Scene files to device triangles: vr_mesh_load of an OBJ (all host
    threads and one) and of a binary PLY, then vr_mesh_upload (PCIe copy of
    the indexed mesh + k_assemble) and LBVH build.  k_assemble's roofline:
    16 B read and 56 B written per triangle plus 12 B per vertex."""
    import tempfile
    from visionaray_amd.model import model
    n = args.grid
    pos, quads = grid_quads(n)
    tmp = Path(tempfile.mkdtemp(prefix="vr_ingest_"))
    t0 = time.time()
    obj, ply = tmp / "grid.obj", tmp / "grid.ply"
    write_obj(obj, pos, quads)
    write_ply(ply, pos, quads)
    obj_mb, ply_mb = obj.stat().st_size / 1e6, ply.stat().st_size / 1e6
    log(f"ingest: {2 * n * n} triangles; OBJ {obj_mb:.0f} MB, PLY {ply_mb:.0f} MB written in {time.time() - t0:.1f}s")
    threads = min(16, os.cpu_count() or 1)

    def load(path, nthreads, reps):
        os.environ["VR_IO_THREADS"] = str(nthreads)
        ts, m = [], None
        for _ in range(reps):
            m = model()
            t = time.perf_counter()
            m.load(str(path))
            ts.append(time.perf_counter() - t)
        return m, float(np.median(ts))

    m_obj, obj_s = load(obj, threads, 3)
    log(f"ingest: OBJ {obj_s * 1e3:.0f} ms with {threads} threads")
    _, obj1_s = load(obj, 1, 1)
    log(f"ingest: OBJ {obj1_s * 1e3:.0f} ms with 1 thread")
    m_ply, ply_s = load(ply, threads, 3)
    log(f"ingest: PLY {ply_s * 1e3:.0f} ms")
    os.environ.pop("VR_IO_THREADS", None)
    ntri = int(m_obj.info["num_prims"])
    nv = int(m_obj.info["num_vertices"])
    assert ntri == 2 * n * n

    V.init(0)
    st = V.Stream()
    tris = V.DeviceArray(ntri, m_obj.primitives.dtype)
    nrm = V.DeviceArray(3 * ntri, np.float32)
    lib = V.lib()
    from visionaray_amd._lib import check

    def upload():
        check(lib.vr_mesh_upload(m_obj._h, tris.ptr, nrm.ptr, st.handle))

    host_tris = m_obj.primitives
    # the same triangles copied as 44-byte records (what a host-side loader
    # hands over); this also warms up the runtime's pageable-copy path
    t = time.perf_counter()
    d_plain = V.DeviceArray.from_host(host_tris)
    plain_ms = (time.perf_counter() - t) * 1e3
    del d_plain
    # first upload: the indexed mesh over PCIe (16 B per triangle + 12 B per
    # vertex), then the kernel
    st.synchronize()
    t = time.perf_counter()
    upload()
    st.synchronize()
    first_ms = (time.perf_counter() - t) * 1e3
    # k_assemble alone: one event pair around each launch; the stream is kept
    # busy by the previous launch so each pair brackets just its kernel
    reps = 20
    evs = [(V.Event(), V.Event()) for _ in range(reps + 1)]
    for a, b in evs:
        a.record(st)
        upload()
        b.record(st)
    st.synchronize()
    kern_ms = float(np.mean([a.elapsed_ms(b) for a, b in evs[1:]]))
    ok = tris.to_host().tobytes() == host_tris.tobytes()
    # file to BVH
    t = time.perf_counter()
    m2 = model()
    os.environ["VR_IO_THREADS"] = str(threads)
    m2.load(str(obj))
    os.environ.pop("VR_IO_THREADS", None)
    d2, _ = m2.upload(stream=st, normals=False)
    bvh = V.lbvh_builder().build(d2, keep_nodes=False, stream=st)
    st.synchronize()
    file_to_bvh_s = time.perf_counter() - t
    bvh.destroy()
    alg_bytes = ntri * (16 + 56) + nv * 12
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    log(f"ingest: upload first {first_ms:.2f} ms, k_assemble {kern_ms * 1e3:.1f} us ({achieved:.0f} GB/s), "
        f"44-B records over PCIe {plain_ms:.1f} ms; file->BVH {file_to_bvh_s * 1e3:.0f} ms; bit-identical {ok}")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = ingest_cpu_baseline(tmp, pos, quads)
    for p in (obj, ply):
        p.unlink()
    tmp.rmdir()
    return {
        "metric": "OBJ scene ingestion Mtriangles/s",
        "value": round(ntri / obj_s / 1e6, 3), "unit": "Mtriangles/s", "n_gpus": 1, "steps": 3, "warmup": 0,
        "ms_per_step": round(obj_s * 1e3, 3), "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n}x{n}-quad height field as OBJ (v/vt/vn, 4-corner v/t/n faces, "
                               f"{obj_mb:.0f} MB) and binary PLY ({ply_mb:.0f} MB); {ntri} triangles",
                   "triangles": ntri, "vertices": nv, "host_threads": threads},
        "obj_ms": round(obj_s * 1e3, 2), "obj_mb_per_s": round(obj_mb / obj_s, 1),
        "obj_1thread_ms": round(obj1_s * 1e3, 2), "ply_ms": round(ply_s * 1e3, 2),
        "ply_mb_per_s": round(ply_mb / ply_s, 1),
        "upload_first_ms": round(first_ms, 3), "upload_44B_records_ms": round(plain_ms, 3),
        "file_to_bvh_ms": round(file_to_bvh_s * 1e3, 1), "device_equals_host": ok,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_assemble",
                     "kernel_ms": round(kern_ms, 5), "algorithmic_bytes_per_launch": int(alg_bytes)},
        "cpu_baseline": cpu,
    }


def ingest_cpu_baseline(tmp: Path, pos, quads, rows: int = 60):
    """The oracle's pure-Python restatement of load_obj on a bounded sample:
    the first `rows` rows of the same grid (OBJ of the same form)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import mesh_oracle as MO
    n = int(round(np.sqrt(len(quads))))
    keep = quads[: rows * n]
    used = np.unique(keep)
    remap = np.zeros(len(pos), np.int64)
    remap[used] = np.arange(len(used))
    p = tmp / "sample.obj"
    write_obj(p, pos[used], remap[keep])
    t = time.perf_counter()
    mod = MO.load([str(p)])
    s = time.perf_counter() - t
    p.unlink()
    return {"value": round(len(mod.prims) / s / 1e6, 5), "unit": "Mtriangles/s", "cores": 1, "kind": "port",
            "sample": f"first {rows} of {n} quad rows ({len(mod.prims)} triangles), same OBJ form; the "
                      "reference's Boost.Spirit loader cannot be built here (no Boost)"}


if __name__ == "__main__":
    main()
