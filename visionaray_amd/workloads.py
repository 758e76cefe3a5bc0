"""Measurement procedures for the batched, differentiable and dynamic
configurations (bench.py --config c2 | c3 | c4).  Each returns one JSON-able
dict; times come from HIP events on the stream the kernels run on."""
from __future__ import annotations

import time

import numpy as np


def _events(V):
    return V.Event(), V.Event()


# ---------------------------------------------------------------------------
# c2: 64 scenes, one persistent launch (graph replay) vs one launch per scene
# ---------------------------------------------------------------------------

def bench_batched(args, V, S, log):
    count = args.scenes or 64
    V.init(0)
    st = V.Stream()
    t0 = time.time()
    scenes = []
    total_tris = 0
    for sc in S.config2_scenes(count=count):
        w, h = sc["size"]
        eye, center, up, fovy, aspect = sc["camera"]
        bvh = V.lbvh_builder().build(sc["tris"], keep_nodes=False, stream=st)
        d_mats = V.DeviceArray.from_host(sc["mats"])
        kp = V.make_kernel_params(num_bounces=4, epsilon=1e-4, ambient_color=(0.6, 0.7, 0.9, 1.0))
        cam = V.pinhole_camera(eye, center, up, fovy, aspect)
        rt = V.render_target(w, h)
        sp = V.make_sched_params(V.jittered_blend_type(spp=4), cam, rt)
        scenes.append((V.pathtracing.kernel(bvh, d_mats, kp), sp))
        total_tris += sc["tris"].size
    st.synchronize()
    log(f"c2: {count} scenes, {total_tris} triangles, prepared in {time.time() - t0:.1f}s")
    steps = max(1, args.steps)

    def per_scene(shape):
        """one frame() launch per scene under a launch shape; (ms, rays) per step"""
        V.set_launch_shape(**shape)
        sched = V.hip_sched()
        for k, sp in scenes:
            sched.frame(k, sp, stream=st)
        st.synchronize()
        sched.reset_counters()
        e0, e1 = _events(V)
        e0.record(st)
        for _ in range(steps):
            for k, sp in scenes:
                sched.frame(k, sp, stream=st)
        e1.record(st)
        ms = e0.elapsed_ms(e1) / steps
        p0, s0 = sched.rays()
        return ms, (p0 + s0) / steps

    # plain launches, the reference's GPU scheme: one thread per pixel in
    # 16 x 16 pixel blocks with the binary-tree traversal, one launch per
    # scene (cuda_sched::frame per scene)
    plain_ms, plain_rays = per_scene(dict(block_size=256, stack_lds=16, tile=16, schedule=V.SCHEDULE_PER_PIXEL))
    # the same launches on the persistent scheduler
    pers_ms, pers_rays = per_scene(dict(block_size=256, stack_lds=16, tile=1))
    # autotune the persistent shape on the scene with the most pixels x triangles
    big = max(range(len(scenes)),
              key=lambda i: scenes[i][1].rt.width * scenes[i][1].rt.height * scenes[i][0].bvh.info()["num_prims"])
    shape, tuned_ms = V.autotune(scenes[big][0], scenes[big][1], reps=2, stream=st)
    log(f"c2: autotuned shape {shape} ({tuned_ms:.2f} ms on scene {big})")
    # one persistent launch for all scenes with that shape, captured in a graph
    b = V.batch(scenes)
    b.render(st)
    b.capture_graph(st)
    b.launch_graph(st)
    st.synchronize()
    b.reset_counters()
    e0, e1 = _events(V)
    e0.record(st)
    for _ in range(steps):
        b.launch_graph(st)
    e1.record(st)
    batch_ms = e0.elapsed_ms(e1) / steps
    p1, s1 = b.rays()
    batch_rays = (p1 + s1) / steps
    V.set_launch_shape(256, 16, 1)
    plain_rate = plain_rays / (plain_ms * 1e-3) / 1e6
    pers_rate = pers_rays / (pers_ms * 1e-3) / 1e6
    batch_rate = batch_rays / (batch_ms * 1e-3) / 1e6
    log(f"c2: per-scene plain launches {plain_ms:.2f} ms ({plain_rate:.0f} Mrays/s), per-scene persistent "
        f"{pers_ms:.2f} ms ({pers_rate:.0f} Mrays/s), batched graph {batch_ms:.2f} ms ({batch_rate:.0f} Mrays/s): "
        f"x{batch_rate / plain_rate:.2f} vs plain, x{batch_rate / pers_rate:.2f} vs persistent per scene")
    return {
        "metric": "batched aggregate Mrays/s vs per-scene plain launches",
        "value": round(batch_rate, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": round(batch_ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"C2 batched: {count} scenes, {total_tris} triangles (1k-2M log-uniform), "
                               "resolutions from {256^2, 512^2, 1024^2, 1920x1080}, 4 spp, depth 4",
                   "scenes": count, "triangles": int(total_tris), "batch_shape": shape},
        "per_scene_plain_launches": {"ms": round(plain_ms, 4), "mrays_per_s": round(plain_rate, 3),
                                     "schedule": "one thread per pixel, 16x16 pixel blocks, binary-tree "
                                                 "traversal (the reference's cuda_sched scheme)"},
        "per_scene_persistent_launches": {"ms": round(pers_ms, 4), "mrays_per_s": round(pers_rate, 3)},
        "batched_graph": {"ms": round(batch_ms, 4), "mrays_per_s": round(batch_rate, 3)},
        "speedup": round(batch_rate / plain_rate, 3),
        "speedup_vs_persistent_per_scene": round(batch_rate / pers_rate, 3),
        "rays_per_frame": int(batch_rays),
    }


# ---------------------------------------------------------------------------
# c3: adjoint gradient vs central finite differences
# ---------------------------------------------------------------------------

def bench_adjoint(args, V, S, log):
    import sys
    from pathlib import Path
    V.init(0)
    st = V.Stream()
    width = height = 512
    spp, depth = 64, 4
    tris, mats = S.config3_scene()
    rng = np.random.default_rng(11)
    target_m = mats.copy()
    target_m["cd"] = np.clip(mats["cd"] + rng.normal(0, 0.05, mats["cd"].shape), 0.01, 0.99).astype(np.float32)
    bvh = V.lbvh_builder().build(tris, keep_nodes=False, stream=st)
    eye, center, up, fovy, _ = S.config0_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, 1.0)
    kp = V.make_kernel_params(num_bounces=depth, epsilon=1e-4)
    rt = V.render_target(width, height)
    sp = V.make_sched_params(V.jittered_blend_type(spp=spp), cam, rt)
    d_mats = V.DeviceArray.from_host(mats)
    kern = V.pathtracing.kernel(bvh, d_mats, kp)
    # target image
    sched = V.hip_sched()
    sched.frame(V.pathtracing.kernel(bvh, V.DeviceArray.from_host(target_m), kp), sp, stream=st)
    target = V.DeviceArray(width * height * 4, np.float32)
    st.synchronize()
    target.upload(rt.color.to_host())
    dl = V.DeviceArray(width * height * 4, np.float32)
    grad = V.DeviceArray(6 * mats.size, np.float64, zero=True)
    loss = V.DeviceArray(1, np.float64, zero=True)
    lib = V.lib()

    d_rr = V.DeviceArray.from_host(mats)  # the roulette table held at the unperturbed point

    def forward_loss(kernel, frozen=False):
        sched.frame_id = 0
        if frozen:
            sched.frame_frozen_rr(kernel, d_rr, sp, stream=st)
        else:
            sched.frame(kernel, sp, stream=st)
        V.check(lib.vr_memset(loss.ptr, 0, 8))
        V.check(lib.vr_image_l2_loss(rt.color.ptr, target.ptr, width * height, dl.ptr, loss.ptr, st.handle))

    # adjoint: forward + loss + backward
    for _ in range(2):
        V.check(lib.vr_memset(grad.ptr, 0, grad.nbytes))
        forward_loss(kern)
        sched.backward(kern, sp, dl, grad, stream=st, frame_id=0)
    st.synchronize()
    V.check(lib.vr_memset(grad.ptr, 0, grad.nbytes))
    e0, e1 = _events(V)
    e0.record(st)
    forward_loss(kern)
    e_mid = V.Event()
    e_mid.record(st)
    sched.backward(kern, sp, dl, grad, stream=st, frame_id=0)
    e1.record(st)
    adjoint_ms = e0.elapsed_ms(e1)
    forward_ms = e0.elapsed_ms(e_mid)
    L0 = float(loss.to_host()[0])
    g = grad.to_host().reshape(-1, 6)
    ga = np.concatenate([g[:, 0:3].ravel(), g[mats["type"] == 1, 3:6].ravel()])
    log(f"c3: loss {L0:.6f}; adjoint (forward + loss + backward) {adjoint_ms:.2f} ms, forward {forward_ms:.2f} ms")

    # central differences with common random numbers, h = 1e-3.  The check
    # differentiates the frozen-roulette estimator (vr_render_pt_frozen_rr:
    # roulette probabilities from the throughput at the unperturbed albedos,
    # so every roulette decision is common to both sides); its expectation is
    # the image's for any albedos, and its derivative is what the adjoint
    # computes.  A subset is also differenced with the plain frame, whose
    # roulette decisions flip under the perturbation (noise ~ 1/sqrt(h)).
    hstep = 1e-3
    params = [(m, "cd", c) for m in range(mats.size) for c in range(3)] + \
             [(m, "ce", c) for m in np.nonzero(mats["type"] == 1)[0] for c in range(3)]
    nfd = len(params) if not args.fd_params else min(args.fd_params, len(params))
    d_pert = V.DeviceArray(mats.size, mats.dtype)
    kpert = V.pathtracing.kernel(bvh, d_pert, kp)

    def central(i, frozen):
        m, key, c = params[i]
        vals = []
        for sgn in (1.0, -1.0):
            mm = mats.copy()
            mm[key][m, c] += np.float32(sgn * hstep)
            d_pert.upload(mm)
            forward_loss(kpert, frozen)
            vals.append(float(loss.to_host()[0]))
        return (vals[0] - vals[1]) / (2 * hstep)

    fd = np.zeros(len(params))
    t0 = time.perf_counter()
    e0, e1 = _events(V)
    e0.record(st)
    for i in range(nfd):
        fd[i] = central(i, True)
        if i % 100 == 0:
            log(f"c3: finite differences {i}/{nfd}")
    e1.record(st)
    fd_ms = e0.elapsed_ms(e1)
    fd_wall = time.perf_counter() - t0
    fd_ms_all = fd_ms * len(params) / nfd

    def cosine(x, y):
        return float(x @ y / (np.linalg.norm(x) * np.linalg.norm(y) + 1e-300))

    cos = cosine(ga[:nfd], fd[:nfd])
    speedup = fd_ms_all / adjoint_ms
    log(f"c3: FD {nfd} params in {fd_ms:.0f} ms (wall {fd_wall:.1f}s) -> all {fd_ms_all:.0f} ms; "
        f"speedup x{speedup:.1f}, cosine {cos:.6f} (frozen-roulette FD)")
    sub = np.arange(0, nfd, max(1, nfd // 96))
    fd_plain = np.array([central(int(i), False) for i in sub])
    cos_plain = cosine(ga[sub], fd_plain)
    cos_sub = cosine(ga[sub], fd[sub])
    log(f"c3: on {sub.size} params: cosine {cos_plain:.4f} vs FD of the plain frame (roulette decisions flip), "
        f"{cos_sub:.6f} vs frozen-roulette FD")

    # FP64 CPU adjoint on a band of rows (the image gradient masked to it)
    fp64 = None
    try:
        sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
        import oracle as O
        rows = (240, 248)
        img = rt.color.to_host().reshape(height, width, 4)
        forward_loss(kern)
        st.synchronize()
        img = rt.color.to_host().reshape(height, width, 4)
        dlh = dl.to_host().reshape(height, width, 4).copy()
        mask = np.zeros_like(dlh)
        mask[rows[0]:rows[1]] = dlh[rows[0]:rows[1]]
        d_mask = V.DeviceArray.from_host(mask.reshape(-1))
        gm = V.DeviceArray(6 * mats.size, np.float64, zero=True)
        sched.backward(kern, sp, d_mask, gm, stream=st, frame_id=0)
        st.synchronize()
        gpu_band = gm.to_host().reshape(-1, 6)
        nodes_ref = None
        t1 = time.time()
        nodes_ref, idx, _, _ = O.lbvh_build(tris)
        p = O.render_params(width, height, spp=spp, num_bounces=depth, epsilon=1e-4)
        gcd, gce = O.render_pt_adjoint(nodes_ref, tris, idx, mats, cam.astype(O.CAMERA_DTYPE), p, mask,
                                       rows=rows)
        want = np.concatenate([gcd, gce], axis=1)
        rel = float(np.linalg.norm(gpu_band - want) / max(np.linalg.norm(want), 1e-300))
        cos64 = float((gpu_band.ravel() @ want.ravel()) / (np.linalg.norm(gpu_band) * np.linalg.norm(want) + 1e-300))
        fp64 = {"rows": list(rows), "rel_err": rel, "cosine": cos64, "cpu_s": round(time.time() - t1, 2)}
        log(f"c3: FP64 CPU adjoint on rows {rows}: rel err {rel:.2e}, cosine {cos64:.6f}")
    except Exception as e:  # the oracle is optional for this leg
        fp64 = {"error": str(e)}

    return {
        "metric": "adjoint vs FD speedup",
        "value": round(speedup, 2), "unit": "x", "n_gpus": 1, "steps": 1, "warmup": 2,
        "ms_per_step": round(adjoint_ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 gradient reduction)", "data": "synthetic",
        "config": {"workload": f"C3 differentiable: room + level-8 icosphere, {tris.size} triangles, "
                               f"{mats.size} materials ({int((mats['type'] == 1).sum())} emissive), "
                               f"{width}x{height}, {spp} spp, depth {depth}, L2 loss to perturbed-albedo target",
                   "params": len(params), "fd_params_timed": nfd, "h": hstep},
        "adjoint_ms": round(adjoint_ms, 4), "forward_ms": round(forward_ms, 4),
        "fd_ms_all_params": round(fd_ms_all, 2), "cosine_vs_fd": round(cos, 6),
        "fd_estimator": "central, h=1e-3, common random numbers and common roulette decisions "
                        "(vr_render_pt_frozen_rr at the unperturbed albedos)",
        "plain_frame_fd_subset": {"params": int(sub.size), "cosine": round(cos_plain, 6),
                                  "cosine_frozen_same_subset": round(cos_sub, 6)},
        "loss": L0,
        "fp64_cpu_adjoint": fp64,
    }


# ---------------------------------------------------------------------------
# c4: dynamic geometry, refit vs rebuild per frame, batched forward + adjoint
# ---------------------------------------------------------------------------

def bench_dynamic(args, V, S, log):
    count = args.scenes or 32
    frames = max(2, args.steps)
    V.init(0)
    st = V.Stream()
    t0 = time.time()
    items = []
    total = 0
    for i, sc in enumerate(S.config4_scenes(count=count)):
        w, h = sc["size"]
        eye, center, up, fovy, aspect = sc["camera"]
        base = sc["tris"]
        moved = [V.DeviceArray.from_host(S.displace(base, f, seed=i)) for f in range(frames)]
        bvh = V.lbvh_builder().build(moved[0], keep_nodes=False, stream=st)
        d_mats = V.DeviceArray.from_host(sc["mats"])
        kp = V.make_kernel_params(num_bounces=3, epsilon=1e-4, ambient_color=(0.5, 0.6, 0.8, 1.0))
        cam = V.pinhole_camera(eye, center, up, fovy, aspect)
        rt = V.render_target(w, h)
        sp = V.make_sched_params(V.jittered_blend_type(spp=32), cam, rt)
        dl = V.DeviceArray.from_host(np.full(w * h * 4, 0.01, np.float32))
        items.append(dict(bvh=bvh, moved=moved, kern=V.pathtracing.kernel(bvh, d_mats, kp), sp=sp, dl=dl))
        total += base.size
    st.synchronize()
    log(f"c4: {count} scenes, {total} triangles, {frames} frames prepared in {time.time() - t0:.1f}s")
    b = V.batch([(it["kern"], it["sp"], it["dl"]) for it in items])
    grad = V.DeviceArray(6 * b.num_materials, np.float64, zero=True)

    def timed(fn):
        e0, e1 = _events(V)
        e0.record(st)
        fn()
        e1.record(st)
        return e0.elapsed_ms(e1)

    # rebuild and refit of every scene (LBVH in place of the same handle)
    rebuild_ms, refit_ms, fwd_ms, bwd_ms = [], [], [], []
    b.render(st)
    b.backward(grad, st)
    st.synchronize()
    b.reset_counters()
    for f in range(frames):
        refit_ms.append(timed(lambda: [it["bvh"].refit(it["moved"][f], stream=st) for it in items]))
        fwd_ms.append(timed(lambda: b.render(st)))
        bwd_ms.append(timed(lambda: b.backward(grad, st)))
    prim, sec = b.rays()
    # rebuild cost for comparison: a new LBVH of the same moved geometry into
    # each scene's handle (vr_lbvh_rebuild, no allocation)
    for f in range(frames):
        rebuild_ms.append(timed(lambda: [it["bvh"].rebuild(it["moved"][f], stream=st) for it in items]))
    frame_ms = np.mean(refit_ms) + np.mean(fwd_ms) + np.mean(bwd_ms)
    rays_s = (prim + sec) / (frames * frame_ms * 1e-3) / 1e6
    grads_s = count / (frame_ms * 1e-3)
    log(f"c4: per frame refit {np.mean(refit_ms):.2f} ms vs rebuild {np.mean(rebuild_ms):.2f} ms; forward "
        f"{np.mean(fwd_ms):.2f} ms, backward {np.mean(bwd_ms):.2f} ms; {rays_s:.0f} Mrays/s, {grads_s:.1f} "
        f"scene gradients/s")
    return {
        "metric": "dynamic batched grads/s and Mrays/s",
        "value": round(grads_s, 3), "unit": "scene gradients/s", "n_gpus": 1, "steps": frames, "warmup": 1,
        "ms_per_step": round(float(frame_ms), 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"C4 dynamic: {count} scenes, {total} triangles (50k-1M), per-frame sinusoidal "
                               "displacement 0.01, 256x256, 32 spp, depth 3, 64 materials per scene",
                   "scenes": count, "triangles": int(total)},
        "refit_ms": round(float(np.mean(refit_ms)), 4), "rebuild_ms": round(float(np.mean(rebuild_ms)), 4),
        "forward_ms": round(float(np.mean(fwd_ms)), 4), "backward_ms": round(float(np.mean(bwd_ms)), 4),
        "mrays_per_s": round(float(rays_s), 3),
        "param_grads_per_s": round(float(grads_s * 6 * 64), 1),
    }
