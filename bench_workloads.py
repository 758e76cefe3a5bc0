"""Measurement procedures of bench.py for the batched, differentiable and
dynamic configurations (C2, C3, C4: --config c2 | c3 | c4; the default run
also reports C2 and C3 in the headline line).  Each returns one JSON-able
dict; times come from HIP events on the stream the kernels run on."""
from __future__ import annotations

import time

import numpy as np


def _events(V):
    return V.Event(), V.Event()


# ---------------------------------------------------------------------------
# c2: 64 scenes, one persistent launch (graph replay) vs one launch per scene
# ---------------------------------------------------------------------------

def bench_batched(args, V, S, log, steps=None):
    """C2: 64 scenes of 1k-2M triangles in one persistent hipGraph-captured
    launch, against per-scene launches in the reference's scheme and on the
    persistent scheduler; `steps` timed frames (default --steps)."""
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
    steps = max(1, args.steps if steps is None else steps)

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

def bench_adjoint(args, V, S, log, quick=False):
    """C3: 1.31 M triangles, 256 materials (792 parameters), 512 x 512, 64 spp,
    depth 4, L2 loss to a perturbed-albedo target.  The speedup compares one
    adjoint (forward + loss + backward, HIP events) with central differences of
    all 792 parameters at h = 1e-3 (2 x 792 forward frames + losses), timed on
    `fd_timed` parameters and extrapolated.  Agreement is checked against the
    plain frame's central differences with common random numbers on a
    parameter subset; see the returned keys for the spp of each check.
    quick: the subset sizes of the default bench run."""
    V.init(0)
    st = V.Stream()
    width = height = 512
    spp, depth, hstep = 64, 4, 1e-3
    check_spp = 256
    tris, mats = S.config3_scene()
    rng = np.random.default_rng(11)
    target_m = mats.copy()
    target_m["cd"] = np.clip(mats["cd"] + rng.normal(0, 0.05, mats["cd"].shape), 0.01, 0.99).astype(np.float32)
    bvh = V.lbvh_builder().build(tris, keep_nodes=False, stream=st)
    eye, center, up, fovy, _ = S.config0_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, 1.0)
    kp = V.make_kernel_params(num_bounces=depth, epsilon=1e-4)
    lib = V.lib()
    params = [(m, "cd", c) for m in range(mats.size) for c in range(3)] + \
             [(m, "ce", c) for m in np.nonzero(mats["type"] == 1)[0] for c in range(3)]
    d_mats = V.DeviceArray.from_host(mats)
    d_pert = V.DeviceArray(mats.size, mats.dtype)
    d_rr = V.DeviceArray.from_host(mats)  # roulette table held at the unperturbed point
    kern = V.pathtracing.kernel(bvh, d_mats, kp)
    kpert = V.pathtracing.kernel(bvh, d_pert, kp)

    class Frame:
        """render target, target image and loss buffers at one spp"""

        def __init__(self, n_spp):
            self.rt = V.render_target(width, height)
            self.sp = V.make_sched_params(V.jittered_blend_type(spp=n_spp), cam, self.rt)
            self.sched = V.hip_sched()
            self.sched.frame(V.pathtracing.kernel(bvh, V.DeviceArray.from_host(target_m), kp), self.sp, stream=st)
            st.synchronize()
            self.target = V.DeviceArray.from_host(self.rt.color.to_host())
            self.dl = V.DeviceArray(width * height * 4, np.float32)
            self.loss = V.DeviceArray(1, np.float64, zero=True)
            self.grad = V.DeviceArray(6 * mats.size, np.float64, zero=True)

        def forward_loss(self, kernel, frozen=False):
            self.sched.frame_id = 0
            if frozen:
                self.sched.frame_frozen_rr(kernel, d_rr, self.sp, stream=st)
            else:
                self.sched.frame(kernel, self.sp, stream=st)
            V.check(lib.vr_memset(self.loss.ptr, 0, 8))
            V.check(lib.vr_image_l2_loss(self.rt.color.ptr, self.target.ptr, width * height, self.dl.ptr,
                                         self.loss.ptr, st.handle))

        def adjoint(self):
            V.check(lib.vr_memset(self.grad.ptr, 0, self.grad.nbytes))
            self.forward_loss(kern)
            self.sched.backward(kern, self.sp, self.dl, self.grad, stream=st, frame_id=0)

        def gradient(self):
            self.adjoint()
            st.synchronize()
            g = self.grad.to_host().reshape(-1, 6)
            return np.concatenate([g[:, 0:3].ravel(), g[mats["type"] == 1, 3:6].ravel()])

        def central(self, i, frozen=False):
            m, key, c = params[i]
            vals = []
            for sgn in (1.0, -1.0):
                mm = mats.copy()
                mm[key][m, c] += np.float32(sgn * hstep)
                d_pert.upload(mm)
                self.forward_loss(kpert, frozen)
                st.synchronize()
                vals.append(float(self.loss.to_host()[0]))
            return (vals[0] - vals[1]) / (2 * hstep)

    def cosine(x, y):
        return float(x @ y / (np.linalg.norm(x) * np.linalg.norm(y) + 1e-300))

    F = Frame(spp)
    # ---- timed: the adjoint, then central differences (enqueued back to back)
    for _ in range(2):
        F.adjoint()
    st.synchronize()
    e0, e_mid, e1 = V.Event(), V.Event(), V.Event()
    e0.record(st)
    V.check(lib.vr_memset(F.grad.ptr, 0, F.grad.nbytes))
    F.forward_loss(kern)
    e_mid.record(st)
    F.sched.backward(kern, F.sp, F.dl, F.grad, stream=st, frame_id=0)
    e1.record(st)
    adjoint_ms = e0.elapsed_ms(e1)
    forward_ms = e0.elapsed_ms(e_mid)
    L0 = float(F.loss.to_host()[0])
    ga = F.gradient()
    nfd = len(params) if not args.fd_params else min(args.fd_params, len(params))
    if quick and not args.fd_params:
        nfd = 24
    timed = np.linspace(0, len(params) - 1, nfd).astype(int)
    uploads = []
    for i in timed:  # the perturbed tables, uploaded between frames as the FD loop does
        for sgn in (1.0, -1.0):
            m, key, c = params[i]
            mm = mats.copy()
            mm[key][m, c] += np.float32(sgn * hstep)
            uploads.append(mm)
    e0, e1 = V.Event(), V.Event()
    e0.record(st)
    for mm in uploads:
        d_pert.upload(mm)
        F.forward_loss(kpert)
    e1.record(st)
    st.synchronize()
    fd_ms = e0.elapsed_ms(e1)
    fd_ms_all = fd_ms * len(params) / nfd
    speedup = fd_ms_all / adjoint_ms
    log(f"c3: adjoint {adjoint_ms:.2f} ms (forward + loss {forward_ms:.2f} ms); FD of {nfd} params "
        f"{fd_ms:.0f} ms -> all {len(params)}: {fd_ms_all:.0f} ms; speedup x{speedup:.1f}")

    # ---- agreement: every 16th albedo parameter and every emission parameter
    sub = np.array(list(range(0, 768, 16 if quick else 4)) + list(range(768, len(params))))
    fd_plain = np.array([F.central(int(i)) for i in sub])
    fd_frozen = np.array([F.central(int(i), frozen=True) for i in sub])
    cos_plain = cosine(ga[sub], fd_plain)
    cos_frozen = cosine(ga[sub], fd_frozen)
    log(f"c3: {sub.size} params at {spp} spp: cosine {cos_plain:.5f} vs plain CRN FD, {cos_frozen:.6f} vs "
        "frozen-roulette FD")
    G = Frame(check_spp)
    ga_check = G.gradient()
    fd_check = np.array([G.central(int(i)) for i in sub])
    cos_check = cosine(ga_check[sub], fd_check)
    log(f"c3: {sub.size} params at {check_spp} spp: cosine {cos_check:.5f} vs plain CRN FD")
    return {
        "metric": "adjoint vs FD speedup",
        "value": round(speedup, 2), "unit": "x", "n_gpus": 1, "steps": 1, "warmup": 2,
        "ms_per_step": round(adjoint_ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 gradient reduction)", "data": "synthetic",
        "config": {"workload": f"C3 differentiable: room + level-8 icosphere, {tris.size} triangles, "
                               f"{mats.size} materials ({int((mats['type'] == 1).sum())} emissive), "
                               f"{width}x{height}, {spp} spp, depth {depth}, L2 loss to perturbed-albedo target",
                   "params": len(params), "h": hstep},
        "adjoint_ms": round(adjoint_ms, 4), "forward_loss_ms": round(forward_ms, 4),
        "fd_params_timed": int(nfd), "fd_ms_timed": round(fd_ms, 2), "fd_ms_all_params": round(fd_ms_all, 2),
        "cosine_vs_plain_fd": {
            "value": round(cos_check, 6), "spp": check_spp, "params": int(sub.size), "h": hstep,
            "fd": "central differences of the plain frame's L2 loss, common random numbers (frame 0 seeds)",
            "adjoint": f"the same loss at {check_spp} spp, frame 0",
        },
        "cosine_at_config_spp": {
            "spp": spp, "params": int(sub.size),
            "vs_plain_fd": round(cos_plain, 6),
            "vs_frozen_roulette_fd": round(cos_frozen, 6),
            "note": "the L2 loss holds each pixel's variance, which Russian roulette makes depend on the albedos "
                    "through the roulette probabilities; the adjoint holds those probabilities constant (it is "
                    "the exact derivative of the frozen-roulette estimator), the plain differences include that "
                    "term, which falls as 1/spp",
        },
        "loss": L0,
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
