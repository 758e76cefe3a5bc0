"""Batched scenes in one persistent launch: every scene's image must equal its
own single-scene frame bit for bit (same seeds, same kernel code); graph
replays must equal eager frames; the batched adjoint equals per-scene adjoints."""
import numpy as np
import pytest

from visionaray_amd import scenes as S

pytestmark = pytest.mark.gpu


def make_scenes(V, count=6, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(200, 20000))
        tris = S.random_clusters(n, seed=100 + i, clusters=4, sigma=0.3)
        nm = 8
        tris["geom_id"] = rng.integers(0, nm, size=n)
        mats = S.lambert_materials(nm, seed=i, emissive_fraction=0.25)
        w, h = [(32, 32), (48, 40), (64, 16), (16, 64)][i % 4]
        v = tris["v1"]
        c = v.mean(0)
        eye = c + np.array([0.0, 0.5, 4.0])
        cam = V.pinhole_camera(eye, c, (0, 1, 0), 0.8, w / h)
        kp = V.make_kernel_params(num_bounces=3 + i % 3, epsilon=1e-4, ambient_color=(0.5, 0.6, 0.8, 1.0))
        bvh = V.lbvh_builder().build(tris)
        d_mats = V.DeviceArray.from_host(mats)
        out.append(dict(bvh=bvh, mats=mats, d_mats=d_mats, kp=kp, cam=cam, w=w, h=h, spp=1 + i % 3))
    return out


def single_frame(V, sc, frame_id):
    rt = V.render_target(sc["w"], sc["h"])
    sched = V.hip_sched()
    sched.frame_id = frame_id
    sched.frame(V.pathtracing.kernel(sc["bvh"], sc["d_mats"], sc["kp"]),
                V.make_sched_params(V.jittered_blend_type(spp=sc["spp"]), sc["cam"], rt))
    return rt.color_host(), sched.rays()


def batch_entries(V, scs):
    ents, rts = [], []
    for sc in scs:
        rt = V.render_target(sc["w"], sc["h"])
        rts.append(rt)
        ents.append((V.pathtracing.kernel(sc["bvh"], sc["d_mats"], sc["kp"]),
                     V.make_sched_params(V.jittered_blend_type(spp=sc["spp"]), sc["cam"], rt)))
    return ents, rts


def test_batch_equals_single_scene_frames(V):
    scs = make_scenes(V)
    ents, rts = batch_entries(V, scs)
    b = V.batch(ents)
    b.render()
    V.synchronize()
    tot = [0, 0]
    for sc, rt in zip(scs, rts):
        want, (p, s) = single_frame(V, sc, 0)
        assert np.array_equal(rt.color_host(), want)
        tot[0] += p
        tot[1] += s
    assert b.rays() == tuple(tot)


def test_batch_graph_replays(V):
    """Each graph replay renders the next frame; replays interleaved with eager
    work must match eager single-scene frames 0, 1, 2."""
    scs = make_scenes(V, count=4, seed=1)
    ents, rts = batch_entries(V, scs)
    b = V.batch(ents)
    st = V.Stream()
    b.capture_graph(st)
    for frame in range(3):
        b.launch_graph(st)
        st.synchronize()
        for sc, rt in zip(scs, rts):
            want, _ = single_frame(V, sc, frame)  # eager work between replays
            assert np.array_equal(rt.color_host(), want), f"replay {frame}"


def test_batch_adjoint_equals_per_scene(V):
    scs = make_scenes(V, count=3, seed=2)
    ents, rts = batch_entries(V, scs)
    dls = []
    for sc in scs:
        rng = np.random.default_rng(sc["w"])
        dl = rng.normal(size=(sc["h"], sc["w"], 4)).astype(np.float32)
        dl[..., 3] = 0
        dls.append(V.DeviceArray.from_host(dl.reshape(-1)))
    b = V.batch([(k, sp, dl) for (k, sp), dl in zip(ents, dls)])
    b.render()
    grad = V.DeviceArray(6 * b.num_materials, np.float64, zero=True)
    b.backward(grad)
    got = grad.to_host().reshape(-1, 6)
    base = 0
    for sc, (k, sp), dl in zip(scs, ents, dls):
        g1 = V.DeviceArray(6 * sc["mats"].size, np.float64, zero=True)
        sched = V.hip_sched()
        sched.frame(k, sp)
        sched.backward(k, sp, dl, g1)
        want = g1.to_host().reshape(-1, 6)
        part = got[base: base + sc["mats"].size]
        assert np.allclose(part, want, rtol=1e-4, atol=1e-6 * max(1.0, np.abs(want).max()))
        base += sc["mats"].size


def _stats(got, want, sigma):
    exact = np.mean(np.all(got == want, axis=-1))
    diff = np.abs(got[..., :3] - want[..., :3])
    within = np.mean(np.all(diff <= 4.0 * sigma[..., :3] + 1e-6, axis=-1))
    rel = np.sqrt(np.mean(diff ** 2)) / max(np.sqrt(np.mean(want[..., :3] ** 2)), 1e-12)
    return exact, within, rel


def test_batch_matches_oracle(V, O):
    """Batched frames and the batched adjoint against the oracle, not only
    against the single-scene path: three clustered scenes of uneven size in
    one launch -- one above 65,536 triangles, so every bottom-up pass of its
    LBVH build runs -- and a dynamic scene whose BVH was refitted after a
    per-frame displacement (config 4's refit path).  Frames: per pixel
    bit-identical to the oracle except where device sin/cos differ in the last
    bit (>= 80 % exact, >= 99 % within 4 sigma of the Monte Carlo estimate,
    sigma from later frames of the same batch); adjoint: against the FP64
    oracle adjoint of the same frame."""
    specs = [  # triangles, width, height, spp, bounces
        (90_000, 48, 40, 2, 4),
        (3_000, 32, 32, 3, 3),
        (20_000, 40, 24, 2, 5),
        (60_000, 32, 32, 2, 3),   # refitted after displacement
    ]
    scs = []
    for i, (n, w, h, spp, nb) in enumerate(specs):
        rng = np.random.default_rng(40 + i)
        tris = S.random_clusters(n, seed=400 + i, clusters=6, sigma=0.3)
        nm = 8
        tris["geom_id"] = rng.integers(0, nm, size=n)
        mats = S.lambert_materials(nm, seed=40 + i, emissive_fraction=0.25)
        c = np.median(tris["v1"], axis=0)
        cam = V.pinhole_camera(c + np.array([0.3, 0.6, 3.5]), c, (0, 1, 0), 0.9, w / h)
        kp = V.make_kernel_params(num_bounces=nb, epsilon=1e-4, ambient_color=(0.5, 0.6, 0.8, 1.0))
        bvh = V.lbvh_builder().build(tris)
        nodes, idx, _, _ = O.lbvh_build(tris)
        if i == 3:
            moved = S.displace(tris, frame=1, amplitude=0.01, seed=3)
            d_moved = V.DeviceArray.from_host(moved)
            bvh.refit(d_moved)
            nodes = O.refit(nodes, moved, idx)
            tris = moved
        scs.append(dict(tris=tris, nodes=nodes, idx=idx, mats=mats, d_mats=V.DeviceArray.from_host(mats), kp=kp,
                        cam=cam, w=w, h=h, spp=spp, nb=nb, bvh=bvh))
    ents, rts = batch_entries(V, scs)
    dls = [V.DeviceArray(sc["w"] * sc["h"] * 4, np.float32, zero=True) for sc in scs]
    b = V.batch([(k, sp, dl) for (k, sp), dl in zip(ents, dls)])
    b.render()
    imgs = [rt.color_host() for rt in rts]
    # d loss / d image of an L2 loss against a dimmer target, uploaded into the
    # batch's own d loss buffers before the backward pass of frame 0
    host_dl = []
    for img, dl in zip(imgs, dls):
        g = np.zeros_like(img)
        g[..., :3] = 2.0 * (img[..., :3] - 0.5 * img[..., :3])
        dl.upload(g.reshape(-1))
        host_dl.append(g)
    grad = V.DeviceArray(6 * b.num_materials, np.float64, zero=True)
    b.backward(grad)
    got_grad = grad.to_host().reshape(-1, 6)
    frames = [[] for _ in scs]
    for _ in range(8):  # frames 1..8 for the per-pixel sigma
        b.render()
        for k, rt in enumerate(rts):
            frames[k].append(rt.color_host())
    base = 0
    for k, sc in enumerate(scs):
        cam = sc["cam"].astype(O.CAMERA_DTYPE)
        p = O.render_params(sc["w"], sc["h"], spp=sc["spp"], num_bounces=sc["nb"], epsilon=1e-4,
                            ambient=sc["kp"]["ambient"][0])
        want, _ = O.render_pt(sc["nodes"], sc["tris"], sc["idx"], sc["mats"], cam, p)
        exact, within, rel = _stats(imgs[k], want, np.stack(frames[k]).std(axis=0))
        print(f"[batch vs oracle] scene {k} ({sc['tris'].size} tris): exact {exact:.4f}, "
              f"within 4 sigma {within:.4f}, rel RMSE {rel:.2e}")
        assert exact >= 0.8 and within >= 0.99 and rel < 0.05
        gcd, gce = O.render_pt_adjoint(sc["nodes"], sc["tris"], sc["idx"], sc["mats"], cam, p, host_dl[k])
        wantg = np.concatenate([gcd, gce], axis=1)
        part = got_grad[base: base + sc["mats"].size]
        err = np.linalg.norm(part - wantg) / np.linalg.norm(wantg)
        cos = float(part.ravel() @ wantg.ravel() / np.linalg.norm(part) / np.linalg.norm(wantg))
        print(f"[batch adjoint vs FP64 oracle] scene {k}: rel err {err:.2e}, cosine {cos:.6f}")
        assert err < 1e-4 and cos > 0.99999
        base += sc["mats"].size


def test_batch_backward_rejects_missing_inputs(V):
    """vr_batch_backward needs every scene's d loss / d image and at most 16
    bounces: a batch without them is an error, and no kernel runs."""
    scs = make_scenes(V, count=2, seed=3)
    ents, _ = batch_entries(V, scs)
    grad = V.DeviceArray(6 * 16, np.float64, zero=True)
    b = V.batch(ents)  # no d loss / d image
    with pytest.raises(V.VRError) as e:
        b.backward(grad)
    assert e.value.code == 1 and "d_dl_dimage" in str(e.value)
    dl = [V.DeviceArray(sc["w"] * sc["h"] * 4, np.float32, zero=True) for sc in scs]
    deep = V.make_kernel_params(num_bounces=17, epsilon=1e-4)
    k0, sp0 = ents[0]
    b2 = V.batch([(V.pathtracing.kernel(k0.bvh, k0.materials, deep), sp0, dl[0]), (*ents[1], dl[1])])
    with pytest.raises(V.VRError):
        b2.backward(grad)
    assert np.all(grad.to_host() == 0)
