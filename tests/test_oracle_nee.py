"""The oracle's next-event estimation (path tracing with an area-light list,
detail/pathtracing.inl:116-196) against the reference's own pathtracing
kernel compiled from its headers (tests/golden/nee.npz, made by
tests/golden/make_golden.py nee): the same image, pixel for pixel."""
import numpy as np

from conftest import load_golden


def oracle_nee_image(O, g, lights):
    p = g["params"]
    w, h, spp = int(p[0]), int(p[1]), int(p[2])
    cam = np.zeros(1, O.CAMERA_DTYPE)
    c = g["cam_basis"]
    cam["eye"], cam["U"], cam["V"], cam["W"] = c[0:3], c[3:6], c[6:9], c[9:12]
    nodes, idx, _, _ = O.lbvh_build(g["tris"])
    rp = O.render_params(w, h, spp=spp, num_bounces=int(p[4]), epsilon=float(p[5]), bg=tuple(p[6:9]),
                         ambient=tuple(p[9:12]))
    img, counts = O.render_pt(nodes, g["tris"], idx, g["mats"], cam, rp, lights=lights)
    return img, counts


def test_oracle_nee_matches_reference_image(O):
    g = load_golden("nee")
    img, counts = oracle_nee_image(O, g, g["lights"])
    assert np.array_equal(img, g["pt"])
    # shadow rays are counted with the secondary queries
    assert counts[1] > counts[0]


def test_oracle_nee_planted_mistakes_are_caught(O):
    g = load_golden("nee")
    # a light list in another order picks other lights for the same draws
    img, _ = oracle_nee_image(O, g, g["lights"][::-1].copy())
    assert not np.array_equal(img, g["pt"])
    # a dimmer light changes the image
    lights = g["lights"].copy()
    lights["kl"] *= np.float32(0.5)
    img, _ = oracle_nee_image(O, g, lights)
    assert not np.array_equal(img, g["pt"])
    # no light list: the plain path tracer, a different estimator
    img, _ = oracle_nee_image(O, g, None)
    assert not np.array_equal(img, g["pt"])
