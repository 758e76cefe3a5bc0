"""Scene generators: sizes of the benchmark configurations and determinism."""
import numpy as np

from visionaray_amd import scenes as S


def test_config0_size_and_layout():
    t = S.config0_scene()
    assert t.size == 24 + 20 * 4 ** 5 == 20504
    assert S.TRI_DTYPE.itemsize == 44
    assert np.array_equal(t["prim_id"], np.arange(t.size))


def test_config1_size_formula():
    # 2 * 1448^2 heightfield + 2000 * 20 * 4^3 icosphere triangles
    assert 2 * 1448 ** 2 + 2000 * 20 * 4 ** 3 == 6_753_408
    t = S.config1_scene(cells=64, spheres=10)
    assert t.size == 2 * 64 ** 2 + 10 * 1280
    assert np.array_equal(t["geom_id"], t["prim_id"])


def test_icosphere_counts_and_radius():
    for level in range(4):
        v, f = S.icosphere_mesh(level)
        assert f.shape[0] == 20 * 4 ** level
        assert np.allclose(np.linalg.norm(v, axis=1), 1.0)


def test_generators_are_deterministic():
    a = S.config0_scene(level=3)
    b = S.config0_scene(level=3)
    assert a.tobytes() == b.tobytes()
    assert S.random_clusters(100, seed=4).tobytes() == S.random_clusters(100, seed=4).tobytes()
    assert S.heightfield(32, seed=1).tobytes() == S.heightfield(32, seed=1).tobytes()


def test_materials_emissive_fraction():
    m = S.config1_materials(10000)
    assert (m["type"] == 1).sum() == 100
    cd = m["cd"][m["type"] == 0]
    assert cd.min() >= 0.2 and cd.max() <= 0.8
