"""The C++ facade (include/visionaray_amd.hpp) used the way a reference
caller writes it: tests/cpp/facade_demo, built by __graft_entry__.build()."""
import json
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
DEMO = ROOT / "tests" / "cpp" / "facade_demo"


def test_facade_demo():
    assert DEMO.exists(), "tests/cpp/facade_demo is not built (run __graft_entry__.build())"
    r = subprocess.run([str(DEMO)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["lbvh_nodes"] == 3 and out["sah_nodes"] == 1
    want = [[0, 1.0], [1, 1.0], [-1, None], [-1, None]]
    for key in ("lbvh", "sah"):
        for (pid, t), (wp, wt) in zip(out[key], want):
            assert pid == wp
            if wt is not None:
                assert t == wt
    assert [h[0] >= 0 for h in out["any"]] == [True, True, False, False]
    assert out["refit_t1"] == 3.0
    assert [h[0] for h in out["inst"]] == [0, 1, -1, -1]
    assert [h[2] for h in out["inst"]] == [7, 7, -1, -1] and out["inst"][0][1] == 1.0
    assert out["inst_top_nodes"] == 1  # two instances: one leaf
    assert np.isfinite(out["pt_sum"]) and out["pt_sum"] > 0
    assert np.isfinite(out["ao_sum"]) and out["ao_sum"] > 0
    assert out["frames"] == 2
