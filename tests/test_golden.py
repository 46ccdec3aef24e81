"""The oracle reproduces the committed golden trajectories bit for bit
(tests/golden/make_golden.py wrote them)."""
import glob
import json
import os

import numpy as np
import pytest

# trajectory fixtures (the generator's vectors live in stl_*.npz)
GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith("stl_"))


def load(path):
    z = np.load(path, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return z, meta


def test_fixtures_present():
    assert len(GOLDEN) >= 4


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_reproduces_golden(oracle, path):
    z, m = load(path)
    s = oracle.Simulator(m["equation"], m["reconstruction"], m["flux"], m["integrator"], m["boundary"],
                         tuple(m["n"]), cfl=m["cfl"], gamma=m["gamma"])
    s.set_interior(z["initial"])
    dts = [s.step() for _ in range(m["steps"])]
    np.testing.assert_array_equal(np.array(dts), z["dts"])
    np.testing.assert_array_equal(s.get_interior(), z["final"])
    assert s.time == float(z["time"])
