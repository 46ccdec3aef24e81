"""The HIP path reproduces the committed golden trajectories bit for bit."""
import glob
import json
import os

import numpy as np
import pytest

import alsvinn_amd

pytestmark = pytest.mark.gpu
# trajectory fixtures (the generator's vectors live in stl_*.npz)
GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith("stl_"))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_hip_reproduces_golden(path):
    z = np.load(path, allow_pickle=False)
    m = json.loads(str(z["meta"]))
    print("golden", os.path.basename(path), flush=True)
    h = alsvinn_amd.Simulator(m["equation"], m["reconstruction"], m["flux"], m["integrator"], m["boundary"],
                              tuple(m["n"]), cfl=m["cfl"], gamma=m["gamma"])
    h.set_interior(z["initial"])
    dts = [h.perform_step() for _ in range(m["steps"])]
    np.testing.assert_array_equal(np.array(dts), z["dts"])
    np.testing.assert_array_equal(h.get_interior(), z["final"])
    assert h.get_current_time() == float(z["time"])
