"""Writes the golden trajectories under tests/golden/ from the CPU oracle.

Each fixture holds the initial interior state, the configuration, the dt of
every step and the final interior state.  They pin the oracle against
accidental change (tests/test_golden.py) and the HIP path against the oracle
without running it (tests/test_gpu_golden.py).  Regenerate only when the
oracle is deliberately changed:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402
from tests.cases import kh2d, smooth_state  # noqa: E402

CASES = [
    # name, equation, recon, flux, integrator, bc, n, cfl, steps, initial
    ("burgers_sine_64", "burgers", "weno2", "godunov", "rungekutta3", "periodic", (64, 1, 1), 0.45, 20, "sine"),
    ("euler2_kh_32", "euler2", "weno2", "hll3", "rungekutta3", "periodic", (32, 32, 1), 0.4, 5, "kh"),
    ("euler3_smooth_16", "euler3", "weno2", "hll3", "rungekutta3", "periodic", (16, 16, 16), 0.3, 3, "smooth"),
    ("euler1_eno3_neumann_48", "euler1", "eno3", "hll", "rungekutta2", "neumann", (48, 1, 1), 0.4, 6, "smooth"),
]


def initial(kind, equation, n):
    if kind == "sine":
        x = (np.arange(n[0]) + 0.5) / n[0]
        return np.sin(2 * np.pi * x).reshape(1, 1, 1, n[0])
    if kind == "kh":
        return kh2d(n[0], seed=11)
    return smooth_state(equation, n, seed=17, jump=True)


def make(case):
    name, eq, recon, flux, integ, bc, n, cfl, steps, kind = case
    q0 = initial(kind, eq, n)
    s = pyoracle.Simulator(eq, recon, flux, integ, bc, n, cfl=cfl)
    s.set_interior(q0)
    dts = [s.step() for _ in range(steps)]
    meta = dict(equation=eq, reconstruction=recon, flux=flux, integrator=integ, boundary=bc, n=list(n), cfl=cfl,
                steps=steps, gamma=1.4)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), initial=q0, final=s.get_interior(), dts=np.array(dts),
                        time=np.array(s.time), meta=np.array(json.dumps(meta)))


if __name__ == "__main__":
    for c in CASES:
        make(c)
        print("wrote", c[0])
