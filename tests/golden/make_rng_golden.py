"""Writes tests/golden/stl_mt19937_64_uniform.npz from the C++ standard
library (through the CPU oracle, which calls std::mt19937_64 and
std::uniform_real_distribution<double> directly): the first 1000 raw
outputs of a default-seeded std::mt19937_64 and the first 1000
uniform_real_distribution<double>(0, 1) draws.  Regenerate with
python tests/golden/make_rng_golden.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import pyoracle  # noqa: E402

if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "stl_mt19937_64_uniform.npz"),
                        raw=pyoracle.mt19937_64(5489, 0, 1000), uniform=pyoracle.uniform01(5489, 0, 1000))
    print("wrote stl_mt19937_64_uniform.npz")
