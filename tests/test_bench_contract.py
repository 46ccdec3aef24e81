"""bench.py's contract on the CPU: the one-GPU rule, the algorithmic byte
counts its roofline divides by, and that the profile summaries it reports
traffic and VALU figures from were measured on the current kernel sources
(a summary left behind by a kernel change would otherwise be dropped from
the line silently)."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr or "--gpus 1" in r.stderr


def test_algorithmic_bytes():
    # RK3: stage 1 reads U0 and writes L0, stages 2 and 3 read two arrays and
    # write one; nv FP64 variables per cell
    assert bench.stage_bytes_per_cell(5) == [80, 120, 120]
    assert bench.stage_bytes_per_cell(4) == [64, 96, 96]
    assert bench.algorithmic_bytes_per_cell_step(5) == 320
    assert sum(bench.stage_bytes_per_cell(4)) * 1 == 256


def test_profiles_match_the_kernel_sources():
    cases = [("stage_kernel_kh3d_256_traffic.json", 6, bench.SHAPE_NAMES[3]),
             ("stage_kernel_kh2d_1024_traffic.json", 3, bench.SHAPE_NAMES[2]),
             ("batch_stage_kernel_uq_traffic.json", 3, bench.SHAPE_NAMES[2]),
             ("batch_stage_kernel_uq_traffic.json", 4, bench.SHAPE_NAMES[2])]
    for name, shape, names in cases:
        prof, why = bench.load_profile(name, shape, names)
        assert prof is not None, why
        assert prof["bytes_per_launch"] > 0 and prof["valu_insts_per_launch"] > 0
        v = bench.valu_bound(prof, 1.0, name)
        assert 0 < v["frac"]
    top = json.load(open(os.path.join(ROOT, "profiles", "stage_kernel_kh3d_256_traffic.json")))
    assert len(top["stage_valu_busy_frac"]) == 3
    assert all(0 < b < 1 for b in top["stage_valu_busy_frac"])
    # the kernel trace summary of the same workload names the same digest
    md = open(os.path.join(ROOT, "profiles", "stage_kernel_kh3d_256_kernel_trace.md")).read()
    assert bench.kernel_source_digest() in md
