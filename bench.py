"""Benchmark of the hip platform on one MI355X.

Default workload: 3D compressible Euler (gamma 1.4) Kelvin-Helmholtz on the
unit cube, 256^3 cells, periodic, FP64, HLL3 flux + WENO2 + RK3-TVD, CFL 0.3,
seeded band perturbations over (x, z) -- the case whose working set (5
variables x 3 buffers x 671 MB) is far beyond the 256 MB last-level cache,
so the fused stage kernel's bytes come from HBM.  --workload kh2d runs the
2D 1024^2 case (its 3 x 33.5 MB working set stays cache-resident, flagged in
the output), --workload uq the batched Monte Carlo set.
One step = one full RK3 time step (3 fused stage launches + on-device CFL
dt) over the grid; the metric is FP64 cell-updates/s = cells * steps / wall
time, the state resident in HBM before the timed region.  The tile shape is
autotuned and the step-chunk hipGraph captured before the warmup steps, so
the timed steps only replay.

Prints one JSON line with the roofline of the fused stage kernel (HIP event
timing on the solver's stream) and the CPU baseline (the oracle -- a
reference-structure restatement of the reference CPU path -- timed on a
bounded number of steps of the same workload on the host cores).
"""
import argparse
import hashlib
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
SHAPE_NAMES = {1: ["256", "64"], 2: ["32x16", "16x15", "64x4", "32x15", "24x20"],
               3: ["32x8 columns x 32 planes", "16x15 columns x 32 planes", "32x7 columns x 32 planes",
                   "16x15 columns x 32 planes, 2 barriers", "16x15 columns x 16 planes, 2 barriers",
                   "32x7 columns x 32 planes, 2 barriers", "16x15 columns x 32 planes, 2 barriers, shared reciprocals"]}

WORKLOADS = {
    # the bench line
    "kh2d": dict(equation="euler2", n=(1024, 1024, 1), nv=4, cfl=0.4, t_end=2.0, profile="kh2d_1024", mall_resident=True,
                 desc="2D compressible Euler Kelvin-Helmholtz 1024x1024, periodic, HLL3 + WENO2 + RK3-TVD, "
                      "CFL 0.40, gamma 1.4, FP64"),
    # 3D KH: working set beyond the 256 MB last-level cache (the %HBM case)
    "kh3d": dict(equation="euler3", n=(256, 256, 256), nv=5, cfl=0.3, t_end=1.0, profile="kh3d_256", mall_resident=False,
                 desc="3D compressible Euler Kelvin-Helmholtz 256^3, periodic, HLL3 + WENO2 + RK3-TVD, "
                      "CFL 0.30, gamma 1.4, FP64"),
}

# Sources of the stage kernels: a profile summary under profiles/ records the
# digest of the sources it was measured on, and its traffic / instruction
# figures are reported only while the sources still match.
KERNEL_SOURCES = ["avh_physics.hpp", "avh_stage.hpp", "avh_stage3d.hpp", "avh_stage_inst.hpp", "avh_batch.hpp",
                  "avh_stage_burgers.hip", "avh_stage_euler1.hip", "avh_stage_euler2.hip", "avh_stage_euler3.hip"]


def kernel_source_digest():
    h = hashlib.sha256()
    csrc = os.path.join(os.path.dirname(os.path.abspath(__file__)), "alsvinn_amd", "csrc")
    for name in KERNEL_SOURCES:
        h.update(name.encode())
        with open(os.path.join(csrc, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def load_profile(name, shape, shape_names):
    """The profile summary `name` for tile shape `shape` if it was measured
    on the current kernel sources and that shape, else None (and the
    reason).  A summary holds its main shape's figures at the top level and
    those of further shapes under "other_shapes"."""
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    if not os.path.exists(tfile):
        return None, "no profile summary %s" % name
    prof = json.load(open(tfile))
    if prof.get("kernel_source_digest") != kernel_source_digest():
        return None, "%s was measured on other kernel sources (digest %s, now %s)" % (
            name, prof.get("kernel_source_digest"), kernel_source_digest())
    if prof.get("tile_shape", shape) == shape:
        return prof, None
    other = prof.get("other_shapes", {}).get(str(shape))
    if other:
        return other, None
    return None, "%s was measured on tile shape %s, this run used %s" % (
        name, shape_names[prof["tile_shape"]], shape_names[shape])


def valu_bound(prof, avg_ms, source):
    """The stage kernel's second bound, FP64 VALU issue: the profile's VALU
    wave-instructions per launch over the launch time measured here, against
    1024 SIMDs x one wave64 instruction per 4 cycles at the rated 2.4 GHz."""
    insts = prof.get("valu_insts_per_launch") if prof else None
    if not insts:
        return None
    peak = 1024 * 2.4e9 / 4
    out = {"insts_per_launch": insts, "achieved_per_s": insts / (avg_ms * 1e-3), "peak_per_s": peak,
           "frac": insts / (avg_ms * 1e-3) / peak, "source": source}
    busy = prof.get("stage_valu_busy_frac")
    if busy:
        # from the profile's own counter run: SQ_ACTIVE_INST_VALU (quarter-rate
        # instructions weighted by four) over the SIMD time, per stage launch
        out["pipe_busy_frac_profiled"] = busy
    return out


def host_cpu():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "affinity": allowed}


# Batched Monte Carlo of the 2D KH problem: 1024 samples at 256^2, 256 at
# 512^2 and 64 at 128^2 advanced together, on-device mean/variance at T = 2.
UQ_GROUPS = [((256, 256, 1), 1024), ((512, 512, 1), 256), ((128, 128, 1), 64)]
UQ_DESC = ("batched Monte Carlo 2D compressible Euler Kelvin-Helmholtz: 1024 samples at 256^2 + 256 at 512^2 "
           "+ 64 at 128^2 in the same launches, 40 U[0,1] parameters per sample, periodic, HLL3 + WENO2 + "
           "RK3-TVD, CFL 0.4, gamma 1.4, FP64, on-device Welford mean/variance at T = 2")


def algorithmic_bytes_per_cell_step(nv):
    # per RK3 step: stage 1 reads U0, writes L0 (2 arrays); stage 2 reads L0,
    # U0, writes U2 (3); stage 3 reads U2, U0, writes U3 (3): 8 arrays of nv
    # doubles per cell
    return 8 * nv * 8


def stage_bytes_per_cell(nv):
    return [2 * nv * 8, 3 * nv * 8, 3 * nv * 8]


def initial_state(wl):
    from alsvinn_amd.initial_data import kelvin_helmholtz_2d, kelvin_helmholtz_3d, kh_params
    if wl["equation"] == "euler2":
        return kelvin_helmholtz_2d(wl["n"][0], kh_params(2024))
    return kelvin_helmholtz_3d(wl["n"][0], kh_params(2024, count=80))


def cpu_baseline(wl, seconds_budget=15.0):
    # The host cores this process may use: OMP_NUM_THREADS when the
    # environment sets it (the CPU share allotted beside one GPU), else every
    # core of the affinity mask.
    cpu = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if threads <= 0:
        threads = cpu["affinity"] or 1
        os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import pyoracle

    sim = pyoracle.Simulator(wl["equation"], "weno2", "hll3", "rungekutta3", "periodic", wl["n"], cfl=wl["cfl"])
    sim.set_interior(initial_state(wl))
    sim.step()  # first touch / thread start-up outside the timing
    steps = 0
    t0 = time.perf_counter()
    while True:
        sim.step()
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds_budget or steps >= 20000:
            break
    cells = wl["n"][0] * wl["n"][1] * wl["n"][2]
    return {
        "value": cells * steps / el,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d RK3 steps of the same workload (%s) on the reference-structure CPU restatement "
                  "(OpenMP, -O3 -ffp-contract=off), %.1f s" % (steps, "x".join(map(str, wl["n"])), el),
        "host": dict(cpu, omp_num_threads=threads),
    }


def run_uq_to_end(args, br, cells, total):
    """The whole Monte Carlo run: every sample to T = 2, statistics on the
    device; reports the wall time and the statistics' mean over the domain."""
    import numpy as np
    t0 = time.perf_counter()
    done = 0
    while True:
        d, left = br.run(max_steps=500, chunk=100)
        done += d
        print("bench: %d batched steps, %d samples running, %.1f s" % (done, left, time.perf_counter() - t0),
              file=sys.stderr, flush=True)
        if left == 0:
            break
    wall = time.perf_counter() - t0
    br.check_constraints()
    steps = [[br.progress(g, s).steps for s in range(m)] for g, (n, m) in enumerate(UQ_GROUPS)]
    updates = sum(n[0] * n[1] * sum(st) for (n, m), st in zip(UQ_GROUPS, steps))
    group_updates = [n[0] * n[1] * sum(st) for (n, m), st in zip(UQ_GROUPS, steps)]
    stats = {}
    eps = args.uq_target_error
    for g, (n, m) in enumerate(UQ_GROUPS):
        mean, var, cnt = br.statistics(g, 0, primitive=True)
        # Monte Carlo error of the statistics of rho (L1 over the domain,
        # relative): standard error of the mean sqrt(var / M), and of the
        # variance sqrt(2 / (M - 1)) var (normal approximation)
        l1 = lambda a: float(np.abs(a[0]).mean())
        err_mean = l1(np.sqrt(var / cnt)) / l1(mean)
        err_var = l1(np.sqrt(2.0 / (cnt - 1)) * var) / l1(var)
        # samples are independent and cost the same, so the error falls as
        # 1/sqrt(M) at a wall time proportional to M: the samples and wall
        # time this group needs for relative errors <= eps (its share of the
        # run's wall time by cell updates)
        need = int(np.ceil(cnt * max(err_mean, err_var) ** 2 / eps ** 2))
        share = wall * group_updates[g] / updates
        stats["%dx%d" % n[:2]] = {"samples": int(cnt), "steps_min_max": [min(steps[g]), max(steps[g])],
                                  "domain_mean_of_mean_rho_ux_uy_p": [float(v) for v in mean.mean(axis=(1, 2, 3))],
                                  "domain_mean_of_variance_rho_ux_uy_p": [float(v) for v in var.mean(axis=(1, 2, 3))],
                                  "rho_mean_rel_std_error": err_mean, "rho_variance_rel_std_error": err_var,
                                  "wall_share_s": share,
                                  "samples_for_target_error": need,
                                  "wall_to_target_error_s": share * need / cnt}
    return {
        "metric": "UQ wall time (all samples to T = 2, on-device mean/variance)",
        "value": wall,
        "unit": "s",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": UQ_DESC, "cells": cells, "samples": total, "parallelism": "single GPU"},
        "cell_updates": updates,
        "cell_updates_per_s": updates / wall,
        "target_rel_error": eps,
        "wall_to_target_error_s": sum(v["wall_to_target_error_s"] for v in stats.values()),
        "wall_to_target_error_note": ("per resolution: measured wall time share x (samples needed for relative "
                                      "MC errors of mean and variance of rho <= target) / samples run; "
                                      "samples are independent, so the cost is linear in their number"),
        "statistics": stats,
    }


def run_to_end(sim, wl, tune):
    """The whole run of a single-grid workload: from the initial data to its
    end time (the last step lands on it, as with one writer save there), the
    state resident on the device throughout."""
    T = wl["t_end"]
    sim.add_fixed_interval_writer(T)
    sim.synchronize()
    t0 = time.perf_counter()
    p = sim.run(T, chunk=512)
    wall = time.perf_counter() - t0
    if p.violated:
        raise SystemExit("constraint violation before T")
    cells = wl["n"][0] * wl["n"][1] * wl["n"][2]
    return {
        "metric": "wall time to T (whole run, one GPU)",
        "value": wall,
        "unit": "s",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["desc"], "cells": cells, "end_time": T, "parallelism": "single GPU"},
        "steps": int(p.steps),
        "final_time": sim.get_current_time(),
        "cell_updates_per_s": cells * p.steps / wall,
        "tile_autotune": tune,
    }


def run_uq(args):
    import alsvinn_amd
    from alsvinn_amd import BatchRunner, generate_parameters

    total = sum(m for _, m in UQ_GROUPS)
    print("bench: building %d UQ samples" % total, file=sys.stderr, flush=True)
    params = generate_parameters([("a", 40, 0.0, 1.0)], 0, total)["a"]
    br = BatchRunner("euler2", "weno2", "hll3", "rungekutta3", "periodic", UQ_GROUPS, 2.0, cfl=0.4, stats_saves=1)
    k = 0
    for g, (n, m) in enumerate(UQ_GROUPS):
        br.set_kelvin_helmholtz(g, params[k:k + m])
        k += m
    cells = sum(n[0] * n[1] * m for n, m in UQ_GROUPS)
    if args.uq_to_end:
        return run_uq_to_end(args, br, cells, total)
    if args.tile_shape >= 0:
        shape = args.tile_shape
        br.set_tile_shape(shape)
        tune = {"chosen": SHAPE_NAMES[2][shape], "fixed": True}
    else:
        shape, tune_ms = br.autotune()
        tune = {"chosen": SHAPE_NAMES[2][shape], "ms_per_shape": dict(zip(SHAPE_NAMES[2], tune_ms))}
    br.run(max_steps=args.warmup, chunk=max(1, args.warmup))
    print("bench: warmup done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    done, left = br.run(max_steps=args.steps, chunk=args.steps)
    el = time.perf_counter() - t0
    if left != total:
        raise SystemExit("samples finished or failed inside the timed steps (%d of %d active)" % (left, total))
    value = cells * args.steps / el
    launch_ms = br.time_step_kernels(args.profile_steps)
    stage_ms = launch_ms[:3]
    sb = stage_bytes_per_cell(4)
    avg_ms = sum(stage_ms) / 3
    avg_bytes = cells * sum(sb) / len(sb)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    pname = "batch_stage_kernel_uq_traffic.json"
    prof, stale = load_profile(pname, shape, SHAPE_NAMES[2])
    traffic = prof.get("bytes_per_launch") if prof else None
    valu = valu_bound(prof, avg_ms, pname)
    cpu = None
    if not args.no_cpu_baseline:
        print("bench: timing the CPU baseline", file=sys.stderr, flush=True)
        wl = dict(equation="euler2", n=(256, 256, 1), nv=4, cfl=0.4)
        cpu = cpu_baseline(wl, args.cpu_seconds)
        cpu["sample"] = ("one 256^2 sample of the batch: " + cpu["sample"] +
                         "; samples are independent, so the batch's CPU rate is this per-cell rate")
    return {
        "metric": "FP64 cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": UQ_DESC,
            "cells": cells,
            "samples": total,
            "step": "one RK3 time step of every sample (3 batched stage launches, per-sample on-device CFL dt, "
                    "end-of-step and statistics kernels)",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_note": stale,
            "kernel": "avh::batch_stage_kernel (fused reconstruct + flux + RK stage, all samples)",
            "algorithmic_bytes_per_launch": avg_bytes,
            "stage_ms": stage_ms,
            "end_of_step_kernel_ms": launch_ms[3],
            "statistics_kernel_ms": launch_ms[4] if len(launch_ms) > 4 else None,
            "kernel_source_digest": kernel_source_digest(),
            "tile_autotune": tune,
            "valu": valu,
        },
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default kh3d 1500, kh2d 20000, uq 300: a few seconds of GPU work each)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 20; uq: 5)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["uq"], default="kh3d")
    ap.add_argument("--tile-shape", type=int, default=-1,
                    help="fix the stage kernel's tile shape (index into the autotuner's list) instead of "
                         "autotuning, e.g. to profile one kernel instance")
    ap.add_argument("--uq-to-end", action="store_true",
                    help="with --workload uq: run every sample to T = 2 and report the wall time, the Monte "
                         "Carlo errors of the statistics and the wall time to --uq-target-error")
    ap.add_argument("--to-end", action="store_true",
                    help="with --workload kh2d / kh3d: run the problem to its end time (T = 2 / 1) and report the "
                         "wall time, the step count and the average rate")
    ap.add_argument("--uq-target-error", type=float, default=0.01,
                    help="relative Monte Carlo error of the mean and variance of rho (L1 over the domain)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (torch copy kernel), to calibrate "
                         "the FETCH_SIZE / WRITE_SIZE counters against a known byte count")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8,
                          "dispatches": 3}), flush=True)
        return

    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path to run; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.steps is None:
        args.steps = {"uq": 300, "kh3d": 1500, "kh2d": 20000}[args.workload]
    if args.warmup is None:
        args.warmup = 5 if args.workload == "uq" else 20
    if args.workload == "uq":
        print(json.dumps(run_uq(args)), flush=True)
        return

    import alsvinn_amd

    wl = WORKLOADS[args.workload]
    NV = wl["nv"]
    print("bench: building %s state" % args.workload, file=sys.stderr, flush=True)
    sim = alsvinn_amd.Simulator(wl["equation"], "weno2", "hll3", "rungekutta3", "periodic", wl["n"], cfl=wl["cfl"])
    sim.set_interior(initial_state(wl))
    dim = 3 if wl["n"][2] > 1 else (2 if wl["n"][1] > 1 else 1)
    if args.tile_shape >= 0:
        shape = args.tile_shape
        sim.set_tile_shape(shape)
        tune = {"chosen": SHAPE_NAMES[dim][shape], "fixed": True}
    else:
        shape, tune_ms = sim.autotune()
        tune = {"chosen": SHAPE_NAMES[dim][shape], "ms_per_shape": dict(zip(SHAPE_NAMES[dim], tune_ms))}
    sim.prepare_graph()  # capture + instantiate now: the timed steps only replay
    if args.to_end:
        print(json.dumps(run_to_end(sim, wl, tune)), flush=True)
        return

    sim.enqueue_steps(args.warmup)
    sim.synchronize()
    print("bench: warmup done", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    sim.enqueue_steps(args.steps)
    sim.synchronize()
    el = time.perf_counter() - t0
    prog = sim.progress()
    if prog.violated:
        raise SystemExit("constraint violation during the timed steps")
    cells = wl["n"][0] * wl["n"][1] * wl["n"][2]
    value = cells * args.steps / el
    ms_per_step = el * 1e3 / args.steps

    # per-launch timing of the fused stage kernel (HIP events on the solver's stream)
    import ctypes
    L = alsvinn_amd.load()
    ms = (ctypes.c_double * 4)()
    nk = ctypes.c_int()
    alsvinn_amd._abi.check(L.avh_time_step_kernels(sim._h, args.profile_steps, ms, ctypes.byref(nk)))
    launch_ms = [ms[i] for i in range(nk.value)]
    stage_ms = launch_ms[:3]          # the three fused stage launches
    # RK3: the last stage's last workgroup ends the step (no end-of-step kernel)
    finish_ms = launch_ms[3] if len(launch_ms) > 3 else None
    sb = stage_bytes_per_cell(NV)
    avg_ms = sum(stage_ms) / len(stage_ms)
    avg_bytes = cells * sum(sb) / len(sb)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9

    traffic = None
    valu = None
    pname = "stage_kernel_%s_traffic.json" % WORKLOADS[args.workload]["profile"]
    prof, stale = load_profile(pname, shape, SHAPE_NAMES[dim])
    if prof:
        traffic = prof.get("bytes_per_launch")
        valu = valu_bound(prof, avg_ms, pname)

    cpu = None
    if not args.no_cpu_baseline:
        print("bench: timing the CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(wl, args.cpu_seconds)

    out = {
        "metric": "FP64 cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": wl["desc"],
            "cells": cells,
            "step": "one RK3 time step (3 fused stage launches, on-device CFL dt)",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_note": stale,
            "mall_resident": wl["mall_resident"],
            "kernel": ("avh::march_kernel3 (fused reconstruct + flux + RK stage, z-marching columns)"
                       if dim == 3 else "avh::stage_kernel (fused reconstruct + flux + RK stage)"),
            "kernel_source_digest": kernel_source_digest(),
            "algorithmic_bytes_per_launch": avg_bytes,
            "stage_ms": stage_ms,
            "end_of_step_kernel_ms": finish_ms,
            "tile_autotune": tune,
            "valu": valu,
        },
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
