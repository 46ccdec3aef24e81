"""alsvinncli on the hip platform: `python -m alsvinn_amd.alsvinncli config.xml`.

The reference driver (alsvinncli/src/main.cpp:193-230): set the initial
value, call the writers once at the start time, then perform steps until
the end time, the writers saving at their fixed intervals.  Output goes to
`<basename>_<k>.nc` in the working directory, as the reference writes it.
"""
import argparse
import sys
import time

from .config import ConfigError, make_simulator, read_config
from .netcdf import CONSERVED, NetCDFWriter


def run(path, device=0, out=sys.stdout):
    cfg = read_config(path)
    sim = make_simulator(cfg, device=device)
    writer = None
    w = cfg["writer"]
    if w is not None:
        if w["type"] != "netcdf":
            raise ConfigError("writer type %r is not supported (netcdf)" % w["type"])
        writer = NetCDFWriter(w["basename"], {"fvm_name": cfg["name"], "platform": "hip",
                                              "configured_platform": cfg["platform"]})
        if "number_of_saves" in w:
            sim.add_fixed_interval_writer(cfg["end_time"] / w["number_of_saves"], 0.0, w["write_initial"])
    names = CONSERVED[cfg["equation"]]
    t0 = time.perf_counter()
    files = []
    steps = 0
    while True:
        p = sim.run(cfg["end_time"], pause_on_save=writer is not None)
        steps = p.steps
        if writer is not None and p.saves_pending:
            q = sim.get_interior()
            for _ in range(p.saves_pending):
                files.append(writer.write({names[v]: q[v] for v in range(len(names))}, p.time))
            sim.ack_saves()
        if p.time >= cfg["end_time"]:
            break
    print("alsvinncli (hip): %s, %d timesteps to t = %.17g in %.3f s, wrote %d file(s)"
          % (cfg["name"], steps, p.time, time.perf_counter() - t0, len(files)), file=out)
    return sim, files


def add_decomposition_flags(ap, sample=False):
    """The reference's MPI decomposition options (alsvinncli/src/main.cpp:55-67,
    alsuqcli/src/main.cpp:40-51), accepted so that existing command lines run
    unchanged.  The hip platform runs on one GPU: each count must be 1."""
    for axis in "xyz":
        ap.add_argument("--multi-" + axis, type=int, default=1,
                        help="processors in the %s direction (one GPU: must be 1)" % axis)
    if not sample:
        for axis in "xyz":
            ap.add_argument("-" + axis, "--automatic-" + axis, action="store_true",
                            help="divide all cores in the %s direction (one GPU: no effect)" % axis)
    else:
        ap.add_argument("--multi-sample", type=int, default=1,
                        help="processors in the sample direction (one GPU: must be 1)")


def check_decomposition(ap, a, sample=False):
    names = ["multi_x", "multi_y", "multi_z"] + (["multi_sample"] if sample else [])
    for name in names:
        if getattr(a, name) != 1:
            ap.error("--%s %d: the hip platform runs on one GPU, every decomposition count must be 1"
                     % (name.replace("_", "-"), getattr(a, name)))


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run an alsvinn configuration on the hip platform")
    ap.add_argument("input", help="XML configuration file")
    ap.add_argument("--device", type=int, default=0)
    add_decomposition_flags(ap)
    a = ap.parse_args(argv)
    check_decomposition(ap, a)
    run(a.input, a.device)
    return 0


if __name__ == "__main__":
    sys.exit(main())
