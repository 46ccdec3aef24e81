"""alsuqcli on the hip platform: `python -m alsvinn_amd.alsuqcli config.xml`.

The reference runner (alsuq/src/run/Runner.cpp:39-90) draws the samples'
parameters (STLMersenne + Uniform), runs one simulator per sample and folds
each into the statistics at their save times; here every sample advances in
the same batched launches, with the fvm writer's save times as a time-step
adjuster of every sample (so each sample takes the reference's steps) and
the meanvar statistics on the device.  Statistics files are
`<basename>_mean_<k>.nc` and `<basename>_variance_<k>.nc` (alsuq/src/config/
Setup.cpp:250-259); per-sample solution files are not written.
"""
import argparse
import sys
import time

from .config import ConfigError, grid_shape, initial_state, read_config
from .netcdf import CONSERVED, NetCDFWriter
from .uq import BatchRunner, generate_parameters


def run(path, device=0, out=sys.stdout):
    cfg = read_config(path)
    uq = cfg["uq"]
    if uq is None:
        raise ConfigError("missing <uq>")
    if len(uq["stats"]) > 1:
        raise ConfigError("one <stat> per run is supported")
    st = uq["stats"][0] if uq["stats"] else None
    if st is not None and st["type"] != "netcdf":
        raise ConfigError("writer type %r is not supported (netcdf)" % st["type"])
    M = uq["samples"]
    params = generate_parameters(uq["parameters"], uq["sample_start"], M) if uq["parameters"] else {}
    n = grid_shape(cfg)
    br = BatchRunner(cfg["equation"], cfg["reconstruction"], cfg["flux"], cfg["integrator"], cfg["boundary"],
                     [(n, M)], cfg["end_time"], lower=tuple(cfg["lower"]), upper=tuple(cfg["upper"]),
                     gamma=cfg["gamma"], cfl=cfg["cfl"], stats_saves=st["number_of_saves"] if st else 0,
                     stats_write_initial=st["write_initial"] if st else False, device=device)
    w = cfg["writer"]
    if w is not None and "number_of_saves" in w:
        br.add_fixed_interval_writer(cfg["end_time"] / w["number_of_saves"], 0.0, w["write_initial"])
    t0 = time.perf_counter()
    for s in range(M):
        br.set_state(0, s, initial_state(cfg, {k: list(v[s]) for k, v in params.items()}))
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    steps, left = br.run()
    t_run = time.perf_counter() - t0
    br.check_constraints()
    files = []
    if st is not None:
        names = CONSERVED[cfg["equation"]]
        interval = cfg["end_time"] / st["number_of_saves"]
        attrs = {"fvm_name": cfg["name"], "platform": "hip", "samples": M}
        wm = NetCDFWriter(st["basename"] + "_mean", attrs)
        wv = NetCDFWriter(st["basename"] + "_variance", attrs)
        for k in range(br.nsaves):
            mean, var, cnt = br.statistics(0, k)
            t = (k if st["write_initial"] else k + 1) * interval
            files.append(wm.write({names[v]: mean[v] for v in range(len(names))}, t))
            files.append(wv.write({names[v]: var[v] for v in range(len(names))}, t))
    print("alsuqcli (hip): %s, %d samples, %d batched steps in %.3f s (initial data %.3f s), wrote %d file(s)"
          % (cfg["name"], M, steps, t_run, t_init, len(files)), file=out)
    return br, files


def main(argv=None):
    from .alsvinncli import add_decomposition_flags, check_decomposition
    ap = argparse.ArgumentParser(description="Run an alsuq configuration on the hip platform")
    ap.add_argument("input", help="XML configuration file")
    ap.add_argument("--device", type=int, default=0)
    add_decomposition_flags(ap, sample=True)
    a = ap.parse_args(argv)
    check_decomposition(ap, a, sample=True)
    run(a.input, a.device)
    return 0


if __name__ == "__main__":
    sys.exit(main())
