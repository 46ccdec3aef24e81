"""XML front end: unchanged alsvinncli / alsuqcli configuration files run on
the hip platform.

  read_config(path)             the <fvm> and <uq> sections, with the
                                reference's node names, defaults and "auto"
                                rules (alsfvm/src/config/SimulatorSetup.cpp:
                                102-585, alsuq/src/config/Setup.cpp:53-313)
  python_initial_data(...)      the reference's Python initial data, both
                                forms (alsfvm/src/init/PythonInitialData.cpp:
                                84-281): a snippet run once per cell midpoint
                                with x, y, z, i, j, k and the primitive
                                variables in scope, or an init_global(...)
                                function filling whole arrays
  make_simulator(cfg, params)   a Simulator with the initial state and the
                                fvm writer adjuster

The <platform> node may say cpu, cuda or hip: every configuration runs on
the hip platform (its value is kept in the output attributes).
"""
import os
import textwrap
import xml.etree.ElementTree as ET

import numpy as np

from .initial_data import conserved_from_primitive
from .simulator import Simulator, resolve_cfl, resolve_integrator

PRIMITIVE = {
    "burgers": ["u"],
    "euler1": ["rho", "ux", "p"],
    "euler2": ["rho", "ux", "uy", "p"],
    "euler3": ["rho", "ux", "uy", "uz", "p"],
}


class ConfigError(ValueError):
    pass


def _text(node, path, default=None, required=True):
    n = node.find(path)
    if n is None or n.text is None:
        if required and default is None:
            raise ConfigError("missing <%s>" % path.replace("/", "."))
        return default
    return n.text.strip()


def _vector(s, cast):
    return [cast(v) for v in s.split()]


def _bool(s):
    # boost::property_tree get<bool>: "true"/"1" are true
    return s.strip().lower() in ("1", "true")


def _parameters(node):
    """<parameters><parameter><name/><length/><value/> or <values><value/>...
    (SimulatorSetup.cpp:445-480): a length-1 parameter is a scalar; a single
    listed value fills the whole length."""
    out = {}
    if node is None:
        return out
    for p in node.findall("parameter"):
        name = _text(p, "name")
        length = int(_text(p, "length"))
        if length == 1 and p.find("value") is not None:
            out[name] = [float(_text(p, "value"))]
            continue
        vals = [float(v.text) for v in p.findall("values/value")]
        if not vals and p.find("value") is not None:
            vals = [float(_text(p, "value"))]
        if len(vals) == 1:
            vals = vals * length
        out[name] = vals
    return out


def read_config(path):
    """Parse a reference configuration file into a plain dict."""
    root = ET.parse(path).getroot()
    if root.tag != "config":
        raise ConfigError("the root node must be <config>")
    fvm = root.find("fvm")
    if fvm is None:
        raise ConfigError("missing <fvm>")
    cfg = {"path": os.path.abspath(path), "base": os.path.dirname(os.path.abspath(path))}
    cfg["name"] = _text(fvm, "name", "")
    cfg["platform"] = _text(fvm, "platform", "cpu", required=False)
    g = fvm.find("grid")
    if g is None:
        raise ConfigError("missing <fvm.grid>")
    cfg["lower"] = _vector(_text(g, "lowerCorner"), float)
    cfg["upper"] = _vector(_text(g, "upperCorner"), float)
    cfg["dimension"] = _vector(_text(g, "dimension"), int)
    cfg["boundary"] = _text(fvm, "boundary")
    if cfg["boundary"] not in ("periodic", "neumann"):
        raise ConfigError("Unknown boundary conditions " + cfg["boundary"])
    cfg["flux"] = _text(fvm, "flux")
    cfg["end_time"] = float(_text(fvm, "endTime"))
    cfg["equation"] = _text(fvm, "equation")
    if cfg["equation"] not in PRIMITIVE:
        raise ConfigError("unsupported equation " + cfg["equation"])
    eqp = fvm.find("equationParameters")
    cfg["equation_parameters"] = {}
    if eqp is not None:
        for c in eqp:
            cfg["equation_parameters"][c.tag] = float(c.text)
    cfg["gamma"] = cfg["equation_parameters"].get("gamma", 1.4)
    cfg["reconstruction"] = _text(fvm, "reconstruction")
    cfg["cfl"] = resolve_cfl(_text(fvm, "cfl"), cfg["reconstruction"])
    cfg["integrator"] = resolve_integrator(_text(fvm, "integrator"), cfg["reconstruction"])
    diff = fvm.find("diffusion")
    if diff is not None:
        name = _text(diff, "name", "none", required=False)
        if name != "none":
            raise ConfigError("diffusion %r is not supported (only none)" % name)
    funcs = fvm.find("functionals")
    if funcs is not None and len(funcs.findall("functional")):
        raise ConfigError("<functionals> are not supported")
    init = fvm.find("initialData")
    if init is None:
        raise ConfigError("missing <fvm.initialData>")
    py = _text(init, "python", required=False)
    if py is None:
        raise ConfigError("only <python> initial data is supported")
    cfg["python"] = os.path.join(cfg["base"], py)
    cfg["init_parameters"] = _parameters(init.find("parameters"))
    w = fvm.find("writer")
    cfg["writer"] = None
    if w is not None:
        wr = {"type": _text(w, "type"), "basename": _text(w, "basename")}
        if w.find("numberOfSaves") is not None:
            wr["number_of_saves"] = int(_text(w, "numberOfSaves"))
            wr["write_initial"] = _bool(_text(w, "writeInitialTimestep", "false", required=False))
        if w.find("numberOfCoarseSaves") is not None or w.find("timeRadius") is not None:
            raise ConfigError("coarse-grained and time-integrated writers are not supported")
        cfg["writer"] = wr
    uq = root.find("uq")
    cfg["uq"] = None
    if uq is not None:
        u = {"samples": int(float(_text(uq, "samples"))),
             "sample_start": int(_text(uq, "sampleStart", "0", required=False)),
             "generator": _text(uq, "generator", "auto", required=False)}
        if u["generator"] not in ("auto", "stlmersenne"):
            raise ConfigError("generator %r is not supported (auto, stlmersenne)" % u["generator"])
        params = []
        for p in uq.findall("parameters/parameter"):
            typ = _text(p, "type")
            if typ != "uniform":
                raise ConfigError("distribution %r is not supported (uniform)" % typ)
            params.append((_text(p, "name"), int(_text(p, "length")), float(_text(p, "lower", "0", required=False)),
                           float(_text(p, "upper", "1", required=False))))
        u["parameters"] = params
        stats = []
        for s in uq.findall("stats/stat"):
            st = {"name": _text(s, "name"), "type": _text(s, "writer/type"), "basename": _text(s, "writer/basename")}
            if st["name"] != "meanvar":
                raise ConfigError("statistics %r are not supported (meanvar)" % st["name"])
            if s.find("numberOfSaves") is None:
                raise ConfigError("statistics need <numberOfSaves>")
            st["number_of_saves"] = int(_text(s, "numberOfSaves"))
            st["write_initial"] = _bool(_text(s, "writeInitialTimestep", "false", required=False))
            stats.append(st)
        u["stats"] = stats
        cfg["uq"] = u
    return cfg


def grid_shape(cfg):
    n = list(cfg["dimension"]) + [1] * (3 - len(cfg["dimension"]))
    return tuple(n[:3])


def cell_midpoints(lower, upper, n):
    """Grid::getCellMidpoints: origin + length * i + length / 2 with length
    (top - origin) / N, 1 on single-cell y / z axes (alsfvm/src/grid/
    Grid.cpp:24-41, 148-171)."""
    out = []
    for d in range(3):
        length = (upper[d] - lower[d]) / n[d]
        if d > 0 and n[d] == 1:
            length = 1.0
        out.append(lower[d] + length * np.arange(n[d]) + length / 2.0)
    return out


def python_initial_data(program, equation, n, lower, upper, parameters):
    """Primitive variables {name: (nz, ny, nx)} from the reference's Python
    initial data program (snippet or init_global form)."""
    names = PRIMITIVE[equation]
    nx, ny, nz = n
    ns = {"__name__": "__alsvinn_init__"}
    exec("from math import *\ntry:\n    from numpy import *\nexcept Exception:\n    pass\n", ns)
    for k, v in parameters.items():
        ns[k] = float(v[0]) if len(v) == 1 else np.array(v, dtype=np.float64)
    prim = {v: np.zeros((nz, ny, nx)) for v in names}
    if "init_global" in program:
        exec(program, ns)
        # arrays of shape (nx, ny, nz) with x the fastest index, as the
        # reference hands them over
        ns["init_global"](*[prim[v].T for v in names], nx, ny, nz, lower[0], lower[1], lower[2], upper[0], upper[1],
                          upper[2])
        return prim
    body = "".join("%s = 0.0\n" % v for v in names) + program + "\n" + \
        "".join("output['%s'] = %s\n" % (v, v) for v in names)
    src = "def initial_data(x, y, z, i, j, k, output):\n" + textwrap.indent(body, "    ")
    exec(src, ns)
    f = ns["initial_data"]
    xs, ys, zs = cell_midpoints(lower, upper, n)
    out = {}
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                f(float(xs[i]), float(ys[j]), float(zs[k]), i, j, k, out)
                for v in names:
                    prim[v][k, j, i] = float(out[v])
    return prim


def conserved_state(equation, prim, gamma):
    """Interior conserved state (nv, nz, ny, nx) from primitive variables
    (Euler::computeConserved)."""
    if equation == "burgers":
        return prim["u"][None].copy()
    names = PRIMITIVE[equation]
    u = [prim[v] for v in names[1:-1]]
    return conserved_from_primitive(prim["rho"], u, prim["p"], gamma)


def initial_state(cfg, parameters=None):
    params = dict(cfg["init_parameters"])
    # equation parameters are initial-data parameters too (SimulatorSetup.cpp:430-440)
    for k, v in cfg["equation_parameters"].items():
        params[k] = [v]
    params.update(parameters or {})
    with open(cfg["python"]) as fh:
        program = fh.read()
    prim = python_initial_data(program, cfg["equation"], grid_shape(cfg), cfg["lower"] + [0] * (3 - len(cfg["lower"])),
                               cfg["upper"] + [0] * (3 - len(cfg["upper"])), params)
    return conserved_state(cfg["equation"], prim, cfg["gamma"])


def make_simulator(cfg, parameters=None, device=0):
    n = grid_shape(cfg)
    sim = Simulator(cfg["equation"], cfg["reconstruction"], cfg["flux"], cfg["integrator"], cfg["boundary"], n,
                    lower=tuple(cfg["lower"]), upper=tuple(cfg["upper"]), gamma=cfg["gamma"], cfl=cfg["cfl"],
                    device=device)
    sim.set_interior(initial_state(cfg, parameters))
    return sim
