"""Host-side mirror of the reference's alsuq surface for the hip platform.

  generate_parameters(...)   SampleGenerator + STLMersenne + Uniform
                             (alsuq/src/samples/SampleGenerator.cpp:29-56,
                              alsuq/src/generator/STLMersenne.cpp:25-85)
  MeanVariance               device meanvar statistics fed one solver state at a
                             time (alsuq/src/stats/MeanVariance.cpp:30-67)
  BatchRunner                Runner (alsuq/src/run/Runner.cpp:39-90) with all
                             samples of up to 4 resolution groups advanced in
                             the same launches, meanvar statistics on device
                             (alsuq/src/stats/MeanVariance.cpp:30-67)

Everything runs through libalsvinn_hip.so.
"""
import ctypes

import numpy as np

from . import _abi
from .initial_data import kh_interfaces
from .simulator import (BOUNDARIES, EQUATIONS, FLUXES, INTEGRATORS, NVARS, RECONSTRUCTIONS, _lookup,
                        default_context, resolve_cfl, resolve_integrator)


def generate_parameters(parameters, sample_start, nsamples):
    """parameters: list of (name, length, lower, upper) for uniform
    parameters.  The reference iterates its generator map, i.e. parameter
    names in sorted order, drawing from one shared default-seeded
    std::mt19937_64.  Returns {name: array (nsamples, length)}."""
    ps = sorted(parameters, key=lambda p: p[0])
    dims = (ctypes.c_int * len(ps))(*[int(p[1]) for p in ps])
    lo = (ctypes.c_double * len(ps))(*[float(p[2]) for p in ps])
    hi = (ctypes.c_double * len(ps))(*[float(p[3]) for p in ps])
    width = sum(int(p[1]) for p in ps)
    out = np.empty((nsamples, width))
    _abi.check(_abi.load().avh_uq_generate(len(ps), dims, lo, hi, int(sample_start), int(nsamples),
                                           out.ctypes.data))
    res, o = {}, 0
    for p in ps:
        res[p[0]] = out[:, o:o + int(p[1])].copy()
        o += int(p[1])
    return res


class MeanVariance:
    """Welford mean / variance of solver states on the device, one save
    index per statistics time (conserved variables and rho, u, p)."""

    def __init__(self, equation, n, gamma=1.4, nsaves=1, device=0):
        self.equation = equation
        self.nv = NVARS[equation]
        self.n = tuple(int(v) for v in n) + (1,) * (3 - len(n))
        self.ctx = default_context(device)
        dims = (ctypes.c_int * 3)(*self.n)
        h = ctypes.c_void_p()
        _abi.check(_abi.load().avh_stats_create(self.ctx._h, _lookup(EQUATIONS, equation, "equation"), dims,
                                                float(gamma), int(nsaves), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _abi._lib is not None:
            _abi._lib.avh_stats_destroy(h)
            self._h = None

    def reset(self):
        _abi.check(_abi.load().avh_stats_reset(self._h))

    def accumulate(self, simulator, save=0):
        _abi.check(_abi.load().avh_stats_accumulate(self._h, simulator._h, int(save)))

    def read(self, save=0, primitive=False):
        shape = (self.nv, self.n[2], self.n[1], self.n[0])
        mean = np.empty(shape)
        var = np.empty(shape)
        cnt = ctypes.c_int64()
        _abi.check(_abi.load().avh_stats_read(self._h, int(save), int(bool(primitive)), mean.ctypes.data,
                                              var.ctypes.data, ctypes.byref(cnt)))
        return mean, var, cnt.value


class BatchRunner:
    """groups: list of (n, nsamples) with n = (nx, ny, nz)."""

    def __init__(self, equation, reconstruction, flux, integrator, boundary, groups, end_time,
                 lower=(0.0, 0.0, 0.0), upper=(1.0, 1.0, 1.0), gamma=1.4, cfl=0.475,
                 stats_saves=1, stats_write_initial=False, device=0):
        integrator = resolve_integrator(integrator, reconstruction)
        cfg = _abi.AvhConfig()
        cfg.equation = _lookup(EQUATIONS, equation, "equation")
        cfg.reconstruction = _lookup(RECONSTRUCTIONS, reconstruction, "reconstruction")
        cfg.flux = _lookup(FLUXES, flux, "flux")
        cfg.integrator = _lookup(INTEGRATORS, integrator, "integrator")
        cfg.boundary = _lookup(BOUNDARIES, boundary, "boundary")
        for d in range(3):
            cfg.n[d] = 1
            cfg.lower[d] = float(lower[d]) if d < len(lower) else 0.0
            cfg.upper[d] = float(upper[d]) if d < len(upper) else 1.0
        cfg.gamma = float(gamma)
        cfg.cfl = resolve_cfl(cfl, reconstruction)
        self.cfg = cfg
        self.equation = equation
        self.nv = NVARS[equation]
        self.end_time = float(end_time)
        self.groups = [(tuple(int(v) for v in n) + (1,) * (3 - len(n)), int(m)) for n, m in groups]
        garr = (_abi.AvhGroup * len(self.groups))()
        for i, (n, m) in enumerate(self.groups):
            for d in range(3):
                garr[i].n[d] = n[d]
            garr[i].nsamples = m
        self.nsaves = int(stats_saves) + (1 if (stats_saves and stats_write_initial) else 0)
        self.ctx = default_context(device)
        h = ctypes.c_void_p()
        _abi.check(_abi.load().avh_batch_create(self.ctx._h, ctypes.byref(cfg), len(self.groups), garr,
                                                self.end_time, int(stats_saves), int(bool(stats_write_initial)),
                                                ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _abi._lib is not None:
            _abi._lib.avh_batch_destroy(h)
            self._h = None

    def interior_shape(self, group):
        n = self.groups[group][0]
        return (self.nv, n[2], n[1], n[0])

    def add_fixed_interval_writer(self, interval, start_time=0.0, write_initial_timestep=False):
        _abi.check(_abi.load().avh_batch_add_fixed_interval_adjuster(
            self._h, float(interval), float(start_time), int(bool(write_initial_timestep))))

    def set_state(self, group, sample, interior):
        a = np.ascontiguousarray(interior, dtype=np.float64).reshape(self.interior_shape(group))
        _abi.check(_abi.load().avh_batch_set_state(self._h, int(group), int(sample), a.ctypes.data))

    def get_state(self, group, sample):
        out = np.empty(self.interior_shape(group))
        _abi.check(_abi.load().avh_batch_get_state(self._h, int(group), int(sample), out.ctypes.data))
        return out

    def set_kelvin_helmholtz(self, group, params):
        """params: (nsamples, 40) perturbation parameters per sample."""
        n = self.groups[group][0]
        params = np.asarray(params, dtype=np.float64)
        iface = np.empty((params.shape[0], 2, n[0]))
        for s in range(params.shape[0]):
            lo, hi = kh_interfaces(n[0], params[s])
            iface[s, 0] = lo
            iface[s, 1] = hi
        iface = np.ascontiguousarray(iface)
        _abi.check(_abi.load().avh_batch_set_kh2d(self._h, int(group), iface.ctypes.data))

    def run(self, max_steps=1 << 62, chunk=64):
        done = ctypes.c_int64()
        left = ctypes.c_int64()
        _abi.check(_abi.load().avh_batch_advance(self._h, int(max_steps), int(chunk), ctypes.byref(done),
                                                 ctypes.byref(left)))
        return done.value, left.value

    def violations(self, group):
        """(number of samples that failed the constraint check, first one or -1)."""
        n = ctypes.c_int64()
        first = ctypes.c_int64()
        _abi.check(_abi.load().avh_batch_violations(self._h, int(group), ctypes.byref(n), ctypes.byref(first)))
        return n.value, first.value

    def check_constraints(self):
        """Raise ConstraintError, as the reference's simulator does, if any
        sample failed the constraint check."""
        for g in range(len(self.groups)):
            n, first = self.violations(g)
            if n:
                p = self.progress(g, first)
                raise _abi.ConstraintError(
                    "Simulation state does not obey constraints!  Group %d sample %d (and %d more) at time %.17g, "
                    "number of timesteps performed %d" % (g, first, n - 1, p.time, p.steps))

    def set_end_time(self, t_end):
        """Every sample runs to t_end (again active if t < t_end)."""
        self.end_time = float(t_end)
        _abi.check(_abi.load().avh_batch_set_end_time(self._h, float(t_end)))

    def kinetic_energy(self, group, component=1):
        """Per sample: sum of 0.5 rho u_c^2 dV at the current state."""
        out = np.empty(self.groups[group][1])
        _abi.check(_abi.load().avh_batch_kinetic_energy(self._h, int(group), int(component), out.ctypes.data))
        return out

    def structure_cube(self, group, var=0, number_of_h=63, p=2.0):
        """Per sample structure function S_p(h), h = 0..number_of_h-1 (S_p(0) = 0)."""
        out = np.empty((self.groups[group][1], int(number_of_h)))
        _abi.check(_abi.load().avh_batch_structure_cube(self._h, int(group), int(var), int(number_of_h), float(p),
                                                        out.ctypes.data))
        return out

    def set_tile_shape(self, shape):
        """Tile shape of the stage kernel (-1: autotune on first use)."""
        _abi.check(_abi.load().avh_batch_set_tile_shape(self._h, int(shape)))

    def autotune(self):
        """Time every tile shape now; returns (best shape, [ms per shape])."""
        best = ctypes.c_int()
        n = ctypes.c_int()
        ms = (ctypes.c_double * 8)()
        _abi.check(_abi.load().avh_batch_autotune(self._h, ctypes.byref(best), ms, ctypes.byref(n)))
        return best.value, [ms[i] for i in range(n.value)]

    def set_graph_steps(self, steps):
        """Steps per captured hipGraph (0: ordinary launches)."""
        _abi.check(_abi.load().avh_batch_set_graph_steps(self._h, int(steps)))

    def time_step_kernels(self, steps):
        """Advance `steps` steps with an event after every launch; returns the
        mean milliseconds of each launch of a step, in launch order."""
        ms = (ctypes.c_double * 8)()
        nk = ctypes.c_int()
        _abi.check(_abi.load().avh_batch_time_step_kernels(self._h, int(steps), ms, ctypes.byref(nk)))
        return [ms[i] for i in range(nk.value)]

    def progress(self, group, sample):
        p = _abi.AvhProgress()
        _abi.check(_abi.load().avh_batch_get_progress(self._h, int(group), int(sample), ctypes.byref(p)))
        return p

    def statistics(self, group, save=0, primitive=False):
        """(mean, variance, count) of statistics save `save`."""
        shape = self.interior_shape(group)
        mean = np.empty(shape)
        var = np.empty(shape)
        cnt = ctypes.c_int64()
        _abi.check(_abi.load().avh_batch_get_statistics(self._h, int(group), int(save), int(bool(primitive)),
                                                        mean.ctypes.data, var.ctypes.data, ctypes.byref(cnt)))
        return mean, var, cnt.value
