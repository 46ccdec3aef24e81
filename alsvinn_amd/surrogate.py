"""Learned surrogate for Monte Carlo UQ of the Kelvin-Helmholtz problem.

  EnsembleMLP            the ensemble network of libalsvinn_hip.so (bf16
                         matrix-core training and inference, fp32 Adam)
  kh_observables         runs samples through the batched solver and reads
                         the observables: the kinetic energy of u_y at given
                         times and the structure function S_2(h) of rho at the
                         end time (structure_cube, p = 2)
  surrogate_statistics   mean / variance of the observables from surrogate
                         predictions, with solver fallback where the ensemble
                         disagrees, and the control-variate estimator that
                         corrects the surrogate with solver samples

The reference has no learned model; the statistics it replaces are those of
alsuq's Monte Carlo runner (alsuq/src/run/Runner.cpp:39-90,
alsuq/src/stats/StructureCube.cpp:37-76).  Everything runs through
libalsvinn_hip.so; there is no CPU path.
"""
import ctypes
import time

import numpy as np

from . import _abi
from .simulator import default_context
from .uq import BatchRunner, generate_parameters


class EnsembleMLP:
    """n_in -> width (GELU) x hidden -> n_out, `members` independently
    initialised and shuffled copies trained together."""

    def __init__(self, n_in, n_out, width=256, hidden=4, members=5, batch=256, seed=0, beta1=0.9, beta2=0.999,
                 eps=1e-8, device=0):
        cfg = _abi.AvhMlpConfig()
        cfg.n_in, cfg.width, cfg.hidden, cfg.n_out = int(n_in), int(width), int(hidden), int(n_out)
        cfg.members, cfg.batch = int(members), int(batch)
        cfg.beta1, cfg.beta2, cfg.eps = float(beta1), float(beta2), float(eps)
        cfg.seed = int(seed)
        self.cfg = cfg
        self.n_in, self.n_out, self.members, self.batch = cfg.n_in, cfg.n_out, cfg.members, cfg.batch
        self.ctx = default_context(device)
        h = ctypes.c_void_p()
        _abi.check(_abi.load().avh_mlp_create(self.ctx._h, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int64()
        _abi.check(_abi.load().avh_mlp_num_params(self._h, ctypes.byref(n)))
        self.num_params = n.value
        self.norm = (np.zeros(self.n_in), np.ones(self.n_in), np.zeros(self.n_out), np.ones(self.n_out))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _abi._lib is not None:
            _abi._lib.avh_mlp_destroy(h)
            self._h = None

    @staticmethod
    def _f32(a, shape=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if shape is not None:
            a = a.reshape(shape)
        return a

    def set_normalization(self, in_mean, in_std, out_mean, out_std):
        arrs = [self._f32(in_mean, (self.n_in,)), self._f32(in_std, (self.n_in,)),
                self._f32(out_mean, (self.n_out,)), self._f32(out_std, (self.n_out,))]
        _abi.check(_abi.load().avh_mlp_set_normalization(self._h, *[a.ctypes.data for a in arrs]))
        self.norm = tuple(a.astype(np.float64) for a in arrs)

    def fit_normalization(self, X, Y, floor=1e-12):
        X = np.asarray(X, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        self.set_normalization(X.mean(0), np.maximum(X.std(0), floor), Y.mean(0), np.maximum(Y.std(0), floor))

    def get_params(self, member):
        out = np.empty(self.num_params, dtype=np.float32)
        _abi.check(_abi.load().avh_mlp_get_params(self._h, int(member), out.ctypes.data))
        return out

    def set_params(self, member, params):
        p = self._f32(params, (self.num_params,))
        _abi.check(_abi.load().avh_mlp_set_params(self._h, int(member), p.ctypes.data))

    def gradients(self, X, Y):
        """Loss and gradients of every member on `batch` rows (no update)."""
        X = self._f32(X, (self.batch, self.n_in))
        Y = self._f32(Y, (self.batch, self.n_out))
        g = np.empty((self.members, self.num_params), dtype=np.float32)
        loss = np.empty(self.members, dtype=np.float32)
        _abi.check(_abi.load().avh_mlp_gradients(self._h, X.ctypes.data, Y.ctypes.data, self.batch, g.ctypes.data,
                                                 loss.ctypes.data))
        return loss, g

    def adam_step(self, lr):
        _abi.check(_abi.load().avh_mlp_adam_step(self._h, float(lr)))

    def train(self, X, Y, epochs, lr=1e-3, final_lr_fraction=0.05):
        """Returns the mean minibatch loss per epoch and member."""
        X = self._f32(X)
        Y = self._f32(Y)
        rows = X.shape[0]
        assert X.shape == (rows, self.n_in) and Y.shape == (rows, self.n_out)
        hist = np.empty((int(epochs), self.members), dtype=np.float32)
        _abi.check(_abi.load().avh_mlp_train(self._h, X.ctypes.data, Y.ctypes.data, rows, int(epochs), float(lr),
                                             float(final_lr_fraction), hist.ctypes.data))
        return hist

    def predict_members(self, X):
        X = self._f32(X)
        rows = X.shape[0]
        out = np.empty((self.members, rows, self.n_out), dtype=np.float32)
        _abi.check(_abi.load().avh_mlp_predict(self._h, X.ctypes.data, rows, out.ctypes.data))
        return out

    def predict(self, X):
        """Ensemble mean and spread (population standard deviation over the
        members) per row and output."""
        p = self.predict_members(X).astype(np.float64)
        return p.mean(0), p.std(0)


def kh_observables(params, n=256, end_time=2.0, times=(0.5, 1.0, 1.5, 2.0), number_of_h=63, cfl=0.4,
                   chunk=1024, chunk_callback=None):
    """Observables of 2D Kelvin-Helmholtz samples (one row of 40 parameters
    each) from the batched solver: Y[:, :len(times)] = kinetic energy of u_y
    at each time, Y[:, len(times):] = S_2(h) of rho at the end time for
    h = 1 .. number_of_h - 1.  Every sample lands exactly on each time (a
    fixed-interval adjuster at the spacing of `times`)."""
    params = np.asarray(params, dtype=np.float64)
    times = [float(t) for t in times]
    spacing = times[0]
    for a, b in zip([0.0] + times[:-1], times):
        if not np.isclose(b - a, spacing):
            raise ValueError("observation times must be equally spaced from 0")
    out = np.empty((params.shape[0], len(times) + number_of_h - 1))
    for c0 in range(0, params.shape[0], chunk):
        p = params[c0:c0 + chunk]
        br = BatchRunner("euler2", "weno2", "hll3", "rungekutta3", "periodic", [((n, n, 1), p.shape[0])], end_time,
                         cfl=cfl, stats_saves=0)
        br.add_fixed_interval_writer(spacing)
        br.set_kelvin_helmholtz(0, p)
        for k, t in enumerate(times):
            br.set_end_time(t)
            done, left = br.run()
            if left:
                raise RuntimeError("samples did not reach t = %g" % t)
            out[c0:c0 + p.shape[0], k] = br.kinetic_energy(0, 1)
        out[c0:c0 + p.shape[0], len(times):] = br.structure_cube(0, 0, number_of_h, 2.0)[:, 1:]
        del br
        if chunk_callback:
            chunk_callback(c0 + p.shape[0])
    return out


def kh_parameters(sample_start, nsamples):
    """The 40 uniform perturbation parameters per sample, drawn like the
    reference's STLMersenne + Uniform generator."""
    return generate_parameters([("a", 40, 0.0, 1.0)], sample_start, nsamples)["a"]


def surrogate_statistics(model, X_query, solver=None, threshold=0.05, X_solved=None, Y_solved=None):
    """Statistics of the observables over the query samples.

    fallback:  the ensemble mean where the ensemble spread is within
               `threshold` of |mean| for every observable, the solver
               (`solver(X) -> Y`) elsewhere; mean and population variance
               over all query samples.
    control variate (when solver samples X_solved, Y_solved are given):
               E[f] = E_query[g] + mean over solved samples of (f - g),
               E[f^2] likewise, var = E[f^2] - E[f]^2; unbiased for any g.

    Returns a dict with both estimates (where available), the fallback
    fraction and the wall time of each part."""
    res = {}
    t0 = time.perf_counter()
    g, spread = model.predict(X_query)
    res["predict_seconds"] = time.perf_counter() - t0
    fallback = np.any(spread > threshold * np.abs(g), axis=1)
    res["fallback_fraction"] = float(fallback.mean())
    Y = g.copy()
    res["solver_seconds"] = 0.0
    if solver is not None and fallback.any():
        t0 = time.perf_counter()
        Y[fallback] = solver(np.asarray(X_query)[fallback])
        res["solver_seconds"] = time.perf_counter() - t0
    res["fallback_mean"] = Y.mean(0)
    res["fallback_var"] = Y.var(0)
    if X_solved is not None:
        gs, _ = model.predict(X_solved)
        f = np.asarray(Y_solved, dtype=np.float64)
        m1 = g.mean(0) + (f - gs).mean(0)
        m2 = (g * g).mean(0) + (f * f - gs * gs).mean(0)
        res["cv_mean"] = m1
        res["cv_var"] = m2 - m1 * m1
    return res
