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

    def set_weight_decay(self, weight_decay):
        """Decoupled weight decay (AdamW) of the weights in later steps."""
        _abi.check(_abi.load().avh_mlp_set_weight_decay(self._h, float(weight_decay)))

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

    def train_early_stopping(self, X, Y, X_val, Y_val, max_epochs=300, check_every=10, lr=1e-3, patience=5):
        """Adam at a constant learning rate in rounds of `check_every` epochs;
        after each round every member's mean squared error on the validation
        rows (in units of the outputs' standard deviations) is measured, and
        each member keeps the parameters of its best round.  Training stops
        when no member improved for `patience` rounds.  Returns the
        validation error per round and member, and the best round's per
        member."""
        Xv = self._f32(X_val)
        Yv = np.asarray(Y_val, dtype=np.float64)
        sd = np.maximum(Yv.std(0), 1e-300)
        best = [None] * self.members
        best_err = np.full(self.members, np.inf)
        best_round = np.zeros(self.members, dtype=int)
        hist = []
        stale = 0
        for rnd in range(max(1, max_epochs // check_every)):
            self.train(X, Y, epochs=check_every, lr=lr, final_lr_fraction=1.0)
            pm = self.predict_members(Xv).astype(np.float64)
            err = np.mean(((pm - Yv[None]) / sd) ** 2, axis=(1, 2))
            hist.append(err)
            improved = err < best_err
            for m in np.nonzero(improved)[0]:
                best[m] = self.get_params(int(m))
                best_err[m] = err[m]
                best_round[m] = rnd
            stale = 0 if improved.any() else stale + 1
            if stale >= patience:
                break
        for m in range(self.members):
            if best[m] is not None:
                self.set_params(m, best[m])
        return np.array(hist), best_round

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
            br.check_constraints()
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


def spread_calibration(model, X_solved, Y_solved, threshold=0.05):
    """Whether the ensemble spread is a trustworthy error indicator, per
    observable: among solver samples the network did not train on whose
    spread passes the fallback test (spread <= threshold |mean| for every
    observable), the fraction whose prediction is within threshold |mean|
    of the solver's value.  An ensemble whose members agree with each other
    but not with the solver (all of them near the mean of a map they did not
    learn) passes the spread test and fails this one."""
    g, spread = model.predict(X_solved)
    y = np.asarray(Y_solved, dtype=np.float64)
    accepted = ~np.any(spread > threshold * np.abs(g), axis=1)
    if not accepted.any():
        return np.zeros(g.shape[1]), 0
    within = np.abs(g[accepted] - y[accepted]) <= threshold * np.abs(g[accepted])
    return within.mean(0), int(accepted.sum())


def surrogate_statistics(model, X_query, solver=None, threshold=0.05, X_solved=None, Y_solved=None,
                         min_coverage=None):
    """Statistics of the observables over the query samples.

    fallback:  the ensemble mean where the ensemble spread is within
               `threshold` of |mean| for every observable, the solver
               (`solver(X) -> Y`) elsewhere; mean and population variance
               over all query samples.  With `min_coverage` (and solver
               samples X_solved, Y_solved) the spread is trusted only if it
               is calibrated: every observable's coverage
               (spread_calibration) is at least min_coverage; otherwise
               every query sample goes to the solver.
    control variate (when solver samples X_solved, Y_solved are given):
               E[f] = mean_S(f) - beta (mean_S(g) - mean_Q(g)) with the
               variance-minimising beta = cov_S(f, g) / var_S(g) per
               observable (S: solved samples, Q: query samples), E[f^2]
               likewise with f^2, g^2; var = E[f^2] - E[f]^2.  Its
               variance is var(f) (1 - rho^2) / |S|, never above plain
               Monte Carlo's on the same solved samples.

    Returns a dict with both estimates (where available), the fallback
    fraction and the wall time of each part."""
    res = {}
    t0 = time.perf_counter()
    g, spread = model.predict(X_query)
    res["predict_seconds"] = time.perf_counter() - t0
    fallback = np.any(spread > threshold * np.abs(g), axis=1)
    res["spread_fallback_fraction"] = float(fallback.mean())
    if min_coverage is not None and X_solved is not None:
        cov, n_acc = spread_calibration(model, X_solved, Y_solved, threshold)
        res["spread_coverage"] = cov
        res["spread_coverage_samples"] = n_acc
        res["spread_trusted"] = bool(n_acc > 0 and np.all(cov >= min_coverage))
        if not res["spread_trusted"]:
            fallback = np.ones_like(fallback)
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
        b1 = optimal_beta(f, gs)
        b2 = optimal_beta(f * f, gs * gs)
        m1 = f.mean(0) - b1 * (gs.mean(0) - g.mean(0))
        m2 = (f * f).mean(0) - b2 * ((gs * gs).mean(0) - (g * g).mean(0))
        res["cv_beta"] = b1
        res["cv_mean"] = m1
        res["cv_var"] = m2 - m1 * m1
    return res


def optimal_beta(f, g):
    """cov(f, g) / var(g) per column (0 where g is constant)."""
    fc = f - f.mean(0)
    gc = g - g.mean(0)
    vg = (gc * gc).mean(0)
    return np.where(vg > 0, (fc * gc).mean(0) / np.where(vg > 0, vg, 1.0), 0.0)


def control_variate_efficiency(model, X_solved, Y_solved):
    """Per observable var(f) / var(f - beta g) = 1 / (1 - rho(f, g)^2) on
    solver samples the network did not train on, beta the variance-minimising
    coefficient: the factor by which the control-variate estimator cuts the
    number of solver samples for a given error of the mean (1 = no gain; it
    is never below 1)."""
    g, _ = model.predict(X_solved)
    f = np.asarray(Y_solved, dtype=np.float64)
    b = optimal_beta(f, g)
    return f.var(0) / np.maximum((f - b * g).var(0), 1e-300)


def _rel_err(est, ref):
    return np.abs(est - ref) / np.maximum(np.abs(ref), 1e-300)


def plain_mc_error(pool, ref_mean, ref_var, nsamples, reps=16, seed=0):
    """Mean over observables of the relative error of plain Monte Carlo mean
    and variance with `nsamples` samples, averaged over `reps` random subsets
    of the solver pool."""
    rng = np.random.default_rng(seed)
    em, ev = [], []
    for _ in range(reps):
        sub = pool[rng.choice(pool.shape[0], nsamples, replace=False)]
        em.append(_rel_err(sub.mean(0), ref_mean).mean())
        ev.append(_rel_err(sub.var(0), ref_var).mean())
    return float(np.mean(em)), float(np.mean(ev))


def uq_experiment(n=256, end_time=2.0, times=(0.5, 1.0, 1.5, 2.0), number_of_h=63, n_train=4096, n_holdout=512,
                  n_ref=16384, n_query=16384, epochs=300, lr=1e-3, members=5, width=256, hidden=4, batch=256,
                  threshold=0.05, chunk=2048, data=None, log=None, seed=0, weight_decay=0.01, check_every=10,
                  min_coverage=0.9):
    """Surrogate-assisted UQ of the Kelvin-Helmholtz observables against plain
    Monte Carlo.  Sample ids: training [0, n_train), held out next, then the
    plain-solver reference, then the surrogate's query samples.  `data` may
    hold solver observables from an earlier call ({"train", "holdout", "ref"}
    arrays plus "solver_seconds_per_sample") to skip those runs.  Returns a
    dict of errors and wall times."""
    log = log or (lambda *a: None)
    data = dict(data or {})
    n_out = len(times) + number_of_h - 1
    ids = {"train": (0, n_train), "holdout": (n_train, n_holdout), "ref": (n_train + n_holdout, n_ref)}
    spent = {}
    for key, (start, count) in ids.items():
        if key in data and data[key].shape == (count, n_out):
            continue
        t0 = time.perf_counter()
        data[key] = kh_observables(kh_parameters(start, count), n=n, end_time=end_time, times=times,
                                   number_of_h=number_of_h, chunk=chunk,
                                   chunk_callback=lambda k, key=key: log("solver %s: %d samples" % (key, k)))
        spent[key] = time.perf_counter() - t0
    if spent:
        data["solver_seconds_per_sample"] = sum(spent.values()) / sum(ids[k][1] for k in spent)
    t_sample = float(data["solver_seconds_per_sample"])
    Xtr = kh_parameters(*ids["train"])
    Xho = kh_parameters(*ids["holdout"])
    q0 = n_train + n_holdout + n_ref
    Xq = kh_parameters(q0, n_query)
    Ytr, Yho, Yref = data["train"], data["holdout"], data["ref"]
    ref_mean, ref_var = Yref.mean(0), Yref.var(0)

    model = EnsembleMLP(40, n_out, width=width, hidden=hidden, members=members, batch=batch, seed=seed)
    model.fit_normalization(Xtr, Ytr)
    model.set_weight_decay(weight_decay)
    # early stopping on the last eighth of the training samples (the held-out
    # samples are kept for the control variate and the error estimates)
    nv = max(batch, n_train // 8)
    t0 = time.perf_counter()
    vhist, best_round = model.train_early_stopping(Xtr[:-nv], Ytr[:-nv], Xtr[-nv:], Ytr[-nv:], max_epochs=epochs,
                                                   check_every=check_every, lr=lr)
    t_train = time.perf_counter() - t0
    log("trained %d rounds of %d epochs in %.1f s, validation error %.3g -> best %.3g (rounds %s)"
        % (len(vhist), check_every, t_train, vhist[0].mean(), vhist.min(0).mean(), list(best_round)))
    g_ho, s_ho = model.predict(Xho)
    holdout_nmse = (np.mean((g_ho - Yho) ** 2, axis=0) / np.maximum(Yho.var(0), 1e-300))

    def solver(X):
        start = time.perf_counter()
        # fallback samples are fresh query samples: their parameters are given
        out = kh_observables(X, n=n, end_time=end_time, times=times, number_of_h=number_of_h,
                             chunk=min(chunk, 1024), chunk_callback=lambda k: log("fallback solver: %d samples" % k))
        log("fallback solver: %d samples in %.1f s" % (X.shape[0], time.perf_counter() - start))
        return out

    st = surrogate_statistics(model, Xq, solver=solver, threshold=threshold, X_solved=Xho, Y_solved=Yho,
                              min_coverage=min_coverage)
    res = {
        "n": n, "end_time": end_time, "n_train": n_train, "n_holdout": n_holdout, "n_ref": n_ref,
        "n_query": n_query, "max_epochs": epochs, "epochs_trained": int(len(vhist) * check_every),
        "best_epoch_per_member": [int((r + 1) * check_every) for r in best_round], "weight_decay": weight_decay,
        "solver_seconds_per_sample": t_sample, "train_seconds": t_train,
        "best_validation_error": float(vhist.min(0).mean()), "holdout_nmse_mean": float(holdout_nmse.mean()),
        "holdout_nmse_max": float(holdout_nmse.max()), "fallback_fraction": st["fallback_fraction"],
        "spread_fallback_fraction": st["spread_fallback_fraction"], "spread_trusted": st.get("spread_trusted"),
        "spread_coverage_min": float(np.min(st["spread_coverage"])) if "spread_coverage" in st else None,
        "spread_coverage_samples": st.get("spread_coverage_samples"),
        "predict_seconds": st["predict_seconds"], "fallback_solver_seconds": st["solver_seconds"],
    }
    # errors of each estimator against the plain-solver reference
    for name, m, v in (("fallback", st["fallback_mean"], st["fallback_var"]), ("cv", st["cv_mean"], st["cv_var"])):
        res[name + "_mean_err"] = float(_rel_err(m, ref_mean).mean())
        res[name + "_var_err"] = float(_rel_err(v, ref_var).mean())
    eff = control_variate_efficiency(model, Xho, Yho)
    res["cv_efficiency_median"] = float(np.median(eff))
    res["cv_efficiency_min"] = float(eff.min())
    res["cv_efficiency_max"] = float(eff.max())
    res["holdout_nmse"] = holdout_nmse.tolist()
    res["cv_efficiency"] = eff.tolist()
    budget = n_train + n_holdout
    res["plain_mean_err_at_budget"], res["plain_var_err_at_budget"] = plain_mc_error(Yref, ref_mean, ref_var,
                                                                                      min(budget, n_ref // 2))
    # fallback: the solver samples it used in total, and plain Monte Carlo's
    # error with that many samples (subsets of the reference pool)
    n_fb = budget + int(round(st["fallback_fraction"] * n_query))
    res["fallback_solver_samples"] = n_fb
    res["plain_mean_err_at_fallback_samples"], res["plain_var_err_at_fallback_samples"] = plain_mc_error(
        Yref, ref_mean, ref_var, min(n_fb, n_ref // 2))
    wall_sur = budget * t_sample + t_train + st["predict_seconds"] + st["solver_seconds"]
    res["surrogate_wall_seconds"] = wall_sur
    # control variate: its variance is var(f - g) / n_holdout (+ var(g) /
    # n_query), so it stands for n_holdout * var(f) / var(f - g) plain solver
    # samples per observable; the speedup is that plain run's solver time over
    # the whole surrogate pipeline's wall time (training samples included)
    n_eq = n_holdout * float(np.median(eff))
    res["cv_plain_equivalent_samples_median"] = n_eq
    res["cv_speedup_median"] = n_eq * t_sample / (budget * t_sample + t_train + st["predict_seconds"])
    return res, data


def main(argv=None):
    """Staged driver for the Kelvin-Helmholtz surrogate study (one GPU call
    per stage, observables cached as .npz in --data):

      --stage train  solver observables of the training and held-out samples
      --stage ref    a part of the plain-solver reference (--ref-start, --ref-count)
      --stage eval   train the ensemble, query, fall back, compare; prints JSON
    """
    import argparse
    import glob
    import json
    import os

    ap = argparse.ArgumentParser(description=main.__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--stage", choices=["train", "ref", "eval"], required=True)
    ap.add_argument("--data", required=True, help="directory of the cached observables")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--end-time", type=float, default=2.0)
    ap.add_argument("--n-train", type=int, default=4096)
    ap.add_argument("--n-holdout", type=int, default=512)
    ap.add_argument("--n-ref", type=int, default=16384)
    ap.add_argument("--ref-start", type=int, default=0)
    ap.add_argument("--ref-count", type=int, default=4096)
    ap.add_argument("--n-query", type=int, default=16384)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--weight-decay", type=float, default=0.01)
    ap.add_argument("--chunk", type=int, default=2048)
    ap.add_argument("--min-coverage", type=float, default=0.9,
                    help="trust the ensemble spread only if, on held-out solver samples, at least this fraction "
                         "of the accepted predictions is within the threshold of the solver's value")
    args = ap.parse_args(argv)
    os.makedirs(args.data, exist_ok=True)
    t0 = time.perf_counter()

    def log(*a):
        print("[%.1f s]" % (time.perf_counter() - t0), *a, flush=True)

    times = tuple(args.end_time * k / 4 for k in (1, 2, 3, 4))
    kw = dict(n=args.n, end_time=args.end_time, times=times, number_of_h=63, chunk=args.chunk)
    if args.stage == "train":
        out = {}
        for key, start, count in (("train", 0, args.n_train), ("holdout", args.n_train, args.n_holdout)):
            ts = time.perf_counter()
            out[key] = kh_observables(kh_parameters(start, count), chunk_callback=lambda k: log(key, k), **kw)
            out[key + "_seconds"] = time.perf_counter() - ts
        np.savez(os.path.join(args.data, "kh_train.npz"), **out)
        log("saved", {k: np.shape(v) for k, v in out.items()})
    elif args.stage == "ref":
        start = args.n_train + args.n_holdout + args.ref_start
        ts = time.perf_counter()
        Y = kh_observables(kh_parameters(start, args.ref_count), chunk_callback=lambda k: log("ref", k), **kw)
        np.savez(os.path.join(args.data, "kh_ref_%06d.npz" % args.ref_start), ref=Y,
                 seconds=time.perf_counter() - ts)
        log("saved ref part", args.ref_start, Y.shape)
    else:
        tr = np.load(os.path.join(args.data, "kh_train.npz"))
        parts = sorted(glob.glob(os.path.join(args.data, "kh_ref_*.npz")))
        refs = [np.load(p) for p in parts]
        Yref = np.concatenate([r["ref"] for r in refs])[:args.n_ref]
        seconds = float(tr["train_seconds"] + tr["holdout_seconds"] + sum(float(r["seconds"]) for r in refs))
        count = tr["train"].shape[0] + tr["holdout"].shape[0] + sum(r["ref"].shape[0] for r in refs)
        data = {"train": tr["train"], "holdout": tr["holdout"], "ref": Yref,
                "solver_seconds_per_sample": seconds / count}
        res, _ = uq_experiment(n_train=tr["train"].shape[0], n_holdout=tr["holdout"].shape[0], n_ref=Yref.shape[0],
                               n_query=args.n_query, epochs=args.epochs, lr=args.lr, data=data, log=log,
                               weight_decay=args.weight_decay, min_coverage=args.min_coverage, **kw)
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
