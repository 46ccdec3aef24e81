"""The surrogate network's HIP kernels (bf16 matrix cores, fp32 accumulate)
against the network oracle (oracle/mlp_oracle.py) in its bf16-storage mode:
loss and gradients of every member, Adam steps, predictions on row counts
that are not a multiple of the minibatch, and training on a smooth map.
Padded feature dimensions are exercised (40 inputs -> 48, 66 outputs -> 80).

Tolerances: the kernels accumulate in fp32 in the matrix cores' order, the
oracle in float64; a product that lands next to a bf16 rounding boundary can
round the other way, so gradients are compared in norm (1 %), losses to
1e-3 relative."""
import sys

import numpy as np
import pytest

from alsvinn_amd.surrogate import EnsembleMLP

pytestmark = pytest.mark.gpu

N_IN, N_OUT, WIDTH, HIDDEN, MEMBERS, BATCH = 40, 66, 64, 3, 3, 64


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _setup(seed=0):
    from oracle import mlp_oracle as mo
    rng = np.random.default_rng(seed)
    net = EnsembleMLP(N_IN, N_OUT, width=WIDTH, hidden=HIDDEN, members=MEMBERS, batch=BATCH, seed=seed)
    oracles = []
    for m in range(MEMBERS):
        o = mo.MlpOracle(N_IN, WIDTH, HIDDEN, N_OUT)
        o.random_init(rng)
        o.set_flat(o.flat().astype(np.float32).astype(np.float64))
        net.set_params(m, o.flat())
        oracles.append(o)
    X = rng.random((BATCH, N_IN)).astype(np.float32)
    Y = (rng.normal(size=(BATCH, N_OUT)) * 3 + 2).astype(np.float32)
    net.fit_normalization(X, Y)
    norm = tuple(np.asarray(v, np.float32).astype(np.float64) for v in net.norm)
    return net, oracles, X, Y, norm


def test_params_round_trip():
    net, oracles, X, Y, norm = _setup(1)
    for m in range(MEMBERS):
        np.testing.assert_array_equal(net.get_params(m), oracles[m].flat().astype(np.float32))
    assert net.num_params == oracles[0].num_params() == 40 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 66 + 66


def test_gradients_match_oracle():
    net, oracles, X, Y, norm = _setup(2)
    _say("case mlp gradients")
    loss, g = net.gradients(X, Y)
    for m in range(MEMBERS):
        lo, go, _ = oracles[m].loss_and_grads(X, Y, norm, exact=False)
        le, ge, _ = oracles[m].loss_and_grads(X, Y, norm, exact=True)
        _say("member", m, "loss", loss[m], lo, le, "grad rel err",
             np.linalg.norm(g[m] - go) / np.linalg.norm(go), np.linalg.norm(g[m] - ge) / np.linalg.norm(ge))
        assert loss[m] == pytest.approx(lo, rel=1e-3)
        assert np.linalg.norm(g[m] - go) <= 1e-2 * np.linalg.norm(go)
        # and within bf16 storage error of the exact float64 gradient
        assert np.linalg.norm(g[m] - ge) <= 5e-2 * np.linalg.norm(ge)
        # per layer, so a transposed or misplaced block cannot hide in the norm
        o = 0
        for (r, c) in oracles[m].shapes:
            for size in (r * c, r):
                sl = slice(o, o + size)
                assert np.linalg.norm(g[m][sl] - go[sl]) <= 2e-2 * np.linalg.norm(go[sl]) + 1e-7
                o += size


def test_adam_steps_match_oracle():
    from oracle import mlp_oracle as mo
    net, oracles, X, Y, norm = _setup(3)
    p = [net.get_params(m).astype(np.float64) for m in range(MEMBERS)]
    mst = [np.zeros_like(x) for x in p]
    vst = [np.zeros_like(x) for x in p]
    for t in (1, 2, 3):
        _, g = net.gradients(X, Y)
        net.adam_step(1e-3)
        for m in range(MEMBERS):
            p[m], mst[m], vst[m] = mo.adam(p[m], g[m].astype(np.float64), mst[m], vst[m], t, 1e-3)
            got = net.get_params(m).astype(np.float64)
            np.testing.assert_allclose(got, p[m], rtol=2e-5, atol=2e-7)
            p[m] = got


def test_adamw_steps_match_oracle():
    """Decoupled weight decay on the weights (not the biases)."""
    from oracle import mlp_oracle as mo
    net, oracles, X, Y, norm = _setup(7)
    net.set_weight_decay(0.05)
    mask = oracles[0].weight_mask()
    p = [net.get_params(m).astype(np.float64) for m in range(MEMBERS)]
    mst = [np.zeros_like(x) for x in p]
    vst = [np.zeros_like(x) for x in p]
    for t in (1, 2, 3):
        _, g = net.gradients(X, Y)
        net.adam_step(1e-3)
        for m in range(MEMBERS):
            p[m], mst[m], vst[m] = mo.adam(p[m], g[m].astype(np.float64), mst[m], vst[m], t, 1e-3,
                                           weight_decay=0.05, decay_mask=mask)
            got = net.get_params(m).astype(np.float64)
            np.testing.assert_allclose(got, p[m], rtol=2e-5, atol=2e-7)
            p[m] = got


def test_predict_matches_oracle_forward():
    net, oracles, X, Y, norm = _setup(4)
    rng = np.random.default_rng(5)
    Xq = rng.random((100, N_IN)).astype(np.float32)  # not a multiple of the minibatch
    out = net.predict_members(Xq)
    assert out.shape == (MEMBERS, 100, N_OUT)
    for m in range(MEMBERS):
        ref = oracles[m].predict(Xq, norm, exact=False)
        scale = np.abs(ref).max()
        assert np.max(np.abs(out[m] - ref)) <= 1e-2 * scale
    mean, spread = net.predict(Xq)
    np.testing.assert_allclose(mean, out.astype(np.float64).mean(0), rtol=1e-6)
    assert np.all(spread > 0)


def test_training_fits_a_smooth_map():
    rng = np.random.default_rng(6)

    def f(X):
        return np.stack([np.sin(2 * np.pi * X[:, 0]) + X[:, 1] ** 2, X[:, 2] * X[:, 3] + 0.5 * X[:, 4],
                         np.cos(np.pi * X[:, 5]) * X[:, 6], X[:, :8].sum(1)], axis=1)
    X = rng.random((2048, 8)).astype(np.float32)
    Y = f(X).astype(np.float32)
    Xt = rng.random((512, 8)).astype(np.float32)
    Yt = f(Xt)
    net = EnsembleMLP(8, 4, width=128, hidden=3, members=5, batch=128, seed=7)
    net.fit_normalization(X, Y)
    hist = net.train(X, Y, epochs=60, lr=2e-3)
    _say("loss first/last epoch", hist[0], hist[-1])
    assert np.all(hist[-1] < 0.05 * hist[0])
    mean, spread = net.predict(Xt)
    nmse = np.mean((mean - Yt) ** 2, axis=0) / np.var(Yt, axis=0)
    _say("held-out normalised MSE", nmse, "mean spread / std", spread.mean(0) / Yt.std(0))
    assert np.all(nmse < 0.02)
    # independently initialised members disagree a little, not wildly
    rel = spread.mean(0) / Yt.std(0)
    assert np.all(rel > 1e-4) and np.all(rel < 0.2)
