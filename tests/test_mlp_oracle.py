"""Pins of the surrogate-network oracle (oracle/mlp_oracle.py) by
independent means: central finite differences of its float64 loss, torch
autograd of the same network, torch.optim.Adam, and torch's bfloat16
conversion (round to nearest even)."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


def _net(seed=0, n_in=5, width=16, hidden=3, n_out=4):
    rng = np.random.default_rng(seed)
    net = mo.MlpOracle(n_in, width, hidden, n_out)
    net.random_init(rng)
    X = rng.random((16, n_in))
    Y = rng.normal(size=(16, n_out)) * 2 + 1
    norm = (X.mean(0), X.std(0) + 0.1, Y.mean(0), Y.std(0) + 0.1)
    return net, X, Y, norm


def test_bf16_matches_torch_round_to_nearest_even():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(size=5000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 5000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.0078125, 1.01171875, 3.0e38, -3.0e38,
                                  np.inf, -np.inf, 1e-40], dtype=np.float32)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mo.bf16(x), ref)


def test_gradients_match_finite_differences():
    net, X, Y, norm = _net()
    loss, g, _ = net.loss_and_grads(X, Y, norm, exact=True)
    p0 = net.flat()
    rng = np.random.default_rng(2)
    for i in rng.choice(p0.size, 60, replace=False):
        h = 1e-6
        p = p0.copy()
        p[i] += h
        net.set_flat(p)
        lp = net.loss_and_grads(X, Y, norm, exact=True)[0]
        p[i] -= 2 * h
        net.set_flat(p)
        lm = net.loss_and_grads(X, Y, norm, exact=True)[0]
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-9)
    net.set_flat(p0)


def test_forward_and_gradients_match_torch_autograd():
    net, X, Y, norm = _net(seed=3)
    loss, g, yhat = net.loss_and_grads(X, Y, norm, exact=True)
    mi, si, mo_, so = [torch.tensor(v, dtype=torch.float64) for v in norm]
    x = (torch.tensor(X) - mi) / si
    params = []
    for W, b in net.layers:
        Wt = torch.tensor(W, requires_grad=True)
        bt = torch.tensor(b, requires_grad=True)
        params += [Wt, bt]
    h = x
    for l in range(len(net.layers)):
        h = h @ params[2 * l].T + params[2 * l + 1]
        if l < len(net.layers) - 1:
            h = torch.nn.functional.gelu(h, approximate="none")
    yn = (torch.tensor(Y) - mo_) / so
    L = torch.mean((h - yn) ** 2)
    L.backward()
    ref = torch.cat([torch.cat([params[2 * l].grad.ravel(), params[2 * l + 1].grad]) for l in range(len(net.layers))])
    assert loss == pytest.approx(L.item(), rel=1e-13)
    np.testing.assert_allclose(yhat, h.detach().numpy(), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(g, ref.numpy(), rtol=1e-10, atol=1e-14)


def test_adamw_matches_torch():
    """Decoupled weight decay on the weights only (biases in a group of
    their own without decay), against torch.optim.AdamW."""
    net, X, Y, norm = _net(5)
    rng = np.random.default_rng(6)
    p = net.flat().copy()
    mask = net.weight_mask()
    assert mask.sum() == sum(o * i for o, i in net.shapes) and mask.size == net.num_params()
    pw = torch.tensor(p[mask == 1].copy(), requires_grad=True)
    pb = torch.tensor(p[mask == 0].copy(), requires_grad=True)
    opt = torch.optim.AdamW([{"params": [pw], "weight_decay": 0.05}, {"params": [pb], "weight_decay": 0.0}],
                            lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in range(1, 8):
        g = rng.normal(size=p.size)
        pw.grad = torch.tensor(g[mask == 1])
        pb.grad = torch.tensor(g[mask == 0])
        opt.step()
        p, m, v = mo.adam(p, g, m, v, t, 3e-3, weight_decay=0.05, decay_mask=mask)
        np.testing.assert_allclose(p[mask == 1], pw.detach().numpy(), rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(p[mask == 0], pb.detach().numpy(), rtol=1e-13, atol=1e-15)


def test_adam_matches_torch():
    rng = np.random.default_rng(4)
    p = rng.normal(size=50)
    pt = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in range(1, 8):
        g = rng.normal(size=50)
        pt.grad = torch.tensor(g)
        opt.step()
        p, m, v = mo.adam(p, g, m, v, t, 3e-3)
        np.testing.assert_allclose(p, pt.detach().numpy(), rtol=1e-13, atol=1e-15)


def test_rounded_mode_stays_close_to_exact():
    """bf16 storage perturbs the loss and gradients by bf16's 2^-8 relative
    precision, not more (a check that the rounded mode rounds where it should
    and nowhere else by accident)."""
    net, X, Y, norm = _net(seed=5, width=32)
    l0, g0, _ = net.loss_and_grads(X, Y, norm, exact=True)
    l1, g1, _ = net.loss_and_grads(X, Y, norm, exact=False)
    assert l1 == pytest.approx(l0, rel=3e-2)
    assert np.linalg.norm(g1 - g0) <= 5e-2 * np.linalg.norm(g0)
