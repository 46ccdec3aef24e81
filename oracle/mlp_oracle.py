"""CPU ORACLE -- test infrastructure, not product code.

The surrogate ensemble network in numpy, for checking the HIP kernels
(alsvinn_amd/csrc/avh_mlp.hip).  The reference has no learned model, so
there is no reference file to follow; the network is defined by its
mathematics:

    x_0 = (x - mean_in) / std_in
    z_l = x_l W_l^T + b_l,  x_{l+1} = gelu(z_l) = z_l Phi(z_l)   (hidden layers)
    yhat = x_H W_H^T + b_H
    loss = mean over rows and outputs of (yhat - (y - mean_out) / std_out)^2

and Adam (Kingma & Ba, bias-corrected, eps outside the square root).

Two modes:
  exact=True   float64 throughout, no rounding: pinned against central
               finite differences and torch autograd / torch.optim.Adam in
               tests/test_mlp_oracle.py;
  exact=False  the kernels' storage rounding: inputs, weights, activations
               and backward signals rounded to bf16 (round to nearest even)
               where the kernels store them, products accumulated in float64
               then rounded to float32 (the kernels accumulate in fp32 on the
               matrix cores), biases, pre-activations and gradients float32.

Only tests/ import this module.
"""
import math

import numpy as np

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    _erf = np.vectorize(math.erf)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


class MlpOracle:
    """One member: layers = [(W [out][in], b [out]), ...]."""

    def __init__(self, n_in, width, hidden, n_out):
        self.n_in, self.width, self.hidden, self.n_out = n_in, width, hidden, n_out
        self.shapes = [(width, n_in)] + [(width, width)] * (hidden - 1) + [(n_out, width)]
        self.layers = [(np.zeros(s), np.zeros(s[0])) for s in self.shapes]

    def num_params(self):
        return sum(o * i + o for o, i in self.shapes)

    def weight_mask(self):
        """1 for the weights, 0 for the biases, in the flat parameter order
        (per layer W row-major, then b)."""
        return np.concatenate([np.concatenate([np.ones(o * i), np.zeros(o)]) for o, i in self.shapes])

    def set_flat(self, flat):
        flat = np.asarray(flat, dtype=np.float64)
        o = 0
        layers = []
        for (r, c) in self.shapes:
            W = flat[o:o + r * c].reshape(r, c)
            o += r * c
            b = flat[o:o + r]
            o += r
            layers.append((W.copy(), b.copy()))
        assert o == flat.size
        self.layers = layers

    def flat(self):
        return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in self.layers])

    def random_init(self, rng):
        layers = []
        for (r, c) in self.shapes:
            bound = 1.0 / math.sqrt(c)
            layers.append((rng.uniform(-bound, bound, (r, c)), rng.uniform(-bound, bound, r)))
        self.layers = layers

    # ------------------------------------------------------------------
    def _mm(self, A, B, exact):
        """A @ B with float64 accumulation; float32 result unless exact."""
        C = np.asarray(A, dtype=np.float64) @ np.asarray(B, dtype=np.float64)
        return C if exact else C.astype(np.float32)

    def loss_and_grads(self, X, Y, norm, exact=True):
        """norm = (in_mean, in_std, out_mean, out_std).  Returns (loss, flat
        gradient, outputs yhat in normalised units)."""
        mi, si, mo, so = [np.asarray(v, dtype=np.float64) for v in norm]
        if exact:
            rnd = lambda v: np.asarray(v, dtype=np.float64)  # noqa: E731
            f = np.float64
        else:
            rnd = lambda v: bf16(np.asarray(v, dtype=np.float32)).astype(np.float64)  # noqa: E731
            f = np.float32
        B = X.shape[0]
        x = rnd((np.asarray(X, f) - mi.astype(f)) / si.astype(f))
        yn = ((np.asarray(Y, f) - mo.astype(f)) / so.astype(f)).astype(np.float64)
        xs, zs = [x], []
        for l, (W, b) in enumerate(self.layers):
            Wr = rnd(W)
            z = (self._mm(xs[-1], Wr.T, exact).astype(f) + np.asarray(b, f)).astype(np.float64)
            if l < len(self.layers) - 1:
                zs.append(z)
                xs.append(rnd(np.asarray(gelu(z), f)))
            else:
                yhat = z
        d = yhat - yn
        loss = float(np.mean(d * d))
        g = rnd(np.asarray(2.0 * d / (B * self.n_out), f))
        grads = [None] * len(self.layers)
        for l in range(len(self.layers) - 1, -1, -1):
            W, b = self.layers[l]
            dW = self._mm(g.T, xs[l], exact)
            db = np.sum(g, axis=0)
            grads[l] = (np.asarray(dW, np.float64), np.asarray(db, np.float64))
            if l == 0:
                break
            dx = self._mm(g, rnd(W), exact).astype(np.float64)
            g = rnd(np.asarray(dx * gelu_grad(zs[l - 1]), f))
        flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
        return loss, flat, yhat

    def predict(self, X, norm, exact=True):
        mi, si, mo, so = [np.asarray(v, dtype=np.float64) for v in norm]
        _, _, yhat = self.loss_and_grads(X, np.zeros((X.shape[0], self.n_out)) + mo, norm, exact)
        return yhat * so + mo


def adam(params, grads, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0, decay_mask=None):
    """One bias-corrected Adam step (t counts from 1); returns new arrays.
    weight_decay: decoupled decay (AdamW, Loshchilov and Hutter 2019),
    p -= lr * weight_decay * p on the entries where decay_mask is 1."""
    m = beta1 * m + (1 - beta1) * grads
    v = beta2 * v + (1 - beta2) * grads * grads
    p = params - lr * (m / (1 - beta1 ** t)) / (np.sqrt(v / (1 - beta2 ** t)) + eps)
    if weight_decay:
        mask = np.ones_like(params) if decay_mask is None else decay_mask
        p = p - lr * weight_decay * params * mask
    return p, m, v
