"""The learned rate surrogate, host side (no GPU).

The network's CPU oracle is the NumPy restatement below (forward pass and
backward pass of the mean squared error).  It is pinned by central finite
differences in float64, which a wrong transpose, a dropped derivative factor or
a wrong index would fail.  The training samples are pinned against the
oracle's restatement of the reference's rate function (itself pinned to the
compiled reference)."""
import numpy as np
import pytest

import oracle_bind as ob

IN, H, OUT = 3, 64, 2


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + OUT * H].reshape(OUT, H); o += OUT * H
    b3 = p[o:o + OUT]; o += OUT
    assert o == len(p)
    return W1, b1, W2, b2, W3, b3


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h1, h2, h2 @ W3.T + b3


def loss_grad(p, X, T):
    """Mean squared error sum_{s,k} (y-t)^2 / n and its gradient."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = len(X)
    h1, h2, y = forward(p, X)
    loss = ((y - T) ** 2).sum() / n
    dy = 2 * (y - T) / n
    d2 = (dy @ W3) * (1 - h2 ** 2)
    d1 = (d2 @ W2) * (1 - h1 ** 2)
    g = np.concatenate([(d1.T @ X).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0), (dy.T @ h2).ravel(),
                        dy.sum(0)])
    return loss, g


def test_numpy_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, 4546)
    X = rng.normal(size=(17, 3))
    T = rng.normal(size=(17, 2))
    _, g = loss_grad(p, X, T)
    idx = rng.choice(len(p), 60, replace=False)
    idx = np.concatenate([idx, [0, 191, 192, 4545, 4543, 256 + 4095]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (loss_grad(p + e, X, T)[0] - loss_grad(p - e, X, T)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def test_samples_follow_the_reference_rate_function():
    from smoldyn_amd import surrogate as sg
    X, Y, redraws = sg.samples(3000, seed=5)
    L = ob.lib()
    step = np.exp(X[:, 1])
    a = step * np.exp(Y[:, 0])
    b = step * np.exp(Y[:, 1])
    assert np.allclose(b / a, X[:, 2], rtol=1e-12)
    for k in range(0, 3000, 7):
        r = L.orxn_numrxnrate(step[k], a[k], b[k])
        assert r == pytest.approx(np.exp(X[k, 0] + 3 * X[k, 1]), rel=1e-12)
    soa = step / a
    assert soa.min() >= 0.05 - 1e-12 and soa.max() <= 5 + 1e-12
    assert redraws < 30


def test_polish_and_judge_and_numerical_inversion():
    """The product's bisection equals the oracle's bindingradius bit for bit;
    the secant correction from a 1% error brings the rate error below 1e-3."""
    from smoldyn_amd import surrogate as sg
    X, Y, _ = sg.samples(500, seed=6)
    ratedt = np.exp(X[:, 0] + 3 * X[:, 1])
    step = np.exp(X[:, 1])
    ratio = X[:, 2]
    a_num = sg.invert_numerical(ratedt, step, ratio)
    L = ob.lib()
    for k in range(0, 500, 11):
        ref = L.orxn_bindingradius(ratedt[k], 1.0, step[k] ** 2 / 2, ratio[k], 1)
        assert a_num[k] == ref
    a_true = step * np.exp(Y[:, 0])
    a0 = a_true * (1 + 0.01 * np.random.default_rng(1).standard_normal(500))
    err0 = sg.rate_error(ratedt, step, ratio, a0)
    a1 = np.zeros(500)
    import ctypes as C
    sg._lib().smolamdRxnparamPolish(500, sg._d(ratedt), sg._d(step), sg._d(ratio), sg._d(a0), 1e-3, sg._d(a1))
    err1 = sg.rate_error(ratedt, step, ratio, a1)
    assert np.median(err1) < np.median(err0) / 50
    assert np.mean(err1 <= 1e-3) > 0.995
