"""The learned rate surrogate on the GPU.

Gradients of the device forward/backward kernels are compared with the NumPy
oracle of the same network (test_surrogate_host.py pins that oracle by finite
differences); training on the device must reduce the loss; BF16 inference must
agree with FP32 inference; the committed trained network must be accepted by
the judge (the reference's rate function) for nearly all queries of the
parameter ranges it was trained on, every returned radius must meet the 1e-3
rate tolerance, and the surrogate path must beat the bisection it replaces."""
import numpy as np
import pytest

from test_surrogate_host import forward, loss_grad

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sg():
    from smoldyn_amd import surrogate
    return surrogate


def _queries(sg, n, seed):
    X, Y, _ = sg.samples(n, seed=seed)
    ratedt = np.exp(X[:, 0] + 3 * X[:, 1])
    step = np.exp(X[:, 1])
    return ratedt, step, X[:, 2].copy(), step * np.exp(Y[:, 0])


@pytest.mark.parametrize("n", [1, 37, 1000])
def test_device_gradient_matches_numpy_oracle(sg, n):
    rng = np.random.default_rng(n)
    p = (rng.normal(0, 0.3, sg.nparam())).astype(np.float32)
    X = rng.normal(size=(n, 3)).astype(np.float32)
    T = rng.normal(size=(n, 2)).astype(np.float32)
    net = sg.Net(p)
    g, loss = net.gradient(X, T)
    ref_loss, ref_g = loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    scale = np.abs(ref_g).max()
    assert np.abs(g - ref_g).max() <= 2e-5 * scale + 1e-7
    # every parameter block carries gradient (no block left at zero)
    assert np.count_nonzero(g) > 0.95 * len(g)


def test_device_training_reduces_loss(sg):
    X, Y, _ = sg.samples(40000, seed=11)
    norm = np.concatenate([X.mean(0), X.std(0), Y.mean(0), Y.std(0)])
    Xs = ((X - norm[:3]) / norm[3:6]).astype(np.float32)
    Ts = ((Y - norm[6:8]) / norm[8:10]).astype(np.float32)
    net = sg.Net(sg.init_params(0))
    _, l0 = net.gradient(Xs[:4096], Ts[:4096])
    losses = net.train(Xs, Ts, epochs=6, batch=512, lr=3e-3, lr_final=3e-4)
    _, l1 = net.gradient(Xs[:4096], Ts[:4096])
    assert np.all(np.isfinite(losses))
    assert l1 < 0.02 * l0, (l0, l1, losses)


def test_bf16_inference_agrees_with_fp32(sg):
    m = sg.RateSurrogate.load()
    X, _, _ = sg.samples(5000, seed=12)
    Xs = m.standardize(X)
    yb, _ = m.net().predict(Xs, bf16=True)
    yf, _ = m.net().predict(Xs, bf16=False)
    _, _, ynp = forward(m.params.astype(np.float64), Xs)
    assert np.abs(yf - ynp).max() < 1e-4
    # BF16 keeps 8 significant bits: standardized outputs within 0.03
    assert np.abs(yb - yf).max() < 0.03


def test_trained_surrogate_meets_rate_tolerance(sg):
    m = sg.RateSurrogate.load()
    ratedt, step, ratio, a_true = _queries(sg, 4096, seed=13)
    a, ok, _ = m.invert(ratedt, step, ratio, tol=1e-3)
    assert ok.mean() > 0.99, ok.mean()
    err = sg.rate_error(ratedt, step, ratio, a)
    # accepted radii meet the tolerance; fallbacks are the bisection's radii
    assert err[ok].max() <= 1e-3
    assert np.array_equal(a[~ok], sg.invert_numerical(ratedt[~ok], step[~ok], ratio[~ok]))
    # a rate error of 1e-3 moves the radius by about as much where the rate
    # is near the diffusion limit (rate ~ radius)
    rel = np.abs(a / a_true - 1)
    assert np.median(rel) < 1e-5 and rel.max() < 5e-3


def test_surrogate_gain_over_bisection(sg):
    """Inverting a sweep's rates: network + one secant step + judge against
    the reference's bisection, in time and in rate accuracy."""
    import time
    m = sg.RateSurrogate.load()
    ratedt, step, ratio, _ = _queries(sg, 4096, seed=14)
    m.invert(ratedt[:64], step[:64], ratio[:64])  # first call initializes the device
    best_s, best_n = 1e9, 1e9
    for _ in range(3):
        a, ok, t = m.invert(ratedt, step, ratio)
        best_s = min(best_s, t["total_s"])
        t0 = time.perf_counter()
        a_num = sg.invert_numerical(ratedt, step, ratio)
        best_n = min(best_n, time.perf_counter() - t0)
    err_s = sg.rate_error(ratedt, step, ratio, a)
    err_n = sg.rate_error(ratedt, step, ratio, a_num)
    print(f"4096 inversions: surrogate {best_s * 1e3:.2f} ms, bisection {best_n * 1e3:.2f} ms, "
          f"speedup {best_n / best_s:.1f}x; median rate error {np.median(err_s):.2e} vs {np.median(err_n):.2e}")
    assert best_s < best_n
    assert np.median(err_s) <= np.median(err_n)


def test_update_sim_uses_surrogate(sg):
    """smolUpdateSim takes the surrogate's binding radius for a reversible
    reaction with ratio placement, and the radius meets the rate tolerance."""
    import smoldyn_amd as sa
    m = sg.RateSurrogate.load()
    m.install(1e-3)
    try:
        sim = sa.Simulation([0, 0, 0], [20, 20, 20])
        sim.times(0, 1, 0.01)
        sim.seed(3)
        for s, d in (("A", 1.0), ("B", 1.0), ("C", 0.5)):
            sim.species(s)
            sim.mobility(s, d)
        sim.reaction("fwd", ["A", "B"], ["C"], 10.0)
        sim.reaction("back", ["C"], ["A", "B"], 1.0)
        sim.product_placement("back", "ratio", 1.2)
        for d in range(3):
            sim.boundary(d, -1, "p")
        sim.add_molecules("A", 100)
        sim.add_molecules("B", 100)
        sim.update()
        acc, rej = sg.RateSurrogate.stats()
        assert acc >= 1 and rej == 0
    finally:
        sg.RateSurrogate.uninstall()
