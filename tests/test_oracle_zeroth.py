"""Zeroth-order reactions in the oracle (zeroreact, smolreact.c:2905-2993):
events per step are Poisson with mean rate * dt * system volume (rxnsetrate's
order-0 branch, smolreact.c:1325-1330), each makes its product at a uniform
position in the open system box (systemrandpos, smolwall.c:42-46).  The
counter stream draws the count by chunks (smolrng_poisson_chunk); these tests
pin what the law fixes: mean and variance of the count, uniform positions,
fresh serial numbers in creation order, and the birth-death steady state."""
import numpy as np
import pytest

import simspec as ss


def _spec(rate, n0=10, decay=None, seed=5):
    rx = [("make", [], ["A"], rate)]
    if decay:
        rx.append(("decay", ["A"], [], decay))
    return dict(low=[0, 0, 0], high=[20, 30, 40], walls="rrr", dt=0.01, seed=seed,
                species={"A": 1.0}, reactions=rx, molecules=[("A", n0)])


@pytest.mark.parametrize("lam", [0.3, 3.7, 50.0])
def test_count_is_poisson_with_mean_rate_dt_volume(lam):
    vol = 20 * 30 * 40
    osim, _ = ss.build_oracle(_spec(lam / (0.01 * vol)))
    c = []
    for _ in range(2000):
        assert osim.step() == 0
        log = osim.rxnlog()
        c.append(int(np.sum(log[:, 0] == 0)))
    c = np.array(c, dtype=float)
    se = np.sqrt(lam / len(c))
    assert abs(c.mean() - lam) < 5 * se, (c.mean(), lam)
    # variance of a Poisson count equals its mean (sampling sd of the variance ~ sqrt(2/n) lam)
    assert abs(c.var() - lam) < 6 * lam * np.sqrt(2.0 / len(c)) + 0.05, (c.var(), lam)


def test_positions_uniform_and_serials_fresh():
    lam = 20.0
    osim, _ = ss.build_oracle(_spec(lam / (0.01 * 24000), n0=0))
    pos, sers = [], []
    for _ in range(150):
        nxt = osim.L.osim_next_serno(osim.p)
        assert osim.step() == 0
        log = osim.rxnlog()
        z = log[log[:, 0] == 0]
        # products take consecutive new serial numbers in creation order
        assert list(z[:, 4]) == list(range(nxt, nxt + len(z)))
        assert np.all(z[:, 2:4] == 0)
        L = osim.get_list(0)
        new = np.isin(L["serno"], z[:, 4])
        pos.append(L["pos"][new])
        sers.extend(z[:, 4])
    p = np.concatenate(pos)
    assert len(p) == len(sers)
    hi = np.array([20.0, 30.0, 40.0])
    assert np.all(p > 0) and np.all(p < hi)
    u = p / hi  # uniform on (0,1) per axis: Kolmogorov-Smirnov distance
    for d in range(3):
        x = np.sort(u[:, d])
        n = len(x)
        ks = max(np.max(np.arange(1, n + 1) / n - x), np.max(x - np.arange(n) / n))
        assert ks < 1.63 / np.sqrt(n), (d, ks)


def test_birth_death_steady_state():
    """0 -> A at rate k (per volume), A -> 0 at rate kd: the mean count
    relaxes to k V / kd (here 400)."""
    vol = 24000.0
    k, kd = 400 * 0.5 / vol, 0.5
    osim, _ = ss.build_oracle(_spec(k, n0=0, decay=kd))
    for _ in range(1500):  # 7.5 relaxation times
        assert osim.step() == 0
    n = []
    for _ in range(3000):
        assert osim.step() == 0
        n.append(osim.count(osim.index("A")))
    n = np.array(n, dtype=float)
    # samples are correlated over ~1/kd = 200 steps: ~15 independent blocks
    assert abs(n.mean() - 400) < 5 * np.sqrt(400 / 15), n.mean()


def test_reference_stream_refused():
    """Zeroth-order draws are restated for the counter stream only."""
    osim, _ = ss.build_oracle(_spec(0.001), sfmt=True)
    assert osim.step() == 12
