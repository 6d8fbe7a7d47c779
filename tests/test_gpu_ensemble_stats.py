"""Statistics of batched ensembles on the device.

Replicates of one model run as ensemble members with different seeds (counter
stream).  Their species counts are compared, at the 3-sigma level of the
difference of means, with
  - replicates of the same model in the oracle driven by the reference's own
    random stream (SFMT-19937 in program order), i.e. the reference's
    ensemble; and
  - the equilibrium that the requested rate constants predict, k_f A B / V =
    k_b C, where k_f is the rate the reaction-parameter solver turned into
    the binding radius (the reference's rxnparam-predicted rate)."""
import numpy as np
import pytest

import simspec as ss

pytestmark = pytest.mark.gpu


def _spec(seed, n=300, side=20.0, kf=None, kb=None, dt=0.01):
    reactions = [("fwd", ["A", "B"], ["C"], kf if kf else 2.0)]
    placement = {}
    if kb:
        reactions.append(("back", ["C"], ["A", "B"], kb))
        placement = {"back": ("ratio", 1.5)}
    return dict(low=[0.0] * 3, high=[side] * 3, walls="rrr", dt=dt, seed=seed,
                species={"A": 1.0, "B": 1.0, "C": 0.5}, reactions=reactions, placement=placement,
                molecules=[("A", n // 2), ("B", n // 2)])


def _ensemble_counts(specs, checkpoints):
    import smoldyn_amd as sa
    members = [ss.build_device(sp, update=False) for sp in specs]
    ens = sa.Ensemble(members)
    out = np.zeros((len(checkpoints), len(specs)))
    done = 0
    for j, t in enumerate(checkpoints):
        ens.run_steps(t - done)
        done = t
        out[j] = [ens.count(k, "C") for k in range(len(specs))]
    return out


def _within_3sigma(a, b):
    ma, mb = a.mean(), b.mean()
    se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    return abs(ma - mb) <= 3 * se, ma, mb, se


def test_ensemble_counts_match_reference_stream_ensemble():
    R = 48
    checkpoints = [25, 100, 200]
    gpu = _ensemble_counts([_spec(1000 + k) for k in range(R)], checkpoints)
    ref = np.zeros_like(gpu)
    for k in range(R):
        osim, _ = ss.build_oracle(_spec(5000 + k), sfmt=True)
        done = 0
        for j, t in enumerate(checkpoints):
            for _ in range(t - done):
                osim.step()
            done = t
            ref[j, k] = osim.count(osim.index("C"))
    for j, t in enumerate(checkpoints):
        ok, mg, mr, se = _within_3sigma(gpu[j], ref[j])
        assert ok, "step %d: device mean C %.2f vs reference-stream mean %.2f (3 sigma = %.2f)" % (t, mg, mr, 3 * se)
    assert gpu[-1].mean() > 5  # the reaction ran


def test_ensemble_equilibrium_matches_rate_constants():
    """Reversible A+B<->C: time-averaged equilibrium C per member vs the
    mass-action equilibrium of the requested rates."""
    R, n, side, kf, kb = 32, 800, 20.0, 4.0, 2.0
    V = side ** 3
    a0 = n // 2
    # kf (a0 - c)^2 / V = kb c
    p = kb * V / kf
    c_eq = ((2 * a0 + p) - np.sqrt((2 * a0 + p) ** 2 - 4 * a0 * a0)) / 2
    import smoldyn_amd as sa
    members = [ss.build_device(_spec(7000 + k, n=n, side=side, kf=kf, kb=kb), update=False) for k in range(R)]
    ens = sa.Ensemble(members)
    ens.run_steps(400)  # several relaxation times 1/(kb + kf (A+B)/V)
    acc = np.zeros(R)
    nsamp = 60
    for _ in range(nsamp):
        ens.run_steps(10)
        acc += [ens.count(k, "C") for k in range(R)]
    means = acc / nsamp
    se = means.std(ddof=1) / np.sqrt(R)
    assert abs(means.mean() - c_eq) <= 3 * se + 1e-9, "equilibrium C %.2f vs predicted %.2f (3 sigma %.2f)" % (
        means.mean(), c_eq, 3 * se)
