"""Molecule lists that fill up grow, as the reference's lists do
(molexpandlist, smolmolec.c:1911-1950).

A birth process (A -> A + A) with predation (A + B -> B) grows A well past
the capacity its lists get at setup (the molecules at setup plus 4096).
smolRunTimeStep and smolRunSim rebuild the device engine with twice the
capacity when a list is more than half full; the trajectory stays the
oracle's bit for bit across the rebuilds (positions, live order, serials,
boxes and reaction events)."""
import numpy as np
import pytest

import smoldyn_amd as sa
import simspec as ss

pytestmark = pytest.mark.gpu


def _steps_to(tmin, tmax, dt):
    """Steps smolRunSim takes from tmin to tmax (the clock accumulates dt)."""
    t, n = tmin, 0
    while True:
        t += dt
        n += 1
        if t >= tmax:
            return n


def _birth_spec(nsteps, seed=91):
    dt = 0.01
    return dict(low=[0, 0, 0], high=[40, 40, 40], walls="ppp", dt=dt, seed=seed, tmax=nsteps * dt,
                species={"A": 1.0, "B": 1.0},
                reactions=[("birth", ["A"], ["A", "A"], 2.0), ("eat", ["A", "B"], ["B"], 30.0)],
                molecules=[("A", 400), ("B", 200)])


def test_lists_grow_through_single_steps():
    nsteps = 160
    spec = _birth_spec(nsteps)
    osim, _ = ss.build_oracle(spec)
    dsim = ss.build_device(spec)
    cap0 = 600 + 4096
    for k in range(nsteps):
        osim.step()
        dsim.run_time_step()
        assert ss.canonical_pairs(osim.rxnlog()[:, :6]) == ss.canonical_pairs(dsim.reaction_log()), "step %d" % k
        if k % 20 == 19:
            ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))
    ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))
    d = dsim.diagnostics()
    print("molecules %d, rebuilds %d" % (d["nmol"], d["rebuilds"]), flush=True)
    assert d["nmol"] > cap0 and d["rebuilds"] >= 1
    assert dsim.count("A") == osim.count(osim.index("A"))


def test_lists_grow_inside_smolrunsim():
    nsteps = 160
    spec = _birth_spec(nsteps, seed=93)
    osim, _ = ss.build_oracle(spec)
    dsim = ss.build_device(spec)
    assert dsim.run() == sa.ECok
    nsteps = _steps_to(0.0, spec["tmax"], spec["dt"])
    for _ in range(nsteps):
        osim.step()
    ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))
    d = dsim.diagnostics()
    print("molecules %d, rebuilds %d, step %d" % (d["nmol"], d["rebuilds"], d["step"]), flush=True)
    assert d["step"] == nsteps and d["rebuilds"] >= 1 and d["nmol"] > 600 + 4096
