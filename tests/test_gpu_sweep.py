"""The learned-surrogate sweep (batched A+B<->C members whose binding radii
come from the rate surrogate) on the GPU.

Every member's forward rate is inverted by the surrogate inside
smolUpdateSim and judged by the reaction-parameter solver (restated from
libSteve/rxnparam.c, bit-identical to it): an accepted radius meets the
1e-3 rate tolerance, a rejected one is the bisection's.  The batched members
then equal their single runs, and a member equals the oracle built on the
same binding radius (the binding_radius route, RxnSetValue "bindrad",
smolreact.c:1767-1769) step for step, bit for bit."""
import numpy as np
import pytest

import simspec as ss

pytestmark = pytest.mark.gpu

NMEM, NMOL, SIDE, STEPS = 16, 1000, 25.0, 60


@pytest.fixture(scope="module")
def sweep():
    import smoldyn_amd as sa
    from smoldyn_amd import surrogate as sg
    specs = [ss.c5_member(k, n=NMOL, side=SIDE) for k in range(NMEM)]
    model = sg.RateSurrogate.load()
    model.install(1e-3)
    a0, r0 = sg.RateSurrogate.stats()
    try:
        sims = [ss.build_device(sp, update=False) for sp in specs]
        ens = sa.Ensemble(sims)
        # single runs of every member, updated with the same surrogate: their
        # radii are the members' (the inversion is deterministic)
        singles = [ss.build_device(sp) for sp in specs]
        radii = np.array([np.sqrt(s.reaction_params(2, 0)["bindrad2"]) for s in singles])
        acc, rej = sg.RateSurrogate.stats()
    finally:
        sg.RateSurrogate.uninstall()
    return dict(specs=specs, sims=sims, ens=ens, singles=singles, radii=radii, judged=(acc - a0) + (rej - r0),
                sg=sg)


def test_every_sweep_radius_is_judged(sweep):
    sg, specs, a = sweep["sg"], sweep["specs"], sweep["radii"]
    # every member's inversion went through the surrogate and its judge
    # (members and their single runs)
    assert sweep["judged"] == 2 * NMEM
    ratedt = np.array([sp["reactions"][0][3] * sp["dt"] for sp in specs])
    step = np.array([sp["step"] for sp in specs])
    ratio = np.array([sp["ratio"] for sp in specs])
    err = sg.rate_error(ratedt, step, ratio, a)
    bis = sg.invert_numerical(ratedt, step, ratio)
    ok = err <= 1e-3
    # accepted radii meet the tolerance; any other radius is the bisection's
    assert np.all(ok | (a == bis)), (err, a, bis)
    # the radius reproduces the member's target (the rate was made from it)
    target = np.array([sp["target_a"] for sp in specs])
    assert np.max(np.abs(a / target - 1)) < 5e-3


def test_sweep_members_equal_single_runs_and_the_oracle(sweep):
    ens, singles, specs, radii = sweep["ens"], sweep["singles"], sweep["specs"], sweep["radii"]
    ens.run_steps(STEPS)
    for k, sim in enumerate(singles[:4]):
        sim.run_steps(STEPS)
        for ll in range(sim.nlists()):
            a, b = ens.molecules(k, ll), sim.molecules(ll)
            assert np.array_equal(a["serno"], b["serno"]) and np.array_equal(a["ident"], b["ident"])
            assert np.array_equal(a["pos"].view(np.uint64), b["pos"].view(np.uint64))
    # the oracle on the surrogate's radius: forward rate left undefined, the
    # binding radius given directly (members 2 and 3 have radii 1.49 and 0.31:
    # hundreds of events in the window)
    for k in (2, 3):
        osp = dict(specs[k])
        osp["reactions"] = [("fwd", ["A", "B"], ["C"], -1.0), ("back", ["C"], ["A", "B"], 1.0)]
        osp["bindrad"] = {"fwd": float(radii[k])}
        osim, rx = ss.build_oracle(osp)
        assert osim.rxn_info(rx["fwd"])["bindrad2"] == singles[k].reaction_params(2, 0)["bindrad2"]
        nev = 0
        for _ in range(STEPS):
            osim.step()
            nev += len(osim.rxnlog())
        ost = ss.oracle_state(osim)
        for ll in range(len(ost)):
            d = ens.molecules(k, ll)
            assert np.array_equal(ost[ll]["serno"], d["serno"]), (k, ll)
            assert np.array_equal(ost[ll]["pos"].view(np.uint64), d["pos"].view(np.uint64)), (k, ll)
        assert nev > 0
