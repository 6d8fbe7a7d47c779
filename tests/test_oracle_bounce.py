"""Bounce product placement in the oracle (rxnsetproduct smolreact.c:1470-1494,
doreact :2753-2805, scaling methods): a reacting pair keeps its serial
numbers and moves apart along the line joining the two molecules, each
displacement in proportion to that molecule's share of the summed diffusion
coefficient; with rparam >= 0 the new separation is rparam, with rparam < 0
it is 2a - d (the pair reflected off the binding radius a).  Pinned by these
closed forms on a single pair that does not diffuse (zero injected
displacements)."""
import numpy as np
import pytest

import simspec as ss


def _pair(rparam, p1, p2, da=1.0, db=0.25):
    spec = dict(low=[0, 0, 0], high=[20, 20, 20], walls="rrr", dt=0.01, seed=3,
                species={"A": da, "B": db},
                reactions=[("hit", ["A", "B"], ["A", "B"], 50.0)],
                placement={"hit": ("bounce", rparam)},
                molecules=[("A", [p1]), ("B", [p2])])
    osim, rx = ss.build_oracle(spec)
    a = np.sqrt(osim.rxn_info(rx["hit"])["bindrad2"])
    before = {}
    for ll in range(osim.nlist()):
        L = osim.get_list(ll)
        for j in range(len(L["ident"])):
            before[int(L["serno"][j])] = L["pos"][j].copy()
    osim.inject(np.zeros((2, 3)))
    assert osim.step() == 0
    log = osim.rxnlog()
    after = {}
    for ll in range(osim.nlist()):
        L = osim.get_list(ll)
        for j in range(len(L["ident"])):
            after[int(L["serno"][j])] = (int(L["ident"][j]), L["pos"][j].copy())
    return a, log, before, after


@pytest.mark.parametrize("rparam", [1.1, 0.7])
def test_bounce_to_given_separation(rparam):
    p1, p2 = [10.0, 10.0, 10.0], [10.12, 9.95, 10.08]
    a, log, before, after = _pair(rparam, p1, p2)
    assert np.linalg.norm(np.subtract(p2, p1)) < a  # the pair is within the binding radius
    assert len(log) == 1 and log[0][0] == 2
    sA, sB = sorted(before)
    # products keep the reactants' serial numbers and species
    assert set(after) == {sA, sB} and after[sA][0] == 1 and after[sB][0] == 2
    qa, qb = after[sA][1], after[sB][1]
    v = np.subtract(p2, p1)
    assert abs(np.linalg.norm(qb - qa) - rparam) < 1e-12
    assert np.linalg.norm(np.cross(qb - qa, v)) < 1e-12  # along the original line
    da, db = 1.0, 0.25
    ma, mb = np.linalg.norm(qa - p1), np.linalg.norm(qb - p2)
    assert abs(ma / mb - da / db) < 1e-9  # displacements in the ratio of the diffusion coefficients


def test_bounce_reflects_off_binding_radius():
    p1, p2 = [10.0, 10.0, 10.0], [10.2, 10.1, 9.9]
    a, log, before, after = _pair(-1.0, p1, p2)
    sA, sB = sorted(before)
    d = np.linalg.norm(np.subtract(p2, p1))
    assert d < a
    qa, qb = after[sA][1], after[sB][1]
    assert abs(np.linalg.norm(qb - qa) - (2 * a - d)) < 1e-12
