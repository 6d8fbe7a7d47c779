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


def _moving_pair(rparam, p1, p2, g1, g2, da=1.0, db=0.25, dt=0.01):
    """One A-B pair moved by injected deviates g1, g2 (units of each
    molecule's rms step) in one step; returns the binding radius, the event
    log and each serial's (species, position) before and after."""
    spec = dict(low=[0, 0, 0], high=[20, 20, 20], walls="rrr", dt=dt, seed=3,
                species={"A": da, "B": db},
                reactions=[("hit", ["A", "B"], ["A", "B"], 50.0)],
                placement={"hit": ("bounce", rparam)},
                molecules=[("A", [p1]), ("B", [p2])])
    osim, rx = ss.build_oracle(spec)
    a = np.sqrt(osim.rxn_info(rx["hit"])["bindrad2"])
    L = osim.get_list(0)
    before = {int(L["serno"][j]): (int(L["ident"][j]), L["pos"][j].copy()) for j in range(len(L["ident"]))}
    g = np.array([g1 if L["ident"][j] == 1 else g2 for j in range(len(L["ident"]))], dtype=np.float64)
    osim.inject(g)
    assert osim.step() == 0
    L = osim.get_list(0)
    after = {int(L["serno"][j]): (int(L["ident"][j]), L["pos"][j].copy()) for j in range(len(L["ident"]))}
    return a, osim.rxnlog(), before, after


def test_bounce_reflection_method_reflects_each_path_at_contact():
    """Bounce by reflection (rparam -2, doreact smolreact.c:2756-2762,
    Geo_SphereReflectSphere Geometry.c:1264-1293): the pair moves on straight
    lines from its places before diffusion to its places after; at the first
    moment p the separation equals the binding radius, and each molecule's
    remaining motion (1 - p) of its displacement is mirrored in the plane normal
    to the separation at contact, reversed for a molecule moving away."""
    dt, da, db = 0.01, 1.0, 0.25
    sa_, sb_ = np.sqrt(2 * da * dt), np.sqrt(2 * db * dt)
    p1 = np.array([10.0, 10.0, 10.0])
    a, _, _, _ = _moving_pair(-2.0, p1.tolist(), [12.0, 10.0, 10.0], [0, 0, 0], [0, 0, 0])
    p2 = p1 + np.array([a + 0.06, 0.03, -0.02])
    g1, g2 = np.array([1.6, 0.25, 0.1]), np.array([-1.2, -0.3, 0.4])
    a, log, before, after = _moving_pair(-2.0, p1.tolist(), p2.tolist(), g1, g2)
    assert len(log) == 1 and log[0][0] == 2
    sA = next(s for s, v in before.items() if v[0] == 1)
    sB = next(s for s, v in before.items() if v[0] == 2)
    a0, b0 = before[sA][1], before[sB][1]
    a1, b1 = a0 + sa_ * g1, b0 + sb_ * g2
    assert np.linalg.norm(b0 - a0) > a > np.linalg.norm(b1 - a1)  # the pair comes into contact in the step
    # first contact: |(b0 - a0) + p ((b1 - a1) - (b0 - a0))| = a, the smaller root
    s0, ds = b0 - a0, (b1 - a1) - (b0 - a0)
    p = min(r.real for r in np.roots([ds @ ds, 2 * s0 @ ds, s0 @ s0 - a * a]) if abs(r.imag) < 1e-12)
    assert 0 < p < 1
    ap, bp = a0 + p * (a1 - a0), b0 + p * (b1 - b0)
    n = (bp - ap) / np.linalg.norm(bp - ap)
    assert abs(np.linalg.norm(bp - ap) - a) < 1e-12

    def mirrored(u, toward):
        m = u - 2 * (u @ n) * n
        return m if toward else -m

    uA, uB = a1 - a0, b1 - b0
    expA = ap + (1 - p) * mirrored(uA, uA @ n >= 0)
    expB = bp + (1 - p) * mirrored(uB, uB @ n <= 0)
    # the products keep the reactants' serial numbers and species
    assert after[sA][0] == 1 and after[sB][0] == 2
    assert np.abs(after[sA][1] - expA).max() < 1e-12
    assert np.abs(after[sB][1] - expB).max() < 1e-12
    # each molecule's path after contact keeps its length
    assert abs(np.linalg.norm(after[sA][1] - ap) - (1 - p) * np.linalg.norm(uA)) < 1e-12


def test_bounce_reflection_without_motion_falls_back():
    """A pair that did not move gives no contact time (the quadratic has no
    motion term), and the reference falls back to reflection off the binding
    radius: the new separation is 2a - d."""
    p1, p2 = [10.0, 10.0, 10.0], [10.2, 10.1, 9.9]
    a, log, before, after = _moving_pair(-2.0, p1, p2, [0, 0, 0], [0, 0, 0])
    sA = next(s for s, v in before.items() if v[0] == 1)
    sB = next(s for s, v in before.items() if v[0] == 2)
    d = np.linalg.norm(np.subtract(p2, p1))
    assert len(log) == 1 and d < a
    assert abs(np.linalg.norm(after[sB][1] - after[sA][1]) - (2 * a - d)) < 1e-12
