"""Adsorption onto and desorption from a panel in the oracle, pinned by
closed forms and by the laws of its random draws.

Reference behaviour restated by oracle/osurf.c:
* dosurfinteract's adsorb branch (smolsurface.c:5005-5020): the molecule
  takes the crossing point of its segment with the panel, is nudged onto the
  face it came from (fixpt2panel with the surfaces' epsilon) and takes the
  bound state its rate names; the first molsort moves it to the list of
  that state (smolmolec.c:2943-3023).
* the collision outcome is drawn from the cumulative probabilities of
  rate_internal values (srfcalcprob smolsurface.c:3070-3146, srfdatasrc 2:
  the value is the probability); what is left over reflects.
* the desorb branch (smolsurface.c:5022-5042): a new solution molecule with
  the same serial number, at the bound position nudged to the face it leaves
  to, plus desorbdist(step) along that face's normal (SurfaceParam.c:202-220:
  step * R(y) with y uniform on [0,1] and R the rational fit of reversible
  adsorption when the species also adsorbs, of irreversible desorption
  otherwise); the desorption probability of a rate k is 1 - exp(-k dt).

Segments are injected (zero lateral noise where a closed form needs it), so
every crossing point is known: a molecule at height 50 + h moving by -2h
crosses the plane z = 50 half way.  A wrong crossing fraction, a molecule
left on the wrong face, a wrong list or state, an adsorption probability
applied to the wrong face, a desorption distance on the wrong side of the
panel, with the wrong step length or the wrong fit, each fails an
assertion."""
import numpy as np
import pytest

import simspec as ss

D_A, DT, Z0 = 1.0, 0.01, 50.0
STEP = np.sqrt(2 * D_A * DT)  # rms step per axis of A in solution (difstep)


def _spec(n, p_front, p_back, k_des=0.0, seed=21):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(10, 90, (n, 2))
    side = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)  # alternate above and below the plane
    h = 0.05
    pos = np.column_stack([xy, Z0 + side * h])
    rates = [("A", "soln", "fsoln", "front", p_front, True), ("A", "soln", "bsoln", "back", p_back, True)]
    if k_des > 0:
        rates += [("A", "front", "front", "fsoln", k_des, False), ("A", "back", "back", "bsoln", k_des, False)]
    spec = dict(low=[0, 0, 0], high=[100, 100, 100], walls="rrr", dt=DT, seed=seed,
                species={"A": D_A},
                surfaces=[{"name": "mem", "actions": [], "rates": rates,
                           "panels": [("rect", "+2", [0.0, 0.0, Z0, 100.0, 100.0], "p0")]}],
                molecules=[("A", pos)])
    return spec, h


def _state(osim):
    out = {}
    for ll in range(osim.nlist()):
        L = osim.get_list(ll)
        for j in range(len(L["ident"])):
            out[int(L["serno"][j])] = dict(list=ll, ms=int(L["mstate"][j]), pos=L["pos"][j].copy(),
                                           ident=int(L["ident"][j]))
    return out


def _cross_once(spec, h, lateral=True, seed=3):
    """One step in which every molecule moves by -2h along z toward the
    plane (so it reaches the mirror height on the other side) plus a lateral
    step; returns the oracle, the states before and after, and the injected
    displacement of each serial."""
    osim, _ = ss.build_oracle(spec)
    before = _state(osim)
    rng = np.random.default_rng(seed)
    L = osim.get_list(0)  # diffuse order: the solution list in live order
    g = np.zeros((len(L["ident"]), 3))
    disp = {}
    for j in range(len(L["ident"])):
        s = int(L["serno"][j])
        up = L["pos"][j][2] > Z0
        if lateral:
            g[j, :2] = rng.normal(size=2)
        g[j, 2] = (-2 * h if up else 2 * h) / STEP
        disp[s] = STEP * g[j]
    osim.inject(g)
    assert osim.step() == 0
    return osim, before, _state(osim), disp


def test_adsorbed_molecule_sits_at_the_crossing_point_on_its_face():
    spec, h = _spec(400, 1.0, 1.0)
    osim, before, after, disp = _cross_once(spec, h)
    lists = set()
    for s, b in before.items():
        a = after[s]
        up = b["pos"][2] > Z0
        assert a["ms"] == (1 if up else 2)  # front state from above, back state from below
        lists.add((a["ms"], a["list"]))
        assert a["list"] != b["list"]  # moved to the list of its bound state
        half = b["pos"][:2] + 0.5 * disp[s][:2]  # the segment crosses z = 50 half way
        np.testing.assert_allclose(a["pos"][:2], half, rtol=0, atol=1e-12)
        dz = a["pos"][2] - Z0
        assert abs(dz) <= 1e-12 and (dz >= 0 if up else dz <= 0)  # on the face it came from
    assert len({l for _, l in lists}) == 1  # both bound states share the default list of immobile states
    assert osim.count(1) == 400


@pytest.mark.parametrize("p_front,p_back", [(0.3, 0.7), (0.05, 0.5)])
def test_adsorption_fraction_is_the_face_probability_and_the_rest_reflect(p_front, p_back):
    n = 6000
    spec, h = _spec(n, p_front, p_back, seed=23)
    osim, before, after, disp = _cross_once(spec, h, seed=4)
    nf = nb = kf = kb = 0
    for s, b in before.items():
        a = after[s]
        up = b["pos"][2] > Z0
        nf += up
        nb += not up
        if a["ms"] != 0:
            kf += up
            kb += not up
            continue
        # not adsorbed: reflected off the plane back to its own side, the
        # lateral motion kept (surfacereflect mirrors the normal component)
        end = b["pos"] + disp[s]
        np.testing.assert_allclose(a["pos"][:2], end[:2], rtol=0, atol=1e-12)
        assert abs(a["pos"][2] - (2 * Z0 - end[2])) < 1e-12
    for k, m, p in ((kf, nf, p_front), (kb, nb, p_back)):
        sd = np.sqrt(m * p * (1 - p))
        assert abs(k - m * p) < 4 * sd, (k, m, p)


def _rev_fit(y):
    return (0.729614 * y - 0.70252 * y * y) / (1.0 - 1.47494 * y + 0.484371 * y * y)


def _rev_fit_inverse(r):
    """y from R(y) = (0.729614 y - 0.70252 y^2) / (1 - 1.47494 y + 0.484371 y^2)
    (SurfaceParam.c:212, reversible adsorption), R increasing on [0, 1]."""
    a2 = -0.70252 - 0.484371 * r
    a1 = 0.729614 + 1.47494 * r
    a0 = -r
    disc = np.sqrt(a1 * a1 - 4 * a2 * a0)
    return (-a1 + disc) / (2 * a2)


def test_desorbed_molecule_leaves_along_the_normal_by_the_desorption_distance():
    n = 4000
    spec, h = _spec(n, 1.0, 1.0, k_des=1e6, seed=29)
    osim, before, bound, _ = _cross_once(spec, h, seed=5)
    assert all(v["ms"] != 0 for v in bound.values())
    L = osim.get_list(0)
    osim.inject(np.zeros((max(len(L["ident"]), 1), 3)))
    assert osim.step() == 0  # probability 1 - exp(-1e6 * 0.01) = 1: every bound molecule desorbs
    after = _state(osim)
    assert set(after) == set(bound)  # the new solution molecule keeps the serial number
    ys = []
    for s, b in bound.items():
        a = after[s]
        assert a["ms"] == 0 and a["ident"] == 1
        front = b["ms"] == 1
        assert np.array_equal(a["pos"][:2], b["pos"][:2])  # the plane's normal is the z axis
        x = (a["pos"][2] - Z0) * (1 if front else -1)  # distance off the face it leaves to
        assert 0 <= x <= STEP * _rev_fit(1.0) * (1 + 1e-12)
        ys.append(_rev_fit_inverse(x / STEP))
    ys = np.sort(np.array(ys))
    # y is uniform on [0, 1]: Kolmogorov-Smirnov at the 0.1 % level
    cdf = np.arange(1, n + 1) / n
    dks = max(np.max(cdf - ys), np.max(ys - (cdf - 1.0 / n)))
    assert dks < 1.95 / np.sqrt(n), dks


def test_desorption_probability_is_one_minus_exp():
    """Bound molecules desorbing at rate k leave with probability
    1 - exp(-k dt) per step (SurfaceParam.c:159-163)."""
    n, k = 6000, 30.0
    spec, h = _spec(n, 1.0, 1.0, k_des=k, seed=31)
    osim, _, bound, _ = _cross_once(spec, h, seed=6)
    osim.inject(np.zeros((1, 3)))
    assert osim.step() == 0
    after = _state(osim)
    left = sum(1 for s in bound if after[s]["ms"] == 0)
    p = 1 - np.exp(-k * DT)
    assert abs(left - n * p) < 4 * np.sqrt(n * p * (1 - p)), (left, n * p)


def _sp(k1, k2, algo):
    """The oracle's surfaceprob (pinned to the reference's) for A in solution."""
    import ctypes as C
    from oracle_bind import lib
    p2 = C.c_double(0)
    return lib().osurf_surfaceprob(k1, k2, DT, D_A, C.byref(p2), algo)


def _rate_spec(n, k_ads, k_des, seed):
    """Rates (`rate`, not `rate_internal`): adsorption from the front is
    reversible (the front state desorbs back to the front at k_des),
    adsorption from the back is irreversible (nothing leaves the back state)."""
    spec, h = _spec(n, 0.0, 0.0, seed=seed)
    spec["surfaces"][0]["rates"] = [("A", "soln", "fsoln", "front", k_ads, False),
                                    ("A", "soln", "bsoln", "back", k_ads, False),
                                    ("A", "front", "front", "fsoln", k_des, False)]
    return spec, h


def test_rate_given_adsorption_picks_reversible_or_irreversible_lookup():
    """srfcalcprob (smolsurface.c:3122-3124): a solution molecule adsorbs with
    the reversible-adsorption probability (SPArevAds, lookuprevads) when its
    bound state desorbs back to that face, else with the irreversible one
    (SPAirrAds, lookupirrevadsorb); kappa = k dt / sqrt(2 D dt)."""
    k_ads, k_des, n = 3.0, 100.0, 12000
    p_rev, p_irr = _sp(k_ads, k_des, 8), _sp(k_ads, 0.0, 4)
    assert abs(p_rev - 0.2885) < 1e-3 and abs(p_irr - 0.4107) < 1e-3  # far apart at these rates
    spec, h = _rate_spec(n, k_ads, k_des, seed=37)
    osim, before, after, _ = _cross_once(spec, h, seed=8)
    nf = nb = kf = kb = 0
    for s, b in before.items():
        up = b["pos"][2] > Z0
        nf += up
        nb += not up
        kf += up and after[s]["ms"] == 1
        kb += (not up) and after[s]["ms"] == 2
    for k, m, p in ((kf, nf, p_rev), (kb, nb, p_irr)):
        assert abs(k - m * p) < 4 * np.sqrt(m * p * (1 - p)), (k, m, p)


def test_rate_given_reversible_desorption_probability():
    """A bound state that desorbs to the face its adsorption came from leaves
    with SPArevDes's probability (the reversible problem read backwards:
    surfaceprob(k_des, k_ads), SurfaceParam.c:165-174), not 1 - exp(-k dt)."""
    k_ads, k_des, n = 3.0, 100.0, 12000
    p_des = _sp(k_des, k_ads, 11)
    assert abs(p_des - 0.5425) < 1e-3 and abs((1 - np.exp(-k_des * DT)) - p_des) > 0.08
    spec, h = _rate_spec(n, k_ads, k_des, seed=39)
    osim, _, bound, _ = _cross_once(spec, h, seed=9)
    osim.inject(np.zeros((max(len(osim.get_list(0)["ident"]), 1), 3)))  # solution molecules stay put
    assert osim.step() == 0
    after = _state(osim)
    front = [s for s in bound if bound[s]["ms"] == 1]
    back = [s for s in bound if bound[s]["ms"] == 2]
    left = sum(1 for s in front if after[s]["ms"] == 0)
    m = len(front)
    assert abs(left - m * p_des) < 4 * np.sqrt(m * p_des * (1 - p_des)), (left, m, p_des)
    assert all(after[s]["ms"] == 2 for s in back)  # the back state has no exit


def test_flip_between_bound_states():
    """A bound molecule flips to the other face's state with the two-state
    relaxation probability k1/(k1+k2) (1 - exp(-(k1+k2) dt)) when both flips
    are given (SPArevFlip, SurfaceParam.c:186-196; srfcalcprob
    smolsurface.c:3131-3134) and moves to that face of its panel
    (molchangeident + fixpt2panel, smolsurface.c:5044-5048): same place in
    the plane, on the other side of z = 50 by at most a few ulps."""
    n, k_fb, k_bf = 8000, 40.0, 10.0
    spec, h = _spec(n, 1.0, 1.0, seed=41)
    spec["surfaces"][0]["rates"] += [("A", "front", "front", "back", k_fb, False),
                                     ("A", "back", "back", "front", k_bf, False)]
    osim, _, bound, _ = _cross_once(spec, h, seed=10)
    assert all(v["ms"] != 0 for v in bound.values())
    osim.inject(np.zeros((1, 3)))
    assert osim.step() == 0
    after = _state(osim)
    kt = k_fb + k_bf
    for src, dst, k in ((1, 2, k_fb), (2, 1, k_bf)):
        mols = [s for s in bound if bound[s]["ms"] == src]
        flipped = [s for s in mols if after[s]["ms"] == dst]
        p = k / kt * (1 - np.exp(-kt * DT))
        m = len(mols)
        assert abs(len(flipped) - m * p) < 4 * np.sqrt(m * p * (1 - p)), (src, len(flipped), m * p)
        for s in flipped:
            a, b = after[s]["pos"], bound[s]["pos"]
            assert np.array_equal(a[:2], b[:2])
            dz = a[2] - Z0
            assert (dz > 0 if dst == 1 else dz < 0) and abs(dz) < 1e-12
            assert after[s]["list"] == bound[s]["list"]  # both bound states share a list
        for s in set(mols) - set(flipped):
            assert np.array_equal(after[s]["pos"], bound[s]["pos"]) and after[s]["ms"] == src
