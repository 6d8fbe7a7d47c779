"""Product placement of first-order reactions C -> A + B in the oracle
(rxnsetproduct smolreact.c:1504-1573, doreact :2806-2860).

The reference places the products of a dissociation at the reactant's
position plus prdpos[prd] turned by one random rotation per event:
prdpos[0] = +u*dcA, prdpos[1] = -u*dcB along x with dcA = D_A/(D_A+D_B),
dcB = D_B/(D_A+D_B), and the unbinding separation u is
  ratio      u = rparam * a   (a: binding radius of the reverse A + B -> C)
  unbindrad  u = rparam
  pgem       u = unbindingradius(rparam, dt, D_A + D_B, a)
  pgemmax    u = unbindingradius(rparam, dt, D_A + D_B, a) when positive
and fixed placement puts product k at the reactant's position plus the
given vector, unrotated.  These are checked on single C molecules that do
not move (zero injected displacements) and dissociate in the first step,
so each product's offset from its reactant is observed directly.  The
unbindingradius values come from the oracle's rxnparam, which is pinned
bit for bit against the reference's compiled rxnparam.c
(tests/test_oracle_ref.py) and its golden fixture.  A dropped weight, a
swapped pair of weights, a sign error (products on the same side) or a
wrong separation fails one of the assertions."""
import numpy as np
import pytest

import oracle_bind as ob
import simspec as ss

DA, DB = 1.0, 0.25
DT = 0.01


def _dissociate(placement, ncs=40, seed=5):
    """ncs C molecules far apart; C -> A + B with probability 1 in the first
    step; returns (binding radius a, {C serial: C position}, rows of the
    reaction log, {serial: (species, position)} after the step)."""
    rng = np.random.default_rng(seed)
    cs = np.column_stack([np.arange(ncs) * 20.0 + 10.0, rng.uniform(20, 80, ncs), rng.uniform(20, 80, ncs)])
    spec = dict(low=[0, 0, 0], high=[ncs * 20.0 + 20.0, 100, 100], walls="rrr", dt=DT, seed=seed,
                species={"A": DA, "B": DB, "C": 0.5},
                reactions=[("fwd", ["A", "B"], ["C"], 20.0), ("back", ["C"], ["A", "B"], 1e6)],
                placement={"back": placement},
                molecules=[("C", cs)])
    osim, rx = ss.build_oracle(spec)
    a = float(np.sqrt(osim.rxn_info(rx["fwd"])["bindrad2"]))
    before = {}
    for ll in range(osim.nlist()):
        L = osim.get_list(ll)
        for j in range(len(L["ident"])):
            before[int(L["serno"][j])] = L["pos"][j].copy()
    osim.inject(np.zeros((ncs, 3)))
    assert osim.step() == 0
    log = osim.rxnlog()
    after = {}
    for ll in range(osim.nlist()):
        L = osim.get_list(ll)
        for j in range(len(L["ident"])):
            after[int(L["serno"][j])] = (int(L["ident"][j]), L["pos"][j].copy())
    assert len(log) == ncs and all(int(r[0]) == 1 for r in log)  # every C dissociated
    return a, before, log, after, osim, rx


def _offsets(before, log, after):
    """(offset of product A, offset of product B) from the reactant, per event."""
    va, vb = [], []
    for r in log:
        c = int(r[2])
        pa, pb = int(r[4]), int(r[5])
        assert after[pa][0] == 1 and after[pb][0] == 2  # products in listed order: A then B
        va.append(after[pa][1] - before[c])
        vb.append(after[pb][1] - before[c])
    return np.array(va), np.array(vb)


def _check_split(va, vb, u):
    la, lb = np.linalg.norm(va, axis=1), np.linalg.norm(vb, axis=1)
    dca, dcb = DA / (DA + DB), DB / (DA + DB)
    # separation and the split in proportion to each product's diffusion coefficient
    np.testing.assert_allclose(np.linalg.norm(va - vb, axis=1), u, rtol=1e-12)
    np.testing.assert_allclose(la, u * dca, rtol=1e-12)
    np.testing.assert_allclose(lb, u * dcb, rtol=1e-12)
    # opposite sides of the reactant: vb = -(dcB/dcA) va
    np.testing.assert_allclose(vb, -(dcb / dca) * va, rtol=0, atol=1e-12 * u)


def test_ratio_placement_separation_is_ratio_times_binding_radius():
    a, before, log, after, osim, rx = _dissociate(("ratio", 1.2))
    va, vb = _offsets(before, log, after)
    _check_split(va, vb, 1.2 * a)
    assert abs(osim.rxn_info(rx["back"])["unbindrad"] - 1.2 * a) < 1e-15


def test_unbindrad_placement_separation_is_the_parameter():
    a, before, log, after, _, _ = _dissociate(("unbindrad", 3.7))
    va, vb = _offsets(before, log, after)
    _check_split(va, vb, 3.7)


@pytest.mark.parametrize("kind,pgem", [("pgem", 0.2), ("pgem", 0.5), ("pgemmax", 0.2)])
def test_pgem_placement_separation_is_the_unbinding_radius(kind, pgem):
    a, before, log, after, _, _ = _dissociate((kind, pgem))
    u = ob.lib().orxn_unbindingradius(pgem, DT, DA + DB, a)
    assert u > a  # geminate probability below its maximum: products start outside the binding radius
    va, vb = _offsets(before, log, after)
    _check_split(va, vb, u)


def test_product_directions_are_isotropic():
    """The random rotation leaves the products' axis uniform on the sphere:
    mean of the unit vector 0 (sd 1/sqrt(3n) per axis) and mean of its
    squared components 1/3 (sd sqrt(4/45/n))."""
    n = 3000
    a, before, log, after, _, _ = _dissociate(("ratio", 1.2), ncs=n, seed=9)
    va, _ = _offsets(before, log, after)
    uvec = va / np.linalg.norm(va, axis=1)[:, None]
    assert np.all(np.abs(uvec.mean(axis=0)) < 4.0 / np.sqrt(3 * n))
    assert np.all(np.abs((uvec ** 2).mean(axis=0) - 1.0 / 3.0) < 4.0 * np.sqrt(4.0 / 45.0 / n))


def test_fixed_placement_adds_the_vectors_unrotated():
    vecs = [("A", [0.3, -0.2, 0.1]), ("B", [-0.4, 0.5, 0.25])]
    a, before, log, after, _, _ = _dissociate(("fixed", 0.0, vecs))
    for r in log:
        c = int(r[2])
        for k, (sp, v) in enumerate(vecs):
            got = after[int(r[4 + k])][1]
            want = before[c] + np.array(v)
            assert np.array_equal(got, want)  # posx + v in FP64, bit for bit
