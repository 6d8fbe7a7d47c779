"""Parity of the HIP engine with the CPU oracle, through the C-ABI.

Bit-exact contract: same FP64 positions, same live-list order and serial
numbers, same virtual-box indices and the same reaction events (reactant and
product serials) as the oracle, which follows the reference's operation order
(and reproduces the reference's own regression outputs)."""
import numpy as np
import pytest

import simspec as ss

pytestmark = pytest.mark.gpu


def _run_pair(spec, nsteps, check_every=1, inject=None):
    osim, _ = ss.build_oracle(spec)
    dsim = ss.build_device(spec)
    ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim), check_box=False)
    rng = np.random.default_rng(1234)
    nev = 0
    for k in range(nsteps):
        if inject:
            n = sum(len(x["ident"]) for x in ss.oracle_state(osim))
            g = rng.standard_normal((n, 3))
            osim.inject(g)
            osim.step()
            dsim.inject_step(g)
        else:
            osim.step()
            dsim.run_time_step()
        lo = ss.canonical_pairs(osim.rxnlog()[:, :6])
        ld = ss.canonical_pairs(dsim.reaction_log())
        assert lo == ld, "step %d reaction events differ (%d vs %d)" % (k, len(lo), len(ld))
        nev += len(lo)
        if (k + 1) % check_every == 0 or k == nsteps - 1:
            ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))
            _compare_counts(spec, osim, dsim)
    return nev


def _compare_counts(spec, osim, dsim):
    """Species counts kept by the device's counters equal the oracle's."""
    import smoldyn_amd as sa
    for name in spec["species"]:
        assert dsim.count(name, sa.MSall) == osim.count(osim.index(name)), name


def test_c1_injected_displacements_one_step():
    """C1 (A+B->C, reflective cube) at 2k molecules, injected Gaussian deviates."""
    spec = ss.c1_spec(n=1000)
    _run_pair(spec, 1, inject=True)


def test_c1_injected_multi_step_reactions():
    spec = ss.c1_spec(n=2000, seed=3)
    nev = _run_pair(spec, 20, inject=True)
    assert nev > 0


def test_c1_counter_stream_trajectory():
    spec = ss.c1_spec(n=5000, seed=5)
    nev = _run_pair(spec, 60, check_every=10)
    assert nev > 10


def test_c2_reversible_periodic_small():
    """C2 model (reversible A+B<->C, periodic, ratio 1.2 unbinding) scaled to 20k."""
    spec = ss.c2_spec(n=10000, side=1000.0 * (20000 / 1e6) ** (1 / 3), kback=20.0, seed=9)
    nev = _run_pair(spec, 80, check_every=10)
    assert nev > 10


def test_two_lists_first_order_products_and_wrap():
    """Lotka-Volterra-like model: two molecule lists, first-order reactions with
    two products at one point, bimolecular reactions across lists and across
    the periodic boundary (neighbor wrap codes)."""
    spec = dict(low=[-100, -100, -10], high=[100, 100, 10], walls="ppp", dt=0.001, seed=21,
                species={"rabbit": 100.0, "fox": 100.0},
                lists={"rlist": ["rabbit"], "flist": ["fox"]},
                reactions=[("r1", ["rabbit"], ["rabbit", "rabbit"], 10.0),
                           ("r2", ["rabbit", "fox"], ["fox", "fox"], 8000.0),
                           ("r3", ["fox"], [], 10.0)],
                molperbox=1.0,
                molecules=[("rabbit", 1000), ("fox", 1000)])
    nev = _run_pair(spec, 60, check_every=5)
    assert nev > 50


def test_injection_identity_with_counter_stream():
    """Injecting the Gaussian deviates the counter stream would draw gives the
    counter-stream trajectory: both paths through the diffusion kernel agree."""
    spec = ss.c1_spec(n=500, seed=8)
    osim, _ = ss.build_oracle(spec)
    dsim = ss.build_device(spec)
    dsim.run_time_step()
    osim.step()
    ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))


def test_c2_full_size_bit_exact_three_steps():
    """The 1M-molecule headline workload: three steps compared with the oracle
    bit for bit (positions, order, serials, boxes, reaction events)."""
    spec = ss.c2_spec(n=500000, seed=2026, kback=1.0)
    nev = _run_pair(spec, 3, check_every=3)
    assert nev > 100


def test_c2_full_size_invariants_over_many_steps():
    """Properties that hold at any size: mass balance of A+B<->C, every
    molecule inside the periodic box, unique serial numbers, box index equals
    pos2box of the position after the walls, and counts from the device's
    counters equal the counts in the molecule lists."""
    spec = ss.c2_spec(n=500000, seed=77, kback=5.0)
    dsim = ss.build_device(spec)
    dsim.run_steps(400)
    st = dsim.molecules(0)
    ident = st["ident"]
    a, b, c = (int((ident == k).sum()) for k in (1, 2, 3))
    assert a + c == 500000 and b + c == 500000
    assert c > 0
    assert (dsim.count("A"), dsim.count("B"), dsim.count("C")) == (a, b, c)
    pos = st["pos"]
    assert (pos >= 0).all() and (pos <= 1000).all()
    assert len(np.unique(st["serno"])) == len(st["serno"])
    n, side, bmin, bsize = dsim.boxes()
    idx = np.clip(((pos - bmin) / bsize).astype(np.int64), 0, side - 1)
    # products created in the last step carry the box of their final position;
    # all others the box computed after diffusion, which is their position now
    assert np.array_equal((idx[:, 0] * side[1] + idx[:, 1]) * side[2] + idx[:, 2], st["box"])
    d = dsim.diagnostics()
    assert d["err"] == 0


def test_molecules_added_between_steps():
    """smolAddSolutionMolecules and smolamdAddMoleculesAt after stepping
    started (addmol without a molsort, smolmolec.c:2717-2740): the new
    molecules take the next serial numbers, join their list at the next
    step's first molsort in reverse creation order, skip that step's
    diffusion and react from then on; the injected step after an addition
    takes deviates for the old molecules only.  Identical to the oracle."""
    spec = ss.c2_spec(n=6000, side=40.0, seed=71, kback=20.0)
    osim, _ = ss.build_oracle(spec)
    dsim = ss.build_device(spec)
    rng = np.random.default_rng(5)
    lo, hi = [5.0, 5.0, 5.0], [20.0, 35.0, 30.0]

    def steps(n):
        for _ in range(n):
            osim.step()
            dsim.run_time_step()
            assert ss.canonical_pairs(osim.rxnlog()[:, :6]) == ss.canonical_pairs(dsim.reaction_log())

    steps(10)
    ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))
    osim.add_mols(50, osim.index("A"), lo, hi)
    dsim.add_molecules("A", 50, lo, hi)
    pos = rng.uniform(0, 40.0, (20, 3))
    osim.add_mols_at(osim.index("B"), pos)
    dsim.add_molecules_at("B", pos)
    assert dsim.count("A") == osim.count(osim.index("A"))
    steps(1)
    ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))
    steps(15)
    ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))
    # an injected step right after an addition
    dsim.add_molecules("C", 30)
    osim.add_mols(30, osim.index("C"), spec["low"], spec["high"])
    n = sum(len(x["ident"]) for x in ss.oracle_state(osim))  # live molecules: the new ones are not yet
    g = rng.standard_normal((n, 3))
    osim.inject(g)
    osim.step()
    dsim.inject_step(g)
    ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))
    steps(10)
    ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))
    _compare_counts(spec, osim, dsim)


# ---------------------------------------------------------------------------
# surfaces


def test_surfaces_sphere_membrane_patches_bit_exact():
    """Collisions with a reflecting sphere, a membrane that reflects B and C
    and transmits A, and disk patches where A crosses with probability 0.5
    and becomes D; reversible A+B<->C with products placed across surfaces;
    pairs separated by a reflecting panel do not react."""
    spec = ss.c3_spec(n=3000, side=60.0, seed=41, kback=20.0)
    nev = _run_pair(spec, 300, check_every=25)
    assert nev > 20


def test_surfaces_adsorption_desorption_bit_exact():
    """Adsorption of A onto disk patches from either side, desorption of the
    bound molecules (reversible-adsorption desorption distance), the list
    changes of the first molsort (bound molecules live in their own list) and
    the serial numbers desorption consumes: identical to the oracle."""
    spec = ss.c3_adsorb_spec(n=3000, side=60.0, seed=45, p_ads=0.5, k_des=5.0, kback=20.0, ndisk=4, rdisk=6.0)
    nev = _run_pair(spec, 300, check_every=10)
    assert nev > 20
    dsim = ss.build_device(spec)
    dsim.run_steps(200)
    import smoldyn_amd as sa
    assert dsim.count("A", sa.MSfront) + dsim.count("A", sa.MSback) > 0
    assert dsim.count("A", sa.MSall) == dsim.count("A", sa.MSsoln) + dsim.count("A", sa.MSfront) + \
        dsim.count("A", sa.MSback)


def test_surfaces_adsorption_injected_steps():
    """Large injected steps with adsorbing patches: several crossings per step."""
    spec = ss.c3_adsorb_spec(n=1500, side=40.0, seed=46, p_ads=0.5, k_des=20.0, kback=20.0, ndisk=4, rdisk=4.0)
    spec["dt"] = 0.2
    _run_pair(spec, 40, check_every=4, inject=True)


def test_surfaces_rate_given_interactions_bit_exact():
    """Surface actions given by rates: the host converts them to collision
    probabilities as srfcalcprob does (reversible transmission through the
    membrane, reversible adsorption from the front with reversible
    desorption, irreversible adsorption from the back, reversible flips of
    the bound states); trajectories, bound states and lists equal the
    oracle's, whose conversion is pinned to the reference's surfaceprob."""
    import smoldyn_amd as sa
    spec = ss.c3_rates_spec(n=3000, side=60.0, seed=47)
    _run_pair(spec, 300, check_every=10)
    dsim = ss.build_device(spec)
    dsim.run_steps(300)
    assert dsim.count("A", sa.MSfront) > 0 and dsim.count("A", sa.MSback) > 0


def test_surfaces_flips_between_lists_bit_exact():
    """The front and back states of A in lists of their own: a flip moves the
    molecule between the two lists at the first molsort (molchangeident's
    list change, smolmolec.c:2902-2906), like adsorption does."""
    spec = ss.c3_rates_spec(n=3000, side=60.0, seed=49, k_flip=(60.0, 90.0))
    sp = list(spec["species"])
    spec["lists"] = {"sol": [(x, "soln") for x in sp], "Afront": [("A", "front")], "Aback": [("A", "back")],
                     "fixed": [("A", "up"), ("A", "down")] + [(x, m) for x in sp[1:]
                                                               for m in ("front", "back", "up", "down")]}
    _run_pair(spec, 200, check_every=10)


def test_surfaces_absorption_bit_exact():
    """Absorbing faces: molecules killed by their surface check after
    diffusion leave their lists in the first molsort (swap from tail), which
    also orders the desorbed molecules appended in the same step (every kill
    rotates the dead list's resurrected entries); products absorbed by their
    own surface check leave in the list sort that ends the step."""
    import smoldyn_amd as sa
    spec = ss.c3_absorb_spec()
    n0 = sum(c for _, c in spec["molecules"])
    _run_pair(spec, 300, check_every=10)
    dsim = ss.build_device(spec)
    dsim.run_steps(300)
    assert dsim.count("B", sa.MSall) < n0 // 2  # B below the membrane is absorbed


def test_surfaces_absorption_injected_large_steps():
    spec = ss.c3_absorb_spec(n=1500, side=40.0, seed=52, rdisk=4.0)
    spec["dt"] = 0.1
    _run_pair(spec, 40, check_every=4, inject=True)


def test_absorbing_walls_bit_exact():
    """Absorbing walls: kills of the wall pass after diffusion leave in the
    first molsort, kills of the pass after the reactions (products and every
    molecule that did not react) in the list sort that ends the step."""
    import smoldyn_amd as sa
    spec = ss.walls_absorb_spec()
    _run_pair(spec, 150, check_every=10)
    dsim = ss.build_device(spec)
    dsim.run_steps(150)
    assert dsim.count("A", sa.MSall) + dsim.count("C", sa.MSall) < 2000


def test_absorbing_walls_injected_large_steps():
    spec = ss.walls_absorb_spec(n=2000, side=20.0, seed=62)
    spec["dt"] = 0.2
    _run_pair(spec, 30, check_every=3, inject=True)


def test_surfaces_rate_given_injected_large_steps():
    spec = ss.c3_rates_spec(n=1500, side=40.0, seed=48, rdisk=4.0)
    spec["dt"] = 0.1
    _run_pair(spec, 40, check_every=4, inject=True)


def test_surfaces_injected_large_steps():
    """Large injected steps make molecules cross several boxes and several
    panels per step (line2nextbox walk, repeated interactions)."""
    spec = ss.c3_spec(n=1500, side=40.0, seed=43, kback=20.0)
    spec["dt"] = 0.5
    nev = _run_pair(spec, 40, check_every=5, inject=True)
    assert nev >= 0


def test_surfaces_jump_panels():
    """Jump panels: molecules crossing one wall of the box reappear at the
    opposite one (the surface form of periodic boundaries)."""
    side = 50.0
    spec = dict(low=[0.0] * 3, high=[side] * 3, walls="ttt", dt=0.01, seed=47,
                species={"A": 3.0, "B": 3.0, "C": 1.0},
                reactions=[("fwd", ["A", "B"], ["C"], 50.0)],
                surfaces=[{"name": "walls", "actions": [("both", "all", "jump")],
                           "panels": [("rect", "+0", [0, 0, 0, side, side], "lo0"),
                                      ("rect", "-0", [side, 0, 0, side, side], "hi0"),
                                      ("rect", "+1", [0, 0, 0, side, side], "lo1"),
                                      ("rect", "-1", [0, side, 0, side, side], "hi1"),
                                      ("rect", "+2", [0, 0, 0, side, side], "lo2"),
                                      ("rect", "-2", [0, 0, side, side, side], "hi2")],
                           "jumps": [("lo0", "front", "hi0", "front", True),
                                     ("lo1", "front", "hi1", "front", True),
                                     ("lo2", "front", "hi2", "front", True)]}],
                molecules=[("A", 2000), ("B", 2000)])
    nev = _run_pair(spec, 200, check_every=20)
    assert nev > 10


def test_surfaces_all_shapes_bit_exact():
    """Triangles (a tetrahedron, a pair of jump triangles), a cylinder with
    hemisphere caps, a sphere and rectangle walls; reflect, jump and
    rate-based crossings with a species change; reversible A+B<->C with
    pairs tested across these panels: identical to the oracle."""
    spec = ss.c3_shapes_spec(n=3000, seed=35)
    nev = _run_pair(spec, 300, check_every=25)
    assert nev > 20


def test_surfaces_all_shapes_injected_large_steps():
    """Large injected steps across the same shapes: several crossings of
    triangles, cylinders and hemispheres per step."""
    spec = ss.c3_shapes_spec(n=1500, seed=36)
    spec["dt"] = 0.5
    _run_pair(spec, 40, check_every=4, inject=True)


def test_surfaces_all_shapes_adsorption_bit_exact():
    """Adsorption onto triangles, a cylinder and hemispheres (placement on
    the panel at the crossing point on the face of the bound state, the
    in-panel test of each shape) and desorption back off them along each
    shape's normal: identical to the oracle."""
    spec = ss.c3_shapes_spec(n=3000, seed=38)
    ads = [("A", "soln", "fsoln", "front", 0.5, True), ("A", "soln", "bsoln", "back", 0.5, True),
           ("A", "front", "front", "fsoln", 5.0, False), ("A", "back", "back", "bsoln", 5.0, False)]
    spec["surfaces"][1]["rates"] = ads
    spec["surfaces"][2]["rates"] = ads
    spec["dt"] = 0.01
    nev = _run_pair(spec, 200, check_every=10)
    assert nev > 10
    dsim = ss.build_device(spec)
    dsim.run_steps(100)
    import smoldyn_amd as sa
    assert dsim.count("A", sa.MSfront) + dsim.count("A", sa.MSback) > 0


# ---------------------------------------------------------------------------
# ensembles


def _same_lists(a, b, what):
    assert len(a["ident"]) == len(b["ident"]), what
    assert np.array_equal(a["serno"], b["serno"]), what
    assert np.array_equal(a["ident"], b["ident"]), what
    assert np.array_equal(a["pos"].view(np.uint64), b["pos"].view(np.uint64)), what
    assert np.array_equal(a["box"], b["box"]), what


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_members_equal_single_runs(graph):
    """Members of uneven size, diffusion coefficient, binding radius and box
    size stepped together give exactly their single-simulation results."""
    import smoldyn_amd as sa
    specs = [ss.c4_member(k, seed=700, scale=0.03) for k in range(7)]
    members = [ss.build_device(sp, update=False) for sp in specs]
    ens = sa.Ensemble(members)
    singles = [ss.build_device(sp) for sp in specs]
    for block in range(3):
        ens.run_steps(40, graph=graph)
        for s in singles:
            s.run_steps(40, graph=graph)
        for k, s in enumerate(singles):
            _same_lists(ens.molecules(k), s.molecules(0), "member %d after %d steps" % (k, 40 * (block + 1)))
            for sp in ("A", "B", "C"):
                assert ens.count(k, sp) == s.count(sp)
    assert sum(ens.diagnostics(k)["molsteps"] for k in range(len(specs))) == \
        sum(s.diagnostics()["molsteps"] for s in singles)


def test_ensemble_against_oracle():
    """One ensemble member checked step by step against the oracle."""
    import smoldyn_amd as sa
    specs = [ss.c4_member(k, seed=900, scale=0.02) for k in range(3)]
    members = [ss.build_device(sp, update=False) for sp in specs]
    ens = sa.Ensemble(members)
    osim, _ = ss.build_oracle(specs[1])
    for step in range(60):
        osim.step()
        ens.run_steps(1, graph=False)
        if step % 20 == 19:
            o = ss.oracle_state(osim)[0]
            d = ens.molecules(1)
            _same_lists(dict(ident=o["ident"], serno=o["serno"], pos=o["pos"], box=o["box"]), d,
                        "member 1 step %d" % step)


def test_ensemble_event_overflow_reports_error():
    """A member with far more reacting pairs in one step than its event space
    (dense, large binding radius) makes the run return an error code; the
    other steps do nothing and the process keeps running (regression test)."""
    import smoldyn_amd as sa
    spec = dict(low=[0.0] * 3, high=[12.0] * 3, walls="rrr", dt=0.01, seed=77,
                species={"A": 1.0, "B": 1.0, "C": 0.5},
                reactions=[("fwd", ["A", "B"], ["C"], 2000.0)],
                molecules=[("A", 6000), ("B", 6000)])
    small = ss.c4_member(0, seed=701, scale=0.03)
    members = [ss.build_device(small, update=False), ss.build_device(spec, update=False)]
    ens = sa.Ensemble(members)
    with pytest.raises(sa.SmoldynError):
        ens.run_steps(5)
    assert ens.diagnostics(1)["err"] != 0


def test_confspread_reactions():
    """Conformational spread (the reference's allostery model in 3D, mobile):
    red + blue -> green + blue within radius 2 with probability
    1 - exp(-k dt), green -> red; the products keep the reactants' places and
    serial numbers.  Events and lists bit-exact with the oracle."""
    spec = dict(low=[0, 0, 0], high=[40, 40, 40], walls="ppp", dt=0.01, seed=41,
                species={"red": 1.0, "green": 1.0, "blue": 0.5},
                reactions=[("back", ["green"], ["red"], 2.0), ("fwd", ["red", "blue"], ["green", "blue"], 10.0)],
                confspread={"fwd": 2.0},
                molecules=[("red", 1500), ("blue", 1500)])
    nev = _run_pair(spec, 60, check_every=10)
    assert nev > 50


def test_fixed_product_offsets():
    """Fixed product placement (off-axis offsets, not rotated) for C -> A + B,
    reversible with A + B -> C, in a reflective box: products at the reaction
    position plus the given vectors, walls applied to them."""
    spec = dict(low=[0, 0, 0], high=[30, 30, 30], walls="rrr", dt=0.01, seed=43,
                species={"A": 1.0, "B": 1.0, "C": 0.5},
                reactions=[("fwd", ["A", "B"], ["C"], 20.0), ("back", ["C"], ["A", "B"], 5.0)],
                placement={"back": ("fixed", 0.0, [("A", [0.6, 0.3, -0.2]), ("B", [-0.4, -0.2, 0.1])])},
                molecules=[("A", 1500), ("B", 1500)])
    nev = _run_pair(spec, 60, check_every=10)
    assert nev > 50


def test_zeroth_order_production():
    """Zeroth-order production (0 -> A at 12.8 events per step: two Poisson
    chunks; 0 -> B at 0.64) with A + B <-> C and C -> 0 in a periodic box:
    event counts, product positions and serial numbers bit-exact with the
    oracle's counter stream."""
    vol = 40.0 ** 3
    spec = dict(low=[0, 0, 0], high=[40, 40, 40], walls="ppp", dt=0.01, seed=47,
                species={"A": 1.0, "B": 1.0, "C": 0.5},
                reactions=[("makeA", [], ["A"], 12.8 / (0.01 * vol)), ("makeB", [], ["B"], 0.64 / (0.01 * vol)),
                           ("fwd", ["A", "B"], ["C"], 20.0), ("back", ["C"], ["A", "B"], 2.0),
                           ("decay", ["C"], [], 0.5)],
                placement={"back": ("ratio", 1.2)},
                molecules=[("A", 1500), ("B", 1500)])
    nev = _run_pair(spec, 80, check_every=10)
    assert nev > 1000


def test_offset_placement_rotated_off_axis():
    """Offset product placement with offsets off the x axis, C -> A + B
    reversible with A + B -> C: each event turns both offset vectors by one
    random rotation (the counter stream's quaternion draw)."""
    spec = dict(low=[0, 0, 0], high=[30, 30, 30], walls="ppp", dt=0.01, seed=53,
                species={"A": 1.0, "B": 1.0, "C": 0.5},
                reactions=[("fwd", ["A", "B"], ["C"], 20.0), ("back", ["C"], ["A", "B"], 5.0)],
                placement={"back": ("offset", 0.0, [("A", [0.5, 0.4, 0.0]), ("B", [-0.3, 0.0, 0.2])])},
                molecules=[("A", 1500), ("B", 1500)])
    nev = _run_pair(spec, 60, check_every=10)
    assert nev > 50


def test_bounce_placement():
    """Bounce product placement (A + B -> A + B to separation 1.0, and
    A + A -> A + A reflected off the binding radius) in a periodic box,
    wrapped pairs included: the molecules keep their serial numbers and move
    apart along the line joining them."""
    spec = dict(low=[0, 0, 0], high=[25, 25, 25], walls="ppp", dt=0.01, seed=59,
                species={"A": 1.0, "B": 0.5},
                reactions=[("hitAB", ["A", "B"], ["A", "B"], 30.0), ("hitAA", ["A", "A"], ["A", "A"], 15.0)],
                placement={"hitAB": ("bounce", 1.0), "hitAA": ("bounce", -1.0)},
                molecules=[("A", 1500), ("B", 1500)])
    nev = _run_pair(spec, 60, check_every=10)
    assert nev > 100


@pytest.mark.parametrize("walls,precision", [("ppp", 64), ("rrr", 64), ("ppp", 32)])
def test_bounce_reflection_method(walls, precision):
    """Bounce by reflection (rparam -2): each pair's paths from the places
    before diffusion (posx, kept per record for this) are reflected at the
    moment of contact; pairs across the periodic boundary move posx with
    the faster molecule (morebireact).  Bit-exact with the oracle over the
    tile and counting-sort steps.  FP32-offset records: the first step's
    events equal the FP64 run's (bounced products sit at the binding radius,
    so later steps part at the FP32 rounding), and the species counts of the
    conserving reactions hold."""
    spec = dict(low=[0, 0, 0], high=[25, 25, 25], walls=walls, dt=0.01, seed=61,
                species={"A": 1.0, "B": 0.5},
                reactions=[("hitAB", ["A", "B"], ["A", "B"], 40.0), ("hitAA", ["A", "A"], ["A", "A"], 15.0)],
                placement={"hitAB": ("bounce", -2.0), "hitAA": ("bounce", -2.0)},
                molecules=[("A", 1500), ("B", 1500)])
    if precision == 64:
        nev = _run_pair(spec, 60, check_every=10)
        assert nev > 100
        return
    d64, d32 = ss.build_device(spec), ss.build_device(dict(spec, precision=32))
    d64.run_time_step()
    d32.run_time_step()
    assert ss.canonical_pairs(d64.reaction_log()) == ss.canonical_pairs(d32.reaction_log())
    nev = 0
    for _ in range(30):
        d32.run_time_step()
        nev += len(d32.reaction_log())
    assert nev > 50 and d32.count("A") == 1500 and d32.count("B") == 1500


@pytest.mark.parametrize("kind,param", [("pgem", 0.3), ("pgemmax", 0.2), ("unbindrad", 1.4)])
def test_geminate_and_unbinding_placements(kind, param):
    """C -> A + B reversible with A + B -> C, the unbinding separation from
    the geminate-rebinding probability (pgem: forward rate scaled by
    1 - pgem; pgemmax: only when the unbinding radius comes out positive;
    rxnsetrate smolreact.c:1383-1404, rxnsetproduct :1553-1570) or given
    directly (unbindrad): products and serial numbers bit-exact with the
    oracle, whose placements are pinned by closed forms and the reference's
    bounds2 output."""
    spec = dict(low=[0, 0, 0], high=[30, 30, 30], walls="ppp", dt=0.01, seed=67,
                species={"A": 1.0, "B": 0.7, "C": 0.5},
                reactions=[("fwd", ["A", "B"], ["C"], 25.0), ("back", ["C"], ["A", "B"], 8.0)],
                placement={"back": (kind, param)},
                molecules=[("A", 2000), ("B", 2000)])
    nev = _run_pair(spec, 60, check_every=10)
    assert nev > 100


def _desorb_offset_spec(n=2000, seed=61, delta=5e-14):
    """Molecules of A above a plane P (z = 50) that adsorbs A from its front
    (probability 0.7) and from which A desorbs at rate 50; a second plane Q
    lies delta above P, transmits A from its front and reflects A from its
    back.  A desorbed molecule starts between the planes, so its path meets Q
    at a fraction (delta - ulp)/x of its length: below VERYCLOSE (1e-12) for
    most desorption distances x."""
    rng = np.random.default_rng(seed)
    pos = np.column_stack([rng.uniform(5, 95, (n, 2)), rng.uniform(50.02, 50.3, n)])
    return dict(low=[0, 0, 0], high=[100, 100, 100], walls="rrr", dt=0.01, seed=seed,
                species={"A": 1.0},
                surfaces=[{"name": "p", "actions": [],
                           "rates": [("A", "soln", "fsoln", "front", 0.7, True),
                                     ("A", "front", "front", "fsoln", 50.0, False)],
                           "panels": [("rect", "+2", [0.0, 0.0, 50.0, 100.0, 100.0], "p0")]},
                          {"name": "q", "actions": [("front", "A", "transmit"), ("back", "A", "reflect")],
                           "panels": [("rect", "+2", [0.0, 0.0, 50.0 + delta, 100.0, 100.0], "q0")]}],
                molecules=[("A", pos)])


def test_desorbed_molecule_ignores_crossings_at_its_start():
    """checksurfaces1mol with crossminimum 0 (smolsurface.c:5041,5087): the
    surface check of a desorbed molecule skips crossings within VERYCLOSE of
    its start, so a desorbing molecule passes the plane Q that lies a few ulps
    above the panel it leaves.  Counting the crossing instead reflects it back
    onto P (about a hundred times more molecules stay bound after 200 steps,
    which the oracle run with the filter removed shows); the device must match
    the oracle, filter included, bit for bit."""
    spec = _desorb_offset_spec()
    _run_pair(spec, 200, check_every=20)


def test_surface_outcome_not_run_halts_and_reports():
    """A rate-based crossing that changes A into B, which lives in another
    molecule list (B does not diffuse, so it sits in the default list of
    immobile states), is an outcome the device does not run: the step reports
    the surface error and stepping stops (the step counter stays at the
    failing step) instead of continuing with a half-processed molecule."""
    import smoldyn_amd as sa
    rng = np.random.default_rng(71)
    pos = np.column_stack([rng.uniform(5, 45, (500, 2)), rng.uniform(25.001, 25.02, 500)])
    spec = dict(low=[0, 0, 0], high=[50, 50, 50], walls="rrr", dt=0.01, seed=71,
                species={"A": 1.0, "B": 0.0},
                surfaces=[{"name": "m", "actions": [("both", "B", "reflect")],
                           "rates": [("A", "soln", "fsoln", "bsoln", 1.0, True, "B"),
                                     ("A", "soln", "bsoln", "fsoln", 1.0, True, "B")],
                           "panels": [("rect", "+2", [0.0, 0.0, 25.0, 50.0, 50.0], "m0")]}],
                molecules=[("A", pos)])
    dsim = ss.build_device(spec)
    with pytest.raises(sa.SmoldynError) as e:
        dsim.run_steps(50)
    assert "surface outcome not run on the device" in str(e.value)
    d = dsim.diagnostics()
    assert d["err"] & 16 and d["step"] < 50
    with pytest.raises(sa.SmoldynError):
        dsim.run_steps(5)
    assert dsim.diagnostics()["step"] == d["step"]
