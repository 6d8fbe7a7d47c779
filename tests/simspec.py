"""Build the same simulation in the CPU oracle and in the device engine.

A spec is a dict: low/high bounds, walls ('r','p','t','a' per dimension, or
'lh' for the low and high sides), dt,
seed, species {name: difc}, optional lists {list: [species or (species, state)]}, reactions
[(name, [reactants], [products], rate)], product placements
{name: (kind, param)}, molecules [(species, n)] placed uniformly from the
counter stream, or explicit positions [(species, array)], and optional
surfaces [{name, actions [(face, species, action)], rates [(species, state,
state1, state2, value, internal)], panels [(shape, axis, params, name)],
jumps [(panel1, face1, panel2, face2, bidirectional)]}] set up the way the
libsmoldyn calls set them (smolSetSurfaceAction passes no species change).
"""
import numpy as np

import oracle_bind as ob


_STATES = {"soln": 0, "front": 1, "back": 2, "up": 3, "down": 4}  # MolecState codes (smoldyn.h)


def build_oracle(spec, sfmt=False):
    """sfmt: draw from the reference's own random stream (SFMT-19937 in
    program order) instead of the counter stream the device uses."""
    sim = ob.OSim(3, spec["low"], spec["high"])
    for d, w in enumerate(spec["walls"]):
        sim.wall(d, 0, w[0])
        sim.wall(d, 1, w[-1])
    sim.times(0.0, spec.get("tmax", 1e9), spec["dt"])
    if sfmt:
        sim.rngmode(ob.RNG_SFMT, 0)
        sim.seed(spec["seed"])
    else:
        sim.rngmode(ob.RNG_COUNTER, spec["seed"])
    idx = {}
    for name, d in spec["species"].items():
        idx[name] = sim.species(name)
        sim.difc(idx[name], d, 0)
    for ll, (lname, members) in enumerate(spec.get("lists", {}).items()):
        sim.add_list(lname)
        for m in members:
            if isinstance(m, tuple):  # (species, state): one state's list
                sim.set_list(idx[m[0]], ll, _STATES[m[1]])
            else:
                sim.set_list(idx[m], ll, 6)
    if "molperbox" in spec:
        sim.partitions("molperbox", spec["molperbox"])
    if "boxsize" in spec:
        sim.partitions("boxsize", spec["boxsize"])
    rx = {}
    for name, rct, prd, rate in spec.get("reactions", []):
        rx[name] = sim.reaction(name, [idx[r] for r in rct], [idx[p] for p in prd], rate)
    for name, pl in spec.get("placement", {}).items():
        sim.revparam(rx[name], pl[0], pl[1])
        for prod, vec in (pl[2] if len(pl) > 2 else []):
            rct, prd = next((r[1], r[2]) for r in spec["reactions"] if r[0] == name)
            sim.prdpos(rx[name], prd.index(prod), vec)
    for name, radius in spec.get("confspread", {}).items():
        sim.confspread(rx[name], radius)
    for name, radius in spec.get("bindrad", {}).items():
        sim.bindrad(rx[name], radius)
    for srf in spec.get("surfaces", []):
        s = sim.add_surface(srf["name"])
        for face, sp, act in srf.get("actions", []):
            sim.surface_action(s, -5 if sp == "all" else idx[sp], 0, ob.FACE[face], ob.SA[act], -1)
        for rt in srf.get("rates", []):
            sp, st, st1, st2, value, internal = rt[:6]
            newid = idx[rt[6]] if len(rt) > 6 else -5
            sim.surface_rate(s, -5 if sp == "all" else idx[sp], ob.MS[st], ob.MS[st1], ob.MS[st2], value,
                             2 if internal else 1, newid)
        pid = {}
        for k, (shape, axis, params, pname) in enumerate(srf.get("panels", [])):
            pid[pname] = sim.add_panel(s, ob.PS[shape], params, axis)
        for p1, f1, p2, f2, bidir in srf.get("jumps", []):
            sim.panel_jump(pid[p1], ob.FACE[f1], pid[p2], ob.FACE[f2], bidir)
    for sp, n in spec.get("molecules", []):
        if isinstance(n, int):
            sim.add_mols(n, idx[sp], spec.get("mol_low", spec["low"]), spec.get("mol_high", spec["high"]))
        else:
            sim.add_mols_at(idx[sp], n)
    sim.update()
    return sim, rx


def build_device(spec, update=True):
    import smoldyn_amd as sa
    sim = sa.Simulation(spec["low"], spec["high"])
    for d, w in enumerate(spec["walls"]):
        sim.boundary(d, 0, w[0])
        sim.boundary(d, 1, w[-1])
    sim.times(0.0, spec.get("tmax", 1e9), spec["dt"])
    sim.seed(spec["seed"])
    if "precision" in spec:
        sim.precision(spec["precision"])
    for lname in spec.get("lists", {}):
        sim.add_list(lname)
    for name, d in spec["species"].items():
        sim.species(name)
        sim.mobility(name, d)
    for lname, members in spec.get("lists", {}).items():
        for m in members:
            if isinstance(m, tuple):
                sim.set_list(m[0], lname, _STATES[m[1]])
            else:
                sim.set_list(m, lname)
    if "molperbox" in spec:
        sim.partitions("molperbox", spec["molperbox"])
    if "boxsize" in spec:
        sim.partitions("boxsize", spec["boxsize"])
    for name, rct, prd, rate in spec.get("reactions", []):
        sim.reaction(name, rct, prd, rate)
    for name, pl in spec.get("placement", {}).items():
        if len(pl) > 2:  # offset / fixed: one call per product position (RxnSetRevparam)
            for prod, vec in pl[2]:
                sim.product_placement(name, pl[0], pl[1], prod, vec)
        else:
            sim.product_placement(name, pl[0], pl[1])
    for name, radius in spec.get("confspread", {}).items():
        sim.product_placement(name, "confspread")
        sim.binding_radius(name, radius)
    for name, radius in spec.get("bindrad", {}).items():
        sim.binding_radius(name, radius)
    for srf in spec.get("surfaces", []):
        sim.add_surface(srf["name"])
        for face, sp, act in srf.get("actions", []):
            sim.surface_action(srf["name"], face, sp, act)
        for rt in srf.get("rates", []):
            sp, st, st1, st2, value, internal = rt[:6]
            sim.surface_rate(srf["name"], sp, st, st1, st2, value, rt[6] if len(rt) > 6 else None, internal)
        for shape, axis, params, pname in srf.get("panels", []):
            sim.add_panel(srf["name"], shape, params, axis, pname)
        for p1, f1, p2, f2, bidir in srf.get("jumps", []):
            sim.panel_jump(srf["name"], p1, f1, p2, f2, bidir)
    for sp, n in spec.get("molecules", []):
        if isinstance(n, int):
            sim.add_molecules(sp, n, spec.get("mol_low"), spec.get("mol_high"))
        else:
            sim.add_molecules_at(sp, n)
    if update:
        sim.update()
    return sim


def c3_spec(n=2000, side=100.0, seed=31, kback=5.0, ndisk=8, rdisk=None, bindrad_rate=0.2):
    """Surfaces workload, scaled: A and B inside a reflecting sphere (radius
    0.4*side, centre of the cube), a membrane plane through the centre that
    reflects B and transmits A, disk patches on the membrane where A crosses
    with probability 0.5 and turns into species D (rate_internal with a new
    species), reversible A+B<->C with unbinding placement, D=2, dt=0.002 ..."""
    c = side / 2
    r = 0.4 * side
    h = r / 3 ** 0.5 * 0.95
    rng = np.random.default_rng(seed)
    rdisk = rdisk if rdisk is not None else side / 40
    panels = [("rect", "+2", [0.0, 0.0, c, side, side], "membrane")]
    disks = []
    for k in range(ndisk):
        ang = rng.uniform(0, 2 * np.pi)
        rad = rng.uniform(0, r * 0.7)
        disks.append(("disk", None, [c + rad * np.cos(ang), c + rad * np.sin(ang), c, rdisk, 0, 0, 1, 20], "d%d" % k))
    return dict(low=[0.0] * 3, high=[side] * 3, walls="rrr", dt=0.002, seed=seed,
                species={"A": 2.0, "B": 2.0, "C": 1.0, "D": 2.0},
                reactions=[("fwd", ["A", "B"], ["C"], bindrad_rate * 100.0), ("back", ["C"], ["A", "B"], kback)],
                placement={"back": ("ratio", 1.5)},
                surfaces=[
                    {"name": "ball", "actions": [("both", "all", "reflect")],
                     "panels": [("sph", None, [c, c, c, r, 20, 20], "s0")]},
                    {"name": "membrane", "actions": [("both", "B", "reflect"), ("both", "C", "reflect")],
                     "panels": panels},
                    {"name": "patches", "actions": [("both", "B", "reflect")],
                     "rates": [("A", "soln", "fsoln", "bsoln", 0.5, True, "D"),
                               ("A", "soln", "bsoln", "fsoln", 0.5, True, "D")],
                     "panels": disks},
                ],
                mol_low=[c - h] * 3, mol_high=[c + h] * 3,
                molecules=[("A", n), ("B", n)])


def c3_shapes_spec(n=2000, side=60.0, seed=35, kback=20.0, p_cross=0.5):
    """Every 3D panel shape at once: reflecting walls built from 6 rectangles;
    a tetrahedron of 4 triangles that reflects B and C while A crosses with
    probability p_cross and becomes D; a cylinder with hemisphere caps that
    reflects everything but A, which crosses with probability p_cross; a
    sphere inside the cylinder; a pair of triangles whose front faces jump
    to each other.  Reversible A+B<->C with unbinding placement, so bound
    pairs are tested for panels between them."""
    s = side
    v = [(8.0, 8.0, 8.0), (32.0, 8.0, 8.0), (20.0, 32.0, 8.0), (20.0, 16.0, 32.0)]
    faces = [(0, 1, 2), (0, 3, 1), (1, 3, 2), (2, 3, 0)]
    tet = [("tri", None, [c for k in f for c in v[k]], "t%d" % j) for j, f in enumerate(faces)]
    cx, cy, z0, z1, rc = 40.0, 40.0, 15.0, 45.0, 6.0
    tube = [("cyl", None, [cx, cy, z0, cx, cy, z1, rc, 20, 20], "c0"),
            ("hemi", None, [cx, cy, z0, rc, 0, 0, 1, 20, 20], "h0"),
            ("hemi", None, [cx, cy, z1, rc, 0, 0, -1, 20, 20], "h1")]
    walls = [("rect", "+0", [0, 0, 0, s, s], "w0"), ("rect", "-0", [s, 0, 0, s, s], "w1"),
             ("rect", "+1", [0, 0, 0, s, s], "w2"), ("rect", "-1", [0, s, 0, s, s], "w3"),
             ("rect", "+2", [0, 0, 0, s, s], "w4"), ("rect", "-2", [0, 0, s, s, s], "w5")]
    jumps = [("tri", None, [10, 40, 40, 10, 50, 40, 10, 45, 50], "j0"),
             ("tri", None, [50, 10, 10, 50, 20, 10, 50, 15, 20], "j1")]
    return dict(low=[0.0] * 3, high=[s] * 3, walls="rrr", dt=0.002, seed=seed,
                species={"A": 2.0, "B": 2.0, "C": 1.0, "D": 2.0},
                reactions=[("fwd", ["A", "B"], ["C"], 20.0), ("back", ["C"], ["A", "B"], kback)],
                placement={"back": ("ratio", 1.5)},
                surfaces=[
                    {"name": "walls", "actions": [("both", "all", "reflect")], "panels": walls},
                    {"name": "tet", "actions": [("both", "B", "reflect"), ("both", "C", "reflect")],
                     "rates": [("A", "soln", "fsoln", "bsoln", p_cross, True, "D"),
                               ("A", "soln", "bsoln", "fsoln", p_cross, True, "D")],
                     "panels": tet},
                    {"name": "tube", "actions": [("both", "B", "reflect"), ("both", "C", "reflect"),
                                                 ("both", "D", "reflect")],
                     "rates": [("A", "soln", "fsoln", "bsoln", p_cross, True),
                               ("A", "soln", "bsoln", "fsoln", p_cross, True)],
                     "panels": tube + [("sph", None, [cx, cy, 30.0, 3.0, 20, 20], "s0")]},
                    {"name": "jumper", "actions": [("front", "all", "jump"), ("back", "all", "reflect")],
                     "panels": jumps, "jumps": [("j0", "front", "j1", "front", True)]},
                ],
                molecules=[("A", n), ("B", n)])


def c3_adsorb_spec(n=2000, side=100.0, seed=33, ndisk=8, p_ads=0.3, k_des=2.0, kback=5.0, rdisk=None):
    """Surfaces workload with surface-bound molecules: c3_spec's sphere and
    membrane, disk patches on the membrane where A crossing from either side
    adsorbs with probability p_ads (to the front or back state) and bound A
    desorbs at rate k_des; bound molecules sit in their own list."""
    spec = c3_spec(n=n, side=side, seed=seed, kback=kback, ndisk=ndisk, rdisk=rdisk)
    patches = spec["surfaces"][2]
    patches["rates"] = [("A", "soln", "fsoln", "front", p_ads, True),
                        ("A", "soln", "bsoln", "back", p_ads, True),
                        ("A", "front", "front", "fsoln", k_des, False),
                        ("A", "back", "back", "bsoln", k_des, False)]
    return spec


def c3_absorb_spec(n=3000, side=60.0, seed=51, p_ads=0.5, k_des=20.0, kback=20.0, ndisk=4, rdisk=6.0):
    """c3_adsorb_spec with absorbing faces: C that reaches the sphere from
    inside dies, and so does B that reaches the membrane from below (B
    products placed across the membrane by the back reaction included).
    Desorption is fast, so absorptions and desorptions share steps (the
    first molsort's dead-list order)."""
    spec = c3_adsorb_spec(n=n, side=side, seed=seed, p_ads=p_ads, k_des=k_des, kback=kback, ndisk=ndisk,
                          rdisk=rdisk)
    spec["surfaces"][0]["actions"] = spec["surfaces"][0]["actions"] + [("back", "C", "absorb")]
    spec["surfaces"][1]["actions"] = spec["surfaces"][1]["actions"] + [("back", "B", "absorb")]
    return spec


def c3_rates_spec(n=2000, side=60.0, seed=47, kback=20.0, ndisk=4, rdisk=6.0, k_ads=8.0, k_des=20.0,
                  k_flip=(10.0, 30.0), k_trans=(5.0, 2.0)):
    """Surface interactions given as rates (`rate`, converted to collision
    probabilities at setup as srfcalcprob does): c3_spec's membrane transmits
    A both ways at rates k_trans (reversible transmission, surfacetransmit);
    its disk patches adsorb A from the front reversibly (k_ads, with
    desorption back to the front at k_des: lookuprevads) and from the back
    irreversibly (lookupirrevadsorb), and the bound states flip between front
    and back at rates k_flip (reversible flips)."""
    spec = c3_spec(n=n, side=side, seed=seed, kback=kback, ndisk=ndisk, rdisk=rdisk)
    spec["surfaces"][1]["rates"] = [("A", "soln", "fsoln", "bsoln", k_trans[0], False),
                                    ("A", "soln", "bsoln", "fsoln", k_trans[1], False)]
    spec["surfaces"][2]["rates"] = [("A", "soln", "fsoln", "front", k_ads, False),
                                    ("A", "soln", "bsoln", "back", k_ads, False),
                                    ("A", "front", "front", "fsoln", k_des, False),
                                    ("A", "front", "front", "back", k_flip[0], False),
                                    ("A", "back", "back", "front", k_flip[1], False)]
    return spec


def walls_absorb_spec(n=4000, side=40.0, seed=61, kback=20.0, dt=0.02):
    """Absorbing walls (no surfaces): A+B<->C in a box whose x-low wall and
    both z walls absorb, x-high reflects and y is periodic; molecules die
    beyond an absorbing wall, or inside it with the Brownian-bridge
    probability, in both wall passes of a step (after diffusion, and after
    the reactions for every molecule)."""
    return dict(low=[0.0] * 3, high=[side] * 3, walls=["ar", "p", "aa"], dt=dt, seed=seed,
                species={"A": 2.0, "B": 2.0, "C": 1.0},
                reactions=[("fwd", ["A", "B"], ["C"], 20.0), ("back", ["C"], ["A", "B"], kback)],
                placement={"back": ("ratio", 1.5)},
                molecules=[("A", n // 2), ("B", n // 2)])


def c4_member(k, seed=500, n_lo=1000, n_hi=100000, side_lo=50.0, side_hi=400.0, dt=0.01, scale=1.0):
    """Member k of the batched parameter sweep: A+B<->C in a reflective cube,
    N log-uniform in [n_lo, n_hi] (split evenly between A and B), D
    log-uniform in [0.1, 10] (A; B at half, C at a quarter), binding radius uniform in
    [0.2, 2.0] (forward rate from the reaction-parameter solver), side
    uniform in [side_lo, side_hi], unbinding radius 1.5x the binding radius,
    back rate 1.  scale multiplies N (for smaller tests)."""
    import smoldyn_amd as sa
    rng = np.random.default_rng(seed * 100003 + k)
    n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))) * scale)
    n = max(n, 20)
    d = float(np.exp(rng.uniform(np.log(0.1), np.log(10.0))))
    a = float(rng.uniform(0.2, 2.0))
    side = float(rng.uniform(side_lo, side_hi))
    sa.lib()
    # forward rate whose binding radius is a with unbinding radius 1.5 a
    # (numrxnrate(step, a, b) = rate * dt, rxnparam.c:558-583)
    dsum = 1.5 * d
    step = (2 * dsum * dt) ** 0.5
    rate = sa.lib().smolamdNumRxnRate(step, a, 1.5 * a) / dt
    return dict(low=[0.0] * 3, high=[side] * 3, walls="rrr", dt=dt, seed=seed + k,
                species={"A": d, "B": d / 2, "C": d / 4},
                reactions=[("fwd", ["A", "B"], ["C"], rate), ("back", ["C"], ["A", "B"], 1.0)],
                placement={"back": ("ratio", 1.5)},
                molecules=[("A", n // 2), ("B", n - n // 2)])


def c5_member(k, seed=700, dt=0.01, n=10000, side=100.0):
    """Member k of the learned-surrogate sweep: A+B<->C in a reflective cube of
    side 100 with n/2 A + n/2 B, D = 1/1/0.5, dt 0.01 (rms step of the pair
    0.2); rms step / binding radius log-uniform in [0.05, 5], unbinding /
    binding radius ratio uniform in [1, 3]; the target forward rate is the
    reaction-parameter solver's rate at those radii, and the setup inverts it
    back into a binding radius (the surrogate's job).  Back rate 1."""
    import smoldyn_amd as sa
    rng = np.random.default_rng(seed * 100003 + k)
    step = (2 * 2.0 * dt) ** 0.5
    while True:
        soa = float(np.exp(rng.uniform(np.log(0.05), np.log(5.0))))
        ratio = float(rng.uniform(1.0, 3.0))
        a = step / soa
        rate = sa.lib().smolamdNumRxnRate(step, a, ratio * a) / dt
        if rate > 0:
            break
    return dict(low=[0.0] * 3, high=[side] * 3, walls="rrr", dt=dt, seed=seed + k,
                species={"A": 1.0, "B": 1.0, "C": 0.5},
                reactions=[("fwd", ["A", "B"], ["C"], rate), ("back", ["C"], ["A", "B"], 1.0)],
                placement={"back": ("ratio", ratio)},
                molecules=[("A", n // 2), ("B", n - n // 2)], target_a=a, ratio=ratio, step=step)


def oracle_state(osim):
    out = []
    for ll in range(osim.nlist()):
        out.append(osim.get_list(ll))
    return out


def device_state(dsim):
    return [dsim.molecules(ll) for ll in range(dsim.nlists())]


def compare_states(o, d, check_box=True):
    """Bit-exact comparison of live lists (order, species, serials, FP64 positions, boxes)."""
    assert len(o) == len(d)
    for ll, (a, b) in enumerate(zip(o, d)):
        assert len(a["ident"]) == len(b["ident"]), "list %d length %d vs %d" % (ll, len(a["ident"]), len(b["ident"]))
        assert np.array_equal(a["serno"], b["serno"]), "list %d serial order differs" % ll
        assert np.array_equal(a["ident"], b["ident"]), "list %d species differ" % ll
        da = a["pos"].view(np.uint64)
        db = b["pos"].view(np.uint64)
        if not np.array_equal(da, db):
            bad = np.argwhere(da != db)
            k = bad[0][0]
            raise AssertionError("list %d position bits differ at %d: %r vs %r (max |diff| %g)" % (
                ll, k, a["pos"][k], b["pos"][k], np.abs(a["pos"] - b["pos"]).max()))
        if check_box:
            assert np.array_equal(a["box"], b["box"]), "list %d box indices differ" % ll


def canonical_pairs(rows):
    """Reacted-pair set: sorted tuples (order, reaction, sorted reactant serials, product serials)."""
    out = []
    for r in rows:
        rct = tuple(sorted(int(x) for x in r[2:4]))
        out.append((int(r[0]), int(r[1]), rct, int(r[4]), int(r[5])))
    return sorted(out)


def c1_spec(n=10000, seed=11, **kw):
    """Irreversible A+B->C in a reflective cube of side 100; D=1/1/0.5, dt 0.01,
    k giving a binding radius of about 0.5 (rxnparam)."""
    spec = dict(low=[0, 0, 0], high=[100, 100, 100], walls="rrr", dt=0.01, seed=seed,
                species={"A": 1.0, "B": 1.0, "C": 0.5},
                reactions=[("fwd", ["A", "B"], ["C"], 9.640578597)],
                molecules=[("A", n), ("B", n)])
    spec.update(kw)
    return spec


def c2_spec(n=500000, side=1000.0, seed=7, kback=1.0, **kw):
    """Reversible A+B<->C in a periodic cube; D=1/1/0.5, dt 0.005, forward rate
    giving binding radius ~1.0 with unbinding radius 1.2x (product_placement
    ratio 1.2 on the reverse reaction)."""
    spec = dict(low=[0, 0, 0], high=[side] * 3, walls="ppp", dt=0.005, seed=seed,
                species={"A": 1.0, "B": 1.0, "C": 0.5},
                reactions=[("fwd", ["A", "B"], ["C"], 99.3218287), ("back", ["C"], ["A", "B"], kback)],
                placement={"back": ("ratio", 1.2)},
                molecules=[("A", n), ("B", n)])
    spec.update(kw)
    return spec
