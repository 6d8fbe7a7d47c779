"""Benchmark of the MI355X Smoldyn particle step.

Workload (the headline configuration): reversible A+B<->C in a periodic cube of
side 1000 with 500k A + 500k B (1M molecules), D = 1/1/0.5, dt = 0.005,
forward rate 99.3218287 (binding radius ~1.0 by the reaction-parameter solver)
with unbinding radius 1.2x the binding radius, virtual boxes of ~4 molecules,
FP64 positions.  One "step" is one reference time step (diffusion, walls,
re-binning, first-order reactions, bimolecular pair search and reaction,
list maintenance) of that whole system, with inputs resident in HBM.

Prints one JSON line: molecule-steps/s (sum over the timed steps of the live
molecules at step start, divided by the timed wall time), the dominant
kernel's HBM roofline fraction (HIP-event timing, algorithmic bytes), and the
CPU oracle's throughput on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg):
    """Progress line on stderr (the JSON result goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)

# Algorithmic HBM bytes per molecule per launch, as the survey of the
# reference fixes them for an FP64 structure-of-arrays step of three passes
# (counted once per molecule, whatever the layout moves besides):
#   k_diffuse        pass A: read pos 24 + species 2, write pos 24 + key 4  = 54
#   k_hsum+k_hscan   cell histogram and scan, 16 B per cell / 4 per cell   =  4
#   k_scatter        pass B: read key 4 + pos 24 + species 2 + rank 4,
#                    write the 30 B payload reordered                      = 64
#   k_pairs          pass C: read own 26 B + cell offsets 1, write result 4 = 31
#   k_exact, k_resolve  work per candidate pair and per event, not per molecule
# 153 B per molecule-step in all.  The tile binning (one simulation without
# surfaces) runs pass A as k_tile1 (54 B) and pass B with the cell scan as
# k_tile2 (64 B; its per-key scan is in LDS) with k_tile_scan between them
# (per tile of ~500 records, no per-molecule bytes).  The layout's own bytes (design bytes: the record also carries its
# serial number, live-list position and rank, and the pair screen's packed
# FP32 copy) are reported beside them.
ALGO_BYTES = {"diffuse": 54, "scan": 4, "scatter": 64, "relist": 0, "pairs": 31, "exact": 0, "resolve": 0,
              "tile1": 54, "tilescan": 0, "tile2": 64}
DESIGN_BYTES = {"diffuse": 72, "scan": 16, "scatter": 112, "relist": 0, "pairs": 24, "exact": 0, "resolve": 0,
                "tile1": 72, "tilescan": 0, "tile2": 112}
STEP_ALGO_BYTES = 153
KERNEL_NAMES = ["diffuse", "scan", "scatter", "relist", "pairs", "exact", "resolve", "tile1", "tilescan", "tile2"]
HIP_KERNEL_SYMBOL = {"diffuse": "k_diffuse", "scan": "k_hsum+k_hscan", "scatter": "k_scatter", "relist": "k_relist",
                     "pairs": "k_pairs", "exact": "k_exact", "resolve": "k_resolve", "tile1": "k_tile1",
                     "tilescan": "k_tile_scan", "tile2": "k_tile2"}


def c2_spec(n_per_species=500000, side=1000.0, seed=2026, kback=1.0):
    return dict(low=[0.0] * 3, high=[side] * 3, walls="ppp", dt=0.005, seed=seed,
                species={"A": 1.0, "B": 1.0, "C": 0.5},
                reactions=[("fwd", ["A", "B"], ["C"], 99.3218287), ("back", ["C"], ["A", "B"], kback)],
                placement={"back": ("ratio", 1.2)},
                molecules=[("A", n_per_species), ("B", n_per_species)])


def build_device(spec, precision=64):
    import smoldyn_amd as sa
    sim = sa.Simulation(spec["low"], spec["high"])
    for d, w in enumerate(spec["walls"]):
        sim.boundary(d, 0, w)
        sim.boundary(d, 1, w)
    sim.times(0.0, 1e30, spec["dt"])
    sim.seed(spec["seed"])
    sim.precision(precision)
    for name, d in spec["species"].items():
        sim.species(name)
        sim.mobility(name, d)
    for name, rct, prd, rate in spec["reactions"]:
        sim.reaction(name, rct, prd, rate)
    for name, (kind, param) in spec["placement"].items():
        sim.product_placement(name, kind, param)
    for sp, n in spec["molecules"]:
        sim.add_molecules(sp, n)
    sim.update()
    return sim


def c4_spec(k, seed=500, dt=0.01):
    """Member k of the batched parameter sweep: A+B<->C in a reflective cube
    of side uniform in [50, 400], N log-uniform in [1e3, 1e5] (A and B), D_A
    log-uniform in [0.1, 10] (B at half, C at a quarter), binding radius
    uniform in [0.2, 2.0] with unbinding radius 1.5x (forward rate from the
    reaction-parameter solver), back rate 1, dt 0.01."""
    import smoldyn_amd as sa
    rng = np.random.default_rng(seed * 100003 + k)
    n = int(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
    d = float(np.exp(rng.uniform(np.log(0.1), np.log(10.0))))
    a = float(rng.uniform(0.2, 2.0))
    side = float(rng.uniform(50.0, 400.0))
    dsum = 1.5 * d
    rate = sa.lib().smolamdNumRxnRate((2 * dsum * dt) ** 0.5, a, 1.5 * a) / dt
    return dict(low=[0.0] * 3, high=[side] * 3, walls="rrr", dt=dt, seed=seed + k,
                species={"A": d, "B": d / 2, "C": d / 4},
                reactions=[("fwd", ["A", "B"], ["C"], rate), ("back", ["C"], ["A", "B"], 1.0)],
                placement={"back": ("ratio", 1.5)},
                molecules=[("A", n // 2), ("B", n - n // 2)])


def c3_surfaces_spec(n_per_species=250000, side=500.0, radius=200.0, ndisk=64, rdisk=10.0, seed=31):
    """Surfaces workload: A and B uniform in a reflecting sphere of radius 200
    centred in a cube of side 500; a membrane plane through the centre
    (z = 250) that transmits A and reflects B and C; 64 disks of radius 10 at
    random non-overlapping centres on it, where A crossing either face
    adsorbs with probability 0.2 (rate_internal) and bound A desorbs at rate
    0.1; A+B->C with rate 21.42267826 (binding radius 0.5), D_A = D_B = 2,
    D_C = 1, dt 0.002."""
    rng = np.random.default_rng(seed)
    c = side / 2
    centres = []
    while len(centres) < ndisk:
        ang, rad = rng.uniform(0, 2 * np.pi), radius * np.sqrt(rng.uniform(0, 0.8))
        p = (c + rad * np.cos(ang), c + rad * np.sin(ang))
        if all((p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2 > (2 * rdisk) ** 2 for q in centres):
            centres.append(p)
    disks = [("disk", [x, y, c, rdisk, 0, 0, 1, 20], "d%d" % k) for k, (x, y) in enumerate(centres)]

    def in_sphere(n):
        v = rng.standard_normal((n, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        return c + v * (radius * 0.999 * rng.uniform(0, 1, n) ** (1 / 3))[:, None]

    return dict(low=[0.0] * 3, high=[side] * 3, walls="rrr", dt=0.002, seed=seed,
                species={"A": 2.0, "B": 2.0, "C": 1.0},
                reactions=[("fwd", ["A", "B"], ["C"], 21.42267826)], placement={},
                surfaces=[("ball", [("both", "all", "reflect")], [], [("sph", None, [c, c, c, radius, 20, 20], "s0")]),
                          ("membrane", [("both", "B", "reflect"), ("both", "C", "reflect")], [],
                           [("rect", "+2", [0.0, 0.0, c, side, side], "plane")]),
                          ("patches", [("both", "B", "reflect"), ("both", "C", "reflect")],
                           [("A", "soln", "fsoln", "front", 0.2, True, None),
                            ("A", "soln", "bsoln", "back", 0.2, True, None),
                            ("A", "front", "front", "fsoln", 0.1, False, None),
                            ("A", "back", "back", "bsoln", 0.1, False, None)],
                           [(sh, None, prm, nm) for sh, prm, nm in disks])],
                molecules_at=[("A", in_sphere(n_per_species)), ("B", in_sphere(n_per_species))])


def surfaces_leg(steps, warmup):
    """The surfaces workload on the device: throughput and kernel times."""
    import smoldyn_amd as sa
    spec = c3_surfaces_spec()
    sim = sa.Simulation(spec["low"], spec["high"])
    for d, w in enumerate(spec["walls"]):
        sim.boundary(d, 0, w)
        sim.boundary(d, 1, w)
    sim.times(0.0, 1e30, spec["dt"])
    sim.seed(spec["seed"])
    for name, d in spec["species"].items():
        sim.species(name)
        sim.mobility(name, d)
    for name, rct, prd, rate in spec["reactions"]:
        sim.reaction(name, rct, prd, rate)
    for name, actions, rates, panels in spec["surfaces"]:
        sim.add_surface(name)
        for face, sp, act in actions:
            sim.surface_action(name, face, sp, act)
        for sp, st, st1, st2, value, internal, newsp in rates:
            sim.surface_rate(name, sp, st, st1, st2, value, newsp, internal)
        for shape, axis, prm, pname in panels:
            sim.add_panel(name, shape, prm, axis, pname)
    for sp, pos in spec["molecules_at"]:
        sim.add_molecules_at(sp, pos)
    t0 = time.perf_counter()
    sim.update()
    setup = time.perf_counter() - t0
    sim.run_steps(warmup)
    d0 = sim.diagnostics()
    t0 = time.perf_counter()
    sim.run_steps(steps)
    wall = time.perf_counter() - t0
    d1 = sim.diagnostics()
    if d1["err"]:
        raise SystemExit("device engine error bits 0x%x (surfaces workload)" % d1["err"])
    kt = sim.profile_steps(100)
    times = dict(zip(KERNEL_NAMES, [float(x) for x in kt]))
    net = dict(times)  # each kernel's own dispatch start -> end
    return {"workload": "500k molecules (250k A + 250k B) in a reflecting sphere r=200, transmitting membrane "
                        "with 64 disk patches (A adsorbs with probability 0.2 per crossing, desorbs at rate 0.1), "
                        "A+B->C a=0.5, dt 0.002",
            "value": (d1["molsteps"] - d0["molsteps"]) / wall, "unit": "molecule-steps/s",
            "ms_per_step": wall * 1e3 / steps, "timed_steps": steps, "kernel_ms": net, "setup_s": setup,
            "counts": {s: sim.count(s, sa.MSall) for s in spec["species"]},
            "bound_A": sim.count("A", sa.MSfront) + sim.count("A", sa.MSback)}


def build_member(spec, molperbox=None, update=False):
    import smoldyn_amd as sa
    sim = sa.Simulation(spec["low"], spec["high"])
    for d, w in enumerate(spec["walls"]):
        sim.boundary(d, 0, w)
        sim.boundary(d, 1, w)
    sim.times(0.0, 1e30, spec["dt"])
    sim.seed(spec["seed"])
    for name, d in spec["species"].items():
        sim.species(name)
        sim.mobility(name, d)
    if molperbox:
        sim.partitions("molperbox", molperbox)
    for name, rct, prd, rate in spec["reactions"]:
        sim.reaction(name, rct, prd, rate)
    for name, (kind, param) in spec["placement"].items():
        sim.product_placement(name, kind, param)
    for sp, n in spec["molecules"]:
        sim.add_molecules(sp, n)
    if update:
        sim.update()
    return sim


def ensemble_leg(members, steps, warmup, tune=True, sim_steps=2000, plain=True):
    """Batched sweep (one set of launches per step for all members) against
    the plain-GPU form (the same members, each its own engine, run one after
    another), both with HIP-graph stepping.  Throughput in simulations per
    hour for sim_steps-step simulations, from a timed window of `steps`."""
    import smoldyn_amd as sa
    from smoldyn_amd.autotune import size_class, tune_box_size
    specs = [c4_spec(k) for k in range(members)]
    counts = [sum(n for _, n in sp["molecules"]) for sp in specs]
    best, timing = ({}, {})
    if tune:
        best, timing = tune_box_size(lambda k, m: build_member(specs[k], m), range(members), counts)
    log("autotune: %s" % best)
    mpb = [best.get(size_class(c)) for c in counts]
    sims = [build_member(specs[k], mpb[k]) for k in range(members)]
    t0 = time.perf_counter()
    ens = sa.Ensemble(sims)
    setup = time.perf_counter() - t0
    ens.run_steps(warmup)
    m0 = sum(ens.diagnostics(k)["molsteps"] for k in range(members))
    t0 = time.perf_counter()
    ens.run_steps(steps)
    t_batch = time.perf_counter() - t0
    molsteps = sum(ens.diagnostics(k)["molsteps"] for k in range(members)) - m0
    err = max(ens.diagnostics(k)["err"] for k in range(members))
    del ens
    del sims
    log("batched %.2f ms/step; plain-GPU members next" % (t_batch * 1e3 / steps))
    t_plain = 0.0
    for k in range(members if plain else 0):
        sim = build_member(specs[k], mpb[k], update=True)
        sim.run_steps(warmup)
        t0 = time.perf_counter()
        sim.run_steps(steps)
        t_plain += time.perf_counter() - t0
        del sim
    scale = sim_steps / steps
    sph = members * 3600.0 / (t_batch * scale)
    sph_plain = members * 3600.0 / (t_plain * scale) if plain else None
    return {"members": members, "molecules": int(sum(counts)), "timed_steps": steps, "sim_steps": sim_steps,
            "sims_per_hour": sph, "plain_gpu_sims_per_hour": sph_plain,
            "speedup_vs_plain_gpu": sph / sph_plain if plain else None,
            "batched_ms_per_step": t_batch * 1e3 / steps,
            "plain_gpu_ms_per_member_step": t_plain * 1e3 / steps / members if plain else None,
            "molecule_steps_per_s": molsteps / t_batch, "setup_s": setup, "err": err,
            "molperbox_by_size_class": {str(c): v for c, v in best.items()},
            "autotune_ms_per_step": {str(c): {str(m): t for m, t in tv.items()} for c, tv in timing.items()}}


def c5_spec(k, seed=700, dt=0.01, n=10000, side=100.0):
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


def sweep_leg(members, steps, warmup, sim_steps=2000):
    """Learned-surrogate sweep: `members` parameter sets batched in one
    ensemble whose binding radii come from the surrogate (judged by the
    reaction-parameter solver, bisection as fallback).  Reports the batched
    throughput and what the surrogate gains on the sweep's inversions."""
    import smoldyn_amd as sa
    from smoldyn_amd import surrogate as sg
    specs = [c5_spec(k) for k in range(members)]
    model = sg.RateSurrogate.load()
    ratedt = np.array([sp["reactions"][0][3] * sp["dt"] for sp in specs])
    step = np.array([sp["step"] for sp in specs])
    ratio = np.array([sp["ratio"] for sp in specs])
    model.invert(ratedt[:64], step[:64], ratio[:64])  # device initialization
    a_s, ok, ts = model.invert(ratedt, step, ratio)
    t0 = time.perf_counter()
    a_n = sg.invert_numerical(ratedt, step, ratio)
    t_bisect = time.perf_counter() - t0
    err_s = sg.rate_error(ratedt, step, ratio, a_s)
    err_n = sg.rate_error(ratedt, step, ratio, a_n)
    log("sweep inversions: surrogate %.2f ms, bisection %.2f ms" % (ts["total_s"] * 1e3, t_bisect * 1e3))
    model.install(1e-3)
    try:
        t0 = time.perf_counter()
        sims = [build_member(sp) for sp in specs]
        ens = sa.Ensemble(sims)
        setup = time.perf_counter() - t0
        acc, rej = sg.RateSurrogate.stats()
    finally:
        sg.RateSurrogate.uninstall()
    log("sweep ensemble built in %.1f s" % setup)
    ens.run_steps(warmup)
    m0 = sum(ens.diagnostics(k)["molsteps"] for k in range(members))
    t0 = time.perf_counter()
    ens.run_steps(steps)
    t_batch = time.perf_counter() - t0
    molsteps = sum(ens.diagnostics(k)["molsteps"] for k in range(members)) - m0
    err = max(ens.diagnostics(k)["err"] for k in range(members))
    del ens, sims
    return {"members": members, "molecules": int(sum(sum(n for _, n in sp["molecules"]) for sp in specs)),
            "timed_steps": steps, "batched_ms_per_step": t_batch * 1e3 / steps,
            "sims_per_hour": members * 3600.0 / (t_batch * sim_steps / steps), "sim_steps": sim_steps,
            "molecule_steps_per_s": molsteps / t_batch, "setup_s": setup, "err": err,
            "setup_surrogate_accepted": acc, "setup_surrogate_rejected": rej,
            "inversion": {"queries": members, "surrogate_ms": ts["total_s"] * 1e3,
                          "surrogate_network_kernel_ms": ts["network_kernel_ms"],
                          "surrogate_accepted_fraction": float(ok.mean()), "bisection_ms": t_bisect * 1e3,
                          "speedup": t_bisect / ts["total_s"],
                          "median_rate_error_surrogate": float(np.median(err_s)),
                          "median_rate_error_bisection": float(np.median(err_n)),
                          "max_rate_error_surrogate": float(err_s.max()),
                          "max_rate_error_bisection": float(err_n.max())}}


def commands_leg(spec, steps, every=10, warmup=20):
    """The same workload through smolRunSim with a 'N every molcount' command
    (species counts recorded on the device between steps, rows written when
    the run returns): wall-clock ms per step against graph-replayed steps
    with no commands, and the last row against smolGetMoleculeCount."""
    import smoldyn_amd as sa
    sim = build_device(spec)
    sim.run_steps(warmup)
    t = 0.0
    for _ in range(warmup):
        t += spec["dt"]
    t_start = t
    for _ in range(steps):
        t += spec["dt"]
    _check = sa._check
    _check(sim.L.smolSetTimeStop(sim.p, t), "smolSetTimeStop")
    sim.add_output_data("counts")
    sim.add_command("molcount counts", "N", step=every)
    t0 = time.perf_counter()
    code = sim.run()
    wall = time.perf_counter() - t0
    rows = sim.output_data("counts")
    last = [sim.count(name, sa.MSall) for name in spec["species"]]
    t1 = time.perf_counter()
    sim.run_steps(steps)
    wall_plain = time.perf_counter() - t1
    return {"steps": steps, "every": every, "rows": int(rows.shape[0]), "code": int(code),
            "first_time": float(rows[0][0]) if len(rows) else None, "start_time": t_start,
            "ms_per_step": wall / steps * 1e3, "ms_per_step_no_commands": wall_plain / steps * 1e3,
            "last_row_equals_counts": bool(len(rows) and list(rows[-1][1:]) == [float(x) for x in last])}


def precision_leg(spec, steps, warmup, check_steps=100):
    """FP32 box-offset records on the same workload: throughput and per-kernel
    times, then the error against FP64 records from the same start: the max
    position difference after one step with the same injected deviates
    (bound: two roundings of an offset, 2 h 2^-24 for box size h), and the
    reaction events of `check_steps` steps with the same Philox streams."""
    sim = build_device(spec, 32)
    sim.run_steps(warmup)
    d0 = sim.diagnostics()
    t0 = time.perf_counter()
    sim.run_steps(steps)
    wall = time.perf_counter() - t0
    d1 = sim.diagnostics()
    if d1["err"]:
        raise SystemExit("device engine error bits 0x%x (FP32 records)" % d1["err"])
    kt = sim.profile_steps(100)
    times = dict(zip(KERNEL_NAMES, [float(x) for x in kt]))
    net = dict(times)  # each kernel's own dispatch start -> end
    value = (d1["molsteps"] - d0["molsteps"]) / wall
    del sim
    s64, s32 = build_device(spec, 64), build_device(spec, 32)
    _, _, _, bsize = s64.boxes()
    h = float(np.max(bsize))
    n = sum(len(s64.molecules(ll)["ident"]) for ll in range(s64.nlists()))
    g = np.random.default_rng(5).standard_normal((n, 3))
    s64.inject_step(g)
    s32.inject_step(g)
    err1 = 0.0
    for ll in range(s64.nlists()):
        a, b = s64.molecules(ll), s32.molecules(ll)
        if len(a["ident"]) and np.array_equal(a["serno"], b["serno"]):
            err1 = max(err1, float(np.abs(a["pos"] - b["pos"]).max()))
        elif len(a["ident"]):
            err1 = float("inf")
    same, diverged = 0, False
    tot = np.zeros((2, 2))
    for _ in range(check_steps):
        s64.run_time_step()
        s32.run_time_step()
        l64, l32 = s64.reaction_log(), s32.reaction_log()
        if not diverged and sorted(map(tuple, l64[:, 2:].tolist())) == sorted(map(tuple, l32[:, 2:].tolist())):
            same += 1
        else:
            diverged = True
        for j, lg in enumerate((l64, l32)):
            tot[j, 0] += np.sum(lg[:, 0] == 2)
            tot[j, 1] += np.sum(lg[:, 0] == 1)
    rate_err = [float(abs(tot[1, c] - tot[0, c]) / tot[0, c]) if tot[0, c] else None for c in range(2)]
    return {"value": value, "unit": "molecule-steps/s", "ms_per_step": wall * 1e3 / steps, "kernel_ms": net,
            "speedup_vs_fp64": None, "box_size": h,
            "max_position_error_one_step": err1, "position_error_bound_one_step": 2 * h * 2.0 ** -24,
            "identical_event_steps": same, "checked_steps": check_steps,
            "events_fp64": {"forward": int(tot[0, 0]), "reverse": int(tot[0, 1])},
            "events_fp32": {"forward": int(tot[1, 0]), "reverse": int(tot[1, 1])},
            "rate_rel_diff": {"forward": rate_err[0], "reverse": rate_err[1]}}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by their reading on a plain copy),
    or None when no summary covers it."""
    for path in sorted((ROOT / "profiles").glob("*.md")):
        for line in path.read_text().splitlines():
            line = line.strip()
            if line.startswith("TRAFFIC " + kernel + " "):
                rec = json.loads(line.split(" ", 2)[2])
                rec["source"] = "profiles/" + path.name
                return rec
    return None


def cpu_baseline(spec, steps):
    """The CPU oracle (a restatement of the reference CPU path, 1 thread) on the
    same workload: setup excluded, `steps` steps timed."""
    sys.path.insert(0, str(ROOT / "tests"))
    import simspec as ss
    osim, _ = ss.build_oracle(dict(spec, tmax=1e30))
    nmol = sum(len(x["ident"]) for x in ss.oracle_state(osim))
    t0 = time.perf_counter()
    molsteps = 0
    for _ in range(steps):
        molsteps += nmol
        osim.step()
        nmol = sum(osim.L.osim_list_count(osim.p, ll) for ll in range(osim.nlist()))
    dt = time.perf_counter() - t0
    return molsteps / dt, dt


# The survey's own timings of the compiled reference (1 core, Xeon @2.70 GHz,
# gcc 11.4 -O3) on two probe models, in molecule-steps/s, and the oracle's
# throughput on the same models relative to them (bench.py --calibrate-cpu on
# the same kind of host core): 1.06 on the C1-like model, 0.62 on the
# reversible workload scaled to 100k molecules.
SURVEY_REFERENCE_PROBES = {"c1_like": 0.99e6, "c2_scaled_100k": 0.52e6}
ORACLE_OVER_REFERENCE = {"c1_like": 1.06, "c2_scaled_100k": 0.62}


def calibrate_cpu():
    """The CPU oracle on the survey's two probe models (the reference's
    timings of them are SURVEY_REFERENCE_PROBES): C1-like, 10k A + 10k B in a
    reflective cube of side 100, 1000 steps; and the reversible workload
    scaled to 100k molecules (periodic cube of side 464.16), 100 steps.
    Prints the oracle's molecule-steps/s and its ratio to the reference's."""
    sys.path.insert(0, str(ROOT / "tests"))
    import simspec as ss
    probes = {"c1_like": (ss.c1_spec(n=10000, seed=11), 1000),
              "c2_scaled_100k": (c2_spec(n_per_species=50000, side=464.16), 100)}
    out = {}
    for name, (spec, steps) in probes.items():
        osim, _ = ss.build_oracle(dict(spec, tmax=1e30))
        nmol = sum(len(x["ident"]) for x in ss.oracle_state(osim))
        molsteps = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            molsteps += nmol
            osim.step()
            nmol = sum(osim.L.osim_list_count(osim.p, ll) for ll in range(osim.nlist()))
        dt = time.perf_counter() - t0
        v = molsteps / dt
        out[name] = {"steps": steps, "seconds": dt, "oracle_molecule_steps_per_s": v,
                     "reference_molecule_steps_per_s": SURVEY_REFERENCE_PROBES[name],
                     "oracle_over_reference": v / SURVEY_REFERENCE_PROBES[name]}
        log("%s: oracle %.3g molecule-steps/s over %d steps (%.1f s)" % (name, v, steps, dt))
    print(json.dumps(out))


def calib_copy(mib, reps=5):
    """A plain device-to-device copy of known size: reads and writes `mib` MiB
    per launch.  Profiling counters read on it give the scale factor between
    counter bytes and true HBM bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calib_copy_mib": mib, "launches": reps, "read_bytes": n * 4, "write_bytes": n * 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--steady-steps", type=int, default=500,
                    help="timed steps of the steady-state window, after 1000 steps from setup (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--n", type=int, default=500000, help="molecules per reactant species")
    ap.add_argument("--ensemble-members", type=int, default=256,
                    help="members of the batched parameter sweep (0: skip the ensemble leg)")
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--ensemble-batched-only", action="store_true",
                    help="skip the one-engine-per-member comparison (profiling runs)")
    ap.add_argument("--surface-steps", type=int, default=1000,
                    help="timed steps of the surfaces workload (0: skip)")
    ap.add_argument("--command-steps", type=int, default=2000,
                    help="steps of the smolRunSim leg with a molcount command (0 skips it)")
    ap.add_argument("--precision-steps", type=int, default=2000,
                    help="timed steps of the FP32 box-offset leg (0: skip)")
    ap.add_argument("--sweep-members", type=int, default=4096,
                    help="members of the learned-surrogate sweep, 1e4 molecules each (0: skip)")
    ap.add_argument("--sweep-steps", type=int, default=100)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--calibrate-cpu", action="store_true",
                    help="only time the CPU oracle on the survey's probe models (ratio to the reference's timings)")
    args = ap.parse_args()
    if args.calibrate_cpu:
        calibrate_cpu()
        return
    if args.calib_copy_mib > 0:
        calib_copy(args.calib_copy_mib)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import smoldyn_amd as sa
    spec = c2_spec(n_per_species=args.n)
    log("setting up the 1M-molecule workload")
    t_setup = time.perf_counter()
    sim = build_device(spec)
    setup_s = time.perf_counter() - t_setup
    use_graph = not args.no_graph
    if args.warmup > 0:
        sim.run_steps(args.warmup, graph=use_graph)
    d0 = sim.diagnostics()
    # timed region: all steps captured as HIP-graph replays on one stream; the
    # device accumulates the live molecule count at the start of every step
    log("timing %d steps" % args.steps)
    t0 = time.perf_counter()
    dev_ms = sim.run_steps(args.steps, graph=use_graph)
    wall = time.perf_counter() - t0
    d1 = sim.diagnostics()
    if d1["err"]:
        raise SystemExit("device engine error bits 0x%x" % d1["err"])
    molsteps = d1["molsteps"] - d0["molsteps"]
    value = molsteps / wall
    # per-kernel device time: each kernel's own start / stop HIP events
    # (stamped at its dispatch's start and end) on the engine's stream
    kt = sim.profile_steps(args.profile_steps) if args.profile_steps > 0 else np.zeros(len(KERNEL_NAMES))
    nmol = sim.diagnostics()["nmol"]
    times = dict(zip(KERNEL_NAMES, [float(x) for x in kt]))
    net = dict(times)
    # the roofline's kernel: the longest of the per-molecule (HBM-streaming)
    # kernels; the per-event kernels (exact test, resolve) have no
    # per-molecule bytes and are reported beside it
    longest = max(net, key=lambda k: net[k])
    dom = max((k for k in net if ALGO_BYTES[k] > 0), key=lambda k: net[k])
    achieved = ALGO_BYTES[dom] * nmol / (net[dom] * 1e-3) / 1e9 if net[dom] > 0 else 0.0
    step_bytes = STEP_ALGO_BYTES * nmol
    counts = {s: sim.count(s) for s in ("A", "B", "C")}
    tr = measured_traffic(HIP_KERNEL_SYMBOL[dom])
    per_kernel = {k: {"ms": net[k], "algo_bytes_per_molecule": ALGO_BYTES[k],
                      "design_bytes_per_molecule": DESIGN_BYTES[k],
                      "algo_GBs": ALGO_BYTES[k] * nmol / (net[k] * 1e-3) / 1e9 if net[k] > 0 else 0.0,
                      "frac": (ALGO_BYTES[k] * nmol / (net[k] * 1e-3) / 1e9 / HBM_PEAK_GBS) if net[k] > 0 else 0.0}
                  for k in net}
    # steady state: continue to 1000 steps after setup, then time 500 more
    steady = None
    if args.steady_steps > 0:
        done = args.warmup + args.steps + args.profile_steps
        if done < 1000:
            sim.run_steps(1000 - done, graph=use_graph)
            done = 1000
        s0 = sim.diagnostics()
        t0 = time.perf_counter()
        sim.run_steps(args.steady_steps, graph=use_graph)
        sw = time.perf_counter() - t0
        s1 = sim.diagnostics()
        steady = {"after_steps": done, "timed_steps": args.steady_steps,
                  "value": (s1["molsteps"] - s0["molsteps"]) / sw, "ms_per_step": sw * 1e3 / args.steady_steps,
                  "live_molecules": s1["nmol"]}
    out = {
        "metric": "molecule-steps/s at 1M molecules",
        "value": value,
        "unit": "molecule-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "reversible A+B<->C, periodic cube 1000, 500k A + 500k B, dt 0.005, "
                               "a~1.0, b/a 1.2, ~4 molecules per virtual box, FP64",
                   "molecules": 2 * args.n, "graph": use_graph},
        "device_ms_per_step": dev_ms / args.steps,
        "kernel_ms": net,
        "roofline": {"bound": "hbm", "kernel": HIP_KERNEL_SYMBOL[dom], "kernel_ms": net[dom],
                     "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "algo_bytes_per_molecule": ALGO_BYTES[dom], "design_bytes_per_molecule": DESIGN_BYTES[dom],
                     "molecules_per_launch": nmol,
                     "traffic": tr["bytes_per_launch"] if tr else None,
                     "traffic_note": ("rocprofv3 FETCH_SIZE+WRITE_SIZE per launch at %d molecules, scaled by a "
                                      "plain-copy calibration (%s)" % (tr["molecules"], tr["source"])) if tr else None,
                     "longest_kernel": {"kernel": HIP_KERNEL_SYMBOL[longest], "ms": net[longest],
                                        "bound": "hbm" if ALGO_BYTES[longest] > 0 else
                                        "latency: per-event work, no per-molecule bytes"}},
        "kernels": per_kernel,
        "step_algo_bytes": step_bytes,
        "step_frac_of_hbm": step_bytes / (dev_ms / args.steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "steady_state": steady,
        "final_counts": counts,
        "setup_s": setup_s,
    }
    if args.surface_steps > 0:
        log("surfaces workload, %d steps" % args.surface_steps)
        out["surfaces"] = surfaces_leg(args.surface_steps, 20)
    if args.command_steps > 0:
        log("run-time commands: molcount every 10 steps through smolRunSim, %d steps" % args.command_steps)
        out["commands"] = commands_leg(spec, args.command_steps)
    if args.precision_steps > 0:
        log("FP32 box-offset records, %d steps" % args.precision_steps)
        out["fp32_offsets"] = precision_leg(spec, args.precision_steps, args.warmup)
        out["fp32_offsets"]["speedup_vs_fp64"] = out["fp32_offsets"]["value"] / value
    if args.ensemble_members > 0:
        del sim
        log("batched ensemble leg, %d members" % args.ensemble_members)
        out["ensemble"] = ensemble_leg(args.ensemble_members, args.ensemble_steps, 20, tune=not args.no_autotune,
                                       plain=not args.ensemble_batched_only)
    if args.sweep_members > 0:
        if "sim" in locals():
            del sim
        log("learned-surrogate sweep leg, %d members" % args.sweep_members)
        out["surrogate_sweep"] = sweep_leg(args.sweep_members, args.sweep_steps, 10)
    if not args.no_cpu:
        log("CPU oracle baseline, %d steps" % args.cpu_steps)
        cv, cdt = cpu_baseline(spec, args.cpu_steps)
        ratio = ORACLE_OVER_REFERENCE["c2_scaled_100k"]
        out["cpu_baseline"] = {"value": cv, "unit": "molecule-steps/s", "cores": 1, "kind": "port",
                               "sample": "same 1M-molecule workload, %d steps after setup (%.1f s), CPU oracle "
                                         "(C restatement of the reference path), 1 thread" % (args.cpu_steps, cdt),
                               "calibration": {"oracle_over_reference": ratio,
                                               "probe": "reversible workload scaled to 100k molecules, against the "
                                                        "compiled reference's 0.52e6 molecule-steps/s",
                                               "reference_estimate": cv / ratio}}
        out["speedup_vs_cpu"] = value / cv
        out["speedup_vs_reference_estimate"] = value / (cv / ratio)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
