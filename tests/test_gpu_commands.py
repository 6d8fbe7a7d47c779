"""Run-time commands on the device: molcount rows recorded between steps
without host syncs equal the oracle's species counts after the same steps,
and the reference's own command tests (tests/test_data_in_array.py,
tests/test_data_outputfile.py) restated on the device engine."""
import numpy as np
import pytest

import oracle_bind as ob
import simspec as ss
import smoldyn_amd as sa

pytestmark = pytest.mark.gpu


def _steps_to(tmin, tmax, dt):
    t, n = tmin, 0
    while True:
        t += dt
        n += 1
        if t >= tmax:
            return n


def _oracle_counts(spec, nsteps):
    osim, _ = ss.build_oracle(spec)
    names = list(spec["species"])
    out = []
    for k in range(nsteps + 1):
        if k:
            osim.step()
        out.append([osim.count(osim.index(n)) for n in names])
    return out


def test_molcount_rows_equal_oracle_counts():
    """Reversible A+B<->C (periodic, ratio 1.2 unbinding) at 10k molecules:
    'N 3 molcount d' and 'E molcountspecieslist s A(all) C' through smolRunSim;
    every row (time and counts) equals the oracle's state after that many
    steps, the times bit for bit."""
    dt = 0.005
    spec = ss.c2_spec(n=5000, side=1000.0 * (10000 / 1e6) ** (1 / 3), kback=20.0, seed=13, tmax=40 * dt)
    dsim = ss.build_device(spec, update=False)
    dsim.add_output_data("d")
    dsim.add_output_data("s")
    dsim.add_command("molcount d", "N", step=3)
    dsim.add_command("molcountspecieslist s A(all) C", "E")
    assert dsim.run() == sa.ECok
    d = dsim.output_data("d")
    s = dsim.output_data("s")
    t, c, it = ob.command_timeline("times 0 %r %r\ncmd N 0 0 3 0 a\ncmd E 0 0 0 0 b\nrun\n" % (spec["tmax"], dt))
    nsteps = _steps_to(0.0, spec["tmax"], dt)
    counts = _oracle_counts(spec, nsteps)
    exp_d = [[t[j]] + counts[min(it[j], nsteps)] for j in range(len(t)) if c[j] == 0]
    exp_s = [[t[j], counts[min(it[j], nsteps)][0], counts[min(it[j], nsteps)][2]] for j in range(len(t)) if c[j] == 1]
    assert d.shape == (len(exp_d), 4) and s.shape == (len(exp_s), 3)
    assert np.array_equal(d, np.array(exp_d, dtype=np.float64))
    assert np.array_equal(s, np.array(exp_s, dtype=np.float64))
    assert d[-1][1] < 5000  # reactions happened


def test_single_steps_record_before_and_after():
    """smolRunTimeStep runs the commands before and after its step: rows at
    steps 0,1,1,2,2,3, each equal to the oracle's counts."""
    dt = 0.01
    spec = ss.c1_spec(n=3000, seed=17, tmax=1.0)
    dsim = ss.build_device(spec, update=False)
    dsim.add_output_data("d")
    dsim.add_command("molcount d", "E")
    for _ in range(3):
        dsim.run_time_step()
    d = dsim.output_data("d")
    counts = _oracle_counts(spec, 3)
    clock = [0.0, dt, dt + dt, dt + dt + dt]
    steps = [0, 1, 1, 2, 2, 3]
    assert np.array_equal(d, np.array([[clock[k]] + counts[k] for k in steps], dtype=np.float64))


def test_stop_command_ends_the_run():
    spec = ss.c1_spec(n=2000, seed=19, tmax=1.0)
    dsim = ss.build_device(spec, update=False)
    dsim.add_output_data("d")
    dsim.add_command("molcount d", "E")
    dsim.add_command("stop", "@", on=0.25)
    assert dsim.run() == sa.ECok
    d = dsim.output_data("d")
    clock = [0.0]
    while clock[-1] < 0.25:
        clock.append(clock[-1] + 0.01)
    k = len(clock) - 1
    # rows at steps 0..k and once more in the end-of-run pass
    assert list(d[:, 0]) == clock + [clock[-1]]
    assert dsim.diagnostics()["step"] == k


def test_reference_data_in_array_model():
    """tests/test_data_in_array.py: R -> 0 at rate 0.1 for 400 R plus one B
    at the centre in a reflective cube of side 100, 'E molcount mydata',
    stop 10, dt 0.01.  1001 rows of 3 columns, first row [0, 400, 1]; the
    last row at time 10 with R near 400 exp(-1) (the reference's SFMT draws
    give 142; the counter stream gives another draw of the same law)."""
    s = sa.Simulation([0, 0, 0], [100, 100, 100])
    for d in range(3):
        s.boundary(d, 0, "r")
        s.boundary(d, 1, "r")
    s.times(0.0, 10.0, 0.01)
    s.seed(100)
    for name in ("R", "B"):
        s.species(name)
        s.mobility(name, 1.0)
    s.reaction("r1", ["R"], [], 0.1)
    s.add_molecules("R", 400)
    s.add_molecules("B", 1, [50, 50, 50], [50, 50, 50])
    s.add_output_data("mydata")
    s.add_command("molcount mydata", "E")
    assert s.run() == sa.ECok
    data = s.output_data("mydata", erase=False)
    assert data.shape == (1001, 3)
    assert list(data[0]) == [0.0, 400.0, 1.0]
    assert abs(data[-1][0] - 10.0) < 1e-3 * 10
    p = np.exp(-1.0)
    sd = np.sqrt(400 * p * (1 - p))
    assert abs(data[-1][1] - 400 * p) < 4 * sd, data[-1]
    assert np.all(np.diff(data[:, 1]) <= 0) and np.all(data[:, 2] == 1)


def test_reference_data_outputfile_model(tmp_path):
    """tests/test_data_outputfile.py: C <-> A + B (kf 0.1, kb 100) in a
    periodic cube of side 100 with boxes of 10, 1000 A and 1000 B, 'i 0 100 10
    molcount box.dat', dt 0.1 to 100: an 11 x 4 file whose column means are
    [50.036, 590.727, 590.727, 409.272] within 10 %."""
    s = sa.Simulation([0, 0, 0], [100, 100, 100])
    for d in range(3):
        s.boundary(d, 0, "p")
        s.boundary(d, 1, "p")
    s.times(0.0, 100.0, 0.1)
    s.seed(7)
    s.partitions("boxsize", 10.0)
    for name in ("A", "B", "C"):
        s.species(name)
        s.mobility(name, 1.0)
    s.reaction("r1fwd", ["C"], ["A", "B"], 0.1)
    s.reaction("r1rev", ["A", "B"], ["C"], 100.0)
    s.add_molecules("A", 1000)
    s.add_molecules("B", 1000)
    s.set_output_path(str(tmp_path) + "/")
    s.add_output_file("box.dat")
    s.add_command("molcount box.dat", "i", on=0, off=100, step=10)
    assert s.run() == sa.ECok
    s.close()
    data = np.loadtxt(tmp_path / "box.dat")
    assert data.shape == (11, 4)
    expected = [50.036, 590.727, 590.727, 409.272]
    assert np.isclose(data.mean(axis=0), expected, atol=1e-1, rtol=1e-1).all(), data.mean(axis=0)
    assert round(float(data[:, 0].mean()), 3) == 50.036


def test_listmols_equals_oracle_listmols(tmp_path):
    """'N 5 listmols lm.txt' (smolcmd.c cmdlistmols) on the device writes the
    same text, byte for byte, as the oracle's listmols (whose format is pinned
    by the reference's S95 regression outputs) after the same steps; the data
    table of 'listmols2' at iteration 7 holds the same molecules."""
    dt = 0.005
    spec = ss.c2_spec(n=1500, side=1000.0 * (3000 / 1e6) ** (1 / 3), kback=20.0, seed=23, tmax=12 * dt)
    dsim = ss.build_device(spec, update=False)
    dsim.set_output_path(str(tmp_path) + "/")
    dsim.add_output_file("lm.txt")
    dsim.add_output_data("d")
    dsim.add_command("listmols lm.txt", "N", step=5)
    dsim.add_command("listmols2 d", "&", on=7)
    assert dsim.run() == sa.ECok
    d = dsim.output_data("d")
    dsim.close()
    osim, _ = ss.build_oracle(spec)
    text, rows7 = "", None
    for k in range(11):
        if k:
            osim.step()
        if k % 5 == 0:
            text += osim.listmols()
        if k == 7:
            rows7 = []
            for ll in range(osim.nlist()):
                L = osim.get_list(ll)
                for j in range(len(L["ident"])):
                    rows7.append([1, L["ident"][j], L["mstate"][j], *L["pos"][j], L["serno"][j]])
    assert (tmp_path / "lm.txt").read_text() == text
    assert np.array_equal(d, np.array(rows7, dtype=np.float64))


def _serno(s):
    """molserno2string (smolmolec.c:625-630)"""
    s = int(s)
    return "%d" % s if s < 0xFFFFFFFF else "%d.%d" % (s >> 32, s & 0xFFFFFFFF)


@pytest.mark.parametrize("serials", [None, "list"])
def test_reaction_log_lines_equal_oracle_events(tmp_path, serials):
    """reaction_log (RxnSetLog smolreact.c:2013-2061, doreact's log lines
    smolreact.c:2879-2896) through smolRunSim on the reversible workload at
    6000 molecules: every logged line is "time name x y z reactant-serials
    [?] product-serials" of an event the oracle makes in the same step, in
    the same order, with the step's start time, the reaction position and the
    serial numbers bit for bit (%g).  With a serial list, only events that
    make a listed product are logged, the products before the first listed
    one printed as '?'."""
    dt = 0.005
    spec = ss.c2_spec(n=3000, side=1000.0 * (6000 / 1e6) ** (1 / 3), kback=20.0, seed=29, tmax=60 * dt)
    osim, rx = ss.build_oracle(spec)
    names = {rx[k][0] * 100 + rx[k][1]: k for k in rx}
    lines, t = [], 0.0
    allrows = []
    for k in range(_steps_to(0.0, spec["tmax"], dt)):
        osim.step()
        rows, pos = osim.rxnlog(), osim.rxnlog_pos()
        allrows.extend((t, r, p) for r, p in zip(rows, pos))
        t += dt
    listed = None
    if serials:
        prods = sorted({int(r[5]) for _, r, _ in allrows if r[0] == 1})  # second products of dissociations
        listed = prods[::3][:40]
    for t, r, p in allrows:
        order = int(r[0])
        nprod = 1 if order == 2 else 2
        prd = [int(r[4]), int(r[5])][:nprod]
        first = 0 if listed is None else next((q for q in range(nprod) if (prd[q] & 0xFFFFFFFF) in listed), -1)
        if first < 0:
            continue
        parts = ["%g" % t, names[order * 100 + int(r[1])]] + ["%g" % v for v in p]
        parts += [_serno(s) for s in r[2:2 + order]] + ["?"] * first + [_serno(s) for s in prd[first:]]
        lines.append(" ".join(parts))
    dsim = ss.build_device(spec, update=False)
    dsim.set_output_path(str(tmp_path) + "/")
    dsim.add_output_file("rxn.log")
    dsim.set_reaction_log("rxn.log", "all", listed)
    assert dsim.run() == sa.ECok
    dsim.close()
    got = (tmp_path / "rxn.log").read_text().splitlines()
    assert len(lines) > (20 if listed is None else 5)
    assert got == lines


def test_c_caller_runs_the_reference_way_and_matches_the_oracle(tmp_path):
    """tests/libsmoldyn_caller.c (a C program built with gcc against
    include/libsmoldyn_amd.h) steps its model with smolRunTimeStep, runs it to
    time 1 with smolRunSimUntil and to its end with smolRunSim, and reads a
    molcount data table: its counts equal
    the oracle's on the same model after the same steps."""
    import subprocess
    from test_capi_host import build_c_caller
    exe = build_c_caller(tmp_path)
    r = subprocess.run([str(exe), "run"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    steps, final, until, rows = {}, None, None, []
    lines = r.stdout.splitlines()
    for k, ln in enumerate(lines):
        w = ln.split()
        if w[0] == "step":
            steps[int(w[1])] = tuple(int(x) for x in w[2:5])
        elif w[0] == "final":
            final = tuple(int(x) for x in w[1:4])
        elif w[0] == "until":
            until = tuple(int(x) for x in w[1:4])
        elif w[0] == "data":
            rows = [[float(x) for x in lines[k + 1 + j].split()] for j in range(int(w[1]))]
    spec = dict(low=[0.0] * 3, high=[40.0] * 3, walls="rrr", dt=0.01, seed=17, molperbox=4,
                species={"A": 1.0, "B": 1.0, "C": 0.5},
                reactions=[("fwd", ["A", "B"], ["C"], 20.0), ("back", ["C"], ["A", "B"], 2.0)],
                placement={"back": ("ratio", 1.5)},
                surfaces=[{"name": "ball", "actions": [("both", "all", "reflect")],
                           "panels": [("sph", None, [20.0, 20.0, 20.0, 15.0, 20, 20], "s0")]}],
                molecules=[("A", 2000), ("B", 2000)])
    want = _oracle_counts(spec, 200)
    for k, c in steps.items():
        assert c == tuple(want[k]), (k, c, want[k])
    assert len(steps) == 20 and until == tuple(want[_steps_to(0.0, 1.0, 0.01)]) and final == tuple(want[200])
    assert len(rows) >= 20 and all(r_[1] + r_[3] == 2000 and r_[2] + r_[3] == 2000 for r_ in rows)
