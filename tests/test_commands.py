"""Run-time commands and output (CPU): the command scheduler of the library
against the oracle's restatement of the reference's (libSteve/SimCommand.c,
Smoldyn/smolsim.cpp smolsimulate), the oracle pinned by the reference's own
tests and by the closed form of each timing code; and the host side of the
observation commands (data tables, files, formats) through smolRunCommand,
which needs no device before the first update."""
import numpy as np
import pytest

import oracle_bind as ob
import smoldyn_amd as sa


def _acc(t0, dt, n):
    """Times of the reference's clock: t0 with dt added n times."""
    out = [t0]
    t = t0
    for _ in range(n):
        t += dt
        out.append(t)
    return out


def _execs(tl, pos=None):
    t, c, it = tl
    m = c >= 0 if pos is None else c == pos
    return t[m], it[m]


def _runcodes(tl):
    return [-1000 - int(x) for x in tl[1] if x <= -1000]


# ---- the oracle pinned by the reference's tests

def test_every_iteration_gives_1001_rows():
    """tests/test_data_in_array.py:33-36: 'E' over [0,10] at dt 0.01 gives 1001
    rows, the first at time 0, the last at time 10 (rel 1e-3)."""
    tl = ob.command_timeline("times 0 10 0.01\ncmd E 0 0 0 0 obs\nrun\n")
    t, it = _execs(tl, 0)
    assert len(t) == 1001
    assert t[0] == 0.0 and list(it) == list(range(1001))
    assert abs(t[-1] - 10.0) < 1e-3 * 10.0
    assert _runcodes(tl) == [1]


def test_interval_command_rows_and_mean_time():
    """tests/test_data_outputfile.py:33-38: 'i 0 100 10' at dt 0.1 up to 100
    gives 11 rows whose time column averages 50.036."""
    tl = ob.command_timeline("times 0 100 0.1\ncmd i 0 100 10 0 obs\nrun\n")
    t, _ = _execs(tl, 0)
    assert len(t) == 11
    assert round(float(t.mean()), 3) == 50.036


# ---- closed forms of the timing codes (oracle)

def test_every_n_iterations():
    tl = ob.command_timeline("times 0 1 0.01\ncmd N 0 0 7 0 obs\nrun\n")
    t, it = _execs(tl, 0)
    clock = _acc(0.0, 0.01, 200)
    nst = int(1 / 0.01 + 0.5)
    assert list(it) == list(range(0, nst + 1, 7))
    assert list(t) == [clock[k] for k in it]


def test_at_time_runs_once_at_the_first_time_reached():
    tl = ob.command_timeline("times 0 1 0.1\ncmd @ 0.3 0 0 0 obs\nrun\n")
    t, it = _execs(tl, 0)
    clock = _acc(0.0, 0.1, 20)
    k = next(i for i, x in enumerate(clock) if x >= 0.3)
    assert list(t) == [clock[k]] and list(it) == [k]


def test_before_and_after():
    tl = ob.command_timeline("times 0 1 0.1\ncmd b 0 0 0 0 first\ncmd B 0 0 0 0 first2\n"
                             "cmd a 0 0 0 0 last\ncmd A 0 0 0 0 last2\nrun\n")
    tb, ib = _execs(tl, 0)
    tB, iB = _execs(tl, 1)
    ta, _ = _execs(tl, 2)
    tA, _ = _execs(tl, 3)
    assert list(tb) == [0.0] and list(ib) == [0]
    assert list(tB) == [0.0] and list(iB) == [0]
    final = tl[0][-1]
    assert list(ta) == [final] and list(tA) == [final]
    # 'a' waits for the end-of-run pass: it is the last execution
    assert tl[1][-2] in (2, 3)


def test_iteration_window_and_single_iteration():
    tl = ob.command_timeline("times 0 1 0.01\ncmd I 5 40 10 0 win\ncmd & 17 0 0 0 one\nrun\n")
    _, iw = _execs(tl, 0)
    _, io = _execs(tl, 1)
    assert list(iw) == [5, 15, 25, 35]
    assert list(io) == [17]


def test_geometric_times():
    tl = ob.command_timeline("times 0 10 0.01\ncmd x 0.5 9 0.25 2 geo\nrun\n")
    t, _ = _execs(tl, 0)
    clock = _acc(0.0, 0.01, 1200)
    on, d, expect = 0.5, 0.25, []
    while True:
        expect.append(next(x for x in clock if x >= on))
        if on + d > 9:
            break
        on += d
        d *= 2
    assert list(t) == expect


def test_stop_ends_the_run_and_others_run_once_more():
    tl = ob.command_timeline("times 0 1 0.1\ncmd E 0 0 0 0 obs\ncmd @ 0.35 0 0 0 stop\nrun\n")
    t, _ = _execs(tl, 0)
    ts, _ = _execs(tl, 1)
    clock = _acc(0.0, 0.1, 20)
    k = next(i for i, x in enumerate(clock) if x >= 0.35)
    assert list(ts) == [clock[k]]
    # obs at iterations 0..k, then once more in the end-of-run pass at the same time
    assert list(t) == clock[:k + 1] + [clock[k]]
    assert _runcodes(tl) == [7]


def test_single_steps_run_commands_before_and_after():
    """smolRunTimeStep is smolsimulate with a break half a step ahead: the
    commands run before and after each step."""
    tl = ob.command_timeline("times 0 1 0.1\ncmd E 0 0 0 0 obs\nrunstep\nrunstep\nrunstep\n")
    t, it = _execs(tl, 0)
    clock = _acc(0.0, 0.1, 3)
    assert list(t) == [clock[0], clock[1], clock[1], clock[2], clock[2], clock[3]]
    assert list(it) == list(range(6))
    assert _runcodes(tl) == [10, 10, 10]


# ---- the library's scheduler equals the oracle's

TIMINGS = list("baBAEe@ixN&Ij")


def _random_script(rng):
    tmin = float(rng.choice([0.0, 1.5, -2.0]))
    dt = float(rng.choice([0.1, 0.01, 0.05, 0.3, 1.0 / 3.0, 0.125]))
    nst = int(rng.integers(5, 60))
    tmax = tmin + nst * dt + float(rng.choice([0.0, dt / 3, -dt / 3]))
    span = tmax - tmin
    lines = ["times %r %r %r" % (tmin, tmax, dt)]
    ncmd = 0

    def add():
        nonlocal ncmd
        t = str(rng.choice(TIMINGS))
        name = ("stop%d" if rng.random() < 0.08 else "obs%d") % ncmd
        on = off = step = mult = 0.0
        if t == "@":
            on = tmin + float(rng.uniform(-dt, span * 1.1))
        elif t in "ix":
            on = tmin + float(rng.uniform(-1.0, span))
            off = on + float(rng.uniform(0, span))
            step = float(rng.choice([dt / 2, dt, 2.5 * dt, span / 4]))
            if t == "x":
                mult = float(rng.choice([0.0, 1.5, 2.0]))
        elif t == "N":
            step = float(rng.integers(1, 10))
        elif t == "&":
            on = float(rng.integers(0, nst + 3))
        elif t in "Ij":
            on = float(rng.integers(-1, nst))
            off = on + float(rng.integers(0, nst))
            step = float(rng.integers(1, 6))
        if rng.random() < 0.5:
            lines.append("cmd %s %r %r %r %r %s" % (t, on, off, step, mult, name))
        else:
            nums = {"@": [on], "&": [on], "N": [step], "n": [step], "i": [on, off, step], "I": [on, off, step],
                    "j": [on, off, step], "x": [on, off, step, mult]}.get(t, [])
            lines.append("cmdstr %s %s %s" % (t, " ".join("%r" % v for v in nums), name))
        ncmd += 1

    for _ in range(int(rng.integers(1, 6))):
        add()
    for _ in range(int(rng.integers(1, 6))):
        op = rng.random()
        if op < 0.3:
            lines.append("run")
        elif op < 0.5:
            for _ in range(int(rng.integers(1, 4))):
                lines.append("runstep")
        elif op < 0.7:
            lines.append("rununtil %r" % (tmin + float(rng.uniform(0, span * 1.2))))
        elif op < 0.8:
            lines.append("tmax %r" % (tmax + float(rng.uniform(0, span))))
        elif op < 0.9:
            lines.append("dt %r" % float(rng.choice([dt, dt / 2, 2 * dt])))
        else:
            add()
    lines.append("run")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("seed", range(8))
def test_library_scheduler_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    for _ in range(60):
        script = _random_script(rng)
        to, co, io = ob.command_timeline(script)
        td, cd, idv = sa.command_timeline(script)
        assert len(to) == len(td), script
        assert np.array_equal(co, cd), script
        assert np.array_equal(io, idv), script
        assert np.array_equal(to, td), script  # bit-identical clocks


def test_scheduler_edge_codes():
    """Add-time checks of scmdaddcommand: a non-positive interval or period is
    refused, a non-integer iteration too; an unknown timing code adds nothing."""
    script = ("times 0 1 0.1\ncmd i 0 1 0 0 bad1\ncmd N 0 0 0 0 bad2\ncmd N 0 0 1.5 0 bad3\n"
              "cmd & 2.5 0 0 0 bad4\ncmd q 0 0 0 0 bad5\ncmd E 0 0 0 0 good\nrun\n")
    for tl in (ob.command_timeline(script), sa.command_timeline(script)):
        assert set(int(x) for x in tl[1] if x >= 0) == {0}  # only "good", at list position 0
        assert len(_execs(tl, 0)[0]) == 11


# ---- host side of the observation commands (no device needed)

def _sim():
    s = sa.Simulation([0, 0, 0], [10, 10, 10])
    s.times(0.0, 1.0, 0.1)
    for name in ("A", "B", "C"):
        s.species(name)
        s.mobility(name, 1.0)
    s.add_molecules("A", 7)
    s.add_molecules("B", 3)
    return s


def test_molcount_into_a_data_table():
    s = _sim()
    s.add_output_data("d")
    s.run_command("molcount d")
    s.run_command("molcountspecies B d")
    s.run_command("molcountspecieslist d A C all")
    d = s.output_data("d", erase=False)
    # rows of different widths: ncol is the widest, short rows padded with zeros
    assert d.shape == (3, 4)
    assert list(d[0]) == [0.0, 7.0, 3.0, 0.0]
    assert list(d[1][:2]) == [0.0, 3.0]
    assert list(d[2]) == [0.0, 7.0, 0.0, 10.0]
    d2 = s.output_data("d", erase=True)
    assert np.array_equal(d, d2)
    assert s.output_data("d").shape == (0, 0)


def test_output_data_errors():
    s = _sim()
    with pytest.raises(sa.SmoldynError):
        s.output_data("d")  # no data tables in the sim
    s.add_output_data("d")
    with pytest.raises(sa.SmoldynError):
        s.output_data("nosuch")
    # an unknown name is a warning recorded by the command; the call itself
    # returns ECok (libsmoldyn.cpp:795, LCHECK with ECwarning does not fail)
    sa.last_error()
    assert s.run_command("molcount nosuch") == sa.ECok
    assert sa.last_error()[0] == sa.ECwarning


def test_commands_outside_the_device_path_are_refused():
    s = _sim()
    with pytest.raises(sa.SmoldynError):
        s.add_command("molpos A out.txt", "E")
    with pytest.raises(sa.SmoldynError):
        s.add_command_str("E pause")
    s.add_command("molcount", "E")
    s.add_command_str("i 0 1 0.5 molcount")
    s.add_command_str("N 10 stop")


def test_files_header_rows_suffix_and_overwrite(tmp_path):
    s = _sim()
    s.set_output_path(str(tmp_path) + "/")
    s.add_output_file("counts.txt")
    s.add_output_file("s.dat", suffix=3)
    s.open_output_files(overwrite=False)
    s.run_command("molcountheader counts.txt")
    s.run_command("molcount counts.txt")
    s.run_command("molcountspecies A(all) s.dat")
    s.close()
    assert (tmp_path / "counts.txt").read_text() == "time A B C\n0 7 3 0\n"
    assert (tmp_path / "s_003.dat").read_text() == "0 7\n"
    s2 = _sim()
    s2.set_output_path(str(tmp_path) + "/")
    s2.add_output_file("counts.txt")
    with pytest.raises(sa.SmoldynError):
        s2.open_output_files(overwrite=False)  # exists: refused without overwrite
    s2.open_output_files(overwrite=True)
    s2.close()
    assert (tmp_path / "counts.txt").read_text() == ""


def test_file_and_data_in_one_command(tmp_path):
    s = _sim()
    s.set_output_path(str(tmp_path) + "/")
    s.add_output_file("f.txt")
    s.add_output_data("d")
    s.open_output_files(overwrite=True)
    s.run_command("molcount f.txt d")
    d = s.output_data("d")
    s.close()
    assert list(d[0]) == [0.0, 7.0, 3.0, 0.0]
    assert (tmp_path / "f.txt").read_text() == "0 7 3 0\n"


def test_listmols_formats(tmp_path):
    """cmdlistmols: "name(state) x y z serial" per molecule in list order;
    listmols2: execution count, ident, state, position, serial."""
    s = sa.Simulation([0, 0, 0], [10, 10, 10])
    s.times(0.0, 1.0, 0.1)
    for name in ("A", "B"):
        s.species(name)
        s.mobility(name, 1.0)
    s.add_molecules_at("A", [[1, 2, 3], [4.5, 5, 6]])
    s.add_molecules_at("B", [[7, 8, 9.25]])
    s.set_output_path(str(tmp_path) + "/")
    s.add_output_file("lm.txt")
    s.add_output_data("m")
    s.open_output_files(overwrite=True)
    s.run_command("listmols lm.txt")
    s.run_command("listmols2 m")
    d = s.output_data("m")
    s.close()
    assert (tmp_path / "lm.txt").read_text() == ("A(solution) 1 2 3 1\nA(solution) 4.5 5 6 2\n"
                                                 "B(solution) 7 8 9.25 3\n")
    assert d.tolist() == [[0, 1, 0, 1, 2, 3, 1], [0, 1, 0, 4.5, 5, 6, 2], [0, 2, 0, 7, 8, 9.25, 3]]
