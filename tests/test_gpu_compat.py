"""Model scripts written for the reference's Python package run on the
device engine through ``import smoldyn`` (the compat package) and give the
oracle's results bit for bit.

The models are those of the reference's examples/S15_python/cmddata.py
(R -> 0 with molcount every step), box.py (C <-> A + B with molecule lists,
box partitions and an output file) and runtimestep.py (single steps through
the low-level class), in 3-D and with a fixed seed."""
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_bind as ob
import simspec as ss
import smoldyn_amd as sa

sys.path.insert(0, str(Path(sa.__file__).resolve().parent / "compat"))
import smoldyn  # noqa: E402
import smoldyn._smoldyn as S  # noqa: E402

pytestmark = pytest.mark.gpu


def _steps_to(tmin, tmax, dt):
    t, n = tmin, 0
    while True:
        t += dt
        n += 1
        if t >= tmax:
            return n


def _oracle(high, walls, dt, seed, species, adds, reactions=()):
    """The oracle built in the order the script defines the model: species,
    reactions, then each addSolutionMolecules call with its own bounds."""
    osim = ob.OSim(3, [0.0] * 3, list(high))
    for d, w in enumerate(walls):
        osim.wall(d, 0, w)
        osim.wall(d, 1, w)
    osim.times(0.0, 1e9, dt)
    osim.rngmode(ob.RNG_COUNTER, seed)
    idx = {}
    for name, dc in species.items():
        idx[name] = osim.species(name)
        osim.difc(idx[name], dc, 0)
    for name, rct, prd, rate in reactions:
        osim.reaction(name, [idx[r] for r in rct], [idx[q] for q in prd], rate)
    for name, n, lo, hi in adds:
        osim.add_mols(n, idx[name], lo, hi)
    osim.update()
    return osim


def test_cmddata_model_rows_equal_oracle():
    """cmddata.py: 400 R decaying at 0.1/s and one B, molcount every step
    into a data table, s.run(stop=2, dt=0.01); every row equals the oracle's
    counts after that many steps."""
    print("compat: cmddata model", flush=True)
    s = smoldyn.Simulation(low=[0, 0, 0], high=[100, 100, 100], seed=21)
    R = s.addSpecies("R", difc=1, color="red")
    B = s.addSpecies("B", difc=1, color="blue")
    s.addReaction("r1", subs=[R], prds=[], rate=0.1)
    R.addToSolution(400)
    B.addToSolution(1, pos=[50, 50, 50])
    s.addOutputData("mydata")
    s.addCommand("molcount mydata", "E")
    s.run(stop=2.0, dt=0.01)
    data = np.array(s.getOutputData("mydata", 0))

    osim = _oracle([100] * 3, "rrr", 0.01, 21, {"R": 1.0, "B": 1.0},
                   [("R", 400, [0] * 3, [100] * 3), ("B", 1, [50] * 3, [50] * 3)], [("r1", ["R"], [], 0.1)])
    nsteps = _steps_to(0.0, 2.0, 0.01)
    t, c, it = ob.command_timeline("times 0 2.0 0.01\ncmd E 0 0 0 0 a\nrun\n")
    counts = []
    for k in range(nsteps + 1):
        if k:
            osim.step()
        counts.append([osim.count(osim.index("R")), osim.count(osim.index("B"))])
    exp = np.array([[t[j]] + counts[min(it[j], nsteps)] for j in range(len(t)) if c[j] == 0])
    assert data.shape == exp.shape
    assert np.array_equal(data, exp)
    assert data[-1][1] < 400
    ss.compare_states(ss.oracle_state(osim), ss.device_state(s.engine))


def test_box_model_state_equals_oracle(tmp_path, monkeypatch):
    """box.py: C <-> A + B (kf 0.1, kb 100) with a list for each species'
    solution state (the other states share the list made for unassigned
    states), boxes of side 10 and 'i' molcount to an output file; the state
    after s.run(10, dt=0.5) equals the oracle's."""
    print("compat: box model", flush=True)
    monkeypatch.chdir(tmp_path)
    s = smoldyn.Simulation(low=[0, 0, 0], high=[100, 100, 100], boundary_type="ppp", seed=22)
    s.addBox(size=10)
    a = s.addSpecies("A", state="soln", difc=1, color="red", mol_list="Alist")
    b = s.addSpecies("B", state="soln", color="green", difc=1, mol_list="Blist")
    c = s.addSpecies("C", state="soln", difc=1.0, color="blue", mol_list="Clist")
    s.addMolecules(a, 1000)
    s.addMolecules(b, 1000)
    s.addBidirectionalReaction("r1", subs=[c], prds=(a, b), kf=0.1, kb=100)
    s.setOutputFile("box.dat", True)
    s.addCommand("molcount box.dat", cmd_type="i", on=0, off=10, step=2)
    s.run(10, dt=0.5, log_level=1)
    assert any(p.name.startswith("box") for p in tmp_path.iterdir())

    spec = dict(low=[0, 0, 0], high=[100, 100, 100], walls="ppp", dt=0.5, seed=22, boxsize=10.0,
                lists={"Alist": [("A", "soln")], "Blist": [("B", "soln")], "Clist": [("C", "soln")]},
                species={"A": 1.0, "B": 1.0, "C": 1.0},
                reactions=[("r1fwd", ["C"], ["A", "B"], 0.1), ("r1rev", ["A", "B"], ["C"], 100.0)],
                molecules=[("A", 1000), ("B", 1000)])
    osim, _ = ss.build_oracle(spec)
    for _ in range(_steps_to(0.0, 10.0, 0.5)):
        osim.step()
    assert osim.count(osim.index("C")) > 0
    ss.compare_states(ss.oracle_state(osim), ss.device_state(s.engine))


def test_runtimestep_through_low_level_class():
    """runtimestep.py: S.Simulation.setSimTimes(s, ...), then s.runTimeStep()
    ten times; the state after each step equals the oracle's."""
    print("compat: runTimeStep", flush=True)
    s = smoldyn.Simulation(low=[0, 0, 0], high=[100, 100, 100], seed=23)
    red = s.addSpecies("red", difc=3, color="red")
    s.addMolecules(red, 100)
    s.addMolecules(red, 30, highpos=(20, 30, 20))
    S.Simulation.setSimTimes(s, 0, 100, 0.1)
    osim = _oracle([100] * 3, "rrr", 0.1, 23, {"red": 3.0},
                   [("red", 100, [0] * 3, [100] * 3), ("red", 30, [0] * 3, [20, 30, 20])])
    for k in range(10):
        assert s.runTimeStep() == S.ErrorCode.ok
        osim.step()
        ss.compare_states(ss.oracle_state(osim), ss.device_state(s.engine))


def test_model_file_through_fromfile(tmp_path):
    """Simulation.fromFile on the reference's regression model 'molecule'
    (uniform, ranged and fixed placements in a reflective box), then
    s.run(stop, dt) with the file's own times: the listmols file its
    end-of-run command writes equals the oracle's byte for byte."""
    from smolcfg import read_config

    print("compat: fromFile molecule", flush=True)
    text = (Path(__file__).resolve().parent / "golden" / "s95" / "molecule.txt").read_text()
    f = tmp_path / "molecule.txt"
    f.write_text("define OUTFILE lm.txt\n" + text)
    s = smoldyn.Simulation.fromFile(f)
    assert s.setOutputPath(str(tmp_path) + "/") == S.ErrorCode.ok
    s.run(stop=10, dt=0.01, overwrite=True)
    s.engine.close()

    class _CounterOSim(ob.OSim):
        def seed(self, seed):
            self.rngmode(ob.RNG_COUNTER, seed)

    osim = read_config(text, lambda d, lo, hi: _CounterOSim(d, lo, hi))
    osim.update()
    assert osim.run() == 1
    expect = osim.listmols()
    got = (tmp_path / "lm.txt").read_text()
    assert len(expect.splitlines()) > 0
    assert got == expect
