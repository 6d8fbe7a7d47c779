"""The reference's own regression models (examples/S95_regression lotvolt:
Lotka-Volterra in a periodic box with two molecule lists; molecule: uniform,
ranged and fixed placements in a reflective box) read by
smolPrepareSimFromFile and run by smolRunSim on the device: the listmols file
written by the model's own end-of-run command equals, byte for byte, the
oracle's listmols after the same run on the same counter stream (the oracle
reads the file with the test reader that reproduces the reference's golden
outputs on the reference's stream)."""
from pathlib import Path

import pytest

import oracle_bind as ob
import smoldyn_amd as sa
from smolcfg import read_config

pytestmark = pytest.mark.gpu

S95 = Path(__file__).resolve().parent / "golden" / "s95"


class _CounterOSim(ob.OSim):
    """The oracle on the device's counter stream: random_seed sets its key."""

    def seed(self, s):
        self.rngmode(ob.RNG_COUNTER, s)


@pytest.mark.parametrize("name,extra", [("lotvolt", "max_mol 40000\n"), ("molecule", "")])
def test_reference_model_file_on_device(tmp_path, name, extra):
    text = (S95 / (name + ".txt")).read_text()
    f = tmp_path / (name + ".txt")
    # the device sizes its lists once: lotvolt's foxes outgrow the default margin
    f.write_text("define OUTFILE lm.txt\n" + extra + text)
    dsim = sa.Simulation.from_file(f)
    dsim.set_output_path(str(tmp_path) + "/")
    assert dsim.run() == sa.ECok
    dsim.close()
    osim = read_config(text, lambda d, lo, hi: _CounterOSim(d, lo, hi))
    osim.update()
    assert osim.run() == 1
    expect = osim.listmols()
    got = (tmp_path / "lm.txt").read_text()
    assert len(got.splitlines()) == len(expect.splitlines())
    assert got == expect


def test_reference_surf3_unmodified_on_device(tmp_path):
    """examples/S95_regression surf3 as the reference ships it, absorbing
    action for green included, read from the file and run for 2500 steps on
    the device: the listmols output equals the oracle's on the same counter
    stream."""
    text = (S95 / "surf3.txt").read_text()
    assert "action both green absorb" in text
    f = tmp_path / "surf3.txt"
    f.write_text("define OUTFILE lm.txt\n" + text)
    dsim = sa.Simulation.from_file(f)
    dsim.set_output_path(str(tmp_path) + "/")
    assert dsim.run() == sa.ECok
    dsim.close()
    osim = read_config(text, lambda d, lo, hi: _CounterOSim(d, lo, hi))
    osim.update()
    assert osim.run() == 1
    expect = osim.listmols()
    got = (tmp_path / "lm.txt").read_text()
    assert len(expect.splitlines()) < 1000  # green molecules were absorbed
    assert got == expect


def test_reference_surf3_model_on_device(tmp_path):
    """examples/S95_regression surf3: rectangles, triangles, a sphere,
    cylinders, hemispheres and disks, with reflect, transmit and jump
    actions, read from the file and run for 2500 steps on the device; the
    listmols output equals the oracle's on the same counter stream.  The
    model's absorbing action for green becomes a transmitting one (the
    device path does not run absorption); the oracle reproduces the
    unmodified model's golden output byte for byte."""
    text = (S95 / "surf3.txt").read_text()
    assert "action both green absorb" in text
    text = text.replace("action both green absorb", "action both green transmit")
    f = tmp_path / "surf3.txt"
    f.write_text("define OUTFILE lm.txt\n" + text)
    dsim = sa.Simulation.from_file(f)
    dsim.set_output_path(str(tmp_path) + "/")
    assert dsim.run() == sa.ECok
    dsim.close()
    osim = read_config(text, lambda d, lo, hi: _CounterOSim(d, lo, hi))
    osim.update()
    assert osim.run() == 1
    expect = osim.listmols()
    got = (tmp_path / "lm.txt").read_text()
    assert len(got.splitlines()) == 1000
    assert got == expect
