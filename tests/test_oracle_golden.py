"""The oracle reproduces the reference's own regression outputs
(examples/S95_regression, golden files of release 2.75) byte for byte.

These runs use the reference's random source (SFMT-19937 in program order), so
they pin the whole CPU path: Gaussian-table construction and shuffle, molecule
placement, diffusion, reflective and periodic walls, virtual-box grid and
neighbor lists (including wrap codes), live-list and dead-list ordering,
first- and second-order reactions with their greedy pair order, product
placement and serial numbers.  surf1 (1D) and ReflectToNewSpecies (2D) pin the
surface path: rectangle, triangle and sphere panels, reflect / transmit /
absorb / jump actions, species change on reflection, box panel lists and the
box walk along a trajectory (line2nextbox), and bimolecular reactions with
surfaces present (rxnXsurface).  surf3 (3D) pins every 3D panel shape:
rectangles, triangles, a sphere, cylinders, hemispheres and disks (their
crossing tests, normals, reflections, box-panel tests), with reflect, absorb,
transmit and jump actions.  bounds1 (2D) pins absorbing walls: the kill of
a molecule beyond the wall, the Brownian-bridge coin exp(-2 d_i d / step^2)
for one that stayed inside, both wall passes of a step, and the list sort at
the end of every step (simdocommands, smolsim.cpp:2571-2581) that removes
molecules killed by the second pass before the next diffusion.  bounds2 (2D)
pins pgemmax product placement (unbindingradius of the geminate probability),
a circle that reflects and jump rectangles on the box sides.  allostery (2D, immobile molecules) pins
conformational-spread reactions: the confspread radius, the reaction
probability 1-exp(-k dt), and products that keep their reactants' places and
serial numbers."""
import gzip
from pathlib import Path

import pytest

from oracle_bind import OSim
from smolcfg import read_config

S95 = Path(__file__).resolve().parent / "golden" / "s95"


def run_model(name):
    sim = read_config((S95 / (name + ".txt")).read_text(), lambda d, lo, hi: OSim(d, lo, hi))
    sim.update()
    code = sim.run()
    return code, sim.listmols()


@pytest.mark.parametrize("name", ["bounce2", "molecule", "lotvolt", "surf1", "ReflectToNewSpecies", "allostery",
                                  "surf3", "bounds1", "bounds2"])
def test_s95_output_identical(name):
    code, out = run_model(name)
    assert code == 1
    gold = gzip.open(S95 / (name + "_275_out.txt.gz")).read().decode()
    assert out == gold
