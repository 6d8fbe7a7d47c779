"""The reference's Python API (``import smoldyn``) served by the engine
library, checked on the host: models are defined through the user-level
classes and the low-level simulation, and each call lands in the C-ABI with
the reference's meaning (source/python/smoldyn/smoldyn.py, module.cpp).
Nothing here steps a simulation; the runs are GPU tests."""
import sys
from pathlib import Path

import pytest

import smoldyn_amd as sa

sys.path.insert(0, str(Path(sa.__file__).resolve().parent / "compat"))
import smoldyn  # noqa: E402
import smoldyn._smoldyn as S  # noqa: E402


def test_enums_match_libsmoldyn():
    assert [int(S.MolecState[n]) for n in ("soln", "front", "back", "up", "down", "bsoln", "all", "none")] == \
        [sa.MSsoln, sa.MSfront, sa.MSback, sa.MSup, sa.MSdown, sa.MSbsoln, sa.MSall, sa.MSnone]
    assert {m.name: int(m) for m in S.RevParam} == sa.RP
    assert int(S.ErrorCode.ok) == sa.ECok and int(S.ErrorCode.nonexist) == sa.ECnonexist
    assert (int(S.PanelFace.front), int(S.PanelFace.back), int(S.PanelFace.both)) == (0, 1, 3)
    shapes = sa.Simulation.SHAPES
    assert [int(S.PanelShape[n]) for n in ("rect", "tri", "sph", "cyl", "hemi", "disk")] == \
        [shapes[n] for n in ("rect", "tri", "sph", "cyl", "hemi", "disk")]
    acts = sa.Simulation.ACTIONS
    assert [int(S.SrfAction[n]) for n in ("reflect", "trans", "absorb", "jump")] == \
        [acts[n] for n in ("reflect", "transmit", "absorb", "jump")]


def test_user_api_defines_model():
    s = smoldyn.Simulation(low=[0, 0, 0], high=[100, 100, 100], boundary_type="p", seed=5)
    a = s.addSpecies("A", difc=1, color="red", mol_list="Alist")
    b = s.addSpecies("B", difc={"soln": 2.0})
    c = s.addSpecies("C", difc=0.5)
    r = s.addBidirectionalReaction("r1", subs=[a, b], prds=[c], kf=10.0, kb=5.0)
    r.reverse.productPlacement("r", 1.2)
    s.addReaction("decay", subs=[c], prds=[], rate=0.1)
    a.addToSolution(100)
    s.addMolecules(b, 50, lowpos=[0, 0, 0], highpos=[10, 10, 10])
    b.addToSolution(1, pos=[50, 50, 50])
    s.addOutputData("d")
    s.addCommand("molcount d", "E")
    s.addCommand("molcount d", "N", step=10)
    s.addCommand("molcount d", "i", on=0, off=1, step=0.1)
    s.addCommandStr("e molcount d")
    s.setGraphics("opengl", 10)
    assert s.speciesIndexNT("A") == 1 and s.speciesIndexNT("C") == 3
    assert s.getSpeciesName(2) == "B"
    assert s.getMolListIndex("Alist") >= 0
    assert s.getReactionIndexNT([-1], "r1fwd") == 0
    assert s.getReactionIndexNT([-1], "r1rev") == 0  # the first first-order reaction
    assert s.getReactionIndexNT([-1], "decay") == 1
    cnt = s.count()
    assert cnt["dim"] == 3 and cnt["species"] == 4 and cnt["surface"] == 0
    assert a.difc == {"soln": 1} and b.difc == {"soln": 2.0} and a.mol_list == "Alist"
    assert s.getMoleculeCount("B", S.MolecState.soln) == 51


def test_low_level_calls_return_error_codes():
    s = S.Simulation([0, 0, 0], [10, 10, 10], ["r", "pp", "r"])
    assert s.addSpecies("A") == S.ErrorCode.ok
    assert s.addSpecies("A") != S.ErrorCode.ok  # already defined
    assert s.setSpeciesMobility("nosuch", S.MolecState.soln, 1.0) == S.ErrorCode.nonexist
    assert s.setBoundaryType(0, 1, "x") == S.ErrorCode.syntax
    assert s.setSimTimes(0, 1, 0.01) == S.ErrorCode.ok
    assert (s.start, s.stop, s.dt) == (0.0, 1.0, 0.01)
    assert s.setPartitions("boxsize", 2.0) == S.ErrorCode.ok
    code, msg = S.getError(clear=True)
    assert isinstance(msg, str)
    with pytest.raises(ValueError):
        s.addCommand("molcount d", "z")


def test_surfaces_through_user_api():
    s = smoldyn.Simulation(low=[0, 0, 0], high=[50, 50, 50])
    a = s.addSpecies("A", difc=1)
    sph = smoldyn.Sphere(center=[25, 25, 25], radius=10, slices=20)
    r1 = s.addRectangle(corner=[0, 0, 5], dimensions=[50, 50], axis="+z")
    r2 = smoldyn.Rectangle(corner=[0, 0, 45], dimensions=[50, 50], axis="-z")
    wall = s.addSurface("wall", panels=[sph])
    wall.setAction("both", [a], "reflect")
    plates = s.addSurface("plates", panels=[r1, r2])
    plates.both.setAction(a, "trans")
    r1.jumpTo("front", r2, "front", bidirectional=True)
    assert r1.name == "rect0" and r2.name == "rect1" and sph.name == "sph0"
    assert s.getSurfaceIndexNT("plates") == 1
    assert s.getPanelIndexNT("plates", "rect1") == 1
    assert s.count()["surface"] == 2
    with pytest.raises(NameError):
        wall.setAction("front", "A", "absorb")


def test_features_outside_the_engine_fail_loudly():
    s = smoldyn.Simulation(low=[0, 0, 0], high=[10, 10, 10])
    a = s.addSpecies("A", difc=1)
    with pytest.raises(NotImplementedError):
        s.addCompartment("c", surface="x", point=[1, 1, 1])
    with pytest.raises(NotImplementedError):
        s.addPort("p", surface=None, panel=None)
    with pytest.raises(NotImplementedError):
        s.connect(print, None, 1, [])
    with pytest.raises(NameError):
        s.addSpecies("B", state="gas")
    with pytest.raises(AssertionError):
        s.run(stop=1.0, dt=0.0)
