"""Configuration files through the C-ABI (smolLoadSimFromFile,
smolReadConfigString), on the CPU: the reader builds the simulation with the
libsmoldyn calls a caller would make, so what it builds is read back through
the same API before any device exists."""
import gzip
from pathlib import Path

import pytest

import smoldyn_amd as sa

S95 = Path(__file__).resolve().parent / "golden" / "s95"


def _load(tmp_path, text):
    f = tmp_path / "model.txt"
    f.write_text(text)
    return sa.Simulation.from_file(f, update=False)


def test_reference_models_load():
    """The reference's own regression models that the device runs read
    without error; molecule counts, lists and commands come through."""
    for name, counts in [("lotvolt", {"rabbit": 1000, "fox": 1000}),
                         ("molecule", {"red": 50, "green": 50, "blue": 50, "yellow": 50, "magenta": 50,
                                       "cyan": 50})]:
        s = sa.Simulation.from_file(S95 / (name + ".txt"), update=False)
        for sp, n in counts.items():
            assert s.count(sp, sa.MSall) == n, (name, sp)
        s.close()


def test_statements_and_blocks(tmp_path):
    text = """# a model with most statement kinds
define SIDE 50
dim 3
boundaries x 0 SIDE p
boundaries y 0 SIDE
low_wall z 0 r
high_wall z 50 r
random_seed 7
time_start 0
time_stop 1   /* a block comment
over two lines */ 
time_step 0.01
species A B C
difc all 1
difc C 0.5
molecule_lists la lbc
mol_list A la
mol_list B lbc
mol_list C lbc
molperbox 4
reaction fwd A + B -> C 20
reaction back C -> A + B 2
product_placement back x 0.2
reaction pair A + C <-> B + B 5 1
mol 100 A u u u
mol 20 B 10-20 5 u
ifdefine NOT_SET
  mol 1000 C u u u
else
  mol 3 C 25 25 25
endif
graphics opengl
color A red
output_data counts
output_files out.txt
output_precision 6
output_format csv
cmd N 10 molcount counts
cmd a listmols out.txt
start_surface wall
  action both all reflect
  panel rect +2 0 0 25 50 50 r1
  panel sph 25 25 25 5 20 20
end_surface
end_file
mol 5 A u u u
"""
    s = _load(tmp_path, text)
    assert [s.count(x, sa.MSall) for x in "ABC"] == [100, 20, 3]
    s.read_config("mol", "4 C 1 2 3")
    assert s.count("C", sa.MSall) == 7
    s.run_command("molcount counts")
    d = s.output_data("counts")
    assert d.tolist() == [[0.0, 100.0, 20.0, 7.0]]
    s.close()


@pytest.mark.parametrize("stmt,where", [("difm A 1 0 0 0 1 0 0 0 1", "model.txt:6"),
                                        ("reaction compartment=c r A -> B 1", "model.txt:6"),
                                        ("start_surface s\n panel cyl 0 0 0 0 0 0 2 20 20\nend_surface",
                                         "model.txt:7")])
def test_unsupported_statements_name_their_line(tmp_path, stmt, where):
    text = "dim 3\nboundaries 0 0 10\nboundaries 1 0 10\nboundaries 2 0 10\n%s\n" % stmt
    text = "species A B\n" + text
    with pytest.raises(sa.SmoldynError) as e:
        _load(tmp_path, text)
    assert where in str(e.value)


def test_read_config_string():
    s = sa.Simulation([0, 0, 0], [10, 10, 10])
    s.read_config("species", "X Y")
    s.read_config("difc", "X 2")
    s.read_config("mol", "12 X u u u")
    assert s.count("X", sa.MSall) == 12
    with pytest.raises(sa.SmoldynError):
        s.read_config("start_surface", "s")
    with pytest.raises(sa.SmoldynError):
        s.read_config("no_such_statement", "1")
    s.close()


@pytest.mark.parametrize("stmt,ok", [("action both all reflect", True), ("action all both reflect", True),
                                     ("action front A transmit", True), ("action A f transmit", True),
                                     ("action b A reflect", True), ("action both nosuch reflect", False),
                                     ("action A sideways reflect", False)])
def test_surface_action_formats(tmp_path, stmt, ok):
    """surfreadstring's action statement: the new format (species face
    action) when its species is known and its face readable, else the old
    one (face species action); faces may be shortened (f, b) and 'all'
    means both faces."""
    text = ("species A\ndim 3\nboundaries 0 0 10\nboundaries 1 0 10\nboundaries 2 0 10\n"
            "start_surface s\n%s\npanel tri 1 1 1 9 1 1 5 9 1\nend_surface\n" % stmt)
    if ok:
        _load(tmp_path, text).close()
    else:
        with pytest.raises(sa.SmoldynError) as e:
            _load(tmp_path, text)
        assert "model.txt:7" in str(e.value)


@pytest.mark.parametrize("stmt,ok", [("panel rectangle +0 0 0 0 10 10", True), ("panel Rect +0 0 0 0 10 10", True),
                                     ("panel TRIANG 1 1 1 9 1 1 5 9 1", True),
                                     ("panel sph 5 5 5 2 10 10", True), ("panel Sphere 5 5 5 2", True),
                                     ("panel cyli 2 5 5 8 5 5 1 10 10", True),
                                     ("panel hemis 5 5 5 2 0 0 1 10 10", True), ("panel Di 5 5 5 2 0 0 1", True),
                                     ("panel rectangles +0 0 0 0 10 10", False), ("panel all 5 5 5 2", False),
                                     ("action Front A Reflect", True), ("action A BACK trans", True),
                                     ("action A fr mult", True), ("action A both periodic", True),
                                     ("action A both per", False)])
def test_surface_names_take_leading_parts_any_case(tmp_path, stmt, ok):
    """Shape, face and action names are read as the reference's strbegin
    does (libSteve/string2.c:81-90, surfstring2ps / surfstring2face /
    surfstring2act in smolsurface.c:86-155): any non-empty leading part of
    the word, in any case; 'periodic' (a jump) only spelled out."""
    text = ("species A\ndim 3\nboundaries 0 0 10\nboundaries 1 0 10\nboundaries 2 0 10\n"
            "start_surface s\n%s\npanel tri 1 1 1 9 1 1 5 9 1\nend_surface\n" % stmt)
    if ok:
        _load(tmp_path, text).close()
    else:
        with pytest.raises(sa.SmoldynError) as e:
            _load(tmp_path, text)
        assert "model.txt:7" in str(e.value)


@pytest.mark.parametrize("code,ok", [("x 0.2", True), ("pgemmax 0.2", True), ("b 1.5", True), ("unbindrad 1.5", True),
                                     ("r 1.2", True), ("p 0.3", True), ("bounce", False), ("pgemmaxw 0.2", False),
                                     ("b2 1.5", False), ("none", False)])
def test_product_placement_codes(tmp_path, code, ok):
    """Placement names as rxnstring2rp reads them (smolreact.c:79-106): one
    letter codes where 'b' is unbindrad; bounce (only for bimolecular
    reactions) and pgemmaxw are not accepted for this first-order reaction,
    the latter because the reader does not know the name."""
    text = ("species A B C\ndim 3\nboundaries 0 0 10\nboundaries 1 0 10\nboundaries 2 0 10\ntime_step 0.01\n"
            "difc all 1\nreaction fwd A + B -> C 20\nreaction back C -> A + B 2\nproduct_placement back %s\n" % code)
    if ok:
        _load(tmp_path, text).close()
    else:
        with pytest.raises(sa.SmoldynError) as e:
            _load(tmp_path, text)
        assert "model.txt:10" in str(e.value)
