"""Host-side checks of the C-ABI library that need no GPU: every symbol the
public header declares is exported, argument errors follow the reference's
ErrorCode behaviour, and the setup's reaction-parameter solver agrees with
the reference (through the oracle, pinned to the compiled reference)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_bind as ob
import smoldyn_amd as sa

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        for m in re.finditer(r"\b(smol[A-Za-z0-9_]+)\s*\(", h.read_text()):
            if not m.group(1).startswith("smolrng"):  # header-inline counter-RNG helpers
                names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = sa.lib()
    syms = declared_symbols()
    assert len(syms) >= 40
    for s in sorted(syms):
        assert hasattr(L, s), s


def test_errors_follow_reference_codes():
    L = sa.lib()
    lo = np.zeros(3)
    hi = np.ones(3)
    assert not L.smolNewSim(4, lo.ctypes.data_as(C.POINTER(C.c_double)), hi.ctypes.data_as(C.POINTER(C.c_double)))
    code, fn, msg = sa.last_error()
    assert code == sa.ECbounds and fn == "smolNewSim"
    sim = sa.Simulation([0, 0, 0], [10, 10, 10])
    sim.species("A")
    with pytest.raises(sa.SmoldynError):
        sim.species("A")
    assert L.smolGetSpeciesIndexNT(sim.p, b"nope") == sa.ECnonexist
    assert L.smolGetSpeciesIndexNT(sim.p, b"all") == sa.ECall
    assert L.smolSetPartitions(sim.p, b"bogus", 1.0) == sa.ECsyntax
    sim.add_molecules("A", 25)
    assert sim.count("A") == 25
    assert sim.reaction_index("zz")[1] == sa.ECnonexist


def test_lookup_errors_propagate_as_the_reference_does():
    """An inner lookup records its code and message (smolGetSpeciesIndexNT,
    libsmoldyn.cpp:851-866) and the caller keeps them under its own name
    (LCHECK with ECsame, e.g. smolAddSolutionMolecules :1057-1058); 'all'
    selects every species where the reference allows it."""
    L = sa.lib()
    sim = sa.Simulation([0, 0, 0], [10, 10, 10])
    sim.species("A")
    sim.species("B")
    lo = np.zeros(3)
    dp = lo.ctypes.data_as(C.POINTER(C.c_double))
    assert L.smolAddSolutionMolecules(sim.p, b"Q", 3, dp, dp) == sa.ECnonexist
    assert sa.last_error() == (sa.ECnonexist, "smolAddSolutionMolecules", "species 'Q' not found")
    assert L.smolAddSolutionMolecules(sim.p, b"all", 3, dp, dp) == sa.ECall
    assert sa.last_error() == (sa.ECall, "smolAddSolutionMolecules", "species cannot be 'all'")
    assert L.smolAddSolutionMolecules(sim.p, None, 3, dp, dp) == sa.ECmissing
    assert sa.last_error() == (sa.ECmissing, "smolAddSolutionMolecules", "missing species name")
    assert L.smolGetMoleculeCount(sim.p, b"Q", sa.MSsoln) == sa.ECnonexist
    assert sa.last_error()[1] == "smolGetMoleculeCount"
    assert L.smolGetSpeciesIndex(sim.p, b"Q") == sa.ECnonexist
    assert sa.last_error() == (sa.ECnonexist, "smolGetSpeciesIndex", "species 'Q' not found")
    # 'all' is every species for the list assignment and the count
    sim.add_list("L1")
    assert L.smolSetMolList(sim.p, b"all", sa.MSall, b"L1") == sa.ECok
    assert L.smolGetMolListIndex(sim.p, b"L1") == 0
    assert L.smolSetMolList(sim.p, b"A", sa.MSall, b"nolist") == sa.ECnonexist
    assert sa.last_error() == (sa.ECnonexist, "smolSetMolList", "list name not recognized")
    sim.add_molecules("A", 4)
    sim.add_molecules("B", 5)
    assert sim.count() == 9 and sim.count("B") == 5
    assert L.smolAddReaction(sim.p, b"r", b"A", sa.MSsoln, b"Z", sa.MSsoln, 0, None, None, 1.0) == sa.ECnonexist
    assert sa.last_error() == (sa.ECnonexist, "smolAddReaction", "species 'Z' not found")


@pytest.mark.parametrize("rate,dt,difc,b,rel", [(9.640578597, 0.01, 2.0, -1, 0), (99.3218287, 0.005, 2.0, 1.2, 1),
                                                (21.42267826, 0.002, 4.0, -1, 0), (8000.0, 0.001, 200.0, -1, 0),
                                                (5.0, 0.0, 1.0, 2.0, 0)])
def test_setup_binding_radius_matches_oracle(rate, dt, difc, b, rel):
    L = sa.lib()
    O = ob.lib()
    assert L.smolamdBindingRadius(rate, dt, difc, b, rel) == O.orxn_bindingradius(rate, dt, difc, b, rel)


def test_setup_rate_tables_match_oracle():
    L = sa.lib()
    O = ob.lib()
    for s in (0.01, 0.2, 0.9, 3.0, 40.0):
        for b in (-1, 0.0, 0.4, 1.0, 1.2, 5.0):
            assert L.smolamdNumRxnRate(s, 1.0, b) == O.orxn_numrxnrate(s, 1.0, b)
    for pg in (0.1, 0.2, 0.6):
        assert L.smolamdUnbindingRadius(pg, 0.005, 3.0, 1.0) == O.orxn_unbindingradius(pg, 0.005, 3.0, 1.0)


def build_c_caller(tmp):
    """tests/libsmoldyn_caller.c compiled with gcc (C99, warnings as errors)
    against include/libsmoldyn_amd.h and linked to the built library."""
    exe = Path(tmp) / "libsmoldyn_caller"
    lib = sa.LIBPATH.parent
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", str(ROOT / "include"), "-o", str(exe),
                    str(ROOT / "tests" / "libsmoldyn_caller.c"), "-L", str(lib), "-l:" + sa.LIBPATH.name,
                    "-Wl,-rpath," + str(lib)], check=True)
    return exe


def test_c_caller_compiles_against_the_header_and_sets_up(tmp_path):
    exe = build_c_caller(tmp_path)
    r = subprocess.run([str(exe), "setup"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert "lookup smolAddSolutionMolecules | species 'Q' not found" in lines
    assert "species 1 2 3" in lines and lines[-1] == "setup ok"
