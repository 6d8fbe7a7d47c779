"""Tile binning against the counting sort and the oracle, through the C-ABI.

A single simulation without surfaces bins its records with two tile passes
(k_tile1 / k_tile2) once its records are in key order; SMOLDYN_AMD_TILE=0
keeps the four-launch counting sort for every step.  Both must give the same
states bit for bit: same positions, live-list order, serials, boxes and
reaction events.  Crowded tiles (more records than a tile keeps in
registers) and full immigrant buckets (the spill list) are driven by
starting every molecule in a few boxes."""
import os

import numpy as np
import pytest

import simspec as ss

pytestmark = pytest.mark.gpu


def _device(spec, tile):
    old = os.environ.get("SMOLDYN_AMD_TILE")
    if not tile:
        os.environ["SMOLDYN_AMD_TILE"] = "0"
    try:
        return ss.build_device(spec)
    finally:
        if old is None:
            os.environ.pop("SMOLDYN_AMD_TILE", None)
        else:
            os.environ["SMOLDYN_AMD_TILE"] = old


def _same(a, b):
    ss.compare_states(ss.device_state(a), ss.device_state(b))
    assert ss.canonical_pairs(a.reaction_log()) == ss.canonical_pairs(b.reaction_log())


@pytest.mark.parametrize("precision", [64, 32])
def test_tile_binning_equals_counting_sort(precision):
    """Reversible A+B<->C scaled to 100k molecules: 20 single steps, then
    graph-replayed steps, tile binning vs the counting sort."""
    spec = ss.c2_spec(n=50000, side=1000.0 * (100000 / 1e6) ** (1 / 3), kback=20.0, seed=71, precision=precision)
    print("tile vs counting sort, precision %d" % precision, flush=True)
    a, b = _device(spec, True), _device(spec, False)
    nev = 0
    for _ in range(20):
        a.run_time_step()
        b.run_time_step()
        _same(a, b)
        nev += len(a.reaction_log())
    assert nev > 0
    a.run_steps(60)
    b.run_steps(60)
    ss.compare_states(ss.device_state(a), ss.device_state(b))
    for _ in range(3):
        a.run_time_step()
        b.run_time_step()
        _same(a, b)


def test_tile_binning_crowded_tiles_and_spill_against_oracle():
    """Every molecule starts in a small cube: tiles hold far more records than
    k_tile2 keeps in registers and the molecules spreading out overflow the
    immigrant buckets.  Bit-exact with the oracle at every step."""
    spec = ss.c2_spec(n=6000, side=120.0, kback=20.0, seed=73, mol_low=[55.0] * 3, mol_high=[65.0] * 3)
    spec["species"] = {"A": 40.0, "B": 40.0, "C": 20.0}
    print("crowded tiles against the oracle", flush=True)
    osim, _ = ss.build_oracle(spec)
    dsim = _device(spec, True)
    nev = 0
    for k in range(40):
        osim.step()
        dsim.run_time_step()
        lo = ss.canonical_pairs(osim.rxnlog()[:, :6])
        assert lo == ss.canonical_pairs(dsim.reaction_log()), "step %d" % k
        nev += len(lo)
        if k % 5 == 4:
            ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))
    assert nev > 0
    ss.compare_states(ss.oracle_state(osim), ss.device_state(dsim))


def test_tile_binning_after_added_molecules():
    """Molecules added between steps: the next step bins with the counting
    sort, later ones with tiles; equal to an engine that never uses tiles."""
    spec = ss.c2_spec(n=20000, side=1000.0 * (40000 / 1e6) ** (1 / 3), kback=20.0, seed=75)
    a, b = _device(spec, True), _device(spec, False)
    rng = np.random.default_rng(3)
    side = spec["high"][0]
    for k in range(12):
        if k % 4 == 1:
            pos = rng.uniform(0, side, (300, 3))
            for s in (a, b):
                s.add_molecules_at("A", pos)
        a.run_time_step()
        b.run_time_step()
        _same(a, b)
