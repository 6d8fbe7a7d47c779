"""FP32 box-offset records against the FP64 oracle.

In this mode each position is stored as FP32 offsets from the FP64 origin of
its virtual box; diffusion, walls, distance tests and product placement stay
FP64.  Storing an offset o (|o| < box size h) rounds it by at most
h * 2^-24, so after s steps from an FP32-rounded start a position differs from
the FP64 trajectory by at most (s + 1) h 2^-24 as long as both make the same
reaction decisions.  The tests check that bound, that species, serials and
live order are the oracle's, and that the reaction events are the oracle's;
the rate statement compares event counts of FP32 and FP64 runs that share
their random streams."""
import numpy as np
import pytest

import simspec as ss

pytestmark = pytest.mark.gpu

U = 2.0 ** -24


def _box_size(spec, dsim):
    _, _, _, bsize = dsim.boxes()
    return float(np.max(bsize))


def _compare_within(o, d, bound):
    assert len(o) == len(d)
    worst = 0.0
    for ll, (a, b) in enumerate(zip(o, d)):
        assert np.array_equal(a["serno"], b["serno"]), "list %d serial order differs" % ll
        assert np.array_equal(a["ident"], b["ident"]), "list %d species differ" % ll
        if len(a["ident"]):
            err = np.abs(a["pos"] - b["pos"]).max()
            worst = max(worst, float(err))
    assert worst <= bound, "max position error %g above the bound %g" % (worst, bound)
    return worst


def _run(spec, nsteps):
    osim, _ = ss.build_oracle(spec)
    dsim = ss.build_device(dict(spec, precision=32))
    h = _box_size(spec, dsim)
    errs = [_compare_within(ss.oracle_state(osim), ss.device_state(dsim), 1.01 * h * U)]
    rng = np.random.default_rng(99)
    nev = 0
    for k in range(nsteps):
        n = sum(len(x["ident"]) for x in ss.oracle_state(osim))
        g = rng.standard_normal((n, 3))
        osim.inject(g)
        osim.step()
        dsim.inject_step(g)
        lo = ss.canonical_pairs(osim.rxnlog()[:, :6])
        ld = ss.canonical_pairs(dsim.reaction_log())
        assert lo == ld, "step %d reaction events differ" % k
        nev += len(lo)
        errs.append(_compare_within(ss.oracle_state(osim), ss.device_state(dsim), (k + 2) * 1.01 * h * U + 1e-12))
    print("box size %g: max position error per step %s (bound per step %g)" % (h, ["%.2e" % e for e in errs], h * U))
    return nev, errs


def test_f32_reflective_one_step():
    _run(ss.c1_spec(n=2000, seed=21), 1)


def test_f32_reflective_reactions_many_steps():
    nev, _ = _run(ss.c1_spec(n=3000, seed=22), 30)
    assert nev > 0


def test_f32_periodic_reversible_many_steps():
    spec = ss.c2_spec(n=3000, side=60.0, seed=23)
    nev, _ = _run(spec, 30)
    assert nev > 0


def test_f32_rate_against_fp64_same_streams():
    """FP32 and FP64 engines with the same seed draw the same Philox numbers,
    so they make the same reaction decisions until a pair distance falls
    within the rounding of the binding radius.  Over 300 steps of the
    reversible workload (6000 molecules) the forward and reverse event counts
    agree to within their statistical error."""
    spec = ss.c2_spec(n=3000, side=40.0, seed=24)
    d64 = ss.build_device(spec)
    d32 = ss.build_device(dict(spec, precision=32))
    same = 0
    diverged = False
    tot = np.zeros((2, 2))
    for k in range(300):
        d64.run_time_step()
        d32.run_time_step()
        l64 = ss.canonical_pairs(d64.reaction_log())
        l32 = ss.canonical_pairs(d32.reaction_log())
        if not diverged and l64 == l32:
            same += 1
        else:
            diverged = True
        for j, lg in enumerate((d64.reaction_log(), d32.reaction_log())):
            lg = np.asarray(lg)
            if len(lg):
                tot[j, 0] += np.sum(lg[:, 0] == 2)
                tot[j, 1] += np.sum(lg[:, 0] == 1)
    print("identical event sets for the first %d of 300 steps; forward/reverse events FP64 %s, FP32 %s" % (
        same, tot[0], tot[1]))
    assert same >= 50
    for c in range(2):
        n64, n32 = tot[0, c], tot[1, c]
        assert abs(n64 - n32) <= 3 * np.sqrt(n64 + n32) + 1, (n64, n32)
