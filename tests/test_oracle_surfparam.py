"""Rates -> per-collision probabilities of surface interactions
(source/libSteve/SurfaceParam.c, math2.c) in the oracle (oracle/osurfparam.c)
and in the library's host setup (smoldyn_amd/csrc/surfparam.cpp), pinned
three ways:

* against the compiled reference (oracle/_ref) on random inputs that reach
  every branch (where the reference tree is present), and against the same
  reference's values committed as a fixture (tests/golden/surfparam_ref.json,
  made by tests/golden/make_surfparam_ref.py) everywhere;
* by what the mathematics fixes: the tabulated adsorption probabilities at
  the table's own nodes, the closed forms of irreversible desorption and of
  two-state flips, exp(x^2) erfc(x) - 1 against scipy's erfcx, the
  closed form of a two-sided transmitting surface (with scipy's erfcx), and
  detailed balance of reversible adsorption (p_des = p_ads kr / (kappa
  sqrt(2 pi)));
* the library's restatement bit for bit against the oracle's (compiled here
  from its source with g++, the library's floating-point flags).

Bit-exact: every comparison is of the IEEE bit patterns (NaN matched as NaN)."""
import ctypes as C
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_bind as ob

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "smoldyn_amd" / "csrc"
GOLD = ROOT / "tests" / "golden" / "surfparam_ref.json"
ALGOS = [0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13]


def same(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(both_nan | (a.view(np.uint64) == b.view(np.uint64))))


def o_prob(k1, k2, dt, difc, algo):
    p2 = C.c_double(0)
    return ob.lib().osurf_surfaceprob(k1, k2, dt, difc, C.byref(p2), algo), p2.value


def o_transmit(k1, k2, d1, d2, dt):
    p1, p2 = C.c_double(-2), C.c_double(-2)
    rc = ob.lib().osurf_transmit(k1, k2, d1, d2, dt, C.byref(p1), C.byref(p2))
    return rc, p1.value, p2.value


def mixed_maximal(k1, k2):
    return (k1 == -1) != (k2 == -1) and (k1 not in (0, -1) or k2 not in (0, -1))


# --- the reference's values -------------------------------------------------

def test_oracle_matches_the_reference_fixture():
    data = json.loads(GOLD.read_text())
    for k1, k2, dt, difc, algo, w1, w2 in data["surfaceprob"]:
        assert same(o_prob(k1, k2, dt, difc, int(algo)), (w1, w2)), (k1, k2, dt, difc, algo)
    for k1, k2, d1, d2, dt, rc, w1, w2 in data["transmit"]:
        assert same(o_transmit(k1, k2, d1, d2, dt), (rc, w1, w2)), (k1, k2, d1, d2, dt)
    L = ob.lib()
    for v, w in data["irrev"]:
        assert same(L.osurf_lookupirrevadsorb(v), w), v
    for v1, v2, w1, w2 in data["rev"]:
        p2 = C.c_double(0)
        assert same((L.osurf_lookuprevads(v1, v2, C.byref(p2)), p2.value), (w1, w2)), (v1, v2)
    for x, w in data["experfcm1"]:
        assert same(L.osurf_experfcm1(x), w), x
    # the fixture reaches every algorithm's branches with real values
    probs = np.array([r[5] for r in data["surfaceprob"]])
    assert np.sum((probs > 0) & (probs < 1)) > 300


@pytest.mark.skipif(not ob.reflib_available(), reason="compiled reference (oracle/_ref) not present")
def test_oracle_matches_the_compiled_reference_on_random_inputs():
    R = ob.reflib()
    L = ob.lib()
    rng = np.random.default_rng(77)
    special = [-1.0, 0.0]
    for _ in range(6000):
        algo = int(rng.choice(ALGOS))
        k1 = float(rng.choice(special)) if rng.random() < 0.15 else float(np.exp(rng.uniform(-8, 9)))
        k2 = float(rng.choice(special)) if rng.random() < 0.15 else float(np.exp(rng.uniform(-8, 9)))
        if algo == 3 and mixed_maximal(k1, k2):
            continue  # the reference searches with random draws there
        dt = float(np.exp(rng.uniform(-9, 0)))
        difc = float(np.exp(rng.uniform(-4, 4)))
        p2 = C.c_double(0)
        want = (R.surfaceprob(k1, k2, dt, difc, C.byref(p2), algo), p2.value)
        assert same(o_prob(k1, k2, dt, difc, algo), want), (k1, k2, dt, difc, algo)
    for _ in range(3000):
        k1 = float(rng.choice(special)) if rng.random() < 0.2 else float(np.exp(rng.uniform(-8, 8)))
        k2 = float(rng.choice(special)) if rng.random() < 0.2 else float(np.exp(rng.uniform(-8, 8)))
        if mixed_maximal(k1, k2):
            continue
        d1, d2, dt = (float(np.exp(rng.uniform(-3, 3))) for _ in range(3))
        a, b, p1, p2 = C.c_double(k1), C.c_double(k2), C.c_double(-2), C.c_double(-2)
        rc = R.surfacetransmit(C.byref(a), C.byref(b), C.byref(p1), C.byref(p2), d1, d2, dt)
        assert same(o_transmit(k1, k2, d1, d2, dt), (rc, p1.value, p2.value)), (k1, k2, d1, d2, dt)
    for x in np.concatenate([rng.normal(0, 5, 3000), rng.uniform(-0.06, 0.06, 1000)]):
        assert same(L.osurf_experfcm1(float(x)), R.experfcm1D(float(x))), x


# --- what the mathematics fixes ----------------------------------------------

def _table(name):
    text = (CSRC / "surfparam_tables.inc").read_text()
    body = text.split("SMOLSP_%s[" % name)[1].split("{")[1].split("}")[0]
    return np.array([float(v) for v in body.replace("\n", " ").split(",") if v.strip()])


def test_irreversible_adsorption_lookup_interpolates_its_table():
    L = ob.lib()
    kap, pon = _table("irr_kappa"), _table("irr_pon")
    assert np.all(np.diff(kap) > 0) and pon[0] == 0 and pon[-1] == 1
    for k, p in zip(kap[1:-1], pon[1:-1]):
        assert abs(L.osurf_lookupirrevadsorb(k) - p) < 1e-14  # the interpolant passes through its nodes
    assert L.osurf_lookupirrevadsorb(0.0) == 0 and L.osurf_lookupirrevadsorb(kap[-1]) == 1.0
    # monotone between the nodes, and the known small-kappa slope sqrt(2 pi)
    # of irreversible adsorption (the IrrAdsEC limit)
    xs = np.linspace(1e-4, kap[-1] - 1e-6, 2000)
    ys = np.array([L.osurf_lookupirrevadsorb(x) for x in xs])
    assert np.all(np.diff(ys) > 0)
    assert abs(L.osurf_lookupirrevadsorb(1e-3) / 1e-3 / np.sqrt(2 * np.pi) - 1) < 0.03


def test_reversible_adsorption_table_nodes_and_detailed_balance():
    L = ob.lib()
    x, z = _table("rev_x"), _table("rev_adsprob").reshape(21, 21)
    # interior nodes of the (kappa/(1+kappa), kr/(1+kr)) grid return the table
    for i in range(2, 19):
        for j in range(2, 19):
            kap, kr = x[i] / (1 - x[i]), x[j] / (1 - x[j])
            p2 = C.c_double(0)
            p1 = L.osurf_lookuprevads(kap, kr, C.byref(p2))
            assert abs(p1 - min(max(z[i, j], 0), 1)) < 2e-9, (i, j)  # the node is hit up to rounding of the axis
            # desorption probability from detailed balance (SurfaceParam.c:648)
            assert abs(p2.value - p1 * kr / kap / np.sqrt(2 * np.pi)) < 1e-9 * max(1.0, p2.value)
    # adsorption probability falls as the desorption rate grows (rows of the
    # table decrease along kr) and rises with kappa
    assert np.all(np.diff(z, axis=1) < 0) and np.all(np.diff(z, axis=0) > 0)


def test_desorption_and_flip_closed_forms():
    for k1, k2, dt in [(3.0, 0.0, 0.01), (3.0, 5.0, 0.01), (0.2, 40.0, 0.1), (50.0, 50.0, 0.002)]:
        kk = k2 if k2 > 0 else k1
        want = k1 / kk * (1 - np.exp(-kk * dt))
        for algo in (10, 12):  # IrrDes, IrrFlip: loss through channels of summed rate kk
            p1, _ = o_prob(k1, k2, dt, 1.0, algo)
            assert abs(p1 - want) <= 1e-15 * max(1, want)
        # RevFlip: two-state relaxation, p1 + p2 = 1 - exp(-(k1+k2) dt), ratio k1:k2
        p1, p2 = o_prob(k1, k2 if k2 > 0 else 1.0, dt, 1.0, 13)
        kb = k2 if k2 > 0 else 1.0
        assert abs(p1 + p2 - (1 - np.exp(-(k1 + kb) * dt))) < 1e-15 and abs(p1 / p2 - k1 / kb) < 1e-12
    assert o_prob(-1, 0.5, 0.01, 1.0, 10)[0] == 1  # a maximal rate always acts


def test_experfcm1_matches_scipy():
    special = pytest.importorskip("scipy.special")
    L = ob.lib()
    for x in np.concatenate([np.linspace(-0.05, 0.05, 101), np.linspace(-5, 30, 351)]):
        want = special.erfcx(x) - 1
        got = L.osurf_experfcm1(float(x))
        # the series near 0 uses sqrt(pi) truncated to 1.7724538509 (rel 3e-12);
        # elsewhere erfc has the Numerical Recipes fractional error < 1.2e-7
        tol = 1e-10 * abs(x) if abs(x) <= 0.05 else 1.3e-7 * special.erfcx(x)
        assert abs(got - want) <= tol + 1e-300, (x, got, want)


def test_two_sided_transmission_closed_form():
    """p_i = kappa_i sqrt(pi dt / D_i) cterm with
    cterm = (2c/sqrt(pi) + erfcx(c) - 1) / c^2, c = sqrt(dt) sum_i kappa_i / sqrt(D_i)
    (SurfaceParam.c:341-345), checked with scipy's erfcx where the reference's
    erfc approximation (fractional error 1.2e-7) allows; cterm -> 1 as
    the kappas vanish."""
    special = pytest.importorskip("scipy.special")
    for d1, d2, dt in [(1.0, 1.0, 0.01), (0.3, 5.0, 0.001), (4.0, 0.5, 0.2)]:
        for k1, k2 in [(2.0, 5.0), (10.0, 1.0), (0.7, 30.0)]:
            rc, p1, p2 = o_transmit(k1, k2, d1, d2, dt)
            c = np.sqrt(dt) * (k1 / np.sqrt(d1) + k2 / np.sqrt(d2))
            cterm = (2 * c / np.sqrt(np.pi) + special.erfcx(c) - 1) / c ** 2
            q1, q2 = k1 * np.sqrt(np.pi * dt / d1) * cterm, k2 * np.sqrt(np.pi * dt / d2) * cterm
            if q1 > 1:  # capped, the other side keeping the ratio
                q1, q2 = 1.0, k2 / k1 * np.sqrt(d1 / d2)
            if q2 > 1:
                q1, q2 = k1 / k2 * np.sqrt(d2 / d1), 1.0
            # erfc's fractional error, amplified by the cancellation in cterm
            tol = 1.3e-7 * special.erfcx(c) / (c * c * cterm)
            assert rc == 0 and abs(p1 - q1) <= tol * q1 and abs(p2 - q2) <= tol * q2
        rc, p1, p2 = o_transmit(1e-4, 3e-4, d1, d2, dt)
        assert abs(p1 / (1e-4 * np.sqrt(np.pi * dt / d1)) - 1) < 1e-3
        # both sides maximal: the side of faster diffusion is scaled by the flux ratio
        rc, p1, p2 = o_transmit(-1, -1, d1, d2, dt)
        assert (p1, p2) == ((1, np.sqrt(d1 / d2)) if d1 <= d2 else (np.sqrt(d2 / d1), 1))
    assert o_transmit(-1, 2.0, 1.0, 1.0, 0.01)[0] == 3  # the reference's randomized search: refused


# --- the library's restatement -------------------------------------------------

SHIM = r'''
#include "rxnparam.cpp"
#include "surfparam.cpp"
extern "C" {
double lib_surfaceprob(double k1, double k2, double dt, double difc, double* p2, int algo) {
  return smolamd::surfparam::surfaceprob(k1, k2, dt, difc, p2, algo);
}
int lib_transmit(double k1, double k2, double d1, double d2, double dt, double* p1, double* p2) {
  return smolamd::surfparam::transmit_probs(k1, k2, d1, d2, dt, p1, p2);
}
double lib_experfcm1(double x) { return smolamd::surfparam::experfcm1(x); }
}
'''


@pytest.fixture(scope="module")
def libsp(tmp_path_factory):
    d = tmp_path_factory.mktemp("surfparam")
    (d / "shim.cpp").write_text(SHIM)
    so = d / "libsp.so"
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off", "-I", str(CSRC),
                    "-o", str(so), str(d / "shim.cpp")], check=True)
    L = C.CDLL(str(so))
    D, dp = C.c_double, C.POINTER(C.c_double)
    L.lib_surfaceprob.restype, L.lib_surfaceprob.argtypes = D, [D, D, D, D, dp, C.c_int]
    L.lib_transmit.restype, L.lib_transmit.argtypes = C.c_int, [D, D, D, D, D, dp, dp]
    L.lib_experfcm1.restype, L.lib_experfcm1.argtypes = D, [D]
    return L


def test_library_restatement_equals_the_oracle(libsp):
    rng = np.random.default_rng(91)
    data = json.loads(GOLD.read_text())
    cases = [tuple(r[:5]) for r in data["surfaceprob"]]
    for _ in range(5000):
        k1 = -1.0 if rng.random() < 0.1 else float(np.exp(rng.uniform(-8, 9)))
        k2 = -1.0 if rng.random() < 0.1 else float(np.exp(rng.uniform(-8, 9)))
        cases.append((k1, k2, float(np.exp(rng.uniform(-9, 0))), float(np.exp(rng.uniform(-4, 4))),
                      int(rng.choice(ALGOS))))
    for k1, k2, dt, difc, algo in cases:
        if algo == 3 and mixed_maximal(k1, k2):
            continue
        p2 = C.c_double(0)
        got = (libsp.lib_surfaceprob(k1, k2, dt, difc, C.byref(p2), int(algo)), p2.value)
        assert same(got, o_prob(k1, k2, dt, difc, int(algo))), (k1, k2, dt, difc, algo)
    for k1, k2, d1, d2, dt, *_ in data["transmit"]:
        p1, p2 = C.c_double(-2), C.c_double(-2)
        rc = libsp.lib_transmit(k1, k2, d1, d2, dt, C.byref(p1), C.byref(p2))
        assert same((rc, p1.value, p2.value), o_transmit(k1, k2, d1, d2, dt))
    assert libsp.lib_transmit(-1, 2.0, 1.0, 1.0, 0.01, None, None) == 3
    for x in np.concatenate([rng.normal(0, 8, 2000), rng.uniform(-0.06, 0.06, 500)]):
        assert same(libsp.lib_experfcm1(float(x)), ob.lib().osurf_experfcm1(float(x)))
