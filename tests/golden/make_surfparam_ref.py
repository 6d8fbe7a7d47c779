"""Write surfparam_ref.json: values of the reference's own surface parameter
functions (surfaceprob, surfacetransmit's kappa -> probability branch,
lookupirrevadsorb, lookuprevads, experfcm1D) from the compiled reference
libSteve (oracle/_ref/libsteve_ref.so, built by `make -C oracle ref` from
/root/reference).  Run in the build container."""
import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import oracle_bind as ob  # noqa: E402

# SurfParamAlgo codes (SurfaceParam.h:11) whose branches srfcalcprob reaches,
# plus the closed-form adsorption fits; aliases (IrrTransT, IrrTransQ,
# IrrAdsT) and RevAdsND are left out
ALGOS = [0, 3, 4, 6, 7, 8, 10, 11, 12, 13]
KS = [-1.0, 0.0, 0.02, 1.5, 40.0, 3000.0]
DTS = [1e-3, 0.05]
DIFCS = [0.2, 3.0]
TRANS_K = [-1.0, 0.0, 1e-4, 0.3, 2.0, 25.0, 900.0]
TRANS_D = [(1.0, 1.0), (0.5, 4.0), (6.0, 0.25)]


def maximal_mixed(k1, k2):
    """One side maximal and the other finite and nonzero: the reference
    searches with random draws there."""
    return (k1 == -1) != (k2 == -1) and (k1 not in (0, -1) or k2 not in (0, -1))


def surfaceprob(R, k1, k2, dt, difc, algo):
    p2 = C.c_double(0)
    p1 = R.surfaceprob(k1, k2, dt, difc, C.byref(p2), algo)
    return p1, p2.value


def transmit(R, k1, k2, d1, d2, dt):
    a, b = C.c_double(k1), C.c_double(k2)
    p1, p2 = C.c_double(-2), C.c_double(-2)
    rc = R.surfacetransmit(C.byref(a), C.byref(b), C.byref(p1), C.byref(p2), d1, d2, dt)
    return rc, p1.value, p2.value


def main():
    R = ob.reflib()
    assert R is not None, "build oracle/_ref first"
    out = {"surfaceprob": [], "transmit": [], "irrev": [], "rev": [], "experfcm1": []}
    for algo in ALGOS:
        for k1 in KS:
            for k2 in KS:
                if algo == 3 and maximal_mixed(k1, k2):
                    continue
                for dt in DTS:
                    for difc in DIFCS:
                        out["surfaceprob"].append([k1, k2, dt, difc, algo, *surfaceprob(R, k1, k2, dt, difc, algo)])
    for k1 in TRANS_K:
        for k2 in TRANS_K:
            if maximal_mixed(k1, k2):
                continue
            for d1, d2 in TRANS_D:
                for dt in DTS:
                    out["transmit"].append([k1, k2, d1, d2, dt, *transmit(R, k1, k2, d1, d2, dt)])
    rng = np.random.default_rng(2024)
    for v in np.concatenate([[-0.1, 0.0, 0.858559, 0.9], rng.uniform(0, 0.9, 120)]):
        out["irrev"].append([float(v), R.lookupirrevadsorb(float(v), 1)])
    for _ in range(200):
        v1, v2 = (float(x) for x in np.exp(rng.uniform(-7, 6, 2)))
        p2 = C.c_double(0)
        out["rev"].append([v1, v2, R.lookuprevads(v1, v2, 1, C.byref(p2)), p2.value])
    for x in np.concatenate([[-25.0, -0.05, 0.0, 0.05, 0.0500001, 19.99, 20.0, 35.0], rng.normal(0, 3, 60),
                             rng.uniform(-0.05, 0.05, 20)]):
        out["experfcm1"].append([float(x), R.experfcm1D(float(x))])
    Path(__file__).with_name("surfparam_ref.json").write_text(json.dumps(out, separators=(",", ":")))


if __name__ == "__main__":
    main()
