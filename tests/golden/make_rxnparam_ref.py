"""Write rxnparam_ref.json: values of the reference's own rxnparam functions,
from the compiled reference libSteve (oracle/_ref/libsteve_ref.so, built by
`make -C oracle ref` from /root/reference).  Run in the build container."""
import json
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import oracle_bind as ob  # noqa: E402


def main():
    R = ob.reflib()
    assert R is not None, "build oracle/_ref first"
    out = {"bindingradius": [], "numrxnrate": [], "unbindingradius": []}
    for rate in (9.640578597, 21.42267826, 99.3218287, 8000.0, 1.0):
        for dt, difc in ((0.01, 2.0), (0.002, 4.0), (0.005, 2.0), (0.001, 200.0), (0.0, 1.0)):
            for b, rel in ((-1, 0), (1.2, 1), (2.0, 0)):
                out["bindingradius"].append([rate, dt, difc, b, rel, R.bindingradius(rate, dt, difc, b, rel)])
    for s in (0.01, 0.1, 0.2, 0.5, 1.0, 3.0, 30.0):
        for b in (-1, 0.0, 0.5, 1.0, 1.2, 3.0, 40.0):
            out["numrxnrate"].append([s, 1.0, b, R.numrxnrate(s, 1.0, b)])
    for pgem in (0.1, 0.2, 0.5):
        for a in (0.5, 1.0):
            out["unbindingradius"].append([pgem, 0.005, 3.0, a, R.unbindingradius(pgem, 0.005, 3.0, a)])
    Path(__file__).with_name("rxnparam_ref.json").write_text(json.dumps(out, indent=0))


if __name__ == "__main__":
    main()
