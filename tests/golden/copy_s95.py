"""Copy the reference's regression inputs and golden outputs used by
tests/test_oracle_golden.py into tests/golden/s95/ (configs as text, outputs
gzipped).  These are data files the reference's own regression suite holds
(examples/S95_regression/*.txt and output_275/*_275_out.txt)."""
import gzip
import shutil
from pathlib import Path

SRC = Path("/root/reference/examples/S95_regression")
DST = Path(__file__).resolve().parent / "s95"
MODELS = ["bounce2", "molecule", "lotvolt", "bounds1", "surf1", "ReflectToNewSpecies", "allostery", "surf3"]


def main():
    DST.mkdir(exist_ok=True)
    for m in MODELS:
        shutil.copy(SRC / (m + ".txt"), DST / (m + ".txt"))
        data = (SRC / "output_275" / (m + "_275_out.txt")).read_bytes()
        with gzip.open(DST / (m + "_275_out.txt.gz"), "wb", compresslevel=9) as f:
            f.write(data)


if __name__ == "__main__":
    main()
