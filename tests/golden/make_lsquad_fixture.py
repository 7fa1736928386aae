"""Write tests/golden/lsquad.json from the reference's level-symmetric tables.

The reference reads its angular quadrature from data/lsquad/LS_<N>.txt
(src/angular_quadrature.cpp:114-140): a "Total Directions = K" header, then
K rows of "mu eta xi w".  This script copies those rows (data, not code) for
the orders the tests use into one JSON fixture, so that tests never need the
reference tree at run time.

    python tests/golden/make_lsquad_fixture.py /root/reference/data/lsquad
"""
import json
import os
import re
import sys

ORDERS = (2, 4, 6, 8, 10, 12, 14, 16)


def main(src_dir):
    out = {}
    for n in ORDERS:
        with open(os.path.join(src_dir, f"LS_{n}.txt")) as fh:
            lines = fh.read().splitlines()
        k = int(re.search(r"Total Directions = ([0-9]+)", lines[0]).group(1))
        rows = [[float(v) for v in ln.split()[:4]] for ln in lines[1:1 + k]]
        assert len(rows) == k, (n, len(rows), k)
        out[str(n)] = rows
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lsquad.json")
    with open(dst, "w") as fh:
        json.dump({"source": "reference data/lsquad/LS_<N>.txt rows: mu eta xi w", "sets": out}, fh)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data/lsquad")
