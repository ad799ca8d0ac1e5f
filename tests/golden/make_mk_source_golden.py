"""Writes tests/golden/mk_source_int1.json: a subsample of the reference's own charge-exchange
rate table I(U, c_p) (MK_source/MK_source/MK_source/MK_source.cpp:38-103), computed by the
reference program compiled from its source into oracle/_ref/mk_source (oracle/Makefile.ref).

The program writes Int_1.bin (1000 x 100 doubles, U-major) into its working directory; the table
grid is U_i = 0.01 + i (200 - 0.01)/999, c_j = 0.01 + j (100 - 0.01)/99.  Its OpenMP reduction
makes the last bits depend on the thread count, so the fixture is compared to 1e-12 relative.
"""
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "mk_source")


def main():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-f", "Makefile.ref"], check=True)
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([EXE], cwd=d, check=True, stdout=subprocess.DEVNULL)
        tab = np.fromfile(os.path.join(d, "Int_1.bin"), dtype=np.float64).reshape(1000, 100)
    pts = []
    for i in (0, 1, 7, 50, 123, 400, 777, 999):
        for j in (0, 1, 3, 20, 57, 99):
            U = 0.01 + i * (200.0 - 0.01) / (1000 - 1)
            c = 0.01 + j * (100.0 - 0.01) / (100 - 1)
            pts.append({"i": i, "j": j, "U": U, "cp": c, "I": float(tab[i, j])})
    out = {"source": "MK_source/MK_source/MK_source/MK_source.cpp:38-103 (compiled, run)",
           "a_2": 0.1307345665, "N1": 100, "N2": 100, "points": pts}
    with open(os.path.join(HERE, "mk_source_int1.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote %d points" % len(pts))


if __name__ == "__main__":
    main()
