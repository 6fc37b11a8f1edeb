"""Writes r2r_golden.npz: inputs and outputs of the eight r2r plans of the
reference (smaiso-wcomp.cpp:499-545) on small grids.  The outputs come from
direct O(N^2) sums written from the FFTW manual's definitions of REDFT10,
REDFT01, RODFT10 and RODFT01 (oracle.direct_r2r3d), not from any FFT.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import smaiso_oracle as O  # noqa: E402

PLANS = {"CT": O.CT, "ICT": O.ICT, "STx": O.ST[0], "STy": O.ST[1], "STz": O.ST[2],
         "ISTx": O.IST[0], "ISTy": O.IST[1], "ISTz": O.IST[2]}
SHAPES = [(8, 6, 4), (4, 5, 3), (16, 4, 2), (2, 16, 4), (4, 2, 16)]


def main():
    out = {}
    rng = np.random.default_rng(20210327)
    for shape in SHAPES:
        for name, kinds in PLANS.items():
            x = rng.uniform(-1, 1, shape)
            key = f"{name}@{'x'.join(map(str, shape))}"
            out["in_" + key] = x
            out["out_" + key] = O.direct_r2r3d(x, kinds)
    np.savez_compressed(os.path.join(HERE, "r2r_golden.npz"), **out)
    print("wrote", len(out) // 2, "cases")


if __name__ == "__main__":
    main()
