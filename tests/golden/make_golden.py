"""Write tests/golden/lebedev_reference.npz from the reference's Lebedev tables.

Run in the build container (the only place the reference is mounted):
    python tests/golden/make_golden.py
The tables (lebedev_{000,003,005,007,009,011}.txt: azimuth and polar angle in
degrees, weight) are data the reference reads in SpinSymm.py:47-58; they are
stored here unchanged as arrays so that tests can run without the reference.
"""
import os

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lebedev_reference.npz")

if __name__ == "__main__":
    arrays = {}
    for npts, tag in [(1, "000"), (6, "003"), (14, "005"), (26, "007"), (38, "009"), (50, "011")]:
        arrays[f"n{npts}"] = np.loadtxt(os.path.join(REF, f"lebedev_{tag}.txt")).reshape(-1, 3)
    np.savez(OUT, **arrays)
    print("wrote", OUT)
