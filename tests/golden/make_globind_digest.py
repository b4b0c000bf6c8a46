"""Writes globind_ne30_p3.json: a digest of the reference's Glob fixture.

The reference ships the single-rank Int64 Glob array of the ne=30, p=3 cubed
sphere as ``TestKernels/GlobInd`` (691,200 bytes, shape (16, 5400)
column-major).  This script reads that data file (it is data, not code) and
stores its SHA-256, its first element and its node multiplicity histogram,
which the CPU tests compare against this build's numbering.
"""
import hashlib
import json
import os
import sys

import numpy as np

SRC = "/root/reference/TestKernels/GlobInd"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "globind_ne30_p3.json")


def main():
    raw = open(SRC, "rb").read()
    a = np.frombuffer(raw, dtype="<i8")
    glob = a.reshape(5400, 16)  # row IF = column IF of the Julia (16, 5400) array
    vals, counts = np.unique(glob, return_counts=True)
    hist = {int(k): int(v) for k, v in zip(*np.unique(counts, return_counts=True))}
    out = {"source": "TestKernels/GlobInd (reference data file)", "ne": 30, "p": 3, "dtype": "int64 little endian",
           "shape_julia": [16, 5400], "sha256": hashlib.sha256(raw).hexdigest(), "num_g": int(vals.max()),
           "first_element": [int(x) for x in glob[0]], "last_element": [int(x) for x in glob[-1]],
           "multiplicity_histogram": hist}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out)[:300])


if __name__ == "__main__":
    sys.exit(main())
