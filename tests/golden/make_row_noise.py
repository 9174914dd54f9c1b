"""Writes row_noise.json: for each full-size Galerkin row sample of tests/row_sample.py, the relative Frobenius
distance between the FP64 oracle's rows and its long-double rows, per matrix and per time-domain weight n.

This is the FP64 noise floor of the oracle on those rows (per-frequency rounding amplified by the R^-n factor of
the inverse transform).  The GPU tests bound the distance of the GPU rows from the FP64 oracle's rows by a multiple
of this profile.  The oracle's sums do not depend on the number of threads, so the FP64 rows are the same on any
host; a CRC-32 of each matrix's FP64 rows is stored with the profile, and the tests check that they recompute
exactly those rows.  Run from the repository root:
    python tests/golden/make_row_noise.py        (about 15 minutes on 8 cores)"""
import json
import os
import sys
import time
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from row_sample import GALERKIN_CONFIGS, galerkin_case, oracle_rows, rel  # noqa: E402


def main():
    out = {}
    for name in GALERKIN_CONFIGS:
        V, F, prm, row = galerkin_case(name)
        t0 = time.time()
        f64 = oracle_rows(O, V, F, prm, row, True, O.MODE_FP64)
        t1 = time.time()
        ld = oracle_rows(O, V, F, prm, row, True, O.MODE_LD)
        t2 = time.time()
        rec = {"row_vertex": row, "nv": len(V), "L": prm["L"], "fp64_s": t1 - t0, "ld_s": t2 - t1, "per_weight": {},
               "stack": {}}
        rec["fp64_crc32"] = {}
        for k in f64:
            rec["per_weight"][k] = [rel(f64[k][n], ld[k][n]) for n in range(prm["L"])]
            rec["stack"][k] = rel(f64[k], ld[k])
            rec["fp64_crc32"][k] = zlib.crc32(np.ascontiguousarray(f64[k], np.float64).tobytes())
        out[name] = rec
        print(name, f"fp64 {t1 - t0:.0f} s, ld {t2 - t1:.0f} s",
              {k: f"{v:.1e}" for k, v in rec["stack"].items()}, flush=True)
    with open(os.path.join(HERE, "row_noise.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
