"""Writes row_noise.json: for each full-size Galerkin row sample of tests/row_sample.py, the relative Frobenius
distance between the FP64 oracle's rows and its long-double rows, per matrix and per time-domain weight n.

This is the FP64 noise floor of the oracle on those rows (per-frequency rounding amplified by the R^-n factor of
the inverse transform).  The GPU tests require the GPU rows to stay within 3x this profile of the FP64 oracle's
rows, which keeps them within ~4x of it from the long-double rows.  Run from the repository root:
    python tests/golden/make_row_noise.py        (about 15 minutes on 8 cores)"""
import json
import os
import sys
import time

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
        for k in f64:
            rec["per_weight"][k] = [rel(f64[k][n], ld[k][n]) for n in range(prm["L"])]
            rec["stack"][k] = rel(f64[k], ld[k])
        out[name] = rec
        print(name, f"fp64 {t1 - t0:.0f} s, ld {t2 - t1:.0f} s",
              {k: f"{v:.1e}" for k, v in rec["stack"].items()}, flush=True)
    with open(os.path.join(HERE, "row_noise.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
