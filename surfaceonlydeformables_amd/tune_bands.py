"""Tuner of CQBEM_QUAD_BANDED's distance thresholds for a workload, on the GPU: kappa (G and H, per patch tile) and
kappa_b (B_trans and B_angular, per face).

For each candidate value (run as kappa = kappa_b) it assembles G, H, B_trans and B_angular with the banded rule and
compares them per time-domain weight with a refined rule (the reference order at twice the subdivision), next to the
reference rule's own distance from it.  A value is accepted for G/H (for B) when the error it adds to the reference
rule's, at any weight of G or H (of B_trans or B_angular), is at most `tol` times the reference rule's largest
per-weight error.  G and H depend on kappa only and B on kappa_b only, so the two are chosen apart: kappa the accepted
value with the fewest G/H evaluations, kappa_b the accepted value with the shortest B phase.  The result, with every candidate's measurements, is written as JSON:

    python -m surfaceonlydeformables_amd.tune_bands --workload icosphere-l4 --out band_table.json

Needs a GPU and device memory for three copies of G and H."""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

WORKLOADS = {
    # name: (mesh, E, nu, rho, dt, L)
    "icosphere-l3": ("ico", 3, 5e5, 0.3, 1000.0, 1e-3, 32),
    "icosphere-l4": ("ico", 4, 5e5, 0.3, 1000.0, 1e-3, 32),
    "cube-8": ("cube", 8, 1e5, 0.3, 1000.0, 1e-3, 32),
}


def workload(name):
    from . import mesh as M
    kind, size, E, nu, rho, dt, L = WORKLOADS[name]
    if kind == "ico":
        V, F = M.icosphere(size, 0.2)
        V = M.center(M.jitter(V, F, 0.01, 0), F)
    else:
        V, F = M.cube(size, 0.5)
        V = M.center(V, F)
    _, c1, c2 = M.material_speeds(E, nu, rho)
    return V, F, dict(c1=c1, c2=c2, rho=rho, nu=nu, dt=dt, R=1e-10 ** (1 / (2 * L)), L=L)


def per_weight_rel(torch, a, b):
    """Relative Frobenius distance of a from b per weight (device tensors (L, n, n))."""
    d = torch.linalg.vector_norm((a - b).reshape(a.shape[0], -1), dim=1)
    return (d / torch.linalg.vector_norm(b.reshape(b.shape[0], -1), dim=1)).cpu().numpy()


MATRICES = ("G", "H", "B_trans", "B_angular")


def assemble(ctx, torch, mesh, prm, names=MATRICES, **kw):
    from . import abi
    n3, L = 3 * mesh.nv, prm["L"]
    shapes = {"G": (L, n3, n3), "H": (L, n3, n3), "B_trans": (L, n3, 3), "B_angular": (L, n3, 3, 3)}
    bits = {"G": abi.G, "H": abi.H, "B_trans": abi.B_TRANS, "B_angular": abi.B_ANGULAR}
    out = {k: torch.empty(shapes[k], dtype=torch.float64, device="cuda:0") for k in names}
    torch.cuda.synchronize()
    which = 0
    for k in names:
        which |= bits[k]
    p = abi.make_params(prm["c1"], prm["c2"], prm["rho"], prm["nu"], prm["dt"], prm["R"], L, which=which, **kw)
    t0 = time.perf_counter()
    ctx.assemble_device(mesh, p, {k: t.data_ptr() for k, t in out.items()})
    wall = time.perf_counter() - t0
    return out, wall, ctx.last_timings()


def tune(name, kappas, tol, order=5, subdivision=2):
    import torch
    from . import abi
    V, F, prm = workload(name)
    mesh = abi.make_mesh(V, F)
    ctx = abi.Context(0)
    try:
        fine, _, _ = assemble(ctx, torch, mesh, prm, order=order, subdivision=2 * subdivision)
        ref, _, _ = assemble(ctx, torch, mesh, prm, order=order, subdivision=subdivision)  # warm-up
        ref, ref_wall, ref_t = assemble(ctx, torch, mesh, prm, order=order, subdivision=subdivision)
        e_ref = {k: per_weight_rel(torch, ref[k], fine[k]) for k in MATRICES}
        del ref
        rows = []
        for kappa in kappas:
            out, wall, t = assemble(ctx, torch, mesh, prm, order=order, subdivision=subdivision, quad=abi.QUAD_BANDED,
                                    band_kappa=kappa, band_kappa_b=kappa)
            e = {k: per_weight_rel(torch, out[k], fine[k]) for k in MATRICES}
            del out
            added = {k: float(np.max(e[k] - e_ref[k]) / np.max(e_ref[k])) for k in MATRICES}
            rows.append({"kappa": kappa, "accepted_gh": all(added[k] <= tol for k in ("G", "H")),
                         "accepted_b": all(added[k] <= tol for k in ("B_trans", "B_angular")),
                         "b_phase_ms": t["regular_b"],
                         "added_error_over_reference_max": added,
                         "evaluations_fraction": t["gh_evaluations"] / ref_t["gh_evaluations"],
                         "gh_phase_ms": t["regular_gh"], "wall_ms": wall * 1e3,
                         "rel_err_vs_refined_per_weight": {k: [float(x) for x in v] for k, v in e.items()}})
            print(json.dumps({k: v for k, v in rows[-1].items() if k != "rel_err_vs_refined_per_weight"}), flush=True)
        ok = [r for r in rows if r["accepted_gh"]]
        best = min(ok, key=lambda r: r["evaluations_fraction"]) if ok else None
        ok_b = [r for r in rows if r["accepted_b"]]
        best_b = min(ok_b, key=lambda r: (r["b_phase_ms"], -r["kappa"])) if ok_b else None
        return {"workload": name, "order": order, "subdivision": subdivision, "tolerance": tol,
                "reference_rule": {"gh_phase_ms": ref_t["regular_gh"], "wall_ms": ref_wall * 1e3,
                                   "rel_err_vs_refined_per_weight": {k: [float(x) for x in v] for k, v in e_ref.items()}},
                "best_kappa": best["kappa"] if best else None, "best_kappa_b": best_b["kappa"] if best_b else None,
                "candidates": rows}
    finally:
        ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="icosphere-l4")
    ap.add_argument("--kappa", type=float, nargs="+", default=[1e-300, 0.5, 1.0, 2.0, 3.0, 4.0, 6.0, 8.0, 12.0, 16.0])
    ap.add_argument("--tol", type=float, default=0.05)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    res = tune(args.workload, args.kappa, args.tol)
    print(json.dumps({"best_kappa": res["best_kappa"], "best_kappa_b": res["best_kappa_b"]}), flush=True)
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(res, fh, indent=1)


if __name__ == "__main__":
    main()
