"""CQBEM precompute benchmark on one MI355X.

Workload (one "step" = one full precompute of one object):
  icosphere subdivision level 4 (2562 vertices, 5120 faces), radius 0.2 m, 1% seeded jitter,
  E = 5e5 Pa, nu = 0.3, rho = 1000 kg/m^3, dt = 1e-3 s, BDF2, L = 32, R = 1e-10^(1/(2L)),
  collocation, quadrature order 5 with subdivision 2; G, H, B_trans and B_angular, all weights
  n = 0..L-1, outputs resident in HBM.

Prints one JSON line: G+H time-domain entries per second (2 (3 nv)^2 L per step), p50/p99 step
latency, the dominant kernel's roofline (VALU FP64) and the CPU oracle's reference-precision
baseline on a bounded row sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CQBEM G+H entries/s mixed vs FP64 at 1e-6 rel err; p50/p99 assembly latency"
PEAK_FP64_TFLOPS = 78.6
# Algorithmic FLOPs per far-field kernel evaluation (collocation point, source point, frequency),
# the frozen per-unit model: G 168 + H 348.  Transcendentals (2 complex exponentials) not counted.
FLOPS_PER_EVAL_GH = 168 + 348


def workload(precision: str):
    from surfaceonlydeformables_amd import abi
    from surfaceonlydeformables_amd import mesh as M
    V, F = M.icosphere(4, 0.2)
    V = M.center(M.jitter(V, F, 0.01, 0), F)
    E, nu, rho, dt, L = 5e5, 0.3, 1000.0, 1e-3, 32
    _, c1, c2 = M.material_speeds(E, nu, rho)
    R = 1e-10 ** (1.0 / (2 * L))
    prec = abi.FP64 if precision == "fp64" else abi.MIXED
    params = abi.make_params(c1, c2, rho, nu, dt, R, L, bdf=abi.BDF2, order=5, subdivision=2, precision=prec)
    desc = ("icosphere-L4 r=0.2 m 1% jitter, E=5e5 Pa nu=0.3 rho=1000, dt=1e-3 s BDF2 L=32 R=1e-10^(1/64), "
            "collocation order 5 x subdivision 2, G+H+B_trans+B_angular")
    return V, F, params, desc


def cpu_baseline(V, F, params, seconds_target=15.0):
    """Oracle in the reference's own precision typedefs, all host threads, on a row sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nv, L = len(V), params.L
    Lf = L // 2 + 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    op = O.params(params.c1, params.c2, params.rho, params.nu, params.dt, params.R, L, bdf=params.bdf,
                  order=params.order, subdiv=params.subdivision, mode=O.MODE_REFMIXED, nthreads=threads)
    s = O.frequencies(L, params.R, params.dt, params.bdf)
    ls = [0, Lf // 2]
    rows = max(threads, 16)
    t0 = time.perf_counter()
    for l in ls:
        O.laplace(V, F, op, s[l], mask=15, rows=(0, rows))
    t_quad = time.perf_counter() - t0
    # inverse transform of the two G/H matrices, timed on a slice of entries
    M_sample = 1 << 19
    spec = (np.random.default_rng(0).standard_normal((Lf, M_sample))
            + 1j * np.random.default_rng(1).standard_normal((Lf, M_sample)))
    t1 = time.perf_counter()
    O.transform(spec, L, params.R)
    t_tr = time.perf_counter() - t1
    n3 = 3 * nv
    full = t_quad * (nv / rows) * (Lf / len(ls)) + t_tr * (2 * n3 * n3 / M_sample)
    entries = 2 * n3 * n3 * L
    return {"value": entries / full, "unit": "entries/s", "cores": threads, "kind": "port",
            "sample": (f"oracle (reference float/double typedefs), G+H+B_trans+B_angular Laplace rows 0..{rows - 1} "
                       f"of {nv} at frequencies {ls} of {Lf} ({t_quad:.1f} s) + inverse transform of "
                       f"{M_sample} entries ({t_tr:.1f} s), extrapolated to the full job: {full:.0f} s")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="one untimed assembly followed by a 1 GiB device copy, for HBM counter runs "
                         "(the copy calibrates FETCH_SIZE / WRITE_SIZE); prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the CQBEM precompute of one object runs on one MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from surfaceonlydeformables_amd import abi

    V, F, params, desc = workload(args.precision)
    nv, L = len(V), params.L
    n3 = 3 * nv
    mesh = abi.make_mesh(V, F)
    ctx = abi.Context(0)
    dev = torch.device("cuda:0")
    outs = {"G": torch.empty((L, n3, n3), dtype=torch.float64, device=dev),
            "H": torch.empty((L, n3, n3), dtype=torch.float64, device=dev),
            "B_trans": torch.empty((L, n3, 3), dtype=torch.float64, device=dev),
            "B_angular": torch.empty((L, n3, 3, 3), dtype=torch.float64, device=dev),
            "B": torch.empty((L, n3, 6), dtype=torch.float64, device=dev)}
    ptrs = {k: t.data_ptr() for k, t in outs.items()}

    if args.traffic_probe:
        ctx.assemble_device(mesh, params, ptrs)
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"traffic probe: {ctx.last_timings()['gh_evaluations']:.0f} G/H evaluations, 1 GiB copy", flush=True)
        ctx.close()
        return

    for _ in range(args.warmup):
        ctx.assemble_device(mesh, params, ptrs)
    torch.cuda.synchronize()
    lat, gh_ms, gh_evals, gh_launches = [], 0.0, 0.0, 0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ctx.assemble_device(mesh, params, ptrs)
        lat.append((time.perf_counter() - t0) * 1e3)
        tm = ctx.last_timings()
        gh_ms += tm["regular_gh"]
        gh_evals += tm["gh_evaluations"]
        gh_launches += int(tm["gh_launches"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1e3 / args.steps
    entries = 2.0 * n3 * n3 * L
    value = entries * args.steps / wall
    last = ctx.last_timings()

    achieved = FLOPS_PER_EVAL_GH * gh_evals / (gh_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "kernel": "k_colloc_gh_pair<true,true>", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
            "launch_ms": gh_ms / max(gh_launches, 1),
            "flops_per_launch": FLOPS_PER_EVAL_GH * gh_evals / max(gh_launches, 1)}
    prof = os.path.join(ROOT, "profiles", "k_colloc_gh_pair_icosphere_l4_traffic.json")
    if os.path.exists(prof):
        with open(prof) as fh:
            roof["traffic"] = json.load(fh).get("hbm_bytes_per_launch")

    line = {"metric": METRIC, "value": value, "unit": "entries/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f64+f32",
            "data": "synthetic",
            "config": {"workload": desc, "nv": nv, "nf": len(F), "L": L, "precision": args.precision,
                       "entries_per_step": entries},
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
            "phases_ms_last_step": {k: last[k] for k in ("regular_gh", "regular_b", "singular", "transform", "host")},
            "roofline": roof}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(V, F, params)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
