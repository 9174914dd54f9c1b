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


WORKLOADS = ("ico4-collocation", "cube8-collocation", "ico3-galerkin", "ico4-galerkin")


def workload(name: str, precision: str):
    """The single-object workloads: ico4-collocation (the default line), the cube k = 8 collocation case, the
    icosphere-L3 Galerkin case (r = 0.1 m, E = 1e6 Pa, nu = 0.35, rho = 1100, dt = 5e-4 s, L = 64) and
    icosphere-L4 in Galerkin form."""
    from surfaceonlydeformables_amd import abi
    from surfaceonlydeformables_amd import mesh as M
    prec = abi.FP64 if precision == "fp64" else abi.MIXED
    if name == "cube8-collocation":
        V, F = M.cube(8, 0.5)
        V = M.center(V, F)
        E, nu, rho, dt, L, galerkin = 1e5, 0.3, 1000.0, 1e-3, 32, False
        shape = "cube k=8 (386 V) side 1 m, E=1e5 Pa nu=0.3 rho=1000, dt=1e-3 s BDF2 L=32, collocation"
    elif name == "ico3-galerkin":
        V, F = M.icosphere(3, 0.1)
        V = M.center(M.jitter(V, F, 0.01, 0), F)
        E, nu, rho, dt, L, galerkin = 1e6, 0.35, 1100.0, 5e-4, 64, True
        shape = "icosphere-L3 (642 V) r=0.1 m 1% jitter, E=1e6 Pa nu=0.35 rho=1100, dt=5e-4 s BDF2 L=64, Galerkin"
    else:
        V, F = M.icosphere(4, 0.2)
        V = M.center(M.jitter(V, F, 0.01, 0), F)
        E, nu, rho, dt, L = 5e5, 0.3, 1000.0, 1e-3, 32
        galerkin = name == "ico4-galerkin"
        shape = ("icosphere-L4 (2562 V) r=0.2 m 1% jitter, E=5e5 Pa nu=0.3 rho=1000, dt=1e-3 s BDF2 L=32, "
                 + ("Galerkin" if galerkin else "collocation"))
    _, c1, c2 = M.material_speeds(E, nu, rho)
    R = 1e-10 ** (1.0 / (2 * L))
    params = abi.make_params(c1, c2, rho, nu, dt, R, L, bdf=abi.BDF2, order=5, subdivision=2, precision=prec,
                             galerkin=galerkin)
    desc = f"{shape}, R=1e-10^(1/(2L)), order 5 x subdivision 2, G+H+B_trans+B_angular"
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
    ls = sorted({0, Lf // 4, Lf // 2, (3 * Lf) // 4, Lf - 1})  # across the spectrum: real s at l = 0, complex after
    rows = max(2 * threads, 32)
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
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")), "")
    except OSError:
        pass
    # the same quadrature on one core (reference precision), and on all cores in FP64 (the oracle's mode), each
    # on a smaller row sample, extrapolated the same way
    def quad_rate(mode, nthreads, nrows):
        q = O.params(params.c1, params.c2, params.rho, params.nu, params.dt, params.R, L, bdf=params.bdf,
                     order=params.order, subdiv=params.subdivision, mode=mode, nthreads=nthreads)
        t = time.perf_counter()
        for l in ls:
            O.laplace(V, F, q, s[l], mask=15, rows=(0, nrows))
        return (time.perf_counter() - t) * (nv / nrows) * (Lf / len(ls)) + t_tr * (2 * n3 * n3 / M_sample)
    full_1 = quad_rate(O.MODE_REFMIXED, 1, 2)
    full_64 = quad_rate(O.MODE_FP64, threads, rows)
    return {"value": entries / full, "unit": "entries/s", "cores": threads, "kind": "port", "cpu_model": model,
            "host_cpus": os.cpu_count(), "one_core_value": entries / full_1,
            "fp64_all_cores_value": entries / full_64,
            "sample": (f"oracle (reference float/double typedefs), G+H+B_trans+B_angular Laplace rows 0..{rows - 1} "
                       f"of {nv} at frequencies {ls} of {Lf} ({t_quad:.1f} s) + inverse transform of "
                       f"{M_sample} entries ({t_tr:.1f} s), extrapolated to the full job: {full:.0f} s; "
                       f"one_core_value: rows 0..1 at the same frequencies on one thread; fp64_all_cores_value: rows "
                       f"0..{rows - 1} at the same frequencies in the oracle's FP64 mode; both add the same "
                       f"(all-thread) transform sample")}


def scene_jobs(n_objects=16, seed=0):
    """The seeded multi-object scene: icospheres of level 1-4 or k x k-per-face cubes (k in [4, 16]),
    radius / half-size log-uniform in [0.05, 0.3] m, E log-uniform in [1e4, 1e6] Pa, nu in [0.2, 0.45],
    rho = 1000, L in {16, 32, 64}, dt = 1e-3 s, BDF2, collocation order 5 x subdivision 2."""
    from surfaceonlydeformables_amd import abi
    from surfaceonlydeformables_amd import mesh as M
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n_objects):
        size = float(np.exp(rng.uniform(np.log(0.05), np.log(0.3))))
        if rng.uniform() < 0.5:
            V, F = M.icosphere(int(rng.integers(1, 5)), size)
        else:
            V, F = M.cube(int(rng.integers(4, 17)), size)
        V = M.center(M.jitter(V, F, 0.01, seed * 1000 + k), F)
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
        nu = float(rng.uniform(0.2, 0.45))
        L = int(rng.choice([16, 32, 64]))
        _, c1, c2 = M.material_speeds(E, nu, 1000.0)
        params = abi.make_params(c1, c2, 1000.0, nu, 1e-3, 1e-10 ** (1.0 / (2 * L)), L, bdf=abi.BDF2, order=5,
                                 subdivision=2)
        jobs.append((V, F, params))
    return jobs


def run_jobs(ctx, jobs, arrivals, torch, outs=None):
    """Feeds jobs to the serving API at the given arrival times (seconds from start; BUSY answers wait
    for a running job), keeps each job's outputs until it completes (allocated per job, or outs[k] when given),
    and returns per-job latencies (completion - arrival) and the makespan."""
    from surfaceonlydeformables_amd import abi
    dev = torch.device("cuda:0")
    queue, running, lat = [], {}, [None] * len(jobs)
    meshes = [abi.make_mesh(V, F) for V, F, _ in jobs]
    stats = {"max_running": 0, "refused": {}}
    head_outs = None  # outputs of the queue's first job, kept across refused submits
    t0 = time.perf_counter()
    nxt = 0
    while nxt < len(jobs) or queue or running:
        now = time.perf_counter() - t0
        while nxt < len(jobs) and arrivals[nxt] <= now:
            queue.append(nxt)
            nxt += 1
        while queue:
            k = queue[0]
            V, F, p = jobs[k]
            n3 = 3 * len(V)
            if head_outs is None and outs is not None:
                head_outs = outs[k]
            if head_outs is None:
                try:
                    head_outs = {"G": torch.empty((p.L, n3, n3), dtype=torch.float64, device=dev),
                                 "H": torch.empty((p.L, n3, n3), dtype=torch.float64, device=dev),
                                 "B": torch.empty((p.L, n3, 6), dtype=torch.float64, device=dev)}
                except torch.OutOfMemoryError:
                    if not running:
                        raise
                    stats["refused"]["torch out of memory"] = stats["refused"].get("torch out of memory", 0) + 1
                    break  # the outputs of running jobs hold the memory: wait for one to finish
            jid = ctx.submit(meshes[k], p, {n: t.data_ptr() for n, t in head_outs.items()})
            if jid < 0:
                why = abi.lib().cqbem_last_error(ctx._h).decode()
                stats["refused"][why] = stats["refused"].get(why, 0) + 1
                if not running:  # device memory held in torch's cache: give it back before the retry
                    torch.cuda.empty_cache()
                break
            running[jid] = (k, head_outs)
            head_outs = None
            queue.pop(0)
            stats["max_running"] = max(stats["max_running"], len(running))
        for jid in [j for j in running if ctx.poll(j)]:
            ctx.wait(jid)
            k, _ = running.pop(jid)
            lat[k] = time.perf_counter() - t0 - arrivals[k]
        if not queue and nxt < len(jobs) and not running:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
        else:
            time.sleep(0.0005)
    torch.cuda.synchronize()
    run_jobs.stats = stats
    return lat, time.perf_counter() - t0


def scene_outputs(jobs, torch):
    """Every object's G, H and B outputs on the device (allocated once, outside every timed region)."""
    dev = torch.device("cuda:0")
    outs = []
    for V, F, p in jobs:
        n3 = 3 * len(V)
        outs.append({"G": torch.empty((p.L, n3, n3), dtype=torch.float64, device=dev),
                     "H": torch.empty((p.L, n3, n3), dtype=torch.float64, device=dev),
                     "B": torch.empty((p.L, n3, 6), dtype=torch.float64, device=dev)})
    torch.cuda.synchronize()
    return outs


def scene_graph(ctx, jobs, outs, torch):
    """All objects through cqbem_assemble_scene (shared tile launches, one HIP graph per wave) into `outs`; returns
    the wall time of the call and its phase timings."""
    from surfaceonlydeformables_amd import abi
    meshes = [abi.make_mesh(V, F) for V, F, _ in jobs]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.assemble_scene(meshes, [p for *_, p in jobs], [{n: t.data_ptr() for n, t in o.items()} for o in outs])
    wall = time.perf_counter() - t0
    return wall, ctx.last_timings()


def scene_main(args):
    """Multi-object scene (16 objects, seed 0): cqbem_assemble_scene (all objects' tiles in shared launches,
    graph-captured waves) as the line's value, beside the same jobs submitted at once to the serving slots and one
    after another.  Every path writes the same preallocated outputs and is timed after a warm-up pass (the first
    pass fills the library's device memory cache)."""
    import torch
    from surfaceonlydeformables_amd import abi
    jobs = scene_jobs()
    for _, _, p in jobs:  # precision and regular-face rule of every object, as the command line asks
        p.precision = abi.MIXED if args.precision == "mixed" else abi.FP64
        p.quad_policy = abi.QUAD_BANDED if args.quad == "banded" else abi.QUAD_REFERENCE
    entries = sum(2.0 * (3 * len(V)) ** 2 * p.L for V, F, p in jobs)
    outs = scene_outputs(jobs, torch)  # 257 GB: the same buffers for every path
    ctx = abi.Context(0)
    first_wall, _ = scene_graph(ctx, jobs, outs, torch)  # warm-up: the first call allocates the spectra region
    graph_wall, graph_t = scene_graph(ctx, jobs, outs, torch)
    os.environ["CQBEM_SCENE_GRAPH"] = "0"  # the same launches, on streams with concurrent branches, not captured
    try:
        direct_wall, direct_t = scene_graph(ctx, jobs, outs, torch)
    finally:
        del os.environ["CQBEM_SCENE_GRAPH"]
    ctx.close()  # its device memory cache holds the scene's spectra region
    if args.scene_only:
        print(json.dumps({"scene_first_ms": first_wall * 1e3, "scene_graph_ms": graph_wall * 1e3, "graph": graph_t,
                          "scene_uncaptured_ms": direct_wall * 1e3, "uncaptured": direct_t}), flush=True)
        return
    ctx = abi.Context(0)
    run_jobs(ctx, jobs, [0.0] * len(jobs), torch, outs)  # warm-up
    lat, makespan = run_jobs(ctx, jobs, [0.0] * len(jobs), torch, outs)
    # the same jobs one after another, each alone on the device (a warm-up pass, then the timed pass)
    meshes = [abi.make_mesh(V, F) for V, F, _ in jobs]
    serial = 0.0
    for rep in range(2):
        serial = 0.0
        for k, (V, F, p) in enumerate(jobs):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.assemble_device(meshes[k], p, {n: t.data_ptr() for n, t in outs[k].items()})
            serial += time.perf_counter() - t0
    ctx.close()
    del outs
    torch.cuda.empty_cache()
    print(json.dumps({"metric": "multi-object scene: G+H entries/s (cqbem_assemble_scene)",
                      "value": entries / graph_wall, "unit": "entries/s", "objects": len(jobs),
                      "scene_graph_ms": graph_wall * 1e3, "scene_waves": graph_t["gh_launches"],
                      "scene_waves_ms": graph_t["regular_gh"], "scene_b_graph_ms": graph_t["regular_b"],
                      "scene_host_ms": graph_t["host"], "scene_first_call_ms": first_wall * 1e3,
                      "scene_uncaptured_ms": direct_wall * 1e3, "scene_uncaptured_waves_ms": direct_t["regular_gh"],
                      "scene_uncaptured_b_ms": direct_t["regular_b"], "slots_stats": run_jobs.stats,
                      "slots_makespan_ms": makespan * 1e3, "serial_sum_ms": serial * 1e3,
                      "slots_latency_ms": {"p50": float(np.percentile(lat, 50)) * 1e3,
                                           "p99": float(np.percentile(lat, 99)) * 1e3},
                      "scene": [{"nv": len(V), "L": p.L} for V, F, p in jobs], "precision": args.precision,
                      "quadrature": args.quad, "dtype": "f64" if args.precision == "fp64" else "f32+f64",
                      "data": "synthetic"}),
          flush=True)


SERVING_BUDGET = 200e9  # bytes of HBM the serving context may hold (scratch of running jobs + resident outputs)
SERVING_STEPS = 500      # time steps an object runs on its resident matrices before it is released


def serving_jobs(n, rng, precision):
    """n jobs of the serving mix: 90% the scene distribution (icosphere L1-4 or cube k 4-16, L in {16, 32, 64}),
    10% icosphere level 5 (10242 vertices) at L = 16, materials and sizes as the scene's."""
    from surfaceonlydeformables_amd import abi
    from surfaceonlydeformables_amd import mesh as M
    prec = abi.MIXED if precision == "mixed" else abi.FP64
    ico5 = M.icosphere(5, 1.0)
    jobs = []
    for k in range(n):
        size = float(np.exp(rng.uniform(np.log(0.05), np.log(0.3))))
        if rng.uniform() < 0.1:
            V, F = ico5[0] * size, ico5[1]
            L = 16
        else:
            if rng.uniform() < 0.5:
                V, F = M.icosphere(int(rng.integers(1, 5)), size)
            else:
                V, F = M.cube(int(rng.integers(4, 17)), size)
            L = int(rng.choice([16, 32, 64]))
        V = M.center(M.jitter(V, F, 0.01, 1000 + k), F)
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
        nu = float(rng.uniform(0.2, 0.45))
        _, c1, c2 = M.material_speeds(E, nu, 1000.0)
        params = abi.make_params(c1, c2, 1000.0, nu, 1e-3, 1e-10 ** (1.0 / (2 * L)), L, bdf=abi.BDF2, order=5,
                                 subdivision=2, precision=prec)
        jobs.append((V, F, params))
    return jobs


def serve_stream(jobs, arrivals, torch, drain_s):
    """The request loop of the serving mix on one 200 GB context.  Arriving jobs queue; any queued job is submitted
    (cqbem_submit_resident) as soon as a stream slot and the budget admit it, by its footprint -- scratch plus its
    resident outputs -- next to the jobs running and the objects resident.  Policy for jobs whose double G + H
    outputs alone exceed the budget (icosphere L5 at L = 16: 242 GB): their G and H are stored in float (121 GB).
    A finished job's object runs SERVING_STEPS history steps on its resident matrices (a stepping thread with its own
    context and stream), then its matrices are released.  Returns per-job records."""
    import threading
    from surfaceonlydeformables_amd import abi
    dev = torch.device("cuda:0")
    ctx = abi.Context(0, hbm_budget_bytes=int(SERVING_BUDGET))
    stepper = abi.Context(0)
    meshes = [abi.make_mesh(V, F) for V, F, _ in jobs]
    rec = [{"nv": len(V), "L": p.L, "arrival": arrivals[k], "done": None, "f32": False} for k, (V, F, p) in
           enumerate(jobs)]
    resident, lock, stop = [], threading.Lock(), threading.Event()
    released = []

    def step_loop():
        hist = {}
        while not stop.is_set() or resident:
            with lock:
                objs = list(resident)
            if not objs:
                time.sleep(0.001)
                continue
            for obj in objs:
                k, jid, outs, steps = obj["k"], obj["jid"], obj["outs"], obj["steps"]
                n3, L = 3 * rec[k]["nv"], rec[k]["L"]
                if k not in hist:
                    hist[k] = tuple(torch.randn(s, dtype=torch.float64, device=dev) for s in
                                    ((L - 1, n3), (L - 1, n3), (L - 1, 6), (n3,)))
                    torch.cuda.synchronize()
                hp, hu, ha, hb = hist[k]
                fn = stepper.history_f32 if rec[k]["f32"] else stepper.history
                fn(n3, L, outs["G"].data_ptr(), outs["H"].data_ptr(), outs["B"].data_ptr(), hp.data_ptr(),
                   hu.data_ptr(), ha.data_ptr(), L - 1, hb.data_ptr())
                obj["steps"] = steps + 1
                if obj["steps"] >= SERVING_STEPS:
                    ctx.release(jid)
                    with lock:
                        resident.remove(obj)
                    del hist[k]
                    released.append(k)

    th = threading.Thread(target=step_loop, daemon=True)
    th.start()
    queue, running, nxt = [], {}, 0
    t0 = time.perf_counter()
    try:
        while True:
            now = time.perf_counter() - t0
            while nxt < len(jobs) and arrivals[nxt] <= now:
                queue.append(nxt)
                nxt += 1
            for k in list(queue):
                V, F, p = jobs[k]
                n3 = 3 * len(V)
                if 2 * p.L * n3 * n3 * 8 > SERVING_BUDGET:  # the policy for over-budget double outputs
                    p.storage = abi.STORE_F32
                    rec[k]["f32"] = True
                mt = torch.float32 if rec[k]["f32"] else torch.float64
                out_bytes = 2 * p.L * n3 * n3 * (4 if rec[k]["f32"] else 8) + p.L * n3 * 6 * 8
                if ctx.reserved_bytes() + out_bytes > SERVING_BUDGET:
                    continue  # cannot be admitted before resident objects are released
                try:
                    outs = {"G": torch.empty((p.L, n3, n3), dtype=mt, device=dev),
                            "H": torch.empty((p.L, n3, n3), dtype=mt, device=dev),
                            "B": torch.empty((p.L, n3, 6), dtype=torch.float64, device=dev)}
                except torch.cuda.OutOfMemoryError:
                    break
                torch.cuda.synchronize()
                jid = ctx.submit(meshes[k], p, {n: t.data_ptr() for n, t in outs.items()}, resident=True)
                if jid < 0:
                    del outs
                    continue  # not admitted now; a smaller queued job may be
                running[jid] = (k, outs)
                queue.remove(k)
            for jid in [j for j in running if ctx.poll(j)]:
                ctx.wait(jid)
                k, outs = running.pop(jid)
                rec[k]["done"] = time.perf_counter() - t0
                with lock:
                    resident.append({"k": k, "jid": jid, "outs": outs, "steps": 0})
            now = time.perf_counter() - t0
            if nxt >= len(jobs) and not queue and not running:
                break
            if now > arrivals[-1] + drain_s:
                break
            time.sleep(0.0005)
        end = time.perf_counter() - t0
    finally:
        stop.set()
        th.join()
        torch.cuda.synchronize()
        stepper.close()
        ctx.close()
    return rec, end


def serving_main(args):
    """The serving mix at `--serving` jobs/s (one rate) for `--duration` seconds of arrivals: p50 / p99 latency
    from arrival to resident matrices (jobs still unfinished when the drain ends count with their waiting time so
    far, a lower bound), aggregate G + H entries/s of the finished jobs over the run, admission counts."""
    import torch
    rng = np.random.default_rng(int(args.serving * 1000) + 1)
    n = max(1, int(round(args.serving * args.duration)))
    arrivals = np.cumsum(rng.exponential(1.0 / args.serving, n)).tolist()
    jobs = serving_jobs(n, rng, args.precision)
    from surfaceonlydeformables_amd import abi
    warm = scene_jobs(2, seed=5)  # small objects: loads the kernels and fills the context's pool
    for _, _, p in warm + jobs:
        p.precision = abi.MIXED if args.precision == "mixed" else abi.FP64
        p.quad_policy = abi.QUAD_BANDED if args.quad == "banded" else abi.QUAD_REFERENCE
    serve_stream(warm, [0.0, 0.0], torch, 60.0)
    rec, end = serve_stream(jobs, arrivals, torch, drain_s=30.0)
    lat = [(r["done"] if r["done"] is not None else end) - r["arrival"] for r in rec]
    done = [r for r in rec if r["done"] is not None]
    entries = sum(2.0 * (3 * r["nv"]) ** 2 * r["L"] for r in done)
    print(json.dumps({"metric": "serving mix: G+H entries/s and p50/p99 job latency (arrival to resident matrices)",
                      "rate_jobs_per_s": args.serving, "jobs": n, "finished": len(done),
                      "icosphere_l5_jobs": sum(r["nv"] == 10242 for r in rec),
                      "stored_f32": sum(r["f32"] for r in rec), "value": entries / end, "unit": "entries/s",
                      "latency_ms": {"p50": float(np.percentile(lat, 50)) * 1e3, "p99": float(np.percentile(lat, 99)) * 1e3},
                      "run_s": end, "budget_GB": SERVING_BUDGET / 1e9, "steps_per_object": SERVING_STEPS,
                      "precision": args.precision, "quadrature": args.quad,
                      "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic"}), flush=True)


def time_stepping_main(args):
    """CQ time steps of one body on the icosphere-L4 matrices (L = 32) through the C++ facade: the collection
    driver assembles, post-processes, uploads the matrices once, then steps an ElastodynamicsCQBEMObject
    (history sums, u_p p and the H_0^-1 product on the GPU, drift elimination on the host).
    The first three steps carry a pressure pulse (constrained steps), the rest none.  With --compression-ratio the
    matrices are wavelet-compressed first (upload_ms then includes the compression of 64 matrices)."""
    import subprocess
    import tempfile
    from surfaceonlydeformables_amd import mesh as M
    V, F, params, desc = workload("ico4-collocation", "fp64")
    nv, n3, L = len(V), 3 * len(V), params.L
    steps = L + max(8, 2 * args.steps)  # step_ms is the mean once the history holds L - 1 steps
    with tempfile.TemporaryDirectory() as tmp:
        obj = os.path.join(tmp, "mesh.obj")
        M.write_obj(obj, V, F)
        cfg = {"enable_cuda": True, "mesh": {"file": obj},
               "material": {"rho": params.rho, "mu": params.rho * params.c2 ** 2, "nu": params.nu},
               "simulation": {"dt": params.dt, "gaussian_quadrature_order": 5, "quadrature_subdivision": 2,
                              "multistep_method": "BDF2", "use_galerkin": False, "cq_length": L,
                              "compression_ratio": args.compression_ratio},
               "object": {"translational_vel": [0.1, 0.0, 0.0], "rotational_vel": [0.0, 1.0, 0.0]}}
        scene = os.path.join(tmp, "scene.json")
        with open(scene, "w") as fh:
            json.dump(cfg, fh)
        Vc = V - V.mean(axis=0)
        p = np.zeros((steps, nv, 3))
        mask = Vc[:, 0] > 0.4 * Vc[:, 0].max()
        p[:3, mask] = -300.0 * Vc[mask] / np.linalg.norm(Vc[mask], axis=1)[:, None]
        tr = os.path.join(tmp, "tractions.bin")
        p.reshape(steps, n3).tofile(tr)
        drv = os.path.join(ROOT, "surfaceonlydeformables_amd", "collection_driver")
        print(f"time stepping: {desc}, {steps} steps", file=sys.stderr, flush=True)
        r = subprocess.run([drv, "--simulate", scene, str(steps), tr, "-"], capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr)
            sys.exit(1)
        info = json.loads(r.stdout.strip().splitlines()[-1])
    # the reference's history loop on the host (Eigen GEMVs over 2 (L - 1) dense matrices): timed with numpy on a
    # 4-matrix sample held in host memory, scaled to 2 (L - 1) matrices
    rng = np.random.default_rng(0)
    sample = [rng.standard_normal((n3, n3)) for _ in range(4)]
    x = rng.standard_normal(n3)
    t0 = time.perf_counter()
    for A in sample:
        A @ x
    per = (time.perf_counter() - t0) / len(sample)
    print(json.dumps({"metric": "CQ time step of one body on resident matrices (history + H_0 solve + drift "
                                "elimination)", "value": info["step_ms"], "unit": "ms", "higher_is_better": False,
                      "compression_ratio": args.compression_ratio,
                      "steps": steps, "step_phases_ms": info["step_phases_ms"], "first_step_ms": info["first_step_ms"],
                      "upload_ms": info["upload_ms"],
                      "assembly_ms": info["ms"], "post_process_ms": info["factor_ms"],
                      "post_process_phases_ms": info.get("factor_phases_ms"),
                      "history_matrices_GB": (2 * (L - 1) * n3 * n3 * 8 / 1e9 if not args.compression_ratio else
                                              2 * (L - 1) * int(n3 * n3 * args.compression_ratio) * 12 / 1e9),
                      "cpu_history_ms": {"value": per * 2 * (L - 1) * 1e3, "threads": os.cpu_count(),
                                         "sample": f"numpy {n3} x {n3} GEMV, 4 matrices timed, x {2 * (L - 1)}"},
                      "config": {"workload": desc, "nv": nv, "L": L}, "dtype": "f64", "data": "synthetic"}),
          flush=True)


REFERENCE_SCENES = ("bunny_friction", "cube_drop", "cube_collide", "glued_cubes")


def reference_scene_main(args):
    """The elastodynamicsCQBEM groups of one of the reference's shipped scenes (data/<name>.json, meshes from
    tests/golden) through the C++ facade: init (mesh preparation, assembly, post-processing), the upload of the
    resident matrices (and their wavelet compression when the scene sets compression_ratio), then time steps of one
    body with a 3-step pressure pulse.  The scene's matrix cache folder is dropped, so every run assembles."""
    import subprocess
    import tempfile
    from surfaceonlydeformables_amd import mesh as M
    golden = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(golden, "scenes", args.reference_scene + ".json")) as fh:
        scene = json.load(fh)
    drv = os.path.join(ROOT, "surfaceonlydeformables_amd", "collection_driver")
    results = []
    for grp in scene["physics_object_groups"]:
        cfg = grp["physics_object_group"]
        if cfg.get("type") != "elastodynamicsCQBEM":
            continue
        cfg = json.loads(json.dumps(cfg))
        cfg["mesh"]["file"] = os.path.join(golden, os.path.basename(cfg["mesh"]["file"]))
        cfg["simulation"].pop("coeffs_folder_path", None)
        cfg["enable_cuda"] = True
        cfg["gravity"] = scene.get("gravitational_constant", [0.0, -9.8, 0.0])
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "group.json")
            with open(path, "w") as fh:
                json.dump(cfg, fh)
            # vertex count after the scene's mesh preparation (split included)
            r = subprocess.run([drv, "--mesh", path], capture_output=True, text=True)
            if r.returncode != 0:
                sys.stderr.write(r.stderr)
                sys.exit(1)
            Vp = np.array(json.loads(r.stdout)["V"]).reshape(-1, 3)
            nv = len(Vp)
            steps = max(16, 4 * args.steps)
            Vc = Vp - Vp.mean(axis=0)
            p = np.zeros((steps, nv, 3))
            mask = Vc[:, 1] < Vc[:, 1].min() + 0.2 * (Vc[:, 1].max() - Vc[:, 1].min())
            p[:3, mask, 1] = 1e3
            tr = os.path.join(tmp, "tractions.bin")
            p.reshape(steps, 3 * nv).tofile(tr)
            print(f"reference scene {args.reference_scene}: group with {nv} vertices, {steps} steps", file=sys.stderr,
                  flush=True)
            r = subprocess.run([drv, "--simulate", path, str(steps), tr, "-"], capture_output=True, text=True)
            if r.returncode != 0:
                sys.stderr.write(r.stderr)
                sys.exit(1)
            info = json.loads(r.stdout.strip().splitlines()[-1])
        sim = cfg["simulation"]
        results.append({"mesh": os.path.basename(cfg["mesh"]["file"]), "nv": nv, "L": info["L"],
                        "formulation": "Galerkin" if sim.get("use_galerkin") else "collocation",
                        "multistep_method": sim.get("multistep_method"),
                        "traction_discontinuity": bool(cfg["mesh"].get("enable_traction_discontinuity", False)),
                        "compression_ratio": sim.get("compression_ratio", 0.0),
                        "init_assembly_ms": info["ms"], "init_post_process_ms": info["factor_ms"],
                        "resident_upload_ms": info["upload_ms"], "step_ms": info["step_ms"],
                        "step_phases_ms": info["step_phases_ms"],
                        "G+H_entries_per_s": 2.0 * (3 * nv) ** 2 * info["L"] / (info["ms"] / 1e3)})
    print(json.dumps({"metric": f"reference scene {args.reference_scene}: facade init and time step of its "
                                "elastodynamicsCQBEM bodies", "groups": results, "dtype": "f64",
                      "data": "the reference's scene file and mesh"}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--quad", choices=["reference", "banded"], default="reference",
                    help="regular-face rule: the reference's uniform rule, or the distance-banded approximation")
    ap.add_argument("--workload", choices=WORKLOADS, default=WORKLOADS[0],
                    help="single-object workload (default: the icosphere-L4 collocation line)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="one untimed assembly followed by a 1 GiB device copy, for HBM counter runs "
                         "(the copy calibrates FETCH_SIZE / WRITE_SIZE); prints no JSON line")
    ap.add_argument("--scene", action="store_true", help="multi-object scene throughput instead of the single-object line")
    ap.add_argument("--scene-only", action="store_true", help="with --scene: only the cqbem_assemble_scene runs")
    ap.add_argument("--serving", type=float, default=0.0, help="Poisson serving stream at this many jobs/s")
    ap.add_argument("--duration", type=float, default=10.0, help="serving stream length in seconds")
    ap.add_argument("--time-stepping", action="store_true",
                    help="time steps of one icosphere-L4 body on its resident matrices instead of the assembly line")
    ap.add_argument("--reference-scene", choices=REFERENCE_SCENES,
                    help="the deformable bodies of one of the reference's shipped scenes through the C++ facade")
    ap.add_argument("--compression-ratio", type=float, default=0.0,
                    help="with --time-stepping: wavelet-compressed matrices keeping this fraction of coefficients")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the CQBEM precompute of one object runs on one MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.scene:
        return scene_main(args)
    if args.serving > 0:
        return serving_main(args)
    if args.time_stepping:
        return time_stepping_main(args)
    if args.reference_scene:
        return reference_scene_main(args)

    import torch
    from surfaceonlydeformables_amd import abi

    V, F, params, desc = workload(args.workload, args.precision)
    if args.quad == "banded":
        if params.galerkin:
            raise SystemExit("--quad banded applies to collocation workloads")
        params.quad_policy = abi.QUAD_BANDED
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
        gh_ms += tm["regular_b" if params.galerkin else "regular_gh"]  # Galerkin: the pair kernel at the test points
        gh_evals += tm["gh_evaluations"]
        gh_launches += int(tm["gh_launches"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1e3 / args.steps
    entries = 2.0 * n3 * n3 * L
    value = entries * args.steps / wall
    last = ctx.last_timings()

    roof = roofline(gh_ms, gh_evals, gh_launches, with_traffic=args.workload == WORKLOADS[0], galerkin=params.galerkin)

    # time stepping on the resident matrices: the history product of one step with a full history
    history = history_step(ctx, ptrs, n3, L, torch, dev)

    phase_keys = ("regular_gh", "regular_b", "singular", "transform", "host")
    if params.galerkin:  # non-singular G/H/B with the test projection in one phase, of which the G/H pair kernel
        phases = {"nonsingular": last["regular_gh"], "nonsingular_gh_pair_kernel": last["regular_b"],
                  "singular_pairs": last["singular"], "transform": last["transform"], "host": last["host"]}
    else:
        phases = {k: last[k] for k in phase_keys}
    line = {"metric": METRIC, "value": value, "unit": "entries/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64",
            "data": "synthetic",
            "config": {"workload": desc, "nv": nv, "nf": len(F), "L": L, "precision": args.precision,
                       "quadrature": "reference rule" if args.quad == "reference" else
                       "distance-banded (far tiles: order 5 x subdivision 1, default kappa)",
                       "entries_per_step": entries},
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
            "phases_ms_last_step": phases, "roofline": roof, "time_step_history": history}
    if not params.galerkin and last["transform"] > 0:
        # the G/H weights' write-out: the inverse transform reads each entry's Lf complex spectrum values once and
        # writes its L real weights once (the B transform and post-processing in the same phase are ~0.1% of it)
        lf = L // 2 + 1
        wbytes = 2.0 * n3 * n3 * (16.0 * lf + 8.0 * L)
        line["write_out"] = {"kernel": "k_transform_pair<2>", "bound": "hbm", "ms": last["transform"], "bytes": wbytes,
                             "achieved_TBps": wbytes / (last["transform"] * 1e-3) / 1e12, "peak_TBps": 8.0,
                             "frac": wbytes / (last["transform"] * 1e-3) / 8e12}
    if args.quad == "banded":
        cmp = mixed_against_fp64(ctx, mesh, params, outs, ms_per_step, torch, dev)
        cmp["speedup_vs_reference_rule"] = cmp.pop("speedup_vs_fp64")
        cmp["reference_rule_ms_per_step"] = cmp.pop("fp64_ms_per_step")
        cmp["note"] = ("banded rule against the reference rule, both FP64; its accuracy against a refined rule is in "
                       "surfaceonlydeformables_amd/band_tables/")
        cmp.pop("target")
        line["banded_vs_reference_rule"] = cmp
        if roof:
            roof["note"] = "banded: evaluations counted per tile's rule"
            roof["kernel"] = "k_colloc_gh_pair<true,true,false,true>"
    if args.precision == "mixed":
        line["mixed_vs_fp64"] = mixed_against_fp64(ctx, mesh, params, outs, ms_per_step, torch, dev)
        if roof and os.environ.get("CQBEM_MIXED_G") == "float":
            roof["note"] = "mixed, float G: the timed G/H phase is the double H pair kernel plus the float G kernel"
            roof["kernel"] = "k_colloc_gh_pair<false,true,false,false> + k_colloc_g32"
        elif roof:
            roof["note"] = "mixed: G and H in the double pair kernel, the far-field B in float"
    if args.precision == "fp64" and args.quad == "reference" and not params.galerkin:
        # the metric's comparison, in the same run after the timed FP64 steps: the mixed-precision job, alone and
        # with the distance-banded rule, against the FP64 outputs just assembled
        line["mixed_vs_fp64"] = variant_against_fp64(ctx, mesh, params, outs, ms_per_step, torch, abi.MIXED, False)
        line["mixed_banded_vs_fp64"] = variant_against_fp64(ctx, mesh, params, outs, ms_per_step, torch, abi.MIXED,
                                                            True)
    if args.workload == WORKLOADS[0] and args.precision == "fp64" and args.quad == "reference":
        del outs, ptrs
        torch.cuda.empty_cache()
        line["galerkin_icosphere_l3"] = galerkin_line(ctx, torch, dev)
    if not args.no_cpu_baseline and not params.galerkin:
        line["cpu_baseline"] = cpu_baseline(V, F, params)
    ctx.close()
    print(json.dumps(line), flush=True)


def galerkin_line(ctx, torch, dev, steps=1):
    """The icosphere-L3 Galerkin workload (L = 64, 1926 x 1926 complex matrices per frequency) in the same run: one
    warm-up and `steps` timed assemblies into device outputs, its phases and the pair kernel's roofline at the test
    points."""
    from surfaceonlydeformables_amd import abi
    V, F, params, desc = workload("ico3-galerkin", "fp64")
    n3, L = 3 * len(V), params.L
    outs = {"G": torch.empty((L, n3, n3), dtype=torch.float64, device=dev),
            "H": torch.empty((L, n3, n3), dtype=torch.float64, device=dev),
            "B": torch.empty((L, n3, 6), dtype=torch.float64, device=dev)}
    ptrs = {k: t.data_ptr() for k, t in outs.items()}
    mesh = abi.make_mesh(V, F)
    ctx.assemble_device(mesh, params, ptrs)
    torch.cuda.synchronize()
    gh_ms = gh_evals = gh_launches = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.assemble_device(mesh, params, ptrs)
        tm = ctx.last_timings()
        gh_ms += tm["regular_b"]
        gh_evals += tm["gh_evaluations"]
        gh_launches += tm["gh_launches"]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    del outs
    torch.cuda.empty_cache()
    return {"workload": desc, "steps": steps, "warmup": 1, "ms_per_step": ms, "entries_per_s": 2.0 * n3 * n3 * L * 1e3 / ms,
            "phases_ms_last_step": {"nonsingular": tm["regular_gh"], "nonsingular_gh_pair_kernel": tm["regular_b"],
                                    "singular_pairs": tm["singular"], "transform": tm["transform"], "host": tm["host"]},
            "roofline": roofline(gh_ms, gh_evals, int(gh_launches), with_traffic=False, galerkin=True)}


def variant_against_fp64(ctx, mesh, params, outs, fp64_ms, torch, precision, banded, steps=3):
    """The FP64 job again as a variant (precision, distance-banded rule or not): one warm-up and `steps` timed
    assemblies into outputs of its own, then the relative Frobenius error of every output against the FP64 outputs
    `outs`, over the whole stack and per weight n (norms on the device)."""
    from surfaceonlydeformables_amd import abi
    pv = abi.make_params(params.c1, params.c2, params.rho, params.nu, params.dt, params.R, params.L, bdf=params.bdf,
                         order=params.order, subdivision=params.subdivision, galerkin=params.galerkin,
                         which=params.which, precision=precision,
                         quad=abi.QUAD_BANDED if banded else abi.QUAD_REFERENCE)
    var = {k: torch.empty_like(t) for k, t in outs.items()}
    ptrs = {k: t.data_ptr() for k, t in var.items()}
    torch.cuda.synchronize()
    ctx.assemble_device(mesh, pv, ptrs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.assemble_device(mesh, pv, ptrs)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    stack, worst = {}, {}
    for k in outs:
        a, b = var[k].reshape(params.L, -1), outs[k].reshape(params.L, -1)
        pw = [float(torch.linalg.norm(a[n] - b[n]) / torch.linalg.norm(b[n])) for n in range(params.L)]
        stack[k] = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        worst[k] = max(pw)
    del var
    torch.cuda.empty_cache()
    return {"ms_per_step": ms, "speedup_vs_fp64": fp64_ms / ms,
            "gh_entries_per_s": 2.0 * (3 * mesh.nv) ** 2 * params.L * 1e3 / ms,
            "rel_err_stack": stack, "rel_err_max_per_weight": worst,
            "gh_within_1e-6": stack["G"] <= 1e-6 and stack["H"] <= 1e-6,
            "note": ("far-field B_trans and B_angular in float (Kahan-compensated face sums); G and H in the double "
                     "pair kernel" + ("; G/H tiles beyond the band threshold on the far rule" if banded else "")
                     + "; H, singular faces and the inverse transform in double")}


def mixed_against_fp64(ctx, mesh, params, outs, mixed_ms, torch, dev, steps=3):
    """The same job in FP64 with the reference rule (a few steps, timed for the ratio) and the relative Frobenius
    error of every output of the mixed (or banded) job against it, per matrix over the whole stack and per weight n
    (norms on the device)."""
    from surfaceonlydeformables_amd import abi
    p64 = abi.make_params(params.c1, params.c2, params.rho, params.nu, params.dt, params.R, params.L, bdf=params.bdf,
                          order=params.order, subdivision=params.subdivision, galerkin=params.galerkin,
                          which=params.which)
    ref = {k: torch.empty_like(t) for k, t in outs.items()}
    ptrs = {k: t.data_ptr() for k, t in ref.items()}
    torch.cuda.synchronize()
    ctx.assemble_device(mesh, p64, ptrs)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.assemble_device(mesh, p64, ptrs)
    fp64_ms = (time.perf_counter() - t0) * 1e3 / steps
    torch.cuda.synchronize()
    per_weight, stack = {}, {}
    for k in outs:
        a, b = outs[k].reshape(params.L, -1), ref[k].reshape(params.L, -1)
        per_weight[k] = [float(torch.linalg.norm(a[n] - b[n]) / torch.linalg.norm(b[n])) for n in range(params.L)]
        stack[k] = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
    del ref
    torch.cuda.empty_cache()
    return {"fp64_ms_per_step": fp64_ms, "speedup_vs_fp64": fp64_ms / mixed_ms, "rel_err_stack": stack,
            "rel_err_per_weight": per_weight, "target": "<= 1e-6 relative error, >= 3x FP64",
            "note": "far-field B_trans and B_angular in float (Kahan-compensated face sums), far-field G too with "
                    "CQBEM_MIXED_G=float; H, singular faces and the inverse transform in double"}


def roofline(gh_ms, gh_evals, gh_launches, with_traffic, galerkin=False):
    """VALU roofline of the dominant kernel from the library's HIP-event timers; the HBM traffic per launch is
    the committed counter measurement of the icosphere-L4 shape.  Galerkin: the same pair kernel at the test
    points, per (test point, regular source face, source point, frequency)."""
    achieved = FLOPS_PER_EVAL_GH * gh_evals / (gh_ms * 1e-3) / 1e12
    kernel = "k_colloc_gh_pair<true,true,true,false>" if galerkin else "k_colloc_gh_pair<true,true,false,false>"
    roof = {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
            "launch_ms": gh_ms / max(gh_launches, 1),
            "flops_per_launch": FLOPS_PER_EVAL_GH * gh_evals / max(gh_launches, 1)}
    prof = os.path.join(ROOT, "profiles", "k_colloc_gh_pair_icosphere_l4_traffic.json")
    if with_traffic and os.path.exists(prof):
        with open(prof) as fh:
            roof["traffic"] = json.load(fh).get("hbm_bytes_per_launch")
    return roof


def history_step(ctx, ptrs, n3, L, torch, dev):
    nh = L - 1
    hp = torch.randn((nh, n3), dtype=torch.float64, device=dev)
    hu = torch.randn((nh, n3), dtype=torch.float64, device=dev)
    ha = torch.randn((nh, 6), dtype=torch.float64, device=dev)
    hb = torch.empty(n3, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()  # the histories come from torch's stream; the library runs on its own
    hist_args = (n3, L, ptrs["G"], ptrs["H"], ptrs["B"], hp.data_ptr(), hu.data_ptr(), ha.data_ptr(), nh, hb.data_ptr())
    ctx.history(*hist_args)
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.history(*hist_args)
    hist_ms = (time.perf_counter() - t0) * 1e3 / reps
    hist_bytes = 2.0 * nh * n3 * n3 * 8 + nh * n3 * 6 * 8
    return {"kernel": "k_history_v2 (4 rows per wavefront, 16-byte loads)", "ms": hist_ms, "matrices_read_GB": hist_bytes / 1e9,
            "achieved_TBps": hist_bytes / (hist_ms * 1e-3) / 1e12, "peak_TBps": 8.0,
            "note": f"b = sum_i G_i p + H_i u + B_i acc over the {nh} resident history weights (one CQ time step)"}

if __name__ == "__main__":
    main()
