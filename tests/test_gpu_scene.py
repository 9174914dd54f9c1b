"""cqbem_assemble_scene: all objects of a scene in shared, graph-captured launches (the G/H regular-face tiles of
every collocation object in one grid per wave, then the singular faces and transforms; B in a last graph).

Every object's outputs must be bit-identical to its own cqbem_assemble: the tiles run the single-object kernel's
computation unchanged, and row slabs and waves do not change any sum.  Checked for a mixed scene (sizes, L, BDF1 and
BDF2, a vertex-split cube, a CQBEM_MIXED object, a float-stored object, a Galerkin object that takes the per-object
path), with both tile dispatch forms (one block per tile, and persistent blocks on a device tile counter), under a small
budget that splits the objects into row slabs over several waves, and with the launches run on their streams
instead of captured (CQBEM_SCENE_GRAPH=0)."""
import numpy as np
import pytest

from surfaceonlydeformables_amd import abi
from surfaceonlydeformables_amd import mesh as M

pytestmark = pytest.mark.gpu

RHO = 1000.0


def scene():
    rng = np.random.default_rng(4)
    objs = []
    specs = [("ico", 2, 32, abi.BDF2, "fp64"), ("cube", 4, 16, abi.BDF1, "fp64"), ("ico", 1, 64, abi.BDF2, "fp64"),
             ("cube", 3, 32, abi.BDF2, "mixed"), ("ico", 2, 16, abi.BDF2, "f32"), ("split", 2, 16, abi.BDF1, "fp64"),
             ("ico", 1, 8, abi.BDF2, "galerkin")]
    for k, (kind, size, L, bdf, mode) in enumerate(specs):
        vmap = None
        if kind == "ico":
            V, F = M.icosphere(size, float(np.exp(rng.uniform(np.log(0.05), np.log(0.3)))))
            V = M.jitter(V, F, 0.01, k)
        elif kind == "cube":
            V, F = M.cube(size, 0.2)
        else:
            V0, F0 = M.cube(size, 0.2)
            V, F, vmap = abi.split_traction_discontinuity(M.center(V0, F0), F0, 20.0)
        V = M.center(V, F) if vmap is None else V
        _, c1, c2 = M.material_speeds(float(np.exp(rng.uniform(np.log(1e4), np.log(1e6)))), 0.3, RHO)
        p = abi.make_params(c1, c2, RHO, 0.3, 1e-3, 1e-10 ** (1 / (2 * L)), L, bdf=bdf,
                            order=2 if mode == "galerkin" else 5, subdivision=1 if mode == "galerkin" else 2,
                            galerkin=mode == "galerkin", precision=abi.MIXED if mode == "mixed" else abi.FP64,
                            storage=abi.STORE_F32 if mode == "f32" else abi.STORE_F64)
        objs.append((V, F, vmap, p))
    return objs


def outputs(torch, nv, p):
    n3, L = 3 * nv, p.L
    mt = torch.float32 if p.storage == abi.STORE_F32 else torch.float64
    o = {"G": torch.full((L, n3, n3), float("nan"), dtype=mt, device="cuda:0"),
         "H": torch.full((L, n3, n3), float("nan"), dtype=mt, device="cuda:0"),
         "B_trans": torch.full((L, n3, 3), float("nan"), dtype=torch.float64, device="cuda:0"),
         "B_angular": torch.full((L, n3, 3, 3), float("nan"), dtype=torch.float64, device="cuda:0"),
         "B": torch.full((L, n3, 6), float("nan"), dtype=torch.float64, device="cuda:0")}
    torch.cuda.synchronize()  # buffers filled on torch's stream; the library runs on its own
    return o


@pytest.mark.parametrize("variant", ["tiles", "persistent", "tiles_small_budget", "uncaptured"])
def test_scene_matches_solo_assembly(variant, monkeypatch):
    import torch
    monkeypatch.setenv("CQBEM_SCENE_KERNEL", "persistent" if variant == "persistent" else "tiles")
    monkeypatch.setenv("CQBEM_SCENE_GRAPH", "0" if variant == "uncaptured" else "1")
    objs = scene()
    meshes = [abi.make_mesh(V, F, vmap) for V, F, vmap, p in objs]
    solo_ctx = abi.Context(0)
    solo = []
    try:
        for m, (V, F, vmap, p) in zip(meshes, objs):
            o = outputs(torch, len(V), p)
            solo_ctx.assemble_device(m, p, {k: t.data_ptr() for k, t in o.items()})
            torch.cuda.synchronize()
            solo.append({k: t.cpu().numpy() for k, t in o.items()})
            del o
    finally:
        solo_ctx.close()
    budget = 0
    if variant == "tiles_small_budget":  # about 40 rows of the largest object's spectra: several slabs and waves
        nv = max(len(V) for V, F, vmap, p in objs)
        fixed = sum(abi.job_footprint(m, p) - 2 * (p.L // 2 + 1) * 9 * len(V) ** 2 * 16
                    for m, (V, F, vmap, p) in zip(meshes, objs) if not p.galerkin)
        budget = int(fixed + 40 * 2 * 33 * 9 * nv * 16)
    ctx = abi.Context(0, hbm_budget_bytes=budget)
    try:
        outs = [outputs(torch, len(V), p) for V, F, vmap, p in objs]
        print(f"scene of {len(objs)} objects, {variant}", flush=True)
        ctx.assemble_scene(meshes, [p for *_, p in objs], [{k: t.data_ptr() for k, t in o.items()} for o in outs])
        torch.cuda.synchronize()
        waves = ctx.last_timings()["gh_launches"]
        print(f"  {waves:.0f} waves", flush=True)
        if variant == "tiles_small_budget":
            assert waves > 1
        for k, o in enumerate(outs):
            for name, t in o.items():
                a = t.cpu().numpy()
                assert np.isfinite(a).all(), (k, name)
                assert np.array_equal(a, solo[k][name]), (k, name)
    finally:
        ctx.close()
