"""cqbem_assemble_noalloc: the assembly without device allocation, against cqbem_assemble.

Two calls in a row on one context (the second with a smaller mesh, so that the scratch array still holds the
first call's data) must each equal, bit for bit, the allocating entry point's outputs for the same job; the
device's free memory must not change across them (the library's own pool and torch's cache are left as they
are, up to 4 MiB of runtime internals); a job whose buffers exceed the scratch maximum is refused with
CQBEM_ERR_NOMEM before anything runs."""
import numpy as np
import pytest

from surfaceonlydeformables_amd import abi
from surfaceonlydeformables_amd import mesh as M

pytestmark = pytest.mark.gpu

RHO = 1000.0


def job(kind, L, galerkin=False):
    if kind == "ico2":
        V, F = M.icosphere(2, 0.2)
        V = M.center(M.jitter(V, F, 0.01, 3), F)
    elif kind == "cube2":
        V, F = M.cube(2, 0.15)
        V = M.center(V, F)
    else:
        V, F = M.icosphere(1, 0.1)
        V = M.center(M.jitter(V, F, 0.01, 5), F)
    _, c1, c2 = M.material_speeds(5e5, 0.3, RHO)
    order, sub = (2, 1) if galerkin else (5, 2)
    p = abi.make_params(c1, c2, RHO, 0.3, 1e-3, 1e-10 ** (1 / (2 * L)), L, order=order, subdivision=sub,
                        galerkin=galerkin)
    return V, F, p


def outputs(torch, nv, L):
    n3 = 3 * nv
    shapes = {"G": (L, n3, n3), "H": (L, n3, n3), "B_trans": (L, n3, 3), "B_angular": (L, n3, 3, 3), "B": (L, n3, 6)}
    outs = {k: torch.full(s, float("nan"), dtype=torch.float64, device="cuda:0") for k, s in shapes.items()}
    torch.cuda.synchronize()  # buffers filled on torch's stream; the library runs on its own
    return outs


def test_noalloc_matches_assemble_on_consecutive_calls():
    import torch
    ctx = abi.Context(0)
    try:
        cases = [job("ico2", 16), job("cube2", 8), job("ico1", 8, galerkin=True)]
        ref = []
        for V, F, p in cases:  # the allocating entry point first (it also loads every kernel)
            o = outputs(torch, len(V), p.L)
            ctx.assemble_device(abi.make_mesh(V, F), p, {k: t.data_ptr() for k, t in o.items()})
            torch.cuda.synchronize()
            ref.append(o)
        got = [outputs(torch, len(V), p.L) for V, F, p in cases]
        free0 = torch.cuda.mem_get_info()[0]
        for k, (V, F, p) in enumerate(cases):
            print(f"noalloc call {k}: nv={len(V)} L={p.L} galerkin={p.galerkin}", flush=True)
            ctx.assemble_device_noalloc(abi.make_mesh(V, F), p, {n: t.data_ptr() for n, t in got[k].items()})
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        # an allocation of the job's own would take tens of MiB (the spectra); leave 4 MiB for runtime internals
        assert free0 - free1 < 4 << 20, (free0, free1)
        for k in range(len(cases)):
            for n in ref[k]:
                a, b = got[k][n].cpu().numpy(), ref[k][n].cpu().numpy()
                assert np.isfinite(a).all(), (k, n)
                assert np.array_equal(a, b), (k, n)
    finally:
        ctx.close()


def test_noalloc_refuses_jobs_beyond_the_scratch():
    import torch
    V, F = M.icosphere(4, 0.2)  # Galerkin spectra of the whole matrix: far beyond the scratch maximum
    _, c1, c2 = M.material_speeds(5e5, 0.3, RHO)
    p = abi.make_params(c1, c2, RHO, 0.3, 1e-3, 1e-10 ** (1 / 32), 16, order=2, subdivision=1, galerkin=True,
                        which=abi.G)
    ctx = abi.Context(0)
    try:
        n3 = 3 * len(V)
        g = torch.empty((16, n3, n3), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        with pytest.raises(abi.CQBEMError) as err:
            ctx.assemble_device_noalloc(abi.make_mesh(V, F), p, {"G": g.data_ptr()})
        assert err.value.code == abi.ERR_NOMEM
        assert "MiB" in str(err.value)
    finally:
        ctx.close()
