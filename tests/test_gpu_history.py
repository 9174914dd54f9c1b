"""History term of a CQ time step on the resident matrices (cqbem_history) and cqbem_matvec, against numpy.

b = sum_{i=1..m} (G_i p[nh-i] - H_i u[nh-i] + B_i acc[nh-i]), m = min(nh, L-1): the sum
ElastodynamicsCQBEMObject::estimate_next_state forms each step (elastodynamics_cqbem_object.cpp:42-49),
with H not negated here.  Floating-point sums in a different order: relative error <= 1e-13."""
import numpy as np
import pytest

from surfaceonlydeformables_amd import abi
from surfaceonlydeformables_amd import mesh as M

pytestmark = pytest.mark.gpu


def reference(G, H, B, p, u, acc, L):
    nh = p.shape[0]
    m = min(nh, L - 1)
    b = np.zeros(G.shape[1])
    for i in range(1, m + 1):
        b += G[i] @ p[nh - i] - H[i] @ u[nh - i] + B[i] @ acc[nh - i]
    return b


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("n3,L,nh", [(9, 4, 2), (27, 8, 7), (129, 6, 11), (387, 16, 15), (390, 5, 0)])
def test_history_random(ctx, n3, L, nh):
    import torch
    rng = np.random.default_rng(n3)
    G, H = rng.standard_normal((L, n3, n3)), rng.standard_normal((L, n3, n3))
    B = rng.standard_normal((L, n3, 6))
    p, u, acc = rng.standard_normal((nh, n3)), rng.standard_normal((nh, n3)), rng.standard_normal((nh, 6))
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in
         dict(G=G, H=H, B=B, p=p if nh else np.zeros((1, n3)), u=u if nh else np.zeros((1, n3)),
              acc=acc if nh else np.zeros((1, 6))).items()}
    b = torch.full((n3,), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()  # buffers filled on torch's stream; the library runs on its own
    print(f"history n3={n3} L={L} nh={nh}", flush=True)
    ctx.history(n3, L, t["G"].data_ptr(), t["H"].data_ptr(), t["B"].data_ptr(), t["p"].data_ptr(),
                t["u"].data_ptr(), t["acc"].data_ptr(), nh, b.data_ptr())
    got = b.cpu().numpy()
    want = reference(G, H, B, p, u, acc, L) if nh else np.zeros(n3)
    if nh:
        assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
    else:
        assert np.array_equal(got, want)


def test_history_on_assembled_matrices(ctx):
    """The matrices cqbem_assemble leaves in HBM feed the history product directly; a NULL G skips its term."""
    import torch
    V, F = M.icosphere(1, 0.2)
    V = M.center(M.jitter(V, F, 0.01, 1), F)
    _, c1, c2 = M.material_speeds(5e5, 0.3, 1000.0)
    L = 16
    params = abi.make_params(c1, c2, 1000.0, 0.3, 1e-3, 1e-10 ** (1 / (2 * L)), L)
    n3 = 3 * len(V)
    dev = torch.device("cuda:0")
    outs = {"G": torch.empty((L, n3, n3), dtype=torch.float64, device=dev),
            "H": torch.empty((L, n3, n3), dtype=torch.float64, device=dev),
            "B": torch.empty((L, n3, 6), dtype=torch.float64, device=dev)}
    ctx.assemble_device(abi.make_mesh(V, F), params, {k: v.data_ptr() for k, v in outs.items()})
    rng = np.random.default_rng(7)
    nh = 20
    p, u, acc = rng.standard_normal((nh, n3)), rng.standard_normal((nh, n3)), rng.standard_normal((nh, 6))
    tp, tu, ta = (torch.from_numpy(x).to(dev) for x in (p, u, acc))
    b = torch.empty(n3, dtype=torch.float64, device=dev)
    G, H, B = (outs[k].cpu().numpy() for k in ("G", "H", "B"))
    ctx.history(n3, L, outs["G"].data_ptr(), outs["H"].data_ptr(), outs["B"].data_ptr(), tp.data_ptr(), tu.data_ptr(),
                ta.data_ptr(), nh, b.data_ptr())
    want = reference(G, H, B, p, u, acc, L)
    assert np.linalg.norm(b.cpu().numpy() - want) <= 1e-13 * np.linalg.norm(want)
    ctx.history(n3, L, 0, outs["H"].data_ptr(), outs["B"].data_ptr(), 0, tu.data_ptr(), ta.data_ptr(), nh, b.data_ptr())
    want = reference(np.zeros_like(G), H, B, p, u, acc, L)
    assert np.linalg.norm(b.cpu().numpy() - want) <= 1e-13 * np.linalg.norm(want)


@pytest.mark.parametrize("rows", ["1", "2", "4"])
@pytest.mark.parametrize("n3,L,nh", [(2, 3, 2), (10, 4, 3), (130, 6, 9), (390, 5, 4), (1026, 9, 12)])
def test_history_even_sizes_all_terms(ctx, monkeypatch, rows, n3, L, nh):
    """Even 3 nv with 16-byte aligned operands takes the vectorized kernel (CQBEM_HISTORY_ROWS = 2 or 4 rows per
    wavefront, 4 the default; 1 is the one-row kernel), with rows past the end when n3 is not a multiple of the row count; each
    of G, H may be NULL.  A vector 8 bytes off alignment falls back to the one-row kernel."""
    import torch
    monkeypatch.setenv("CQBEM_HISTORY_ROWS", rows)
    rng = np.random.default_rng(n3 + 1)
    G, H = rng.standard_normal((L, n3, n3)), rng.standard_normal((L, n3, n3))
    B = rng.standard_normal((L, n3, 6))
    p, u, acc = rng.standard_normal((nh, n3)), rng.standard_normal((nh, n3)), rng.standard_normal((nh, 6))
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in dict(G=G, H=H, B=B, p=p, u=u, acc=acc).items()}
    b = torch.full((n3,), float("nan"), dtype=torch.float64, device=dev)
    zero = np.zeros_like(G)
    torch.cuda.synchronize()  # buffers filled on torch's stream; the library runs on its own
    for use_g, use_h in ((True, True), (True, False), (False, True), (False, False)):
        print(f"history rows={rows} n3={n3} L={L} nh={nh} G={use_g} H={use_h}", flush=True)
        ctx.history(n3, L, t["G"].data_ptr() if use_g else 0, t["H"].data_ptr() if use_h else 0, t["B"].data_ptr(),
                    t["p"].data_ptr() if use_g else 0, t["u"].data_ptr() if use_h else 0, t["acc"].data_ptr(), nh,
                    b.data_ptr())
        want = reference(G if use_g else zero, H if use_h else zero, B, p, u, acc, L)
        assert np.linalg.norm(b.cpu().numpy() - want) <= 1e-13 * np.linalg.norm(want)
    # p one double into a larger buffer: 8 bytes off a 16-byte boundary
    pp = torch.zeros(nh * n3 + 1, dtype=torch.float64, device=dev)
    pp[1:] = t["p"].reshape(-1)
    b.fill_(float("nan"))
    torch.cuda.synchronize()  # buffers filled on torch's stream; the library runs on its own
    ctx.history(n3, L, t["G"].data_ptr(), t["H"].data_ptr(), t["B"].data_ptr(), pp.data_ptr() + 8, t["u"].data_ptr(),
                t["acc"].data_ptr(), nh, b.data_ptr())
    want = reference(G, H, B, p, u, acc, L)
    assert np.linalg.norm(b.cpu().numpy() - want) <= 1e-13 * np.linalg.norm(want)


def test_history_full_size_against_device_gemv(ctx):
    """At the icosphere-L4 size (3 nv = 7686; L = 8 to keep it at 7.6 GB), b against torch's FP64 matrix-vector
    products on the device (rocBLAS), summed in another order: relative error <= 1e-13.  H_0^-1 b's matvec too."""
    import torch
    n3, L, nh = 7686, 8, 10
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(11)
    G = torch.randn((L, n3, n3), dtype=torch.float64, device=dev, generator=g)
    H = torch.randn((L, n3, n3), dtype=torch.float64, device=dev, generator=g)
    B = torch.randn((L, n3, 6), dtype=torch.float64, device=dev, generator=g)
    p = torch.randn((nh, n3), dtype=torch.float64, device=dev, generator=g)
    u = torch.randn((nh, n3), dtype=torch.float64, device=dev, generator=g)
    a = torch.randn((nh, 6), dtype=torch.float64, device=dev, generator=g)
    b = torch.full((n3,), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()  # the inputs come from torch's stream; the library runs on its own
    print(f"history full size n3={n3} L={L} nh={nh}", flush=True)
    ctx.history(n3, L, G.data_ptr(), H.data_ptr(), B.data_ptr(), p.data_ptr(), u.data_ptr(), a.data_ptr(), nh, b.data_ptr())
    want = torch.zeros(n3, dtype=torch.float64, device=dev)
    for i in range(1, L):
        want += torch.mv(G[i], p[nh - i]) - torch.mv(H[i], u[nh - i]) + torch.mv(B[i], a[nh - i])
    assert (torch.linalg.norm(b - want) / torch.linalg.norm(want)).item() <= 1e-13
    y = torch.full((n3,), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    print(f"matvec full size n={n3}", flush=True)
    ctx.matvec(n3, G[0].data_ptr(), p[0].data_ptr(), y.data_ptr())
    want = torch.mv(G[0], p[0])
    assert (torch.linalg.norm(y - want) / torch.linalg.norm(want)).item() <= 1e-13
    del G, H
    torch.cuda.empty_cache()


@pytest.mark.parametrize("rows", ["1", "2", "4"])
@pytest.mark.parametrize("n", [1, 2, 63, 258, 1030, 1031])
def test_matvec(ctx, monkeypatch, rows, n):
    """cqbem_matvec (the u_p p and H_0^-1 b products of a time step) against numpy; relative error <= 1e-14.  Even n
    takes the vectorized kernel with 2 or 4 rows per wavefront (CQBEM_HISTORY_ROWS), odd n the one-row kernel."""
    import torch
    monkeypatch.setenv("CQBEM_HISTORY_ROWS", rows)
    rng = np.random.default_rng(n)
    A, x = rng.standard_normal((n, n)), rng.standard_normal(n)
    dev = torch.device("cuda:0")
    tA, tx = torch.from_numpy(A).to(dev), torch.from_numpy(x).to(dev)
    y = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()  # buffers filled on torch's stream; the library runs on its own
    print(f"matvec rows={rows} n={n}", flush=True)
    ctx.matvec(n, tA.data_ptr(), tx.data_ptr(), y.data_ptr())
    want = A @ x
    assert np.linalg.norm(y.cpu().numpy() - want) <= 1e-14 * np.linalg.norm(A) * np.linalg.norm(x)
