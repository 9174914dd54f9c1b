"""GPU parity of the Galerkin assembly (non-singular pairs + Sauter-Schwab touching pairs) against the
CPU oracle, per frequency and in the time domain.  Small quadrature orders keep the oracle fast; one
case runs the shipped order 5 x subdivision 2 rule at a single frequency.

Two non-singular decompositions are checked against the same bounds: "points" (default: collocation
rows at every test quadrature point, then projection onto the test basis) and "star" (one lane per
test face and source vertex star; CQBEM_GALERKIN_KERNEL=star).  Touching pairs: the quad kernels for G
and H (default) and the one-lane-per-point kernels (CQBEM_GALERKIN_SINGULAR=lane)."""
import numpy as np
import pytest

from surfaceonlydeformables_amd import abi
from surfaceonlydeformables_amd import mesh as M

pytestmark = pytest.mark.gpu

RHO = 1100.0


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


def mesh(name):
    if name == "ico0":
        V, F = M.icosphere(0, 0.1)
    elif name == "ico1":
        V, F = M.icosphere(1, 0.1)
        V = M.jitter(V, F, 0.01, 0)
    else:
        V, F = M.cube(2, 0.1)
    return M.center(V, F), F


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


LAPLACE_CASES = [(m, o, s, l, "points") for (m, o, s) in [("ico0", 5, 2), ("ico1", 2, 1), ("cube2", 3, 1)] for l in (0, 7)]
LAPLACE_CASES += [("ico1", 2, 1, 7, "star"), ("cube2", 3, 1, 0, "star")]


@pytest.mark.parametrize("name,order,sub,l,variant", LAPLACE_CASES)
def test_galerkin_laplace_parity(ctx, oracle, name, order, sub, l, variant, monkeypatch):
    monkeypatch.setenv("CQBEM_GALERKIN_KERNEL", variant)
    V, F = mesh(name)
    _, c1, c2 = M.material_speeds(1e6, 0.35, RHO)
    L, dt = 16, 5e-4
    R = 1e-10 ** (1 / (2 * L))
    print(f"galerkin laplace {name} order {order} sub {sub} l={l} {variant}", flush=True)
    g = ctx.laplace(abi.make_mesh(V, F), abi.make_params(c1, c2, RHO, 0.35, dt, R, L, order=order, subdivision=sub,
                                                        galerkin=True), l)
    s = oracle.frequencies(L, R, dt, 2)[l]
    o = oracle.laplace(V, F, oracle.params(c1, c2, RHO, 0.35, dt, R, L, order=order, subdiv=sub, galerkin=True,
                                           nthreads=0), s)
    assert rel(g["U"], o["U"]) < 1e-13
    assert rel(g["P"], o["P"]) < 1e-10
    assert rel(g["Bt"], o["Bt"]) < 1e-13
    assert rel(g["Ba"], o["Ba"]) < 1e-13


@pytest.mark.parametrize("name,L,bdf,variant", [("ico1", 16, 2, "points"), ("cube2", 9, 1, "points"), ("ico1", 16, 2, "star"),
                                                ("cube2", 9, 1, "points/lane")])
def test_galerkin_time_domain_parity(ctx, oracle, name, L, bdf, variant, monkeypatch):
    """Touching pairs run the quad kernels (G, H) by default; "/lane" selects the one-lane-per-point kernels
    (CQBEM_GALERKIN_SINGULAR=lane)."""
    monkeypatch.setenv("CQBEM_GALERKIN_KERNEL", variant.split("/")[0])
    monkeypatch.setenv("CQBEM_GALERKIN_SINGULAR", "lane" if variant.endswith("/lane") else "quad")
    V, F = mesh(name)
    _, c1, c2 = M.material_speeds(1e6, 0.35, RHO)
    dt = 5e-4
    R = 1e-10 ** (1 / (2 * L))
    print(f"galerkin time domain {name} L={L} {variant}", flush=True)
    res = ctx.assemble(abi.make_mesh(V, F), abi.make_params(c1, c2, RHO, 0.35, dt, R, L, bdf=bdf, order=2,
                                                           subdivision=1, galerkin=True))
    op = lambda mode: oracle.params(c1, c2, RHO, 0.35, dt, R, L, bdf=bdf, order=2, subdiv=1, galerkin=True, mode=mode,
                                    nthreads=0)
    f64 = oracle.assemble(V, F, op(oracle.MODE_FP64))
    ld = oracle.assemble(V, F, op(oracle.MODE_LD))
    for mine, theirs in (("G", "G"), ("H", "H"), ("B_trans", "Bt"), ("B_angular", "Ba"), ("B", "B6")):
        a = res[mine].reshape(L, -1); b = f64[theirs].reshape(L, -1); c = ld[theirs].reshape(L, -1)
        assert rel(a, b) < 1e-8, mine
        for n in range(L):
            e_gpu, e_f64 = rel(a[n], c[n]), rel(b[n], c[n])
            assert e_gpu <= max(2 * e_f64, 1e-12), (mine, n, e_gpu, e_f64)


@pytest.mark.parametrize("name,order,sub", [("ico1", 3, 1), ("cube2", 2, 2)])
def test_touching_pair_symmetry_matches_full_integration(ctx, name, order, sub, monkeypatch):
    """The default path integrates coincident pairs over regions 1, 3, 5 and adds the exchanged half (G even, H odd
    in y - x; B folds it into its weights, Lx - Ly for its odd terms and Lx + Ly for its even ones), and takes each
    vertex-adjacent mirror pair's G block from its primary's transpose; the lane kernels (CQBEM_GALERKIN_SINGULAR=
    lane) integrate all regions of every ordered pair.  Same matrices to rounding: a wrong sign, corner swap or
    missing half is an O(1) error in the touching blocks."""
    V, F = mesh(name)
    _, c1, c2 = M.material_speeds(1e6, 0.35, RHO)
    L, dt = 16, 5e-4
    R = 1e-10 ** (1 / (2 * L))
    prm = abi.make_params(c1, c2, RHO, 0.35, dt, R, L, order=order, subdivision=sub, galerkin=True)
    res = {}
    for variant in ("quad", "lane"):
        monkeypatch.setenv("CQBEM_GALERKIN_SINGULAR", variant)
        print(f"touching-pair symmetry {name} order {order} sub {sub}: {variant}", flush=True)
        res[variant] = ctx.laplace(abi.make_mesh(V, F), prm, 5)
    for k in ("U", "P", "Bt", "Ba"):
        e = rel(res["quad"][k], res["lane"][k])
        print(f"  {k}: quad vs lane {e:.2e}", flush=True)
        assert e < 1e-13, (k, e)


def test_galerkin_single_layer_symmetry(ctx):
    """Size-independent property: the Galerkin G weights are symmetric matrices for every n."""
    V, F = M.icosphere(3, 0.1)
    V = M.center(M.jitter(V, F, 0.01, 0), F)
    _, c1, c2 = M.material_speeds(1e6, 0.35, RHO)
    L = 8
    R = 1e-10 ** (1 / (2 * L))
    print("galerkin symmetry ico3", flush=True)
    res = ctx.assemble(abi.make_mesh(V, F), abi.make_params(c1, c2, RHO, 0.35, 5e-4, R, L, order=5, subdivision=1,
                                                           galerkin=True, which=abi.G))
    # exact integrals are symmetric; the Sauter-Schwab rule of order 5 x 1 leaves ~1e-6 asymmetry in
    # edge/vertex-adjacent pairs, measured against the largest weight
    gmax = np.abs(res["G"]).max()
    for n in range(L):
        Gn = res["G"][n]
        assert np.abs(Gn - Gn.T).max() < 1e-6 * gmax, n


def test_galerkin_chunks_are_bitwise_identical():
    """A small HBM budget cuts the test faces into chunks of 3 faces for the per-test-point rows; the faces
    of one colour class share no vertex, so every output row still receives its face contributions in
    colour-class order and the result is bit-identical to the unchunked run."""
    import torch
    V, F = mesh("ico1")
    _, c1, c2 = M.material_speeds(1e6, 0.35, RHO)
    L = 16
    params = abi.make_params(c1, c2, RHO, 0.35, 5e-4, 1e-10 ** (1 / (2 * L)), L, order=2, subdivision=1, galerkin=True)
    m = abi.make_mesh(V, F)
    nv, nf, npts, nl = len(V), len(F), 3, L // 2 + 1
    per_face = npts * 3 * nl * (2 * 3 * nv + 12) * 16  # G, H and B rows of one test face's points
    full_fp = abi.job_footprint(m, params, outputs_on_device=True)  # all faces in one chunk
    n3 = 3 * nv
    dev = torch.device("cuda:0")
    shapes = {"G": (L, n3, n3), "H": (L, n3, n3), "B": (L, n3, 6)}
    full = abi.Context(0)
    small = abi.Context(0, hbm_budget_bytes=full_fp - (nf - 1 - 3) * per_face)
    print("galerkin chunks ico1", flush=True)
    try:
        res = []
        for ctx in (full, small):
            outs = {k: torch.full(s, float("nan"), dtype=torch.float64, device=dev) for k, s in shapes.items()}
            torch.cuda.synchronize()  # buffers filled on torch's stream; the library runs on its own
            ctx.assemble_device(m, params, {k: t.data_ptr() for k, t in outs.items()})
            res.append({k: t.cpu().numpy() for k, t in outs.items()})
        for k in shapes:
            assert np.isfinite(res[1][k]).all(), k
            assert np.array_equal(res[0][k], res[1][k]), k
    finally:
        full.close()
        small.close()
