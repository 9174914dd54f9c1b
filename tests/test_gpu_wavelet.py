"""Haar compression of dense matrices on the GPU (cqbem_wavelet_compress) and the CSR product
(cqbem_csr_matvec), against oracle/cqbem_wavelet.py (the reference's include/wavelets.hpp, M = 1).

The transform forms the same rounded products and sums in the same order as the oracle, so the kept
coefficients must match bit for bit: the same rows, columns (ascending) and values.  Ties with the smallest
kept magnitude are resolved in row-major order by both.  The product sums in another order: relative
error <= 1e-14."""
import numpy as np
import pytest

from surfaceonlydeformables_amd import abi
from surfaceonlydeformables_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


def gpu_compress(ctx, A, k):
    import torch
    n = A.shape[0]
    npad = abi.wavelet_padded(n)
    rows = 2 * npad - 2
    dev = torch.device("cuda:0")
    tA = torch.from_numpy(np.ascontiguousarray(A)).to(dev)
    rowptr = torch.full((rows + 1,), -1, dtype=torch.int64, device=dev)
    cap = k if k > 0 else npad * npad  # k <= 0 keeps every nonzero coefficient (at most np^2)
    col = torch.full((cap,), -1, dtype=torch.int32, device=dev)
    val = torch.full((cap,), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()  # buffers filled on torch's stream; the library runs on its own
    print(f"wavelet compress n={n} k={k}", flush=True)
    nnz = ctx.wavelet_compress(n, tA.data_ptr(), k, rowptr.data_ptr(), col.data_ptr(), val.data_ptr())
    return nnz, rowptr, col, val


def check_same(ctx, A, k):
    from cqbem_wavelet import compress
    nnz, rowptr, col, val = gpu_compress(ctx, A, k)
    want = compress(A, k)
    assert nnz == want.nnz
    assert np.array_equal(rowptr.cpu().numpy(), want.indptr)
    if nnz:
        assert np.array_equal(col.cpu().numpy()[:nnz], want.indices)
        assert np.array_equal(val.cpu().numpy()[:nnz], want.data)
    return nnz, rowptr, col, val, want


@pytest.mark.parametrize("n", [2, 3, 37, 100, 258])
@pytest.mark.parametrize("frac", [0.0, 0.02, 0.3, 2.0])
def test_compress_random(ctx, n, frac):
    A = np.random.default_rng(n).standard_normal((n, n))
    check_same(ctx, A, int(n * n * frac))


@pytest.mark.parametrize("ratio", [1e-9, 0.1 / (37 * 37)])
def test_compress_count_below_one_keeps_everything(ctx, ratio):
    """(int)(n^2 ratio) = 0 keeps every nonzero coefficient, as the reference's decompose_mat does without a
    positive count (include/wavelets.hpp:168-184): the product is then the dense product."""
    from cqbem_wavelet import padded, product
    n = 37
    A = np.random.default_rng(11).standard_normal((n, n))
    k = int(n * n * ratio)
    assert k == 0
    nnz, rowptr, col, val, want = check_same(ctx, A, k)
    assert nnz == np.count_nonzero(want.data) and nnz > n * n // 2
    x = np.random.default_rng(12).standard_normal(n)
    assert np.allclose(product(want, x), A @ x, rtol=0, atol=1e-12 * np.abs(A @ x).max())


def test_compress_ties_and_zeros(ctx):
    """Integer entries give many coefficients of equal magnitude; a zero block gives zero coefficients, which
    are never kept."""
    rng = np.random.default_rng(5)
    A = rng.integers(-2, 3, size=(40, 40)).astype(np.float64)
    A[20:, 20:] = 0.0
    for k in (1, 50, 333, 900, 4096):
        check_same(ctx, A, k)
    nnz, *_ = check_same(ctx, np.zeros((9, 9)), 40)
    assert nnz == 0


def test_compress_assembled_G_and_product(ctx):
    """An assembled G_1 (icosphere level 1) at the reference's compression ratio 0.1, and its product."""
    import torch
    from cqbem_wavelet import product
    V, F = M.icosphere(1, 0.2)
    V = M.center(M.jitter(V, F, 0.01, 1), F)
    _, c1, c2 = M.material_speeds(5e5, 0.3, 1000.0)
    L = 8
    params = abi.make_params(c1, c2, 1000.0, 0.3, 1e-3, 1e-10 ** (1 / (2 * L)), L)
    n3 = 3 * len(V)
    dev = torch.device("cuda:0")
    G = torch.empty((L, n3, n3), dtype=torch.float64, device=dev)
    H = torch.empty((L, n3, n3), dtype=torch.float64, device=dev)
    B = torch.empty((L, n3, 6), dtype=torch.float64, device=dev)
    ctx.assemble_device(abi.make_mesh(V, F), params, {"G": G.data_ptr(), "H": H.data_ptr(), "B": B.data_ptr()})
    A = G[1].cpu().numpy()
    k = int(n3 * n3 * 0.1)
    nnz, rowptr, col, val, want = check_same(ctx, A, k)
    rows = len(want.indptr) - 1
    x = np.random.default_rng(2).standard_normal(rows)
    tx = torch.from_numpy(x).to(dev)
    y = torch.full((rows,), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()  # buffers filled on torch's stream; the library runs on its own
    ctx.csr_matvec(rows, rowptr.data_ptr(), col.data_ptr(), val.data_ptr(), tx.data_ptr(), y.data_ptr())
    ref = want @ x
    assert np.linalg.norm(y.cpu().numpy() - ref) <= 1e-14 * np.linalg.norm(want.data) * np.linalg.norm(x)
    # the compressed G_1 still acts like G_1 on a smooth vector field
    u = np.tile([1.0, 0.5, -0.25], len(V))
    err = np.linalg.norm(product(want, u) - A @ u) / np.linalg.norm(A @ u)
    assert err < 0.2
