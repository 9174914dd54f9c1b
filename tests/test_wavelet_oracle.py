"""Pins of oracle/cqbem_wavelet.py (the Haar compression of include/wavelets.hpp) by what the mathematics
fixes, independently of the restatement's own formulas:
- the 2x2 case in closed form: alpha = (a - b - c + d)/2, beta = (a + b - c - d)/2, gamma = (a - b + c - d)/2,
  coarse = (a + b + c + d)/2 at (0,0), (0,1), (1,0), (1,1);
- each level is an orthogonal 2D Haar step, so the squared coefficients sum to ||A||_F^2 (Parseval);
- with every coefficient kept, reconstruct(M_c decompose(x)) = A x, for sizes that are and are not
  powers of two (the nonstandard-form product identity the reference's time stepping relies on);
- the vector transform of (x0, x1) is ((x0 - x1)/sqrt 2, (x0 + x1)/sqrt 2);
- keeping the k largest coefficients keeps exactly k, all at least as large as every dropped one, and
  the product error is bounded by the dropped coefficients (||M - M_k||_F ||x||)."""
import numpy as np
import pytest

from cqbem_wavelet import candidates, compress, decompose_vec, padded, product, reconstruct


def test_two_by_two_closed_form():
    a, b, c, d = 1.5, -2.0, 0.25, 3.0
    M = compress(np.array([[a, b], [c, d]]), 4).toarray()
    want = np.array([[a - b - c + d, a + b - c - d], [a - b + c - d, a + b + c + d]]) / 2
    assert np.allclose(M, want, rtol=1e-15, atol=0)


@pytest.mark.parametrize("n", [2, 5, 16, 37])
def test_parseval(n):
    A = np.random.default_rng(n).standard_normal((n, n))
    _, _, v = candidates(A)
    assert len(v) == padded(n) ** 2
    assert abs(np.sum(v * v) - np.sum(A * A)) <= 1e-13 * np.sum(A * A)


@pytest.mark.parametrize("n", [2, 8, 13, 64, 100])
def test_full_product_is_the_dense_product(n):
    rng = np.random.default_rng(100 + n)
    A, x = rng.standard_normal((n, n)), rng.standard_normal(n)
    M = compress(A, padded(n) ** 2)
    y = product(M, x)
    assert np.linalg.norm(y - A @ x) <= 1e-13 * np.linalg.norm(A) * np.linalg.norm(x)


def test_vector_transform_closed_form():
    x0, x1 = 3.0, -1.25
    c = decompose_vec(np.array([x0, x1]))
    assert np.allclose(c, [(x0 - x1) / np.sqrt(2), (x0 + x1) / np.sqrt(2)], rtol=1e-15, atol=0)
    # a vector whose product with the identity goes back to itself
    x = np.random.default_rng(3).standard_normal(11)
    assert np.allclose(product(compress(np.eye(11), 256), x), x, rtol=0, atol=1e-14)


@pytest.mark.parametrize("frac", [0.05, 0.3])
def test_truncation(frac):
    n = 48
    rng = np.random.default_rng(7)
    # a smooth kernel matrix (what G and H look like) plus a diagonal
    t = np.linspace(0, 1, n)
    A = 1.0 / (0.05 + np.abs(t[:, None] - t[None, :])) + np.eye(n)
    k = int(n * n * frac)
    M = compress(A, k)
    full = compress(A, padded(n) ** 2)
    assert M.nnz == k
    kept = np.abs(M.data).min()
    dropped = (full - M)
    dropped.eliminate_zeros()
    assert np.abs(dropped.data).max() <= kept
    x = rng.standard_normal(n)
    err = np.linalg.norm(product(M, x) - A @ x)
    assert err <= np.sqrt(np.sum(dropped.data ** 2)) * np.linalg.norm(x) * (1 + 1e-12)


def test_ties_keep_row_major_order():
    A = np.ones((4, 4))  # all level-0 coefficients vanish except through the coarse parts
    A[0, 0] = 2.0
    r, c, v = candidates(A)
    nzv = np.abs(v[v != 0])
    k = 3
    M = compress(A, k)
    assert M.nnz == k
    T = np.sort(nzv)[::-1][k - 1]
    coo = M.tocoo()
    # every kept magnitude >= T, and among the ties the first in row-major order
    assert np.all(np.abs(coo.data) >= T)
    ties = sorted((ri, ci) for ri, ci, vi in zip(r, c, v) if vi != 0 and abs(vi) == T)
    kept_ties = sorted((ri, ci) for ri, ci, vi in zip(coo.row, coo.col, coo.data) if abs(vi) == T)
    assert kept_ties == ties[:len(kept_ties)]


@pytest.mark.parametrize("k", [0, -1])
def test_nonpositive_count_keeps_every_nonzero(k):
    """decompose_mat keeps every nonzero coefficient when it gets no positive count (wavelets.hpp:168-184), as a
    compression ratio with (int)(n^2 ratio) = 0 gives; the product is then the dense product."""
    n = 13
    A = np.random.default_rng(2).standard_normal((n, n))
    A[3:5, :] = 0.0
    M = compress(A, k)
    full = compress(A, padded(n) ** 2)
    assert M.nnz == full.nnz > 0
    assert np.array_equal(M.indptr, full.indptr) and np.array_equal(M.indices, full.indices)
    assert np.array_equal(M.data, full.data)
    x = np.random.default_rng(4).standard_normal(n)
    assert np.allclose(product(M, x), A @ x, rtol=0, atol=1e-12 * np.abs(A @ x).max())
