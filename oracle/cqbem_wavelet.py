"""numpy restatement of the reference's Haar wavelet compression -- test infrastructure only.

Only tests/ import this module; the product path never does.

Follows include/wavelets.hpp with M = 1 (CompressedMatrix<> / CompressedVector<> default, the Haar filters
h_2 = (c, c), g_2 = (c, -c), c = M_SQRT1_2, :15-17):
- decompose_mat (:154-219): Beylkin's nonstandard form of the zero-padded matrix; every sum
  g_k g_m s(2i+k, 2l+m) etc. is formed as four rounded products (c*c)*s added in the order
  (k, m) = (0,0), (0,1), (1,0), (1,1), as the reference's loop does;
- the num_elems largest-magnitude nonzero coefficients are kept (:165-176); among coefficients tied with
  the smallest kept magnitude, the first in row-major order are kept;
- decompose_vec (:221-251) and reconstruct (:253-281), and the product reconstruct(M_c decompose(x)).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

C = np.float64(0.70710678118654752440)


def padded(n: int) -> int:
    p = 1
    while p < n:
        p <<= 1
    return p


def candidates(A):
    """All np^2 coefficients of the nonstandard form: (rows, cols, values) of the (2np-2)^2 matrix."""
    n = A.shape[0]
    npad = padded(n)
    s = np.zeros((npad, npad))
    s[:n, :n] = A
    cc = C * C
    R, K, V = [], [], []
    off, sz = 0, npad // 2
    while sz >= 1:
        p00, p01 = cc * s[0::2, 0::2], cc * s[0::2, 1::2]
        p10, p11 = cc * s[1::2, 0::2], cc * s[1::2, 1::2]
        alpha = (((0.0 + p00) - p01) - p10) + p11
        beta = (((0.0 + p00) + p01) - p10) - p11
        gamma = (((0.0 + p00) - p01) + p10) - p11
        coarse = (((0.0 + p00) + p01) + p10) + p11
        I, L = np.meshgrid(np.arange(sz), np.arange(sz), indexing="ij")
        for rr, kk, vv in ((off + I, off + L, alpha), (off + I, off + sz + L, beta), (off + sz + I, off + L, gamma)):
            R.append(rr.ravel()), K.append(kk.ravel()), V.append(vv.ravel())
        s = coarse
        off += 2 * sz
        sz //= 2
    R.append(np.array([2 * npad - 3])), K.append(np.array([2 * npad - 3])), V.append(np.array([s[0, 0]]))
    return np.concatenate(R), np.concatenate(K), np.concatenate(V)


def compress(A, num_elems: int):
    """The kept coefficients as a CSR matrix of shape (2np-2, 2np-2), columns ascending per row."""
    npad = padded(A.shape[0])
    r, c, v = candidates(A)
    nz = v != 0
    r, c, v = r[nz], c[nz], v[nz]
    if 0 < num_elems < len(v):  # num_elems <= 0 keeps every nonzero coefficient (wavelets.hpp:168-184)
        keep = np.zeros(len(v), dtype=bool)
        if num_elems > 0:
            mag = np.abs(v)
            T = -np.partition(-mag, num_elems - 1)[num_elems - 1]
            keep = mag > T
            need = num_elems - int(keep.sum())
            ties = np.flatnonzero(mag == T)
            ties = ties[np.lexsort((c[ties], r[ties]))][:need]
            keep[ties] = True
        r, c, v = r[keep], c[keep], v[keep]
    M = sp.csr_matrix((v, (r, c)), shape=(2 * npad - 2, 2 * npad - 2))
    M.sort_indices()
    return M


def decompose_vec(x):
    n = len(x)
    npad = padded(n)
    s = np.zeros(npad)
    s[:n] = x
    out = np.zeros(2 * npad - 2)
    off, sz = 0, npad // 2
    while sz >= 1:
        d = (0.0 + C * s[0::2]) + (-C) * s[1::2]
        ns = (0.0 + C * s[0::2]) + C * s[1::2]
        out[off:off + sz], out[off + sz:off + 2 * sz] = d, ns
        s = ns
        off += 2 * sz
        sz //= 2
    return out


def reconstruct(c, n: int):
    npad = (len(c) + 2) // 2
    res = np.zeros(npad)
    offset, size = -2, 1
    while size <= npad // 2:
        base = 2 * npad - 2 + offset
        d = c[base:base + size]
        s = c[base + size:base + 2 * size] + res[:size]
        res[:2 * size] = 0.0
        res[0:2 * size:2] = 0.0 + ((C * s) + (C * d))
        res[1:2 * size:2] = 0.0 + ((C * s) + (-C * d))
        offset -= 4 * size
        size *= 2
    return res[:n]


def product(M, x):
    """reconstruct(M decompose(x)): the compressed matrix-vector product mapped back."""
    return reconstruct(M @ decompose_vec(x), len(x))
