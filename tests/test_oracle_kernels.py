"""Pins the oracle's Laplace-domain kernels (G_collocation.cu:44-60, H_collocation.cu:45-76,
B_trans_collocation.cu:41-58, B_angular_collocation.cu:44-63 of the reference) against physics
that does not come from the reference's formulas: the Kelvin static limits, the Laplace-domain
Navier equation, the traction operator, and the divergence identities of the fictitious-force kernels.
All derivative checks use the long-double oracle mode and 4th-order central differences."""
import numpy as np
import pytest

RHO, C2, C1 = 1000.0, 13.87, 25.94
MU = RHO * C2 ** 2
LAM = RHO * C1 ** 2 - 2 * MU
NU = LAM / (2 * (LAM + MU))
X = np.array([0.01, -0.02, 0.03])
POINTS = [np.array([0.08, 0.05, -0.03]), np.array([-0.2, 0.1, 0.15]), np.array([0.011, -0.019, 0.0305])]
FREQS = [350 + 200j, 40 - 900j, 2500 + 10j]
LD = 2


def kern(O, which, y, n, s, mode=LD, x=X):
    return O.kernel(which, x, y, n, C1, C2, RHO, s, mode=mode)


def d1(f, y, k, h):
    e = np.zeros(3); e[k] = h
    return (-f(y + 2 * e) + 8 * f(y + e) - 8 * f(y - e) + f(y - 2 * e)) / (12 * h)


def kelvin_u(y):
    r = y - X; d = np.linalg.norm(r); rh = r / d
    return ((3 - 4 * NU) * np.eye(3) + np.outer(rh, rh)) / (16 * np.pi * MU * (1 - NU) * d)


def kelvin_p(y, n):
    r = y - X; d = np.linalg.norm(r); rh = r / d; rn = rh @ n
    return -(rn * ((1 - 2 * NU) * np.eye(3) + 3 * np.outer(rh, rh))
             - (1 - 2 * NU) * (np.outer(rh, n) - np.outer(n, rh))) / (8 * np.pi * (1 - NU) * d ** 2)


@pytest.mark.parametrize("y", POINTS[:2])
def test_static_limits_are_kelvin(oracle, y):
    n = np.array([0.3, -0.5, 0.81]); n /= np.linalg.norm(n)
    r = np.linalg.norm(y - X)
    s = 1e-3 * C2 / r  # |s| r / c ~ 1e-3; Richardson removes the O(s) term
    u_ext = 2 * kern(oracle, 0, y, n, s / 2) - kern(oracle, 0, y, n, s)
    p_ext = 2 * kern(oracle, 1, y, n, s / 2) - kern(oracle, 1, y, n, s)
    ku, kp = kelvin_u(y), kelvin_p(y, n)
    assert np.abs(u_ext - ku).max() < 1e-6 * np.abs(ku).max()
    assert np.abs(p_ext - kp).max() < 1e-5 * np.abs(kp).max()
    assert np.abs(u_ext.imag).max() < 1e-12 * np.abs(ku).max()


@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_conjugation_symmetry(oracle, which):
    n = np.array([0.0, 0.6, 0.8])
    for s in FREQS:
        a = kern(oracle, which, POINTS[0], n, s, mode=0)
        b = kern(oracle, which, POINTS[0], n, np.conj(s), mode=0)
        assert np.allclose(a, np.conj(b), rtol=1e-13, atol=0)
    u = kern(oracle, 0, POINTS[0], n, FREQS[0], mode=0)
    assert np.allclose(u, u.T, rtol=1e-14, atol=0)


@pytest.mark.parametrize("y", POINTS)
@pytest.mark.parametrize("s", FREQS)
def test_navier_residual(oracle, y, s):
    """mu lap u + (lam+mu) grad div u - rho s^2 u = 0 away from the source, for every force direction."""
    r = np.linalg.norm(y - X); h = 2e-3 * min(r, C2 / abs(s))
    u = lambda p: kern(oracle, 0, p, np.zeros(3), s)  # u[i, j]: displacement j due to force i
    lap = np.zeros((3, 3), complex)
    hess = np.zeros((3, 3, 3, 3), complex)  # [a, b, i, j] = d_a d_b u_ij
    for a in range(3):
        for b in range(3):
            hess[a, b] = d1(lambda p: d1(u, p, b, h), y, a, h)
    lap = sum(hess[a, a] for a in range(3))
    graddiv = np.einsum("jaia->ij", hess.transpose(0, 1, 2, 3).reshape(3, 3, 3, 3).transpose(1, 0, 2, 3))
    # graddiv[i, j] = d_j (sum_a d_a u_ia)
    graddiv = np.array([[sum(hess[j, a, i, a] for a in range(3)) for j in range(3)] for i in range(3)])
    res = MU * lap + (LAM + MU) * graddiv - RHO * s ** 2 * u(y)
    scale = max(np.abs(MU * lap).max(), np.abs(RHO * s ** 2 * u(y)).max())
    assert np.abs(res).max() < 2e-6 * scale


@pytest.mark.parametrize("y", POINTS)
@pytest.mark.parametrize("s", FREQS)
def test_traction_of_displacement_kernel(oracle, y, s):
    """p_hat_ij = sigma_jk(u_i.) n_k with sigma = lam div u I + mu (grad u + grad u^T)."""
    n = np.array([0.48, -0.6, 0.64]); n /= np.linalg.norm(n)
    r = np.linalg.norm(y - X); h = 1e-3 * min(r, C2 / abs(s))
    u = lambda p: kern(oracle, 0, p, n, s)
    grad = np.stack([d1(u, y, k, h) for k in range(3)])  # grad[k, i, j] = d_k u_ij
    t = np.zeros((3, 3), complex)
    for i in range(3):
        g = grad[:, i, :]  # g[k, j] = d_k u_j
        div = np.trace(g)
        sigma = LAM * div * np.eye(3) + MU * (g + g.T)
        t[i] = sigma @ n
    p = kern(oracle, 1, y, n, s)
    assert np.abs(p - t).max() < 1e-7 * np.abs(p).max()


@pytest.mark.parametrize("y", POINTS)
@pytest.mark.parametrize("s", FREQS)
def test_fictitious_force_kernels_divergence(oracle, y, s):
    """b_hat_ij(n) = W_ijk n_k with div_k W_ijk = -rho u_ij, and b_hat_ijl(n) = W_ijlk n_k with
    div_k W_ijlk = -rho u_ij (y - x)_l: the surface integrals of the B kernels are the volume
    integrals of the displacement response to the body forces -rho a and -rho (alpha x y)."""
    r = np.linalg.norm(y - X); h = 1e-3 * min(r, C2 / abs(s))
    uy = kern(oracle, 0, y, np.zeros(3), s)
    # scale: the divergence may be exponentially small next to the non-decaying 1/s^2 terms of W
    wt = max(np.abs(kern(oracle, 2, y, np.eye(3)[k], s)).max() for k in range(3)) / r
    div_t = sum(d1(lambda p: kern(oracle, 2, p, np.eye(3)[k], s), y, k, h) for k in range(3))
    assert np.abs(div_t + RHO * uy).max() < 1e-8 * max(np.abs(RHO * uy).max(), wt)
    wa = max(np.abs(kern(oracle, 3, y, np.eye(3)[k], s)).max() for k in range(3)) / r
    div_a = sum(d1(lambda p: kern(oracle, 3, p, np.eye(3)[k], s), y, k, h) for k in range(3))
    expect = -RHO * uy[:, :, None] * (y - X)[None, None, :]
    assert np.abs(div_a - expect).max() < 1e-8 * max(np.abs(expect).max(), wa)


def test_b_kernels_linear_in_normal(oracle):
    n1 = np.array([0.3, 0.4, 0.5]); n2 = np.array([-0.7, 0.2, 0.1])
    for which in (1, 2, 3):
        a = kern(oracle, which, POINTS[0], n1 + 2 * n2, FREQS[0], mode=0)
        b = kern(oracle, which, POINTS[0], n1, FREQS[0], mode=0) + 2 * kern(oracle, which, POINTS[0], n2, FREQS[0], mode=0)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_reference_precision_mode_is_float(oracle):
    """Mode 1 evaluates in float32: it differs from fp64 at the 1e-7 level, not more, not less."""
    a = kern(oracle, 0, POINTS[1], np.zeros(3), FREQS[0], mode=1)
    b = kern(oracle, 0, POINTS[1], np.zeros(3), FREQS[0], mode=0)
    rel = np.abs(a - b).max() / np.abs(b).max()
    assert 1e-9 < rel < 1e-5
