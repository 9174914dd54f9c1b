"""Pins the CQ frequency set and the half-spectrum inverse transform (G.cu:19-24,65,166-179 of the
reference) with known convolution-quadrature weights: for a kernel K(s), the transform of K(s_l)
must equal the aliased power-series coefficients sum_k R^(kL) w_(n+kL) of K(gamma(z)/dt)."""
import mpmath as mp
import numpy as np
import pytest

mp.mp.dps = 40


def gamma_poly(bdf):
    return [mp.mpf(1), mp.mpf(-1)] if bdf == 1 else [mp.mpf(3) / 2, mp.mpf(-2), mp.mpf(1) / 2]


def series_exp_neg(tau, bdf, m):
    """Coefficients of exp(-tau gamma(z)) up to z^m (exact via the ODE g' = -tau gamma'(z) g)."""
    g = gamma_poly(bdf)
    dg = [k * g[k] for k in range(1, len(g))]
    c = [mp.e ** (-tau * g[0])] + [mp.mpf(0)] * m
    for n in range(m):
        acc = mp.mpf(0)
        for k, d in enumerate(dg):
            if n - k >= 0:
                acc += d * c[n - k]
        c[n + 1] = -tau * acc / (n + 1)
    return c


def series_inv_gamma(bdf, m, power=1):
    g = gamma_poly(bdf)
    inv = [mp.mpf(1) / g[0]] + [mp.mpf(0)] * m  # 1/gamma by long division
    for n in range(1, m + 1):
        acc = sum(g[k] * inv[n - k] for k in range(1, len(g)) if n - k >= 0)
        inv[n] = -acc / g[0]
    out = inv
    for _ in range(power - 1):
        out = [sum(out[k] * inv[n - k] for k in range(n + 1)) for n in range(m + 1)]
    return out


def aliased(coef, L, R, n):
    return float(sum(mp.mpf(R) ** (k * L) * coef[n + k * L] for k in range((len(coef) - n - 1) // L + 1)))


CASES = [(32, 2, False), (32, 1, False), (9, 2, False), (8, 1, True), (17, 2, True), (64, 2, False)]


@pytest.mark.parametrize("L,bdf,ref_R", CASES)
@pytest.mark.parametrize("kind", ["exp", "inv_s", "inv_s2"])
def test_transform_known_weights(oracle, L, bdf, ref_R, kind):
    dt = 1e-3
    R = 1e-10 ** (1 / (2 * (L - 1))) if ref_R else 1e-10 ** (1 / (2 * L))
    s = oracle.frequencies(L, R, dt, bdf)
    m = 40 * L
    if kind == "exp":
        T = 3.7 * dt
        spec = np.exp(-s * T)
        coef = series_exp_neg(mp.mpf(T) / mp.mpf(dt), bdf, m)
    elif kind == "inv_s":
        spec = 1 / s
        coef = [c * mp.mpf(dt) for c in series_inv_gamma(bdf, m)]
    else:
        spec = 1 / s ** 2
        coef = [c * mp.mpf(dt) ** 2 for c in series_inv_gamma(bdf, m, power=2)]
    W = oracle.transform(spec[:, None], L, R)[:, 0]
    expect = np.array([aliased(coef, L, R, n) for n in range(L)])
    scale = np.abs(expect).max()
    assert np.abs(W - expect).max() < 2e-9 * scale, (W[:4], expect[:4])


@pytest.mark.parametrize("L,bdf", [(32, 2), (8, 1), (9, 2)])
def test_frequencies(oracle, L, bdf):
    R, dt = 0.7, 5e-4
    s = oracle.frequencies(L, R, dt, bdf)
    assert len(s) == L // 2 + 1
    z = R * np.exp(2j * np.pi * np.arange(L // 2 + 1) / L)
    g = (1 - z) if bdf == 1 else (1.5 - 2 * z + z * z / 2)
    assert np.allclose(s, g / dt, rtol=1e-15, atol=0)
    assert abs(s[0].imag) == 0 and (L % 2 == 1 or abs(s[-1].imag) < 1e-9 * abs(s[-1]))


def test_transform_of_real_constant_is_delta(oracle):
    """A frequency-independent spectrum maps to weight 0 only (the static part of a kernel)."""
    L, R = 32, 0.7
    W = oracle.transform(np.full((L // 2 + 1, 1), 2.5 + 0j), L, R)[:, 0]
    assert abs(W[0] - 2.5) < 1e-14 and np.abs(W[1:]).max() < 1e-9
