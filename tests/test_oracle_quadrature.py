"""Pins the oracle's quadrature rules (integrator.hpp, integrator_collocation.hpp of the reference)
against polynomial exactness, numpy's Gauss-Legendre nodes and closed-form triangle integrals."""
import math

import numpy as np
import pytest


def exact_monomial(a, b):
    # integral of x^a y^b over the unit right triangle
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


# Polynomial degree each rule integrates exactly.  The reference's order-3 rule
# (integrator.hpp:64-68) places its three outer points at (11/15, 2/15, 2/15)
# instead of the textbook (3/5, 1/5, 1/5); with the -27/48 / 25/48 weights it is
# exact only to degree 1.  Oracle and product keep the reference's points.
EXACT_DEGREE = {1: 1, 2: 2, 3: 1, 4: 4, 5: 5}


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_triangle_rule_exactness(oracle, order):
    bary, w = oracle.tri_rule(order)
    assert len(w) == [1, 3, 4, 6, 7][order - 1]
    assert abs(w.sum() - 1) < 1e-14
    assert np.allclose(bary.sum(axis=1), 1.0, atol=1e-15)
    order = EXACT_DEGREE[order]
    for deg in range(order + 1):
        for a in range(deg + 1):
            b = deg - a
            approx = 0.5 * np.sum(w * bary[:, 1] ** a * bary[:, 2] ** b)
            assert abs(approx - exact_monomial(a, b)) < 1e-13, (order, a, b)
    # one degree higher is not integrated exactly (the rule really has that order)
    deg = order + 1
    errs = [abs(0.5 * np.sum(w * bary[:, 1] ** a * bary[:, 2] ** (deg - a)) - exact_monomial(a, deg - a))
            for a in range(deg + 1)]
    assert max(errs) > 1e-12


@pytest.mark.parametrize("order,sub", [(5, 1), (5, 2), (3, 3), (4, 4), (2, 5)])
def test_composite_rule_exactness(oracle, order, sub):
    bary, w = oracle.tri_rule_composite(order, sub)
    assert len(w) == len(oracle.tri_rule(order)[1]) * sub * sub
    assert abs(w.sum() - 1) < 1e-13
    assert (bary > -1e-15).all()
    for deg in range(EXACT_DEGREE[order] + 1):
        for a in range(deg + 1):
            approx = 0.5 * np.sum(w * bary[:, 1] ** a * bary[:, 2] ** (deg - a))
            assert abs(approx - exact_monomial(a, deg - a)) < 1e-13


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_gauss_rule_matches_leggauss(oracle, order):
    x, w = oracle.gauss_rule(order)
    xr, wr = np.polynomial.legendre.leggauss(order)
    assert np.allclose(x, xr, atol=1e-15, rtol=0)
    assert np.allclose(w, wr, atol=1e-15, rtol=0)


TRIS = [
    (np.array([0.1, -0.2, 0.3]), np.array([1.1, 0.0, 0.2]), np.array([0.3, 0.9, -0.4])),
    (np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])),
    (np.array([2.0, 1.0, -1.0]), np.array([2.5, 1.8, -0.7]), np.array([1.4, 1.6, -0.2])),
]


@pytest.mark.parametrize("tri", TRIS)
@pytest.mark.parametrize("order,sub", [(2, 1), (5, 2), (4, 3)])
def test_far_rule_quadratic(oracle, tri, order, sub):
    """The collocation far-field rule integrates 1 + a.y + (b.y)^2 exactly (order >= 2)."""
    y1, y2, y3 = tri
    a = np.array([0.3, -1.2, 0.7]); b = np.array([1.5, 0.4, -0.8])
    area = 0.5 * np.linalg.norm(np.cross(y2 - y1, y3 - y1))
    f = np.array([b @ y1, b @ y2, b @ y3])
    exact = area * (1 + a @ (y1 + y2 + y3) / 3) + area / 12 * (np.sum(f ** 2) + np.sum(f) ** 2)
    approx = oracle.far_poly(y1, y2, y3, a, b, order, sub)
    assert abs(approx - exact) < 1e-13 * max(1, abs(exact))


def vertex_potential(P, B, Cv):
    """Closed form of the integral of 1/|y - P| over triangle (P, B, C): h ln((|PC| + t_C)/(|PB| + t_B)),
    with h the distance of P to line BC and t the signed position along B->C from the foot."""
    e = (Cv - B) / np.linalg.norm(Cv - B)
    foot = B + ((P - B) @ e) * e
    h = np.linalg.norm(P - foot)
    tB, tC = (B - foot) @ e, (Cv - foot) @ e
    return h * math.log((np.linalg.norm(Cv - P) + tC) / (np.linalg.norm(B - P) + tB))


@pytest.mark.parametrize("tri", [
    (np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1.0, 0])),       # right angle at P
    (np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([1.0, 1.0, 0])),       # right angle at B
    (np.array([0.1, 0.2, 0.3]), np.array([1.0, 0.3, 0.1]), np.array([0.4, 1.1, 0.5])),
    (np.array([0.0, 0, 0]), np.array([0.05, 0.01, 0]), np.array([0.02, 0.04, 0.01])),
])
def test_polar_rule_vertex_potential(oracle, tri):
    P, B, Cv = tri
    exact = vertex_potential(P, B, Cv)
    errs = [abs(oracle.polar_inv_r(P, B, Cv, 5, sub) - exact) / exact for sub in (1, 2, 4, 8)]
    assert errs[1] < 1e-7                        # the shipped (order 5, subdivision 2) rule
    assert errs[3] < 1e-12                       # refinement converges to the closed form
    assert errs[0] >= errs[1] >= errs[2] >= errs[3] or errs[0] < 1e-14
