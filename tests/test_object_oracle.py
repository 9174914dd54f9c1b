"""Pins of oracle/cqbem_object.py (the reference's ElastodynamicsCQBEMObject step restated in numpy) by closed
forms, on the CPU: with no traction and no history the body stays undeformed and its frame follows the
semi-implicit Euler free fall, v_k = v_0 + k dt g and t_k = t_0 + k dt v_0 + dt^2 g k (k + 1) / 2, while it spins
at its constant angular velocity w: R_k = exp(k dt [w]_x) R_0 (scipy's matrix exponential).  Without the local
frame update the gravity acceleration enters the history instead: acc_k = -R^T g.  The matrices are random
(well-conditioned H_0), which these closed forms do not depend on.

The deformation branch runs on synthetic CQ weights whose discrete solutions are textbook schemes: backward Euler for
u' + k u = A p, the truncated convolution with the BDF1 weights of 1/s^2 (np.convolve), the untruncated second
difference M u'' = N p, and the gravity history convolved with the translational columns of B.  The drift
elimination is checked on a rigid displacement, which its centroid and polar fit must recognise exactly, and the
volume centroid on a pyramid.  Each of these fails for a reversed or shifted history, a sign flip of a history term,
a transposed rotation, the angular columns in place of the translational ones, or a dropped u_p p term."""
import zlib

import numpy as np
from scipy.linalg import expm

from cqbem_object import ObjectStep, rotation_from_vector
from surfaceonlydeformables_amd import mesh as M


def matrices(update_local_frame=True, L=5, seed=0):
    V, F = M.icosphere(0, 0.2)
    V = M.center(V, F)
    nv = len(V)
    n3 = 3 * nv
    rng = np.random.default_rng(seed)
    m = {"G": rng.standard_normal((L, n3, n3)) * 1e-3, "H": rng.standard_normal((L, n3, n3)) * 1e-3,
         "B": rng.standard_normal((L, n3, 6)), "V": V, "F": F, "original_cm": np.array([0.1, -0.2, 0.3]),
         "u_p": rng.standard_normal((n3, n3)) * 1e-3, "rotation_removal_weights": np.ones(nv),
         "update_local_frame": update_local_frame}
    m["H"][0] = -(np.eye(n3) + 1e-3 * rng.standard_normal((n3, n3)))
    for k in ("translational_acc_mat", "rotational_acc_mat", "B_trans_inv_H", "B_euler_inv_H", "B_trans_inv_G_prime",
              "B_euler_inv_G_prime"):
        m[k] = rng.standard_normal((3, n3)) * 1e-3
    m["B_euler_inv_B_trans"] = rng.standard_normal((3, 3)) * 1e-3
    m["H_inv_B_trans"] = rng.standard_normal((n3, 3)) * 1e-3
    m["H_inv_B_euler"] = rng.standard_normal((n3, 3)) * 1e-3
    return m


def test_rodrigues_matches_expm():
    for w in ([0.0, 0.0, 0.0], [0.3, -0.2, 0.9], [1e-9, 0.0, 2e-9], [3.0, 1.0, -2.0]):
        w = np.array(w)
        W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
        assert np.allclose(rotation_from_vector(w), expm(W), rtol=0, atol=1e-13)


def test_free_fall_and_spin_closed_form():
    m = matrices()
    dt, g = 2e-3, np.array([0.0, -9.8, 0.0])
    v0, t0, w = np.array([0.3, 0.5, -0.2]), np.array([0.0, 1.0, 0.0]), np.array([0.5, 1.0, -0.3])
    R0 = expm(np.array([[0, -0.1, 0.2], [0.1, 0, -0.3], [-0.2, 0.3, 0]]))
    o = ObjectStep(m, dt, g, {"translational_vel": list(v0), "translation": list(t0), "rotational_vel": list(w),
                              "rotation": list(R0.ravel())})
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    t_init = t0 + R0 @ m["original_cm"]
    for k in range(1, 31):
        o.estimate_next_state()
        o.confirm_next_state()
        assert np.allclose(o.translational_vel, v0 + k * dt * g, rtol=0, atol=1e-12)
        assert np.allclose(o.translation, t_init + k * dt * v0 + dt * dt * g * k * (k + 1) / 2, rtol=0, atol=1e-12)
        assert np.allclose(o.rotation, expm(k * dt * W) @ R0, rtol=0, atol=1e-12)
        assert np.abs(o.V_local - m["V"]).max() < 1e-14
        assert np.allclose(o.V_global, m["V"] @ o.rotation.T + o.translation, rtol=0, atol=1e-12)


def test_no_local_frame_gravity_enters_the_history():
    m = matrices(update_local_frame=False)
    dt, g = 2e-3, np.array([0.0, -9.8, 0.0])
    R0 = expm(np.array([[0, -0.4, 0.1], [0.4, 0, -0.2], [-0.1, 0.2, 0]]))
    o = ObjectStep(m, dt, g, {"rotation": list(R0.ravel())})
    o.estimate_next_state()
    # the first step's b is the gravity term alone: -B_0[:, :3] R^T g
    assert np.allclose(o.b, -m["B"][0][:, :3] @ (R0.T @ g), rtol=1e-14, atol=0)
    o.confirm_next_state()
    assert np.allclose(o.hist_acc[-1], np.concatenate([-R0.T @ g, np.zeros(3)]), rtol=0, atol=1e-15)
    # the frame does not move without the local frame update
    assert np.array_equal(o.rotation, R0) and np.array_equal(o.translation, np.zeros(3))


# --- the deformation branch on synthetic CQ weights whose discrete solutions have closed forms ---------------
# The object solves H_0 u_n = G_0 p_n + sum_{i>=1} (G_i p_{n-i} - H_i u_{n-i} + B_i acc_{n-i}) over a history of
# L - 1 steps, with G_0 p_n entering as u_p p_n (u_p = H_0^-1 G_0) and the collection's H_i stored negated.  With
# backward-Euler (BDF1) weights of simple transfer functions these recurrences are textbook schemes.

def synthetic(L, H, G, u_p, B=None, V_F=None):
    """Matrices of a body whose weights are the given lists (H_i as the operator, stored negated)."""
    V, F = V_F if V_F is not None else M.icosphere(0, 0.2)
    V = M.center(V, F)
    n3 = 3 * len(V)
    return {"G": np.array(G), "H": -np.array(H), "B": np.zeros((L, n3, 6)) if B is None else np.array(B),
            "V": V, "F": F, "original_cm": np.zeros(3), "u_p": u_p, "rotation_removal_weights": np.ones(len(V)),
            "update_local_frame": False}


def drive(o, ps):
    """Steps the body with the nodal tractions ps[k] as its constraint forces; returns the displacements."""
    out = []
    for p in ps:
        o.estimate_next_state()
        o.update_V_estimate_matrices()
        o.p = p.copy()
        o.has_constraints = True
        o.confirm_next_state()
        out.append((o.V_local - o.V).reshape(-1))
    return np.array(out)


def test_backward_euler_of_a_damped_first_order_system():
    """H(s) = s + k, G(s) = A: the CQ step is backward Euler for u' + k u = A p (A non-symmetric)."""
    dt, k, L = 1e-2, 3.0, 4
    rng = np.random.default_rng(1)
    n3 = 36
    A, I, Z = rng.standard_normal((n3, n3)), np.eye(n3), np.zeros((n3, n3))
    m = synthetic(L, [(1 / dt + k) * I, -I / dt, Z, Z], [A, Z, Z, Z], A / (1 / dt + k))
    R0 = expm(np.array([[0, -0.4, 0.1], [0.4, 0, -0.2], [-0.1, 0.2, 0]]))
    o = ObjectStep(m, dt, [0.0, -9.8, 0.0], {"use_gravity": False, "rotation": list(R0.ravel())})
    ps = rng.standard_normal((12, n3)) * 1e-3
    u = drive(o, ps)
    ref = np.zeros(n3)
    for n in range(len(ps)):
        ref = (ref / dt + A @ ps[n]) / (1 / dt + k)
        assert np.allclose(u[n], ref, rtol=0, atol=1e-14 * max(1.0, np.abs(ref).max()) + 1e-17)
    assert np.allclose(o.V_global, (m["V"] + u[-1].reshape(-1, 3)) @ R0.T, rtol=0, atol=1e-15)


def test_double_integration_history_order_and_truncation():
    """H = I, G(s) = A / s^2: BDF1 weights dt^2 (j + 1); the step is the discrete convolution of the tractions with
    those weights cut at L terms (np.convolve), so a reversed, shifted or longer history fails."""
    dt, L = 1e-2, 6
    rng = np.random.default_rng(2)
    n3 = 36
    A, I, Z = rng.standard_normal((n3, n3)), np.eye(n3), np.zeros((n3, n3))
    w = dt * dt * np.arange(1, L + 1)
    m = synthetic(L, [I] + [Z] * (L - 1), [wj * A for wj in w], w[0] * A)
    o = ObjectStep(m, dt, [0.0, 0.0, 0.0], {"use_gravity": False})
    ps = rng.standard_normal((15, n3))
    u = drive(o, ps)
    q = np.stack([np.convolve(ps[:, c], w)[:len(ps)] for c in range(n3)], axis=1)
    assert np.allclose(u, q @ A.T, rtol=0, atol=1e-13 * np.abs(u).max())
    # constant tractions: the double integral dt^2 (n+1)(n+2)/2 p while the history is not yet cut
    o2 = ObjectStep(m, dt, [0.0, 0.0, 0.0], {"use_gravity": False})
    p = rng.standard_normal(n3)
    u2 = drive(o2, [p] * (L - 1))
    for n in range(L - 1):
        ref = dt * dt * (n + 1) * (n + 2) / 2 * (A @ p)
        assert np.allclose(u2[n], ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_second_difference_history_is_not_truncated():
    """M u'' = N p with BDF1 weights of s^2 (1, -2, 1) / dt^2 on H: the H history carries the whole past, so
    u_n = dt^2 M^-1 N sum_{j<=n} (j+1) p_{n-j} for every n, long after the L - 1 = 4 step window."""
    dt, L = 1e-2, 5
    rng = np.random.default_rng(3)
    n3 = 36
    Mm = np.eye(n3) + 0.3 * rng.standard_normal((n3, n3)) / np.sqrt(n3)
    N, Z = rng.standard_normal((n3, n3)), np.zeros((n3, n3))
    m = synthetic(L, [Mm / dt**2, -2 * Mm / dt**2, Mm / dt**2, Z, Z], [N, Z, Z, Z, Z],
                  np.linalg.solve(Mm, N) * dt**2)
    o = ObjectStep(m, dt, [0.0, 0.0, 0.0], {"use_gravity": False})
    ps = rng.standard_normal((20, n3))
    u = drive(o, ps)
    w = dt * dt * np.arange(1, len(ps) + 1)
    q = np.stack([np.convolve(ps[:, c], w)[:len(ps)] for c in range(n3)], axis=1)
    ref = q @ np.linalg.solve(Mm, N).T
    assert np.allclose(u, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_gravity_history_is_a_convolution_of_the_translational_columns():
    """Without the local frame update, gravity is the constant acceleration -R^T g in the history and
    -B_trans,0 R^T g now; with B_trans,j = dt^2 (j+1) C the displacement is -dt^2 (m+1)(m+2)/2 C R^T g,
    m = min(n, L - 1).  The angular columns (random D) must not enter."""
    dt, L = 1e-2, 5
    rng = np.random.default_rng(4)
    n3 = 36
    C, D, I, Z = rng.standard_normal((n3, 3)), rng.standard_normal((n3, 3)), np.eye(n3), np.zeros((n3, n3))
    B = [dt * dt * (j + 1) * np.hstack([C, D]) for j in range(L)]
    m = synthetic(L, [I] + [Z] * (L - 1), [Z] * L, Z, B=B)
    g = np.array([0.5, -9.8, 1.0])
    R0 = expm(np.array([[0, -0.4, 0.1], [0.4, 0, -0.2], [-0.1, 0.2, 0]]))
    o = ObjectStep(m, dt, g, {"rotation": list(R0.ravel())})
    for n in range(10):
        o.estimate_next_state()
        o.confirm_next_state()
        k = min(n, L - 1)
        ref = -dt * dt * (k + 1) * (k + 2) / 2 * (C @ (R0.T @ g))
        assert np.allclose((o.V_local - o.V).reshape(-1), ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def skew(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def test_drift_elimination_of_a_rigid_displacement():
    """A displacement that moves the body rigidly (V -> V Q^T + c) is recognised exactly: the volume centroid is c
    and the weighted polar fit of the vertex directions is Q, so the rotated-back shape is V again and the angular
    correction gets no B_euler_inv_H term, whatever that matrix is; the translational one sees V Q^T - V."""
    V, F = M.icosphere(1, 0.2)
    V = M.center(M.jitter(V, F, 0.05, 7), F)
    nv, n3 = len(V), 3 * len(V)
    rng = np.random.default_rng(5)
    L, I, Z = 3, np.eye(n3), np.zeros((n3, n3))
    Q = expm(skew([0.3, -0.4, 0.2]))
    c = np.array([0.01, -0.02, 0.015])
    p = (V @ Q.T + c - V).reshape(-1)

    def run(Zeuler, shift):
        m = synthetic(L, [I, Z, Z], [Z] * L, I, B=rng.standard_normal((L, n3, 6)), V_F=(V, F))
        m["update_local_frame"] = True
        m["rotation_removal_weights"] = 0.5 + rng.random(nv)
        m["translational_acc_mat"] = np.zeros((3, n3))
        m["rotational_acc_mat"] = np.zeros((3, n3))
        for key, shape in (("B_trans_inv_H", (3, n3)), ("B_trans_inv_G_prime", (3, n3)), ("B_euler_inv_G_prime", (3, n3)),
                           ("B_euler_inv_B_trans", (3, 3)), ("H_inv_B_trans", (n3, 3)), ("H_inv_B_euler", (n3, 3))):
            m[key] = np.random.default_rng(zlib.crc32(key.encode())).standard_normal(shape)
        m["B_euler_inv_H"] = Zeuler
        o = ObjectStep(m, 1e-3, [0.0, 0.0, 0.0], {"use_gravity": False})
        pp = p + np.tile(shift, nv)
        drive(o, [pp])
        return m, o, pp

    m, o, pp = run(rng.standard_normal((3, n3)), np.zeros(3))
    tad = m["B_trans_inv_H"] @ (V @ Q.T - V).reshape(-1) - m["B_trans_inv_G_prime"] @ pp
    aad = -m["B_euler_inv_G_prime"] @ pp - m["B_euler_inv_B_trans"] @ tad
    assert np.allclose(o.hist_acc[-1], np.concatenate([tad, aad]), rtol=0, atol=1e-12 * np.abs(tad).max())
    corr = m["H_inv_B_trans"] @ tad + m["H_inv_B_euler"] @ aad
    assert np.allclose(o.V_local, V @ Q.T + c + corr.reshape(-1, 3), rtol=0, atol=1e-12 * np.abs(corr).max())
    # another B_euler_inv_H changes nothing; another rigid shift changes tad only through the G' p term
    # (equal to rounding: B_euler_inv_H multiplies the rotated-back shape minus V, zero up to a few ulps of V)
    Z2 = rng.standard_normal((3, n3)) * 1e3
    m2, o2, pp2 = run(Z2, np.zeros(3))
    tol = 1e-12 * max(np.abs(tad).max(), np.abs(Z2).max() * np.abs(V).max())
    assert np.allclose(o2.hist_acc[-1], o.hist_acc[-1], rtol=0, atol=tol)
    s = np.array([0.02, 0.01, -0.03])
    m3, o3, pp3 = run(rng.standard_normal((3, n3)), s)
    tad3 = m3["B_trans_inv_H"] @ (V @ Q.T - V).reshape(-1) - m3["B_trans_inv_G_prime"] @ pp3
    assert np.allclose(o3.hist_acc[-1][:3], tad3, rtol=0, atol=1e-12 * np.abs(tad3).max())


def test_volume_centroid_of_a_pyramid():
    """A square pyramid of height h has its volume centroid h/4 above the base (the vertex mean is at h/5); the
    tetrahedra from the origin give it wherever the solid lies."""
    from cqbem_object import volume_centroid
    h, off = 2.0, np.array([3.0, -1.0, 0.5])
    V = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0], [0, 0, h]], dtype=np.float64)
    F = np.array([[0, 2, 1], [0, 3, 2], [0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]])
    inside = np.array([0.0, 0.0, h / 4])
    for f in F:  # outward orientation checked, not assumed
        a, b, cc = V[f]
        assert np.cross(b - a, cc - a) @ ((a + b + cc) / 3 - inside) > 0
    assert np.allclose(volume_centroid(V + off, F), off + [0, 0, h / 4], rtol=0, atol=1e-14)
