"""Pins of oracle/cqbem_object.py (the reference's ElastodynamicsCQBEMObject step restated in numpy) by closed
forms, on the CPU: with no traction and no history the body stays undeformed and its frame follows the
semi-implicit Euler free fall, v_k = v_0 + k dt g and t_k = t_0 + k dt v_0 + dt^2 g k (k + 1) / 2, while it spins
at its constant angular velocity w: R_k = exp(k dt [w]_x) R_0 (scipy's matrix exponential).  Without the local
frame update the gravity acceleration enters the history instead: acc_k = -R^T g.  The matrices are random
(well-conditioned H_0), which these closed forms do not depend on."""
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
