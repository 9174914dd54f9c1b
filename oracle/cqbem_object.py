"""numpy restatement of the reference's CQ time step of one deformable body -- test infrastructure only.

Only tests/ import this module; the product path never does.

Follows src/elastodynamics_cqbem_object.cpp (estimate_next_state :25-82, update_V_estimate_matrices
:84-124, confirm_next_state :126-238; with m["compression_ratio"] != 0 the compressed branches, whose
matrices come from oracle/cqbem_wavelet.py) and the PhysicsBaseObject parts it uses
(src/physics_base_object.cpp:9-45 init, :85-95 Rodrigues).  The inputs are the collection's matrices:
G_n, H_n (negated, as the object reads them), B_n, u_p and the 3 x 3nv post-processing products.  The
solves are restated with LAPACK: H_inv as numpy.linalg.solve on -H_0, B_trans_inv / B_euler_inv as
least squares on B_0's halves, and the drift rotation from numpy.linalg.svd (R = V U^T of F = U S V^T).
"""
from __future__ import annotations

from collections import deque

import numpy as np


def rotation_from_vector(w):
    """Rodrigues' formula (physics_base_object.cpp:85-95); zero vector -> identity."""
    w = np.asarray(w, dtype=np.float64)
    t = float(np.sqrt(w @ w))
    n = w / t if t > 0 else w
    K = np.array([[0.0, -n[2], n[1]], [n[2], 0.0, -n[0]], [-n[1], n[0], 0.0]])
    return np.eye(3) + K * np.sin(t) + K @ K * (1.0 - np.cos(t))


def volume_centroid(V, F):
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0
    return (vol[:, None] * (a + b + c) / 4.0).sum(axis=0) / vol.sum()


def _normalized_rows(X):
    n = np.linalg.norm(X, axis=1)
    out = X.copy()
    nz = n > 0
    out[nz] /= n[nz, None]
    return out


class ObjectStep:
    """State and step of ElastodynamicsCQBEMObject for one body."""

    def __init__(self, m, dt, gravity, config=None):
        """m: dict with V (nv x 3, centred), F, G, H (L x n3 x n3, H negated), B (L x n3 x 6), u_p, original_cm,
        translational_acc_mat, rotational_acc_mat, B_trans_inv_H, B_euler_inv_H, B_trans_inv_G_prime,
        B_euler_inv_G_prime, B_euler_inv_B_trans, H_inv_B_trans, H_inv_B_euler, rotation_removal_weights,
        update_local_frame."""
        self.m = m
        self.dt = dt
        self.g = np.asarray(gravity, dtype=np.float64)
        self.L = m["G"].shape[0]
        self.max_time_history = self.L - 1
        n3 = m["G"].shape[1]
        self.n3 = n3
        self.H0 = -m["H"][0]
        B0 = m["B"][0]
        self.Bt0, self.Be0 = B0[:, :3], B0[:, 3:]
        config = config or {}
        self.translational_vel = np.array(config.get("translational_vel", [0.0, 0.0, 0.0]), dtype=np.float64)
        self.rotational_vel = np.array(config.get("rotational_vel", [0.0, 0.0, 0.0]), dtype=np.float64)
        self.rotation = (np.array(config["rotation"], dtype=np.float64).reshape(3, 3) if "rotation" in config
                         else np.eye(3))
        self.translation = np.zeros(3)
        if "translation" in config:
            self.translation = np.array(config["translation"], dtype=np.float64) + self.rotation @ m["original_cm"]
        self.use_gravity = bool(config.get("use_gravity", True))
        self.is_fixed = bool(config.get("is_fixed", False))
        self.V = m["V"]
        self.V_local = self.V.copy()
        self.V_global = self.V_local @ self.rotation.T + self.translation
        self.hist_acc, self.hist_u, self.hist_p = deque(), deque(), deque()
        self.has_constraints = False
        self.ratio = m.get("compression_ratio", 0.0)
        if self.ratio:
            import cqbem_wavelet as W
            self.W = W
            keep = int(n3 * n3 * self.ratio)
            self.Gc = [None] + [W.compress(m["G"][i], keep) for i in range(1, self.L)]
            self.Hc = [None] + [W.compress(m["H"][i], keep) for i in range(1, self.L)]
            self.u_p_c = W.compress(m["u_p"], keep)
            self.H_inv_c = W.compress(np.linalg.inv(self.H0), keep)

    def estimate_next_state(self):
        m = self.m
        self.has_constraints = False
        b = np.zeros(self.n3)
        nh = len(self.hist_acc)
        if self.ratio:
            bc = 0.0
            for i in range(min(nh, self.max_time_history)):
                b += m["B"][i + 1] @ self.hist_acc[nh - i - 1]
                bc = bc + self.Hc[i + 1] @ self.hist_u[nh - i - 1] + self.Gc[i + 1] @ self.hist_p[nh - i - 1]
            if nh:
                b += self.W.reconstruct(bc, self.n3)
        else:
            for i in range(min(nh, self.max_time_history)):
                b += m["B"][i + 1] @ self.hist_acc[nh - i - 1]
                b += m["H"][i + 1] @ self.hist_u[nh - i - 1]
                b += m["G"][i + 1] @ self.hist_p[nh - i - 1]
        if not m["update_local_frame"] and self.use_gravity:
            b -= self.Bt0 @ (self.rotation.T @ self.g)
        self.b = b
        self.p = np.zeros(self.n3)
        if self.ratio:
            self.u_unconstrained = self.W.product(self.H_inv_c, b)
        else:
            self.u_unconstrained = np.linalg.solve(self.H0, b)
        u3 = self.u_unconstrained.reshape(-1, 3)
        if self.is_fixed:
            self.V_estimate_global = (self.V + u3) @ self.rotation.T + self.translation
        else:
            R_est = rotation_from_vector(self.dt * self.rotational_vel) @ self.rotation
            t_est = self.translation + self.dt * self.translational_vel
            if self.use_gravity:
                t_est = t_est + self.dt * self.dt * self.g
            self.V_estimate_global = (self.V + u3) @ R_est.T + t_est

    def update_V_estimate_matrices(self):
        m = self.m
        nv = self.V.shape[0]
        self.rotation_estimate = (self.rotation if self.is_fixed
                                  else rotation_from_vector(self.dt * self.rotational_vel) @ self.rotation)
        self.rot_trans_acc = np.zeros((3 * nv, 3))
        self.rot_rot_acc = np.zeros((3 * nv, 3))
        if not self.is_fixed and m["update_local_frame"]:
            dd = self.dt * self.dt * self.rotation
            for v in range(nv):
                x = self.V_local[v]
                X = np.array([[0.0, -x[2], x[1]], [x[2], 0.0, -x[0]], [-x[1], x[0], 0.0]])
                self.rot_trans_acc[3 * v:3 * v + 3] = dd
                self.rot_rot_acc[3 * v:3 * v + 3] = -dd @ X
        Vc = self.V_estimate_global.copy()
        if not self.is_fixed and not m["update_local_frame"]:
            Vc -= self.dt * self.translational_vel
            if self.use_gravity:
                Vc -= self.dt * self.dt * self.g
        self.V_const = Vc.reshape(-1)

    def confirm_next_state(self):
        m, dt = self.m, self.dt
        if self.has_constraints:
            u_p = self.W.product(self.u_p_c, self.p) if self.ratio else m["u_p"] @ self.p
            u = u_p + self.u_unconstrained
            Vg = np.kron(np.eye(self.V.shape[0]), self.rotation_estimate) @ u_p + self.V_const
            if not self.is_fixed and m["update_local_frame"]:
                Vg += (self.rot_trans_acc @ (m["translational_acc_mat"] @ self.p)
                       + self.rot_rot_acc @ (m["rotational_acc_mat"] @ self.p))
            self.V_global = Vg.reshape(-1, 3)
        else:
            u = self.u_unconstrained.copy()
            u_p = np.zeros(self.n3)
        self.V_local = self.V + u.reshape(-1, 3)
        tacc, racc = np.zeros(3), np.zeros(3)
        if m["update_local_frame"]:
            old_rotation = self.rotation.copy()
            tacc = m["translational_acc_mat"] @ self.p
            racc = m["rotational_acc_mat"] @ self.p
            if not self.is_fixed:
                self.translational_vel = self.translational_vel + dt * self.rotation @ tacc
                if self.use_gravity:
                    self.translational_vel = self.translational_vel + dt * self.g
                self.rotational_vel = self.rotational_vel + dt * self.rotation @ racc
                self.translation = self.translation + dt * self.translational_vel
                self.rotation = rotation_from_vector(dt * self.rotational_vel) @ self.rotation
            if not self.has_constraints:
                self.V_global = self.V_local @ self.rotation.T + self.translation
            # drift elimination (object.cpp:171-211)
            cm = volume_centroid(self.V_local, m["F"])
            V_vec = _normalized_rows(self.V_local - cm)
            init_V_vec = _normalized_rows(self.V)
            Fm = init_V_vec.T @ (m["rotation_removal_weights"][:, None] * V_vec)
            U, _, Vh = np.linalg.svd(Fm)
            R = Vh.T @ U.T
            du = ((self.V_local - cm) - self.V).reshape(-1)
            tad = (m["B_trans_inv_H"] @ du - m["B_trans_inv_G_prime"] @ self.p
                   - np.linalg.lstsq(self.Bt0, self.b, rcond=None)[0])
            du = ((self.V_local - cm) @ R - self.V).reshape(-1)
            aad = (m["B_euler_inv_H"] @ du - m["B_euler_inv_G_prime"] @ self.p
                   - np.linalg.lstsq(self.Be0, self.b, rcond=None)[0] - m["B_euler_inv_B_trans"] @ tad)
            u = u_p + self.u_unconstrained + m["H_inv_B_trans"] @ tad + m["H_inv_B_euler"] @ aad
            tacc, racc = tacc + tad, racc + aad
            if not self.is_fixed:
                self.translational_vel = self.translational_vel + dt * old_rotation @ tad
                self.rotational_vel = self.rotational_vel + dt * old_rotation @ aad
                self.translation = self.translation + dt * dt * old_rotation @ tad
                self.rotation = rotation_from_vector(dt * self.rotational_vel) @ old_rotation
            self.V_local = self.V + u.reshape(-1, 3)
        else:
            if self.use_gravity:
                tacc = -self.rotation.T @ self.g
            if not self.has_constraints:
                self.V_global = self.V_local @ self.rotation.T + self.translation
        if self.max_time_history > 0:
            if len(self.hist_acc) >= self.max_time_history:
                self.hist_acc.popleft(), self.hist_u.popleft(), self.hist_p.popleft()
            self.hist_acc.append(np.concatenate([tacc, racc]))
            if self.ratio:  # compressed histories (object.cpp:216-222)
                self.hist_u.append(self.W.decompose_vec(u))
                self.hist_p.append(self.W.decompose_vec(self.p))
            else:
                self.hist_u.append(u)
                self.hist_p.append(self.p.copy())
