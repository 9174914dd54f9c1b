"""numpy derivation of what the time stepping reads besides G, H and B (test infrastructure).

Restates ElastodynamicsCQBEMMatrixCollection::init's post-processing (src/elastodynamics_cqbem_matrix_collection.cpp:
174-227) from the assembled matrices and the mesh, with LAPACK solves and least squares: the external acceleration
matrices of the rigid frame, the rotation-removal weights, u_p = H_0^-1 G', H_0^-1 B and the least-squares
inverses of B.  The inertia tensor comes from the exact polyhedral (tetrahedra) formula."""
import numpy as np


def tetra_inertia(V, F, rho):
    """Exact inertia tensor of the closed polyhedron about its centroid (signed tetrahedra from the origin)."""
    vol, first, second = 0.0, np.zeros(3), np.zeros((3, 3))
    for f in F:
        a, b, c = V[f[0]], V[f[1]], V[f[2]]
        det = np.dot(a, np.cross(b, c))
        vol += det / 6
        first += det / 24 * (a + b + c)
        s = a + b + c
        second += det / 120 * (np.outer(a, a) + np.outer(b, b) + np.outer(c, c) + np.outer(s, s))
    cm = first / vol
    C = second - vol * np.outer(cm, cm)
    return rho * (np.trace(C) * np.eye(3) - C), vol


def cross_matrix(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def acceleration_matrices(Vc, F, rho, I=None):
    """translational_acc_mat T and rotational_acc_mat I^-1 R (3 x 3nv): the frame accelerations a traction p
    produces, p integrated with the linear basis (vertex weights A/6) and its moment about the centroid."""
    nv = len(Vc)
    if I is None:
        I, vol = tetra_inertia(Vc, F, rho)
    else:
        vol = tetra_inertia(Vc, F, rho)[1]
    A2 = np.linalg.norm(np.cross(Vc[F[:, 1]] - Vc[F[:, 0]], Vc[F[:, 2]] - Vc[F[:, 0]]), axis=1)
    T = np.zeros((3, 3 * nv))
    Rr = np.zeros((3, 3 * nv))
    for f, tri in enumerate(F):
        s = Vc[tri].sum(0)
        for j in tri:
            T[:, 3 * j:3 * j + 3] += np.eye(3) * A2[f] / 6
            Rr[:, 3 * j:3 * j + 3] += A2[f] / 6 * 0.25 * cross_matrix(s + Vc[j])
    T /= rho * vol
    return T, np.linalg.solve(I, Rr)


def rotation_removal_weights(Vc, F):
    nv = len(Vc)
    C = np.cross(Vc[F[:, 1]] - Vc[F[:, 0]], Vc[F[:, 2]] - Vc[F[:, 0]])
    A2 = np.linalg.norm(C, axis=1)
    Nf = C / A2[:, None]
    w = np.zeros(nv)
    for f, tri in enumerate(F):
        w[tri] += A2[f] / 6 * np.dot(Vc[tri[0]], Nf[f])
    return w


def time_stepping_matrices(Vc, F, G, H, B, rho, original_cm, update_local_frame=True):
    """The dict oracle/cqbem_object.ObjectStep reads, from G, H (H_0 with C, not negated) and B = [B_trans |
    B_euler] (L x 3nv x 6) on the centred mesh Vc."""
    lsq = lambda A, Y: np.linalg.lstsq(A, Y, rcond=None)[0]
    T, Rm = acceleration_matrices(Vc, F, rho)
    H0 = H[0]
    Bt, Be = B[0][:, :3], B[0][:, 3:]
    Gp = G[0] + Bt @ T + Be @ Rm if update_local_frame else G[0]
    return {"V": Vc, "F": np.asarray(F), "original_cm": np.asarray(original_cm, dtype=np.float64),
            "G": G, "H": -H, "B": B, "u_p": np.linalg.solve(H0, Gp),
            "rotation_removal_weights": rotation_removal_weights(Vc, F), "update_local_frame": update_local_frame,
            "translational_acc_mat": T, "rotational_acc_mat": Rm,
            "B_trans_inv_H": lsq(Bt, H0), "B_euler_inv_H": lsq(Be, H0), "B_trans_inv_G_prime": lsq(Bt, Gp),
            "B_euler_inv_G_prime": lsq(Be, Gp), "B_euler_inv_B_trans": lsq(Be, Bt),
            "H_inv_B_trans": np.linalg.solve(H0, Bt), "H_inv_B_euler": np.linalg.solve(H0, Be)}
