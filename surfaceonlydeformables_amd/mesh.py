"""Surface meshes and per-face geometry for the CQBEM precompute.

Synthetic closed triangle meshes (icosphere, subdivided cube) with seeded
jitter, OBJ I/O, face normals and the volume-centroid centering that the
reference applies before any assembly
(src/physics_base_matrix_collection.cpp:40-47, igl::per_face_normals,
igl::centroid).
"""
from __future__ import annotations

import numpy as np


def icosphere(level: int, radius: float = 1.0):
    """Icosphere of subdivision `level`: 10*4^level+2 vertices, 20*4^level faces, CCW outward."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
         (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
         (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
         (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def mid(a: int, b: int) -> int:
            key = (min(a, b), max(a, b))
            if key not in cache:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                cache[key] = len(V) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = nf
    return np.asarray(V) * radius, np.asarray(F, dtype=np.int32)


def cube(k: int, half_size: float = 0.5):
    """Cube surface with k x k quads per face split along one fixed diagonal:
    6k^2+2 vertices, 12k^2 faces, CCW outward."""
    index: dict[tuple[int, int, int], int] = {}
    V: list[tuple[float, float, float]] = []

    def vid(p):
        if p not in index:
            index[p] = len(V)
            V.append(tuple(half_size * (2.0 * c / k - 1.0) for c in p))
        return index[p]

    F = []
    # (normal axis, side, u axis, v axis) with u x v = outward normal
    for axis in range(3):
        u_ax, v_ax = (axis + 1) % 3, (axis + 2) % 3
        for side in (0, k):
            for a in range(k):
                for b in range(k):
                    def P(du, dv):
                        p = [0, 0, 0]
                        p[axis] = side
                        p[u_ax] = a + du
                        p[v_ax] = b + dv
                        return vid(tuple(p))
                    p00, p10, p11, p01 = P(0, 0), P(1, 0), P(1, 1), P(0, 1)
                    if side == k:
                        F += [(p00, p10, p11), (p00, p11, p01)]
                    else:
                        F += [(p00, p11, p10), (p00, p01, p11)]
    return np.asarray(V, dtype=np.float64), np.asarray(F, dtype=np.int32)


def mean_edge_length(V, F) -> float:
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    return float(np.linalg.norm(V[e[:, 0]] - V[e[:, 1]], axis=1).mean())


def jitter(V, F, fraction: float, seed: int):
    """Per-coordinate uniform noise in +-fraction*(mean edge length), numpy PCG64(seed)."""
    amp = fraction * mean_edge_length(V, F)
    rng = np.random.default_rng(seed)
    return V + rng.uniform(-amp, amp, size=V.shape)


def face_normals(V, F):
    """Unit face normals (v1-v0) x (v2-v0) / |.| (igl::per_face_normals)."""
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def volume_centroid(V, F):
    """Volume and volume centroid of a closed, outward-oriented mesh (igl::centroid),
    from the signed tetrahedra (origin, v0, v1, v2)."""
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0
    total = vol.sum()
    cen = (vol[:, None] * (a + b + c) / 4.0).sum(axis=0) / total
    return float(total), cen


def center(V, F):
    """Translate so that the volume centroid is at the origin (physics_base_matrix_collection.cpp:41-45)."""
    _, cm = volume_centroid(V, F)
    return V - cm


def read_obj(path: str):
    V, F = [], []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                V.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f":
                idx = [int(p.split("/")[0]) for p in parts[1:]]
                idx = [i - 1 if i > 0 else len(V) + i for i in idx]
                for k in range(1, len(idx) - 1):
                    F.append([idx[0], idx[k], idx[k + 1]])
    return np.asarray(V, dtype=np.float64), np.asarray(F, dtype=np.int32)


def write_obj(path: str, V, F) -> None:
    with open(path, "w") as fh:
        for v in V:
            fh.write("v %.17g %.17g %.17g\n" % tuple(v))
        for f in F:
            fh.write("f %d %d %d\n" % (f[0] + 1, f[1] + 1, f[2] + 1))


def material_speeds(E: float, nu: float, rho: float):
    """(mu, c1, c2) from Young's modulus (elastodynamics_cqbem_matrix_collection.cpp:59-61 with mu = E/(2(1+nu)))."""
    mu = E / (2.0 * (1.0 + nu))
    lam = 2.0 * nu * mu / (1.0 - 2.0 * nu)
    return mu, float(np.sqrt((lam + 2.0 * mu) / rho)), float(np.sqrt(mu / rho))


def reference_history(V, c2: float, dt: float):
    """max_time_history = floor(max pairwise distance / (c2 dt) + 2) and L = that + 1
    (elastodynamics_cqbem_matrix_collection.cpp:63-69)."""
    from scipy.spatial import ConvexHull  # max distance is attained on the hull
    hull = V[ConvexHull(V).vertices]
    d = np.sqrt(((hull[:, None, :] - hull[None, :, :]) ** 2).sum(-1)).max()
    mth = int(np.floor(d / (c2 * dt) + 2))
    return mth, mth + 1


def contour_radius(max_time_history: int) -> float:
    """R = (1e-10)^(1/(2 max_time_history)) (src/elastodynamics_cqbem_G.cu:19)."""
    return 1e-10 ** (1.0 / (2.0 * max_time_history))
