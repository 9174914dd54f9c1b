"""Row samples of the full-size workloads, assembled by the CPU oracle (test infrastructure).

At the full workload sizes the oracle cannot assemble whole matrices in test time, so it assembles the rows of
one vertex (collocation: that vertex's rows; Galerkin: the test faces around it) at every CQ frequency and takes
them to the time domain with its inverse transform.  The C matrix is added to H_0 and B_angular receives its
B_trans (x) x term, as the oracle's full assembly does (both pinned in tests/test_oracle_*.py).

The long-double runs are too slow for a GPU test at these sizes; tests/golden/make_row_noise.py records, per
time-domain weight, how far the FP64 oracle's rows sit from its long-double rows (the FP64 noise floor that the
R^-n factor of the transform amplifies), and the GPU tests bound the GPU's distance to the FP64 rows by that
profile."""
from __future__ import annotations

import numpy as np

from surfaceonlydeformables_amd import mesh as M

# Full-size Galerkin workloads: the ico-L3 configuration (L = 64) and the ico-L4 one (L = 32).
GALERKIN_CONFIGS = {
    "ico3_galerkin_L64": dict(level=3, radius=0.1, E=1e6, nu=0.35, rho=1100.0, dt=5e-4, L=64),
    "ico4_galerkin_L32": dict(level=4, radius=0.2, E=5e5, nu=0.3, rho=1000.0, dt=1e-3, L=32),
}


def galerkin_case(name):
    c = GALERKIN_CONFIGS[name]
    V, F = M.icosphere(c["level"], c["radius"])
    V = M.center(M.jitter(V, F, 0.01, 0), F)
    _, c1, c2 = M.material_speeds(c["E"], c["nu"], c["rho"])
    R = 1e-10 ** (1 / (2 * c["L"]))
    row = len(V) // 3
    return V, F, dict(c1=c1, c2=c2, rho=c["rho"], nu=c["nu"], dt=c["dt"], R=R, L=c["L"]), row


def oracle_rows(O, V, F, prm, row, galerkin, mode, order=5, subdiv=2, bdf=2, nthreads=0):
    """Time-domain rows 3 row .. 3 row + 2 of G, H (with C), B_trans, B_angular (with B_trans (x) x) and B."""
    L, R = prm["L"], prm["R"]
    nv = len(V)
    p = O.params(prm["c1"], prm["c2"], prm["rho"], prm["nu"], prm["dt"], R, L, bdf=bdf, order=order, subdiv=subdiv,
                 galerkin=galerkin, mode=mode, nthreads=nthreads)
    s = O.frequencies(L, R, prm["dt"], bdf)
    lf = len(s)
    sl = slice(3 * row, 3 * row + 3)
    su = np.zeros((lf, 3, 3 * nv), complex)
    sp = np.zeros_like(su)
    st = np.zeros((lf, 3, 3), complex)
    sa = np.zeros((lf, 3, 9), complex)
    for l in range(lf):
        o = O.laplace(V, F, p, s[l], mask=15, rows=(row, row + 1))
        su[l], sp[l], st[l], sa[l] = o["U"][sl], o["P"][sl], o["Bt"][sl], o["Ba"][sl].reshape(3, 9)
    tr = lambda X: O.transform(X.reshape(lf, -1), L, R).reshape((L,) + X.shape[1:])
    G, H, Bt, Ba = tr(su), tr(sp), tr(st), tr(sa).reshape(L, 3, 3, 3)
    if galerkin:
        H[0] += O.C_galerkin(V, F)[sl]
    else:
        H[0][:, sl] += O.C_collocation(V, F, prm["nu"])[sl]
    Ba = Ba + Bt[..., :, None] * V[row][None, None, None, :]
    eul = np.stack([Ba[..., (j + 2) % 3, (j + 1) % 3] - Ba[..., (j + 1) % 3, (j + 2) % 3] for j in range(3)], -1)
    return {"G": G, "H": H, "B_trans": Bt, "B_angular": Ba, "B": np.concatenate([Bt, eul], -1)}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
