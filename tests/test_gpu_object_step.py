"""CQ time stepping of one deformable body (ElastodynamicsCQBEMObject, src/elastodynamics_cqbem_object.cpp)
on the GPU-resident matrices of the C++ facade, through collection_driver --simulate.

Parity: the same trajectory from oracle/cqbem_object.py, a numpy restatement of the reference's step
(estimate_next_state, update_V_estimate_matrices, confirm_next_state with the drift elimination), fed the
matrices the facade dumped.  The product runs the history sums and u_p p on the GPU, the H_0 solve through
rocSOLVER and the drift rotation by a polar iteration; the oracle uses LAPACK solves and an SVD.  Bound:
relative error 1e-9 per quantity and step (the solves' rounding, amplified by cond(H_0) ~ 1e2..1e3).

With simulation.compression_ratio the step runs on wavelet-compressed G_n, H_n, u_p and H_0^-1 (the reference's
compressed branches); the oracle compresses the same dumped matrices with oracle/cqbem_wavelet.py (its H_0^-1
from LAPACK), and the bound is the same.

Closed form: with no traction the body stays undeformed, its frame follows the semi-implicit Euler
free fall, v_k = v_0 + k dt g, t_k = t_0 + k dt v_0 + dt^2 g k (k + 1) / 2, and spins at its constant
angular velocity w: R_k = exp(k dt [w]_x) (scipy's matrix exponential, independent of the Rodrigues
formula both implementations use)."""
import json
import os
import subprocess

import numpy as np
import pytest

from surfaceonlydeformables_amd import mesh as M

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "surfaceonlydeformables_amd", "collection_driver")
DT = 2e-3
GRAVITY = [0.0, -9.8, 0.0]


def simulate(tmp_path, V, F, obj, tractions, **sim):
    obj_path = tmp_path / "mesh.obj"
    M.write_obj(str(obj_path), V, F)
    cfg = {"enable_cuda": True, "mesh": {"file": str(obj_path)},
           "material": {"rho": 1000.0, "mu": 3.0e4, "nu": 0.3},
           "simulation": {"dt": DT, "gaussian_quadrature_order": 5, "quadrature_subdivision": 2,
                          "multistep_method": "BDF2", "use_galerkin": False, **sim},
           "object": obj, "gravity": GRAVITY}
    path = tmp_path / "scene.json"
    path.write_text(json.dumps(cfg))
    tr = tmp_path / "tractions.bin"
    np.ascontiguousarray(tractions, dtype=np.float64).tofile(tr)
    print(f"object step nv={len(V)} steps={len(tractions)} sim={sim} obj={obj}", flush=True)
    r = subprocess.run([DRIVER, "--simulate", str(path), str(len(tractions)), str(tr), str(tmp_path / "out")],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    nv, L = len(V), info["L"]
    n3 = 3 * nv
    post = lambda name, shape: np.fromfile(tmp_path / f"out_post_{name}.bin").reshape(shape)
    m = {"G": np.fromfile(tmp_path / "out_G.bin").reshape(L, n3, n3),
         "H": np.fromfile(tmp_path / "out_H.bin").reshape(L, n3, n3),
         "B": np.fromfile(tmp_path / "out_B.bin").reshape(L, n3, 6),
         "V": post("V", (nv, 3)), "F": np.asarray(F), "original_cm": post("original_cm", (3,)),
         "u_p": post("u_p", (n3, n3)), "rotation_removal_weights": post("rotation_removal_weights", (nv,)),
         "update_local_frame": sim.get("update_local_frame", True),
         "compression_ratio": sim.get("compression_ratio", 0.0)}
    for k in ("translational_acc_mat", "rotational_acc_mat", "B_trans_inv_H", "B_euler_inv_H",
              "B_trans_inv_G_prime", "B_euler_inv_G_prime"):
        m[k] = post(k, (3, n3))
    m["B_euler_inv_B_trans"] = post("B_euler_inv_B_trans", (3, 3))
    m["H_inv_B_trans"] = post("H_inv_B_trans", (n3, 3))
    m["H_inv_B_euler"] = post("H_inv_B_euler", (n3, 3))
    rec = 24 + 6 * nv
    traj = np.fromfile(tmp_path / "out_traj.bin").reshape(len(tractions), rec)
    return m, traj, info


def split(row, nv):
    return {"translation": row[0:3], "rotation": row[3:12].reshape(3, 3), "translational_vel": row[12:15],
            "rotational_vel": row[15:18], "acc": row[18:24], "V_local": row[24:24 + 3 * nv].reshape(nv, 3),
            "V_global": row[24 + 3 * nv:].reshape(nv, 3)}


def pressure_pulse(V, F, steps, first, last, P):
    """Inward normal traction P on the vertices with x > 0.4 max x, during steps first..last."""
    nv = len(V)
    Nf = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    Nv = np.zeros((nv, 3))
    for f, tri in enumerate(F):
        Nv[tri] += Nf[f]
    Nv /= np.linalg.norm(Nv, axis=1)[:, None]
    Vc = M.center(V, F)
    mask = Vc[:, 0] > 0.4 * Vc[:, 0].max()
    p = np.zeros((steps, nv, 3))
    p[first:last + 1, mask] = -P * Nv[mask]
    return p.reshape(steps, 3 * nv)


def compare(m, traj, obj, nv, tractions, tol=1e-9):
    from cqbem_object import ObjectStep
    o = ObjectStep(m, DT, GRAVITY, obj)
    worst = 0.0
    for k in range(len(tractions)):
        o.estimate_next_state()
        if np.any(tractions[k] != 0):
            o.update_V_estimate_matrices()
            o.p = tractions[k].copy()
            o.has_constraints = True
        o.confirm_next_state()
        got = split(traj[k], nv)
        want = {"translation": o.translation, "rotation": o.rotation, "translational_vel": o.translational_vel,
                "rotational_vel": o.rotational_vel, "acc": o.hist_acc[-1], "V_local": o.V_local,
                "V_global": o.V_global}
        for key, w in want.items():
            # relative to the quantity, with floors at physical scales: 1 mm, 1 mm/s, g, 1e-3 rad/s, and for
            # V_local the deformation itself down to 1e-5 of the mesh (before the pulse the accelerations and
            # deformation are the drift elimination's rounding noise: eps x |B^-1 H| x |V| ~ 1e-11 m/s^2)
            floors = {"translation": 1e-3, "translational_vel": 1e-3, "acc": 9.8, "rotational_vel": 1e-3}
            scale = max(np.linalg.norm(w), floors.get(key, 1e-6))
            if key == "V_local":
                d, scale = np.linalg.norm(got[key] - w), max(np.linalg.norm(w - m["V"]), 1e-5 * np.linalg.norm(m["V"]))
            else:
                d = np.linalg.norm(got[key] - w)
            worst = max(worst, d / scale)
            assert d <= tol * scale, (k, key, d, scale)
    return worst, o


@pytest.fixture(scope="module")
def mesh():
    V, F = M.icosphere(1, 0.1)
    return M.jitter(V, F, 0.01, 11), F


def test_free_fall_closed_form(tmp_path, mesh):
    V, F = mesh
    steps = 30
    from scipy.linalg import expm
    v0, t0, w = np.array([0.3, 0.5, -0.2]), np.array([0.0, 1.0, 0.0]), np.array([0.5, 1.0, -0.3])
    obj = {"translational_vel": list(v0), "translation": list(t0), "rotational_vel": list(w)}
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    m, traj, info = simulate(tmp_path, V, F, obj, np.zeros((steps, 3 * len(V))))
    nv = len(V)
    g = np.array(GRAVITY)
    cm0 = m["original_cm"]
    for k in range(steps):
        s = split(traj[k], nv)
        n = k + 1
        assert np.allclose(s["translational_vel"], v0 + n * DT * g, rtol=0, atol=1e-12)
        assert np.allclose(s["translation"], t0 + cm0 + n * DT * v0 + DT * DT * g * n * (n + 1) / 2, rtol=0, atol=1e-12)
        assert np.abs(s["V_local"] - m["V"]).max() < 1e-14
        assert np.allclose(s["rotation"], expm(n * DT * W), rtol=0, atol=1e-12)
        assert np.allclose(s["rotational_vel"], w, rtol=0, atol=1e-12)
        assert np.allclose(s["V_global"], m["V"] @ s["rotation"].T + s["translation"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("case", ["local_frame", "no_local_frame", "fixed", "compressed", "compressed_no_local_frame"])
def test_object_step_matches_oracle(tmp_path, mesh, case):
    V, F = mesh
    steps = 48
    tractions = pressure_pulse(V, F, steps, 2, 7, 300.0)
    obj = {"translational_vel": [0.1, 0.0, 0.05], "rotational_vel": [0.0, 2.0, 1.0], "translation": [0.0, 0.5, 0.0]}
    sim = {}
    if case.endswith("no_local_frame"):
        sim["update_local_frame"] = False
    if case.startswith("compressed"):
        sim["compression_ratio"] = 0.5  # wavelet-compressed G_n, H_n, u_p and H_0^-1
    if case == "fixed":
        obj["is_fixed"] = True
    m, traj, info = simulate(tmp_path, V, F, obj, tractions, **sim)
    worst, o = compare(m, traj, obj, len(V), tractions)
    print(f"{case}: worst relative difference {worst:.2e}, step {info['step_ms']:.3f} ms")
    # the pulse deforms the body, and the deformation is still there after it ends
    assert np.linalg.norm(o.V_local - m["V"]) > 1e-6


CUBE_E, CUBE_NU, CUBE_RHO, CUBE_DT, CUBE_STEPS = 1e7, 0.3, 1000.0, 1e-3, 200
CUBE_OBJ = {"translational_vel": [0.0, -0.5, 0.0], "rotational_vel": [0.3, 0.0, 0.2], "translation": [0.0, 1.0, 0.0]}


def cube_run(tmp_path):
    """The k = 8 cube (386 vertices, side 1 m) through the facade and the GPU object for 200 steps: a 10-step
    upward pressure pulse of 4e4 Pa on its bottom face (what a floor contact applies), gravity, an initial spin.
    E = 1e7 Pa (nu = 0.3, rho = 1000, dt = 1e-3 s, BDF2) with the mesh-derived L = 30: with E = 1e5 Pa waves need
    ~280 steps to cross the cube, and L = 32 truncates the convolution so that the step diverges (both
    implementations do; the LU-solve and explicit-inverse variants of the oracle then part by 1e-3 m at step 150
    while here they agree to 2e-15 m).  Returns the facade's dumped matrices (as ObjectStep reads them), the
    trajectory, the tractions and the run info."""
    V, F = M.cube(8, 0.5)
    steps, nv = CUBE_STEPS, len(V)
    Vc = M.center(V, F)
    p = np.zeros((steps, nv, 3))
    bottom = Vc[:, 1] < Vc[:, 1].min() + 1e-9
    p[20:30, bottom, 1] = 4e4
    tractions = p.reshape(steps, 3 * nv)
    mu = CUBE_E / (2 * (1 + CUBE_NU))
    obj_path = tmp_path / "cube.obj"
    M.write_obj(str(obj_path), V, F)
    cfg = {"enable_cuda": True, "mesh": {"file": str(obj_path)},
           "material": {"rho": CUBE_RHO, "mu": mu, "nu": CUBE_NU},
           "simulation": {"dt": CUBE_DT, "gaussian_quadrature_order": 5, "quadrature_subdivision": 2,
                          "multistep_method": "BDF2", "use_galerkin": False},
           "object": CUBE_OBJ, "gravity": GRAVITY}
    path = tmp_path / "scene.json"
    path.write_text(json.dumps(cfg))
    tr = tmp_path / "tractions.bin"
    tractions.tofile(tr)
    print("object step cube k=8, 200 steps", flush=True)
    r = subprocess.run([DRIVER, "--simulate", str(path), str(steps), str(tr), str(tmp_path / "out")],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    L, n3 = info["L"], 3 * nv
    assert L == 30
    post = lambda name, shape: np.fromfile(tmp_path / f"out_post_{name}.bin").reshape(shape)
    m = {"G": np.fromfile(tmp_path / "out_G.bin").reshape(L, n3, n3),
         "H": np.fromfile(tmp_path / "out_H.bin").reshape(L, n3, n3),
         "B": np.fromfile(tmp_path / "out_B.bin").reshape(L, n3, 6),
         "V": post("V", (nv, 3)), "F": np.asarray(F), "original_cm": post("original_cm", (3,)),
         "u_p": post("u_p", (n3, n3)), "rotation_removal_weights": post("rotation_removal_weights", (nv,)),
         "update_local_frame": True}
    for k in ("translational_acc_mat", "rotational_acc_mat", "B_trans_inv_H", "B_euler_inv_H",
              "B_trans_inv_G_prime", "B_euler_inv_G_prime"):
        m[k] = post(k, (3, n3))
    m["B_euler_inv_B_trans"] = post("B_euler_inv_B_trans", (3, 3))
    m["H_inv_B_trans"] = post("H_inv_B_trans", (n3, 3))
    m["H_inv_B_euler"] = post("H_inv_B_euler", (n3, 3))
    traj = np.fromfile(tmp_path / "out_traj.bin").reshape(steps, 24 + 6 * nv)
    return V, F, m, traj, tractions, info


def step_against(m, traj, tractions, nv, tol=1e-8):
    """Steps oracle/cqbem_object.ObjectStep on the matrices m beside the recorded trajectory; every vertex of
    V_global within tol (m) of it at every step.  Returns the worst difference and the final object."""
    from cqbem_object import ObjectStep
    o = ObjectStep(m, CUBE_DT, GRAVITY, CUBE_OBJ)
    worst = 0.0
    for k in range(len(tractions)):
        o.estimate_next_state()
        if np.any(tractions[k] != 0):
            o.update_V_estimate_matrices()
            o.p = tractions[k].copy()
            o.has_constraints = True
        o.confirm_next_state()
        got = split(traj[k], nv)
        worst = max(worst, np.abs(got["V_global"] - o.V_global).max())
        assert np.abs(got["V_global"] - o.V_global).max() < tol, (k, worst)
    return worst, o


def test_cube_trajectory_200_steps(tmp_path):
    """The time stepping alone: the oracle's object steps on the matrices the facade assembled on the GPU."""
    V, F, m, traj, tractions, info = cube_run(tmp_path)
    worst, o = step_against(m, traj, tractions, len(V))
    deform = np.abs(o.V_local - m["V"]).max()
    print(f"cube 200 steps: worst vertex difference {worst:.2e} m, final deformation {deform:.2e} m, "
          f"step {info['step_ms']:.3f} ms")
    assert deform > 1e-6  # the pulse left the cube ringing


@pytest.mark.timeout(600)
def test_cube_trajectory_200_steps_oracle_assembled(tmp_path, oracle):
    """End to end against the CPU precompute: the oracle assembles G, H and B of the cube in FP64 (same L and R
    as the facade derives, order 5 x subdivision 2), numpy derives u_p, H_0^-1 B and the least-squares products
    from them (tests/post_numpy.py, as the reference's init does), and the oracle's object steps on those.  The
    facade's GPU assembly and GPU object must stay within 1e-8 m of that trajectory at every vertex and step."""
    from post_numpy import time_stepping_matrices
    V, F, m_gpu, traj, tractions, info = cube_run(tmp_path)
    nv, L = len(V), info["L"]
    Vc = M.center(V, F)
    mu = CUBE_E / (2 * (1 + CUBE_NU))
    lam = 2 * CUBE_NU * mu / (1 - 2 * CUBE_NU)
    c1, c2 = float(np.sqrt((lam + 2 * mu) / CUBE_RHO)), float(np.sqrt(mu / CUBE_RHO))
    R = M.contour_radius(L - 1)
    print(f"oracle assembly of the cube: nv={nv} L={L}", flush=True)
    ref = oracle.assemble(Vc, F, oracle.params(c1, c2, CUBE_RHO, CUBE_NU, CUBE_DT, R, L, nthreads=0))
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    errs = {"G": rel(m_gpu["G"], ref["G"]), "H": rel(-m_gpu["H"], ref["H"]), "B": rel(m_gpu["B"], ref["B6"])}
    print("  GPU vs oracle matrices: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    m = time_stepping_matrices(Vc, F, ref["G"], ref["H"], ref["B6"], CUBE_RHO, M.volume_centroid(V, F)[1])
    worst, o = step_against(m, traj, tractions, nv)
    deform = np.abs(o.V_local - m["V"]).max()
    print(f"cube 200 steps, oracle-assembled: worst vertex difference {worst:.2e} m, final deformation "
          f"{deform:.2e} m", flush=True)
    assert deform > 1e-6


def test_compressed_matrix_cache(tmp_path, mesh):
    """With simulation.coeffs_folder_path the compressed matrices go to the reference's .cpmat files
    (H_<n>_<ratio>.cpmat, G_<n>_..., u_p_..., H_inv_...; wavelets.hpp:451-504 layout): the first run writes them,
    a second run reads all of them and steps identically.  The written G_1 and H_1 are the oracle's compressions
    of the dumped (host) G_1 and negated H_1, bit for bit: H_n is compressed from the un-negated device copy with
    the factor -1."""
    import glob
    from cqbem_wavelet import compress
    V, F = mesh
    steps = 12
    tractions = pressure_pulse(V, F, steps, 1, 3, 300.0)
    obj = {"translational_vel": [0.1, 0.0, 0.0]}
    cache = tmp_path / "cache"
    sim = {"compression_ratio": 0.5, "coeffs_folder_path": str(cache)}
    m1, traj1, info1 = simulate(tmp_path, V, F, obj, tractions, **sim)
    L, n3 = info1["L"], 3 * len(V)
    folder = glob.glob(str(cache / "elastodynamicsCQBEM_*"))
    assert len(folder) == 1
    names = sorted(os.path.basename(f) for f in glob.glob(folder[0] + "/*.cpmat"))
    want = sorted([f"{k}_{n}_0.500000.cpmat" for k in ("G", "H") for n in range(L)] +
                  ["u_p_0.500000.cpmat", "H_inv_0.500000.cpmat"])
    assert names == want and info1["compressed_loaded"] == 0
    for name, dense in (("G_1", m1["G"][1]), ("H_1", m1["H"][1])):
        with open(folder[0] + f"/{name}_0.500000.cpmat", "rb") as fh:
            raw = fh.read()
        N = np.frombuffer(raw[:4], np.int32)[0]
        rows, cols, nnz, outer, inner = np.frombuffer(raw[4:44], np.int64)
        off = 44
        val = np.frombuffer(raw[off:off + 8 * nnz], np.float64); off += 8 * nnz
        rp = np.frombuffer(raw[off:off + 4 * outer], np.int32); off += 4 * outer
        col = np.frombuffer(raw[off:off + 4 * nnz], np.int32); off += 4 * nnz
        assert off == len(raw) and N == n3 and rows == cols == outer == inner, name
        ref = compress(dense, int(n3 * n3 * 0.5))
        assert ref.shape[0] == rows and ref.nnz == nnz, name
        assert np.array_equal(rp, ref.indptr[:-1]) and np.array_equal(col, ref.indices), name
        assert np.array_equal(val, ref.data), name
    m2, traj2, info2 = simulate(tmp_path, V, F, obj, tractions, **sim)  # same scene, same cache key
    assert info2["cached"] and info2["compressed_loaded"] == 2 * L + 2
    assert np.array_equal(traj1, traj2)
