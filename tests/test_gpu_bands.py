"""CQBEM_QUAD_BANDED: the distance-banded regular-face rule of the collocation G/H pair kernel and the B kernels.

- With an infinite distance threshold no tile or face is far: bit-identical to the reference rule.
- At three frequencies, every 3x3 block (row i, column j) of U and P is integrated in tile (i, patch of j): it must
  equal, bit for bit, the reference-rule run's block where the host plan (cqbem_band_plan) says near, and the run
  whose every tile outside its patch's sphere is far where the plan says far.
- B per (row, face): the banded B minus the reference-rule B of a row equals the sum, over the faces the face plan
  (cqbem_band_plan_faces) calls far for it, of the far rule minus the reference rule applied to the oracle's point
  kernels (b_hat, b_hat_ijk).
- Accuracy against a refined rule (the reference order at 2x the subdivision), icosphere level 4, L = 32, the
  default thresholds: per time-domain weight the banded G and H add at most 5% of the reference rule's largest
  per-weight error to the reference rule's own error (printed with the evaluation fraction); B keeps the reference
  rule unless band_kappa_b is set, so it adds nothing."""
import numpy as np
import pytest

from surfaceonlydeformables_amd import abi
from surfaceonlydeformables_amd import mesh as M

pytestmark = pytest.mark.gpu

RHO = 1000.0


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


def mesh3():
    V, F = M.icosphere(3, 0.2)
    return M.center(M.jitter(V, F, 0.01, 0), F), F


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def params(L=32, **kw):
    _, c1, c2 = M.material_speeds(5e5, 0.3, RHO)
    return abi.make_params(c1, c2, RHO, 0.3, 1e-3, 1e-10 ** (1 / (2 * L)), L, **kw)


def test_banded_with_no_far_tile_is_the_reference_rule(ctx):
    V, F = mesh3()
    mesh = abi.make_mesh(V, F)
    print("banded, kappa = 1e300", flush=True)
    ref = ctx.assemble(mesh, params(L=16))
    band = ctx.assemble(mesh, params(L=16, quad=abi.QUAD_BANDED, band_kappa=1e300, band_kappa_b=1e300))
    for k in ref:
        assert np.array_equal(ref[k], band[k]), k


# (kappa, l): the plan is the same at every frequency; three frequencies across the spectrum
PLANS = [(1.0, 0), (0.5, 8), (1.0, 16)]


@pytest.mark.parametrize("kappa,l", PLANS)
def test_banded_tiles_follow_the_plan(ctx, kappa, l):
    V, F = mesh3()
    mesh = abi.make_mesh(V, F)
    nv = len(V)
    p_band = params(quad=abi.QUAD_BANDED, band_kappa=kappa)
    p_far = params(quad=abi.QUAD_BANDED, band_kappa=1e-300)
    print(f"band plan kappa={kappa} l={l}", flush=True)
    near = ctx.laplace(mesh, params(), l)
    band = ctx.laplace(mesh, p_band, l)
    far = ctx.laplace(mesh, p_far, l)
    pov, plan = abi.band_plan(mesh, p_band)
    _, plan_far = abi.band_plan(mesh, p_far)
    tile = plan[:, pov]  # (row i, column j) -> far?
    assert np.all(plan_far[plan])  # every far tile of the banded plan is far in the all-far run too
    print(f"  far tiles {plan.mean():.3f}, far blocks {tile.mean():.3f}", flush=True)
    assert 0 < tile.mean() < 1
    for k in ("U", "P"):
        b = band[k].reshape(nv, 3, nv, 3)
        expect = np.where(tile[:, None, :, None], far[k].reshape(nv, 3, nv, 3), near[k].reshape(nv, 3, nv, 3))
        assert np.array_equal(b, expect), k
    for k in ("Bt", "Ba"):  # without band_kappa_b, B keeps the reference rule
        assert np.array_equal(band[k], near[k]), k


def test_banded_b_faces_follow_the_plan(ctx, oracle):
    V, F = M.icosphere(2, 0.2)
    V = M.center(M.jitter(V, F, 0.01, 0), F)
    mesh = abi.make_mesh(V, F)
    L, l, kappa = 16, 4, 0.5
    p_band = params(L=L, quad=abi.QUAD_BANDED, band_kappa_b=kappa)
    print(f"B face plan kappa_b={kappa} l={l}", flush=True)
    band = ctx.laplace(mesh, p_band, l)
    ref = ctx.laplace(mesh, params(L=L), l)
    far = abi.band_plan_faces(mesh, p_band)
    print(f"  far faces {far.mean():.3f}", flush=True)
    assert 0 < far.mean() < 1
    for f in range(len(F)):  # a face at the row vertex is never far
        assert not far[F[f], f].any()
    _, c1, c2 = M.material_speeds(5e5, 0.3, RHO)
    s = oracle.frequencies(L, 1e-10 ** (1 / (2 * L)), 1e-3, 2)[l]
    rules = {+1: oracle.tri_rule_composite(5, 1), -1: oracle.tri_rule_composite(5, 2)}  # far rule, reference rule
    cr = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    area = 0.5 * np.linalg.norm(cr, axis=1)
    nrm = cr / (2 * area)[:, None]
    for i in (0, 40, 97, 161):
        dBt = np.zeros((3, 3), complex)
        dBa = np.zeros((3, 3, 3), complex)
        for f in np.flatnonzero(far[i]):
            for sign, (b, w) in rules.items():
                for q in range(len(w)):
                    y = b[q, 0] * V[F[f, 0]] + b[q, 1] * V[F[f, 1]] + b[q, 2] * V[F[f, 2]]
                    wq = sign * w[q] * area[f]
                    dBt += wq * oracle.kernel(2, V[i], y, nrm[f], c1, c2, RHO, s)
                    dBa += wq * oracle.kernel(3, V[i], y, nrm[f], c1, c2, RHO, s)
        for k, d in (("Bt", dBt), ("Ba", dBa)):
            got = band[k][3 * i:3 * i + 3] - ref[k][3 * i:3 * i + 3]
            d = d.reshape(got.shape)
            scale = np.linalg.norm(ref[k][3 * i:3 * i + 3])
            print(f"  row {i} {k}: far faces {int(far[i].sum())}, far - reference {np.linalg.norm(d) / scale:.2e} of the "
                  f"row, GPU vs oracle {np.linalg.norm(got - d) / scale:.2e}", flush=True)
            assert np.linalg.norm(d) > 1e-9 * scale  # the rules differ visibly at these faces
            assert np.linalg.norm(got - d) <= 1e-12 * scale, (i, k)


def test_banded_accuracy_against_a_refined_rule(ctx):
    import torch
    from surfaceonlydeformables_amd.tune_bands import assemble, per_weight_rel, workload
    V, F, prm = workload("icosphere-l4")
    mesh = abi.make_mesh(V, F)
    L = prm["L"]
    print("banded accuracy, icosphere level 4, L = 32, default threshold", flush=True)
    fine, _, _ = assemble(ctx, torch, mesh, prm, subdivision=4)
    ref, _, t_ref = assemble(ctx, torch, mesh, prm)
    e_ref = {k: per_weight_rel(torch, ref[k], fine[k]) for k in ref}
    del ref
    band, _, t_band = assemble(ctx, torch, mesh, prm, quad=abi.QUAD_BANDED)
    e_band = {k: per_weight_rel(torch, band[k], fine[k]) for k in band}
    del band, fine
    torch.cuda.empty_cache()
    frac = t_band["gh_evaluations"] / t_ref["gh_evaluations"]
    print(f"  regular-face evaluations {frac:.3f} of the reference rule's; G/H phase {t_band['regular_gh']:.1f} ms vs "
          f"{t_ref['regular_gh']:.1f} ms; B phase {t_band['regular_b']:.1f} ms vs {t_ref['regular_b']:.1f} ms", flush=True)
    assert frac < 0.85
    for k in e_ref:
        added = e_band[k] - e_ref[k]
        print(f"  {k}: vs refined rule, reference rule max {e_ref[k].max():.2e}, banded max {e_band[k].max():.2e}; "
              f"largest added error {added.max():.2e} at n={int(added.argmax())}", flush=True)
        assert added.max() <= 0.05 * e_ref[k].max(), k


def test_banded_with_mixed_precision(ctx, monkeypatch):
    """CQBEM_QUAD_BANDED with CQBEM_MIXED: G and H come from the same banded FP64 pair kernel (bit-identical to the
    banded FP64 run); B from the float B kernel, without face bands by default (bit-identical to the unbanded mixed
    run) and with them when band_kappa_b is set (within float rounding of the FP64 run with the same face bands, and
    different from the unbanded float B)."""
    monkeypatch.delenv("CQBEM_MIXED_G", raising=False)
    V, F = mesh3()
    mesh = abi.make_mesh(V, F)
    print("banded + mixed", flush=True)
    both = ctx.assemble(mesh, params(L=16, quad=abi.QUAD_BANDED, band_kappa=1.0, precision=abi.MIXED))
    band = ctx.assemble(mesh, params(L=16, quad=abi.QUAD_BANDED, band_kappa=1.0))
    mixed = ctx.assemble(mesh, params(L=16, precision=abi.MIXED))
    for k in ("G", "H"):
        assert np.array_equal(both[k], band[k]), k
    for k in ("B_trans", "B_angular", "B"):
        assert np.array_equal(both[k], mixed[k]), k
    assert not np.array_equal(both["G"], mixed["G"])  # the bands are active (kappa = 1 has far tiles here)
    kw = dict(L=16, quad=abi.QUAD_BANDED, band_kappa=1.0, band_kappa_b=2.0)
    both_b = ctx.assemble(mesh, params(precision=abi.MIXED, **kw))
    band_b = ctx.assemble(mesh, params(**kw))
    for k in ("B_trans", "B_angular", "B"):
        e_band, e_unbanded = rel(both_b[k], band_b[k]), rel(both_b[k], mixed[k])
        print(f"  {k}: mixed, face bands 2.0: vs FP64 with them {e_band:.2e}, vs unbanded mixed {e_unbanded:.2e}",
              flush=True)
        # float rounding of B (~1e-5 here) exceeds what the face bands change, so: within float rounding of the
        # FP64 result with the same bands, and not the unbanded float result (the bands act in the float kernel)
        assert 0 < e_band < 1e-4 and not np.array_equal(both_b[k], mixed[k]), k
