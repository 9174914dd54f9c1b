"""CQBEM_QUAD_BANDED's tile plan on the CPU (cqbem_band_plan needs no GPU): every vertex's column belongs to one
patch; a tile is far exactly when the row vertex lies kappa x h beyond the patch's bounding sphere, so the far set
shrinks as kappa grows, is empty for an infinite kappa and never holds the tile of a vertex's own patch; the
default kappa is the tuned one; and the refusals (the reference rule, Galerkin, the float G kernel of mixed
precision)."""
import numpy as np
import pytest

from surfaceonlydeformables_amd import abi
from surfaceonlydeformables_amd import mesh as M


@pytest.fixture(scope="module")
def ico4():
    V, F = M.icosphere(4, 0.2)
    return M.center(M.jitter(V, F, 0.01, 0), F), F


def params(**kw):
    _, c1, c2 = M.material_speeds(5e5, 0.3, 1000.0)
    return abi.make_params(c1, c2, 1000.0, 0.3, 1e-3, 1e-10 ** (1 / 64), 32, **kw)


def test_plan_partitions_the_columns_and_shrinks_with_kappa(ico4):
    V, F = ico4
    mesh = abi.make_mesh(V, F)
    fractions = []
    for kappa in (1e-300, 1.0, 3.0, 6.0, 12.0):
        pov, far = abi.band_plan(mesh, params(quad=abi.QUAD_BANDED, band_kappa=kappa))
        n = far.shape[1]
        assert pov.min() == 0 and pov.max() == n - 1
        assert np.bincount(pov, minlength=n).min() > 0  # every patch holds columns
        assert not far[np.arange(len(V)), pov].any()  # a row is never far from its own vertex's patch
        fractions.append(far[:, pov].mean())
    assert all(a >= b for a, b in zip(fractions, fractions[1:]))
    assert fractions[0] > 0.5 and fractions[-1] < fractions[0]
    _, none = abi.band_plan(mesh, params(quad=abi.QUAD_BANDED, band_kappa=1e300))
    assert not none.any()


def test_default_kappa_is_the_tuned_one(ico4):
    V, F = ico4
    mesh = abi.make_mesh(V, F)
    _, dflt = abi.band_plan(mesh, params(quad=abi.QUAD_BANDED))
    _, six = abi.band_plan(mesh, params(quad=abi.QUAD_BANDED, band_kappa=6.0))
    assert np.array_equal(dflt, six)


def test_far_tiles_are_beyond_the_patch(ico4):
    """Far tile (i, p) => every vertex of every face of patch p is at least kappa h_min away from vertex i, with
    h_min the shortest edge (the sphere holds the faces, and kappa h >= kappa h_min)."""
    V, F = ico4
    mesh = abi.make_mesh(V, F)
    kappa = 2.0
    pov, far = abi.band_plan(mesh, params(quad=abi.QUAD_BANDED, band_kappa=kappa))
    e = np.concatenate([np.linalg.norm(V[F[:, k]] - V[F[:, (k + 1) % 3]], axis=1) for k in range(3)])
    rng = np.random.default_rng(0)
    for i in rng.choice(len(V), 40, replace=False):
        for p in np.flatnonzero(far[i]):
            d = np.linalg.norm(V[pov == p] - V[i], axis=1)
            assert d.min() >= kappa * e.min(), (i, p)


def test_face_plan_is_the_face_sphere_test(ico4):
    """The B face plan: face f is far for row vertex i exactly when |x_i - c_f| >= r_f (1 + 1e-12) + kappa h_f (c_f:
    the mean of its corners, r_f: its farthest corner, h_f: its longest edge); checked outside a 1e-9 rounding margin.
    A face at vertex i is never far; the far set shrinks as kappa grows and is empty for an infinite kappa."""
    V, F = ico4
    mesh = abi.make_mesh(V, F)
    kappa = 2.0
    far = abi.band_plan_faces(mesh, params(quad=abi.QUAD_BANDED, band_kappa_b=kappa))
    assert far.shape == (len(V), len(F))
    c = V[F].mean(axis=1)
    r = np.max(np.linalg.norm(V[F] - c[:, None, :], axis=2), axis=1)
    h = np.max(np.stack([np.linalg.norm(V[F[:, k]] - V[F[:, (k + 1) % 3]], axis=1) for k in range(3)]), axis=0)
    thr = r * (1 + 1e-12) + kappa * h
    for i in range(0, len(V), 7):
        d = np.linalg.norm(c - V[i], axis=1)
        assert far[i][d >= thr * (1 + 1e-9)].all() and not far[i][d <= thr * (1 - 1e-9)].any(), i
    for k in range(3):
        assert not far[F[:, k], np.arange(len(F))].any()
    fractions = [abi.band_plan_faces(mesh, params(quad=abi.QUAD_BANDED, band_kappa_b=k)).mean()
                 for k in (1e-300, 1.0, 6.0)]
    assert fractions[0] > fractions[1] > fractions[2] > 0
    assert not abi.band_plan_faces(mesh, params(quad=abi.QUAD_BANDED, band_kappa_b=1e300)).any()
    # kappa_b alone sets it: the G/H threshold does not
    assert np.array_equal(abi.band_plan_faces(mesh, params(quad=abi.QUAD_BANDED, band_kappa=0.5, band_kappa_b=2.0)), far)


def test_plan_refusals(ico4, monkeypatch):
    V, F = ico4
    mesh = abi.make_mesh(V, F)
    for p in (params(), params(quad=abi.QUAD_BANDED, galerkin=True), params(quad=7)):
        with pytest.raises(abi.CQBEMError):
            abi.band_plan(mesh, p)
    # mixed precision keeps G and H in the pair kernel: banded; with the float G kernel it is not
    monkeypatch.delenv("CQBEM_MIXED_G", raising=False)
    abi.band_plan(mesh, params(quad=abi.QUAD_BANDED, precision=abi.MIXED))
    monkeypatch.setenv("CQBEM_MIXED_G", "float")
    with pytest.raises(abi.CQBEMError):
        abi.band_plan(mesh, params(quad=abi.QUAD_BANDED, precision=abi.MIXED))


def _patch_lists(F, pov, npatches):
    """Faces listed by each patch: those with a corner in it (the pair kernel evaluates each once per patch)."""
    return [np.flatnonzero((pov[F] == p).any(axis=1)) for p in range(npatches)]


@pytest.mark.parametrize("case", ["ico4", "split_cube12"])
def test_pair_patches_fit_the_kernel_and_list_few_faces_twice(ico4, case):
    """The pair kernel's column patches (the band plan reports them): every vertex in one patch, at most 256
    vertices (the LDS accumulator slots) and 512 listed faces (four passes of 128) per patch.  Patches grown by
    fewest new faces list each face at most 1.14 times at icosphere level 4 (runs of the Morton order: 1.22).
    The split cube is six separate sheets: a patch whose sheet runs out continues in the next, so no patch is
    left small while another sheet still has vertices."""
    if case == "ico4":
        V, F = ico4
        mesh = abi.make_mesh(V, F)
    else:
        V, F, vm = abi.split_traction_discontinuity(*M.cube(12, 0.5), 20.0)
        mesh = abi.make_mesh(V, F, vm)
    pov, far = abi.band_plan(mesh, params(quad=abi.QUAD_BANDED))
    n = far.shape[1]
    sizes = np.bincount(pov, minlength=n)
    assert sizes.sum() == len(V) and sizes.min() > 0 and sizes.max() <= 256
    lists = _patch_lists(np.asarray(F), pov, n)
    counts = np.array([len(l) for l in lists])
    assert counts.max() <= 512
    listed = counts.sum() / len(F)
    if case == "ico4":
        assert listed <= 1.14, listed
    # only the last patch may be cut short by the end of the vertices, not by a mesh piece running out
    full = (sizes[:-1] == 256) | (counts[:-1] > 512 - 12)
    assert full.all(), (sizes, counts)
