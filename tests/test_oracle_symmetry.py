"""Pins the oracle's symmetry projector (oracle/symmetry.py) by the properties
the reference's own tests check (tests/test_symmetry.py), re-expressed in NumPy:
identity, invariance under every group element, Hermitian preservation,
idempotence, the explicit combined-group average, the masked time-reversal
index map, and the argument errors."""

import numpy as np
import pytest

from oracle.symmetry import flip_k_index, make_project_fn


def _dag(X):
    return np.conj(np.swapaxes(X, -1, -2))


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _c3_maps(nk):
    """C3 on the centred fractional grid (f_x, f_y) -> (-f_y, f_x - f_y)."""
    half = nk // 2
    ii, jj = np.meshgrid(np.arange(nk), np.arange(nk), indexing="ij")
    idx_i = np.stack([ii, (-jj) % nk, (jj - ii + half) % nk])
    idx_j = np.stack([jj, (ii - jj + half) % nk, (-ii) % nk])
    return idx_i, idx_j


def test_identity_without_arguments():
    A = np.ones((2, 2, 3, 3), complex)
    assert np.array_equal(make_project_fn()(A), A)


def test_unitary_group_invariant_hermitian_idempotent():
    rng = np.random.default_rng(0)
    nb = 3
    perm = np.zeros((nb, nb), complex)
    perm[0, 1] = perm[1, 2] = perm[2, 0] = 1.0
    G = np.stack([np.eye(nb), perm, perm @ perm])
    proj = make_project_fn(unitary_group=G)
    A = _rand(rng, (2, 2, nb, nb))
    Ah = 0.5 * (A + _dag(A))
    out = proj(Ah)
    for g in G:
        assert np.allclose(g @ out @ _dag(g), out, atol=1e-13)
    assert np.allclose(out, _dag(out), atol=1e-13)
    assert np.allclose(proj(out), out, atol=1e-13)


@pytest.mark.parametrize("conv", ["flip", "mod"])
def test_time_reversal_invariance(conv):
    rng = np.random.default_rng(1)
    nk, nb = (4, 2) if conv == "flip" else (5, 2)
    U = np.array([[0, -1j], [1j, 0]])
    proj = make_project_fn(time_reversal_U=U, time_reversal_k_convention=conv)
    out = proj(_rand(rng, (nk, nk, nb, nb)))
    fi, fj = flip_k_index(nk, nk, conv)
    assert np.allclose(out, U @ np.conj(out[fi, fj]) @ _dag(U), atol=1e-13)


def test_combined_same_k_and_flip_k_group_explicit():
    rng = np.random.default_rng(2)
    z = np.diag([1.0, -1.0]).astype(complex)
    x = np.array([[0, 1], [1, 0]], complex)
    proj = make_project_fn(unitary_group=np.stack([np.eye(2), z]), spatial_group=x[None],
                           spatial_k_convention="flip")
    A = _rand(rng, (4, 4, 2, 2))
    An = np.flip(A, axis=(0, 1))
    assert np.allclose(proj(A), (A + z @ A @ z + x @ An @ x) / 3.0, atol=1e-14)


def test_spatial_flip_only_axis_zero_idempotent():
    rng = np.random.default_rng(3)
    x = np.array([[0, 1], [1, 0]], complex)
    proj = make_project_fn(unitary_group=np.eye(2)[None], spatial_group=x[None], spatial_k_convention="flip",
                           spatial_k_flip_axes=(0,))
    once = proj(_rand(rng, (6, 4, 2, 2)))
    assert np.allclose(proj(once), once, atol=1e-13)


def test_spatial_index_maps_idempotent_and_invariant():
    rng = np.random.default_rng(11)
    nk, nb = 6, 4
    w = np.exp(2j * np.pi / 3)
    U = np.diag([1, w, w ** 2, 1])
    S = np.stack([np.eye(nb), U, U @ U])
    ii, jj = _c3_maps(nk)
    proj = make_project_fn(spatial_group=S, spatial_k_index_maps=(ii, jj))
    once = proj(_rand(rng, (nk, nk, nb, nb)))
    assert np.allclose(proj(once), once, atol=1e-13)
    for g in (1, 2):
        assert np.allclose(S[g] @ once[ii[g], jj[g]] @ _dag(S[g]), once, atol=1e-13)


def test_time_reversal_index_map_with_invalid_mask():
    rng = np.random.default_rng(3)
    nk, nb = 4, 2
    ii, jj = np.meshgrid(np.arange(nk), np.arange(nk), indexing="ij")
    valid = np.ones((nk, nk), bool)
    valid[0, 0] = False
    proj = make_project_fn(time_reversal_U=np.eye(nb),
                           time_reversal_k_index_map=((-ii) % nk, (-jj) % nk, valid))
    A = _rand(rng, (nk, nk, nb, nb))
    out = proj(A)
    assert np.allclose(out[0, 0], A[0, 0], atol=1e-15)
    assert np.allclose(out[1, 2], 0.5 * (A[1, 2] + np.conj(A[3, 2])), atol=1e-15)


def test_argument_errors():
    ii, jj = _c3_maps(4)
    with pytest.raises(ValueError, match="spatial_group"):
        make_project_fn(spatial_k_index_maps=(ii, jj))
    with pytest.raises(ValueError, match="leading dim"):
        make_project_fn(spatial_group=np.stack([np.eye(2)] * 3), spatial_k_index_maps=(ii[:2], jj[:2]))
    tmap = (ii[0], jj[0], np.ones((4, 4), bool))
    with pytest.raises(ValueError, match="time_reversal_U"):
        make_project_fn(unitary_group=np.eye(2)[None], time_reversal_k_index_map=tmap)
    # symmetry.py:205-206: no group and no time-reversal unitary -> identity, before any check
    A = np.ones((4, 4, 2, 2), complex)
    assert np.array_equal(make_project_fn(time_reversal_k_index_map=tmap)(A), A)
