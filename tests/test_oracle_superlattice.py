"""Pins the oracle's superlattice Fock build (oracle/superlattice.py) by the
reference tests' checks (tests/test_superlattice.py: CSR consistency, zero
density, orbital Hartree with an orbital-blind table == scalar Hartree, the
q = 0 warning) and by independent mathematics: the exchange against a
brute-force circular convolution on the extended grid, and the Hartree shift
of a k-independent density against its explicit pair sum."""

import numpy as np
import pytest

from oracle import superlattice as S


def hex_g_basis(g_cut):
    """Gamma plus the first g_cut shells of a hexagonal reciprocal lattice
    (tests/test_superlattice.py:29-47)."""
    b0 = 4 * np.pi / np.sqrt(3)
    B = np.stack([np.array([np.sqrt(3) / 2, -0.5]) * b0, np.array([0.0, 1.0]) * b0], axis=1)
    fr = np.array([(a, b) for a in range(-g_cut - 1, g_cut + 2) for b in range(-g_cut - 1, g_cut + 2)], np.int64)
    norms = np.linalg.norm(fr @ B.T, axis=1)
    shells = np.sort(np.unique(np.round(norms, 6)))
    return fr[norms <= shells[min(g_cut, len(shells) - 1)] + 1e-3], B


def vfunc(q):
    return 1.0 / np.sqrt(q ** 2 + 0.5 ** 2)


def _layout(nk, g_cut=1):
    g, B = hex_g_basis(g_cut)
    return S.build_extended_layout(g, B, nk, nk, vfunc, 1.0 / nk ** 2), g.shape[0]


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_layout_csr_and_tiles():
    lay, n_G = _layout(6)
    assert n_G == 7 and (lay.N_ext_x, lay.N_ext_y) == (36, 36) and lay.n_delta == 19
    seen = set()
    for d in range(lay.n_delta):
        for s in range(lay.delta_pair_start[d], lay.delta_pair_start[d + 1]):
            i, j = int(lay.delta_pair_i[s]), int(lay.delta_pair_j[s])
            assert lay.pair_to_delta[i, j] == d
            seen.add((i, j))
    assert len(seen) == n_G * n_G
    assert len({tuple(o) for o in lay.g_a_off}) == n_G           # one G per tile
    assert lay.g_a_off.min() >= 0 and lay.g_a_off.max() < lay.N_ext_x // 6


def test_exchange_matches_direct_circular_convolution():
    rng = np.random.default_rng(0)
    nk, dorb = 2, 2
    lay, n_G = _layout(nk)
    D = n_G * dorb
    rho = _rand(rng, (nk, nk, D, D))
    fft = S.superlattice_exchange(rho, lay, n_G, dorb, nk, nk)
    direct = S.superlattice_exchange(rho, lay, n_G, dorb, nk, nk, direct=True)
    assert np.abs(fft - direct).max() < 1e-12 * np.abs(direct).max()
    assert np.abs(S.superlattice_exchange(np.zeros_like(rho), lay, n_G, dorb, nk, nk)).max() == 0.0


def test_orbital_hartree_with_blind_table_equals_scalar_and_q0_warns():
    rng = np.random.default_rng(11)
    nk, dorb = 3, 3
    lay, n_G = _layout(nk)
    D = n_G * dorb
    HH = rng.standard_normal((n_G, n_G))
    HH = HH + HH.T
    np.fill_diagonal(HH, 0.0)
    HHo = np.broadcast_to(HH[:, :, None, None], (n_G, n_G, dorb, dorb)).copy()
    rho = _rand(rng, (nk, nk, D, D)) * 1e-3
    rho = 0.5 * (rho + np.conj(np.swapaxes(rho, -1, -2)))
    kw = dict(h=np.zeros_like(rho), w2d=np.full((nk, nk), 1.0 / nk ** 2), include_exchange=False,
              include_hartree=True)
    Fs = S.make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, HH, 1.6)(rho, **kw)[2]
    Fo = S.make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, HH, 1.6, HH_GG_orbital=HHo)(rho, **kw)[2]
    assert np.abs(Fs - Fo).max() < 1e-14
    HHq = HHo.copy()
    HHq[0, 0, 0, 0] = 0.5
    with pytest.warns(UserWarning, match="q=0"):
        S.make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, HH, 1.0, HH_GG_orbital=HHq)


def test_hartree_shift_of_uniform_density_is_the_pair_sum():
    """For rho_k = diag(n_(A,c)), the same at every k, rho_c(dG = 0) =
    sum_A n_(A,c) and every other dG vanishes, so the shift is
    deg * sum_c HH_orb[A, A, a, c] sum_A' n_(A',c) on the (A, a), (A, a)
    entries only (a q = 0 table: the warning is expected)."""
    rng = np.random.default_rng(4)
    nk, dorb = 2, 2
    lay, n_G = _layout(nk)
    D = n_G * dorb
    n = rng.uniform(size=D)
    rho = np.broadcast_to(np.diag(n).astype(complex), (nk, nk, D, D)).copy()
    HHo = rng.standard_normal((n_G, n_G, dorb, dorb))
    w = np.full((nk, nk), 0.25)
    with pytest.warns(UserWarning, match="q=0"):
        build = S.make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, np.zeros((n_G, n_G)), 2.0,
                                                  HH_GG_orbital=HHo)
    S_ = build(rho, h=np.zeros_like(rho), w2d=w,
               include_exchange=False, include_hartree=True)[0]
    tot = n.reshape(n_G, dorb).sum(axis=0)                  # sum over A' per orbital c
    expect = np.zeros((n_G, dorb, n_G, dorb))
    for A in range(n_G):
        for a in range(dorb):
            expect[A, a, A, a] = 2.0 * np.dot(HHo[A, A, a], tot)
    np.testing.assert_allclose(S_[0, 0], expect.reshape(D, D), atol=1e-13)
