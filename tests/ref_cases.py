"""Small model inputs of the reference's own solver tests, as plain arrays.

Each builder returns ``(weights, hamiltonian, coulomb_q, T)`` plus keyword
arguments for the kernel constructor, so the same inputs feed the oracle
(``oracle.OracleKernel``) and the GPU package (``jax_hf_amd.HartreeFockKernel``).

* ``dirac``: the massive 2D Dirac model of ``tests/test_scf_acceleration.py:26-38``
  (exchange 0.7) and ``tests/test_solver_newton.py:94-107`` (exchange 0.5):
  h = m sz + v (kx sx + ky sy) on the centred nk x nk mesh, unit weights,
  scalar exchange.
* ``multiband``: ``tests/test_solver_newton.py:75-91``: linspace(-1, 1, nb)
  on the diagonal plus a Hermitian random matrix times (cos kx + cos ky),
  seeded by ``default_rng(seed)`` in the reference's draw order.
* ``contact``: ``tests/test_solver_newton.py:55-72``: a two-band model with a
  sz-sz contact term and a unique Hartree-Fock minimum.
* ``gapped``: a two-band insulator (bands near -1 and +1 at every k, gap > 1)
  for the level-shift invariant.
"""

from __future__ import annotations

import numpy as np


def dirac(nk=4, m=0.3, v=1.0, T=0.1, ex=0.7):
    ks = (np.arange(nk) - nk // 2) * (2 * np.pi / nk)
    KX, KY = np.meshgrid(ks, ks, indexing="ij")
    h = np.zeros((nk, nk, 2, 2), np.complex128)
    h[..., 0, 0] = m
    h[..., 1, 1] = -m
    h[..., 0, 1] = v * (KX - 1j * KY)
    h[..., 1, 0] = v * (KX + 1j * KY)
    return (np.ones((nk, nk)), h, np.full((nk, nk, 1, 1), ex, np.complex128), T), {}


def multiband(nk, nb, T, ex, seed):
    rng = np.random.default_rng(seed)
    diag = np.linspace(-1.0, 1.0, nb)
    A = (rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))) * 0.35
    A = 0.5 * (A + A.conj().T)
    h = np.zeros((nk, nk, nb, nb), np.complex128)
    kx = np.arange(nk) / nk * 2 * np.pi
    for i in range(nk):
        for j in range(nk):
            hk = np.diag(diag).astype(np.complex128) + A * (np.cos(kx[i]) + np.cos(kx[j]))
            h[i, j] = 0.5 * (hk + hk.conj().T)
    return (np.ones((nk, nk)), h, np.full((nk, nk, 1, 1), ex, np.complex128), T), {}


def contact(g=0.3, nk=2, T=0.05):
    h = np.zeros((nk, nk, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -1.0, 1.0
    for i in range(nk):
        for j in range(nk):
            h[i, j, 0, 1] = h[i, j, 1, 0] = 0.4 + 0.1 * (i + j)
    sz = np.diag([1.0, -1.0]).astype(np.complex128)
    return ((np.ones((nk, nk)), h, np.full((nk, nk, 1, 1), 0.1, np.complex128), T),
            {"contact_terms": [(g, sz, sz)]})


def gapped(nk=4, T=0.01, ex=0.3, t=0.15):
    """Two bands at -1 and +1 with a weak k-dependent mixing: gap ~2 - 4t."""
    ks = (np.arange(nk) - nk // 2) * (2 * np.pi / nk)
    KX, KY = np.meshgrid(ks, ks, indexing="ij")
    h = np.zeros((nk, nk, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -1.0, 1.0
    h[..., 0, 1] = t * (np.cos(KX) + 1j * np.sin(KY))
    h[..., 1, 0] = np.conj(h[..., 0, 1])
    return (np.ones((nk, nk)), h, np.full((nk, nk, 1, 1), ex, np.complex128), T), {}


def comm_rms(F, P, w):
    """Weighted RMS of [F, P] (the reference tests' ``_comm_rms``)."""
    c = F @ P - P @ F
    return float(np.sqrt(np.sum(w * np.sum(np.abs(c) ** 2, axis=(-2, -1))) / max(np.sum(w), 1e-30)))
