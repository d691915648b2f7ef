"""Seeded synthetic tight-binding inputs for the Fock-build / solve workloads.

Host-side NumPy generators (float64 / complex128) shared by the tests and the
benchmark, so the HIP path and the CPU checker consume identical arrays.

Conventions (same as the reference's centred meshgrid order,
``src/jax_hf/utils.py:490`` and ``src/jax_hf/problem.py:90-98``):

* mesh index ``i`` on axis ``a`` <-> fractional coordinate ``(i - n//2)/n``
  along the reciprocal vector ``b_a``;
* triangular lattice ``a1 = (1, 0)``, ``a2 = (1/2, sqrt(3)/2)`` so
  ``b1 = 2 pi (1, -1/sqrt(3))``, ``b2 = 2 pi (0, 2/sqrt(3))``;
* ``|q|`` is the minimal-image norm of ``q = (m1/n) b1 + (m2/n) b2`` over the
  reciprocal lattice translations ``{-1, 0, 1}^2``.

Draw order of ``tight_binding`` (documented so the arrays are reproducible):
``rng = numpy.random.default_rng(seed)``; for ``R`` in
``[0, a1, -a1, a2, -a2, a1-a2, -(a1-a2)]`` draw ``t_R = 0.5 (X + iY)`` with
``X, Y ~ N(0,1)`` of shape ``(nb, nb)`` (X first); ``t_0`` is Hermitised; then
``h(k) = M(k) + M(k)^dagger`` with ``M(k) = sum_R t_R exp(i k.R)``.
"""

from __future__ import annotations

import numpy as np

A1 = np.array([1.0, 0.0])
A2 = np.array([0.5, np.sqrt(3.0) / 2.0])
B1 = 2.0 * np.pi * np.array([1.0, -1.0 / np.sqrt(3.0)])
B2 = 2.0 * np.pi * np.array([0.0, 2.0 / np.sqrt(3.0)])
HOPS = [np.zeros(2), A1, -A1, A2, -A2, A1 - A2, -(A1 - A2)]


def kmesh(n1: int, n2: int) -> np.ndarray:
    """Cartesian k-points of the centred (n1, n2) mesh, shape (n1, n2, 2)."""
    f1 = (np.arange(n1) - n1 // 2) / n1
    f2 = (np.arange(n2) - n2 // 2) / n2
    return f1[:, None, None] * B1[None, None, :] + f2[None, :, None] * B2[None, None, :]


def qnorm_min_image(n1: int, n2: int) -> np.ndarray:
    """Minimal-image |q| on the centred (n1, n2) q-mesh, shape (n1, n2)."""
    q = kmesh(n1, n2)
    best = np.full((n1, n2), np.inf)
    for g1 in (-1, 0, 1):
        for g2 in (-1, 0, 1):
            G = g1 * B1 + g2 * B2
            best = np.minimum(best, np.linalg.norm(q + G[None, None, :], axis=-1))
    return best


def tight_binding(n1: int, n2: int, nb: int, seed: int, rng=None) -> np.ndarray:
    """h(k) = sum_R t_R e^{ik.R} + h.c. on the centred mesh, (n1, n2, nb, nb) c128."""
    rng = np.random.default_rng(seed) if rng is None else rng
    ts = []
    for _ in HOPS:
        X = rng.standard_normal((nb, nb))
        Y = rng.standard_normal((nb, nb))
        ts.append(0.5 * (X + 1j * Y))
    ts[0] = 0.5 * (ts[0] + ts[0].conj().T)
    k = kmesh(n1, n2)
    M = np.zeros((n1, n2, nb, nb), np.complex128)
    for R, t in zip(HOPS, ts):
        ph = np.exp(1j * (k @ R))
        M += ph[..., None, None] * t[None, None]
    return np.ascontiguousarray(M + np.conj(np.swapaxes(M, -1, -2)))


def coulomb_scalar(n1: int, n2: int, U0: float = 5.0, q0: float = 0.3) -> np.ndarray:
    """Scalar V(q) = U0 / (|q| + q0), shape (n1, n2, 1, 1) float64."""
    return (U0 / (qnorm_min_image(n1, n2) + q0))[..., None, None]


def layer_of_band(nb: int, nlayers: int) -> np.ndarray:
    """Layer index per band: band mod nlayers."""
    return np.arange(nb) % nlayers


def coulomb_layer(n1: int, n2: int, nb: int, nlayers: int = 2, U0: float = 5.0,
                  q0: float = 0.3, d: float = 0.5) -> np.ndarray:
    """V_ll'(q) = U0 e^{-|q| d |l-l'|} / (|q| + q0), shape (n1, n2, nb, nb) float64."""
    q = qnorm_min_image(n1, n2)[..., None, None]
    lay = layer_of_band(nb, nlayers)
    dl = np.abs(lay[:, None] - lay[None, :])[None, None]
    return U0 * np.exp(-q * d * dl) / (q + q0)


def haar_density(n1: int, n2: int, nb: int, rng) -> np.ndarray:
    """P(k) = U diag(f) U^dagger, U Haar (QR of complex Ginibre, R-phase fixed), f~U(0,1)."""
    Z = (rng.standard_normal((n1 * n2, nb, nb)) + 1j * rng.standard_normal((n1 * n2, nb, nb))) / np.sqrt(2.0)
    Q, R = np.linalg.qr(Z)
    d = np.diagonal(R, axis1=-2, axis2=-1)
    Q = Q * (d / np.abs(d))[:, None, :]
    f = rng.uniform(0.0, 1.0, size=(n1 * n2, nb))
    P = np.einsum("kin,kn,kjn->kij", Q, f, np.conj(Q))
    P = 0.5 * (P + np.conj(np.swapaxes(P, -1, -2)))
    return np.ascontiguousarray(P.reshape(n1, n2, nb, nb))


def uniform_weights(n1: int, n2: int) -> np.ndarray:
    return np.full((n1, n2), 1.0 / (n1 * n2))


def config0(nk: int = 24, nb: int = 4, seed: int = 0):
    """Plain CG FP64 case: triangular lattice, scalar V, T=0.1, n=2.1 per k, P0=0."""
    h = tight_binding(nk, nk, nb, seed)
    return dict(weights=uniform_weights(nk, nk), hamiltonian=h,
                coulomb_q=coulomb_scalar(nk, nk), T=0.1, n_electrons=2.1,
                P0=np.zeros_like(h))


def config1(nk: int = 96, nb: int = 16, seed: int = 1):
    """Isolated build_fock + hf_energy: layer-resolved V (layer = band mod 2), Haar P.

    One generator ``default_rng(seed)`` draws the hoppings first, then P.
    """
    rng = np.random.default_rng(seed)
    h = tight_binding(nk, nk, nb, seed, rng=rng)
    P = haar_density(nk, nk, nb, rng)
    return dict(weights=uniform_weights(nk, nk), hamiltonian=h,
                coulomb_q=coulomb_layer(nk, nk, nb, 2), T=0.1, P=P)


def config2(nk: int = 64, nb: int = 8, seed: int = 2, offset: float = 0.1):
    """Newton / DIIS precision form: layer = band mod 4, T=0.05.

    Filling: n = (nb/4 + offset) * sum(w), i.e. quarter filling of the nb bands
    plus ``offset`` electrons per k-point.
    """
    h = tight_binding(nk, nk, nb, seed)
    return dict(weights=uniform_weights(nk, nk), hamiltonian=h,
                coulomb_q=coulomb_layer(nk, nk, nb, 4), T=0.05,
                n_electrons=nb / 4.0 + offset, P0=np.zeros_like(h))


def config3_specs(nsolves: int = 64, nb: int = 16, seed: int = 3,
                  nks=(36, 48, 60, 72, 90)):
    """Density-scan specs: per solve (nk, n_electrons, u).

    nk ~ Uniform{nks}; filling nu ~ U(-0.5, 0.5) around neutrality (n = nb/2 + nu
    per k); displacement u ~ U(0, 0.4) entering as u * diag(+1 / -1 by layer).
    """
    rng = np.random.default_rng(seed)
    nk = rng.choice(np.asarray(nks), size=nsolves)
    nu = rng.uniform(-0.5, 0.5, size=nsolves)
    u = rng.uniform(0.0, 0.4, size=nsolves)
    return [dict(nk=int(nk[i]), n_electrons=nb / 2.0 + float(nu[i]), u=float(u[i]))
            for i in range(nsolves)]


def displacement(nb: int, u: float, nlayers: int = 2) -> np.ndarray:
    """u * diag(+1 for layer 0, -1 for layer 1) (layer = band mod 2)."""
    lay = layer_of_band(nb, nlayers)
    return np.diag(np.where(lay == 0, u, -u)).astype(np.complex128)
