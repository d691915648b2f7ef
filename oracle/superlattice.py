"""NumPy restatement of the reference's superlattice Fock build.

TEST INFRASTRUCTURE ONLY (imported by ``tests/``; never by the product path).

Follows ``src/jax_hf/superlattice.py``: the extended-grid layout (:75-164:
tile offsets of -G, the unique dG list in first-seen row-major order, the CSR
pair tables, ``numpy.fft.fftn`` of w V(|lag|) on the FFT-natural lag grid),
the per-dG scatter / convolve / gather exchange (:167-230), and the build_fock
adapter with the scalar and orbital-resolved Hartree shift (:233-377).

Pinned in ``tests/test_oracle_superlattice.py``: the layout's CSR
consistency (the reference test's check) and tile geometry; the exchange
against a brute-force circular convolution on the extended grid
(Sigma(x) = -sum_x' wV(x - x') X(x'), an O(N^4) direct sum that fixes the FFT
sign, normalisation and lag-grid conventions) and zero density -> zero; the
orbital Hartree path with an orbital-blind table equal to the scalar path
(the reference test's equivalence), the q = 0 warning, and the Hartree shift
of a k-independent diagonal density against its explicit pair sum.
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass

import numpy as np

__all__ = ["Layout", "build_extended_layout", "superlattice_exchange", "make_superlattice_build_fock_fn"]


@dataclass(frozen=True)
class Layout:
    N_ext_x: int
    N_ext_y: int
    n_supp_x: int
    n_supp_y: int
    g_a_off: np.ndarray
    n_delta: int
    pair_to_delta: np.ndarray
    V_lag_fft: np.ndarray
    V_lag: np.ndarray
    delta_pair_count: np.ndarray
    delta_pair_i: np.ndarray
    delta_pair_j: np.ndarray
    delta_pair_start: np.ndarray


def build_extended_layout(g_basis_fractional, g_basis_B, nkx, nky, coulomb_V, w_scalar) -> Layout:
    """superlattice.py:75-164 (V_lag, the unscaled kernel, is kept for the brute-force check)."""
    g = np.asarray(g_basis_fractional, np.int64)
    n_G = g.shape[0]
    pg = -g
    dx = int(pg[:, 0].max() - pg[:, 0].min() + 1)
    dy = int(pg[:, 1].max() - pg[:, 1].min() + 1)
    Nx, Ny = 2 * nkx * dx, 2 * nky * dy
    off = np.stack([pg[:, 0] - pg[:, 0].min(), pg[:, 1] - pg[:, 1].min()], 1)
    seen, p2d = {}, np.zeros((n_G, n_G), np.int64)
    for i in range(n_G):
        for j in range(n_G):
            key = (int(g[i, 0] - g[j, 0]), int(g[i, 1] - g[j, 1]))
            if key not in seen:
                seen[key] = len(seen)
            p2d[i, j] = seen[key]
    nd = len(seen)
    pairs = [[(i, j) for i in range(n_G) for j in range(n_G) if p2d[i, j] == d] for d in range(nd)]
    count = np.array([len(p) for p in pairs], np.int64)
    start = np.concatenate([[0], np.cumsum(count)]).astype(np.int64)
    di = np.array([i for p in pairs for i, _ in p], np.int64)
    dj = np.array([j for p in pairs for _, j in p], np.int64)
    lx = np.where(np.arange(Nx) < Nx // 2, np.arange(Nx), np.arange(Nx) - Nx) / nkx
    ly = np.where(np.arange(Ny) < Ny // 2, np.arange(Ny), np.arange(Ny) - Ny) / nky
    FX, FY = np.meshgrid(lx, ly, indexing="ij")
    B = np.asarray(g_basis_B, float)
    q = np.hypot(B[0, 0] * FX + B[0, 1] * FY, B[1, 0] * FX + B[1, 1] * FY)
    V = np.asarray(coulomb_V(q), float)
    return Layout(Nx, Ny, nkx * dx, nky * dy, off, nd, p2d, np.fft.fftn(float(w_scalar) * V), float(w_scalar) * V,
                  count, di, dj, start)


def superlattice_exchange(rho, lay: Layout, n_G, dorb, nkx, nky, direct=False):
    """superlattice.py:199-228 (direct=True: the circular convolution as an explicit sum)."""
    r = np.asarray(rho, np.complex128).reshape(nkx, nky, n_G, dorb, n_G, dorb)
    out = np.zeros_like(r)
    Nx, Ny = lay.N_ext_x, lay.N_ext_y
    for d in range(lay.n_delta):
        X = np.zeros((Nx, Ny, dorb, dorb), np.complex128)
        pairs = [(int(lay.delta_pair_i[s]), int(lay.delta_pair_j[s]))
                 for s in range(lay.delta_pair_start[d], lay.delta_pair_start[d + 1])]
        for i, j in pairs:
            ox, oy = nkx * lay.g_a_off[i, 0], nky * lay.g_a_off[i, 1]
            X[ox:ox + nkx, oy:oy + nky] = r[:, :, i, :, j, :]
        if direct:
            S = np.zeros_like(X)
            for x in range(Nx):
                for y in range(Ny):
                    Vs = lay.V_lag[(x - np.arange(Nx))[:, None] % Nx, (y - np.arange(Ny))[None, :] % Ny]
                    S[x, y] = -np.einsum("uv,uvab->ab", Vs, X)
        else:
            S = -np.fft.ifftn(np.fft.fftn(X, axes=(0, 1)) * lay.V_lag_fft[:, :, None, None], axes=(0, 1))
        for i, j in pairs:
            ox, oy = nkx * lay.g_a_off[i, 0], nky * lay.g_a_off[i, 1]
            out[:, :, i, :, j, :] = S[ox:ox + nkx, oy:oy + nky]
    return out.reshape(nkx, nky, n_G * dorb, n_G * dorb)


def make_superlattice_build_fock_fn(lay: Layout, n_G, dorb, nkx, nky, HH_GG, hartree_degeneracy,
                                    HH_GG_orbital=None):
    """superlattice.py:233-377."""
    D = n_G * dorb
    HHo = None
    if HH_GG_orbital is not None:
        HHo = np.asarray(HH_GG_orbital, np.float64)
        if HHo.shape != (n_G, n_G, dorb, dorb):
            raise ValueError(f"HH_GG_orbital must have shape {(n_G, n_G, dorb, dorb)}; got {HHo.shape}")
        if np.max(np.abs(HHo[np.arange(n_G), np.arange(n_G)])) > 1e-12:
            warnings.warn("HH_GG_orbital has nonzero diagonal-G blocks; the q=0 layer Hartree will be included.",
                          UserWarning, stacklevel=2)
    HH = np.asarray(HH_GG, np.float64)
    deg = float(hartree_degeneracy)

    def build(P, *, h, refP=None, w2d=None, include_exchange=True, include_hartree=False, project_fn=None, **_):
        rho = np.asarray(P, np.complex128) - (0 if refP is None else np.asarray(refP, np.complex128))
        S = superlattice_exchange(rho, lay, n_G, dorb, nkx, nky) if include_exchange else np.zeros_like(rho)
        if include_hartree:
            rb = np.einsum("ij,ijab->ab", np.asarray(w2d, np.float64), rho).reshape(n_G, dorb, n_G, dorb)
            blocks = np.zeros((n_G, dorb, n_G, dorb), np.complex128)
            if HHo is not None:
                po = np.einsum("AxBx->ABx", rb)
                rd = np.zeros((lay.n_delta, dorb), np.complex128)
                np.add.at(rd, lay.pair_to_delta.ravel(), po.reshape(n_G * n_G, dorb))
                sp = deg * np.einsum("ABag,ABg->ABa", HHo, rd[lay.pair_to_delta])
                for a in range(dorb):
                    blocks[:, a, :, a] = sp[:, :, a]
            else:
                rgg = np.einsum("AxBx->AB", rb)
                rd = np.zeros(lay.n_delta, np.complex128)
                np.add.at(rd, lay.pair_to_delta.ravel(), rgg.ravel())
                sg = deg * HH * rd[lay.pair_to_delta]
                for a in range(dorb):
                    blocks[:, a, :, a] = sg
            S = S + blocks.reshape(D, D)[None, None]
        F = np.asarray(h, np.complex128) + S
        F = 0.5 * (F + np.conj(np.swapaxes(F, -1, -2)))
        if project_fn is not None:
            F = project_fn(F)
            F = 0.5 * (F + np.conj(np.swapaxes(F, -1, -2)))
        return S, np.zeros_like(rho), F

    return build
