"""NumPy complex128 restatement of the reference's symmetry projector.

TEST INFRASTRUCTURE ONLY (imported by ``tests/``; never by the product path).

Follows ``src/jax_hf/symmetry.py``: the k-axis negation under the ``mod`` /
``flip`` conventions (:24-43), unitary-conjugation sums and averages (:46-60),
the combined same-k / flipped-k group (:63-75), the spatial average over
per-element k-index maps (:78-109), the time-reversal averages with a
convention (:140-150) or an index map with a validity mask (:112-137), and
``make_project_fn``'s argument checks and composition order (:153-281).

Pinned in ``tests/test_oracle_symmetry.py`` by the properties the reference's
own tests check (``tests/test_symmetry.py``): identity without arguments,
invariance of the output under every group element, Hermitian preservation,
idempotence (a projection), the explicit (A + zAz + xA(-k)x)/3 of the
combined group, the masked time-reversal index map's (0, 0) and (1, 2)
entries, and the argument errors.
"""

from __future__ import annotations

import numpy as np

__all__ = ["make_project_fn", "flip_k_index"]


def _dag(X):
    return np.conj(np.swapaxes(X, -1, -2))


def flip_k_index(nk1: int, nk2: int, convention: str, axes=(0, 1)):
    """(i, j) source indices of A(-k) on the mesh (symmetry.py:24-43)."""
    ii, jj = np.meshgrid(np.arange(nk1), np.arange(nk2), indexing="ij")
    if convention == "flip":
        si = nk1 - 1 - ii if 0 in axes else ii
        sj = nk2 - 1 - jj if 1 in axes else jj
    elif convention == "mod":
        si = (-ii) % nk1 if 0 in axes else ii
        sj = (-jj) % nk2 if 1 in axes else jj
    else:
        raise ValueError(f"k_convention must be 'mod' or 'flip', got {convention!r}")
    return si, sj


def _sum_conj(A, G):
    """sum_g g A g^dagger (symmetry.py:46-55)."""
    acc = np.zeros_like(A)
    for g in G:
        acc = acc + g @ A @ _dag(g)
    return acc


def make_project_fn(*, unitary_group=None, spatial_group=None, spatial_k_convention="mod",
                    spatial_k_flip_axes=(0, 1), spatial_k_index_maps=None, spatial_k_orbit_mask=None,
                    time_reversal_U=None, time_reversal_k_convention="mod", time_reversal_k_flip_axes=(0, 1),
                    time_reversal_k_index_map=None):
    """symmetry.py:153-281 (same arguments, checks and composition)."""
    has_group = unitary_group is not None or spatial_group is not None or spatial_k_index_maps is not None
    if not has_group and time_reversal_U is None:
        return lambda A: A
    G = None if unitary_group is None else np.asarray(unitary_group, np.complex128)
    S = None if spatial_group is None else np.asarray(spatial_group, np.complex128)
    U = None if time_reversal_U is None else np.asarray(time_reversal_U, np.complex128)
    if spatial_k_index_maps is not None:
        si = np.asarray(spatial_k_index_maps[0], np.int64)
        sj = np.asarray(spatial_k_index_maps[1], np.int64)
        if si.shape != sj.shape:
            raise ValueError("spatial_k_index_maps[0] and [1] must have the same shape; "
                             f"got {si.shape} vs {sj.shape}.")
        if S is None:
            raise ValueError("spatial_k_index_maps requires spatial_group (the orbital unitaries to average over).")
        if S.shape[0] != si.shape[0]:
            raise ValueError("spatial_group and spatial_k_index_maps must have matching "
                             f"leading dim; got {S.shape[0]} and {si.shape[0]}.")
    else:
        si = sj = None
    omask = None if spatial_k_orbit_mask is None else np.asarray(spatial_k_orbit_mask, bool)
    if time_reversal_k_index_map is not None:
        if U is None:
            raise ValueError("time_reversal_k_index_map requires time_reversal_U.")
        ti = np.asarray(time_reversal_k_index_map[0], np.int64)
        tj = np.asarray(time_reversal_k_index_map[1], np.int64)
        tv = np.asarray(time_reversal_k_index_map[2], bool)
    else:
        ti = tj = tv = None
    s_conv, s_axes = str(spatial_k_convention), tuple(spatial_k_flip_axes)
    t_conv, t_axes = str(time_reversal_k_convention), tuple(time_reversal_k_flip_axes)

    def project(A):
        A = np.asarray(A, np.complex128)
        out = A
        nk1, nk2 = A.shape[0], A.shape[1]
        if si is not None:                                   # symmetry.py:78-109, 258-266
            acc = np.zeros_like(out)
            for g in range(S.shape[0]):
                acc = acc + S[g] @ out[si[g], sj[g]] @ _dag(S[g])
            avg = acc / S.shape[0]
            if omask is not None:
                avg = np.where(omask[..., None, None], avg, out)
            out = avg
            if G is not None:
                out = _sum_conj(out, G) / G.shape[0]
        elif G is not None or S is not None:                 # symmetry.py:63-75, 267-274
            Gs = G if G is not None else np.eye(S.shape[-1], dtype=np.complex128)[None]
            acc = _sum_conj(out, Gs)
            n = Gs.shape[0]
            if S is not None:
                fi, fj = flip_k_index(nk1, nk2, s_conv, s_axes)
                acc = acc + _sum_conj(out[fi, fj], S)
                n += S.shape[0]
            out = acc / n
        if U is not None:                                    # symmetry.py:112-150, 275-279
            if ti is not None:
                tr = U @ np.conj(out[ti, tj]) @ _dag(U)
                out = np.where(tv[..., None, None], 0.5 * (out + tr), out)
            else:
                fi, fj = flip_k_index(nk1, nk2, t_conv, t_axes)
                out = 0.5 * (out + U @ np.conj(out[fi, fj]) @ _dag(U))
        return out

    return project
