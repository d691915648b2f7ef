"""Symmetry projectors on the GPU (src/jax_hf/symmetry.py).

``make_project_fn`` keeps the reference's arguments, checks and composition
order (symmetry.py:153-281) and returns a callable on device tensors.  Each
group average is one launch of ``hfx_symmetry_average``: for every k,

    out[k] = mask[k] ? (1/N) sum_g U_g op_g(A[src_g(k)]) U_g^dagger : A[k],

with the k-maps of the reference (identity, the ``mod`` / ``flip``
negations of selected axes, per-element index maps) turned into int32 gather
tables once, on the device, and op_g the complex conjugation of the
time-reversal element.  The tables and unitaries are uploaded at the first
call for a mesh and precision; after that a call only allocates its outputs
and launches kernels on the current stream, so a solver that has called it
once may capture it inside a HIP graph (``hfx_graph_safe``; the CG loop runs
its first iteration eagerly before capturing).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from ._ops import dt_code
from ._util import host_array

__all__ = ["make_project_fn"]

_MAX_ELEMS = 64


def _flip_index(nk1, nk2, convention, axes):
    """Source (i, j) of A(-k) under the mod / flip conventions (symmetry.py:24-43)."""
    ii, jj = np.meshgrid(np.arange(nk1), np.arange(nk2), indexing="ij")
    if convention == "flip":
        return (nk1 - 1 - ii if 0 in axes else ii), (nk2 - 1 - jj if 1 in axes else jj)
    if convention == "mod":
        return ((-ii) % nk1 if 0 in axes else ii), ((-jj) % nk2 if 1 in axes else jj)
    raise ValueError(f"k_convention must be 'mod' or 'flip', got {convention!r}")


class _Stage:
    """One group average: elements (unitary, k-source rule, conjugate?), an
    optional keep-mask, weight 1/N."""

    def __init__(self, elems, mask=None):
        self.elems = elems           # list of (U (nb, nb) complex128, kmap, conj)
        self.mask = None if mask is None else np.asarray(mask, bool)
        if len(elems) > _MAX_ELEMS:
            raise ValueError(f"a symmetry group of at most {_MAX_ELEMS} elements is supported, got {len(elems)}")
        self._dev = {}

    def _tables(self, nk1, nk2, nb, dtype, device):
        key = (nk1, nk2, nb, dtype, str(device))
        t = self._dev.get(key)
        if t is not None:
            return t
        ng = len(self.elems)
        src = np.empty((ng, nk1 * nk2), np.int32)
        ii, jj = np.meshgrid(np.arange(nk1), np.arange(nk2), indexing="ij")
        for g, (_, kmap, _) in enumerate(self.elems):
            if kmap is None:
                si, sj = ii, jj
            elif kmap[0] == "conv":
                si, sj = _flip_index(nk1, nk2, kmap[1], kmap[2])
            else:
                si, sj = np.asarray(kmap[1][g]), np.asarray(kmap[2][g])
                if si.shape != (nk1, nk2) or sj.shape != (nk1, nk2):
                    raise ValueError(f"k-index maps of shape {si.shape} do not match the mesh ({nk1}, {nk2})")
                if si.min() < 0 or sj.min() < 0 or si.max() >= nk1 or sj.max() >= nk2:
                    raise ValueError("k-index maps point outside the mesh")
            src[g] = (si * nk2 + sj).reshape(-1)
        U = np.stack([u for u, _, _ in self.elems]).astype(np.complex128)
        if U.shape[-1] != nb:
            raise ValueError(f"group unitaries are {U.shape[-1]}x{U.shape[-1]}, matrices {nb}x{nb}")
        Ut = torch.from_numpy(U).to(device).to(dtype).contiguous()
        srct = torch.from_numpy(src).to(device)
        maskt = None
        if self.mask is not None:
            if self.mask.shape != (nk1, nk2):
                raise ValueError(f"mask of shape {self.mask.shape} does not match the mesh ({nk1}, {nk2})")
            maskt = torch.from_numpy(self.mask.astype(np.uint8).reshape(-1)).to(device)
        bits = 0
        for g, (_, _, cj) in enumerate(self.elems):
            if cj:
                bits |= 1 << g
        t = (Ut, srct, maskt, bits)
        self._dev[key] = t
        return t

    def apply(self, A, out):
        nk1, nk2, nb = A.shape[0], A.shape[1], A.shape[-1]
        Ut, srct, maskt, bits = self._tables(nk1, nk2, nb, A.dtype, A.device)
        L.call("hfx_symmetry_average", dt_code(A), nb, nk1 * nk2, len(self.elems), L.ptr(A), L.ptr(Ut), L.ptr(srct),
               bits, L.ptr(maskt), 1.0 / len(self.elems), L.ptr(out), L.stream(A.device))
        return out


def make_project_fn(*, unitary_group=None, spatial_group=None, spatial_k_convention="mod",
                    spatial_k_flip_axes=(0, 1), spatial_k_index_maps=None, spatial_k_orbit_mask=None,
                    time_reversal_U=None, time_reversal_k_convention="mod", time_reversal_k_flip_axes=(0, 1),
                    time_reversal_k_index_map=None):
    """Symmetry-averaging projection (symmetry.py:153-281), same arguments.

    Returns ``project(A)`` for device tensors ``(nk1, nk2, nb, nb)`` (complex64
    or complex128; the result has A's dtype and device, A is not modified).
    """
    has_group = unitary_group is not None or spatial_group is not None or spatial_k_index_maps is not None
    if not has_group and time_reversal_U is None:           # symmetry.py:205-206
        def identity(A):
            return A
        identity.hfx_graph_safe = True
        return identity
    G = None if unitary_group is None else host_array(unitary_group).astype(np.complex128)
    S = None if spatial_group is None else host_array(spatial_group).astype(np.complex128)
    U = None if time_reversal_U is None else host_array(time_reversal_U).astype(np.complex128)
    if spatial_k_index_maps is not None:                     # symmetry.py:220-237
        si = host_array(spatial_k_index_maps[0]).astype(np.int64)
        sj = host_array(spatial_k_index_maps[1]).astype(np.int64)
        if si.shape != sj.shape:
            raise ValueError("spatial_k_index_maps[0] and [1] must have the same shape; "
                             f"got {si.shape} vs {sj.shape}.")
        if S is None:
            raise ValueError("spatial_k_index_maps requires spatial_group (the orbital unitaries to average over).")
        if S.shape[0] != si.shape[0]:
            raise ValueError("spatial_group and spatial_k_index_maps must have matching "
                             f"leading dim; got {S.shape[0]} and {si.shape[0]}.")
    if time_reversal_k_index_map is not None and U is None:   # symmetry.py:246-249
        raise ValueError("time_reversal_k_index_map requires time_reversal_U.")
    # conventions are checked when a stage that uses one builds its tables
    # (the reference raises in _flip_k, at the first call)
    s_conv, s_axes = str(spatial_k_convention), tuple(spatial_k_flip_axes)
    t_conv, t_axes = str(time_reversal_k_convention), tuple(time_reversal_k_flip_axes)

    stages = []
    if spatial_k_index_maps is not None:                     # symmetry.py:258-266
        kmap = ("maps", si, sj)
        stages.append(_Stage([(S[g], kmap, False) for g in range(S.shape[0])], mask=spatial_k_orbit_mask))
        if G is not None:
            stages.append(_Stage([(g, None, False) for g in G]))
    elif G is not None or S is not None:                     # symmetry.py:267-274
        same = G if G is not None else np.eye(S.shape[-1], dtype=np.complex128)[None]
        elems = [(g, None, False) for g in same]
        if S is not None:
            elems += [(s, ("conv", s_conv, s_axes), False) for s in S]
        stages.append(_Stage(elems))
    if U is not None:                                        # symmetry.py:275-279
        eye = np.eye(U.shape[-1], dtype=np.complex128)
        if time_reversal_k_index_map is not None:
            ti = host_array(time_reversal_k_index_map[0]).astype(np.int64)
            tj = host_array(time_reversal_k_index_map[1]).astype(np.int64)
            tv = host_array(time_reversal_k_index_map[2]).astype(bool)
            kmap = ("maps", np.stack([ti, ti]), np.stack([tj, tj]))
            stages.append(_Stage([(eye, None, False), (U, kmap, True)], mask=tv))
        else:
            stages.append(_Stage([(eye, None, False), (U, ("conv", t_conv, t_axes), True)]))

    def project(A):
        if not isinstance(A, torch.Tensor) or not A.is_cuda:
            raise TypeError("make_project_fn: the projector takes device tensors (use to_tensor first)")
        A = A.contiguous()
        x = A
        for st in stages:
            out = torch.empty_like(x)
            st.apply(x, out)
            x = out
        return x

    project.hfx_graph_safe = True
    return project
