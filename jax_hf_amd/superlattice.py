"""Superlattice Fock build on the GPU (drop-in for src/jax_hf/superlattice.py).

For a plane-wave basis indexed by superlattice reciprocal vectors G, the Fock
self-energy is a 2D convolution in absolute momentum p = k - G, one
independent convolution per (dG = G_a - G_b, alpha, beta) channel
(superlattice.py:1-17).  ``build_extended_layout`` makes the reference's index
tables on the host and the FFT of the Cartesian Coulomb kernel on the lag grid
on the GPU; the returned functions run ``hfx_superlattice_fock``: scatter of
every channel block onto the extended grid, one planar exchange launch over
all n_delta * dim_orb^2 planes (the FFT convolution of the Fock build), and a gather that adds the k-independent Hartree shift and forms
F = herm(h + Sigma).  The reference streams one dG channel at a time to bound
device memory at O(N_ext^2 dim_orb^2); with 288 GB of HBM all channels fit at
once, and one launch over all planes fills the GPU (extended grids beyond the
LDS plane run the HBM-staged exchange).
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass
from typing import Any, Callable

import numpy as np
import torch

from . import _lib as L
from ._util import default_device, host_array, to_tensor

__all__ = ["ExtendedGridLayout", "build_extended_layout", "make_superlattice_fock_fn",
           "make_superlattice_build_fock_fn"]


@dataclass(frozen=True)
class ExtendedGridLayout:
    """Extended-grid indexing data (superlattice.py:31-72), same fields;
    ``V_lag_fft`` is a complex128 device tensor."""

    N_ext_x: int
    N_ext_y: int
    n_supp_x: int
    n_supp_y: int
    g_a_off: np.ndarray
    n_delta: int
    pair_to_delta: np.ndarray
    V_lag_fft: Any
    delta_pair_count: np.ndarray
    delta_pair_i: np.ndarray
    delta_pair_j: np.ndarray
    delta_pair_start: np.ndarray


def build_extended_layout(g_basis_fractional, g_basis_B, nkx: int, nky: int,
                          coulomb_V: Callable[[np.ndarray], np.ndarray], w_scalar: float, *,
                          device=None) -> ExtendedGridLayout:
    """superlattice.py:75-164: tile offsets of -G, the unique dG list with its
    CSR pair tables, and fft2(w_scalar * V(|lag|)) on the FFT-natural lag grid
    (computed on the GPU by the Fock build's FFT)."""
    g_frac = np.asarray(g_basis_fractional, dtype=np.int64)
    n_G = int(g_frac.shape[0])
    pg = -g_frac
    gx0, gy0 = int(pg[:, 0].min()), int(pg[:, 1].min())
    G_dim_x = int(pg[:, 0].max()) - gx0 + 1
    G_dim_y = int(pg[:, 1].max()) - gy0 + 1
    n_supp_x, n_supp_y = nkx * G_dim_x, nky * G_dim_y
    Nx, Ny = 2 * n_supp_x, 2 * n_supp_y
    g_a_off = np.stack([pg[:, 0] - gx0, pg[:, 1] - gy0], axis=1).astype(np.int64)

    # unique dG in first-seen row-major (i, j) order, as the reference numbers them
    dg = (g_frac[:, None, :] - g_frac[None, :, :]).reshape(-1, 2)
    index: dict = {}
    p2d = np.empty(n_G * n_G, dtype=np.int64)
    for s, v in enumerate(map(tuple, dg)):
        p2d[s] = index.setdefault(v, len(index))
    n_delta = len(index)
    pair_to_delta = p2d.reshape(n_G, n_G)
    count = np.bincount(p2d, minlength=n_delta).astype(np.int64)
    start = np.zeros(n_delta + 1, dtype=np.int64)
    start[1:] = np.cumsum(count)
    order = np.argsort(p2d, kind="stable")                 # pairs grouped by dG, row-major within
    delta_i, delta_j = (order // n_G).astype(np.int64), (order % n_G).astype(np.int64)

    ix, iy = np.arange(Nx), np.arange(Ny)
    fx = np.where(ix < Nx // 2, ix, ix - Nx).astype(float) / float(nkx)
    fy = np.where(iy < Ny // 2, iy, iy - Ny).astype(float) / float(nky)
    FX, FY = np.meshgrid(fx, fy, indexing="ij")
    B = np.asarray(g_basis_B, dtype=float)
    qmag = np.sqrt((B[0, 0] * FX + B[0, 1] * FY) ** 2 + (B[1, 0] * FX + B[1, 1] * FY) ** 2)
    V_lag = np.asarray(coulomb_V(qmag), dtype=float)

    dev = torch.device(device) if device is not None else default_device()
    L.require_gpu(dev)
    Vq = torch.from_numpy(np.ascontiguousarray(V_lag.astype(np.complex128).reshape(Nx, Ny, 1, 1))).to(dev)
    w = torch.full((Nx, Ny), float(w_scalar), dtype=torch.float64, device=dev)
    plan = L.plan_for(Nx, Ny, 1, torch.complex128, dev)
    Vf = torch.empty_like(Vq)
    L.call("hfx_prepare_vr", plan.handle, L.ptr(Vq), 1, L.ptr(w), 0, L.ptr(Vf), L.stream(dev))
    return ExtendedGridLayout(N_ext_x=Nx, N_ext_y=Ny, n_supp_x=n_supp_x, n_supp_y=n_supp_y, g_a_off=g_a_off,
                              n_delta=n_delta, pair_to_delta=pair_to_delta, V_lag_fft=Vf.reshape(Nx, Ny),
                              delta_pair_count=count, delta_pair_i=delta_i, delta_pair_j=delta_j,
                              delta_pair_start=start)


class _SLFock:
    """Device tables and buffers of one (layout, n_G, dim_orb, mesh); one
    ``hfx_superlattice_fock`` call per build."""

    def __init__(self, layout: ExtendedGridLayout, n_G: int, dim_orb: int, nkx: int, nky: int,
                 HH_orb: np.ndarray | None = None, degeneracy: float = 0.0):
        self.layout, self.n_G, self.dorb, self.nkx, self.nky = layout, int(n_G), int(dim_orb), int(nkx), int(nky)
        self.D = self.n_G * self.dorb
        Nx, Ny = int(layout.N_ext_x), int(layout.N_ext_y)
        if Nx % self.nkx or Ny % self.nky:
            raise ValueError(f"extended grid ({Nx}, {Ny}) is not a multiple of the mesh ({nkx}, {nky})")
        goff = np.asarray(layout.g_a_off, dtype=np.int64)
        p2d = np.asarray(layout.pair_to_delta, dtype=np.int64)
        if goff.shape != (self.n_G, 2) or p2d.shape != (self.n_G, self.n_G):
            raise ValueError(f"layout was built for {goff.shape[0]} G vectors, not n_G={n_G}")
        tx, ty = Nx // self.nkx, Ny // self.nky
        if goff.min() < 0 or goff[:, 0].max() >= tx or goff[:, 1].max() >= ty:
            raise ValueError("layout tile offsets fall outside the extended grid")
        di, dj = np.asarray(layout.delta_pair_i), np.asarray(layout.delta_pair_j)
        if di.shape != (self.n_G * self.n_G,) or dj.shape != di.shape or p2d.min() < 0 or \
                p2d.max() >= int(layout.n_delta):
            raise ValueError("layout pair tables do not cover the n_G x n_G pairs")
        self.host = dict(p2d=p2d.astype(np.int32), goff=goff.astype(np.int32),
                         start=np.asarray(layout.delta_pair_start, np.int32),
                         di=np.asarray(layout.delta_pair_i, np.int32), dj=np.asarray(layout.delta_pair_j, np.int32))
        self.HH_orb = None if HH_orb is None else np.array(HH_orb, dtype=np.float64, order="C", copy=True)
        self.deg = float(degeneracy)
        self._dev = {}
        L.register_owner(self)

    def _hfx_release(self):
        self._dev.clear()

    def _state(self, dtype, device):
        key = (dtype, str(device))
        st = self._dev.get(key)
        if st is None:
            st = {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in self.host.items()}
            lay = self.layout
            Nx, Ny = int(lay.N_ext_x), int(lay.N_ext_y)
            st["V"] = to_tensor(lay.V_lag_fft, device, dtype).reshape(Nx, Ny).contiguous()
            st["plan"] = L.plan_for(Nx, Ny, self.dorb, dtype, device)
            npl = (int(lay.n_delta), self.dorb * self.dorb, Nx, Ny)
            st["planes"] = torch.zeros(npl, dtype=dtype, device=device)     # zero outside the pair blocks
            st["splanes"] = torch.empty(npl, dtype=dtype, device=device)
            st["R"] = torch.empty(int(lay.n_delta) * self.dorb, dtype=torch.complex128, device=device)
            if self.HH_orb is not None:
                st["HH"] = torch.from_numpy(self.HH_orb).to(device)
            self._dev[key] = st
        return st

    def _per_k(self, X, name, P):
        """X as a contiguous (nkx, nky, D, D) tensor of P's dtype on P's device;
        a k-independent (D, D) matrix is broadcast over k (as ``h + Sigma``
        broadcasts in superlattice.py:370)."""
        X = to_tensor(X, P.device, P.dtype)
        D = self.D
        if tuple(X.shape) == (D, D):
            X = X.expand(self.nkx, self.nky, D, D)
        if tuple(X.shape) != tuple(P.shape):
            raise ValueError(f"superlattice Fock: {name} of shape {tuple(X.shape)}, expected {tuple(P.shape)} "
                             f"or ({D}, {D})")
        return X.contiguous()

    def __call__(self, P, h, refP=None, w2d=None, exchange=True, hartree=False):
        if P.dtype not in (torch.complex128, torch.complex64):
            raise TypeError(f"superlattice Fock: P must be complex128 or complex64, got {P.dtype}")
        shape = (self.nkx, self.nky, self.D, self.D)
        if tuple(P.shape) != shape:
            raise ValueError(f"superlattice Fock: density of shape {tuple(P.shape)}, expected {shape}")
        if not P.is_contiguous():
            P = P.contiguous()
        h = self._per_k(h, "h", P)
        refP = None if refP is None else self._per_k(refP, "refP", P)
        w = None
        if hartree:
            if w2d is None:
                raise ValueError("superlattice Fock: the Hartree shift needs w2d")
            w = to_tensor(w2d, P.device, torch.float64).reshape(-1).contiguous()
            if w.numel() != self.nkx * self.nky:
                raise ValueError(f"superlattice Fock: w2d has {w.numel()} entries, expected "
                                 f"{self.nkx * self.nky} ({self.nkx} x {self.nky})")
        st = self._state(P.dtype, P.device)
        dt = L.HFX_C128 if P.dtype == torch.complex128 else L.HFX_C64
        Sigma, F = torch.empty_like(P), torch.empty_like(P)
        lay = self.layout
        L.call("hfx_superlattice_fock", dt, self.nkx, self.nky, self.n_G, self.dorb, int(lay.n_delta),
               int(lay.N_ext_x), int(lay.N_ext_y), L.ptr(P), L.ptr(refP), L.ptr(st["p2d"]), L.ptr(st["goff"]), st["plan"].handle, L.ptr(st["V"]),
               L.ptr(st["planes"]) if exchange else None, L.ptr(st["splanes"]), int(bool(hartree)), L.ptr(w),
               L.ptr(st["start"]), L.ptr(st["di"]), L.ptr(st["dj"]), L.ptr(st.get("HH")), self.deg,
               L.ptr(st["R"]), L.ptr(h), L.ptr(Sigma), L.ptr(F), L.stream(P.device))
        return Sigma, F


def make_superlattice_fock_fn(layout: ExtendedGridLayout, n_G: int, dim_orb: int, nkx: int, nky: int):
    """superlattice.py:167-230: rho (nkx, nky, n_G dim_orb, n_G dim_orb) -> the
    exchange self-energy of the same shape (complex128 or complex64 device
    tensors; other array-likes are uploaded)."""
    eng = _SLFock(layout, n_G, dim_orb, nkx, nky)
    zeros = {}

    def fock_fn(rho_kk):
        P = to_tensor(rho_kk)
        key = (P.dtype, str(P.device))
        if key not in zeros:
            zeros[key] = torch.zeros((nkx, nky, eng.D, eng.D), dtype=P.dtype, device=P.device)
        return eng(P, zeros[key], exchange=True, hartree=False)[0]

    return fock_fn


def make_superlattice_build_fock_fn(layout: ExtendedGridLayout, n_G: int, dim_orb: int, nkx: int, nky: int,
                                    HH_GG, hartree_degeneracy: float, HH_GG_orbital=None):
    """superlattice.py:233-377: a callable with ``build_fock``'s keyword
    signature for the ``fock_build_fn`` seam of ``solve`` / ``solve_scf``.
    Sigma = streaming exchange of P - refP + the k-independent superlattice
    Hartree shift; H is zero; F = herm(h + Sigma), then herm(project_fn(F)).
    The CG loop of ``solve`` captures it in its iteration graph
    (``hfx_graph_safe``; the first iteration runs eagerly)."""
    n_G, dim_orb = int(n_G), int(dim_orb)
    if HH_GG_orbital is not None:
        HH_orb = np.asarray(host_array(HH_GG_orbital), dtype=np.float64)
        expected = (n_G, n_G, dim_orb, dim_orb)
        if HH_orb.shape != expected:
            raise ValueError(f"HH_GG_orbital must have shape {expected}; got {HH_orb.shape}")
        if float(np.max(np.abs(HH_orb[np.arange(n_G), np.arange(n_G)]))) > 1e-12:
            warnings.warn("HH_GG_orbital has nonzero diagonal-G blocks; the q=0 layer Hartree will be included.  "
                          "This is intended for a layer-resolved system at finite displacement field; for the "
                          "moire/TBG convention pass HH_GG_orbital[i, i, :, :] = 0.", UserWarning, stacklevel=2)
    else:
        # the scalar formula traces the orbital before applying HH: the orbital
        # table with every (alpha, gamma) entry equal to HH_GG[A, B]
        HH = np.asarray(host_array(HH_GG), dtype=np.float64)
        HH_orb = np.broadcast_to(HH[:, :, None, None], (n_G, n_G, dim_orb, dim_orb))
    eng = _SLFock(layout, n_G, dim_orb, nkx, nky, HH_orb=HH_orb, degeneracy=float(hartree_degeneracy))

    def build_fock_sl(P, *, h, VR=None, refP=None, HH=None, w2d=None, include_exchange=True, include_hartree=False,
                      exchange_hermitian_channel_packing=False, contact_g=None, contact_Oi=None, contact_Oj=None,
                      project_fn=None, exchange_block_specs=None, exchange_check_offdiag=None,
                      exchange_offdiag_atol=1e-12, exchange_offdiag_rtol=0.0):
        P = to_tensor(P)
        h = to_tensor(h, P.device, P.dtype)
        ref = None if refP is None else to_tensor(refP, P.device, P.dtype)
        Sigma, F = eng(P, h, refP=ref, w2d=w2d, exchange=bool(include_exchange), hartree=bool(include_hartree))
        if project_fn is not None:
            from . import _ops as K
            X = to_tensor(project_fn(F), P.device, P.dtype)
            F = torch.empty_like(X)
            K.hermitize(X, F)
        return Sigma, torch.zeros_like(P), F

    # after its first call (device tables, plan, planes) a build only allocates
    # its outputs and enqueues launches, so a solver may capture it in a graph
    build_fock_sl.hfx_graph_safe = True
    return build_fock_sl
