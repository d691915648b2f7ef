"""Physics utilities on the GPU (drop-in for the hot-path parts of utils.py).

Mirrors ``src/jax_hf/utils.py``: ``validate_electron_count`` (:21-43, host
check), ``hermitize`` (:46-48), ``fermidirac`` (:51-53), ``selfenergy_fft``
(:66-137), ``find_chemical_potential`` (:314-442, bisection or safeguarded
Newton, one cooperative launch), ``density_matrix_from_fock`` (:445-474) and
``resample_kgrid`` (:477-535, one launch of ``hfx_resample_kgrid``).
"""

from __future__ import annotations

from typing import Any

import numpy as np
import torch

from . import _lib as L
from . import _ops as K
from ._util import convert, host_array, to_tensor

__all__ = ["validate_electron_count", "hermitize", "fermidirac", "selfenergy_fft",
           "find_chemical_potential", "density_matrix_from_fock", "resample_kgrid"]


def validate_electron_count(weights: Any, nbands: int, n_electrons: Any, *, context: str = "n_electrons") -> None:
    """utils.py:21-43."""
    _check_count(float(np.real(host_array(weights)).sum()), nbands, n_electrons, context)


def validate_kernel_electron_count(kernel, n_electrons: Any, *, context: str = "n_electrons") -> None:
    """validate_electron_count for a HartreeFockKernel, from the weight sum it
    keeps on the host (no device read)."""
    _check_count(float(kernel.weight_sum), kernel.h.shape[-1], n_electrons, context)


def _check_count(total_weight: float, nbands: int, n_electrons: Any, context: str) -> None:
    target = float(np.asarray(host_array(n_electrons)))
    max_electrons = total_weight * int(nbands)
    tol = max(1e-8, 1e-6 * max(1.0, abs(max_electrons)))
    if not np.isfinite(target):
        raise ValueError(f"{context} must be finite, got {target!r}.")
    if target < -tol or target > max_electrons + tol:
        raise ValueError(f"{context}={target} is outside the physically reachable range "
                         f"[0, {max_electrons}] for {int(nbands)} bands and weights.sum()={total_weight}.")


def hermitize(X):
    """0.5 (X + X^dagger) on the last two axes (utils.py:46-48)."""
    X = to_tensor(X)
    if not X.is_complex():
        X = convert(X, torch.complex128 if X.dtype == torch.float64 else torch.complex64)
    Y = torch.empty_like(X)
    K.hermitize(X, Y)
    return Y


def fermidirac(x, T: float):
    """1 / (1 + exp(x / (T + 1e-12))) (utils.py:51-53), evaluated on the device in FP64."""
    x = to_tensor(x)
    xt = convert(x, torch.float64)
    out = torch.empty_like(xt)
    L.call("hfx_fermi", L.ptr(xt), None, float(T), xt.numel(), L.ptr(out), L.stream(xt.device))
    return out


def selfenergy_fft(VR, P, *, block_specs: Any | None = None, check_offdiag: bool | None = None,
                   offdiag_atol: float = 1e-12, offdiag_rtol: float = 0.0, _apply_ifftshift: bool = True,
                   hermitian_channel_packing: bool = False):
    """Sigma(k) = -IFFT2[FFT2(P) VR] (utils.py:66-137), optional ifftshift."""
    from .fock import FockEngine, block_channels
    P = to_tensor(P)
    cdt = P.dtype if P.is_complex() else torch.complex128
    P = convert(P, cdt)
    VRt = to_tensor(VR, P.device, cdt)
    if hermitian_channel_packing and VRt.shape[-2:] != (1, 1):
        raise ValueError("hermitian_channel_packing requires a scalar interaction kernel with shape (..., 1, 1).")
    nk1, nk2, nb = P.shape[0], P.shape[1], P.shape[-1]
    eng = FockEngine(h=P, VR=VRt, refP=None, HH=None, w2d=np.ones((nk1, nk2)), include_exchange=True,
                     include_hartree=False, herm=hermitian_channel_packing)
    out = torch.empty_like(P)
    chans, nchan = None, 0
    if block_specs is not None:
        sel = block_channels(P, block_specs, check_offdiag, offdiag_atol, offdiag_rtol,
                             hermitian_channel_packing, eng.plan)
        if sel is not None:
            chans, nchan = sel
            out.zero_()
    eng.selfenergy(P, out, chans=chans, nchan=nchan, rho=False, ifftshift=_apply_ifftshift)
    return out


def _mu_target_weights(bands, weights):
    nb = bands.shape[-1]
    nk = bands.numel() // nb
    w = to_tensor(weights, bands.device, torch.float64).reshape(-1)
    if w.numel() != nk:
        raise ValueError("weights must have one entry per k-point of bands")
    return nk, nb, w


def find_chemical_potential(bands, weights, n_electrons: float, T: float, *, maxiter: int | None = None,
                            method: str = "bisection"):
    """mu with sum_k w_k sum_j f(e_kj - mu) = n_electrons (utils.py:314-350)."""
    b = to_tensor(bands)
    single = b.dtype in (torch.float32, torch.complex64)
    validate_electron_count(weights, b.shape[-1], n_electrons, context="n_electrons")
    method = str(method).lower()
    if method not in ("bisection", "newton"):
        raise ValueError(f"method must be 'bisection' or 'newton', got {method!r}")
    b64 = convert(b, torch.float64)
    nk, nb, w = _mu_target_weights(b64, weights)
    ws = K.Workspace.get(b64.device, nk)
    mu = torch.zeros((), dtype=torch.float64, device=b64.device)
    if method == "bisection":
        it = maxiter if maxiter is not None else (30 if single else 54)
        K.solve_mu(ws, b64.reshape(-1).contiguous(), w, nk, nb, float(n_electrons), float(T), maxiter=it,
                   mode=K.MU_BISECTION, mu_out_ptr=mu.data_ptr())
    else:
        it = maxiter if maxiter is not None else (15 if single else 25)
        K.solve_mu(ws, b64.reshape(-1).contiguous(), w, nk, nb, float(n_electrons), float(T), maxiter=it,
                   mode=K.MU_NEWTON, mu_out_ptr=mu.data_ptr())
    return mu


def occupations_from_fock(F, weights, n_electrons, T, level_shift=0.0, single=False, herm_in=True):
    """(eps, U, mu, occ) for density_matrix_from_fock / the SCF step (reference_scf.py:87-96)."""
    F = to_tensor(F)
    nk1, nk2, nb = F.shape[0], F.shape[1], F.shape[-1]
    nk = nk1 * nk2
    eps = torch.empty((nk1, nk2, nb), dtype=torch.float64, device=F.device)
    U = torch.empty_like(F)
    K.eigh(F, eps, U, herm_in=herm_in)
    w = to_tensor(weights, F.device, torch.float64).reshape(-1)
    ws = K.Workspace.get(F.device, nk)
    occ = torch.empty_like(eps)
    mu = torch.zeros((), dtype=torch.float64, device=F.device)
    K.solve_mu(ws, eps, w, nk, nb, float(n_electrons), float(T), maxiter=30 if single else 54,
               mode=K.MU_BISECTION, level_shift=float(level_shift), mu_out_ptr=mu.data_ptr(), p_out=occ)
    return eps, U, mu, occ


def density_matrix_from_fock(F, weights, n_electrons: float, T: float, *, eigh_block_specs=None,
                             eigh_check_offdiag=None, eigh_offdiag_atol: float = 1e-12,
                             eigh_offdiag_rtol: float = 0.0):
    """P = herm(U diag(f) U^dagger) from a Fock matrix (utils.py:445-474)."""
    F = to_tensor(F)
    if not F.is_complex():
        F = convert(F, torch.complex128 if F.dtype == torch.float64 else torch.complex64)
    validate_electron_count(weights, F.shape[-1], n_electrons, context="n_electrons")
    single = F.dtype == torch.complex64
    _, U, mu, occ = occupations_from_fock(F, weights, n_electrons, T, single=single)
    P = torch.empty_like(F)
    K.density(U, P, p=occ)
    return P, mu


_RESAMPLE_KIND = {torch.float64: 0, torch.float32: 1, torch.complex128: 2, torch.complex64: 3}


def resample_kgrid(values, nk: int, *, method: str = "linear"):
    """Resample a centred, periodic (n0, n1, ...) k-grid array to (nk, nk, ...)
    (utils.py:512-535): periodic linear interpolation along axis 0, then axis 1,
    with k = 0 at index n // 2.  Returns ``values`` itself when both axes
    already have length nk.  Device tensors of float64 / float32 / complex128 /
    complex64; other array-likes are uploaded first."""
    nk = int(nk)
    x = to_tensor(values)
    if x.dim() < 2:
        raise ValueError("resample_kgrid expects an array with at least 2 dimensions (nk,nk,...).")
    if x.shape[0] == nk and x.shape[1] == nk:
        return x
    method = str(method).lower()
    if method != "linear":
        raise ValueError(f"Unsupported resample method {method!r}. Only 'linear' is supported.")
    if nk <= 0:
        raise ValueError("n_new must be a positive integer.")
    kind = _RESAMPLE_KIND.get(x.dtype)
    if kind is None:
        raise TypeError(f"resample_kgrid: unsupported dtype {x.dtype}")
    m = 1
    for d in x.shape[2:]:
        m *= int(d)
    out = torch.empty((nk, nk) + tuple(x.shape[2:]), dtype=x.dtype, device=x.device)
    L.call("hfx_resample_kgrid", kind, int(x.shape[0]), int(x.shape[1]), m, nk, L.ptr(x), L.ptr(out),
           L.stream(x.device))
    return out
