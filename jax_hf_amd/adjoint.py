"""Hand-written adjoint of build_fock + hf_energy and converged-state parameter gradients.

The reference obtains derivatives by JAX autodiff through ``fock.py:18-111``;
here the backward pass is explicit and runs on the same HIP kernels:

    E(P, h, VR) = sum_k w_k Re Tr[(h + (Sigma + H)/2) P],  Sigma = A[P - refP],
    A_c = IFFT2 diag(-VR_c) FFT2 (one exchange channel c = (i, j)).

With torch's convention for complex gradients (grad = dE/dRe + i dE/dIm):

    grad_h  = w P^dagger
    grad_P  = w (h + Sigma/2 + H)^dagger + A^H[w P^dagger / 2]
              (+ the Hartree term w/2 diag(HH n + HH^T m), n, m the weighted diagonals)
    grad_VR = -conj(FFT2(P - refP)) FFT2(w P^dagger / 2) / N      (summed over channels
              for a scalar VR)

``A^H[G] = -IFFT2[conj(VR) FFT2(G)]`` is the cross-correlation: it is recomputed
by the fused exchange kernel with a conjugated VR (``hfx_selfenergy_adjoint``),
nothing from the forward pass is stored but (P, h, VR).  At a converged
(stationary) state, the free-energy gradient with respect to any parameter
theta of h or V is the explicit derivative dOmega/dtheta = <grad_h, dh/dtheta>
+ <grad_VR, dVR/dtheta> (Hellmann-Feynman), see ``free_energy_gradients``.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from . import _ops as K
from ._util import KIND, convert, to_tensor

__all__ = ["HFEnergy", "hf_energy_autograd", "free_energy_gradients", "vr_from_vq"]


def _scale_k(X, w, alpha):
    nb2 = X.shape[-1] * X.shape[-2]
    out = torch.empty_like(X)
    L.call("hfx_scale_k", X.numel() // nb2, nb2, KIND[X.dtype], L.ptr(X), L.ptr(w), float(alpha), L.ptr(out),
           L.stream(X.device))
    return out


class HFEnergy(torch.autograd.Function):
    """E = hf_energy(P, build_fock(P)) with a hand-written backward (exchange + Hartree).

    Non-tensor argument ``spec``: dict with ``w2d`` (float64 (nk1, nk2) device),
    ``refP`` (tensor or None), ``HH`` (float64 (nb, nb) or None).
    """

    @staticmethod
    def forward(ctx, P, h, VR, spec):
        P = P.detach().contiguous()
        h = h.detach().contiguous()
        VR = VR.detach().contiguous()
        refP, HH, w = spec.get("refP"), spec.get("HH"), spec["w2d"]
        eng = _forward_engine(h, VR, refP, HH, w)
        F = torch.empty_like(P)
        Sig = torch.empty_like(P)
        eng.build(P, F, Sigma_out=Sig, energy=True)
        ctx.save_for_backward(P, h, VR, Sig)
        ctx.spec = spec
        ctx.Hdiag = eng.Hdiag.clone() if eng.Hdiag is not None else None
        return eng.E.clone()

    @staticmethod
    def backward(ctx, gE):
        P, h, VR, Sig = ctx.saved_tensors
        spec = ctx.spec
        w = spec["w2d"]
        refP, HH = spec.get("refP"), spec.get("HH")
        # the incoming gradient stays on the device (no host round trip): the
        # scaling kernels read it as a device scalar
        g = gE.detach().to(device=P.device, dtype=torch.float64).reshape(1).contiguous()
        nk1, nk2, nb = P.shape[0], P.shape[1], P.shape[-1]
        nkt, kind = nk1 * nk2, KIND[P.dtype]
        plan = L.plan_for(nk1, nk2, nb, P.dtype, P.device)
        st = L.stream(P.device)
        wp = L.ptr(w)

        def scaled(X, alpha, gdev, dagger, out=None):
            out = torch.empty_like(X) if out is None else out
            L.call("hfx_scale_k_ex", nkt, X.shape[-1], kind, L.ptr(X), wp if dagger else None, float(alpha),
                   L.ptr(gdev), int(dagger), L.ptr(out), st)
            return out

        grad_h = scaled(P, 1.0, g, True) if ctx.needs_input_grad[1] else None      # g w P^dagger
        grad_P = grad_VR = None
        G = scaled(P, 0.5, None, True)                     # seed of the exchange term: w P^dagger / 2
        if ctx.needs_input_grad[0]:
            X = torch.empty_like(P)
            K.lincomb(X, h, 1.0, Sig, 0.5)               # h + Sigma/2
            if ctx.Hdiag is not None:
                X = _add_diag(X, ctx.Hdiag, 0.5)
            gP = scaled(X, 1.0, None, True)
            AH = torch.empty_like(P)
            vfull = 1 if VR.shape[-1] != 1 else 0
            L.call("hfx_selfenergy_adjoint", plan.handle, L.ptr(G), L.ptr(VR), vfull, L.ptr(AH), st)
            K.lincomb(gP, gP, 1.0, AH, 1.0)
            if HH is not None:
                gP = _add_hartree_grad(gP, P, refP, HH, w)
            grad_P = scaled(gP, 1.0, g, False, out=gP)
        if ctx.needs_input_grad[2]:
            dP = P if refP is None else _sub(P, refP)
            w1 = _ones_like_cached(w)
            Pf = torch.empty_like(P)
            Gf = torch.empty_like(P)
            L.call("hfx_fft2_weighted", plan.handle, L.ptr(dP), L.ptr(w1), L.ptr(Pf), st)
            L.call("hfx_fft2_weighted", plan.handle, L.ptr(G), L.ptr(w1), L.ptr(Gf), st)
            N = nk1 * nk2
            scalar = VR.shape[-1] == 1
            grad_VR = torch.empty(VR.shape, dtype=VR.dtype, device=VR.device)
            L.call("hfx_cmul", 0 if P.dtype == torch.complex128 else 1, nk1 * nk2, nb * nb, L.ptr(Pf), L.ptr(Gf),
                   -1.0 / N, 0.0, int(scalar), L.ptr(grad_VR), st)
            grad_VR = scaled(grad_VR, 1.0, g, False, out=grad_VR)
        return grad_P, grad_h, grad_VR, None


# One forward engine, reused while the inputs are the same storage at the same
# version (the packed channel-major VR and the build buffers are made once).
_FWD = {"key": None, "eng": None}
_ONES: dict = {}
L.register_cache(lambda: (_FWD.update(key=None, eng=None), _ONES.clear()))


def _tkey(t):
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        return (t.data_ptr(), t._version, tuple(t.shape), t.dtype, str(t.device))
    return ("host", id(t))


def _forward_engine(h, VR, refP, HH, w):
    from .fock import FockEngine
    key = (_tkey(h), _tkey(VR), _tkey(refP), _tkey(HH), _tkey(w))
    if _FWD["key"] != key:
        _FWD["eng"] = FockEngine(h=h, VR=VR, refP=refP, HH=HH, w2d=w, include_exchange=True,
                                 include_hartree=HH is not None, herm=False)
        _FWD["key"] = key
    return _FWD["eng"]


def _ones_like_cached(w):
    k = (tuple(w.shape), str(w.device))
    o = _ONES.get(k)
    if o is None:
        o = _ONES[k] = torch.ones_like(w)
    return o


def _sub(A, B):
    out = torch.empty_like(A)
    K.lincomb(out, A, 1.0, B, -1.0)
    return out


def _add_diag(X, d, alpha):
    """X + alpha diag(d) broadcast over k (d float64 (nb,))."""
    nb = X.shape[-1]
    D = torch.zeros((nb, nb), dtype=torch.float64, device=X.device)
    D.view(-1)[:: nb + 1].copy_(d)
    Dc = convert(D, X.dtype).expand(X.shape).contiguous()
    out = torch.empty_like(X)
    K.lincomb(out, X, 1.0, Dc, alpha)
    return out


def _add_hartree_grad(gP, P, refP, HH, w):
    """+ w_k/2 diag(HH^T m), m = sum_k w Re diag(P): the part of dE_H/dP through n.

    E_H = (1/2) sum_a (HH n)_a m_a with n = sum_k w Re diag(P - refP); the
    (HH n)/2 part is already in w (h + Sigma/2 + H/2)^dagger.
    """
    from .fock import FockEngine
    HH64 = to_tensor(HH, P.device, torch.float64)
    eng = FockEngine(h=P, VR=None, refP=refP, HH=HH64, w2d=w, include_exchange=False, include_hartree=True,
                     herm=False)
    L.call("hfx_density_moments", eng.plan.handle, L.ptr(P), None, L.ptr(eng.w2d), L.ptr(eng.rho),
           L.stream(P.device))
    m = np.real(np.diagonal(eng.rho.cpu().numpy()))
    H = HH64.cpu().numpy()
    d = torch.from_numpy(H.T @ m).to(P.device)
    nb = P.shape[-1]
    D = torch.zeros((nb, nb), dtype=torch.float64, device=P.device)
    D.view(-1)[:: nb + 1].copy_(d)
    Dk = _scale_k(convert(D, P.dtype).expand(P.shape).contiguous(), w, 0.5)
    out = torch.empty_like(gP)
    K.lincomb(out, gP, 1.0, Dk, 1.0)
    return out


def hf_energy_autograd(P, h, VR, w2d, refP=None, HH=None):
    """Differentiable E(P, h, VR) (exchange + optional Hartree) on the GPU."""
    w = to_tensor(w2d, P.device, torch.float64)
    return HFEnergy.apply(P, h, VR, dict(w2d=w, refP=refP, HH=HH))


def vr_from_vq(Vq, w2d, device, dtype=torch.complex128):
    """Phase-absorbed VR of a (nk1, nk2, vb, vb) interaction (problem.py:88-98)."""
    from .problem import _prepare_vr
    Vt = to_tensor(np.ascontiguousarray(np.asarray(Vq).astype(np.complex128)), device, torch.complex128)
    w = to_tensor(w2d, device, torch.float64)
    return convert(_prepare_vr(Vt, w, phase=True), dtype)


def free_energy_gradients(kernel, density, *, dVq=None, dh=None):
    """dOmega/dtheta at a converged density via stationarity (Hellmann-Feynman).

    ``dVq``: {name: dVq/dtheta array shaped like coulomb_q}; ``dh``: {name:
    dh/dtheta, (nk1, nk2, nb, nb) or (nb, nb)}.  Returns {name: float} plus
    ``"h"``: the per-k gradient w_k P^dagger (torch tensor).
    """
    P = to_tensor(density, kernel.device, kernel.dtype)
    h = kernel.h.clone().requires_grad_(True)
    VR = kernel._VR_shifted.clone().requires_grad_(True)
    spec = dict(w2d=kernel._w64, refP=kernel._refP, HH=kernel._HH64 if kernel.include_hartree else None)
    E = HFEnergy.apply(P, h, VR, spec)
    gh, gV = torch.autograd.grad(E, (h, VR))
    out = {"h": gh}
    part = K.Workspace.get(kernel.device, P.shape[0] * P.shape[1]).part
    ws = K.Workspace.get(kernel.device, P.shape[0] * P.shape[1])

    def inner(Gr, D):
        nk = Gr.shape[0] * Gr.shape[1]
        K.dots([Gr], [D], part, mode=0)
        K.reduce(ws, part, 1, nk, out=ws.out[0:1])
        return float(ws.out[0].item())

    for name, dV in (dVq or {}).items():
        dVR = vr_from_vq(dV, kernel._w64, kernel.device, kernel.dtype)
        out[name] = inner(gV, dVR)
    for name, d in (dh or {}).items():
        d = np.asarray(d)
        if d.ndim == 2:
            d = np.broadcast_to(d, tuple(kernel.h.shape))
        Dt = to_tensor(np.ascontiguousarray(d.astype(np.complex128)), kernel.device, kernel.dtype)
        out[name] = inner(gh, Dt)
    return out
