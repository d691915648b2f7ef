"""Hartree-Fock kernel: device-resident problem arrays (drop-in for problem.py).

Mirrors ``HartreeFockKernel`` of ``src/jax_hf/problem.py:14-217``: same
constructor arguments, attributes and ``as_args()``.  The interaction
transform ``VR = fftn(w * Vq)`` and its ifftshift phase (problem.py:88-98) are
computed once on the GPU by the same FFT kernels that run the exchange
convolution (``hfx_prepare_vr``), always in FP64 (a complex64 kernel
receives the FP64 result rounded once).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from ._util import complex_dtype_of, convert, default_device, host_array, real_dtype_of, to_tensor

__all__ = ["HartreeFockKernel"]


def _prepare_vr(Vq_c128: torch.Tensor, w64: torch.Tensor, phase: bool) -> torch.Tensor:
    nk1, nk2, vb = Vq_c128.shape[0], Vq_c128.shape[1], Vq_c128.shape[-1]
    plan = L.plan_for(nk1, nk2, vb, torch.complex128, Vq_c128.device)
    out = torch.empty_like(Vq_c128)
    L.call("hfx_prepare_vr", plan.handle, L.ptr(Vq_c128), 1, L.ptr(w64), int(phase), L.ptr(out),
           L.stream(Vq_c128.device))
    return out


class HartreeFockKernel:
    """Precomputed device arrays for the solver loop (problem.py:14-217)."""

    def __init__(self, weights, hamiltonian, coulomb_q, T: float, include_hartree: bool = False,
                 include_exchange: bool = True, reference_density=None, hartree_matrix=None,
                 contact_terms=None, embed_reference_potential: bool = False,
                 center_embedded_hartree: bool = True, *, device=None):
        L.require_gpu()
        L.register_owner(self)
        device = torch.device(device) if device is not None else default_device()
        h = to_tensor(hamiltonian, device)
        cdt = complex_dtype_of(h)
        h = convert(h, cdt)
        if h.dim() != 4 or h.shape[-1] != h.shape[-2]:
            raise ValueError(f"hamiltonian must have shape (nk1, nk2, nb, nb); got {tuple(h.shape)}")
        rdt = real_dtype_of(cdt)
        nk1, nk2, nb = int(h.shape[0]), int(h.shape[1]), int(h.shape[-1])
        self.h = h
        self.dtype = cdt
        self.device = device
        w_np = np.real(host_array(weights)).astype(np.float64)               # :73-76
        if w_np.shape != (nk1, nk2):
            raise ValueError(f"weights must have shape {(nk1, nk2)}, got {w_np.shape}")
        self._w64 = to_tensor(w_np, device, torch.float64)
        self.w2d = convert(self._w64, rdt)
        self.weights_b = self.w2d[..., None, None]
        self.weight_sum = float(w_np.sum())

        Vq = host_array(coulomb_q)                                            # :78-87
        if np.iscomplexobj(Vq) and float(np.max(np.abs(np.imag(Vq)))) <= 1e-8:
            Vq = np.real(Vq)
        Vq_complex = np.iscomplexobj(Vq)
        if Vq.ndim != 4 or Vq.shape[:2] != (nk1, nk2) or Vq.shape[2:] not in ((1, 1), (nb, nb)):
            raise ValueError(f"coulomb_q must have shape (nk1, nk2, 1, 1) or (nk1, nk2, nb, nb); "
                             f"got {Vq.shape}")
        self.exchange_hermitian_channel_packing = bool(Vq.shape[-2:] == (1, 1) and not Vq_complex)
        # Beyond the reference: a real symmetric layer-resolved V also keeps
        # Sigma Hermitian for Hermitian P, so the solvers may use packing.
        self._v_real_symmetric = bool(not Vq_complex and (
            Vq.shape[-2:] == (1, 1) or np.array_equal(Vq, np.swapaxes(Vq, -1, -2))))
        Vq_t = to_tensor(np.ascontiguousarray(Vq.astype(np.complex128)), device, torch.complex128)
        VR64 = _prepare_vr(Vq_t, self._w64, phase=False)                      # :88
        VRs64 = _prepare_vr(Vq_t, self._w64, phase=True)                      # :90-98
        self.VR = convert(VR64, cdt)
        self._VR_shifted = convert(VRs64, cdt)
        del Vq_t

        self.T = float(T)
        self.include_hartree = bool(include_hartree)
        self.include_exchange = bool(include_exchange)
        if not self.include_hartree and not self.include_exchange:          # :102-103
            raise ValueError("HartreeFockKernel must include at least one of Hartree or exchange.")

        if reference_density is not None:                                    # :105-111
            ref = to_tensor(reference_density, device, cdt)
            if tuple(ref.shape) != tuple(h.shape):
                raise ValueError(f"reference_density must have shape {tuple(h.shape)}, got {tuple(ref.shape)}")
            self._refP = ref
        else:
            self._refP = None

        if self.include_hartree:                                             # :113-123
            if hartree_matrix is None:
                raise ValueError("include_hartree=True requires hartree_matrix to be provided")
            if reference_density is None:
                raise ValueError("include_hartree=True requires reference_density to be provided")
            HH = np.real(host_array(hartree_matrix)).astype(np.float64)
            if HH.shape != (nb, nb):
                raise ValueError(f"hartree_matrix must have shape {(nb, nb)}, got {HH.shape}")
            self._HH64 = to_tensor(HH, device, torch.float64)
        else:
            self._HH64 = torch.zeros((nb, nb), dtype=torch.float64, device=device)
        self.HH = convert(self._HH64, rdt)

        if contact_terms is None or len(contact_terms) == 0:                 # :128-158
            g = np.zeros((1,), np.float64)
            Oi = np.zeros((1, nb, nb), np.complex128)
            Oj = np.zeros((1, nb, nb), np.complex128)
        else:
            gs, Ois, Ojs = [], [], []
            for k, term in enumerate(contact_terms):
                if len(term) != 3:
                    raise ValueError(f"contact_terms[{k}] must be (g, O_i, O_j); got len={len(term)}")
                g_t, Oi_t, Oj_t = term
                Oi_a = host_array(Oi_t).astype(np.complex128)
                Oj_a = host_array(Oj_t).astype(np.complex128)
                if Oi_a.shape != (nb, nb):
                    raise ValueError(f"contact_terms[{k}].O_i must have shape {(nb, nb)}, got {Oi_a.shape}")
                if Oj_a.shape != (nb, nb):
                    raise ValueError(f"contact_terms[{k}].O_j must have shape {(nb, nb)}, got {Oj_a.shape}")
                gs.append(float(np.real(host_array(g_t))))
                Ois.append(Oi_a)
                Ojs.append(Oj_a)
            g, Oi, Oj = np.asarray(gs), np.stack(Ois), np.stack(Ojs)
        self._contact_host = (g, Oi, Oj)
        self.contact_g = to_tensor(g, device, rdt)
        self.contact_Oi = to_tensor(Oi, device, cdt)
        self.contact_Oj = to_tensor(Oj, device, cdt)

        self.embed_reference_potential = bool(embed_reference_potential)     # :161-199
        self.center_embedded_hartree = bool(center_embedded_hartree)
        self.embedded_energy_offset = 0.0
        if self.embed_reference_potential:
            if reference_density is None:
                raise ValueError("embed_reference_potential=True requires reference_density.")
            self._embed()

    # ------------------------------------------------------------------
    @property
    def refP(self):
        """Reference density (zeros when none was given, problem.py:105-111)."""
        if self._refP is None:
            return torch.zeros_like(self.h)
        return self._refP

    @property
    def nb(self) -> int:
        return int(self.h.shape[-1])

    def _embed(self):
        """h <- h + Sigma[refP] + diag(J[refP]) and the energy offset (problem.py:164-199)."""
        from .fock import FockEngine
        nb = self.nb
        ref = self._refP
        w = host_array(self._w64)
        const_J = 0.0
        J = np.zeros(nb)
        eng = FockEngine(h=self.h, VR=self._VR_shifted, refP=None, HH=self._HH64, w2d=self._w64,
                         include_exchange=True, include_hartree=False,
                         herm=self.exchange_hermitian_channel_packing)
        if self.include_hartree:
            L.call("hfx_density_moments", eng.plan.handle, L.ptr(ref), None, L.ptr(self._w64),
                   L.ptr(eng.rho), L.stream(self.device))
            diag_R = np.real(np.diagonal(host_array(eng.rho)))
            J = host_array(self._HH64) @ diag_R
            if self.center_embedded_hartree:
                J = J - np.mean(J)
            const_J = float(np.sum(J * diag_R))
        h_eff = self.h.clone()
        const_S = 0.0
        if self.include_exchange:
            S_ref = torch.empty_like(self.h)
            eng.selfenergy(ref, S_ref, refP=None, rho=False)
            # sum_k w_k Tr(Sigma_ref refP) = 2 hf_energy(P=refP, h=0, Sigma=S_ref)
            from .fock import hf_energy
            zero = torch.zeros_like(self.h)
            const_S = 2.0 * float(hf_energy(ref, h=zero, Sigma=S_ref, H=None, weights_b=self._w64))
            h_eff = h_eff + S_ref
        Jm = torch.from_numpy(np.diag(J).astype(np.complex128)).to(self.device)
        h_eff = h_eff + convert(Jm, self.dtype)[None, None]
        self.h = h_eff.contiguous()
        self.embedded_energy_offset = 0.5 * (const_J + const_S)

    def as_args(self):
        """Dynamic solver inputs (problem.py:201-217)."""
        return dict(h=self.h, weights_b=self.weights_b, weight_sum=self.weight_sum,
                    VR=self._VR_shifted, T=self.T, refP=self.refP, HH=self.HH,
                    include_hartree=self.include_hartree, include_exchange=self.include_exchange,
                    exchange_hermitian_channel_packing=self.exchange_hermitian_channel_packing,
                    contact_g=self.contact_g, contact_Oi=self.contact_Oi, contact_Oj=self.contact_Oj)

    def _hfx_release(self):
        """Drop the cached solver states and their captured graphs (rebuilt on
        the next solve)."""
        for key in ("_hfx_cg", "_hfx_polish"):
            cache = self.__dict__.pop(key, None)
            if cache:
                cache.clear()

    def fock_engine(self, *, herm=None, refP="kernel", dtype=None):
        """FockEngine bound to this kernel (internal, used by the solvers)."""
        from .fock import FockEngine
        cdt = dtype or self.dtype
        herm = (self._v_real_symmetric if herm is None else herm)
        h = self.h if cdt == self.dtype else convert(self.h, cdt)
        VR = self._VR_shifted if cdt == self.dtype else convert(self._VR_shifted, cdt)
        ref = self._refP if isinstance(refP, str) else refP
        if ref is not None and ref.dtype != cdt:
            ref = convert(ref, cdt)
        g, Oi, Oj = self._contact_host
        return FockEngine(h=h, VR=VR if self.include_exchange else None, refP=ref, HH=self._HH64,
                          w2d=self._w64, include_exchange=self.include_exchange,
                          include_hartree=self.include_hartree, herm=herm, contact_g=g,
                          contact_Oi=Oi, contact_Oj=Oj)
