"""Roothaan SCF solver on the GPU (drop-in for reference_scf.py).

Mirrors ``src/jax_hf/reference_scf.py:31-365``: ``SCFConfig`` (validated the
same way), ``SCFResult``, ``solve_scf`` with linear mixing, Pulay C-DIIS and
ODA, optional level shift and trust-radius clip.  Per iteration: Fock build
(+ energy) -> batched Jacobi eigh -> one cooperative launch for both mu
bisections -> density -> commutator / density residuals, all libhfx kernels.
The DIIS B matrix is kept incrementally: only the row of the history slot
that was just overwritten is recomputed (a fixed-order device reduction), and
the (N+1)^2 Pulay system is solved on the host, as the reference's own
comment places it.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np
import torch

from . import _lib as L
from . import _ops as K
from ._util import KIND, to_tensor

__all__ = ["SCFConfig", "SCFResult", "solve_scf"]


@dataclass(frozen=True)
class SCFConfig:
    """reference_scf.py:31-63."""
    max_iter: int = 200
    density_tol: float = 1e-7
    comm_tol: float = 1e-6
    mixing: float = 0.5
    level_shift: float = 0.0
    acceleration: str = "linear"
    diis_size: int = 6
    diis_start: int = 2
    diis_damping: float = 1.0
    trust_radius: float = 0.0
    project_fn: Callable[[Any], Any] | None = None

    def __post_init__(self) -> None:
        if self.max_iter <= 0:
            raise ValueError("max_iter must be positive")
        if self.acceleration not in ("linear", "diis", "oda"):
            raise ValueError("SCFConfig.acceleration must be 'linear', 'diis', or 'oda'; "
                             f"got {self.acceleration!r}")
        if self.diis_size <= 0:
            raise ValueError("diis_size must be positive")


@dataclass(frozen=True)
class SCFResult:
    """reference_scf.py:66-75."""
    density_matrix: Any
    fock_matrix: Any
    energy: Any
    chemical_potential: Any
    iterations: int
    converged: bool
    message: str
    history: dict = field(default_factory=dict)


def diis_solve(B: np.ndarray, active: np.ndarray):
    """Masked Pulay C-DIIS augmented solve (reference_scf.py:99-121) on the host."""
    N = B.shape[0]
    Bn = B / max(float(np.max(np.abs(B))), 1e-30)
    af = active.astype(np.float64)
    A = np.zeros((N + 1, N + 1))
    A[:N, :N] = Bn + np.diag(1.0 - af)
    A[:N, N] = -af
    A[N, :N] = -af
    rhs = np.zeros(N + 1)
    rhs[N] = -1.0
    try:
        sol = np.linalg.solve(A, rhs)
    except np.linalg.LinAlgError:
        return np.full(N, np.nan), False
    return sol[:N], bool(np.all(np.isfinite(sol)))


class _SCF:
    def __init__(self, kernel, cfg: SCFConfig, fock_build_fn=None, dtype=None, mu_newton=False):
        from .solver import _Fock
        self.mu_newton = bool(mu_newton)
        self.mu_warm = False
        self.kernel = kernel
        self.cfg = cfg
        self.dev = kernel.device
        self.cdt = dtype or kernel.dtype
        self.single = self.cdt == torch.complex64
        nk1, nk2, nb = kernel.h.shape[0], kernel.h.shape[1], kernel.h.shape[-1]
        self.shape4 = (nk1, nk2, nb, nb)
        self.nk, self.nb = nk1 * nk2, nb
        self.fock = _Fock(kernel, self.cdt, fock_build_fn, None)
        self.w = kernel._w64.reshape(-1).contiguous()
        self.ws_sum = max(kernel.weight_sum, 1e-30)
        self.wsp = K.Workspace.get(self.dev, self.nk)
        self.kind = KIND[self.cdt]
        c = dict(dtype=self.cdt, device=self.dev)
        self.F = torch.zeros(self.shape4, **c)
        self.Pn = torch.zeros(self.shape4, **c)
        self.tmp = torch.zeros(self.shape4, **c)
        self.U = torch.zeros(self.shape4, **c)
        self.Ft = torch.empty(self.shape4, **c)
        self.n_warm = -1                  # eigh calls in the running basis U (-1: no basis yet)
        self.allow_warm = False
        self.eps = torch.zeros((nk1, nk2, nb), dtype=torch.float64, device=self.dev)
        self.occ = torch.zeros_like(self.eps)
        self.st = torch.zeros(K.ST_LEN, dtype=torch.float64, device=self.dev)

    def herm_project(self, X, out):
        """out = herm(project(X)) (reference_scf.py:265, 272, 276)."""
        pf = self.cfg.project_fn
        src = X if pf is None else to_tensor(pf(X), self.dev, self.cdt)
        K.hermitize(src, out)
        return out

    # the running basis is re-orthonormalised (one Newton-Schulz step) every
    # this many warm eigh calls, so that its drift from unitarity stays at
    # rounding level over long runs
    REORTH_EVERY = 16
    # False: every eigh from scratch
    WARM_EIGH = True
    # The warm start engages once the commutator residual is below this.  Far
    # from the fixed point the DIIS history can be exactly degenerate (the
    # density repeats, so rows of B coincide and the augmented solve reports a
    # singular system); a cold eigh reproduces a repeated density bit for bit
    # and keeps that degeneracy exact, as the reference's does, while a basis
    # that moved between the calls perturbs it at rounding level into a nearly
    # singular solve with huge coefficients.
    WARM_BELOW = 1e-3

    def density_from_fock(self, F, P_out, mu_slot):
        """reference_scf.py:87-96: eigh(herm F), two bisections with level shift, P.

        Once the commutator residual is below WARM_BELOW the calls work in the
        previous call's eigenvectors U (warm start: Ft = U^H herm(F) U is
        nearly diagonal, so the Jacobi eigh needs a few sweeps instead of a
        cold run's ~9; then U <- U W); before that every call diagonalises
        herm(F) from scratch.  The eigenvalues, the occupations and
        P = U diag(occ) U^H are those of eigh(herm F) to rounding.

        With ``mu_newton`` (the FP64 polish of the mixed CG form, level shift 0)
        mu comes from the safeguarded Newton solver warm-started at the previous
        mu instead: the same root to rounding, in a few steps instead of 54.
        """
        warm = self.n_warm >= 0 and self.WARM_EIGH and self.allow_warm
        if warm:
            K.extrap_rotate(self.Ft, self.U, [F], [1.0])
            K.eigh_basis(self.Ft, self.eps, self.U, herm_in=True)      # U <- U W
        else:
            K.eigh(F, self.eps, self.U, herm_in=True)
        if self.mu_newton and self.cfg.level_shift == 0.0:
            K.solve_mu_steps(self.wsp, self.eps, self.w, self.nk, self.nb, self.n_e, self.kernel.T,
                             mu0_from_state=self.mu_warm, maxiter=54, st=self.st, p_out=self.occ)
            self.mu_warm = True
            if mu_slot != K.ST_MU:
                self.st[mu_slot] = self.st[K.ST_MU]
        else:
            K.solve_mu(self.wsp, self.eps, self.w, self.nk, self.nb, self.n_e, self.kernel.T,
                       maxiter=30 if self.single else 54, mode=K.MU_BISECTION,
                       level_shift=self.cfg.level_shift, mu_out_ptr=K.st_ptr(self.st, mu_slot), p_out=self.occ)
        K.density(self.U, self.tmp, p=self.occ)
        if warm:
            self.n_warm += 1
            if self.n_warm % self.REORTH_EVERY == 0:
                K.orthonormalize(self.U, 1)
        else:
            self.n_warm = 0
        return self.herm_project(self.tmp, P_out)

    def run(self, P0, n_electrons):
        cfg = self.cfg
        self.n_e = float(n_electrons)
        dev, nk = self.dev, self.nk
        c = dict(dtype=self.cdt, device=dev)
        density = to_tensor(P0, dev, self.cdt).clone()
        Ph = torch.zeros(self.shape4, **c)
        delta = torch.zeros(self.shape4, **c)
        comm = torch.zeros(self.shape4, **c)
        upd = torch.zeros(self.shape4, **c)
        step = torch.zeros(self.shape4, **c)
        part = self.wsp.part
        use_diis = cfg.acceleration == "diis"
        nh = cfg.diis_size if use_diis else 0
        F_hist = [torch.zeros(self.shape4, **c) for _ in range(nh)]
        R_hist = [torch.zeros(self.shape4, **c) for _ in range(nh)]
        B = np.zeros((nh, nh))
        outs = torch.zeros(64, dtype=torch.float64, device=dev)     # [0:2] residuals, [2] E, [8:] B row
        o_res, o_B = outs[0:2], outs[8:8 + max(nh, 1)]
        E_ptr = outs.data_ptr() + 8 * 2
        hE, hD, hC = [], [], []
        k = 0
        converged = False
        while k < cfg.max_iter and not converged:                         # :287-347
            self.herm_project(density, Ph)                                # _build_and_occupy (:264-273)
            self.fock.build(Ph, self.F, E_ptr=E_ptr, project_fn=cfg.project_fn)
            self.density_from_fock(self.F, self.Pn, K.ST_MU)
            K.lincomb(delta, self.Pn, 1.0, density, -1.0)                 # :291
            K.scf_resid(self.F, density, self.Pn, comm, part)             # :292-294
            K.reduce(self.wsp, part, 2, nk, w=self.w, wmask=0b11, out=o_res)
            if use_diis:
                idx = k % nh
                F_hist[idx].copy_(self.F)
                R_hist[idx].copy_(comm)
                for j0 in range(0, nh, 8):                                # B row (:110-111)
                    Ys = R_hist[j0:j0 + 8]
                    K.dots(R_hist[idx], Ys, part, mode=0)
                    K.reduce(self.wsp, part, len(Ys), nk, w=self.w, wmask=(1 << len(Ys)) - 1,
                             out=o_B[j0:j0 + len(Ys)])
            sv = outs.cpu().numpy()
            c_res = float(np.sqrt(sv[0] / self.ws_sum))
            d_res = float(np.sqrt(sv[1] / self.ws_sum))
            E = float(sv[2])
            converged = d_res <= cfg.density_tol and c_res <= cfg.comm_tol  # :295
            self.allow_warm = c_res < self.WARM_BELOW
            engaged = False
            if use_diis:                                                  # :298-314
                row = sv[8:8 + nh]
                B[idx, :] = row
                B[:, idx] = row
                count = min(k + 1, nh)
                active = np.arange(nh) < count
                coeffs, ok = diis_solve(B * np.outer(active, active), active)
                engaged = count >= max(2, cfg.diis_start) and ok
                if engaged:
                    damp = min(max(cfg.diis_damping, 0.0), 1.0)
                    K.lincomb(upd, self.F, 1.0 - damp)                    # F_ex = damp sum c F + (1-damp) F
                    for i in range(nh):
                        if coeffs[i] != 0.0:
                            K.lincomb(upd, F_hist[i], damp * float(coeffs[i]), upd, 1.0)
                    K.hermitize(upd, step)
                    self.density_from_fock(step, upd, 12)
            elif cfg.acceleration == "oda":                               # :315-328
                self.fock.build(self.Pn, step, E_ptr=E_ptr + 8, project_fn=cfg.project_fn)
                K.dots(self.F, [delta], part, mode=1)
                K.reduce(self.wsp, part, 1, nk, w=self.w, wmask=1, out=outs[4:5])
                sv2 = outs.cpu().numpy()
                E_new, c1 = float(sv2[3]), float(sv2[4])
                c2 = (E_new - E) - c1
                if c2 > 1e-12:
                    lam = min(max(-c1 / (2.0 * c2), 0.0), 1.0)
                else:
                    lam = 1.0 if c1 < 0.0 else 0.0
                K.lincomb(step, density, 1.0, delta, lam)
                self.herm_project(step, upd)
                engaged = True
            if not engaged:                                               # linear (:275-276)
                K.lincomb(step, density, 1.0, delta, cfg.mixing)
                self.herm_project(step, upd)
            if cfg.trust_radius > 0.0:                                    # :332-339
                K.lincomb(step, upd, 1.0, density, -1.0)
                K.dots(step, [step], part, mode=0)
                K.reduce(self.wsp, part, 1, nk, w=self.w, wmask=1, out=outs[5:6])
                sn = float(np.sqrt(float(outs[5].item()) / self.ws_sum))
                if sn > cfg.trust_radius:
                    sc = cfg.trust_radius / max(sn, 1e-30)
                    K.lincomb(delta, density, 1.0, step, sc)
                    self.herm_project(delta, upd)
            if converged:
                density.copy_(self.Pn)                                    # :341
            else:
                density.copy_(upd)
            hE.append(E)
            hD.append(d_res)
            hC.append(c_res)
            k += 1
        # final evaluation at the returned density (:363-365)
        self.herm_project(density, Ph)
        F_fin = torch.empty(self.shape4, **c)
        self.fock.build(Ph, F_fin, E_ptr=E_ptr, project_fn=cfg.project_fn)
        self.density_from_fock(F_fin, self.Pn, K.ST_MU)
        sv = outs.cpu().numpy()
        E_fin = torch.tensor(float(sv[2]), dtype=torch.float64, device=dev)
        mu_fin = self.st[K.ST_MU].clone()
        return density, F_fin, E_fin, mu_fin, k, converged, hE, hD, hC


def solve_scf(kernel, P0, n_electrons: float, *, config: SCFConfig | None = None,
              fock_build_fn: Callable | None = None) -> SCFResult:
    """Reference SCF solver (reference_scf.py:124-219)."""
    if config is None:
        config = SCFConfig()
    if kernel.h.shape[-1] > L.PERK_NB_MAX:
        raise ValueError(f"the GPU solvers support nb <= {L.PERK_NB_MAX} bands")
    runner = _SCF(kernel, config, fock_build_fn)
    density, F, E, mu, n_iter, converged, hE, hD, hC = runner.run(P0, n_electrons)
    if converged:
        message = f"converged in {n_iter} iterations"
    else:
        dr = hD[n_iter - 1] if n_iter > 0 else float("nan")
        cr = hC[n_iter - 1] if n_iter > 0 else float("nan")
        message = f"stopped after {n_iter} iterations (density_res={dr:.3e}, comm_res={cr:.3e})"
    dev = kernel.device
    return SCFResult(density_matrix=density, fock_matrix=F, energy=E, chemical_potential=mu,
                     iterations=int(n_iter), converged=bool(converged), message=message,
                     history={"E": torch.tensor(hE, dtype=torch.float64, device=dev),
                              "density_residual": torch.tensor(hD, dtype=torch.float64, device=dev),
                              "commutator_residual": torch.tensor(hC, dtype=torch.float64, device=dev)})
