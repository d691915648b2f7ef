"""Deflated basin search (drop-in for src/jax_hf/deflation.py).

``solve_deflated`` finds distinct Hartree-Fock stationary points: an unbiased
trust-region Newton solve, then repeated Newton solves with a repulsive
Gaussian bias around every solution found so far (the ``deflation_targets``
of ``SolverConfig``, evaluated on the device inside the Newton loop), each
re-polished unbiased, keeping the results that are farther than
``distinct_tol`` (weighted Frobenius distance) from all earlier ones
(deflation.py:56-180).  Distances are reduced on the device by the same
batched-dot and weighted-reduce kernels the solvers use; the seed
perturbations are drawn from ``numpy.random.default_rng(seed)`` in the
reference's order, so a seed gives the same sequence of starting points.
"""

from __future__ import annotations

from dataclasses import replace
from typing import NamedTuple

import numpy as np
import torch

from . import _ops as K
from ._util import to_tensor
from .solver import SolveResult, SolverConfig, solve

__all__ = ["DeflatedResult", "solve_deflated"]


class DeflatedResult(NamedTuple):
    """deflation.py:32-36."""

    solutions: list          # list[SolveResult], distinct minima, ascending energy
    energies: np.ndarray     # physical energies of `solutions`, ascending
    best: SolveResult        # lowest-energy solution
    n_found: int             # == len(solutions)


class _Metric:
    """sum_k w_k ||A_k - B_k||_F^2 on the device (deflation.py:39-47)."""

    def __init__(self, kernel):
        self.w = kernel._w64.reshape(-1).contiguous()
        self.nk = self.w.numel()
        self.ws = K.Workspace.get(kernel.device, self.nk)
        self.out = torch.zeros(8, dtype=torch.float64, device=kernel.device)

    def norm2(self, M: torch.Tensor) -> float:
        K.dots([M], [M], self.ws.part, mode=0)
        K.reduce(self.ws, self.ws.part, 1, self.nk, w=self.w, wmask=1, out=self.out[:1])
        return float(self.out[0].item())

    def dist(self, A: torch.Tensor, B: torch.Tensor) -> float:
        B = to_tensor(B, A.device, A.dtype)
        d = torch.empty_like(A)
        K.lincomb(d, A, 1.0, B, -1.0)
        return float(np.sqrt(max(self.norm2(d), 0.0)))


def _hermitian_perturbation(shape, scale: float, rng, device) -> torch.Tensor:
    """scale * 0.5 (A + A^dagger) for A of standard normal parts (deflation.py:50-53):
    A is drawn on the host in the reference's order, the rest runs on the device."""
    A = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex128)
    At = to_tensor(A, device, torch.complex128)
    H = torch.empty_like(At)
    K.hermitize(At, H)
    K.lincomb(H, H, float(scale))
    return H


def solve_deflated(kernel, P0, n_electrons: float, *, base_config: SolverConfig | None = None,
                   n_solutions: int = 4, sigma: float | None = None, length: float | None = None,
                   distinct_tol: float | None = None, max_retries: int = 4, sigma_growth: float = 3.0,
                   perturb: float = 1e-2, repolish: bool = True, seed: int = 0) -> DeflatedResult:
    """Distinct HF solutions by deflated trust-region Newton (deflation.py:56-180).

    Same arguments and defaults: ``optimizer`` is forced to ``"newton"``;
    ``length`` defaults to 0.5 ||P*_0||_w (1 if that is zero), ``sigma`` to
    max(|E_0|, 1), ``distinct_tol`` to 0.25 length; each new solution gets up
    to ``max_retries`` biased attempts from ``P0`` plus a Hermitian
    perturbation of size ``perturb * (attempt + 1)``, the bias height growing
    by ``sigma_growth`` per failed attempt; the search stops at
    ``n_solutions`` or when an attempt budget finds nothing new.  Solutions
    are returned sorted by energy.
    """
    metric = _Metric(kernel)
    rng = np.random.default_rng(seed)
    base = base_config if base_config is not None else SolverConfig()
    newton_cfg = replace(base, optimizer="newton", deflation_targets=None, deflation_sigma=0.0)
    P0 = to_tensor(P0, kernel.device, torch.complex128)
    ne = float(n_electrons)

    r0 = solve(kernel, P0, ne, config=newton_cfg)
    solutions = [r0]

    if length is None:                                                   # deflation.py:127-133
        d0 = float(np.sqrt(max(metric.norm2(to_tensor(r0.density)), 0.0)))
        length = 0.5 * d0 if d0 > 0.0 else 1.0
    if sigma is None:
        sigma = max(abs(float(r0.energy)), 1.0)
    if distinct_tol is None:
        distinct_tol = 0.25 * length

    while len(solutions) < n_solutions:                                  # deflation.py:136-168
        found_new = False
        cur_sigma = float(sigma)
        for attempt in range(max_retries):
            targets = torch.stack([to_tensor(s.density, kernel.device, torch.complex128) for s in solutions])
            biased_cfg = replace(newton_cfg, deflation_targets=targets, deflation_sigma=cur_sigma,
                                 deflation_length=float(length))
            seed_P = torch.empty_like(P0)
            K.lincomb(seed_P, P0, 1.0, _hermitian_perturbation(tuple(P0.shape), perturb * (attempt + 1), rng,
                                                               kernel.device), 1.0)
            rb = solve(kernel, seed_P, ne, config=biased_cfg)
            cand = solve(kernel, to_tensor(rb.density, kernel.device, torch.complex128), ne,
                         config=newton_cfg) if repolish else rb
            if bool(cand.converged) and all(metric.dist(to_tensor(cand.density), s.density) > distinct_tol
                                            for s in solutions):
                solutions.append(cand)
                found_new = True
                break
            cur_sigma *= sigma_growth
        if not found_new:
            break

    order = np.argsort([float(s.energy) for s in solutions])
    solutions = [solutions[i] for i in order]
    energies = np.array([float(s.energy) for s in solutions])
    return DeflatedResult(solutions=solutions, energies=energies, best=solutions[0], n_found=len(solutions))
