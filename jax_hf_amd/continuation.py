"""Coarse-to-fine continuation (drop-in for src/jax_hf/continuation.py).

``solve_continuation`` runs ``solve`` or ``solve_scf`` on a coarse kernel,
resamples the converged density onto the fine k-mesh with the periodic linear
``resample_kgrid`` (one ``hfx_resample_kgrid`` launch), Hermitises it
(``hfx_hermitize``) and runs the fine stage from that seed
(continuation.py:48-141).  Both stages and the hand-off stay on the device.
"""

from __future__ import annotations

from typing import Any, NamedTuple, Union

import torch

from . import _ops as K
from ._util import convert, to_tensor
from .problem import HartreeFockKernel
from .reference_scf import SCFConfig, SCFResult, solve_scf
from .solver import SolverConfig, SolveResult, solve
from .utils import resample_kgrid

__all__ = ["ContinuationResult", "solve_continuation"]

StageResult = Union[SolveResult, SCFResult]


class ContinuationResult(NamedTuple):
    """(coarse, fine) stage results and the fine-grid seed (continuation.py:35-45)."""

    coarse: StageResult
    fine: StageResult
    P0_fine: torch.Tensor


def _resample_density(density_c: torch.Tensor, nk_fine: int) -> torch.Tensor:
    """Resampled coarse density, Hermitised (continuation.py:48-51)."""
    P = resample_kgrid(density_c, nk_fine, method="linear")
    out = torch.empty_like(P)
    K.hermitize(P, out)
    return out


def _run_stage(kernel: HartreeFockKernel, P0, n_electrons: float, config: Any, *, stage_name: str) -> StageResult:
    """continuation.py:54-69: SCFConfig -> solve_scf, SolverConfig / None -> solve."""
    if isinstance(config, SCFConfig):
        return solve_scf(kernel, P0, float(n_electrons), config=config)
    if isinstance(config, SolverConfig) or config is None:
        return solve(kernel, P0, float(n_electrons), config=config)
    raise TypeError(f"{stage_name}_config must be SolverConfig, SCFConfig, or None; "
                    f"got {type(config).__name__}.")


def _density_of(result: StageResult) -> torch.Tensor:
    if isinstance(result, SCFResult):
        return result.density_matrix
    return result.density


def _validate_kernels(coarse: HartreeFockKernel, fine: HartreeFockKernel) -> None:
    """continuation.py:144-175."""
    if tuple(coarse.h.shape[-2:]) != tuple(fine.h.shape[-2:]):
        raise ValueError("coarse_kernel and fine_kernel must share orbital dimensions "
                         f"(..., nb, nb); got {tuple(coarse.h.shape[-2:])} vs {tuple(fine.h.shape[-2:])}.")
    if coarse.include_hartree != fine.include_hartree:
        raise ValueError("coarse_kernel and fine_kernel must agree on include_hartree.")
    if coarse.include_exchange != fine.include_exchange:
        raise ValueError("coarse_kernel and fine_kernel must agree on include_exchange.")
    if coarse.h.shape[0] != coarse.h.shape[1]:
        raise ValueError(f"coarse_kernel must have a square k-mesh, got shape {tuple(coarse.h.shape[:2])}.")
    if fine.h.shape[0] != fine.h.shape[1]:
        raise ValueError(f"fine_kernel must have a square k-mesh, got shape {tuple(fine.h.shape[:2])}.")


def solve_continuation(coarse_kernel: HartreeFockKernel, fine_kernel: HartreeFockKernel, P0_coarse,
                       n_electrons_coarse: float, n_electrons_fine: float, *,
                       coarse_config: SolverConfig | SCFConfig | None = None,
                       fine_config: SolverConfig | SCFConfig | None = None) -> ContinuationResult:
    """Coarse -> fine continuation (continuation.py:78-141), same arguments.

    The coarse stage starts from ``P0_coarse`` (cast to the coarse kernel's
    dtype); the fine stage from the Hermitised periodic-linear resample of the
    coarse density, cast to the fine kernel's dtype and returned as
    ``P0_fine``.  Electron counts are passed through unchanged.
    """
    _validate_kernels(coarse_kernel, fine_kernel)
    P0c = to_tensor(P0_coarse, coarse_kernel.device, coarse_kernel.dtype)
    if tuple(P0c.shape) != tuple(coarse_kernel.h.shape):
        raise ValueError(f"P0_coarse shape {tuple(P0c.shape)} does not match "
                         f"coarse_kernel hamiltonian shape {tuple(coarse_kernel.h.shape)}.")
    coarse_result = _run_stage(coarse_kernel, P0c, n_electrons_coarse, coarse_config, stage_name="coarse")
    Pc = to_tensor(_density_of(coarse_result), fine_kernel.device)
    P0_fine = convert(_resample_density(Pc, int(fine_kernel.h.shape[0])), fine_kernel.dtype)
    fine_result = _run_stage(fine_kernel, P0_fine, n_electrons_fine, fine_config, stage_name="fine")
    return ContinuationResult(coarse=coarse_result, fine=fine_result, P0_fine=P0_fine)
