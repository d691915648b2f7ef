"""NumPy restatement of the reference's host drivers around the solvers.

TEST INFRASTRUCTURE ONLY (imported by ``tests/``; never by the product path).

* ``resample_kgrid`` follows ``src/jax_hf/utils.py:477-535``: periodic linear
  interpolation of a centred k-grid, axis 0 then axis 1, source coordinate
  u = (j - n_new//2) * (n_old / n_new) + n_old//2 evaluated in the array's
  real precision, neighbours floor(u) and floor(u) + 1 modulo n_old.
* ``solve_continuation`` follows ``src/jax_hf/continuation.py:48-175`` over
  the oracle's ``solve`` / ``solve_scf``.
* ``solve_deflated`` follows ``src/jax_hf/deflation.py:39-180`` over the
  oracle's trust-region Newton with deflation targets.

Pinned in ``tests/test_oracle_drivers.py``: the resample against
``numpy.interp`` with ``period`` (a library routine for periodic linear
interpolation, axis by axis), identity at equal size, exact reproduction of
the coarse samples at coincident points and the 2n -> n round trip,
constants preserved; the continuation against the reference test's manual
coarse-then-fine driver and the cold fine solve; the deflated search
against the bistable model's analytic solutions E = -delta - ex/2 and
E = +delta - ex/2 (tests/test_deflation.py) and its argument handling.
"""

from __future__ import annotations

from dataclasses import replace
from typing import Any, NamedTuple

import numpy as np

from .hf_oracle import SCFConfig, SCFResult, SolverConfig, solve, solve_scf

__all__ = ["resample_kgrid", "ContinuationResult", "solve_continuation", "DeflatedResult", "solve_deflated",
           "wdist"]


def _resample_axis(x: np.ndarray, n_new: int, axis: int) -> np.ndarray:
    """utils.py:477-509."""
    n_old = x.shape[axis]
    if n_new <= 0:
        raise ValueError("n_new must be a positive integer.")
    if n_old == n_new:
        return x
    rdt = np.real(x).dtype.type
    j = np.arange(n_new, dtype=rdt)
    u = (j - rdt(n_new // 2)) * (rdt(n_old) / rdt(n_new))
    u = u + rdt(n_old // 2)
    i0 = np.mod(np.floor(u).astype(np.int64), n_old)
    i1 = np.mod(i0 + 1, n_old)
    shape = [1] * x.ndim
    shape[axis] = n_new
    frac = (u - np.floor(u)).reshape(shape).astype(rdt)
    x0 = np.take(x, i0, axis=axis)
    x1 = np.take(x, i1, axis=axis)
    return (rdt(1) - frac) * x0 + frac * x1


def resample_kgrid(values, nk: int, *, method: str = "linear") -> np.ndarray:
    """utils.py:512-535."""
    x = np.asarray(values)
    nk = int(nk)
    if x.ndim < 2:
        raise ValueError("resample_kgrid expects an array with at least 2 dimensions (nk,nk,...).")
    if x.shape[0] == nk and x.shape[1] == nk:
        return x
    if str(method).lower() != "linear":
        raise ValueError(f"Unsupported resample method {method!r}. Only 'linear' is supported.")
    return _resample_axis(_resample_axis(x, nk, 0), nk, 1)


class ContinuationResult(NamedTuple):
    coarse: Any
    fine: Any
    P0_fine: np.ndarray


def _herm(X):
    return 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))


def _run_stage(kernel, P0, n, config, stage_name):
    if isinstance(config, SCFConfig):
        return solve_scf(kernel, P0, float(n), config=config)
    if isinstance(config, SolverConfig) or config is None:
        return solve(kernel, P0, float(n), config=config)
    raise TypeError(f"{stage_name}_config must be SolverConfig, SCFConfig, or None; got {type(config).__name__}.")


def solve_continuation(coarse_kernel, fine_kernel, P0_coarse, n_electrons_coarse, n_electrons_fine, *,
                       coarse_config=None, fine_config=None) -> ContinuationResult:
    """continuation.py:78-175."""
    if coarse_kernel.h.shape[-2:] != fine_kernel.h.shape[-2:]:
        raise ValueError("coarse_kernel and fine_kernel must share orbital dimensions")
    if coarse_kernel.include_hartree != fine_kernel.include_hartree:
        raise ValueError("coarse_kernel and fine_kernel must agree on include_hartree.")
    if coarse_kernel.include_exchange != fine_kernel.include_exchange:
        raise ValueError("coarse_kernel and fine_kernel must agree on include_exchange.")
    P0c = np.asarray(P0_coarse, np.complex128)
    if P0c.shape != coarse_kernel.h.shape:
        raise ValueError(f"P0_coarse shape {P0c.shape} does not match coarse_kernel hamiltonian shape")
    rc = _run_stage(coarse_kernel, P0c, n_electrons_coarse, coarse_config, "coarse")
    Pc = rc.density_matrix if isinstance(rc, SCFResult) else rc.density
    P0f = _herm(resample_kgrid(Pc, fine_kernel.h.shape[0]))
    rf = _run_stage(fine_kernel, P0f, n_electrons_fine, fine_config, "fine")
    return ContinuationResult(coarse=rc, fine=rf, P0_fine=P0f)


class DeflatedResult(NamedTuple):
    solutions: list
    energies: np.ndarray
    best: Any
    n_found: int


def _wnorm2(M, w2d):
    return float(np.sum(np.asarray(w2d) * np.sum(np.abs(M) ** 2, axis=(-2, -1))))


def wdist(A, B, w2d) -> float:
    """deflation.py:45-47."""
    return float(np.sqrt(max(_wnorm2(np.asarray(A) - np.asarray(B), w2d), 0.0)))


def solve_deflated(kernel, P0, n_electrons, *, base_config=None, n_solutions=4, sigma=None, length=None,
                   distinct_tol=None, max_retries=4, sigma_growth=3.0, perturb=1e-2, repolish=True,
                   seed=0) -> DeflatedResult:
    """deflation.py:56-180."""
    w2d = np.asarray(kernel.w2d)
    rng = np.random.default_rng(seed)
    base = base_config if base_config is not None else SolverConfig()
    ncfg = replace(base, optimizer="newton", deflation_targets=None, deflation_sigma=0.0)
    P0 = np.ascontiguousarray(np.asarray(P0), dtype=np.complex128)
    r0 = solve(kernel, P0, float(n_electrons), config=ncfg)
    sols = [r0]
    if length is None:
        d0 = float(np.sqrt(max(_wnorm2(r0.density, w2d), 0.0)))
        length = 0.5 * d0 if d0 > 0.0 else 1.0
    if sigma is None:
        sigma = max(abs(float(r0.energy)), 1.0)
    if distinct_tol is None:
        distinct_tol = 0.25 * length
    while len(sols) < n_solutions:
        found, cur = False, float(sigma)
        for attempt in range(max_retries):
            bcfg = replace(ncfg, deflation_targets=np.stack([s.density for s in sols]), deflation_sigma=cur,
                           deflation_length=float(length))
            A = rng.standard_normal(P0.shape) + 1j * rng.standard_normal(P0.shape)
            seed_P = P0 + perturb * (attempt + 1) * 0.5 * (A + np.conj(np.swapaxes(A, -1, -2)))
            rb = solve(kernel, seed_P, float(n_electrons), config=bcfg)
            cand = solve(kernel, np.asarray(rb.density, np.complex128), float(n_electrons), config=ncfg) \
                if repolish else rb
            if bool(cand.converged) and all(wdist(cand.density, s.density, w2d) > distinct_tol for s in sols):
                sols.append(cand)
                found = True
                break
            cur *= sigma_growth
        if not found:
            break
    order = np.argsort([float(s.energy) for s in sols])
    sols = [sols[i] for i in order]
    en = np.array([float(s.energy) for s in sols])
    return DeflatedResult(solutions=sols, energies=en, best=sols[0], n_found=len(sols))
