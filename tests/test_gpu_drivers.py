"""GPU host drivers against the oracle (oracle/drivers.py): resample_kgrid
(hfx_resample_kgrid) for every storage kind, solve_continuation
(continuation.py) and solve_deflated (deflation.py), following the reference's
tests/test_continuation.py and tests/test_deflation.py cases.

resample_kgrid rounds every product and sum on its own, like the oracle's
NumPy evaluation in the same precision (the floor indices are identical by
construction); outputs agree within 4 ulp of the largest input in the
array's own precision."""

import numpy as np
import pytest

import oracle as O
from oracle import drivers as D

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _rand(rng, shape, dt):
    x = rng.standard_normal(shape)
    if np.dtype(dt).kind == "c":
        x = x + 1j * rng.standard_normal(shape)
    return x.astype(dt)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64, np.float64, np.float32])
@pytest.mark.parametrize("shape,nk", [((2, 2, 2, 2), 4), ((6, 6, 4, 4), 4), ((3, 3, 16, 16), 4), ((5, 5, 3), 12),
                                      ((12, 12, 2, 2), 5), ((3, 5), 6), ((4, 6, 2, 2), 4), ((96, 96, 2, 2), 144)])
def test_resample_matches_oracle(dev, dt, shape, nk):
    import torch
    from jax_hf_amd import resample_kgrid
    rng = np.random.default_rng(sum(shape) + nk)
    x = _rand(rng, shape, dt)
    out = resample_kgrid(torch.from_numpy(x).to(dev), nk)
    ref = D.resample_kgrid(x, nk)
    assert tuple(out.shape) == ref.shape and _np(out).dtype == ref.dtype
    eps = np.finfo(np.real(ref).dtype).eps
    err = np.abs(_np(out).astype(np.complex128) - ref.astype(np.complex128)).max()
    assert err <= 4 * eps * np.abs(x).max(), err


def test_resample_identity_and_errors(dev):
    import torch
    from jax_hf_amd import resample_kgrid
    x = torch.ones((4, 4, 2, 2), dtype=torch.complex128, device=dev)
    assert resample_kgrid(x, 4) is x
    with pytest.raises(ValueError, match="at least 2"):
        resample_kgrid(torch.ones(4, dtype=torch.float64, device=dev), 2)
    with pytest.raises(ValueError, match="Only 'linear'"):
        resample_kgrid(x, 8, method="cubic")
    with pytest.raises(TypeError, match="dtype"):
        resample_kgrid(torch.ones((2, 2), dtype=torch.int32, device=dev), 4)


def _two_band(J, nk, dt=np.complex64, exchange=0.25, T=0.2):
    h = np.zeros((nk, nk, 2, 2), dt)
    h[..., 0, 0], h[..., 1, 1] = -0.5, 0.5
    args = (np.ones((nk, nk)), h, np.full((nk, nk, 1, 1), exchange, dt), T)
    return J.HartreeFockKernel(*args), O.OracleKernel(*args)


def test_continuation_matches_manual_driver(dev):
    """tests/test_continuation.py:49-84 on complex64 kernels."""
    import jax_hf_amd as J
    from jax_hf_amd.utils import hermitize
    kc, _ = _two_band(J, 2)
    kf, _ = _two_band(J, 4)
    P0 = np.zeros((2, 2, 2, 2), np.complex64)
    cfg = J.SolverConfig(max_iter=50, tol_E=1e-8)
    res = J.solve_continuation(kc, kf, P0, 1.0, 1.0, coarse_config=cfg, fine_config=cfg)
    ec = J.solve(kc, P0, 1.0, config=cfg)
    Pf = hermitize(J.resample_kgrid(ec.density, 4))
    ef = J.solve(kf, Pf, 1.0, config=cfg)
    assert isinstance(res, J.ContinuationResult)
    assert isinstance(res.coarse, J.SolveResult) and isinstance(res.fine, J.SolveResult)
    np.testing.assert_allclose(_np(res.coarse.density), _np(ec.density), atol=1e-6)
    np.testing.assert_allclose(_np(res.fine.density), _np(ef.density), atol=1e-6)
    np.testing.assert_allclose(_np(res.P0_fine), _np(Pf), atol=1e-6)
    assert res.P0_fine.dtype == kf.dtype


def test_continuation_scf_and_mixed_stages(dev):
    """tests/test_continuation.py:86-126."""
    import jax_hf_amd as J
    kc, _ = _two_band(J, 2)
    kf, _ = _two_band(J, 4)
    P0 = np.zeros((2, 2, 2, 2), np.complex64)
    scf = J.SCFConfig(max_iter=40, density_tol=1e-6, comm_tol=1e-5, mixing=0.5)
    r = J.solve_continuation(kc, kf, P0, 1.0, 1.0, coarse_config=scf, fine_config=scf)
    assert isinstance(r.coarse, J.SCFResult) and isinstance(r.fine, J.SCFResult)
    assert r.fine.converged and tuple(r.P0_fine.shape) == tuple(kf.h.shape)
    m = J.solve_continuation(kc, kf, P0, 1.0, 1.0, coarse_config=J.SCFConfig(max_iter=30, mixing=0.5),
                             fine_config=J.SolverConfig(max_iter=30, tol_E=1e-8))
    assert isinstance(m.coarse, J.SCFResult) and isinstance(m.fine, J.SolveResult)
    d = J.solve_continuation(kc, kf, P0, 1.0, 1.0)
    assert isinstance(d.coarse, J.SolveResult) and isinstance(d.fine, J.SolveResult)


def test_continuation_matches_oracle_and_cold_fixed_point(dev):
    """tests/test_continuation.py:135-157 in complex128 against the oracle
    continuation (tol_E 1e-13: the 1e-10 windowed stop of the cold solve fires
    3e-6 above the fixed point in complex128).  Energies agree to 1e-8: the
    energy, unlike the free energy, is not stationary at the fixed point, so
    the 1e-8-level density differences the windowed stop leaves between two
    solvers move it linearly (4e-9 here)."""
    import jax_hf_amd as J
    kc, oc = _two_band(J, 4, np.complex128, 0.6)
    kf, of = _two_band(J, 8, np.complex128, 0.6)
    cfg = dict(max_iter=200, tol_E=1e-13)
    f64 = J.SolverConfig(precision="fp64", **cfg)
    cold = J.solve(kf, np.zeros((8, 8, 2, 2), complex), 1.0, config=f64)
    res = J.solve_continuation(kc, kf, np.zeros((4, 4, 2, 2), complex), 1.0, 1.0,
                               coarse_config=f64, fine_config=f64)
    mixed = J.solve_continuation(kc, kf, np.zeros((4, 4, 2, 2), complex), 1.0, 1.0,
                                 coarse_config=J.SolverConfig(**cfg), fine_config=J.SolverConfig(**cfg))
    ref = D.solve_continuation(oc, of, np.zeros((4, 4, 2, 2), complex), 1.0, 1.0,
                               coarse_config=O.SolverConfig(**cfg), fine_config=O.SolverConfig(**cfg))
    assert cold.converged and res.fine.converged
    np.testing.assert_allclose(_np(res.P0_fine), ref.P0_fine, atol=1e-9)
    assert abs(float(res.fine.energy) - ref.fine.energy) < 1e-8
    assert abs(float(res.fine.energy) - float(cold.energy)) < 1e-8
    np.testing.assert_allclose(_np(res.fine.density), _np(cold.density), atol=1e-6)
    assert mixed.fine.converged
    assert abs(float(mixed.fine.energy) - ref.fine.energy) < 1e-8
    np.testing.assert_allclose(_np(mixed.fine.density), ref.fine.density, atol=1e-6)


def test_continuation_validation(dev):
    """tests/test_continuation.py:165-204."""
    import jax_hf_amd as J
    kc, _ = _two_band(J, 2)
    kf3 = J.HartreeFockKernel(np.ones((4, 4)), np.zeros((4, 4, 3, 3), np.complex64),
                              np.full((4, 4, 1, 1), 0.1, np.complex64), 0.1)
    P0 = np.zeros((2, 2, 2, 2), np.complex64)
    with pytest.raises(ValueError, match="orbital dimensions"):
        J.solve_continuation(kc, kf3, P0, 1.0, 1.0)
    kfh = J.HartreeFockKernel(np.ones((4, 4)), np.zeros((4, 4, 2, 2), np.complex64),
                              np.full((4, 4, 1, 1), 0.1, np.complex64), 0.1, include_hartree=True,
                              include_exchange=True, reference_density=np.zeros((4, 4, 2, 2), np.complex64),
                              hartree_matrix=np.eye(2))
    with pytest.raises(ValueError, match="include_hartree"):
        J.solve_continuation(kc, kfh, P0, 1.0, 1.0)
    kf, _ = _two_band(J, 4)
    with pytest.raises(ValueError, match="P0_coarse shape"):
        J.solve_continuation(kc, kf, np.zeros((3, 3, 2, 2), np.complex64), 1.0, 1.0)
    with pytest.raises(TypeError, match="coarse_config"):
        J.solve_continuation(kc, kf, P0, 1.0, 1.0, coarse_config={"max_iter": 5})


def _bistable(J, nk=2, T=0.02, ex=1.0, delta=0.1):
    """tests/test_deflation.py:50-64 on an nk x nk mesh with weights 1/nk^2
    (Sigma = -ex P at every k, so the analytic solutions are those of nk = 1)."""
    h = np.zeros((nk, nk, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -delta, delta
    args = (np.full((nk, nk), 1.0 / nk ** 2), h, np.full((nk, nk, 1, 1), ex), T)
    return J.HartreeFockKernel(*args), O.OracleKernel(*args), h


def test_deflation_finds_both_bistable_solutions(dev):
    """tests/test_deflation.py:156-182, and the oracle's deflated search."""
    import jax_hf_amd as J
    G, K0, h = _bistable(J)
    P0 = np.zeros_like(h)
    base = dict(optimizer="newton", tol_grad=1e-6, max_iter=300)
    res = J.solve_deflated(G, P0, 1.0, base_config=J.SolverConfig(**base), n_solutions=2, seed=0)
    ref = D.solve_deflated(K0, P0, 1.0, base_config=O.SolverConfig(**base), n_solutions=2, seed=0)
    assert isinstance(res, J.DeflatedResult)
    assert res.n_found == 2 == ref.n_found
    np.testing.assert_allclose(res.energies, [-0.6, -0.4], atol=3e-2)
    np.testing.assert_allclose(res.energies, ref.energies, atol=1e-8)
    w = np.full((2, 2), 0.25)
    for s in res.solutions:
        F, P = _np(s.fock), _np(s.density)
        comm = F @ P - P @ F
        assert np.sqrt(np.sum(w * np.sum(np.abs(comm) ** 2, axis=(-2, -1)))) < 1e-3
    assert np.abs(_np(res.solutions[0].density) - _np(res.solutions[1].density)).max() > 1e-2
    cold = J.solve(G, P0, 1.0, config=J.SolverConfig(**base))
    assert float(res.best.energy) <= float(cold.energy) + 1e-7


def test_deflation_off_is_a_noop_and_cg_rejects_it(dev):
    """tests/test_deflation.py:85-106."""
    import jax_hf_amd as J
    h = np.zeros((2, 2, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -0.5, 0.5
    h[..., 0, 1] = h[..., 1, 0] = 0.3
    G = J.HartreeFockKernel(np.ones((2, 2)), h, np.full((2, 2, 1, 1), 0.3), 0.05)
    P0 = np.zeros_like(h)
    base = dict(optimizer="newton", tol_grad=1e-7, max_iter=200)
    r1 = J.solve(G, P0, 4.0, config=J.SolverConfig(**base))
    r2 = J.solve(G, P0, 4.0, config=J.SolverConfig(deflation_sigma=0.0, **base))
    r3 = J.solve(G, P0, 4.0, config=J.SolverConfig(deflation_targets=_np(r1.density)[None], deflation_sigma=0.0,
                                                   **base))
    for r in (r2, r3):
        np.testing.assert_allclose(_np(r.density), _np(r1.density), atol=1e-12)
        assert abs(float(r.energy) - float(r1.energy)) < 1e-12
    with pytest.raises(ValueError, match="deflation"):
        J.solve(G, P0, 4.0, config=J.SolverConfig(optimizer="cg", deflation_targets=P0[None], deflation_sigma=1.0))
