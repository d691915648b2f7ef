"""Pins the oracle's host drivers (oracle/drivers.py).

resample_kgrid: against numpy.interp with period (periodic linear
interpolation, a library routine, applied axis by axis on the centred
coordinate), identity at equal size, the coarse samples reproduced at
coincident fine points, the 2n -> n round trip, constants preserved and the
argument errors of utils.py:512-535.

solve_continuation / solve_deflated: the reference tests' cases
(tests/test_continuation.py, tests/test_deflation.py) on the oracle solvers:
the continuation equals the manual coarse-then-fine driver and lands on the
cold fine solve's fixed point; deflation on the bistable model finds both
analytic solutions E = -delta - ex/2 (ground) and +delta - ex/2."""

import numpy as np
import pytest

import oracle as O
from oracle import drivers as D


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _interp_axis(x, n_new, axis):
    """Periodic linear interpolation by numpy.interp on the centred coordinate."""
    n_old = x.shape[axis]
    if n_old == n_new:
        return x
    xm = np.moveaxis(x, axis, 0)
    u = (np.arange(n_new) - n_new // 2) * (n_old / n_new) + n_old // 2
    flat = xm.reshape(n_old, -1)
    cols = [np.interp(u, np.arange(n_old), flat[:, c].real, period=n_old)
            + 1j * np.interp(u, np.arange(n_old), flat[:, c].imag, period=n_old) for c in range(flat.shape[1])]
    out = np.stack(cols, axis=1).reshape((n_new,) + xm.shape[1:])
    return np.moveaxis(out, 0, axis)


@pytest.mark.parametrize("n_old,n_new", [(2, 4), (4, 8), (6, 4), (3, 4), (5, 12), (12, 5), (8, 8)])
def test_resample_matches_numpy_interp(n_old, n_new):
    rng = np.random.default_rng(n_old * 31 + n_new)
    x = _rand(rng, (n_old, n_old, 2, 2))
    ref = _interp_axis(_interp_axis(x, n_new, 0), n_new, 1)
    out = D.resample_kgrid(x, n_new)
    assert out.shape == (n_new, n_new, 2, 2)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-13)


def test_resample_non_square_input_and_scalar_trailing():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((3, 5))
    ref = _interp_axis(_interp_axis(x.astype(complex), 6, 0), 6, 1).real
    np.testing.assert_allclose(D.resample_kgrid(x, 6), ref, atol=1e-13)


def test_resample_exact_properties():
    rng = np.random.default_rng(3)
    n = 6
    x = _rand(rng, (n, n, 3, 3))
    assert D.resample_kgrid(x, n) is x
    fine = D.resample_kgrid(x, 2 * n)
    # coincident points: fine indices whose source coordinate u is an integer
    J = np.arange(2 * n)
    u = (J - n) * 0.5 + n // 2
    hit = J[u == np.floor(u)]
    src = (u[hit].astype(int)) % n
    np.testing.assert_array_equal(fine[np.ix_(hit, hit)], x[np.ix_(src, src)])
    np.testing.assert_array_equal(D.resample_kgrid(fine, n), x)                # 2n -> n round trip
    c = np.full((4, 4, 2, 2), 0.3 - 0.7j)
    np.testing.assert_allclose(D.resample_kgrid(c, 7), np.full((7, 7, 2, 2), 0.3 - 0.7j), atol=1e-15)


def test_resample_float32_rounds_in_float32():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3, 3, 2)).astype(np.float32)
    out = D.resample_kgrid(x, 7)
    assert out.dtype == np.float32
    np.testing.assert_allclose(out, _interp_axis(_interp_axis(x.astype(complex), 7, 0), 7, 1).real, atol=1e-6)


def test_resample_errors():
    with pytest.raises(ValueError, match="at least 2"):
        D.resample_kgrid(np.ones(4), 2)
    with pytest.raises(ValueError, match="Only 'linear'"):
        D.resample_kgrid(np.ones((2, 2)), 4, method="cubic")


def _two_band(nk, exchange=0.25, T=0.2):
    h = np.zeros((nk, nk, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -0.5, 0.5
    return O.OracleKernel(np.ones((nk, nk)), h, np.full((nk, nk, 1, 1), exchange), T)


def test_continuation_matches_manual_driver():
    kc, kf = _two_band(2), _two_band(4)
    cfg = O.SolverConfig(max_iter=50, tol_E=1e-8)
    P0 = np.zeros_like(kc.h)
    res = D.solve_continuation(kc, kf, P0, 1.0, 1.0, coarse_config=cfg, fine_config=cfg)
    ec = O.solve(kc, P0, 1.0, config=cfg)
    Pf = D.resample_kgrid(ec.density, 4)
    Pf = 0.5 * (Pf + np.conj(np.swapaxes(Pf, -1, -2)))
    ef = O.solve(kf, Pf, 1.0, config=cfg)
    np.testing.assert_allclose(res.P0_fine, Pf, atol=1e-15)
    np.testing.assert_allclose(res.fine.density, ef.density, atol=1e-12)
    scf = D.solve_continuation(kc, kf, P0, 1.0, 1.0, coarse_config=O.SCFConfig(max_iter=40, mixing=0.5),
                               fine_config=cfg)
    assert isinstance(scf.coarse, O.SCFResult) and scf.fine.converged
    with pytest.raises(TypeError, match="coarse_config"):
        D.solve_continuation(kc, kf, P0, 1.0, 1.0, coarse_config={"max_iter": 5})


def test_continuation_lands_on_cold_fixed_point():
    # tol_E tighter than the reference test's 1e-10: in complex128 the cold
    # solve's windowed-energy stop fires 3e-6 above the fixed point at 1e-10
    kc, kf = _two_band(4, 0.6), _two_band(8, 0.6)
    cfg = O.SolverConfig(max_iter=200, tol_E=1e-13)
    cold = O.solve(kf, np.zeros_like(kf.h), 1.0, config=cfg)
    cont = D.solve_continuation(kc, kf, np.zeros_like(kc.h), 1.0, 1.0, coarse_config=cfg, fine_config=cfg)
    assert cold.converged and cont.fine.converged
    assert abs(cont.fine.energy - cold.energy) < 1e-10
    assert np.abs(cont.fine.density - cold.density).max() < 1e-6


def bistable(nk, T=0.02, ex=1.0, delta=0.1):
    """Two-band model with two self-consistent solutions when ex > 2 delta
    (tests/test_deflation.py:50-64); weights 1/nk^2 so Sigma = -ex P at every k."""
    h = np.zeros((nk, nk, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -delta, delta
    return dict(weights=np.full((nk, nk), 1.0 / nk ** 2), hamiltonian=h, coulomb_q=np.full((nk, nk, 1, 1), ex),
                T=T)


@pytest.mark.parametrize("nk", [1, 2])
def test_deflation_finds_both_bistable_solutions(nk):
    c = bistable(nk)
    K = O.OracleKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    base = O.SolverConfig(optimizer="newton", tol_grad=1e-6, max_iter=300)
    res = D.solve_deflated(K, np.zeros_like(c["hamiltonian"]), 1.0, base_config=base, n_solutions=2, seed=0)
    assert res.n_found == 2
    np.testing.assert_allclose(res.energies, [-0.6, -0.4], atol=3e-2)
    assert np.abs(res.solutions[0].density - res.solutions[1].density).max() > 1e-2
    for s in res.solutions:
        comm = s.fock @ s.density - s.density @ s.fock
        assert np.sqrt(np.sum(K.w2d * np.sum(np.abs(comm) ** 2, axis=(-2, -1)))) < 1e-3
    cold = O.solve(K, np.zeros_like(c["hamiltonian"]), 1.0, config=base)
    assert res.best.energy <= cold.energy + 1e-7


def test_deflation_bias_gradient_identity_and_closed_forms():
    """tests/test_deflation.py:109-150 on the oracle's bias: dPhi/dP_k = w_k S_pen_k
    by central differences along a Hermitian direction (O(eps^2) convergence),
    Phi and the diagonal curvature against their closed forms."""
    rng = np.random.default_rng(0)
    nk, nb = 2, 2
    w2d = np.arange(nk * nk).reshape(nk, nk).astype(np.float64) + 1.0
    sigma, length = 0.7, 0.5
    inv2L2 = 1.0 / (2.0 * length * length)

    def herm(shape):
        A = _rand(rng, shape)
        return 0.5 * (A + np.conj(np.swapaxes(A, -1, -2)))

    P = herm((nk, nk, nb, nb))
    targets = np.stack([herm((nk, nk, nb, nb)) for _ in range(2)])
    Phi, S, diag = O.deflation_bias_terms(P, targets, w2d, sigma, inv2L2)
    dP = herm((nk, nk, nb, nb))
    analytic = float(np.sum(w2d[..., None, None] * np.real(np.conj(S) * dP)))
    fd = [(O.deflation_bias_terms(P + e * dP, targets, w2d, sigma, inv2L2)[0]
           - O.deflation_bias_terms(P - e * dP, targets, w2d, sigma, inv2L2)[0]) / (2 * e) for e in (1e-4, 1e-5)]
    assert abs(fd[1] - analytic) < 1e-4 * (abs(analytic) + 1.0)
    assert abs(fd[1] - analytic) < 0.2 * abs(fd[0] - analytic) + 1e-9
    d2 = np.array([float(np.sum(w2d[..., None, None] * np.abs(P - t) ** 2)) for t in targets])
    phi = sigma * np.exp(-d2 * inv2L2)
    assert abs(Phi - phi.sum()) < 1e-12
    assert abs(diag - 2.0 * float(np.sum(-phi * inv2L2))) < 1e-12
