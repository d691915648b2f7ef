"""Batched scan solves (solve_batch) against one-by-one solve calls.

The stacked CG runs the same per-solve algorithm, so every solve's result
must match the serial solve: FP64 energies within 1e-9 per k-point (the
batched reductions sum in a different order, so the trajectories agree to
rounding, not bitwise), iteration counts within 2, densities within 1e-5;
the mixed form (FP32 batched CG + FP64 DIIS polish) lands on the same FP64
fixed point as the serial mixed solve within 1e-9.  Solves freeze at
different iterations (different fillings and displacement fields).
"""

import numpy as np
import pytest

from jax_hf_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _scan(nks, nb=8, T=0.1, seed=5):
    import jax_hf_amd as J
    rng = np.random.default_rng(seed)
    kernels, ne = [], []
    for i, nk in enumerate(nks):
        h = syn.tight_binding(nk, nk, nb, seed=seed + nk) + syn.displacement(nb, float(rng.uniform(0.0, 0.4)))
        V = syn.coulomb_layer(nk, nk, nb, 2)
        kernels.append(J.HartreeFockKernel(syn.uniform_weights(nk, nk), np.ascontiguousarray(h), V, T))
        ne.append(nb / 2.0 + float(rng.uniform(-0.5, 0.5)))
    P0s = [np.zeros((nk, nk, nb, nb), np.complex128) for nk in nks]
    return kernels, P0s, ne


def test_solve_batch_fp64_matches_serial(dev):
    import jax_hf_amd as J
    kernels, P0s, ne = _scan([12, 12, 12, 16, 16, 12])
    cfg = J.SolverConfig(max_iter=300, tol_E=1e-10)
    batch = J.solve_batch(kernels, P0s, ne, config=cfg)
    iters = set()
    for kern, P0, n, rb in zip(kernels, P0s, ne, batch):
        rs = J.solve(kern, P0, n, config=cfg)
        assert rb.converged == rs.converged
        assert abs(float(rb.energy) - float(rs.energy)) < 1e-9
        assert abs(rb.n_iter - rs.n_iter) <= 2
        assert np.abs(_np(rb.density) - _np(rs.density)).max() < 1e-5
        assert abs(float(rb.mu) - float(rs.mu)) < 1e-6
        w = _np(kern.w2d)
        assert abs(float(np.sum(w[..., None] * _np(rb.p))) - n) < 1e-9     # electron count per solve
        iters.add(rb.n_iter)
    assert len(iters) > 1            # solves froze at different iterations


def test_solve_batch_mixed_reaches_fixed_point(dev):
    import jax_hf_amd as J
    kernels, P0s, ne = _scan([12, 12, 16], nb=4, seed=7)
    cfg = J.SolverConfig(max_iter=300, precision="mixed")
    batch = J.solve_batch(kernels, P0s, ne, config=cfg)
    for kern, P0, n, rb in zip(kernels, P0s, ne, batch):
        rs = J.solve(kern, P0, n, config=cfg)
        assert rb.converged and rs.converged
        assert abs(float(rb.energy) - float(rs.energy)) < 1e-9


def test_solve_batch_mixed_mu_completion_paths(dev, monkeypatch):
    """One Newton mu step per captured batched mu solve forces the completion
    on the readback (CG) and the density repair (pipelined polish); the batch
    still lands on the serial solves' fixed points."""
    import jax_hf_amd as J
    from jax_hf_amd import batch as B
    kernels, P0s, ne = _scan([12, 12, 16], nb=4, seed=7)
    cfg = J.SolverConfig(max_iter=300, precision="mixed")
    serial = [J.solve(k, P0, n, config=cfg) for k, P0, n in zip(kernels, P0s, ne)]
    monkeypatch.setattr(B._BatchCG, "MU_LOOP_STEPS", 1)
    monkeypatch.setattr(B._BatchPolish, "MU_STEPS", 1)
    batch = J.solve_batch(kernels, P0s, ne, config=cfg)
    for rs, rb in zip(serial, batch):
        assert rb.converged and rs.converged
        assert abs(float(rb.energy) - float(rs.energy)) < 1e-9


def test_solve_batch_falls_back_for_unbatchable(dev):
    """A Hartree kernel runs through solve() and gives solve()'s result."""
    import jax_hf_amd as J
    kernels, P0s, ne = _scan([12, 12], nb=4, seed=9)
    nb = 4
    h = syn.tight_binding(12, 12, nb, seed=3)
    HH = np.full((nb, nb), 0.2)
    kh = J.HartreeFockKernel(syn.uniform_weights(12, 12), h, syn.coulomb_scalar(12, 12), 0.1,
                             include_hartree=True, hartree_matrix=HH, reference_density=np.zeros_like(h))
    cfg = J.SolverConfig(max_iter=200, tol_E=1e-9)
    out = J.solve_batch([kernels[0], kh, kernels[1]], [P0s[0], np.zeros_like(h), P0s[1]],
                        [ne[0], 2.0, ne[1]], config=cfg)
    ref = J.solve(kh, np.zeros_like(h), 2.0, config=cfg)
    assert float(out[1].energy) == float(ref.energy)


def test_solve_batch_streams_match_one_stream(dev):
    """Mesh groups interleaved on four streams give bit-identical results to
    the groups run one after another on the caller's stream, and the results
    are ready for the caller's stream on return."""
    import torch
    import jax_hf_amd as J
    kernels, P0s, ne = _scan([12, 16, 12, 20, 16], nb=4, seed=11)
    for prec in ("fp64", "mixed"):
        cfg = J.SolverConfig(max_iter=300, precision=prec, tol_E=1e-10)
        one = J.solve_batch(kernels, P0s, ne, config=cfg, streams=1)
        four = J.solve_batch(kernels, P0s, ne, config=cfg, streams=4)
        for a, b in zip(one, four):
            assert float(a.energy) == float(b.energy)
            assert int(a.n_iter) == int(b.n_iter) and bool(a.converged) == bool(b.converged)
            assert torch.equal(a.density, b.density) and torch.equal(a.fock, b.fock)
            assert torch.equal(a.Q, b.Q) and torch.equal(a.p, b.p)
            assert float(a.mu) == float(b.mu)
