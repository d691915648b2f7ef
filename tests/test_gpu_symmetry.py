"""GPU symmetry projector (jax_hf_amd.symmetry, src/jax_hf/symmetry.py) against
the oracle restatement (oracle/symmetry.py) on the same inputs, for every
form of make_project_fn: 1e-13 (complex128) / 2e-6 (complex64) relative.
nb = 16 runs the conjugations on the matrix cores.  Solver integration: a
CG solve with the projector matches the oracle CG with the oracle projector,
and its captured graph replays bit-identically to the eager loop."""

import numpy as np
import pytest

import oracle as O
from oracle import symmetry as OS
from jax_hf_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _unitary(rng, nb):
    return np.linalg.qr(_rand(rng, (nb, nb)))[0]


def _c3_maps(nk):
    half = nk // 2
    ii, jj = np.meshgrid(np.arange(nk), np.arange(nk), indexing="ij")
    return np.stack([ii, (-jj) % nk, (jj - ii + half) % nk]), np.stack([jj, (ii - jj + half) % nk, (-ii) % nk])


def _cases(nb, nk, rng):
    perm = np.roll(np.eye(nb), 1, axis=1)
    z = np.diag(np.where(np.arange(nb) % 2 == 0, 1.0, -1.0)).astype(complex)
    w = np.exp(2j * np.pi / 3)
    C = np.diag(w ** (np.arange(nb) % 3))
    ii, jj = _c3_maps(nk)
    ti, tj = np.meshgrid(np.arange(nk), np.arange(nk), indexing="ij")
    valid = rng.uniform(size=(nk, nk)) > 0.3
    orbit = rng.uniform(size=(nk, nk)) > 0.2
    Uq = _unitary(rng, nb)
    return {
        "unitary": dict(unitary_group=np.stack([np.eye(nb), perm, perm @ perm])),
        "tr_flip": dict(time_reversal_U=Uq, time_reversal_k_convention="flip"),
        "tr_mod": dict(time_reversal_U=Uq, time_reversal_k_convention="mod"),
        "combined": dict(unitary_group=np.stack([np.eye(nb), z]), spatial_group=Uq[None],
                         spatial_k_convention="flip"),
        "spatial_only_axis0": dict(spatial_group=Uq[None], spatial_k_convention="mod", spatial_k_flip_axes=(0,)),
        "maps_masked_plus_unitary": dict(spatial_group=np.stack([np.eye(nb), C, C @ C]),
                                         spatial_k_index_maps=(ii, jj), spatial_k_orbit_mask=orbit,
                                         unitary_group=np.stack([np.eye(nb), z])),
        "tr_map_masked": dict(time_reversal_U=Uq, time_reversal_k_index_map=((-ti) % nk, (-tj) % nk, valid)),
        "all": dict(unitary_group=np.stack([np.eye(nb), z]), spatial_group=_unitary(rng, nb)[None],
                    time_reversal_U=Uq),
    }


@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("nb,nk", [(4, 6), (16, 9)])
def test_projector_matches_oracle(dev, nb, nk, dt):
    import torch
    from jax_hf_amd.symmetry import make_project_fn
    rng = np.random.default_rng(nb + nk)
    tdt = torch.complex128 if dt == "c128" else torch.complex64
    tol = 1e-13 if dt == "c128" else 2e-6
    A = _rand(rng, (nk, nk, nb, nb))
    if dt == "c64":
        A = A.astype(np.complex64).astype(np.complex128)
    At = torch.from_numpy(A).to(dev).to(tdt)
    for name, kw in _cases(nb, nk, rng).items():
        out = make_project_fn(**kw)(At)
        ref = OS.make_project_fn(**kw)(A)
        assert out.dtype == tdt and out.shape == At.shape
        err = np.abs(_np(out).astype(np.complex128) - ref).max() / np.abs(ref).max()
        assert err <= tol, (name, err)
    assert np.array_equal(_np(At), A.astype(_np(At).dtype))          # input untouched


def test_projector_identity_and_errors(dev):
    import torch
    from jax_hf_amd.symmetry import make_project_fn
    A = torch.ones((3, 3, 2, 2), dtype=torch.complex128, device=dev)
    assert make_project_fn()(A) is A
    ii, jj = _c3_maps(4)
    with pytest.raises(ValueError, match="spatial_group"):
        make_project_fn(spatial_k_index_maps=(ii, jj))
    with pytest.raises(ValueError, match="leading dim"):
        make_project_fn(spatial_group=np.stack([np.eye(2)] * 3), spatial_k_index_maps=(ii[:2], jj[:2]))
    bad = make_project_fn(spatial_group=np.eye(2)[None], spatial_k_convention="rot")
    with pytest.raises(ValueError, match="k_convention"):
        bad(A)
    wrong = make_project_fn(spatial_group=np.stack([np.eye(2)] * 3), spatial_k_index_maps=(ii + 5, jj))
    with pytest.raises(ValueError, match="outside the mesh"):
        wrong(torch.ones((4, 4, 2, 2), dtype=torch.complex128, device=dev))


def _sym_problem():
    c = syn.config0(nk=12, nb=4, seed=0)
    z = np.diag([1.0, 1.0, -1.0, -1.0]).astype(complex)
    kw = dict(unitary_group=np.stack([np.eye(4), z]), time_reversal_U=np.eye(4))
    return c, kw


def test_cg_with_projector_matches_oracle(dev):
    import jax_hf_amd as J
    from jax_hf_amd.symmetry import make_project_fn
    c, kw = _sym_problem()
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    K0 = O.OracleKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    r = J.solve(G, c["P0"], c["n_electrons"],
                config=J.SolverConfig(max_iter=300, tol_E=1e-12, project_fn=make_project_fn(**kw)))
    r0 = O.solve(K0, c["P0"], c["n_electrons"],
                 config=O.SolverConfig(max_iter=300, tol_E=1e-12, project_fn=OS.make_project_fn(**kw)))
    assert r.converged and r0.converged
    assert abs(float(r.energy) - r0.energy) < 1e-10
    assert np.abs(_np(r.density) - r0.density).max() < 1e-6
    # the returned density is symmetric
    P = _np(r.density)
    assert np.abs(OS.make_project_fn(**kw)(P) - P).max() < 1e-10


def test_cg_with_projector_graph_replays_eager_loop(dev):
    import jax_hf_amd as J
    from jax_hf_amd.symmetry import make_project_fn
    c, kw = _sym_problem()
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    cfg = dict(max_iter=40, tol_E=1e-12, project_fn=make_project_fn(**kw))
    rg = J.solve(G, c["P0"], c["n_electrons"], config=J.SolverConfig(use_graph=True, **cfg))
    re = J.solve(G, c["P0"], c["n_electrons"], config=J.SolverConfig(use_graph=False, **cfg))
    assert rg.n_iter == re.n_iter
    np.testing.assert_array_equal(_np(rg.density), _np(re.density))
    np.testing.assert_array_equal(_np(rg.history["E"]), _np(re.history["E"]))
