"""GPU superlattice Fock build (jax_hf_amd.superlattice, hfx_superlattice_fock)
against the oracle (oracle/superlattice.py) on the reference tests' layout
(hexagonal Gamma + first shell, nk = 6, extended grid 36 x 36, 19 dG
channels): layout tables identical, V_lag_fft within 1e-12 of numpy.fft,
the exchange within 1e-12 (complex128) / 2e-6 (complex64) relative, the
build_fock adapter (exchange + scalar / orbital Hartree, refP, h) within
1e-12, zero density -> zero, the reference's orbital == scalar Hartree
equivalence and q = 0 warning, SCF / CG solves through the fock_build_fn
seam matching the oracle solvers with the oracle adapter, and the CG loop's
captured graph with the build replaying the eager loop bit for bit."""

import numpy as np
import pytest

import oracle as O
from oracle import superlattice as OS
from tests.test_oracle_superlattice import hex_g_basis, vfunc

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _rand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _layouts(nk, g_cut=1):
    from jax_hf_amd import superlattice as SL
    g, B = hex_g_basis(g_cut)
    return (SL.build_extended_layout(g, B, nk, nk, vfunc, 1.0 / nk ** 2),
            OS.build_extended_layout(g, B, nk, nk, vfunc, 1.0 / nk ** 2), g.shape[0])


def test_layout_matches_oracle(dev):
    lay, ref, n_G = _layouts(6)
    for f in ("N_ext_x", "N_ext_y", "n_supp_x", "n_supp_y", "n_delta"):
        assert getattr(lay, f) == getattr(ref, f), f
    for f in ("g_a_off", "pair_to_delta", "delta_pair_count", "delta_pair_i", "delta_pair_j", "delta_pair_start"):
        np.testing.assert_array_equal(getattr(lay, f), getattr(ref, f), err_msg=f)
    V = _np(lay.V_lag_fft)
    assert np.abs(V - ref.V_lag_fft).max() <= 1e-12 * np.abs(ref.V_lag_fft).max()


@pytest.mark.parametrize("dt", ["c128", "c64"])
def test_exchange_matches_oracle(dev, dt):
    import torch
    from jax_hf_amd.superlattice import make_superlattice_fock_fn
    nk, dorb = 6, 3
    lay, ref, n_G = _layouts(nk)
    D = n_G * dorb
    rho = _rand(np.random.default_rng(0), (nk, nk, D, D)) * 1e-3
    tdt, tol = (torch.complex128, 1e-12) if dt == "c128" else (torch.complex64, 2e-6)
    if dt == "c64":
        rho = rho.astype(np.complex64).astype(np.complex128)
    fn = make_superlattice_fock_fn(lay, n_G, dorb, nk, nk)
    out = fn(torch.from_numpy(rho).to(dev).to(tdt))
    want = OS.superlattice_exchange(rho, ref, n_G, dorb, nk, nk)
    assert out.dtype == tdt
    assert np.abs(_np(out).astype(np.complex128) - want).max() <= tol * np.abs(want).max()
    zero = fn(torch.zeros((nk, nk, D, D), dtype=tdt, device=dev))
    assert float(torch.abs(zero).max()) == 0.0


@pytest.mark.parametrize("orbital", [False, True])
def test_build_fock_adapter_matches_oracle(dev, orbital):
    import torch
    from jax_hf_amd.superlattice import make_superlattice_build_fock_fn
    nk, dorb = 6, 2
    lay, ref, n_G = _layouts(nk)
    D = n_G * dorb
    rng = np.random.default_rng(5)
    HH = rng.standard_normal((n_G, n_G))
    HH = HH + HH.T
    np.fill_diagonal(HH, 0.0)
    HHo = None
    if orbital:
        HHo = rng.standard_normal((n_G, n_G, dorb, dorb))
        HHo[np.arange(n_G), np.arange(n_G)] = 0.0
    P = _rand(rng, (nk, nk, D, D)) * 1e-2
    refP = _rand(rng, (nk, nk, D, D)) * 1e-3
    h = _rand(rng, (nk, nk, D, D))
    w = rng.uniform(0.5, 1.5, size=(nk, nk)) / nk ** 2
    kw = dict(include_exchange=True, include_hartree=True)
    S0, H0, F0 = OS.make_superlattice_build_fock_fn(ref, n_G, dorb, nk, nk, HH, 1.7, HH_GG_orbital=HHo)(
        P, h=h, refP=refP, w2d=w, **kw)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    S1, H1, F1 = make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, HH, 1.7, HH_GG_orbital=HHo)(
        t(P), h=t(h), refP=t(refP), w2d=t(w), VR=None, HH=None, **kw)
    for a, b in ((S1, S0), (F1, F0)):
        assert np.abs(_np(a) - b).max() <= 1e-12 * np.abs(b).max()
    assert float(torch.abs(H1).max()) == 0.0


def test_orbital_blind_hartree_equals_scalar_and_q0_warns(dev):
    """tests/test_superlattice.py:135-246 on the GPU."""
    import torch
    from jax_hf_amd.superlattice import make_superlattice_build_fock_fn
    nk, dorb = 6, 3
    lay, _, n_G = _layouts(nk)
    D = n_G * dorb
    rng = np.random.default_rng(11)
    HH = rng.standard_normal((n_G, n_G))
    HH = HH + HH.T
    np.fill_diagonal(HH, 0.0)
    HHo = np.broadcast_to(HH[:, :, None, None], (n_G, n_G, dorb, dorb)).copy()
    rho = _rand(rng, (nk, nk, D, D)) * 1e-3
    P = torch.from_numpy(0.5 * (rho + np.conj(np.swapaxes(rho, -1, -2)))).to(dev)
    kw = dict(h=torch.zeros_like(P), refP=torch.zeros_like(P), w2d=torch.full((nk, nk), 1.0 / nk ** 2, device=dev,
              dtype=torch.float64), include_exchange=False, include_hartree=True)
    Fs = make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, HH_GG=HH, hartree_degeneracy=1.6)(P, **kw)[2]
    Fo = make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, HH_GG=HH, hartree_degeneracy=1.6,
                                         HH_GG_orbital=HHo)(P, **kw)[2]
    assert float(torch.abs(Fo - Fs).max()) < 1e-12
    HHq = HHo.copy()
    HHq[np.arange(n_G), np.arange(n_G), 0, 0] = 0.5
    with pytest.warns(UserWarning, match="q=0"):
        bq = make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, HH_GG=np.zeros((n_G, n_G)),
                                             hartree_degeneracy=1.0, HH_GG_orbital=HHq)
    Fq = bq(P, **kw)[2]
    assert float(torch.abs(Fq - Fs).max()) > 1e-9


def _sl_problem(nk=6, dorb=1):
    g, B = hex_g_basis(1)
    n_G = g.shape[0]
    D = n_G * dorb
    k = (np.arange(nk) - nk // 2) / nk
    kx, ky = np.meshgrid(k, k, indexing="ij")
    h = np.zeros((nk, nk, D, D), complex)
    for A in range(n_G):
        p = (np.stack([kx, ky], -1) - g[A]) @ B.T
        for a in range(dorb):
            h[:, :, A * dorb + a, A * dorb + a] = 0.05 * np.sum(p ** 2, axis=-1) + 0.3 * a
    rng = np.random.default_rng(2)
    U = _rand(rng, (D, D)) * 0.2
    h = h + 0.5 * (U + U.conj().T)[None, None]
    args = (np.full((nk, nk), 1.0 / nk ** 2), h, np.full((nk, nk, 1, 1), 1.0), 0.05)
    return args, g, B, n_G, D


@pytest.mark.parametrize("solver", ["scf", "cg"])
def test_solve_through_fock_build_fn_matches_oracle(dev, solver):
    import jax_hf_amd as J
    from jax_hf_amd.superlattice import build_extended_layout, make_superlattice_build_fock_fn
    nk, dorb = 6, 1
    args, g, B, n_G, D = _sl_problem(nk, dorb)
    lay = build_extended_layout(g, B, nk, nk, vfunc, 1.0 / nk ** 2)
    ref = OS.build_extended_layout(g, B, nk, nk, vfunc, 1.0 / nk ** 2)
    HH = np.zeros((n_G, n_G))
    fb = make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, HH, 1.0)
    fb0 = OS.make_superlattice_build_fock_fn(ref, n_G, dorb, nk, nk, HH, 1.0)
    G = J.HartreeFockKernel(*args)
    K0 = O.OracleKernel(*args)
    P0 = np.zeros((nk, nk, D, D), complex)
    ne = 2.0
    if solver == "scf":
        cfg = dict(max_iter=200, density_tol=1e-10, comm_tol=1e-9, mixing=0.5)
        r = J.solve_scf(G, P0, ne, config=J.SCFConfig(**cfg), fock_build_fn=fb)
        r0 = O.solve_scf(K0, P0, ne, config=O.SCFConfig(**cfg), fock_build_fn=fb0)
        assert r.converged and r0.converged
        assert abs(float(r.energy) - float(r0.energy)) < 1e-10
        assert np.abs(_np(r.density_matrix) - r0.density_matrix).max() < 1e-8
    else:
        cfg = dict(max_iter=300, tol_E=1e-12)
        r = J.solve(G, P0, ne, config=J.SolverConfig(precision="fp64", **cfg), fock_build_fn=fb)
        r0 = O.solve(K0, P0, ne, config=O.SolverConfig(**cfg), fock_build_fn=fb0)
        assert r.converged and r0.converged
        assert abs(float(r.energy) - float(r0.energy)) < 1e-9


def test_cg_graph_with_superlattice_build_replays_eager_loop(dev):
    """The CG loop captures the superlattice build (hfx_graph_safe): every
    iteration's energy and the final density equal the eager loop's bit for bit."""
    import jax_hf_amd as J
    from jax_hf_amd.superlattice import build_extended_layout, make_superlattice_build_fock_fn
    nk, dorb = 6, 1
    args, g, B, n_G, D = _sl_problem(nk, dorb)
    lay = build_extended_layout(g, B, nk, nk, vfunc, 1.0 / nk ** 2)
    fb = make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, np.zeros((n_G, n_G)), 1.0)
    G = J.HartreeFockKernel(*args)
    P0 = np.zeros((nk, nk, D, D), complex)
    res = {}
    for use_graph in (False, True):
        cfg = J.SolverConfig(precision="fp64", max_iter=40, tol_E=1e-13, use_graph=use_graph)
        res[use_graph] = J.solve(G, P0, 2.0, config=cfg, fock_build_fn=fb)
    # a second graph solve reuses the cached runner and its captured graphs
    res["again"] = J.solve(G, P0, 2.0, config=J.SolverConfig(precision="fp64", max_iter=40, tol_E=1e-13,
                                                              use_graph=True), fock_build_fn=fb)
    a = res[False]
    assert a.n_iter > 3
    for b in (res[True], res["again"]):
        assert a.n_iter == b.n_iter
        np.testing.assert_array_equal(_np(a.history["E"]), _np(b.history["E"]))
        np.testing.assert_array_equal(_np(a.density), _np(b.density))


def _sl_hartree_setup(nk, dorb, orbital, seed=3):
    """A superlattice problem with the Hartree shift on (refP = 0, the kernel's
    own Hartree matrix unused by the plug-in) and the GPU / oracle builders."""
    from jax_hf_amd.superlattice import build_extended_layout, make_superlattice_build_fock_fn
    args, g, B, n_G, D = _sl_problem(nk, dorb)
    rng = np.random.default_rng(seed)
    HH = np.abs(rng.standard_normal((n_G, n_G))) * 0.05
    HH = HH + HH.T
    np.fill_diagonal(HH, 0.0)
    HHo = None
    if orbital:
        HHo = np.abs(rng.standard_normal((n_G, n_G, dorb, dorb))) * 0.05
        HHo = HHo + HHo.transpose(1, 0, 3, 2)
        HHo[np.arange(n_G), np.arange(n_G)] = 0.0
    lay = build_extended_layout(g, B, nk, nk, vfunc, 1.0 / nk ** 2)
    ref = OS.build_extended_layout(g, B, nk, nk, vfunc, 1.0 / nk ** 2)
    fb = make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, HH, 1.3, HH_GG_orbital=HHo)
    fb0 = OS.make_superlattice_build_fock_fn(ref, n_G, dorb, nk, nk, HH, 1.3, HH_GG_orbital=HHo)
    kw = dict(include_hartree=True, hartree_matrix=np.zeros((D, D)),
              reference_density=np.zeros((nk, nk, D, D), complex))
    return args, kw, fb, fb0, D


@pytest.mark.parametrize("orbital", [False, True])
@pytest.mark.parametrize("form", ["scf", "cg_fp64", "cg_mixed"])
def test_solve_with_superlattice_hartree_matches_oracle(dev, orbital, form):
    """Solves through fock_build_fn with the superlattice Hartree shift (scalar
    and orbital tables, dim_orb 2) against the oracle solvers with the oracle
    adapter; the CG forms run the build inside the captured iteration graph,
    the mixed one in complex64 with the FP64 polish through the same plug-in."""
    import jax_hf_amd as J
    nk, dorb = 6, 2
    args, kw, fb, fb0, D = _sl_hartree_setup(nk, dorb, orbital)
    G = J.HartreeFockKernel(*args, **kw)
    K0 = O.OracleKernel(*args, **kw)
    P0 = np.zeros((nk, nk, D, D), complex)
    ne = 3.0
    scf_cfg = dict(max_iter=300, density_tol=1e-11, comm_tol=1e-10, acceleration="diis")
    ref = O.solve_scf(K0, P0, ne, config=O.SCFConfig(**scf_cfg), fock_build_fn=fb0)
    assert ref.converged
    if form == "scf":
        r = J.solve_scf(G, P0, ne, config=J.SCFConfig(**scf_cfg), fock_build_fn=fb)
        assert r.converged
        assert abs(float(r.energy) - ref.energy) < 1e-10
        assert np.abs(_np(r.density_matrix) - ref.density_matrix).max() < 1e-8
        return
    prec = "fp64" if form == "cg_fp64" else "mixed"
    cfg = dict(max_iter=400, tol_E=1e-12, precision=prec)
    if prec == "fp64":
        cfg.update(polish="diis")
    r = J.solve(G, P0, ne, config=J.SolverConfig(**cfg), fock_build_fn=fb)
    assert r.converged
    # both forms end on the FP64 SCF fixed point through the polish
    assert abs(float(r.energy) - ref.energy) < 1e-9
    assert np.abs(_np(r.density) - ref.density_matrix).max() < 1e-6
    # and the FP64 CG trajectory follows the oracle CG's first iterations
    if prec == "fp64":
        r0 = O.solve(K0, P0, ne, config=O.SolverConfig(max_iter=400, tol_E=1e-12), fock_build_fn=fb0)
        np.testing.assert_allclose(_np(r.history["E"])[:5], r0.history["E"][:5], atol=1e-10)


def test_cg_graph_with_superlattice_hartree_replays_eager_loop(dev):
    """The captured CG iteration with the Hartree-moment launches of the
    superlattice build replays the eager loop bit for bit."""
    import jax_hf_amd as J
    nk, dorb = 6, 2
    args, kw, fb, _, D = _sl_hartree_setup(nk, dorb, True)
    G = J.HartreeFockKernel(*args, **kw)
    P0 = np.zeros((nk, nk, D, D), complex)
    res = {}
    for use_graph in (False, True):
        cfg = J.SolverConfig(precision="fp64", max_iter=30, tol_E=1e-13, use_graph=use_graph)
        res[use_graph] = J.solve(G, P0, 3.0, config=cfg, fock_build_fn=fb)
    a, b = res[False], res[True]
    assert a.n_iter > 3 and a.n_iter == b.n_iter
    np.testing.assert_array_equal(_np(a.history["E"]), _np(b.history["E"]))
    np.testing.assert_array_equal(_np(a.density), _np(b.density))


def test_adapter_checks_shapes_and_broadcasts_h(dev):
    """h / refP must be per-k (nkx, nky, D, D) or one (D, D) matrix (broadcast
    over k as h + Sigma broadcasts); w2d must hold nkx * nky weights."""
    import torch
    from jax_hf_amd.superlattice import make_superlattice_build_fock_fn
    nk, dorb = 6, 2
    lay, ref, n_G = _layouts(nk)
    D = n_G * dorb
    rng = np.random.default_rng(8)
    HH = np.zeros((n_G, n_G))
    P = _rand(rng, (nk, nk, D, D)) * 1e-2
    hk = _rand(rng, (D, D))
    w = np.full((nk, nk), 1.0 / nk ** 2)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    fb = make_superlattice_build_fock_fn(lay, n_G, dorb, nk, nk, HH, 1.0)
    fb0 = OS.make_superlattice_build_fock_fn(ref, n_G, dorb, nk, nk, HH, 1.0)
    kw = dict(include_exchange=True, include_hartree=True)
    S1, _, F1 = fb(t(P), h=t(hk), refP=None, w2d=t(w), **kw)
    S0, _, F0 = fb0(P, h=np.broadcast_to(hk, P.shape), refP=None, w2d=w, **kw)
    assert np.abs(_np(F1) - F0).max() <= 1e-12 * np.abs(F0).max()
    with pytest.raises(ValueError, match="h of shape"):
        fb(t(P), h=t(hk[:-1]), refP=None, w2d=t(w), **kw)
    with pytest.raises(ValueError, match="refP of shape"):
        fb(t(P), h=t(hk), refP=t(P[:, :-1]), w2d=t(w), **kw)
    with pytest.raises(ValueError, match="w2d has"):
        fb(t(P), h=t(hk), refP=None, w2d=t(w[:-1]), **kw)
