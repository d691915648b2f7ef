"""Exchange on meshes outside the fixed kernels: prime factors above 7 (the
generic radix stage) and planes beyond LDS (the HBM-staged three-pass
exchange), against the FP64 oracle's batched FFTs (numpy.fft on any mesh).

Bar: FP64 Sigma / F within 1e-11 relative of the oracle (the FP64 target),
complex64 within 2e-6; VR (the forward transform with the ifftshift phase)
within 1e-12 (FP64: the rounding of a length-3000 transform reaches 2e-13).
"""

import numpy as np
import pytest

import oracle as O
from jax_hf_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# (n1, n2, nb, layer): primes 11 / 13 / 97 (generic stage), 22 x 26 (2 * prime),
# 128 x 128 and 120 x 100 (planes beyond the FP64 LDS plane: staged, 256-thread
# tiles), 4 x 3000 and 3000 x 4 (staged, 1024-thread tiles: one row / column of
# 3000 exceeds the small tile)
MESHES = [(11, 11, 2, False), (13, 7, 3, True), (22, 26, 2, True), (97, 5, 2, False),
          (128, 128, 2, True), (120, 100, 2, False), (4, 3000, 2, False), (3000, 4, 2, True)]


@pytest.mark.parametrize("planar", [True, False])
@pytest.mark.parametrize("n1,n2,nb,layer", MESHES)
def test_any_mesh_fock_fp64_matches_oracle(dev, n1, n2, nb, layer, planar, monkeypatch):
    import jax_hf_amd as J
    from jax_hf_amd import fock as fock_mod
    monkeypatch.setattr(fock_mod, "PLANAR", planar)
    print(f"mesh {n1}x{n2} nb={nb} layer={layer} planar={planar}", flush=True)
    rng = np.random.default_rng(n1 * 1000 + n2)
    h = syn.tight_binding(n1, n2, nb, seed=n1 + n2)
    V = syn.coulomb_layer(n1, n2, nb, 2) if layer else syn.coulomb_scalar(n1, n2)
    w = rng.uniform(0.5, 1.5, size=(n1, n2)) / (n1 * n2)
    P = syn.haar_density(n1, n2, nb, rng)
    refP = 0.3 * syn.haar_density(n1, n2, nb, rng)
    HH = rng.normal(size=(nb, nb))
    HH = HH + HH.T
    K = O.OracleKernel(w, h, V, 0.1, include_hartree=True, reference_density=refP, hartree_matrix=HH)
    G = J.HartreeFockKernel(w, h, V, 0.1, include_hartree=True, reference_density=refP, hartree_matrix=HH)
    assert _rel(_np(G._VR_shifted), K._VR_shifted) < 1e-12
    assert _rel(_np(G.VR), K.VR) < 1e-12
    a, b = G.as_args(), K.as_args()
    S0, H0, F0 = O.build_fock(P, h=b["h"], VR=b["VR"], refP=b["refP"], HH=b["HH"], w2d=K.w2d,
                              include_exchange=True, include_hartree=True,
                              exchange_hermitian_channel_packing=b["exchange_hermitian_channel_packing"])
    S1, H1, F1 = J.build_fock(P, h=a["h"], VR=a["VR"], refP=a["refP"], HH=a["HH"], w2d=G.w2d,
                              include_exchange=True, include_hartree=True,
                              exchange_hermitian_channel_packing=a["exchange_hermitian_channel_packing"])
    assert _rel(_np(S1), S0) < 1e-11
    assert _rel(_np(H1), H0) < 1e-11
    assert _rel(_np(F1), F0) < 1e-11
    E0 = O.hf_energy(P, h=K.h, Sigma=S0, H=H0, weights_b=K.weights_b)
    E1 = float(J.hf_energy(P, h=G.h, Sigma=S1, H=H1, weights_b=G.weights_b))
    assert abs(E1 - E0) <= 1e-11 * max(1.0, abs(E0))


@pytest.mark.parametrize("n1,n2", [(11, 11), (150, 150), (4, 6000)])
def test_any_mesh_fock_complex64(dev, n1, n2):
    """complex64: 150 x 150 and 4 x 6000 exceed the complex64 LDS plane (staged)."""
    import jax_hf_amd as J
    nb = 2
    print(f"c64 mesh {n1}x{n2}", flush=True)
    rng = np.random.default_rng(n1 + 7 * n2)
    h = syn.tight_binding(n1, n2, nb, seed=3)
    V = syn.coulomb_layer(n1, n2, nb, 2)
    w = np.full((n1, n2), 1.0 / (n1 * n2))
    P = syn.haar_density(n1, n2, nb, rng)
    K = O.OracleKernel(w, h, V, 0.1)
    G = J.HartreeFockKernel(w, h.astype(np.complex64), V, 0.1)
    a, b = G.as_args(), K.as_args()
    S0, H0, F0 = O.build_fock(P, h=b["h"], VR=b["VR"], refP=b["refP"], HH=b["HH"], w2d=K.w2d,
                              include_exchange=True, include_hartree=False,
                              exchange_hermitian_channel_packing=b["exchange_hermitian_channel_packing"])
    S1, H1, F1 = J.build_fock(P.astype(np.complex64), h=a["h"], VR=a["VR"], refP=a["refP"], HH=a["HH"],
                              w2d=G.w2d, include_exchange=True, include_hartree=False,
                              exchange_hermitian_channel_packing=a["exchange_hermitian_channel_packing"])
    assert _rel(_np(F1).astype(np.complex128), F0) < 2e-6


@pytest.mark.parametrize("n1,n2", [(13, 11), (130, 128), (3000, 4)])
def test_any_mesh_selfenergy_ifftshift_and_adjoint(dev, n1, n2):
    """The ifftshift-on-store form and the conjugate (cross-correlation) form of
    the staged and generic-radix exchange against the oracle's FFTs."""
    import torch
    from jax_hf_amd import _lib as L
    from jax_hf_amd.fock import FockEngine
    nb = 2
    print(f"ifftshift / adjoint mesh {n1}x{n2}", flush=True)
    rng = np.random.default_rng(n1 * n2)
    P = rng.normal(size=(n1, n2, nb, nb)) + 1j * rng.normal(size=(n1, n2, nb, nb))
    VR = rng.normal(size=(n1, n2, 1, 1)) + 1j * rng.normal(size=(n1, n2, 1, 1))
    d = torch.device("cuda:0")
    Pt = torch.from_numpy(P).to(d)
    eng = FockEngine(h=torch.zeros_like(Pt), VR=torch.from_numpy(VR).to(d), refP=None,
                     HH=None, w2d=np.ones((n1, n2)), include_exchange=True,
                     include_hartree=False, herm=False)
    out = torch.empty_like(Pt)
    eng.selfenergy(Pt, out, rho=False, ifftshift=True)
    assert _rel(_np(out), O.selfenergy_fft(VR, P, apply_ifftshift=True)) < 1e-12
    # adjoint: -IFFT2[FFT2(G) conj(VR)] per channel
    ref = -np.fft.ifft2(np.fft.fft2(P, axes=(0, 1)) * np.conj(VR), axes=(0, 1))
    adj = torch.empty_like(Pt)
    st = L.stream(d)
    L.call("hfx_selfenergy_adjoint", eng.plan.handle, L.ptr(Pt), L.ptr(eng.VR), 0, L.ptr(adj), st)
    assert _rel(_np(adj), ref) < 1e-12


@pytest.mark.parametrize("n1,n2,nb,opt", [(11, 11, 2, "cg"), (13, 13, 3, "scf"), (128, 128, 2, "scf"),
                                          (11, 11, 2, "newton")])
def test_any_mesh_solvers_match_oracle(dev, n1, n2, nb, opt):
    """CG, SCF (DIIS) and Newton on a prime mesh and a staged mesh follow the
    oracle: the CG trajectory's first energies within 1e-10, the SCF fixed
    point and the Newton minimum within 1e-10 per k-point (weights sum to 1)."""
    import jax_hf_amd as J
    print(f"solver {opt} mesh {n1}x{n2} nb={nb}", flush=True)
    h = syn.tight_binding(n1, n2, nb, seed=5)
    V = syn.coulomb_scalar(n1, n2)
    w = np.full((n1, n2), 1.0 / (n1 * n2))
    n_el = 0.45 * nb
    T = 0.1
    K0 = O.OracleKernel(w, h, V, T)
    G = J.HartreeFockKernel(w, h, V, T)
    P0 = np.zeros_like(h)
    if opt == "scf":
        kw = dict(max_iter=300, acceleration="diis", density_tol=1e-10, comm_tol=1e-10)
        r0 = O.solve_scf(K0, P0, n_el, config=O.SCFConfig(**kw))
        r = J.solve_scf(G, P0, n_el, config=J.SCFConfig(**kw))
        assert r.converged and r0.converged
        assert abs(float(r.energy) - float(r0.energy)) < 1e-10
        assert np.abs(_np(r.density_matrix) - np.asarray(r0.density_matrix)).max() < 1e-8
        return
    if opt == "cg":
        kw = dict(max_iter=60, tol_E=1e-12)
        r0 = O.solve(K0, P0, n_el, config=O.SolverConfig(**kw))
        r = J.solve(G, P0, n_el, config=J.SolverConfig(**kw))
        hE = _np(r.history["E"])
        assert np.abs(hE[:5] - r0.history["E"][:5]).max() < 1e-10
        return
    kw = dict(max_iter=60, tol_grad=1e-7, optimizer="newton")
    r0 = O.solve(K0, P0, n_el, config=O.SolverConfig(**kw))
    r = J.solve(G, P0, n_el, config=J.SolverConfig(**kw))
    assert r.converged == r0.converged
    assert abs(float(r.energy) - r0.energy) < 1e-10
