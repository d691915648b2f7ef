"""GPU parity of the Fock build (HIP) against the FP64 oracle and the reference fixture.

Bar: FP64 (complex128) Fock matrix within 1e-11 relative of the oracle
(tolerance from the FP64 target: |F - F_oracle|_max <= 1e-11 |F_oracle|_max);
complex64 within 2e-6 relative; energies within 1e-11 (FP64) / 1e-5 (c64)
relative.
"""

import numpy as np
import pytest

import oracle as O
from jax_hf_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _oracle_fock(K, P):
    a = K.as_args()
    return O.build_fock(P, h=a["h"], VR=a["VR"], refP=a["refP"], HH=a["HH"], w2d=K.w2d,
                        include_exchange=a["include_exchange"], include_hartree=a["include_hartree"],
                        exchange_hermitian_channel_packing=a["exchange_hermitian_channel_packing"],
                        contact_g=a["contact_g"], contact_Oi=a["contact_Oi"], contact_Oj=a["contact_Oj"])


def _gpu_fock(G, P, **kw):
    import jax_hf_amd as J
    a = G.as_args()
    return J.build_fock(P, h=a["h"], VR=a["VR"], refP=a["refP"], HH=a["HH"], w2d=G.w2d,
                        include_exchange=a["include_exchange"], include_hartree=a["include_hartree"],
                        exchange_hermitian_channel_packing=a["exchange_hermitian_channel_packing"],
                        contact_g=a["contact_g"], contact_Oi=a["contact_Oi"], contact_Oj=a["contact_Oj"],
                        **kw)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_library_loads_on_gpu(dev):
    from jax_hf_amd import _lib
    _lib.require_gpu()


@pytest.mark.parametrize("nk,nb,layer", [(1, 2, False), (2, 2, False), (5, 3, False), (12, 4, False),
                                         (24, 4, False), (24, 4, True), (36, 8, True), (48, 16, True),
                                         (9, 2, True), (7, 3, False)])
def test_vr_and_fock_fp64_match_oracle(dev, nk, nb, layer):
    import jax_hf_amd as J
    rng = np.random.default_rng(nk * 100 + nb)
    h = syn.tight_binding(nk, nk, nb, seed=nk)
    V = syn.coulomb_layer(nk, nk, nb, 2) if layer else syn.coulomb_scalar(nk, nk)
    w = rng.uniform(0.5, 1.5, size=(nk, nk)) / nk ** 2
    P = syn.haar_density(nk, nk, nb, rng)
    K = O.OracleKernel(w, h, V, 0.1)
    G = J.HartreeFockKernel(w, h, V, 0.1)
    assert G.exchange_hermitian_channel_packing == K.exchange_hermitian_channel_packing
    assert _rel(_np(G._VR_shifted), K._VR_shifted) < 1e-13
    assert _rel(_np(G.VR), K.VR) < 1e-13
    S0, H0, F0 = _oracle_fock(K, P)
    S1, H1, F1 = _gpu_fock(G, P)
    assert _rel(_np(F1), F0) < 1e-11
    assert _rel(_np(S1), S0) < 1e-11
    E0 = O.hf_energy(P, h=K.h, Sigma=S0, H=H0, weights_b=K.weights_b)
    E1 = float(J.hf_energy(P, h=G.h, Sigma=S1, H=H1, weights_b=G.weights_b))
    assert abs(E1 - E0) <= 1e-11 * max(1.0, abs(E0))


def test_fixture_fock_hartree_refP(dev, meshgrid_case):
    """Hartree + refP + scalar exchange on the reference's converged state."""
    import jax_hf_amd as J
    d = meshgrid_case
    args = dict(include_hartree=True, reference_density=d["reference_density"].astype(np.complex128),
                hartree_matrix=d["hartree_matrix"])
    h = d["hamiltonian"].astype(np.complex128)
    K = O.OracleKernel(d["weights"], h, d["coulomb_q"], float(d["T"]), **args)
    G = J.HartreeFockKernel(d["weights"], h, d["coulomb_q"], float(d["T"]), **args)
    P = d["P_expected"].astype(np.complex128)
    S0, H0, F0 = _oracle_fock(K, P)
    S1, H1, F1 = _gpu_fock(G, P)
    assert _rel(_np(F1), F0) < 1e-11
    np.testing.assert_allclose(_np(H1)[0, 0], H0[0, 0].real, atol=1e-13)
    assert _rel(_np(F1), d["F_expected"]) < 2e-7
    E1 = float(J.hf_energy(P, h=G.h, Sigma=S1, H=H1, weights_b=G.weights_b))
    assert abs(E1 - float(d["E_expected"])) < 1e-6


def test_contact_terms_and_no_exchange(dev):
    import jax_hf_amd as J
    rng = np.random.default_rng(5)
    nk, nb = 4, 2
    w = rng.uniform(0.5, 1.0, size=(nk, nk))
    P = syn.haar_density(nk, nk, nb, rng)
    ref = syn.haar_density(nk, nk, nb, rng)
    h = syn.tight_binding(nk, nk, nb, 3)
    sz = np.diag([1.0, -1.0]).astype(complex)
    sx = np.array([[0, 1], [1, 0]], complex)
    for ex in (True, False):
        kw = dict(include_hartree=True, include_exchange=ex, reference_density=ref,
                  hartree_matrix=np.array([[0.4, 0.1], [0.1, 0.3]]),
                  contact_terms=[(0.3, sz, sx), (-0.2, sx, sz)])
        K = O.OracleKernel(w, h, syn.coulomb_scalar(nk, nk), 0.1, **kw)
        G = J.HartreeFockKernel(w, h, syn.coulomb_scalar(nk, nk), 0.1, **kw)
        S0, _, F0 = _oracle_fock(K, P)
        S1, _, F1 = _gpu_fock(G, P)
        assert _rel(_np(F1), F0) < 1e-12
        assert _rel(_np(S1), S0) < 1e-12


def test_block_specs_gpu(dev):
    import jax_hf_amd as J
    rng = np.random.default_rng(0)
    nk, n0 = 4, 2
    nb = 2 * n0
    P = np.zeros((nk, nk, nb, nb), np.complex128)
    for b in range(2):
        s = slice(b * n0, (b + 1) * n0)
        blk = rng.normal(size=(nk, nk, n0, n0)) + 1j * rng.normal(size=(nk, nk, n0, n0))
        P[..., s, s] = 0.5 * (blk + np.conj(np.swapaxes(blk, -1, -2)))
    VR = rng.normal(size=(nk, nk, nb, nb)) + 1j * rng.normal(size=(nk, nk, nb, nb))
    w = np.ones((nk, nk))
    h = np.zeros_like(P)
    kw = dict(h=h, VR=VR, HH=np.zeros((nb, nb)), w2d=w, include_exchange=True,
              include_hartree=False, exchange_hermitian_channel_packing=False)
    for Pc, spec in ((P, (("sizes", (2, 2)),)), (P + 0.1, (("sizes", (2, 2)),))):
        S0, _, _ = O.build_fock(Pc, refP=np.zeros_like(Pc), exchange_block_specs=spec,
                                exchange_offdiag_atol=1e-6, **kw)
        S1, _, _ = J.build_fock(Pc, refP=None, exchange_block_specs=spec, exchange_offdiag_atol=1e-6, **kw)
        assert _rel(_np(S1), S0) < 1e-12


@pytest.mark.parametrize("nk,nb", [(24, 4), (48, 16)])
def test_fock_complex64_error_stated(dev, nk, nb):
    """Mixed form: complex64 storage / FP32 FFT vs FP64 oracle (bound 2e-6 relative)."""
    import jax_hf_amd as J
    rng = np.random.default_rng(7)
    h = syn.tight_binding(nk, nk, nb, seed=1)
    V = syn.coulomb_layer(nk, nk, nb, 2)
    w = syn.uniform_weights(nk, nk)
    P = syn.haar_density(nk, nk, nb, rng)
    K = O.OracleKernel(w, h, V, 0.1)
    G = J.HartreeFockKernel(w, h.astype(np.complex64), V, 0.1)
    S0, H0, F0 = _oracle_fock(K, P)
    S1, H1, F1 = _gpu_fock(G, P.astype(np.complex64))
    assert _rel(_np(F1).astype(np.complex128), F0) < 2e-6
    E0 = O.hf_energy(P, h=K.h, Sigma=S0, H=H0, weights_b=K.weights_b)
    E1 = float(J.hf_energy(P.astype(np.complex64), h=G.h, Sigma=S1, H=H1, weights_b=G.weights_b))
    assert abs(E1 - E0) < 1e-5 * max(1.0, abs(E0))


def test_selfenergy_ifftshift_and_single_point(dev):
    import torch
    from jax_hf_amd.fock import FockEngine
    rng = np.random.default_rng(3)
    for nk in (1, 6, 5):
        nb = 2
        P = rng.normal(size=(nk, nk, nb, nb)) + 1j * rng.normal(size=(nk, nk, nb, nb))
        VR = rng.normal(size=(nk, nk, 1, 1)) + 1j * rng.normal(size=(nk, nk, 1, 1))
        ref = O.selfenergy_fft(VR, P, apply_ifftshift=True)
        d = torch.device("cuda:0")
        Pt = torch.from_numpy(P).to(d)
        eng = FockEngine(h=torch.zeros_like(Pt), VR=torch.from_numpy(VR).to(d), refP=None,
                         HH=None, w2d=np.ones((nk, nk)), include_exchange=True,
                         include_hartree=False, herm=False)
        out = torch.empty_like(Pt)
        eng.selfenergy(Pt, out, rho=False, ifftshift=True)
        assert _rel(_np(out), ref) < 1e-13


def test_entropy_and_free_energy(dev):
    import jax_hf_amd as J
    rng = np.random.default_rng(1)
    p = rng.uniform(0.0, 1.0, size=(6, 6, 4))
    p[0, 0, 0] = 0.0
    p[0, 0, 1] = 1.0
    w = rng.uniform(0.1, 1.0, size=(6, 6))
    S0 = O.occupation_entropy(p, w)
    assert abs(float(J.occupation_entropy(p, w)) - S0) < 1e-13 * max(1, abs(S0))
    assert abs(float(J.free_energy(-1.25, p, w, 0.3)) - O.free_energy(-1.25, p, w, 0.3)) < 1e-13


@pytest.mark.parametrize("cdt", ["c128", "c64"])
@pytest.mark.parametrize("nk", [12, 16, 24, 32, 36, 48, 60, 64, 72, 90, 96])
def test_exchange_variants_match_oracle(dev, nk, cdt):
    """Every exchange-kernel variant (one-stage-per-radix, two-stage at three
    workgroup sizes) gives the oracle's Sigma and channel moments, with refP,
    full and Hermitian-packed channel lists, and the conjugated-VR adjoint."""
    import torch
    import jax_hf_amd as J
    from jax_hf_amd import _lib as L
    nb = 2
    rng = np.random.default_rng(1000 + nk)
    h = syn.tight_binding(nk, nk, nb, seed=nk)
    V = syn.coulomb_layer(nk, nk, nb, 2)
    w = rng.uniform(0.5, 1.5, size=(nk, nk)) / nk ** 2
    P = syn.haar_density(nk, nk, nb, rng)
    R = syn.haar_density(nk, nk, nb, rng)
    K = O.OracleKernel(w, h, V, 0.1)
    S_ref = O.selfenergy_fft(K._VR_shifted, P - R, apply_ifftshift=False)
    A_ref = O.selfenergy_fft(np.conj(K._VR_shifted), P, apply_ifftshift=False)
    rho_ref = np.einsum("ab,abij->ij", w, P - R)
    npd = np.complex128 if cdt == "c128" else np.complex64
    tol = 1e-12 if cdt == "c128" else 2e-6
    G = J.HartreeFockKernel(w, h.astype(npd), V, 0.1)
    eng = G.fock_engine(herm=False)
    Pt = torch.from_numpy(P.astype(npd)).to(G.device)
    Rt = torch.from_numpy(R.astype(npd)).to(G.device)
    plan = eng.plan
    try:
        for v in (1, 2, 3, 4):
            plan.exchange_variant = v
            for herm in (False, True):
                S = torch.full_like(Pt, float("nan"))
                eng.rho.zero_()
                eng.selfenergy(Pt, S, refP=Rt, herm=herm, rho=True)
                assert _rel(_np(S), S_ref) < tol, (v, herm)
                assert np.abs(_np(eng.rho) - rho_ref).max() < 1e-12 * max(1.0, np.abs(rho_ref).max()) * (
                    1 if cdt == "c128" else 1e6), (v, herm)
            A = torch.full_like(Pt, float("nan"))
            L.call("hfx_selfenergy_adjoint", plan.handle, L.ptr(Pt), L.ptr(eng.VR), 1, L.ptr(A),
                   L.stream(G.device))
            assert _rel(_np(A), A_ref) < tol, v
        S = torch.empty_like(Pt)
        best, times = plan.autotune(Pt, eng.VR, S, reps=2)
        assert best in (1, 2, 3, 4) and plan.exchange_variant == best
        assert all(t > 0 for t in times.values()), times
    finally:
        plan.exchange_variant = 0
