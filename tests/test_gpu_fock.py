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


@pytest.mark.parametrize("planar", [True, False])
@pytest.mark.parametrize("nk,nb,layer", [(1, 2, False), (2, 2, False), (5, 3, False), (12, 4, False),
                                         (24, 4, False), (24, 4, True), (36, 8, True), (48, 16, True),
                                         (9, 2, True), (7, 3, False)])
def test_vr_and_fock_fp64_match_oracle(dev, nk, nb, layer, planar, monkeypatch):
    import jax_hf_amd as J
    from jax_hf_amd import fock as fock_mod
    monkeypatch.setattr(fock_mod, "PLANAR", planar)
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


@pytest.mark.parametrize("planar", [True, False])
def test_fixture_fock_hartree_refP(dev, meshgrid_case, planar, monkeypatch):
    from jax_hf_amd import fock as fock_mod
    monkeypatch.setattr(fock_mod, "PLANAR", planar)
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


@pytest.mark.parametrize("planar", [True, False])
def test_contact_terms_and_no_exchange(dev, planar, monkeypatch):
    from jax_hf_amd import fock as fock_mod
    monkeypatch.setattr(fock_mod, "PLANAR", planar)
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


@pytest.mark.parametrize("planar", [True, False])
def test_block_specs_gpu(dev, planar, monkeypatch):
    from jax_hf_amd import fock as fock_mod
    monkeypatch.setattr(fock_mod, "PLANAR", planar)
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
    workgroup sizes, the fused-column two-stage form, which falls back to the
    two-stage kernel in this layout) gives the oracle's Sigma and channel moments, with refP,
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
        for v in range(1, 8):
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
        assert 1 <= best <= 7 and plan.exchange_variant == best
        assert all(t > 0 for t in times.values()), times
    finally:
        plan.exchange_variant = 0


@pytest.mark.parametrize("cdt", ["c128", "c64"])
@pytest.mark.parametrize("nk", [7, 12, 20, 36, 48, 96])
def test_planar_exchange_matches_oracle(dev, nk, cdt):
    """Slot-plane exchange: pack is an exact gather; Sigma planes and moments
    match the oracle for every kernel variant, full / Hermitian / explicit
    channel lists, scalar and full VR, with refP and ifftshift; the planar
    epilogue equals the (k, i, j)-layout epilogue."""
    import torch
    import jax_hf_amd as J
    from jax_hf_amd import _lib as L
    nb = 3
    rng = np.random.default_rng(2000 + nk)
    h = syn.tight_binding(nk, nk, nb, seed=nk)
    w = rng.uniform(0.5, 1.5, size=(nk, nk)) / nk ** 2
    P = syn.haar_density(nk, nk, nb, rng)
    R = syn.haar_density(nk, nk, nb, rng)
    npd = np.complex128 if cdt == "c128" else np.complex64
    tol = 1e-12 if cdt == "c128" else 2e-6
    st = None
    for layer in (True, False):
        V = syn.coulomb_layer(nk, nk, nb, 2) if layer else syn.coulomb_scalar(nk, nk)
        K = O.OracleKernel(w, h, V, 0.1)
        G = J.HartreeFockKernel(w, h.astype(npd), V, 0.1)
        eng = G.fock_engine(herm=False)
        plan = eng.plan
        st = L.stream(G.device)
        dev_ = G.device
        Pt = torch.from_numpy(P.astype(npd)).to(dev_)
        Rt = torch.from_numpy(R.astype(npd)).to(dev_)
        VRc = eng.vr_planar()
        if eng.vfull:
            assert torch.equal(VRc, eng.VR.permute(2, 3, 0, 1).reshape(nb * nb, nk, nk))
        rho_ref = np.einsum("ab,abij->ij", w, P - R)
        for shift in (0, 1):
            S_ref = O.selfenergy_fft(K._VR_shifted, P - R, apply_ifftshift=bool(shift))
            for herm, chans in ((0, None), (1, None), (0, [4, 0, 7]), (1, [1, 5])):
                if chans is None:
                    cl = list(range(nb * nb)) if not herm else [i * nb + j for i in range(nb) for j in range(i, nb)]
                else:
                    cl = chans
                ct = None if chans is None else torch.tensor(chans, dtype=torch.int32, device=dev_)
                nch = len(cl)
                X = torch.full((nch, nk, nk), float("nan"), dtype=Pt.dtype, device=dev_)
                L.call("hfx_planar_pack", plan.handle, 1, L.ptr(Pt), L.ptr(Rt), L.ptr(ct), nch, herm, L.ptr(X), st)
                D = (Pt - Rt).reshape(nk, nk, nb * nb)
                assert torch.equal(X, D[..., cl].permute(2, 0, 1)), (herm, chans)
                for v in range(1, 8):
                    plan.exchange_variant = v
                    S = torch.full_like(X, float("nan"))
                    eng.rho.zero_()
                    L.call("hfx_selfenergy_planar", plan.handle, 1, L.ptr(X), L.ptr(VRc), eng.vfull, None, L.ptr(ct), nch,
                           herm, shift, L.ptr(eng.w2d), L.ptr(S), L.ptr(eng.rho), st)
                    Sn = _np(S)
                    for t, c in enumerate(cl):
                        i, j = divmod(c, nb)
                        assert _rel(Sn[t], S_ref[..., i, j]) < tol, (v, herm, chans, shift, c)
                        r = _np(eng.rho)
                        assert abs(r[i, j] - rho_ref[i, j]) < 1e-12 * (1 if cdt == "c128" else 1e6) * max(
                            1.0, np.abs(rho_ref).max()), (v, herm, c)
                plan.exchange_variant = 0
                if shift:
                    continue
                # epilogue on the slot planes == epilogue on the (k, i, j) Sigma
                Sa = torch.zeros_like(Pt)
                Sn = torch.from_numpy(_np(S)).to(dev_)
                for t, c in enumerate(cl):
                    i, j = divmod(c, nb)
                    Sa[..., i, j] = Sn[t]
                    if herm and i != j:
                        Sa[..., j, i] = Sn[t].conj()
                hd = torch.tensor(rng.normal(size=nb), dtype=torch.float64, device=dev_)
                sigc = torch.from_numpy(rng.normal(size=(nb, nb)) + 1j * rng.normal(size=(nb, nb))).to(dev_)
                F0, F1 = torch.empty_like(Pt), torch.empty_like(Pt)
                O0, O1 = torch.empty_like(Pt), torch.empty_like(Pt)
                E = torch.zeros(2, dtype=torch.float64, device=dev_)
                L.call("hfx_fock_epilogue", plan.handle, L.ptr(eng.h), L.ptr(Sa), L.ptr(hd), L.ptr(sigc), L.ptr(Pt),
                       L.ptr(eng.w2d), L.ptr(F0), L.ptr(O0), L.ptr(plan.scratch), L.ptr(E), 0.25, st)
                L.call("hfx_fock_epilogue_planar", plan.handle, L.ptr(eng.h), L.ptr(S), L.ptr(ct), nch, herm,
                       L.ptr(hd), L.ptr(sigc), L.ptr(Pt), L.ptr(eng.w2d), L.ptr(F1), L.ptr(O1),
                       L.ptr(plan.scratch), L.ptr(E) + 8, 0.25, st)
                assert torch.equal(F0, F1) and torch.equal(O0, O1), (herm, chans)
                e = _np(E)
                etol = 1e-13 if cdt == "c128" else 1e-5
                assert abs(e[0] - e[1]) <= etol * max(1.0, abs(e[0])), (e, herm, chans)


@pytest.mark.parametrize("cdt", ["c128", "c64"])
@pytest.mark.parametrize("herm", [False, True])
def test_vr_plane_map_build_matches_dense(dev, cdt, herm, monkeypatch):
    """A layer-resolved V with repeated planes: the Fock build reading one copy
    per distinct plane equals the build reading every channel's own plane bit
    for bit, and the oracle to the FP64 / complex64 bar."""
    import torch
    import jax_hf_amd as J
    from jax_hf_amd import fock as fock_mod
    nk, nb = 24, 8
    rng = np.random.default_rng(77)
    npd = np.complex128 if cdt == "c128" else np.complex64
    h = syn.tight_binding(nk, nk, nb, seed=3)
    V = syn.coulomb_layer(nk, nk, nb, 4)
    w = np.full((nk, nk), 1.0 / nk ** 2)
    P = syn.haar_density(nk, nk, nb, rng)
    out = {}
    for dedup in (True, False):
        monkeypatch.setattr(fock_mod, "DEDUP_VR", dedup)
        G = J.HartreeFockKernel(w, h.astype(npd), V, 0.1)
        eng = G.fock_engine(herm=herm)
        eng.vr_planar()
        assert (eng._vmap is not None) == dedup
        if dedup:
            assert int(torch.unique(eng._vmap).numel()) == 4
        Pt = torch.from_numpy(P.astype(npd)).to(dev)
        F = torch.empty_like(Pt)
        S = torch.empty_like(Pt)
        eng.build(Pt, F, Sigma_out=S, energy=True, refP=None)
        out[dedup] = (F.clone(), S.clone(), float(eng.E))
    assert torch.equal(out[True][0], out[False][0]) and torch.equal(out[True][1], out[False][1])
    assert out[True][2] == out[False][2]
    K = O.OracleKernel(w, h, V, 0.1)
    b = K.as_args()
    S0, H0, F0 = O.build_fock(P, h=b["h"], VR=b["VR"], refP=b["refP"], HH=b["HH"], w2d=K.w2d,
                              include_exchange=True, include_hartree=False,
                              exchange_hermitian_channel_packing=False)   # P Hermitian: the same Sigma
    tol = 1e-11 if cdt == "c128" else 2e-6
    assert _rel(_np(out[True][0]).astype(np.complex128), F0) < tol
