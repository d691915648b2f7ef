"""Pins the CPU oracle against the reference's fixture, closed forms and invariants.

None of these checks retypes an oracle formula: each compares the oracle with
data written by the reference (``meshgrid_regression_case_v1.npz``), with a
differently-structured statement of the same mathematics (direct O(Nk^2) sum,
LU Cayley, dense linear algebra, finite differences), or with an invariant.
"""

import numpy as np
import pytest
from scipy.special import expit

import oracle as O
from oracle import hf_oracle as OH
from jax_hf_amd import synthetic as syn


def _fixture_kernel(d):
    return O.OracleKernel(d["weights"], d["hamiltonian"], d["coulomb_q"], float(d["T"]),
                          include_hartree=True, reference_density=d["reference_density"],
                          hartree_matrix=d["hartree_matrix"])


def _fock(K, P, **kw):
    a = K.as_args()
    return O.build_fock(P, h=a["h"], VR=a["VR"], refP=a["refP"], HH=a["HH"], w2d=K.w2d,
                        include_exchange=a["include_exchange"],
                        include_hartree=a["include_hartree"],
                        exchange_hermitian_channel_packing=a["exchange_hermitian_channel_packing"],
                        contact_g=a["contact_g"], contact_Oi=a["contact_Oi"],
                        contact_Oj=a["contact_Oj"], **kw)


# ---------------------------------------------------------------- fixture pins

def test_fixture_fock_matrix(meshgrid_case):
    d = meshgrid_case
    K = _fixture_kernel(d)
    assert K.exchange_hermitian_channel_packing
    _, _, F = _fock(K, d["P_expected"])
    Fe = d["F_expected"].astype(np.complex128)
    assert np.abs(F - Fe).max() / np.abs(Fe).max() < 2e-7


def test_fixture_energy(meshgrid_case):
    d = meshgrid_case
    K = _fixture_kernel(d)
    S, H, _ = _fock(K, d["P_expected"])
    E = O.hf_energy(d["P_expected"], h=K.h, Sigma=S, H=H, weights_b=K.weights_b)
    assert abs(E - float(d["E_expected"])) < 1e-6 * abs(float(d["E_expected"]))


def test_fixture_mu_and_density(meshgrid_case):
    d = meshgrid_case
    w = d["weights"].astype(np.float64)
    P, mu = O.density_matrix_from_fock(d["F_expected"], w, float(d["electrondensity0"]),
                                       float(d["T"]))
    assert abs(mu - float(d["mu_expected"])) < 2e-6
    assert np.abs(P - d["P_expected"]).max() < 5e-6


@pytest.mark.parametrize("which", ["cg", "scf"])
def test_fixture_fixed_point(meshgrid_case, which):
    """The converged state the reference wrote is reached by the oracle solvers."""
    d = meshgrid_case
    K = _fixture_kernel(d)
    n = float(d["electrondensity0"])
    if which == "cg":
        r = O.solve(K, np.zeros_like(K.h), n, config=O.SolverConfig(max_iter=300, tol_E=1e-12))
        P, E, conv = r.density, r.energy, r.converged
    else:
        r = O.solve_scf(K, np.zeros_like(K.h), n,
                        config=O.SCFConfig(max_iter=300, acceleration="diis",
                                           density_tol=1e-10, comm_tol=1e-9))
        P, E, conv = r.density_matrix, r.energy, r.converged
    assert conv
    assert abs(E - float(d["E_expected"])) < 1e-6
    assert np.abs(P - d["P_expected"]).max() < 1e-5


# ------------------------------------------------------------ exchange pins

@pytest.mark.parametrize("shape", [(6, 6), (5, 7), (4, 9)])
def test_exchange_matches_direct_sum_nonuniform_weights(shape):
    rng = np.random.default_rng(11)
    n0, n1 = shape
    nb = 3
    w = rng.uniform(0.5, 1.5, size=(n0, n1))
    V = rng.uniform(0.1, 2.0, size=(n0, n1, nb, nb))
    P = rng.standard_normal((n0, n1, nb, nb)) + 1j * rng.standard_normal((n0, n1, nb, nb))
    h = np.zeros((n0, n1, nb, nb), np.complex128)
    K = O.OracleKernel(w, h, V, 0.1)
    S = O.selfenergy_fft(K._VR_shifted, P, apply_ifftshift=False)
    Sd = O.direct_exchange_sum(w, V, P)
    assert np.abs(S - Sd).max() < 1e-12 * np.abs(Sd).max()
    # The unshifted VR path with the final ifftshift is the same object.
    S2 = O.selfenergy_fft(K.VR, P, apply_ifftshift=True)
    assert np.abs(S2 - Sd).max() < 1e-12 * np.abs(Sd).max()


def test_exchange_single_point_closed_form():
    VR = np.array([[[[0.3]]]], np.complex128)
    P = np.array([[[[1.0 + 0.5j, 0.2], [0.2, 0.8 - 0.1j]]]])
    S = O.selfenergy_fft(VR, P)
    np.testing.assert_allclose(S, -0.3 * P, atol=1e-15)


def test_hermitian_packing_equals_full():
    rng = np.random.default_rng(3)
    nk, nb = 6, 4
    P = syn.haar_density(nk, nk, nb, rng)
    V = syn.coulomb_scalar(nk, nk)
    K = O.OracleKernel(syn.uniform_weights(nk, nk), np.zeros_like(P), V, 0.1)
    full = O.selfenergy_fft(K._VR_shifted, P, apply_ifftshift=False)
    pk = O.selfenergy_fft(K._VR_shifted, P, apply_ifftshift=False, hermitian_channel_packing=True)
    assert np.abs(full - pk).max() < 1e-13


def test_layer_resolved_real_symmetric_v_gives_hermitian_sigma():
    """Identity behind packing for real symmetric V_ll'(q): Sigma_ji = conj(Sigma_ij)."""
    rng = np.random.default_rng(4)
    nk, nb = 6, 4
    P = syn.haar_density(nk, nk, nb, rng)
    V = syn.coulomb_layer(nk, nk, nb, 2)
    K = O.OracleKernel(syn.uniform_weights(nk, nk), np.zeros_like(P), V, 0.1)
    S = O.selfenergy_fft(K._VR_shifted, P, apply_ifftshift=False)
    assert np.abs(S - np.conj(np.swapaxes(S, -1, -2))).max() < 1e-13


def test_block_specs_match_full_when_block_diagonal_and_fall_back():
    rng = np.random.default_rng(0)
    nk, n0 = 4, 2
    nb = 2 * n0
    P = np.zeros((nk, nk, nb, nb), np.complex128)
    for b in range(2):
        s = slice(b * n0, (b + 1) * n0)
        blk = rng.normal(size=(nk, nk, n0, n0)) + 1j * rng.normal(size=(nk, nk, n0, n0))
        P[..., s, s] = 0.5 * (blk + np.conj(np.swapaxes(blk, -1, -2)))
    VR = rng.normal(size=(nk, nk, nb, nb)) + 1j * rng.normal(size=(nk, nk, nb, nb))
    full = O.selfenergy_fft(VR, P)
    blk = O.selfenergy_fft(VR, P, block_specs=(("sizes", (2, 2)),), check_offdiag=True)
    off = np.ones((nb, nb), bool)
    off[:2, :2] = off[2:, 2:] = False
    np.testing.assert_allclose(blk[..., ~off], full[..., ~off], atol=1e-12)
    assert np.all(blk[..., off] == 0)
    P2 = P.copy()
    P2[..., 0, 2] = 0.1
    coupled = O.selfenergy_fft(VR, P2, block_specs=(("sizes", (2, 2)),), check_offdiag=True,
                               offdiag_atol=1e-6)
    np.testing.assert_allclose(coupled, O.selfenergy_fft(VR, P2), atol=1e-14)


def test_hartree_and_contact_terms_closed_form():
    """Hartree = diag(HH n); contact = g O_i tr(O_j rho) - g O_i rho O_j (k-independent)."""
    rng = np.random.default_rng(5)
    nk, nb = 3, 2
    w = rng.uniform(0.5, 1.0, size=(nk, nk))
    P = syn.haar_density(nk, nk, nb, rng)
    ref = syn.haar_density(nk, nk, nb, rng)
    HH = np.array([[0.4, 0.1], [0.1, 0.3]])
    sz = np.diag([1.0, -1.0]).astype(complex)
    sx = np.array([[0, 1], [1, 0]], complex)
    K = O.OracleKernel(w, np.zeros_like(P), np.zeros((nk, nk, 1, 1)), 0.1, include_hartree=True,
                       include_exchange=False, reference_density=ref, hartree_matrix=HH,
                       contact_terms=[(0.3, sz, sx), (-0.2, sx, sz)])
    S, H, F = _fock(K, P)
    dP = P - ref
    n = np.array([sum(w[i, j] * dP[i, j, a, a].real for i in range(nk) for j in range(nk))
                  for a in range(nb)])
    np.testing.assert_allclose(np.diagonal(H[0, 0]).real, HH @ n, atol=1e-14)
    rho = sum(w[i, j] * dP[i, j] for i in range(nk) for j in range(nk))
    sig = (0.3 * sz * np.trace(sx @ rho) - 0.3 * sz @ rho @ sx
           - 0.2 * sx * np.trace(sz @ rho) + 0.2 * sx @ rho @ sz)
    np.testing.assert_allclose(S[1, 2], sig, atol=1e-14)
    A = S + H
    np.testing.assert_allclose(F, 0.5 * (A + np.conj(np.swapaxes(A, -1, -2))), atol=1e-15)


# ------------------------------------------------------- occupations / mu

def test_entropy_closed_form_and_free_energy_derivative():
    w = np.array([[0.25, 0.75]])
    p = np.full((1, 2, 3), 0.5)
    assert abs(O.occupation_entropy(p, w) - 3 * np.log(2.0)) < 1e-14
    rng = np.random.default_rng(1)
    p = rng.uniform(0.1, 0.9, size=(1, 2, 3))
    eps = rng.standard_normal((1, 2, 3))
    T = 0.3

    def om(pp):
        return O.free_energy(float(np.sum(w[..., None] * pp * eps)), pp, w, T)

    h = 1e-6
    for idx in [(0, 0, 0), (0, 1, 2)]:
        e = np.zeros_like(p)
        e[idx] = h
        fd = (om(p + e) - om(p - e)) / (2 * h)
        exact = w[idx[:2]] * (eps[idx] + T * np.log(p[idx] / (1 - p[idx])))
        assert abs(fd - exact) < 1e-8


@pytest.mark.parametrize("method", ["bisection", "newton"])
def test_mu_hits_target_multiband(method):
    rng = np.random.RandomState(42)
    bands = rng.randn(4, 4, 6)
    w = np.ones((4, 4)) / 16
    mu = O.find_chemical_potential(bands, w, 3.0, 0.05, method=method)
    assert abs(np.sum(w[..., None] * O.fermidirac(bands - mu, 0.05)) - 3.0) < 1e-10


def test_mu_cold_limit_and_validation():
    bands = np.array([[[0.0, 1.0, 2.0]]])
    mu = O.find_chemical_potential(bands, np.ones((1, 1)), 2.0, 1e-5)
    assert abs(np.sum(O.fermidirac(bands - mu, 1e-5)) - 2.0) < 1e-4
    with pytest.raises(ValueError, match="physically reachable range"):
        O.find_chemical_potential(np.zeros((1, 1, 1)), np.ones((1, 1)), 2.0, 0.1)


def test_solve_mu_newton_bracket_count_equation():
    rng = np.random.default_rng(2)
    eps = rng.standard_normal((5, 5, 4))
    w = np.full((5, 5), 1 / 25)
    for T in (0.5, 0.05, 0.005):
        mu = O.solve_mu(eps, w, 1.7, 0.0, T, 25)
        assert abs(np.sum(w[..., None] * expit((mu - eps) / T)) - 1.7) < 1e-10


# ------------------------------------------------------ retractions

def _skew_herm(rng, shape):
    d = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    return 0.5 * (d - np.conj(np.swapaxes(d, -1, -2))), None


def test_spectral_cayley_matches_lu_and_is_unitary():
    rng = np.random.default_rng(0)
    d, _ = _skew_herm(rng, (3, 3, 6, 6))
    F = rng.normal(size=(3, 3, 6, 6)) + 1j * rng.normal(size=(3, 3, 6, 6))
    F = 0.5 * (F + np.conj(np.swapaxes(F, -1, -2)))
    V, lam = O.cayley_spectral_setup(d)
    Fe = np.conj(np.swapaxes(V, -1, -2)) @ F @ V
    for tau in [0.0, 0.1, 0.5, 1.0, -0.3]:
        Ulu = O.cayley_retract(d, np.asarray(tau))
        Usp = O.cayley_unitary_from_spectrum(V, lam, tau)
        np.testing.assert_allclose(Usp, Ulu, atol=1e-12)
        np.testing.assert_allclose(Usp @ np.conj(np.swapaxes(Usp, -1, -2)),
                                   np.broadcast_to(np.eye(6), Usp.shape), atol=1e-12)
        dg = np.real(np.diagonal(np.conj(np.swapaxes(Ulu, -1, -2)) @ F @ Ulu, axis1=-2, axis2=-1))
        np.testing.assert_allclose(O.diag_UFU_from_spectrum(V, Fe, lam, tau), dg, atol=1e-12)


def test_qr_retract_is_unitary_and_first_order():
    rng = np.random.default_rng(1)
    d, _ = _skew_herm(rng, (2, 2, 5, 5))
    tau = 1e-4
    U = O.qr_retract(d, np.full((2, 2), tau))
    np.testing.assert_allclose(U @ np.conj(np.swapaxes(U, -1, -2)),
                               np.broadcast_to(np.eye(5), U.shape), atol=1e-12)
    np.testing.assert_allclose(U, np.eye(5) - tau * d, atol=5e-8)


# ------------------------------------------------------------- solvers

def _contact_kernel(nk=2, g=0.3, T=0.05):
    h = np.zeros((nk, nk, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -1.0, 1.0
    for i in range(nk):
        for j in range(nk):
            h[i, j, 0, 1] = h[i, j, 1, 0] = 0.4 + 0.1 * (i + j)
    sz = np.diag([1.0, -1.0]).astype(complex)
    return O.OracleKernel(np.ones((nk, nk)), h, np.full((nk, nk, 1, 1), 0.1), T,
                          contact_terms=[(g, sz, sz)])


def _comm_rms(F, P, w):
    c = F @ P - P @ F
    return float(np.sqrt(np.sum(w * np.sum(np.abs(c) ** 2, axis=(-2, -1))) / np.sum(w)))


def test_cg_newton_scf_agree_on_contact_kernel():
    K = _contact_kernel()
    P0 = np.zeros_like(K.h)
    cg = O.solve(K, P0, 4.0, config=O.SolverConfig(max_iter=400, tol_E=1e-11, tol_grad=1e-9))
    nt = O.solve(K, P0, 4.0, config=O.SolverConfig(max_iter=200, optimizer="newton",
                                                   tol_grad=1e-7))
    scf = O.solve_scf(K, P0, 4.0, config=O.SCFConfig(max_iter=600, mixing=0.3,
                                                     density_tol=1e-10, comm_tol=1e-9))
    assert nt.converged and scf.converged
    assert abs(cg.energy - scf.energy) < 1e-8
    assert abs(nt.energy - scf.energy) < 1e-8
    assert np.linalg.norm(cg.density - scf.density_matrix) < 1e-5
    assert _comm_rms(cg.fock, cg.density, K.w2d) < 1e-4
    assert nt.n_iter < cg.n_iter


def test_cg_gradient_stop_first_crossing():
    h = np.zeros((2, 2, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -0.5, 0.5
    h[..., 0, 1] = h[..., 1, 0] = 0.2
    K = O.OracleKernel(np.ones((2, 2)), h, np.full((2, 2, 1, 1), 0.5), 0.1)
    P0 = np.zeros_like(h)
    P0[..., 0, 0], P0[..., 1, 1] = 0.6, 0.15
    P0[..., 0, 1] = P0[..., 1, 0] = 0.25
    tol = 1e-4
    r = O.solve(K, P0, 3.0, config=O.SolverConfig(max_iter=300, tol_grad=tol))
    assert r.converged and 1 < r.n_iter < 300
    hG = r.history["grad_norm"][:r.n_iter]
    assert hG[-1] <= tol and np.all(hG[:-1] > tol)
    r3 = O.solve(K, P0, 3.0, config=O.SolverConfig(max_iter=3, tol_grad=1e-14))
    assert not r3.converged and r3.n_iter == 3


def test_newton_hvp_matches_finite_difference_of_gradient():
    """Directional derivative of the orbital gradient (Q <- Q Cayley(t X)) vs the HVP."""
    nk, nb = 2, 3
    rng = np.random.default_rng(7)
    h = np.zeros((nk, nk, nb, nb), np.complex128)
    A = (rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))) * 0.35
    A = 0.5 * (A + A.conj().T)
    for i in range(nk):
        for j in range(nk):
            h[i, j] = np.diag(np.linspace(-1, 1, nb)) + A * (np.cos(i) + np.cos(j))
    K = O.OracleKernel(np.ones((nk, nk)), h, np.full((nk, nk, 1, 1), 0.4), 0.2)
    a = K.as_args()
    _, _, F0 = _fock(K, np.zeros_like(h))
    _, Q = np.linalg.eigh(F0)
    p = rng.uniform(0.2, 0.8, size=(nk, nk, nb))

    def grad(Qx):
        P = OH._herm(OH._density_from_Qp(Qx, p))
        _, _, F = _fock(K, P)
        Ft = np.conj(np.swapaxes(Qx, -1, -2)) @ F @ Qx
        dpp = p[..., None, :] - p[..., :, None]
        return OH._skew(dpp * Ft) * (1 - np.eye(nb))

    X = rng.standard_normal((nk, nk, nb, nb)) + 1j * rng.standard_normal((nk, nk, nb, nb))
    X = OH._skew(X) * (1 - np.eye(nb))
    # Newton HVP orbital block (solver.py:916-930) restated with dp = 0.
    P = OH._herm(OH._density_from_Qp(Q, p))
    _, _, F = _fock(K, P)
    Ft = np.conj(np.swapaxes(Q, -1, -2)) @ F @ Q
    dpp = p[..., None, :] - p[..., :, None]
    M = X * dpp
    dP = OH._herm(Q @ M @ np.conj(np.swapaxes(Q, -1, -2)))
    _, _, Fr = _fock(K, dP, )
    resp = Fr - a["h"]
    St = np.conj(np.swapaxes(Q, -1, -2)) @ resp @ Q
    Am = Ft @ X - X @ Ft + St
    HX = OH._skew(dpp * Am) * (1 - np.eye(nb))
    t = 1e-5
    Up = O.cayley_retract(X, np.asarray(-t))
    Um = O.cayley_retract(X, np.asarray(t))
    fd = (grad(Q @ Up) - grad(Q @ Um)) / (2 * t)
    assert np.abs(fd - HX).max() < 1e-6 * max(1.0, np.abs(HX).max())


def test_diis_coefficients_minimise_residual():
    """Pulay coefficients = argmin c^T B c s.t. sum c = 1 (closed form B^-1 1 / 1^T B^-1 1)."""
    rng = np.random.default_rng(9)
    R = rng.standard_normal((4, 2, 2, 3, 3)) + 1j * rng.standard_normal((4, 2, 2, 3, 3))
    R[3] = 0.0      # inactive ring-buffer slots hold zeros (reference_scf.py:280-302)
    w = rng.uniform(0.5, 1.0, (2, 2))
    c, ok = O.diis_coefficients(R, np.array([True, True, True, False]), w)
    assert ok and c[3] == 0.0
    B = np.real(np.einsum("iklab,jklab,kl->ij", np.conj(R[:3]), R[:3], w))
    x = np.linalg.inv(B) @ np.ones(3)
    np.testing.assert_allclose(c[:3], x / x.sum(), atol=1e-12)


def test_synthetic_generators_shapes_and_hermiticity():
    c0 = syn.config0(nk=6, nb=4)
    h = c0["hamiltonian"]
    assert h.shape == (6, 6, 4, 4)
    np.testing.assert_allclose(h, np.conj(np.swapaxes(h, -1, -2)), atol=1e-14)
    V = syn.coulomb_layer(6, 6, 4, 2)
    np.testing.assert_allclose(V, np.swapaxes(V, -1, -2))
    assert abs(V[3, 3, 0, 0] - 5.0 / 0.3) < 1e-12     # q = 0 at the centre index
    # minimal image is symmetric under q -> -q on the centred mesh
    q = syn.qnorm_min_image(6, 6)
    np.testing.assert_allclose(q[1:, 1:], q[1:, 1:][::-1, ::-1], atol=1e-12)
