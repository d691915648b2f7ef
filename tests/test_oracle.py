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


# ------------------------------------------- the reference's solver property tests
# tests/test_scf_acceleration.py:43-101 and tests/test_solver_newton.py:204-241,
# run on the oracle in FP64 (the reference runs the SCF ones in float32).  They pin
# the branches of solve_scf (DIIS, ODA step, trust-radius clip, level shift) and
# of Newton's trust region by properties the mathematics fixes, not by the
# oracle's own formulas.

def _ref_kernel(case):
    import ref_cases as RC
    args, kw = case
    return O.OracleKernel(*args, **kw)


def _dirac_scf(acc, **extra):
    import ref_cases as RC
    K = _ref_kernel(RC.dirac())
    cfg = dict(max_iter=800, density_tol=1e-6, comm_tol=1e-5, mixing=0.5)
    cfg.update(extra)
    return K, O.solve_scf(K, np.zeros_like(K.h), 16.0, config=O.SCFConfig(acceleration=acc, **cfg))


def test_scf_schemes_agree_on_energy():
    """test_scf_acceleration.py:43-54 (spread < 1e-3); FP64 agrees far tighter."""
    E = {acc: _dirac_scf(acc)[1].energy for acc in ("linear", "diis", "oda")}
    assert max(E.values()) - min(E.values()) < 1e-8, E


def test_scf_diis_far_faster_than_linear():
    """test_scf_acceleration.py:57-68: DIIS needs < 1/3 of linear mixing's steps."""
    import ref_cases as RC
    K, lin = _dirac_scf("linear")
    _, diis = _dirac_scf("diis")
    assert lin.converged and diis.converged
    assert diis.iterations < lin.iterations // 3
    np.testing.assert_allclose(diis.energy, lin.energy, atol=1e-4, rtol=1e-5)
    assert RC.comm_rms(diis.fock_matrix, diis.density_matrix, K.w2d) < 1e-4


def test_scf_trust_radius_and_damping_converge():
    """test_scf_acceleration.py:70-88: the clipped and the damped DIIS reach the
    free DIIS energy; every clipped step is within the trust radius."""
    K, free = _dirac_scf("diis")
    _, clipped = _dirac_scf("diis", trust_radius=0.5)
    _, damped = _dirac_scf("diis", diis_damping=0.7)
    assert clipped.converged and damped.converged
    np.testing.assert_allclose(clipped.energy, free.energy, atol=1e-4, rtol=1e-5)
    np.testing.assert_allclose(damped.energy, free.energy, atol=1e-4, rtol=1e-5)
    # a radius below the first step must bind: the clipped run then differs from
    # the free one while reaching the same fixed point
    _, tight = _dirac_scf("diis", trust_radius=0.05)
    assert tight.converged and tight.iterations > free.iterations
    np.testing.assert_allclose(tight.energy, free.energy, atol=1e-8)


def test_scf_oda_energy_matches_and_is_monotone():
    """test_scf_acceleration.py:91-101: ODA reaches linear mixing's energy and its
    energy history never rises (the step is the exact line minimum of the
    quadratic model, so a wrong sign of lambda would climb)."""
    _, lin = _dirac_scf("linear")
    _, oda = _dirac_scf("oda")
    np.testing.assert_allclose(oda.energy, lin.energy, atol=1e-3, rtol=1e-5)
    hE = np.asarray(oda.history["E"])
    assert hE.size > 1 and np.all(np.diff(hE) <= 1e-4)
    assert oda.iterations < lin.iterations


def _two_level_case(nk=2, e1=-0.2, e2=0.3, T=0.1, seed=5):
    """Every k holds the same 2x2 matrix U diag(e1, e2) U^H (U random unitary),
    no interaction: F = h at every iteration."""
    rng = np.random.default_rng(seed)
    Z = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    U, _ = np.linalg.qr(Z)
    hk = U @ np.diag([e1, e2]) @ U.conj().T
    h = np.broadcast_to(hk, (nk, nk, 2, 2)).copy()
    return (np.ones((nk, nk)), h, np.zeros((nk, nk, 1, 1), np.complex128), T), U


@pytest.mark.parametrize("shift", [0.0, 0.4, 2.5])
def test_level_shift_two_level_closed_form(shift):
    """reference_scf.py:87-96 on a two-level system.  Half filling of two levels
    fixes mu = (a + b)/2 for any pair (f(x) + f(-x) = 1), so mu_raw sits midway,
    the shift moves the levels apart to Delta' = Delta + s tanh(Delta / 4T), and
    the lower level's occupation is f(-Delta'/2).  A wrong sign, the occupied
    instead of the empty levels shifted, or an ignored shift all give another
    occupation."""
    e1, e2, T = -0.2, 0.3, 0.1
    (w, h, V, _), U = _two_level_case(e1=e1, e2=e2, T=T)
    K = O.OracleKernel(w, h, V, T)          # V = 0: F = h
    r = O.solve_scf(K, np.zeros_like(h), 4.0, config=O.SCFConfig(
        max_iter=1, mixing=1.0, level_shift=shift))
    D = e2 - e1
    Dp = D + shift * np.tanh(D / (4 * T))
    p1 = 1.0 / (1.0 + np.exp(-Dp / (2 * T)))
    P = U @ np.diag([p1, 1.0 - p1]) @ U.conj().T
    np.testing.assert_allclose(r.density_matrix, np.broadcast_to(P, h.shape), atol=1e-13)


def test_level_shift_keeps_gapped_fixed_point():
    """A shift larger than the gap leaves a cold insulator's fixed point alone
    (only empty levels move up, the occupations stay 0/1), while the converged
    state stays stationary ([F, P] = 0)."""
    import ref_cases as RC
    K = _ref_kernel(RC.gapped())
    P0 = np.zeros_like(K.h)
    kw = dict(max_iter=300, density_tol=1e-11, comm_tol=1e-10, acceleration="diis")
    r0 = O.solve_scf(K, P0, 16.0, config=O.SCFConfig(**kw))
    r1 = O.solve_scf(K, P0, 16.0, config=O.SCFConfig(level_shift=10.0, **kw))
    ev = np.linalg.eigvalsh(r0.fock_matrix)
    assert ev[..., 1].min() - ev[..., 0].max() < 10.0        # the shift exceeds the gap
    assert r0.converged and r1.converged
    assert abs(r1.energy - r0.energy) < 1e-10
    np.testing.assert_allclose(r1.density_matrix, r0.density_matrix, atol=1e-9)
    assert RC.comm_rms(r1.fock_matrix, r1.density_matrix, K.w2d) < 1e-9


@pytest.mark.parametrize("nk,nb,T,ex,seed", [(4, 4, 0.05, 0.3, 7), (5, 5, 0.05, 0.4, 4),
                                             (4, 5, 0.08, 0.5, 11)])
def test_newton_multiband_stationary_and_conserves_particles(nk, nb, T, ex, seed):
    """test_solver_newton.py:216-230."""
    import ref_cases as RC
    K = _ref_kernel(RC.multiband(nk, nb, T, ex, seed))
    ne = float(nk * nk * nb * 0.5)
    r = O.solve(K, np.zeros_like(K.h), ne, config=O.SolverConfig(
        max_iter=200, tol_grad=1e-6, optimizer="newton"))
    assert r.converged
    assert RC.comm_rms(r.fock, r.density, K.w2d) < 1e-4
    np.testing.assert_allclose(float(np.sum(K.w2d[..., None] * r.p)), ne, atol=1e-4)


def test_newton_superlinear_on_dirac():
    """test_solver_newton.py:232-241: Newton reaches CG's energy in fewer than a
    fifth of its outer iterations."""
    import ref_cases as RC
    K = _ref_kernel(RC.dirac(ex=0.5))
    P0 = np.zeros_like(K.h)
    cfg = dict(max_iter=2000, tol_grad=1e-7)
    cg = O.solve(K, P0, 16.0, config=O.SolverConfig(optimizer="cg", **cfg))
    nw = O.solve(K, P0, 16.0, config=O.SolverConfig(optimizer="newton", **cfg))
    assert nw.converged
    np.testing.assert_allclose(nw.energy, cg.energy, atol=1e-6, rtol=1e-7)
    assert nw.n_iter < cg.n_iter // 5


def _hartree_multiband(hh=2.0):
    """multiband(4, 4) with a strong on-band Hartree repulsion: a full SCF step
    overshoots, so the ODA line minimum lies inside (0, 1)."""
    import ref_cases as RC
    (w, h, V, T), _ = RC.multiband(4, 4, 0.05, 0.3, 7)
    return (w, h, V, T), dict(include_hartree=True, hartree_matrix=hh * np.eye(4),
                              reference_density=np.zeros_like(h))


def _thermal_density(K, h, n, T, scale, seed):
    """A particle-conserving thermal density of a randomly perturbed h."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal(h.shape) + 1j * rng.standard_normal(h.shape)
    ev, U = np.linalg.eigh(h + scale * (X + np.conj(np.swapaxes(X, -1, -2))))
    mu = O.find_chemical_potential(ev, K.w2d, n, T)
    return np.einsum("...in,...n,...jn->...ij", U, O.fermidirac(ev - mu, T), np.conj(U))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oda_step_is_the_line_minimum(seed):
    """reference_scf.py:315-328.  One ODA step from P moves along the Roothaan
    step d = P_new - P to the minimum of the energy on the segment [P, P + d]:
    the energy there is below every point of a 201-point grid on the segment
    and its central difference along d vanishes (the energy is quadratic in P,
    so the minimum is where dE/dlambda = 0 unless it is clipped to an end)."""
    (w, h, V, T), kw = _hartree_multiband()
    K = O.OracleKernel(w, h, V, T, **kw)
    a = K.as_args()

    def energy(P):
        S, H, _ = O.build_fock(P, h=a["h"], VR=a["VR"], refP=a["refP"], HH=a["HH"], w2d=K.w2d,
                               include_exchange=True, include_hartree=True,
                               exchange_hermitian_channel_packing=a["exchange_hermitian_channel_packing"])
        return O.hf_energy(P, h=a["h"], Sigma=S, H=H, weights_b=K.weights_b)

    P0 = _thermal_density(K, h, 32.0, T, 0.5 * seed, seed)
    Pn = O.solve_scf(K, P0, 32.0, config=O.SCFConfig(max_iter=1, mixing=1.0)).density_matrix
    Po = O.solve_scf(K, P0, 32.0, config=O.SCFConfig(max_iter=1, acceleration="oda")).density_matrix
    d = Pn - P0
    lam = float(np.real(np.vdot(d, Po - P0)) / np.real(np.vdot(d, d)))
    np.testing.assert_allclose(Po, P0 + lam * d, atol=1e-13)
    assert 0.02 < lam < 0.98
    grid = np.linspace(0.0, 1.0, 201)
    assert energy(Po) <= min(energy(P0 + g * d) for g in grid) + 1e-12
    eps = 1e-4
    slope = (energy(P0 + (lam + eps) * d) - energy(P0 + (lam - eps) * d)) / (2 * eps)
    assert abs(slope) < 1e-7 * max(1.0, abs(energy(Po)))


def test_trust_radius_clips_the_step_to_the_radius():
    """reference_scf.py:332-339: a step longer than trust_radius (weighted norm
    sqrt(sum_k w_k |S_k|^2 / sum w)) is scaled onto the radius, direction kept;
    a shorter one is untouched."""
    import ref_cases as RC
    args, kw = RC.dirac()
    K = O.OracleKernel(*args, **kw)
    P0 = _thermal_density(K, K.h, 16.0, K.T, 0.8, 3)
    free = O.solve_scf(K, P0, 16.0, config=O.SCFConfig(max_iter=1, mixing=0.5)).density_matrix - P0
    n_free = np.sqrt(np.sum(K.w2d * np.sum(np.abs(free) ** 2, axis=(-2, -1))) / np.sum(K.w2d))
    for r in (0.25 * n_free, 0.5 * n_free, 2.0 * n_free):
        s = O.solve_scf(K, P0, 16.0, config=O.SCFConfig(max_iter=1, mixing=0.5, trust_radius=r)
                        ).density_matrix - P0
        n = np.sqrt(np.sum(K.w2d * np.sum(np.abs(s) ** 2, axis=(-2, -1))) / np.sum(K.w2d))
        np.testing.assert_allclose(n, min(r, n_free), rtol=1e-12)
        np.testing.assert_allclose(s, free * (n / n_free), atol=1e-13)


@pytest.mark.parametrize("nb", [1, 3, 8])
def test_qr_retract_has_positive_real_r_diagonal(nb):
    """solver.py:108-119: after the phase normalisation, U^H (I - tau G) is
    upper triangular with a positive real diagonal (the unique QR factor), for
    general and skew-Hermitian G and per-k tau."""
    rng = np.random.default_rng(nb)
    G = rng.standard_normal((2, 3, nb, nb)) + 1j * rng.standard_normal((2, 3, nb, nb))
    for Gx in (G, OH._skew(G)):
        tau = rng.uniform(0.0, 1.5, (2, 3))
        U = O.qr_retract(Gx, tau)
        A = np.eye(nb) - tau[..., None, None] * Gx
        R = np.conj(np.swapaxes(U, -1, -2)) @ A
        np.testing.assert_allclose(np.tril(R, -1), 0, atol=1e-12)
        d = np.diagonal(R, axis1=-2, axis2=-1)
        assert np.all(np.real(d) > 0) and np.abs(np.imag(d)).max() < 1e-12
        np.testing.assert_allclose(U @ R, A, atol=1e-12)
