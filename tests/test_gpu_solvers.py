"""GPU parity of the per-k kernels and the solvers against the FP64 oracle.

Bars: batched eigh -- eigenvalues within 1e-12 (FP64) / 2e-5 (FP32) relative to
LAPACK, reconstruction and orthonormality; mu -- 1e-12; CG / SCF -- converged
energy within 1e-10 per k-point (weights sum to 1, FP64) of the oracle run with
the same configuration, density within 1e-6; the reference fixture's
converged state within its complex64 precision.
"""

import numpy as np
import pytest

import oracle as O
from jax_hf_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _herm(X):
    return 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))


@pytest.mark.parametrize("nb", [1, 2, 3, 4, 5, 8, 13, 16, 17, 24, 32])
@pytest.mark.parametrize("dt", ["c128", "c64"])
def test_eigh_batched(dev, nb, dt):
    import torch
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(nb)
    A = _herm(rng.standard_normal((5, 7, nb, nb)) + 1j * rng.standard_normal((5, 7, nb, nb)))
    A[0, 0] = np.diag(np.arange(nb, dtype=float))           # already diagonal
    A[0, 1] = np.eye(nb)                                     # fully degenerate
    if nb >= 2:
        A[0, 2] = np.diag(np.r_[np.ones(nb - 1), 2.0])       # degenerate block
    cdt = torch.complex128 if dt == "c128" else torch.complex64
    At = torch.from_numpy(np.ascontiguousarray(A)).to(dev).to(cdt)
    lam = torch.empty((5, 7, nb), dtype=torch.float64, device=dev)
    V = torch.empty_like(At)
    K.eigh(At, lam, V)
    w_ref = np.linalg.eigvalsh(A)
    lam_h, V_h = _np(lam), _np(V).astype(np.complex128)
    tol = 1e-12 if dt == "c128" else 2e-5
    scale = np.abs(w_ref).max()
    assert np.abs(lam_h - w_ref).max() <= tol * scale * 10
    assert np.all(np.diff(lam_h, axis=-1) >= -tol * scale)
    rec = V_h @ (lam_h[..., None] * np.conj(np.swapaxes(V_h, -1, -2)))
    assert np.abs(rec - A).max() <= 50 * tol * scale
    eye = np.conj(np.swapaxes(V_h, -1, -2)) @ V_h
    assert np.abs(eye - np.eye(nb)).max() <= 50 * tol


@pytest.mark.parametrize("rel_tol", [1e-10, 1e-6])
def test_eigh_relative_tolerance(dev, rel_tol):
    """The tolerance form stops at off-diagonals <= rel_tol ||A||_F: the basis
    stays unitary to rounding, A is reproduced to ~rel_tol ||A||_F, and from a
    nearly diagonal start (the polish's warm case) it is as exact as the
    full-precision form up to that size."""
    import torch
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(5)
    nb = 16
    D = np.diag(np.linspace(-3, 3, nb))
    E = _herm(rng.standard_normal((40, nb, nb)) + 1j * rng.standard_normal((40, nb, nb)))
    A = D + 1e-3 * E
    At = torch.from_numpy(A).to(dev)
    lam = torch.empty((40, nb), dtype=torch.float64, device=dev)
    V = torch.empty_like(At)
    K.eigh(At, lam, V, rel_tol=rel_tol)
    Vh, lh = _np(V), _np(lam)
    eye = np.conj(np.swapaxes(Vh, -1, -2)) @ Vh
    assert np.abs(eye - np.eye(nb)).max() < 1e-13
    nrm = np.sqrt(np.sum(np.abs(A) ** 2, axis=(1, 2))).max()
    rec = Vh @ (lh[..., None] * np.conj(np.swapaxes(Vh, -1, -2)))
    assert np.abs(rec - A).max() <= 4 * nb * rel_tol * nrm
    assert np.abs(lh - np.linalg.eigvalsh(A)).max() <= 4 * nb * rel_tol * nrm


def test_density_and_matmul(dev):
    import torch
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(0)
    nb = 6
    Q = np.linalg.qr(rng.standard_normal((3, 4, nb, nb)) + 1j * rng.standard_normal((3, 4, nb, nb)))[0]
    p = rng.uniform(0, 1, (3, 4, nb))
    Qt = torch.from_numpy(Q).to(dev)
    P = torch.empty_like(Qt)
    K.density(Qt, P, p=torch.from_numpy(p).to(dev))
    ref = _herm(np.einsum("...in,...n,...jn->...ij", Q, p, np.conj(Q)))
    assert np.abs(_np(P) - ref).max() < 1e-14
    F = _herm(rng.standard_normal((3, 4, nb, nb)) + 1j * rng.standard_normal((3, 4, nb, nb)))
    Ft = torch.from_numpy(F).to(dev)
    out = torch.empty_like(Qt)
    K.matmul(Qt, Ft, out, C=Qt, ops=1)       # Q^H F Q
    ref = np.conj(np.swapaxes(Q, -1, -2)) @ F @ Q
    assert np.abs(_np(out) - ref).max() < 1e-13


@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("nb", [6, 16, 24, 32])
def test_matmul_operand_forms(dev, nb, dt):
    """Every op(A) op(B) [op(C)] form of the batched product against NumPy.
    nb = 16 runs on the matrix cores (16x16x4 MFMA), nb = 6 on the LDS loop,
    nb = 24 / 32 on the 256-thread workgroup form."""
    import torch
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(nb)
    cdt = np.complex128 if dt == "c128" else np.complex64
    tdt = torch.complex128 if dt == "c128" else torch.complex64
    tol = 1e-13 if dt == "c128" else 2e-5
    shp = (5, 7, nb, nb)
    A, B, C = [(rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).astype(cdt) for _ in range(3)]
    H = lambda X: np.conj(np.swapaxes(X, -1, -2))  # noqa: E731
    At, Bt, Ct = (torch.from_numpy(X).to(dev) for X in (A, B, C))
    out = torch.empty_like(At)
    for ops in range(8):                     # bit 0: A^H, bit 1: B^H, bit 2: C^H
        a = H(A) if ops & 1 else A
        b = H(B) if ops & 2 else B
        c = H(C) if ops & 4 else C
        for use_c in (False, True):
            K.matmul(At, Bt, out, C=Ct if use_c else None, ops=ops)
            ref = a.astype(np.complex128) @ b
            if use_c:
                ref = ref @ c
            scale = np.abs(ref).max()
            assert np.abs(_np(out) - ref).max() <= tol * scale, (ops, use_c)


def test_diis_coefficients_match_host_solve(dev):
    """Device C-DIIS (B update + masked augmented solve) against the host
    diis_solve, which is pinned to the closed form B^-1 1 / (1^T B^-1 1)."""
    import torch
    from jax_hf_amd import _ops as K
    from jax_hf_amd.reference_scf import diis_solve
    rng = np.random.default_rng(11)
    nh = 6
    Bd = torch.zeros((nh, nh), dtype=torch.float64, device=dev)
    Bh = np.zeros((nh, nh))
    coef = torch.zeros(8, dtype=torch.float64, device=dev)
    R = rng.standard_normal((12, 40)) * np.logspace(0, -3, 12)[:, None]   # shrinking residuals
    slots = [None] * nh
    for k in range(12):
        idx, count = k % nh, min(k + 1, nh)
        slots[idx] = R[k]
        row = np.array([R[k] @ slots[j] if slots[j] is not None else 0.0 for j in range(nh)])
        Bh[idx, :] = row
        Bh[:, idx] = row
        K.diis_coefficients(torch.from_numpy(row).to(dev), Bd, idx, count, coef)
        active = np.arange(nh) < count
        cs, ok = diis_solve(Bh * np.outer(active, active), active)
        ref = np.zeros(8)
        if count >= 2 and ok:
            ref[:nh] = np.where(active, cs, 0.0)
        else:
            ref[idx] = 1.0
        got = _np(coef)
        assert np.abs(got - ref).max() <= 1e-7 * max(1.0, np.abs(ref).max()), k
        assert np.allclose(_np(Bd), Bh, rtol=0, atol=0)
    # exactly singular system (all-zero B): plain step
    Bd.zero_()
    K.diis_coefficients(torch.zeros(nh, dtype=torch.float64, device=dev), Bd, 2, 4, coef)
    assert np.array_equal(_np(coef), np.eye(8)[2])


@pytest.mark.parametrize("nb", [5, 16, 32])
def test_orthonormalize_fp32_basis(dev, nb):
    """Two Newton-Schulz steps take an FP32-rounded unitary to FP64 unitarity,
    moving it by no more than its rounding (the polar factor is the nearest unitary)."""
    import torch
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(nb)
    Q = np.linalg.qr(rng.standard_normal((40, nb, nb)) + 1j * rng.standard_normal((40, nb, nb)))[0]
    Q32 = Q.astype(np.complex64).astype(np.complex128)
    U = torch.from_numpy(Q32).to(dev)
    K.orthonormalize(U, 2)
    Uh = _np(U)
    eye = np.conj(np.swapaxes(Uh, -1, -2)) @ Uh
    assert np.abs(eye - np.eye(nb)).max() < 1e-14
    assert np.abs(Uh - Q32).max() < 1e-6
    # polar factor of Q32 (SVD) -- the fixed point Newton-Schulz converges to
    W, _, Vh = np.linalg.svd(Q32)
    assert np.abs(Uh - W @ Vh).max() < 1e-13


@pytest.mark.parametrize("dt", ["c128", "c64", "c128h32"])
@pytest.mark.parametrize("nb", [5, 16, 24])
def test_scf_resid_dots_matches_numpy(dev, nb, dt):
    """Residuals and the DIIS B row in one pass (reference_scf.py:110-111, 291-294);
    c128h32: FP64 build with the residual history stored in complex64 (nb = 16)."""
    import torch
    from jax_hf_amd import _ops as K
    if dt == "c128h32" and nb != 16:
        pytest.skip("complex64 residual history is the nb = 16 form")
    rng = np.random.default_rng(3 + nb)
    cdt = np.complex128 if dt != "c64" else np.complex64
    hdt = np.complex64 if dt != "c128" else np.complex128
    tol = 1e-12 if dt == "c128" else 1e-4
    nk = 37
    shp = (nk, nb, nb)
    cplx = lambda: (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).astype(cdt)  # noqa: E731
    F, D, Pn = _herm(cplx()), _herm(cplx()), _herm(cplx())
    Rs = [cplx().astype(hdt) for _ in range(5)]
    self_idx = 2
    t = lambda X: torch.from_numpy(np.ascontiguousarray(X)).to(dev)  # noqa: E731
    Rt = [t(R) for R in Rs]
    part = torch.zeros((8, nk), dtype=torch.float64, device=dev)
    K.scf_resid_dots(t(F), t(D), t(Pn), Rt[self_idx], Rt, self_idx, part)
    p = _np(part)
    F64, D64 = F.astype(np.complex128), D.astype(np.complex128)
    comm = F64 @ D64 - D64 @ F64
    assert np.abs(_np(Rt[self_idx]) - comm).max() <= tol * np.abs(comm).max()
    c2 = np.sum(np.abs(comm) ** 2, axis=(1, 2))
    d2 = np.sum(np.abs(Pn.astype(np.complex128) - D64) ** 2, axis=(1, 2))
    assert np.allclose(p[0], c2, rtol=tol, atol=0) and np.allclose(p[1], d2, rtol=tol, atol=0)
    for v, R in enumerate(Rs):
        ref = c2 if v == self_idx else np.real(np.sum(np.conj(comm) * R, axis=(1, 2)))
        assert np.abs(p[2 + v] - ref).max() <= tol * np.abs(c2).max(), v
        if v == self_idx:
            assert np.array_equal(p[2 + v], p[0])


@pytest.mark.parametrize("T", [0.5, 0.05, 0.005])
def test_mu_solvers_match_oracle(dev, T):
    import torch
    import jax_hf_amd.utils as U
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(3)
    eps = rng.standard_normal((20, 20, 6))
    w = rng.uniform(0.5, 1.5, (20, 20))
    wn = w / w.sum()
    target = 2.3
    mu0 = O.solve_mu(eps, wn, target, 0.1, T, 25)
    et = torch.from_numpy(eps).to(dev)
    ws = K.Workspace.get(dev, 400)
    st = torch.zeros(16, dtype=torch.float64, device=dev)
    p = torch.empty_like(et)
    K.solve_mu(ws, et, torch.from_numpy(wn).to(dev), 400, 6, target, T, mu0=0.1, maxiter=25,
               mode=K.MU_NEWTON_BRACKET, st=st, p_out=p)
    assert abs(float(st[0]) - mu0) < 1e-12
    assert np.abs(_np(p) - 1 / (1 + np.exp(-(mu0 - eps) / T))).max() < 1e-12
    n = target * w.sum()
    for method in ("bisection", "newton"):
        m_ref = O.find_chemical_potential(eps, w, n, T, method=method)
        m_gpu = float(U.find_chemical_potential(torch.from_numpy(eps).to(dev), w, n, T, method=method))
        assert abs(m_gpu - m_ref) < 1e-12


def test_density_matrix_from_fock_fixture(dev, meshgrid_case):
    import jax_hf_amd.utils as U
    d = meshgrid_case
    P, mu = U.density_matrix_from_fock(d["F_expected"].astype(np.complex128), d["weights"],
                                       float(d["electrondensity0"]), float(d["T"]))
    P0, mu0 = O.density_matrix_from_fock(d["F_expected"], d["weights"].astype(float),
                                         float(d["electrondensity0"]), float(d["T"]))
    assert abs(float(mu) - mu0) < 1e-12
    assert np.abs(_np(P) - P0).max() < 1e-12
    assert np.abs(_np(P) - d["P_expected"]).max() < 5e-6


def _fixture_kernels(d):
    import jax_hf_amd as J
    kw = dict(include_hartree=True, reference_density=d["reference_density"].astype(np.complex128),
              hartree_matrix=d["hartree_matrix"])
    h = d["hamiltonian"].astype(np.complex128)
    return (O.OracleKernel(d["weights"], h, d["coulomb_q"], float(d["T"]), **kw),
            J.HartreeFockKernel(d["weights"], h, d["coulomb_q"], float(d["T"]), **kw))


def test_cg_fixture_fixed_point(dev, meshgrid_case):
    import jax_hf_amd as J
    d = meshgrid_case
    K0, G = _fixture_kernels(d)
    n = float(d["electrondensity0"])
    cfg = J.SolverConfig(max_iter=300, tol_E=1e-12)
    r = J.solve(G, np.zeros_like(d["hamiltonian"], dtype=np.complex128), n, config=cfg)
    r0 = O.solve(K0, np.zeros_like(K0.h), n, config=O.SolverConfig(max_iter=300, tol_E=1e-12))
    assert r.converged
    assert abs(float(r.energy) - float(d["E_expected"])) < 1e-6
    assert np.abs(_np(r.density) - d["P_expected"]).max() < 1e-5
    assert abs(float(r.energy) - r0.energy) < 1e-10
    assert np.abs(_np(r.density) - r0.density).max() < 1e-6
    # the first iterations follow the oracle trajectory
    hE = _np(r.history["E"])
    assert np.abs(hE[:5] - r0.history["E"][:5]).max() < 1e-10


def test_cg_config0_fp64_matches_oracle(dev):
    import jax_hf_amd as J
    c = syn.config0(nk=24, nb=4, seed=0)
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    K0 = O.OracleKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    cfg_kw = dict(max_iter=400, tol_E=1e-12)
    r = J.solve(G, c["P0"], c["n_electrons"], config=J.SolverConfig(**cfg_kw))
    r0 = O.solve(K0, c["P0"], c["n_electrons"], config=O.SolverConfig(**cfg_kw))
    assert r.converged and r0.converged
    assert abs(float(r.energy) - r0.energy) < 1e-10          # weights sum to 1: per k-point
    assert np.abs(_np(r.density) - r0.density).max() < 1e-6
    def comm_rms(F, P):
        c = F @ P - P @ F
        return np.sqrt(np.mean(np.sum(np.abs(c) ** 2, axis=(-2, -1))))
    c_gpu, c_ref = comm_rms(_np(r.fock), _np(r.density)), comm_rms(r0.fock, r0.density)
    assert abs(c_gpu - c_ref) < 0.1 * c_ref + 1e-8


def test_cg_mixed_precision_energy_error(dev):
    """Mixed form (FP32 CG, FP64 DIIS polish) lands on the FP64 fixed point:
    |dE| < 1e-8 per k-point against the oracle's DIIS SCF converged to 1e-11
    residuals (the reference CG's own windowed stop ends ~5e-7 above it here,
    so the fixed point, not the early-stopped CG energy, is the reference);
    the FP64 form with the same polish agrees as well."""
    import jax_hf_amd as J
    c = syn.config0(nk=24, nb=4, seed=0)
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    K0 = O.OracleKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    r0 = O.solve_scf(K0, c["P0"], c["n_electrons"],
                     config=O.SCFConfig(max_iter=300, acceleration="diis", density_tol=1e-11, comm_tol=1e-11))
    assert r0.converged
    r = J.solve(G, c["P0"], c["n_electrons"], config=J.SolverConfig(max_iter=400, precision="mixed"))
    assert r.converged
    assert abs(float(r.energy) - r0.energy) < 1e-8            # |dE| < 1e-8 per k-point
    P = _np(r.density)
    assert np.abs(P - np.asarray(r0.density_matrix)).max() < 1e-6
    r64 = J.solve(G, c["P0"], c["n_electrons"],
                  config=J.SolverConfig(max_iter=400, tol_E=1e-9, precision="fp64", polish="diis"))
    assert r64.converged and abs(float(r64.energy) - r0.energy) < 1e-8


@pytest.mark.parametrize("nk,nb", [(24, 4), (24, 16)])
def test_mixed_complex64_diis_stage(dev, nk, nb):
    """The optional complex64 DIIS stage between the FP32 CG and the FP64
    polish (polish32_tol > 0): three phases in the history, the FP64 part a
    few refinement builds, and the same FP64 fixed point (|dE| < 1e-8, density
    1e-6) as the oracle's DIIS SCF."""
    import jax_hf_amd as J
    c = syn.config0(nk=nk, nb=nb, seed=0)
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    K0 = O.OracleKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    r0 = O.solve_scf(K0, c["P0"], c["n_electrons"],
                     config=O.SCFConfig(max_iter=300, acceleration="diis", density_tol=1e-11, comm_tol=1e-11))
    r = J.solve(G, c["P0"], c["n_electrons"], config=J.SolverConfig(max_iter=400, precision="mixed",
                                                                       polish32_tol=1e-4))
    assert r.converged
    ph = [int(x) for x in _np(r.history["phase_iters"])]
    assert len(ph) == 3 and sum(ph) == int(r.n_iter) and ph[1] >= 1
    assert abs(float(r.energy) - r0.energy) < 1e-8
    assert np.abs(_np(r.density) - np.asarray(r0.density_matrix)).max() < 1e-6


@pytest.mark.parametrize("bt", [1, 8])
def test_mixed_fp32_line_search_cap(dev, bt):
    """The mixed form's FP32 phase with at most `mixed_bt_max` line-search
    trials (1: the blind fallback step whenever the first trial is rejected;
    8: the reference's bt_max) reaches the oracle's FP64 fixed point either
    way (|dE| < 1e-8, density 1e-6)."""
    import jax_hf_amd as J
    c = syn.config0(nk=24, nb=4, seed=1)
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    K0 = O.OracleKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    r0 = O.solve_scf(K0, c["P0"], c["n_electrons"],
                     config=O.SCFConfig(max_iter=300, acceleration="diis", density_tol=1e-11, comm_tol=1e-11))
    r = J.solve(G, c["P0"], c["n_electrons"], config=J.SolverConfig(max_iter=400, precision="mixed",
                                                                       polish_tol=1e-9, mixed_bt_max=bt))
    assert r.converged
    assert abs(float(r.energy) - r0.energy) < 1e-8
    assert np.abs(_np(r.density) - np.asarray(r0.density_matrix)).max() < 1e-6


def _contact_kernels(nk=2, g=0.3, T=0.05):
    import jax_hf_amd as J
    h = np.zeros((nk, nk, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -1.0, 1.0
    for i in range(nk):
        for j in range(nk):
            h[i, j, 0, 1] = h[i, j, 1, 0] = 0.4 + 0.1 * (i + j)
    sz = np.diag([1.0, -1.0]).astype(complex)
    args = (np.ones((nk, nk)), h, np.full((nk, nk, 1, 1), 0.1), T)
    return (O.OracleKernel(*args, contact_terms=[(g, sz, sz)]),
            J.HartreeFockKernel(*args, contact_terms=[(g, sz, sz)]))


def test_cg_and_scf_agree_with_contact(dev):
    import jax_hf_amd as J
    K0, G = _contact_kernels()
    P0 = np.zeros((2, 2, 2, 2), np.complex128)
    dm = J.solve(G, P0, 4.0, config=J.SolverConfig(max_iter=200, tol_E=1e-9))
    scf = J.solve_scf(G, P0, 4.0, config=J.SCFConfig(max_iter=400, mixing=0.3, density_tol=1e-7,
                                                     comm_tol=1e-6))
    assert dm.converged and scf.converged
    assert abs(float(dm.energy) - float(scf.energy)) < 1e-5
    assert np.linalg.norm(_np(dm.density) - _np(scf.density_matrix)) < 1e-4
    n = dm.n_iter
    assert abs(float(_np(dm.history["E"])[n - 1]) - float(dm.energy)) < 1e-5


@pytest.mark.parametrize("acc", ["linear", "diis", "oda"])
def test_scf_matches_oracle(dev, meshgrid_case, acc):
    import jax_hf_amd as J
    d = meshgrid_case
    K0, G = _fixture_kernels(d)
    n = float(d["electrondensity0"])
    kw = dict(max_iter=300, acceleration=acc, density_tol=1e-10, comm_tol=1e-9)
    r = J.solve_scf(G, np.zeros_like(K0.h), n, config=J.SCFConfig(**kw))
    r0 = O.solve_scf(K0, np.zeros_like(K0.h), n, config=O.SCFConfig(**kw))
    assert r.converged == r0.converged
    assert abs(float(r.energy) - r0.energy) < 1e-10
    assert np.abs(_np(r.density_matrix) - r0.density_matrix).max() < 1e-8
    assert abs(r.iterations - r0.iterations) <= 2
    assert abs(float(r.chemical_potential) - r0.chemical_potential) < 1e-9


def test_scf_level_shift_and_trust_radius(dev, meshgrid_case):
    import jax_hf_amd as J
    d = meshgrid_case
    K0, G = _fixture_kernels(d)
    n = float(d["electrondensity0"])
    kw = dict(max_iter=300, level_shift=0.2, trust_radius=0.05, density_tol=1e-9, comm_tol=1e-8)
    r = J.solve_scf(G, np.zeros_like(K0.h), n, config=J.SCFConfig(**kw))
    r0 = O.solve_scf(K0, np.zeros_like(K0.h), n, config=O.SCFConfig(**kw))
    assert r.converged and r0.converged
    assert abs(float(r.energy) - r0.energy) < 1e-9
    np.testing.assert_allclose(_np(r.history["E"])[:5], r0.history["E"][:5], atol=1e-10)


def test_gradient_stop_and_unreachable_count(dev):
    import jax_hf_amd as J
    h = np.zeros((2, 2, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -0.5, 0.5
    h[..., 0, 1] = h[..., 1, 0] = 0.2
    G = J.HartreeFockKernel(np.ones((2, 2)), h, np.full((2, 2, 1, 1), 0.5), 0.1)
    P0 = np.zeros_like(h)
    P0[..., 0, 0], P0[..., 1, 1] = 0.6, 0.15
    P0[..., 0, 1] = P0[..., 1, 0] = 0.25
    tol = 1e-4
    r = J.solve(G, P0, 3.0, config=J.SolverConfig(max_iter=300, tol_grad=tol))
    assert r.converged and 1 < r.n_iter < 300
    hG = _np(r.history["grad_norm"])[:r.n_iter]
    assert hG[-1] <= tol and np.all(hG[:-1] > tol)
    r3 = J.solve(G, P0, 3.0, config=J.SolverConfig(max_iter=3, tol_grad=1e-14))
    assert not r3.converged and r3.n_iter == 3
    with pytest.raises(ValueError, match="physically reachable range"):
        J.solve(G, P0, 9.0)
    with pytest.raises(ValueError, match="optimizer"):
        J.solve(G, P0, 3.0, config=J.SolverConfig(optimizer="bfgs"))


# ----------------------------------------------------------------- Newton

def _multiband(nk, nb, T, ex, seed):
    import jax_hf_amd as J
    rng = np.random.default_rng(seed)
    diag = np.linspace(-1.0, 1.0, nb)
    A = (rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))) * 0.35
    A = 0.5 * (A + A.conj().T)
    h = np.zeros((nk, nk, nb, nb), np.complex128)
    kx = np.arange(nk) / nk * 2 * np.pi
    for i in range(nk):
        for j in range(nk):
            hk = np.diag(diag).astype(np.complex128) + A * (np.cos(kx[i]) + np.cos(kx[j]))
            h[i, j] = 0.5 * (hk + hk.conj().T)
    args = (np.ones((nk, nk)), h, np.full((nk, nk, 1, 1), ex), T)
    return O.OracleKernel(*args), J.HartreeFockKernel(*args)


def test_newton_matches_oracle_and_cg(dev):
    import jax_hf_amd as J
    K0, G = _contact_kernels()
    P0 = np.zeros((2, 2, 2, 2), np.complex128)
    nw = J.solve(G, P0, 4.0, config=J.SolverConfig(max_iter=200, tol_grad=1e-7, optimizer="newton"))
    nw0 = O.solve(K0, P0, 4.0, config=O.SolverConfig(max_iter=200, tol_grad=1e-7, optimizer="newton"))
    assert nw.converged and nw0.converged
    assert abs(float(nw.energy) - nw0.energy) < 1e-10
    assert abs(nw.n_iter - nw0.n_iter) <= 1


@pytest.mark.parametrize("nk,nb,T,ex,seed", [(4, 4, 0.05, 0.3, 7), (5, 5, 0.05, 0.4, 4)])
def test_newton_multiband_stationary(dev, nk, nb, T, ex, seed):
    import jax_hf_amd as J
    K0, G = _multiband(nk, nb, T, ex, seed)
    P0 = np.zeros((nk, nk, nb, nb), np.complex128)
    ne = float(nk * nk * nb * 0.5)
    r = J.solve(G, P0, ne, config=J.SolverConfig(max_iter=200, tol_grad=1e-6, optimizer="newton"))
    r0 = O.solve(K0, P0, ne, config=O.SolverConfig(max_iter=200, tol_grad=1e-6, optimizer="newton"))
    assert r.converged
    Fh, Ph = _np(r.fock), _np(r.density)
    comm = Fh @ Ph - Ph @ Fh
    assert np.sqrt(np.mean(np.sum(np.abs(comm) ** 2, axis=(-2, -1)))) < 1e-4
    assert abs(float(np.sum(np.ones((nk, nk))[..., None] * _np(r.p))) - ne) < 1e-6
    assert abs(float(r.energy) - r0.energy) < 1e-9 * nk * nk


def test_newton_cold_start_relaxation(dev):
    import jax_hf_amd as J
    h = np.zeros((2, 2, 2, 2), np.complex128)
    h[..., 0, 0], h[..., 1, 1] = -0.5, 0.5
    args = (np.ones((2, 2)), h, np.full((2, 2, 1, 1), 0.25), 0.1)
    G = J.HartreeFockKernel(*args)
    P0 = np.zeros_like(h)
    cg = J.solve(G, P0, 4.0, config=J.SolverConfig(max_iter=400, tol_E=1e-12, tol_grad=1e-9))
    nw = J.solve(G, P0, 4.0, config=J.SolverConfig(max_iter=200, tol_grad=1e-8, optimizer="newton"))
    np.testing.assert_allclose(float(nw.energy), float(cg.energy), atol=1e-8, rtol=1e-9)


def test_newton_deflation_matches_oracle(dev):
    import jax_hf_amd as J
    K0, G = _contact_kernels()
    P0 = np.zeros((2, 2, 2, 2), np.complex128)
    base = O.solve(K0, P0, 4.0, config=O.SolverConfig(optimizer="newton", tol_grad=1e-8))
    kw = dict(optimizer="newton", tol_grad=1e-8, max_iter=100, deflation_targets=base.density[None],
              deflation_sigma=0.05, deflation_length=0.5)
    r = J.solve(G, P0, 4.0, config=J.SolverConfig(**kw))
    r0 = O.solve(K0, P0, 4.0, config=O.SolverConfig(**kw))
    assert abs(float(r.energy) - r0.energy) < 1e-9
    assert r.n_iter == r0.n_iter


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cg_graph_replay_matches_eager(dev, prec):
    """The captured-graph CG loop replays exactly the eager launch sequence:
    every iteration's energy and gradient norm (one replay each, with the host
    readback between replays) and the final state agree bitwise."""
    import jax_hf_amd as J
    c = syn.config0(nk=24, nb=4, seed=0)
    npd = np.complex128 if prec == "fp64" else np.complex64
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"].astype(npd), c["coulomb_q"], c["T"])
    res = {}
    for use_graph in (False, True):
        cfg = J.SolverConfig(max_iter=40, tol_E=1e-10, use_graph=use_graph, cg_restart=7)
        res[use_graph] = J.solve(G, c["P0"].astype(npd), c["n_electrons"], config=cfg)
    a, b = res[False], res[True]
    assert a.n_iter == b.n_iter and a.n_iter > 10
    np.testing.assert_array_equal(_np(a.history["E"]), _np(b.history["E"]))
    np.testing.assert_array_equal(_np(a.history["grad_norm"]), _np(b.history["grad_norm"]))
    np.testing.assert_array_equal(_np(a.density), _np(b.density))
    assert float(a.energy) == float(b.energy)


@pytest.mark.parametrize("nb", [4, 7, 16])
def test_cg_gauss_jordan_cayley_matches_spectral(dev, nb):
    """The Gauss-Jordan Cayley factor is the spectral one: same converged energy
    (1e-10 per k-point) and density, FP64, including nb = 16 and odd nb."""
    import jax_hf_amd as J
    nk = 12
    h = syn.tight_binding(nk, nk, nb, seed=nb)
    V = syn.coulomb_layer(nk, nk, nb, 2)
    w = syn.uniform_weights(nk, nk)
    G = J.HartreeFockKernel(w, h, V, 0.1)
    P0 = np.zeros_like(h)
    out = {}
    for form in ("spectral", "gauss_jordan"):
        cfg = J.SolverConfig(max_iter=300, tol_E=1e-11, cayley=form)
        out[form] = J.solve(G, P0, nb / 2.0, config=cfg)
    a, b = out["spectral"], out["gauss_jordan"]
    assert a.converged and b.converged
    assert abs(float(a.energy) - float(b.energy)) < 1e-10
    assert np.abs(_np(a.density) - _np(b.density)).max() < 1e-5


def test_repeated_solves_reuse_cached_state(dev):
    """A repeated solve on the same kernel reuses the kept CG runner, polish
    state and their captured graphs: the second result equals the first
    bitwise, the first result's tensors are not overwritten by the second
    solve, and a solve at another electron count gets its own state."""
    import jax_hf_amd as J
    c = syn.config0(nk=24, nb=4, seed=0)
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    cfg = J.SolverConfig(max_iter=400, precision="mixed")
    r1 = J.solve(G, c["P0"], c["n_electrons"], config=cfg)
    snap = {k: _np(getattr(r1, k)).copy() for k in ("Q", "p", "density", "fock")}
    E1 = float(r1.energy)
    r2 = J.solve(G, c["P0"], c["n_electrons"], config=cfg)
    assert float(r2.energy) == E1 and r2.n_iter == r1.n_iter
    for k, v in snap.items():
        np.testing.assert_array_equal(_np(getattr(r1, k)), v)
        np.testing.assert_array_equal(_np(getattr(r2, k)), v)
    r3 = J.solve(G, c["P0"], c["n_electrons"] + 1.0, config=cfg)
    K0 = O.OracleKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    r0 = O.solve_scf(K0, c["P0"], c["n_electrons"] + 1.0,
                     config=O.SCFConfig(max_iter=300, acceleration="diis", density_tol=1e-11, comm_tol=1e-11))
    assert r3.converged and r0.converged and abs(float(r3.energy) - r0.energy) < 1e-8
    # the kept state refers back to its kernel weakly: dropping the kernel frees
    # it (and the state, graphs included) at once, without the cyclic collector
    import gc
    import weakref
    wr = weakref.ref(G)
    was = gc.isenabled()
    gc.disable()
    try:
        del G
        assert wr() is None
    finally:
        if was:
            gc.enable()


@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("nb", [3, 16, 32])
def test_extrap_rotate_and_rotate_density(dev, dt, nb):
    """The fused per-k kernels of the DIIS polish against NumPy: Ft =
    U^H herm(sum_v c_v X_v) U, and U <- U W with P = herm(U diag(occ) U^H)."""
    import torch
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(nb)
    nk = 37
    cdt = torch.complex128 if dt == "c128" else torch.complex64
    tol = 1e-13 if dt == "c128" else 2e-5
    Xs = [rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb)) for _ in range(5)]
    cs = rng.standard_normal(5)
    U = np.linalg.qr(rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb)))[0]
    W = np.linalg.qr(rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb)))[0]
    occ = rng.uniform(0, 1, (nk, nb))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(cdt)   # noqa: E731
    Ut, Ft = t(U), torch.empty((nk, nb, nb), dtype=cdt, device=dev)
    K.extrap_rotate(Ft, Ut, [t(x) for x in Xs], cs)
    S = sum(c * x for c, x in zip(cs, Xs))
    ref = np.conj(np.swapaxes(U, -1, -2)) @ _herm(S) @ U
    assert np.abs(_np(Ft) - ref).max() <= tol * np.abs(ref).max()
    P = torch.empty_like(Ft)
    K.rotate_density(Ut, t(W), torch.from_numpy(occ).to(dev), P)
    Un = U @ W
    assert np.abs(_np(Ut) - Un).max() <= tol * 10
    Pr = np.einsum("kin,kn,kjn->kij", Un, occ, np.conj(Un))
    assert np.abs(_np(P) - Pr).max() <= tol * 10


@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("nb", [3, 8, 16, 24])
@pytest.mark.parametrize("warm", [False, True])
def test_eigh_basis(dev, dt, nb, warm):
    """hfx_eigh_batched_basis: eigh of A = U^H F U with the rotations accumulated
    onto U in place -- the new U diagonalises F (U^H F U = diag(lam) against
    LAPACK's eigenvalues of F), stays unitary, and equals U times the plain
    eigh's eigenvectors of A up to each column's phase (non-degenerate F).
    warm: U already near F's eigenvectors (the polish's running basis)."""
    import torch
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(100 + nb)
    nk = 29
    cdt = torch.complex128 if dt == "c128" else torch.complex64
    tol = 1e-12 if dt == "c128" else 3e-5
    F = _herm(rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb)))
    if warm:
        U = np.linalg.eigh(F + 1e-3 * _herm(rng.standard_normal((nk, nb, nb))))[1]
    else:
        U = np.linalg.qr(rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb)))[0]
    A = np.conj(np.swapaxes(U, -1, -2)) @ F @ U
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(cdt)   # noqa: E731
    At, Ut = t(A), t(U)
    lam = torch.empty((nk, nb), dtype=torch.float64, device=dev)
    Vo = torch.empty_like(Ut)
    K.eigh_basis(At, lam, Ut, herm_in=True, out=Vo)             # out of place: U unchanged
    assert torch.equal(Ut, t(U))
    K.eigh_basis(At, lam, Ut, herm_in=True)                     # in place: the same V
    assert torch.equal(Ut, Vo)
    Un, ln = _np(Ut).astype(np.complex128), _np(lam)
    lref = np.linalg.eigvalsh(F)
    sc = np.abs(lref).max()
    assert np.abs(ln - lref).max() <= tol * sc * 10
    D = np.conj(np.swapaxes(Un, -1, -2)) @ F @ Un
    assert np.abs(D - ln[:, None, :] * np.eye(nb)).max() <= tol * sc * 20
    assert np.abs(np.conj(np.swapaxes(Un, -1, -2)) @ Un - np.eye(nb)).max() <= tol * 20
    V = torch.empty_like(At)
    lam2 = torch.empty_like(lam)
    K.eigh(t(A), lam2, V, herm_in=True)
    Uv = U @ _np(V).astype(np.complex128)
    ph = np.einsum("kin,kin->kn", np.conj(Uv), Un)          # column overlaps, |.| = 1
    assert np.abs(np.abs(ph) - 1).max() <= tol * 1e3
    assert np.abs(_np(lam2) - ln).max() <= tol * sc * 10


@pytest.mark.parametrize("nk,nb,T,ex,seed", [(4, 4, 0.05, 0.3, 7), (5, 5, 0.05, 0.4, 4)])
def test_newton_mixed_precision_matches_fp64(dev, nk, nb, T, ex, seed):
    """precision="mixed" Newton (complex64 Fock-response builds inside the
    Hessian-vector products, everything else FP64) converges to tol_grad 1e-6
    like the FP64 Newton, to the same energy (1e-9 per k-point) and the
    oracle's."""
    import jax_hf_amd as J
    K0, G = _multiband(nk, nb, T, ex, seed)
    P0 = np.zeros((nk, nk, nb, nb), np.complex128)
    ne = float(nk * nk * nb * 0.5)
    cfg = dict(max_iter=200, tol_grad=1e-6, optimizer="newton")
    r64 = J.solve(G, P0, ne, config=J.SolverConfig(precision="fp64", **cfg))
    rmx = J.solve(G, P0, ne, config=J.SolverConfig(precision="mixed", **cfg))
    r0 = O.solve(K0, P0, ne, config=O.SolverConfig(**cfg))
    assert r64.converged and rmx.converged
    assert abs(float(rmx.energy) - float(r64.energy)) < 1e-9 * nk * nk
    assert abs(float(rmx.energy) - r0.energy) < 1e-9 * nk * nk
    assert np.abs(_np(rmx.density) - _np(r64.density)).max() < 1e-5


def test_mixed_solve_mu_completion_paths(dev, monkeypatch):
    """One Newton mu step per captured solve (instead of 6-8) forces the
    completion paths: the CG loop finishes mu on its readback, the pipelined
    polish repairs the density after a speculative update.  The result is the
    same fixed point as with the default step counts."""
    import jax_hf_amd as J
    from jax_hf_amd import solver as SV
    c = syn.config0(nk=24, nb=4, seed=0)
    cfg = J.SolverConfig(max_iter=400, precision="mixed", polish_tol=1e-9)
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    ref = J.solve(G, c["P0"], c["n_electrons"], config=cfg)
    monkeypatch.setattr(SV._CG, "MU_LOOP_STEPS", 1)
    monkeypatch.setattr(SV._DIISPolish, "MU_GRAPH_STEPS", 1)
    G2 = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    r = J.solve(G2, c["P0"], c["n_electrons"], config=cfg)
    assert ref.converged and r.converged
    assert abs(float(r.energy) - float(ref.energy)) < 1e-10
    assert np.abs(_np(r.density) - _np(ref.density)).max() < 1e-7
    # the returned occupations hold the electron count
    w = c["weights"] / c["weights"].sum()
    assert abs(float(np.sum(w[..., None] * _np(r.p))) - c["n_electrons"] / c["weights"].sum()) < 1e-10


@pytest.mark.parametrize("tol", [1e-6, 1e-8])
def test_mixed_solve_meets_reference_residuals(dev, tol):
    """The converged mixed solve meets the reference SCF's own stopping test
    at the returned state (reference_scf.py:287-294): the commutator residual
    |F P - P F| and the density residual |density(F) - P|, both weighted as the
    reference, within polish_tol -- recomputed here on the host from the
    returned density, Fock matrix and orbitals (density(F) = Q diag(p) Q^H,
    the eigen-decomposition of F, which the check also confirms)."""
    import jax_hf_amd as J
    c = syn.config1(nk=32, nb=16, seed=5)
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], 0.1)
    P0 = np.zeros(c["hamiltonian"].shape, dtype=np.complex128)
    r = J.solve(G, P0, 8.0, config=J.SolverConfig(max_iter=400, precision="mixed", polish_tol=tol))
    assert r.converged
    P, F, Q, p = _np(r.density), _np(r.fock), _np(r.Q), _np(r.p)
    w = c["weights"] / c["weights"].sum()
    wn = lambda X: float(np.sqrt(np.sum(w * np.sum(np.abs(X) ** 2, axis=(-2, -1)))))  # noqa: E731
    eps = np.real(np.einsum("...ji,...jk,...ki->...i", np.conj(Q), F, Q))
    assert np.abs(F @ Q - Q * eps[..., None, :]).max() < 1e-9 * np.abs(F).max()
    Pn = np.einsum("...in,...n,...jn->...ij", Q, p, np.conj(Q))
    assert wn(F @ P - P @ F) <= tol
    assert wn(Pn - P) <= tol * (1 + 1e-6)


def test_polish_budget_spent_returns_last_iterate(dev):
    """With a polish budget too small to converge, the mixed solve returns its
    last evaluated iterate: not converged, and energy / Fock matrix those of
    the returned density."""
    import jax_hf_amd as J
    c = syn.config0(nk=24, nb=4, seed=0)
    G = J.HartreeFockKernel(c["weights"], c["hamiltonian"], c["coulomb_q"], c["T"])
    r = J.solve(G, c["P0"], c["n_electrons"], config=J.SolverConfig(max_iter=400, precision="mixed",
                                                                       polish_tol=1e-13, polish_max_iter=2))
    assert not r.converged
    a = G.as_args()
    S, H, F = J.build_fock(r.density, h=a["h"], VR=a["VR"], refP=a["refP"], HH=a["HH"], w2d=G.w2d,
                           include_exchange=a["include_exchange"], include_hartree=a["include_hartree"],
                           exchange_hermitian_channel_packing=a["exchange_hermitian_channel_packing"],
                           contact_g=a["contact_g"], contact_Oi=a["contact_Oi"], contact_Oj=a["contact_Oj"])
    assert np.abs(_np(F) - _np(r.fock)).max() < 1e-12
    E = float(J.hf_energy(r.density, h=G.h, Sigma=S, H=H, weights_b=G.weights_b))
    assert abs(E - float(r.energy)) < 1e-12


@pytest.mark.parametrize("nb", [1, 2, 3, 4, 5, 8, 13, 16, 24, 32])
@pytest.mark.parametrize("dt", ["c128", "c64"])
def test_qr_retract_matches_oracle(dev, nb, dt):
    """_qr_retract (solver.py:108-119) on the GPU against the oracle's
    numpy.linalg.qr with the R-diagonal phase fix: general and skew-Hermitian
    G, per-k tau including 0; unitary to rounding."""
    import torch
    from jax_hf_amd.solver import _qr_retract
    rng = np.random.default_rng(nb)
    G = rng.standard_normal((5, 7, nb, nb)) + 1j * rng.standard_normal((5, 7, nb, nb))
    tdt, tol = (torch.complex128, 1e-12) if dt == "c128" else (torch.complex64, 3e-5)
    for Gx in (0.3 * G, 0.5 * (G - np.conj(np.swapaxes(G, -1, -2)))):
        if dt == "c64":
            Gx = Gx.astype(np.complex64).astype(np.complex128)
        tau = rng.uniform(0.0, 1.2, (5, 7))
        tau[0, 0] = 0.0
        U = _np(_qr_retract(torch.from_numpy(Gx).to(dev).to(tdt), tau)).astype(np.complex128)
        U0 = O.qr_retract(Gx, tau)
        assert np.abs(U - U0).max() < tol * (1 + np.abs(tau).max() * np.abs(Gx).max())
        np.testing.assert_allclose(np.conj(np.swapaxes(U, -1, -2)) @ U,
                                   np.broadcast_to(np.eye(nb), U.shape), atol=tol)
        np.testing.assert_allclose(U[0, 0], np.eye(nb), atol=tol)
    # a scalar tau broadcasts over k
    Us = _np(_qr_retract(torch.from_numpy(0.2 * G).to(dev).to(tdt), 0.7)).astype(np.complex128)
    assert np.abs(Us - O.qr_retract(0.2 * G, np.full((5, 7), 0.7))).max() < 10 * tol
