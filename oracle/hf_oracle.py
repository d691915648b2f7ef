"""NumPy/SciPy float64 restatement of the jax_hf Fock-build and solver path.

TEST INFRASTRUCTURE ONLY (imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``; never by the product path).

Every function cites the reference lines (``src/jax_hf/...``) it follows.
All arithmetic is complex128 / float64 regardless of the input dtype; the
dtype-dependent iteration counts of the reference (bisection 30/54 steps) are
selected by the ``single`` flag where the reference keys them on dtype.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable, NamedTuple

import numpy as np
import scipy.fft as sfft
from scipy.special import expit

__all__ = [
    "OracleKernel", "selfenergy_fft", "build_fock", "hf_energy",
    "occupation_entropy", "free_energy", "fermidirac", "solve_mu",
    "find_chemical_potential", "validate_electron_count",
    "density_matrix_from_fock", "cayley_retract", "cayley_spectral_setup",
    "cayley_unitary_from_spectrum", "diag_UFU_from_spectrum", "qr_retract",
    "SolverConfig", "SolveResult", "solve_cg", "solve_newton", "solve",
    "SCFConfig", "SCFResult", "solve_scf", "diis_coefficients",
    "direct_exchange_sum", "deflation_bias_terms", "FFT_WORKERS",
]

# scipy.fft worker count (the CPU baseline sets this to the host core count).
FFT_WORKERS = 1


def _fftn(x):
    return sfft.fftn(x, axes=(0, 1), workers=FFT_WORKERS)


def _ifftn(x):
    return sfft.ifftn(x, axes=(0, 1), workers=FFT_WORKERS)


def _herm(X):
    """0.5 (X + X^dagger) on the last two axes (fock.py:13-15, solver.py:34-35)."""
    return 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))


def _skew(A):
    """0.5 (A - A^dagger) (solver.py:38-39)."""
    return 0.5 * (A - np.conj(np.swapaxes(A, -1, -2)))


def _dag(X):
    return np.conj(np.swapaxes(X, -1, -2))


# ---------------------------------------------------------------------------
# Problem definition (problem.py:22-217)
# ---------------------------------------------------------------------------

class OracleKernel:
    """Restates ``HartreeFockKernel.__init__`` / ``as_args`` (problem.py:22-217)."""

    def __init__(self, weights, hamiltonian, coulomb_q, T, include_hartree=False,
                 include_exchange=True, reference_density=None, hartree_matrix=None,
                 contact_terms=None, embed_reference_potential=False,
                 center_embedded_hartree=True):
        h = np.asarray(hamiltonian).astype(np.complex128)            # :71
        self.h = h
        w2d = np.real(np.asarray(weights)).astype(np.float64)       # :73
        self.w2d = w2d
        self.weights_b = w2d[..., None, None]                        # :74
        self.weight_sum = float(np.sum(w2d))                         # :75
        Vq = np.asarray(coulomb_q)                                   # :78-84
        if np.iscomplexobj(Vq):
            if float(np.max(np.abs(np.imag(Vq)))) <= 1e-8:
                Vq = np.real(Vq)
        Vq = Vq.astype(np.complex128) if np.iscomplexobj(Vq) else Vq.astype(np.float64)
        self.exchange_hermitian_channel_packing = bool(              # :85-87
            Vq.shape[-2:] == (1, 1) and not np.iscomplexobj(Vq))
        self.VR = _fftn(self.weights_b * Vq.astype(np.complex128))   # :88
        n0, n1 = h.shape[0], h.shape[1]                              # :90-98
        s0, s1 = n0 // 2, n1 // 2
        ph0 = np.exp(2j * np.pi * np.arange(n0) * s0 / n0)
        ph1 = np.exp(2j * np.pi * np.arange(n1) * s1 / n1)
        self._VR_shifted = self.VR * (ph0[:, None] * ph1[None, :])[..., None, None]
        self.T = float(T)
        self.include_hartree = bool(include_hartree)
        self.include_exchange = bool(include_exchange)
        if not self.include_hartree and not self.include_exchange:   # :102-103
            raise ValueError("HartreeFockKernel must include at least one of Hartree or exchange.")
        if reference_density is not None:                            # :105-111
            ref = np.asarray(reference_density).astype(np.complex128)
            if ref.shape != h.shape:
                raise ValueError(f"reference_density must have shape {h.shape}, got {ref.shape}")
            self.refP = ref
        else:
            self.refP = np.zeros_like(h)
        if self.include_hartree:                                     # :113-123
            if hartree_matrix is None:
                raise ValueError("include_hartree=True requires hartree_matrix to be provided")
            if reference_density is None:
                raise ValueError("include_hartree=True requires reference_density to be provided")
            HH = np.real(np.asarray(hartree_matrix)).astype(np.float64)
            if HH.shape != h.shape[-2:]:
                raise ValueError(f"hartree_matrix must have shape {h.shape[-2:]}, got {HH.shape}")
            self.HH = HH
        else:
            self.HH = np.zeros(h.shape[-2:], dtype=np.float64)
        nb = h.shape[-1]                                             # :128-158
        if contact_terms is None or len(contact_terms) == 0:
            self.contact_g = np.zeros((1,), np.float64)
            self.contact_Oi = np.zeros((1, nb, nb), np.complex128)
            self.contact_Oj = np.zeros((1, nb, nb), np.complex128)
        else:
            gs, Ois, Ojs = [], [], []
            for k, term in enumerate(contact_terms):
                if len(term) != 3:
                    raise ValueError(f"contact_terms[{k}] must be (g, O_i, O_j); got len={len(term)}")
                g, Oi, Oj = term
                Oi = np.asarray(Oi).astype(np.complex128)
                Oj = np.asarray(Oj).astype(np.complex128)
                if Oi.shape != (nb, nb) or Oj.shape != (nb, nb):
                    raise ValueError(f"contact_terms[{k}] operators must have shape {(nb, nb)}")
                gs.append(float(np.real(g)))
                Ois.append(Oi)
                Ojs.append(Oj)
            self.contact_g = np.asarray(gs, np.float64)
            self.contact_Oi = np.stack(Ois)
            self.contact_Oj = np.stack(Ojs)
        self.embed_reference_potential = bool(embed_reference_potential)  # :161-199
        self.center_embedded_hartree = bool(center_embedded_hartree)
        self.embedded_energy_offset = 0.0
        if self.embed_reference_potential:
            if reference_density is None:
                raise ValueError("embed_reference_potential=True requires reference_density.")
            ref = self.refP
            cJ = 0.0
            J_diag = np.zeros(nb, np.float64)
            if self.include_hartree:
                diag_R = np.einsum("ij,ijaa->a", w2d, ref).real
                J_diag = self.HH @ diag_R
                if self.center_embedded_hartree:
                    J_diag = J_diag - np.mean(J_diag)
                cJ = float(np.sum(J_diag * diag_R))
            cS = 0.0
            S_ref = None
            if self.include_exchange:
                S_ref = selfenergy_fft(self._VR_shifted, ref, apply_ifftshift=False,
                                       hermitian_channel_packing=self.exchange_hermitian_channel_packing)
                cS = float(np.einsum("ij,ijab,ijba->", w2d, S_ref, ref).real)
            h_eff = self.h
            if S_ref is not None:
                h_eff = h_eff + S_ref
            h_eff = h_eff + np.diag(J_diag.astype(np.complex128))[None, None]
            self.h = h_eff
            self.embedded_energy_offset = 0.5 * (cJ + cS)

    def as_args(self):
        """problem.py:201-217."""
        return dict(h=self.h, weights_b=self.weights_b, weight_sum=self.weight_sum,
                    VR=self._VR_shifted, T=self.T, refP=self.refP, HH=self.HH,
                    include_hartree=self.include_hartree,
                    include_exchange=self.include_exchange,
                    exchange_hermitian_channel_packing=self.exchange_hermitian_channel_packing,
                    contact_g=self.contact_g, contact_Oi=self.contact_Oi,
                    contact_Oj=self.contact_Oj)


# ---------------------------------------------------------------------------
# Exchange self-energy (utils.py:66-311)
# ---------------------------------------------------------------------------

def _selfenergy_full(VR, P, hermitian_channel_packing=False):
    """utils.py:140-173: -ifftn(fftn(P) VR) over axes (0, 1), optionally on triu channels."""
    if hermitian_channel_packing:
        n = P.shape[-1]
        ti, tj = np.triu_indices(n)
        packed = P[..., ti, tj]
        sp = -_ifftn(_fftn(packed) * VR[..., 0, 0][..., None])
        out = np.zeros_like(P, dtype=np.complex128)
        out[..., ti, tj] = sp
        off = ti != tj
        out[..., tj[off], ti[off]] = np.conj(sp[..., off])
        return out
    return -_ifftn(_fftn(P) * VR)


def _normalize_block_specs(block_specs):
    """Subset of linalg.py:19-90: canonical ('sizes', ...) / ('indices', ...) tuples."""
    if block_specs is None:
        return None
    items = (block_specs,) if isinstance(block_specs, dict) else block_specs
    out = []
    for spec in items:
        if isinstance(spec, tuple) and len(spec) == 2 and isinstance(spec[0], str):
            kind = spec[0].strip().lower()
            if kind == "sizes":
                out.append(("sizes", tuple(int(x) for x in spec[1])))
            else:
                out.append(("indices", tuple(tuple(int(i) for i in b) for b in spec[1])))
        elif isinstance(spec, dict):
            if "block_sizes" in spec:
                out.append(("sizes", tuple(int(x) for x in spec["block_sizes"])))
            else:
                out.append(("indices", tuple(tuple(int(i) for i in b) for b in spec["block_indices"])))
        else:
            seq = list(spec)
            if all(isinstance(x, (int, np.integer)) for x in seq):
                out.append(("sizes", tuple(int(x) for x in seq)))
            else:
                out.append(("indices", tuple(tuple(int(i) for i in b) for b in seq)))
    return tuple(out)


def _block_index_sets(kind, data, n):
    if kind == "sizes":
        if sum(data) != n:
            raise ValueError(f"block_sizes must sum to {n}.")
        sets, s = [], 0
        for size in data:
            sets.append(tuple(range(s, s + size)))
            s += size
        return sets
    return [tuple(b) for b in data]


def selfenergy_fft(VR, P, *, block_specs=None, check_offdiag=None, offdiag_atol=1e-12,
                   offdiag_rtol=0.0, apply_ifftshift=True, hermitian_channel_packing=False):
    """utils.py:66-137 (with the block-spec chain of utils.py:252-311)."""
    VR = np.asarray(VR, np.complex128)
    P = np.asarray(P, np.complex128)
    if hermitian_channel_packing and VR.shape[-2:] != (1, 1):
        raise ValueError("hermitian_channel_packing requires a scalar interaction kernel "
                         "with shape (..., 1, 1).")
    specs = _normalize_block_specs(block_specs)
    if not specs:
        res = _selfenergy_full(VR, P, hermitian_channel_packing)
    else:
        n = P.shape[-1]
        check = bool(check_offdiag) if check_offdiag is not None else True
        absP = np.abs(P)
        tol = offdiag_atol + offdiag_rtol * float(np.max(absP)) if check else 0.0
        res = None
        for kind, data in specs:          # first compatible spec wins (utils.py:277-311)
            sets = _block_index_sets(kind, data, n)
            mask = np.ones((n, n), bool)
            for s in sets:
                mask[np.ix_(s, s)] = False
            ok = (float(np.max(absP * mask)) <= tol) if check else True
            if ok:
                res = np.zeros_like(P)
                for s in sets:
                    idx = np.asarray(s)
                    sub = P[..., idx[:, None], idx[None, :]]
                    Vs = VR if VR.shape[-2:] == (1, 1) else VR[..., idx[:, None], idx[None, :]]
                    res[..., idx[:, None], idx[None, :]] = _selfenergy_full(Vs, sub, hermitian_channel_packing)
                break
        if res is None:
            res = _selfenergy_full(VR, P, hermitian_channel_packing)
    if apply_ifftshift:
        res = np.fft.ifftshift(res, axes=(0, 1))
    return res


def direct_exchange_sum(weights, coulomb_q, P):
    """Brute-force O(Nk^2) statement of the centred-mesh exchange convolution.

    Sigma(k) = -sum_{k'} w[q] V[q] P(k') with q index = (k - k' + n//2) mod n on
    each axis (the weight attaches to the q grid index, problem.py:88).  Used to
    pin the FFT + phase conventions of ``OracleKernel``/``selfenergy_fft``.
    """
    P = np.asarray(P, np.complex128)
    n0, n1 = P.shape[:2]
    wV = np.asarray(weights, np.float64)[..., None, None] * np.asarray(coulomb_q)
    out = np.zeros_like(P)
    for a in range(n0):
        for b in range(n1):
            acc = 0.0
            for c in range(n0):
                for d in range(n1):
                    q0 = (a - c + n0 // 2) % n0
                    q1 = (b - d + n1 // 2) % n1
                    acc = acc + wV[q0, q1] * P[c, d]
            out[a, b] = -acc
    return out


# ---------------------------------------------------------------------------
# Fock matrix and energies (fock.py:18-129)
# ---------------------------------------------------------------------------

def build_fock(P, *, h, VR, refP, HH, w2d, include_exchange, include_hartree,
               exchange_hermitian_channel_packing, contact_g=None, contact_Oi=None,
               contact_Oj=None, exchange_block_specs=None, exchange_check_offdiag=None,
               exchange_offdiag_atol=1e-12, exchange_offdiag_rtol=0.0, project_fn=None):
    """fock.py:18-97 -> (Sigma, H, F)."""
    P = np.asarray(P, np.complex128)
    h = np.asarray(h, np.complex128)
    dP = P - refP
    if include_exchange:                                              # :51-64
        Sigma = selfenergy_fft(VR, dP, block_specs=exchange_block_specs,
                               check_offdiag=exchange_check_offdiag,
                               offdiag_atol=exchange_offdiag_atol,
                               offdiag_rtol=exchange_offdiag_rtol, apply_ifftshift=False,
                               hermitian_channel_packing=exchange_hermitian_channel_packing)
    else:
        Sigma = np.zeros_like(h)
    if include_hartree:                                               # :66-74
        diag_real = np.real(np.diagonal(dP, axis1=-2, axis2=-1))
        n_vec = np.sum(np.asarray(w2d)[..., None] * diag_real, axis=(0, 1))
        H = np.diag(np.asarray(HH) @ n_vec).astype(np.complex128)[None, None]
    else:
        H = np.zeros_like(h)
    if contact_g is not None and contact_Oi is not None and contact_Oj is not None:  # :76-92
        rho = np.einsum("ij,ijab->ab", w2d, dP)
        tr = np.einsum("tij,ji->t", contact_Oj, rho)
        s_h = np.einsum("t,tij->ij", contact_g * tr, contact_Oi)
        s_f = np.einsum("t,tij->ij", -np.asarray(contact_g),
                        np.einsum("tik,tkl->til", np.einsum("tij,jk->tik", contact_Oi, rho), contact_Oj))
        Sigma = Sigma + (s_h + s_f)[None, None]
    F = _herm(h + Sigma + H)                                          # :94
    if project_fn is not None:
        F = _herm(np.asarray(project_fn(F), np.complex128))
    return Sigma, H, F


def hf_energy(P, *, h, Sigma, H, weights_b):
    """fock.py:100-111: sum_k w_k Re Tr[(h + (Sigma+H)/2) P]."""
    A = np.asarray(weights_b) * (h + 0.5 * (Sigma + H))
    return float(np.sum(np.real(np.einsum("...ij,...ji->...", A, P))))


def occupation_entropy(p, w_norm):
    """fock.py:114-118 (p clipped to [1e-14, 1-1e-14])."""
    ps = np.clip(p, 1e-14, 1.0 - 1e-14)
    s = ps * np.log(ps) + (1.0 - ps) * np.log1p(-ps)
    return float(-np.sum(np.asarray(w_norm)[..., None] * s))


def free_energy(E, p, w_norm, T):
    """fock.py:121-129: Omega = E - max(T, 1e-14) S."""
    return float(E) - max(float(T), 1e-14) * occupation_entropy(p, w_norm)


# ---------------------------------------------------------------------------
# Occupations and chemical potential (utils.py:21-53, 314-474; solver.py:62-101)
# ---------------------------------------------------------------------------

def validate_electron_count(weights, nbands, n_electrons, *, context="n_electrons"):
    """utils.py:21-43."""
    total = float(np.real(np.asarray(weights)).sum())
    target = float(np.asarray(n_electrons))
    mx = total * int(nbands)
    tol = max(1e-8, 1e-6 * max(1.0, abs(mx)))
    if not np.isfinite(target):
        raise ValueError(f"{context} must be finite, got {target!r}.")
    if target < -tol or target > mx + tol:
        raise ValueError(f"{context}={target} is outside the physically reachable range "
                         f"[0, {mx}] for {int(nbands)} bands and weights.sum()={total}.")


def fermidirac(x, T):
    """utils.py:51-53: expit(-x / (T + 1e-12))."""
    return expit(-np.asarray(x) / (T + 1e-12))


def solve_mu(eps, w_norm, n_target_norm, mu0, T, maxiter=25):
    """solver.py:62-101: Newton with a bracket, fixed ``maxiter`` steps."""
    Tj = max(float(T), 1e-12)
    lo = float(np.min(eps)) - 50.0 * Tj
    hi = float(np.max(eps)) + 50.0 * Tj
    mu = min(max(float(mu0), lo), hi)
    wb = np.asarray(w_norm)[..., None]
    for _ in range(int(maxiter)):
        p = expit((mu - eps) / Tj)
        N = float(np.sum(wb * p))
        Z = float(np.sum(wb * p * (1.0 - p) / Tj))
        g = N - n_target_norm
        if g < 0:
            lo = mu
        if g > 0:
            hi = mu
        mu_new = mu - g / max(Z, 1e-18)
        mu_bis = 0.5 * (lo + hi)
        if mu_new <= lo or mu_new >= hi:
            mu_new = mu_bis
        mu_new = min(max(mu_new, lo), hi)
        if not np.isfinite(mu_new):
            mu_new = mu_bis
        mu = mu_new
    return mu


def _find_mu_bisection(bands, weights, n_electrons, T, maxiter=None, single=False):
    """utils.py:353-382 (30 steps in float32, 54 in float64)."""
    if maxiter is None:
        maxiter = 30 if single else 54
    bmin, bmax = float(np.min(bands)), float(np.max(bands))
    span = bmax - bmin + 10.0 * max(float(T), 1e-6)
    lo, hi = bmin - span, bmax + span
    wb = np.asarray(weights)[..., None]
    for _ in range(int(maxiter)):
        mid = 0.5 * (lo + hi)
        count = float(np.sum(wb * fermidirac(bands - mid, T)))
        if count > n_electrons:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def _find_mu_newton(bands, weights, n_electrons, T, maxiter=None, single=False):
    """utils.py:385-442."""
    if maxiter is None:
        maxiter = 15 if single else 25
    Tj = max(float(T), 1e-12)
    bmin, bmax = float(np.min(bands)), float(np.max(bands))
    span = bmax - bmin + 10.0 * max(Tj, 1e-6)
    lo, hi = bmin - span, bmax + span
    mu = 0.5 * (lo + hi)
    wb = np.asarray(weights)[..., None]

    def cs(m):
        p = expit((m - bands) / Tj)
        return float(np.sum(wb * p)), float(np.sum(wb * p * (1 - p) / Tj))

    for _ in range(int(maxiter)):
        N, Z = cs(mu)
        g = N - n_electrons
        if g < 0:
            lo = mu
        if g > 0:
            hi = mu
        mu_new = mu - g / max(Z, 1e-18)
        mu_bis = 0.5 * (lo + hi)
        if mu_new <= lo or mu_new >= hi:
            mu_new = mu_bis
        mu_new = min(max(mu_new, lo), hi)
        if not np.isfinite(mu_new):
            mu_new = mu_bis
        mu = mu_new
    N, _ = cs(mu)
    return 0.5 * (lo + hi) if abs(N - n_electrons) > 1e-12 else mu


def find_chemical_potential(bands, weights, n_electrons, T, *, maxiter=None,
                            method="bisection", single=False):
    """utils.py:314-350."""
    bands = np.asarray(bands, np.float64)
    validate_electron_count(weights, bands.shape[-1], n_electrons)
    method = str(method).lower()
    if method == "bisection":
        return _find_mu_bisection(bands, weights, float(n_electrons), T, maxiter, single)
    if method == "newton":
        return _find_mu_newton(bands, weights, float(n_electrons), T, maxiter, single)
    raise ValueError(f"method must be 'bisection' or 'newton', got {method!r}")


def density_matrix_from_fock(F, weights, n_electrons, T, single=False):
    """utils.py:445-474 (full-eigh branch)."""
    F = _herm(np.asarray(F, np.complex128))
    validate_electron_count(weights, F.shape[-1], n_electrons)
    eps, U = np.linalg.eigh(F)
    mu = find_chemical_potential(eps, weights, n_electrons, T, single=single)
    occ = fermidirac(eps - mu, T)
    P = np.einsum("...in,...n,...jn->...ij", U, occ, np.conj(U))
    return _herm(P), mu


# ---------------------------------------------------------------------------
# Per-k helpers and retractions (solver.py:34-233)
# ---------------------------------------------------------------------------

def _density_from_Qp(Q, p):
    """solver.py:42-44: P = Q diag(p) Q^dagger."""
    return np.einsum("...in,...n,...jn->...ij", Q, p, np.conj(Q))


def _fock_in_orbital_basis(Q, F):
    """solver.py:47-49: Ft = Q^dagger F Q."""
    return _dag(Q) @ F @ Q


def _logit(p):
    """solver.py:52-55 (clip 1e-6)."""
    ps = np.clip(p, 1e-6, 1.0 - 1e-6)
    return np.log(ps) - np.log1p(-ps)


def qr_retract(G, tau):
    """solver.py:108-119: U = QR(I - tau G) with R-diagonal phase normalisation."""
    n = G.shape[-1]
    U, R = np.linalg.qr(np.eye(n) - np.asarray(tau)[..., None, None] * G)
    ph = np.diagonal(R, axis1=-2, axis2=-1)
    pn = np.where(np.abs(ph) > 1e-30, ph / np.where(np.abs(ph) > 0, np.abs(ph), 1.0), 1.0)
    return U * pn[..., None, :]


def cayley_retract(d, tau):
    """solver.py:122-138: U = (I + tau d/2)^{-1} (I - tau d/2)."""
    n = d.shape[-1]
    A = 0.5 * np.asarray(tau)[..., None, None] * d
    eye = np.eye(n)
    return np.linalg.solve(eye + A, np.broadcast_to(eye - A, A.shape))


def cayley_spectral_setup(d):
    """solver.py:161-166: eigh(herm(i d)) -> (V, lam)."""
    iA = _herm(1j * d)
    lam, V = np.linalg.eigh(iA)
    return V, lam


def _cayley_factor(lam, tau):
    """solver.py:169-175."""
    a = tau * lam * 0.5
    return (1.0 + 1j * a) / (1.0 - 1j * a)


def cayley_unitary_from_spectrum(V, lam, tau):
    """solver.py:178-182: U = V diag(c) V^dagger."""
    c = _cayley_factor(lam, tau)
    return (V * c[..., None, :]) @ _dag(V)


def diag_UFU_from_spectrum(V, Ft_eig, lam, tau):
    """solver.py:185-197: diag(U^dagger Ft U) from Ft_eig = V^dagger Ft V."""
    c = _cayley_factor(lam, tau)
    M = np.conj(c)[..., :, None] * Ft_eig * c[..., None, :]
    A = V @ M
    return np.real(np.sum(A * np.conj(V), axis=-1))


def _ip_matrix(X, Y, w):
    """solver.py:205-208."""
    return float(np.sum(w * np.sum(np.real(np.conj(X) * Y), axis=(-2, -1))))


def _ip_vec(x, y, w):
    """solver.py:211-213."""
    return float(np.sum(np.asarray(w)[..., None] * x * y))


def _norm_matrix(X, w):
    return float(np.sqrt(max(0.0, _ip_matrix(X, X, w))))


def _project_occ(dp, w2d):
    """solver.py:224-233."""
    nb = dp.shape[-1]
    return dp - float(np.sum(w2d[..., None] * dp)) / (float(np.sum(w2d)) * nb)


# ---------------------------------------------------------------------------
# Solver configs / results (solver.py:240-298, reference_scf.py:31-75)
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class SolverConfig:
    """solver.py:240-286 (same fields and defaults)."""
    max_iter: int = 200
    tol_E: float = 1e-7
    tol_grad: float = 0.0
    denom_scale: float = 1e-3
    max_step: float = 0.6
    cg_restart: int = 10
    bt_shrink: float = 0.5
    bt_max: int = 8
    mu_maxiter: int = 25
    plateau_window: int = 5
    optimizer: str = "cg"
    tr_delta0: float = 0.5
    tr_cg_max: int = 20
    deflation_targets: Any = None
    deflation_sigma: float = 0.0
    deflation_length: float = 1.0
    block_sizes: tuple | None = None
    project_fn: Any = None


class SolveResult(NamedTuple):
    """solver.py:289-298."""
    Q: np.ndarray
    p: np.ndarray
    mu: float
    density: np.ndarray
    fock: np.ndarray
    energy: float
    n_iter: int
    converged: bool
    history: dict


def _block_specs(block_sizes):
    """solver.py:301-304."""
    if block_sizes is None:
        return None
    return (("sizes", tuple(int(s) for s in block_sizes)),)


def _fock_caller(kernel, fock_build_fn, block_sizes):
    a = kernel.as_args()
    fn = fock_build_fn if fock_build_fn is not None else build_fock
    specs = _block_specs(block_sizes)

    def call(P, refP=None, project_fn=None):
        return fn(P, h=a["h"], VR=a["VR"], refP=a["refP"] if refP is None else refP,
                  HH=a["HH"], w2d=kernel.w2d, include_exchange=a["include_exchange"],
                  include_hartree=a["include_hartree"],
                  exchange_hermitian_channel_packing=a["exchange_hermitian_channel_packing"],
                  contact_g=a["contact_g"], contact_Oi=a["contact_Oi"],
                  contact_Oj=a["contact_Oj"], exchange_block_specs=specs,
                  project_fn=project_fn)
    return call


def solve_cg(kernel, P0, n_electrons, config=None, fock_build_fn=None, trace=None):
    """solver.py:307-660 (_solve_impl): preconditioned Riemannian CG.

    ``trace`` (optional list) receives per-iteration dicts for trajectory tests.
    """
    cfg = config or SolverConfig()
    a = kernel.as_args()
    h, weights_b, VR = a["h"], a["weights_b"], a["VR"]
    tiny = 1e-30
    project = cfg.project_fn if cfg.project_fn is not None else (lambda A: A)
    fock = _fock_caller(kernel, fock_build_fn, cfg.block_sizes)
    w2d = np.real(weights_b[..., 0, 0])                                 # :360-365
    ws = float(a["weight_sum"])
    w_norm = w2d / max(ws, tiny)
    n_norm = float(n_electrons) / max(ws, tiny)
    T_r = max(float(a["T"]), 1e-12)
    nb = h.shape[-1]
    offdiag = (1.0 - np.eye(nb))[None, None]

    P0h = _herm(np.asarray(P0, np.complex128))                         # :380-394
    _, _, F0 = fock(P0h, project_fn=cfg.project_fn)
    eps, Q = np.linalg.eigh(F0)
    mu = solve_mu(eps, w_norm, n_norm, 0.0, T_r, cfg.mu_maxiter)
    p = expit((mu - eps) / T_r)

    hist_E = np.zeros(cfg.max_iter)
    hist_G = np.zeros(cfg.max_iter)
    k = 0
    d_Q = np.zeros_like(Q)
    d_p = np.zeros_like(p)
    G_prev = np.zeros_like(Q)
    gp_prev = np.zeros_like(p)
    grad_norm = np.inf
    E_prev = np.inf
    dE = np.inf

    def cond():                                                        # :422-449
        if cfg.plateau_window > 0:
            e_now = hist_E[max(k - 1, 0)]
            e_past = hist_E[max(k - 1 - cfg.plateau_window, 0)]
            enc = (e_past - e_now) > cfg.tol_E if k > cfg.plateau_window else dE > cfg.tol_E
        else:
            enc = dE > cfg.tol_E
        nc = grad_norm > cfg.tol_grad if cfg.tol_grad > 0.0 else enc
        return k < cfg.max_iter and nc

    while cond():                                                      # :451-600
        P_cur = _herm(project(_density_from_Qp(Q, p)))
        Sig, Hh, F = fock(P_cur)
        E = hf_energy(P_cur, h=h, Sigma=Sig, H=Hh, weights_b=weights_b)
        Ft = _fock_in_orbital_basis(Q, F)
        eps = np.real(np.diagonal(Ft, axis1=-2, axis2=-1))
        diff_p = p[..., None, :] - p[..., :, None]
        G_Q = _skew(diff_p * Ft) * offdiag
        g_p = eps + T_r * _logit(p) - mu
        grad_norm = _norm_matrix(G_Q, w_norm)
        gap = eps[..., :, None] - eps[..., None, :]
        eps_scale = np.sqrt(np.mean(eps ** 2) + tiny)
        lam = max(T_r, cfg.denom_scale * eps_scale)
        H_Q = G_Q / np.sqrt(gap ** 2 + lam ** 2)
        ps = np.clip(p, 1e-8, 1 - 1e-8)
        h_p = g_p * (ps * (1 - ps) / T_r)
        pr_num = (_ip_matrix(G_Q, H_Q, w_norm) - _ip_matrix(G_prev, H_Q, w_norm)
                  + _ip_vec(g_p, h_p, w_norm) - _ip_vec(gp_prev, h_p, w_norm))
        pr_den = _ip_matrix(G_prev, G_prev, w_norm) + _ip_vec(gp_prev, gp_prev, w_norm)
        beta = max(0.0, pr_num / pr_den) if pr_den > tiny else 0.0
        beta = min(beta, 5.0)
        if k == 0 or k % cfg.cg_restart == 0:
            beta = 0.0
        d_Q = H_Q + beta * d_Q
        d_p = h_p + beta * d_p
        V_d, lam_d = cayley_spectral_setup(d_Q)
        Ft_eig = _dag(V_d) @ Ft @ V_d
        dQn = _norm_matrix(d_Q, w_norm)
        tau0 = min(1.0, cfg.max_step / max(dQn, tiny))

        p0c = np.clip(p, 1e-8, 1 - 1e-8)
        Om0 = free_energy(float(np.sum(w_norm[..., None] * p0c * eps)), p0c, w_norm, T_r)
        tau = tau0
        accepted = False
        trials = []
        for _i in range(cfg.bt_max):                                   # :563-579
            eps_t = diag_UFU_from_spectrum(V_d, Ft_eig, lam_d, tau)
            p_t = np.clip(p - tau * d_p, 1e-8, 1 - 1e-8)
            Om = free_energy(float(np.sum(w_norm[..., None] * p_t * eps_t)), p_t, w_norm, T_r)
            trials.append(Om)
            if Om < Om0:
                accepted = True
                break
            tau = tau * cfg.bt_shrink
        tau_final = tau if accepted else tau0 * cfg.bt_shrink ** cfg.bt_max
        U = cayley_unitary_from_spectrum(V_d, lam_d, tau_final)       # :586-592
        Q = Q @ U
        eps_new = diag_UFU_from_spectrum(V_d, Ft_eig, lam_d, tau_final)
        mu = solve_mu(eps_new, w_norm, n_norm, mu, T_r, cfg.mu_maxiter)
        p = expit((mu - eps_new) / T_r)
        hist_E[k] = E
        hist_G[k] = grad_norm
        dE = abs(E - E_prev)
        E_prev = E
        G_prev = G_Q
        gp_prev = g_p
        if trace is not None:
            trace.append(dict(k=k, E=E, grad_norm=grad_norm, beta=beta, tau0=tau0,
                              tau=tau_final, accepted=accepted, Omega0=Om0, trials=trials,
                              mu=mu, eps_scale=eps_scale))
        k += 1

    P_pre = _herm(project(_density_from_Qp(Q, p)))                     # :606-638
    _, _, F_pre = fock(P_pre)
    eps_f, V_f = np.linalg.eigh(_fock_in_orbital_basis(Q, F_pre))
    Q = Q @ V_f
    mu = solve_mu(eps_f, w_norm, n_norm, mu, T_r, cfg.mu_maxiter)
    p = expit((mu - eps_f) / T_r)
    P_fin = _herm(project(_density_from_Qp(Q, p)))
    Sig, Hh, F_fin = fock(P_fin)
    E_fin = hf_energy(P_fin, h=h, Sigma=Sig, H=Hh, weights_b=weights_b)
    converged = (grad_norm <= cfg.tol_grad) if cfg.tol_grad > 0.0 else (k < cfg.max_iter)
    return SolveResult(Q=Q, p=p, mu=mu, density=P_fin, fock=F_fin, energy=E_fin, n_iter=k,
                       converged=bool(converged), history=dict(E=hist_E, grad_norm=hist_G))


def deflation_bias_terms(P, targets, w2d, sigma, inv2L2):
    """Gaussian deflation bias around targets (solver.py:799-812):
    (Phi = sum_i sigma exp(-d_i^2 inv2L2), S_pen = sum_i 2 phi_i' (P - P*_i),
    diag = 2 sum_i phi_i') with d_i^2 = sum_k w_k ||P_k - P*_{i,k}||_F^2 and
    phi_i' = -phi_i inv2L2."""
    diff = P[None] - targets
    per_k = np.sum(np.abs(diff) ** 2, axis=(-2, -1))
    d2 = np.sum(w2d[None] * per_k, axis=(1, 2))
    phi = sigma * np.exp(-d2 * inv2L2)
    phip = -phi * inv2L2
    S = np.sum((2.0 * phip)[:, None, None, None, None] * diff, axis=0)
    return float(np.sum(phi)), S, float(2.0 * np.sum(phip))


def solve_newton(kernel, P0, n_electrons, config=None, fock_build_fn=None, trace=None):
    """solver.py:686-1078 (_solve_newton_impl): trust-region Newton / Steihaug CG."""
    cfg = config or SolverConfig(optimizer="newton")
    a = kernel.as_args()
    h, weights_b = a["h"], a["weights_b"]
    tiny = 1e-30
    project = cfg.project_fn if cfg.project_fn is not None else (lambda A: A)
    fock = _fock_caller(kernel, fock_build_fn, cfg.block_sizes)
    w2d = np.real(weights_b[..., 0, 0])
    ws = float(a["weight_sum"])
    w_norm = w2d / max(ws, tiny)
    ones_w = np.ones_like(w2d)
    n_norm = float(n_electrons) / max(ws, tiny)
    T_r = max(float(a["T"]), 1e-12)
    tol_g = float(cfg.tol_grad) if cfg.tol_grad > 0.0 else 1e-6        # :1191
    delta_min = 1e-10
    noise_floor = max(1e-12, 4500.0 * np.finfo(np.float64).eps)
    nb = h.shape[-1]
    eye = np.eye(nb)
    offdiag = (1.0 - np.eye(nb))[None, None]
    refP_zero = np.zeros_like(a["refP"])

    has_defl = cfg.deflation_targets is not None and float(cfg.deflation_sigma) > 0.0
    if has_defl:
        targets = np.asarray(cfg.deflation_targets, np.complex128)
        if targets.ndim == 4:
            targets = targets[None]
    inv2L2 = 1.0 / (2.0 * max(float(cfg.deflation_length), 1e-12) ** 2)
    dsig = float(cfg.deflation_sigma)

    def response(dP):                                                  # :770-780
        _, _, Fr = fock(dP, refP=refP_zero)
        return Fr - h

    def jip(aX, ap, bX, bp):                                           # :783-787
        return _ip_matrix(aX, bX, ones_w) + _ip_vec(ap, bp, w2d)

    def jnorm(aX, ap):
        return float(np.sqrt(max(0.0, jip(aX, ap, aX, ap))))

    def deflation_bias(P):
        return deflation_bias_terms(P, targets, w2d, dsig, inv2L2)

    def eval_state(Q, p):                                              # :814-837
        P = _herm(project(_density_from_Qp(Q, p)))
        Sg, Hh, F = fock(P)
        E = hf_energy(P, h=h, Sigma=Sg, H=Hh, weights_b=weights_b)
        if has_defl:
            Phi, S, pen = deflation_bias(P)
            E = E + Phi
            F = F + S
        else:
            pen = 0.0
        Ft = _fock_in_orbital_basis(Q, F)
        diff_p = p[..., None, :] - p[..., :, None]
        G = _skew(diff_p * Ft) * offdiag
        return E, Ft, _norm_matrix(G, w_norm), pen

    P0h = _herm(np.asarray(P0, np.complex128))                         # :840-848
    _, _, F0 = fock(P0h)
    eps0, Q0 = np.linalg.eigh(F0)
    mu = solve_mu(eps0, w_norm, n_norm, 0.0, T_r, cfg.mu_maxiter)
    p = expit((mu - eps0) / T_r)
    E, Ft, grad, pen = eval_state(Q0, p)
    Q = Q0
    if grad <= tol_g:                                                  # :858-877
        for _ in range(30):
            P = _herm(project(_density_from_Qp(Q0, p)))
            _, _, F = fock(P)
            eps_i = np.real(np.diagonal(_fock_in_orbital_basis(Q0, F), axis1=-2, axis2=-1))
            mu = solve_mu(eps_i, w_norm, n_norm, mu, T_r, cfg.mu_maxiter)
            p = expit((mu - eps_i) / T_r)
        E, Ft, grad, pen = eval_state(Q0, p)

    hist_E = np.zeros(cfg.max_iter)
    hist_G = np.zeros(cfg.max_iter)
    k = 0
    Delta = float(cfg.tr_delta0)
    while k < cfg.max_iter and grad > tol_g and Delta > delta_min:     # :886-891
        hist_E[k] = E
        hist_G[k] = grad
        eps = np.real(np.diagonal(Ft, axis1=-2, axis2=-1))
        diff_p = p[..., None, :] - p[..., :, None]
        G = _skew(diff_p * Ft) * offdiag
        eps_scale = np.sqrt(np.mean(eps ** 2) + tiny)
        lam_pc = max(T_r, cfg.denom_scale * eps_scale)
        denomX = np.sqrt((eps[..., :, None] - eps[..., None, :]) ** 2 + lam_pc ** 2)
        pc = np.clip(p, 1e-8, 1 - 1e-8)
        fd_curv = pc * (1 - pc) / T_r

        def prec(aX, ap):
            return aX / denomX, _project_occ(ap * fd_curv, w2d)

        def hvp(X, dp):                                                # :916-930
            M = X * diff_p + dp[..., :, None] * eye
            dP = _herm(Q @ M @ _dag(Q))
            resp = response(dP)
            if has_defl:
                resp = resp + pen * dP
            St = _fock_in_orbital_basis(Q, resp)
            A = Ft @ X - X @ Ft + St
            diff_dp = dp[..., None, :] - dp[..., :, None]
            HX = _skew(diff_dp * Ft + diff_p * A) * offdiag
            A_diag = np.real(np.diagonal(A, axis1=-2, axis2=-1))
            Hp = A_diag + (T_r / (pc * (1 - pc))) * dp
            return HX, _project_occ(Hp, w2d)

        rX, rp = -G, np.zeros_like(p)                                  # :933-991
        dX, dp = prec(rX, rp)
        rz = jip(rX, rp, dX, dp)
        vX, vp = np.zeros_like(Q), np.zeros_like(p)
        HvX, Hvp = np.zeros_like(Q), np.zeros_like(p)
        cg = 0
        done = False
        while cg < cfg.tr_cg_max and not done:
            HdX, Hdp = hvp(dX, dp)
            dHd = jip(dX, dp, HdX, Hdp)
            dd = jip(dX, dp, dX, dp)
            neg = dHd <= 1e-14 * dd
            b = 2.0 * jip(vX, vp, dX, dp)
            cc = jip(vX, vp, vX, vp) - Delta * Delta
            disc = max(b * b - 4.0 * dd * cc, 0.0)
            t_bd = (-b + np.sqrt(disc)) / (2.0 * dd + 1e-30)
            alpha = rz / (1.0 if neg else dHd)
            vXa, vpa = vX + alpha * dX, vp + alpha * dp
            take_bd = neg or jnorm(vXa, vpa) >= Delta
            rXa, rpa = rX - alpha * HdX, rp - alpha * Hdp
            inner = jnorm(rXa, rpa) < 0.05 * grad
            zX, zp = prec(rXa, rpa)
            rz_new = jip(rXa, rpa, zX, zp)
            beta = rz_new / (rz if abs(rz) > tiny else 1.0)
            stop = take_bd or inner
            if take_bd:
                vX, vp = vX + t_bd * dX, vp + t_bd * dp
                HvX, Hvp = HvX + t_bd * HdX, Hvp + t_bd * Hdp
            else:
                vX, vp = vXa, vpa
                HvX, Hvp = HvX + alpha * HdX, Hvp + alpha * Hdp
                rX, rp = rXa, rpa
            if not stop:
                dX, dp = zX + beta * dX, zp + beta * dp
                rz = rz_new
            done = stop
            cg += 1

        gv = _ip_matrix(G, vX, w2d)                                    # :998-1000
        vHv = _ip_matrix(vX, HvX, w2d) + _ip_vec(vp, Hvp, w2d)
        pred = -(gv + 0.5 * vHv)
        V_v, lam_v = cayley_spectral_setup(vX)                         # :1005-1011
        U = cayley_unitary_from_spectrum(V_v, lam_v, -1.0)
        Qt = Q @ U
        eps_t = np.real(np.diagonal(_fock_in_orbital_basis(U, Ft), axis1=-2, axis2=-1))
        mu_t = solve_mu(eps_t, w_norm, n_norm, mu, T_r, cfg.mu_maxiter)
        p_t = expit((mu_t - eps_t) / T_r)
        E_t, Ft_t, g_t, pen_t = eval_state(Qt, p_t)                    # :1014
        noise = noise_floor * max(abs(E), 1.0)                         # :1021-1046
        use_g = pred <= noise
        if use_g:
            accept = (g_t <= grad) and (E_t <= E + noise)
            Delta_new = Delta if accept else 0.5 * Delta
        else:
            ratio = (E - E_t) / pred
            sn = jnorm(vX, vp)
            if ratio < 0.25:
                Delta_new = 0.25 * Delta
            elif ratio > 0.75 and sn > 0.9 * Delta:
                Delta_new = min(2.0 * Delta, 5.0)
            else:
                Delta_new = Delta
            accept = ratio > 0.1
        if trace is not None:
            trace.append(dict(k=k, E=E, grad=grad, pred=pred, E_trial=E_t, accept=accept,
                              Delta=Delta, cg_iters=cg))
        if accept:
            Q, p, Ft, E, mu, grad, pen = Qt, p_t, Ft_t, E_t, mu_t, g_t, pen_t
        Delta = Delta_new
        k += 1

    P_fin = _herm(project(_density_from_Qp(Q, p)))                     # :1062-1066
    Sg, Hh, F_fin = fock(P_fin)
    E_fin = hf_energy(P_fin, h=h, Sigma=Sg, H=Hh, weights_b=weights_b)
    return SolveResult(Q=Q, p=p, mu=mu, density=P_fin, fock=F_fin, energy=E_fin, n_iter=k,
                       converged=bool(grad <= tol_g), history=dict(E=hist_E, grad_norm=hist_G))


def solve(kernel, P0, n_electrons, *, config=None, fock_build_fn=None, trace=None):
    """solver.py:1098-1247 (dispatcher)."""
    cfg = config or SolverConfig()
    validate_electron_count(kernel.w2d, kernel.h.shape[-1], n_electrons)
    opt = str(cfg.optimizer).lower()
    if opt not in ("cg", "newton"):
        raise ValueError(f"Unknown optimizer {cfg.optimizer!r}; expected 'cg' or 'newton'.")
    defl = cfg.deflation_targets is not None and float(cfg.deflation_sigma) > 0.0
    if defl and opt != "newton":
        raise ValueError("deflation is only supported with optimizer='newton' "
                         f"(got optimizer={cfg.optimizer!r}).")
    if opt == "newton":
        return solve_newton(kernel, P0, n_electrons, cfg, fock_build_fn, trace)
    return solve_cg(kernel, P0, n_electrons, cfg, fock_build_fn, trace)


# ---------------------------------------------------------------------------
# Reference SCF (reference_scf.py:31-365)
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class SCFConfig:
    """reference_scf.py:31-63."""
    max_iter: int = 200
    density_tol: float = 1e-7
    comm_tol: float = 1e-6
    mixing: float = 0.5
    level_shift: float = 0.0
    acceleration: str = "linear"
    diis_size: int = 6
    diis_start: int = 2
    diis_damping: float = 1.0
    trust_radius: float = 0.0
    project_fn: Callable | None = None

    def __post_init__(self):
        if self.max_iter <= 0:
            raise ValueError("max_iter must be positive")
        if self.acceleration not in ("linear", "diis", "oda"):
            raise ValueError("SCFConfig.acceleration must be 'linear', 'diis', or 'oda'; "
                             f"got {self.acceleration!r}")
        if self.diis_size <= 0:
            raise ValueError("diis_size must be positive")


@dataclass(frozen=True)
class SCFResult:
    """reference_scf.py:66-75."""
    density_matrix: Any
    fock_matrix: Any
    energy: Any
    chemical_potential: Any
    iterations: int
    converged: bool
    message: str
    history: dict = field(default_factory=dict)


def _wnorm(M, w, ws):
    """reference_scf.py:82-84."""
    return float(np.sqrt(np.sum(w * np.sum(np.abs(M) ** 2, axis=(-2, -1))) / ws))


def _density_from_fock_scf(F, w2d, n, T, level_shift, single=False):
    """reference_scf.py:87-96."""
    ev, U = np.linalg.eigh(_herm(F))
    mu_raw = find_chemical_potential(ev, w2d, n, T, single=single)
    occ_raw = fermidirac(ev - mu_raw, T)
    sh = ev + level_shift * (1.0 - occ_raw)
    mu = find_chemical_potential(sh, w2d, n, T, single=single)
    occ = fermidirac(sh - mu, T)
    return _herm(np.einsum("...in,...n,...jn->...ij", U, occ, np.conj(U))), mu


def diis_coefficients(R_hist, active, w2d):
    """reference_scf.py:99-121: masked Pulay C-DIIS augmented solve."""
    N = R_hist.shape[0]
    B = np.real(np.einsum("iklab,jklab->ij", np.conj(R_hist), R_hist * w2d[None, :, :, None, None]))
    B = B / max(float(np.max(np.abs(B))), 1e-30)
    af = active.astype(np.float64)
    A = np.zeros((N + 1, N + 1))
    A[:N, :N] = B + np.diag(1.0 - af)
    A[:N, N] = -af
    A[N, :N] = -af
    rhs = np.zeros(N + 1)
    rhs[N] = -1.0
    try:
        sol = np.linalg.solve(A, rhs)
    except np.linalg.LinAlgError:
        sol = np.full(N + 1, np.nan)
    return sol[:N], bool(np.all(np.isfinite(sol)))


def solve_scf(kernel, P0, n_electrons, *, config=None, fock_build_fn=None, single=False):
    """reference_scf.py:124-365 (solve_scf + _run_scf)."""
    cfg = config or SCFConfig()
    a = kernel.as_args()
    h, weights_b = a["h"], a["weights_b"]
    w2d = kernel.w2d
    ws = max(float(a["weight_sum"]), 1e-30)
    T = a["T"]
    n = float(n_electrons)
    project = cfg.project_fn if cfg.project_fn is not None else (lambda A: A)
    fock = _fock_caller(kernel, fock_build_fn, None)

    def build_and_occupy(P):                                           # :264-273
        Ph = _herm(np.asarray(project(P), np.complex128))
        Sg, Hh, F = fock(Ph, project_fn=cfg.project_fn)
        E = hf_energy(Ph, h=h, Sigma=Sg, H=Hh, weights_b=weights_b)
        Pn, mu = _density_from_fock_scf(F, w2d, n, T, cfg.level_shift, single)
        return _herm(np.asarray(project(Pn), np.complex128)), F, E, mu

    def linear(P, delta):
        return _herm(np.asarray(project(P + cfg.mixing * delta), np.complex128))

    nh = cfg.diis_size if cfg.acceleration == "diis" else 0
    F_hist = np.zeros((nh,) + h.shape, np.complex128)
    R_hist = np.zeros((nh,) + h.shape, np.complex128)
    hE = np.zeros(cfg.max_iter)
    hD = np.zeros(cfg.max_iter)
    hC = np.zeros(cfg.max_iter)
    density = np.asarray(P0, np.complex128)
    k = 0
    converged = False
    while k < cfg.max_iter and not converged:                          # :287-347
        P_new, F, E, mu = build_and_occupy(density)
        delta = P_new - density
        comm = F @ density - density @ F
        d_res = _wnorm(delta, w2d, ws)
        c_res = _wnorm(comm, w2d, ws)
        converged = d_res <= cfg.density_tol and c_res <= cfg.comm_tol
        if cfg.acceleration == "diis":
            idx = k % nh
            F_hist[idx] = F
            R_hist[idx] = comm
            count = min(k + 1, nh)
            active = np.arange(nh) < count
            coeffs, ok = diis_coefficients(R_hist, active, w2d)
            if ok:
                F_ex = np.einsum("i,iklab->klab", coeffs.astype(np.complex128), F_hist)
            else:
                F_ex = np.full_like(F, np.nan)
            damp = min(max(cfg.diis_damping, 0.0), 1.0)
            F_ex = damp * F_ex + (1.0 - damp) * F
            engaged = count >= max(2, cfg.diis_start) and ok
            if engaged:
                P_d, _ = _density_from_fock_scf(_herm(F_ex), w2d, n, T, cfg.level_shift, single)
                updated = _herm(np.asarray(project(P_d), np.complex128))
            else:
                updated = linear(density, delta)
        elif cfg.acceleration == "oda":
            Sn, Hn, _ = fock(P_new, project_fn=cfg.project_fn)
            E_new = hf_energy(P_new, h=h, Sigma=Sn, H=Hn, weights_b=weights_b)
            c1 = float(np.sum(w2d * np.real(np.einsum("...ij,...ji->...", F, delta))))
            c2 = (E_new - E) - c1
            if c2 > 1e-12:
                lam = min(max(-c1 / (2.0 * c2), 0.0), 1.0)
            else:
                lam = 1.0 if c1 < 0.0 else 0.0
            updated = _herm(np.asarray(project(density + lam * delta), np.complex128))
        else:
            updated = linear(density, delta)
        step = updated - density                                       # :333-339
        sn = _wnorm(step, w2d, ws)
        if cfg.trust_radius > 0.0 and sn > cfg.trust_radius:
            sc = cfg.trust_radius / max(sn, 1e-30)
            updated = _herm(np.asarray(project(density + sc * step), np.complex128))
        density = P_new if converged else updated
        hE[k] = E
        hD[k] = d_res
        hC[k] = c_res
        k += 1
    _, F_fin, E_fin, mu_fin = build_and_occupy(density)                # :364
    if converged:
        msg = f"converged in {k} iterations"
    else:
        dr = hD[k - 1] if k > 0 else float("nan")
        cr = hC[k - 1] if k > 0 else float("nan")
        msg = f"stopped after {k} iterations (density_res={dr:.3e}, comm_res={cr:.3e})"
    return SCFResult(density_matrix=density, fock_matrix=F_fin, energy=E_fin,
                     chemical_potential=mu_fin, iterations=k, converged=bool(converged),
                     message=msg, history={"E": hE[:k], "density_residual": hD[:k],
                                           "commutator_residual": hC[:k]})
