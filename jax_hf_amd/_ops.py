"""Thin typed wrappers over the per-k / reduction entry points of libhfx."""

from __future__ import annotations

import torch

from . import _lib as L

ST_MU, ST_BETA, ST_LAM, ST_GRAD, ST_TAU0, ST_TAU, ST_TAUF, ST_ACC = 0, 1, 2, 3, 4, 5, 6, 7
ST_OMEGA0, ST_DNORM, ST_OMEGA, ST_E = 8, 9, 10, 11
ST_ITER, ST_MULO, ST_MUHI, ST_MUDONE = 12, 13, 14, 15
ST_FROZEN = 16
ST_LEN = 32
OP_SUM, OP_GRAD, OP_BETA, OP_TAU0, OP_TRIAL = 0, 1, 2, 3, 4
OP_NT0, OP_NT1, OP_NT2 = 5, 6, 7
OP_PARTIALS = 8
MU_NEWTON_BRACKET, MU_BISECTION, MU_NEWTON = 0, 1, 2


def dt_code(t: torch.Tensor) -> int:
    return L.HFX_C128 if t.dtype == torch.complex128 else L.HFX_C64


class Workspace:
    """Device scratch shared by the reductions / mu solver on one device."""

    _cache: dict = {}

    @classmethod
    def get(cls, device, nk: int):
        key = (torch.device(device).index, nk)
        w = cls._cache.get(key)
        if w is None:
            w = cls(device, nk)
            cls._cache[key] = w
        return w

    def __init__(self, device, nk: int):
        n = int(L.lib().hfx_solver_scratch_doubles())
        self.scratch = torch.zeros(n, dtype=torch.float64, device=device)
        self.part = torch.zeros(8 * nk, dtype=torch.float64, device=device)
        self.out = torch.zeros(16, dtype=torch.float64, device=device)


L.register_cache(Workspace._cache.clear)


def s(dev):
    return L.stream(dev)


def eigh(A, lam, V, herm_in=False, rel_tol=0.0):
    """Batched Hermitian eigh; rel_tol > 0 stops the Jacobi rotations at |A_pq| <= rel_tol ||A||_F."""
    nk = A.numel() // (A.shape[-1] * A.shape[-1])
    L.call("hfx_eigh_batched_tol", dt_code(A), A.shape[-1], nk, L.ptr(A), L.ptr(lam), L.ptr(V), int(herm_in),
           float(rel_tol), s(A.device))


def eigh_basis(A, lam, U, herm_in=False, rel_tol=0.0, out=None):
    """Eigendecomposition of A = U^H F U from the basis U: out = U W (default:
    U in place), lam ascending."""
    nk = A.numel() // (A.shape[-1] * A.shape[-1])
    L.call("hfx_eigh_batched_basis", dt_code(A), A.shape[-1], nk, L.ptr(A), L.ptr(lam), L.ptr(U),
           L.ptr(U if out is None else out), int(herm_in), float(rel_tol), s(A.device))


def density_diff(Q, p, Pref, P, part):
    """P = herm(Q diag(p) Q^H) and part[k] = sum |P - Pref|^2 per k."""
    nb = Q.shape[-1]
    L.call("hfx_density_diff", dt_code(Q), nb, Q.numel() // (nb * nb), L.ptr(Q), L.ptr(p), L.ptr(Pref), L.ptr(P),
           L.ptr(part), s(Q.device))


def density(Q, P, p=None, eps=None, mu_ptr=None, T=1.0, pout=None):
    nk = Q.numel() // (Q.shape[-1] ** 2)
    L.call("hfx_density_from_Qp", dt_code(Q), Q.shape[-1], nk, L.ptr(Q), L.ptr(p), L.ptr(eps), mu_ptr,
           float(T), L.ptr(pout), L.ptr(P), s(Q.device))


def matmul(A, B, out, C=None, ops=0, Bsub=None, herm_out=False):
    nk = A.numel() // (A.shape[-1] ** 2)
    L.call("hfx_matmul_batched", dt_code(A), A.shape[-1], nk, L.ptr(A), L.ptr(B), L.ptr(C), int(ops),
           L.ptr(Bsub), int(herm_out), L.ptr(out), s(A.device))


def hermitize(X, Y):
    nb = X.shape[-1]
    L.call("hfx_hermitize", L.ptr(X), L.ptr(Y), X.numel() // (nb * nb), nb, dt_code(X), s(X.device))


def reduce(ws: Workspace, part, nval, nk, w=None, wmask=0, out=None, st=None, op=OP_SUM, params=None):
    par = None if params is None else _params_buf(ws, params)
    L.call("hfx_reduce", L.ptr(part), int(nval), int(nk), L.ptr(w), int(wmask), L.ptr(out), L.ptr(st), int(op),
           None if par is None else par, L.ptr(ws.scratch), s(part.device))


_param_bufs: dict = {}


def _params_buf(ws, params):
    """Host-side parameter vector for hfx_reduce (read by the launcher, not the kernel)."""
    import ctypes as C
    vals = list(params) + [0.0] * (10 - len(params))
    arr = (C.c_double * 10)(*[float(v) for v in vals])
    _param_bufs[id(ws)] = arr          # keep alive for the call
    return C.cast(arr, C.c_void_p)


def solve_mu(ws: Workspace, eps, w, nk, nb, target, T, mu0=0.0, mu0_from_state=False, maxiter=25,
             mode=MU_NEWTON_BRACKET, level_shift=0.0, st=None, mu_out_ptr=None, p_out=None):
    L.call("hfx_solve_mu", L.ptr(eps), L.ptr(w), int(nk), int(nb), float(target), float(T), float(mu0),
           int(mu0_from_state), int(maxiter), int(mode), float(level_shift), L.ptr(st), mu_out_ptr,
           L.ptr(p_out), L.ptr(ws.scratch), s(eps.device))


def solve_mu_steps(ws: Workspace, eps, w, nk, nb, target, T, mu0=0.0, mu0_from_state=False, maxiter=25,
                   st=None, p_out=None):
    L.call("hfx_solve_mu_steps", L.ptr(eps), L.ptr(w), int(nk), int(nb), float(target), float(T), float(mu0),
           int(mu0_from_state), int(maxiter), L.ptr(st), L.ptr(p_out), L.ptr(ws.scratch), s(eps.device))


def st_ptr(st, idx):
    return st.data_ptr() + 8 * idx


def scf_resid(F, D, Pn, comm, part):
    nb = F.shape[-1]
    nk = F.numel() // (nb * nb)
    L.call("hfx_scf_resid", dt_code(F), nb, nk, L.ptr(F), L.ptr(D), L.ptr(Pn), L.ptr(comm), L.ptr(part),
           s(F.device))


def diis_coefficients(brow, B, idx, count, coef):
    """Device C-DIIS: B[idx] row/column <- brow, coef <- masked augmented solve (or e_idx)."""
    L.call("hfx_diis_coefficients", L.ptr(brow), L.ptr(B), int(B.shape[0]), int(idx), int(count), L.ptr(coef),
           s(B.device))


def orthonormalize(U, steps=2):
    """In place: U <- U (3 I - U^H U) / 2, `steps` times, per k."""
    nb = U.shape[-1]
    L.call("hfx_orthonormalize", dt_code(U), nb, U.numel() // (nb * nb), L.ptr(U), int(steps), s(U.device))


def scf_resid_dots(F, D, Pn, comm, Rs, self_idx, part):
    """scf_resid + the DIIS B row: part[2 + v][k] = <comm, Rs[v]> per k (Rs[self_idx] is comm)."""
    import ctypes as C
    nb = F.shape[-1]
    nk = F.numel() // (nb * nb)
    ra = (C.c_void_p * len(Rs))(*[r.data_ptr() for r in Rs])
    L.call("hfx_scf_resid_dots", dt_code(F), nb, nk, L.ptr(F), L.ptr(D), L.ptr(Pn), L.ptr(comm),
           C.cast(ra, C.c_void_p), len(Rs), int(self_idx), dt_code(comm), L.ptr(part), s(F.device))


def dots(X, Ys, part, mode=0):
    """part[v][k] = per-k <X, Y_v> (X may be one matrix or a list paired with Ys)."""
    import ctypes as C
    Xs = X if isinstance(X, (list, tuple)) else [X] * len(Ys)
    nb = Xs[0].shape[-1]
    nk = Xs[0].numel() // (nb * nb)
    xa = (C.c_void_p * len(Ys))(*[x.data_ptr() for x in Xs])
    ya = (C.c_void_p * len(Ys))(*[y.data_ptr() for y in Ys])
    L.call("hfx_batched_dots", dt_code(Xs[0]), nb, nk, C.cast(xa, C.c_void_p), C.cast(ya, C.c_void_p),
           len(Ys), int(mode), L.ptr(part), s(Xs[0].device))


def vdots(xs, ys, part):
    import ctypes as C
    nb = xs[0].shape[-1]
    nk = xs[0].numel() // nb
    xa = (C.c_void_p * len(xs))(*[x.data_ptr() for x in xs])
    ya = (C.c_void_p * len(ys))(*[y.data_ptr() for y in ys])
    L.call("hfx_vec_dots", C.cast(xa, C.c_void_p), C.cast(ya, C.c_void_p), len(xs), nk, nb, L.ptr(part),
           s(xs[0].device))


def lincomb_multi(out, Xs, coefs, herm=False):
    """out = sum_v coefs[v] Xs[v] (real coefficients, <= 8 inputs), Hermitised per matrix with herm."""
    import ctypes as C
    from ._util import KIND
    nb = out.shape[-1]
    xa = (C.c_void_p * len(Xs))(*[x.data_ptr() for x in Xs])
    ca = (C.c_double * len(Xs))(*[float(c) for c in coefs])
    L.call("hfx_lincomb_multi", out.numel() // (nb * nb), nb, KIND[out.dtype], len(Xs), C.cast(xa, C.c_void_p),
           C.cast(ca, C.c_void_p), int(herm), L.ptr(out), s(out.device))


def extrap_rotate(Ft, U, Xs, coefs):
    """Ft = U^H herm(sum_v coefs[v] Xs[v]) U per k (<= 8 inputs)."""
    import ctypes as C
    nb = U.shape[-1]
    xa = (C.c_void_p * len(Xs))(*[x.data_ptr() for x in Xs])
    ca = (C.c_double * len(Xs))(*[float(c) for c in coefs])
    L.call("hfx_extrap_rotate", dt_code(U), nb, U.numel() // (nb * nb), len(Xs), C.cast(xa, C.c_void_p),
           C.cast(ca, C.c_void_p), L.ptr(U), L.ptr(Ft), s(U.device))


def rotate_density(U, W, occ, P):
    """U <- U W per k, P = herm(U diag(occ) U^H) with the new U."""
    nb = U.shape[-1]
    L.call("hfx_rotate_density", dt_code(U), nb, U.numel() // (nb * nb), L.ptr(U), L.ptr(W), L.ptr(occ), L.ptr(P),
           s(U.device))


def lincomb(out, X, a=1.0, Y=None, b=0.0):
    """out = a X + b Y (a, b real or complex)."""
    from ._util import KIND
    a, b = complex(a), complex(b)
    L.call("hfx_lincomb", X.numel(), KIND[X.dtype], a.real, a.imag, L.ptr(X), b.real, b.imag, L.ptr(Y),
           L.ptr(out), s(X.device))
