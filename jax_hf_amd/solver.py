"""Direct-minimisation HF solver on the GPU (drop-in for solver.py).

Mirrors ``src/jax_hf/solver.py``: ``SolverConfig`` (same fields/defaults plus
``precision``), ``SolveResult``, ``solve`` / ``solve_direct_minimization``
with the CG (``_solve_impl``, solver.py:307-660) and trust-region Newton
(``_solve_newton_impl``, solver.py:686-1078) algorithms.

The loop body is a fixed sequence of libhfx launches on one stream; all
scalar steps (gradient norm, preconditioner scale, Polak-Ribiere beta, tau0,
line-search acceptance, chemical potential) run on the device and live in a
16-double state vector, so an iteration needs exactly one host readback (the
energy / gradient pair that the stopping rule of solver.py:422-449 needs).

Precision: ``precision="fp64"`` runs everything in complex128;
``"mixed"`` runs the loop in complex64 storage with FP32 FFT / products and
FP64 reductions, mu and entropy, then re-converges in FP64 from the mixed
iterate (warm start) so the final Fock builds and the reported energy are
FP64.  ``"auto"`` follows the kernel dtype.
"""

from __future__ import annotations

import warnings
import weakref
from dataclasses import dataclass, replace
from typing import Any, NamedTuple

import numpy as np
import torch

from . import _lib as L
from . import _ops as K
from ._util import KIND, capture_graphs, convert, host_array, to_tensor
from .reference_scf import diis_solve

__all__ = ["SolverConfig", "SolveResult", "solve", "solve_direct_minimization"]


@dataclass(frozen=True)
class SolverConfig:
    """Solver controls (solver.py:240-286), plus ``precision``, ``mixed_tol_E`` and
    ``use_graph`` (replay the CG iteration as a captured HIP graph; the launches
    are identical to the eager loop) and ``cayley`` (how the CG line search
    forms U(tau): per trial by Gauss-Jordan, or from one eigh of i d)."""
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
    precision: str = "auto"
    # relative windowed-energy stop of the FP32 phase (x max(1, |E|)): the FP32 CG
    # only has to bring the state into the fixed point's basin, the FP64 DIIS
    # polish sets the accuracy (its residuals are the stopping test).  Measured
    # at 96^2 x 16 (T = 0.1): 3e-4 ends the FP32 phase after 2 iterations with
    # the same 5 polish iterations as 1e-4 after 3 (3.14 vs 3.44 ms per solve);
    # at 64^2 x 16: 2 + 6 vs 3 + 5 iterations, the same time
    mixed_tol_E: float = 3e-4
    use_graph: bool = True
    cayley: str = "gauss_jordan"   # or "spectral" (the reference's eigh-based form)
    polish: str | None = None      # "diis": FP64 DIIS SCF to the fixed point after CG (default for mixed)
    polish_tol: float = 1e-8       # density and commutator residual of the polish
    polish_max_iter: int = 60
    # mixed form: a complex64 DIIS stage between the FP32 CG and the FP64 polish,
    # to residuals <= this (or until its rounding floor stops it), so that the
    # FP64 part is a few refinement builds; 0 skips the stage
    polish32_tol: float = 0.0
    # mixed form: line-search trials of the FP32 CG phase (at most bt_max; when
    # none is accepted the fallback step is the next shorter one, as after bt_max
    # trials).  The captured iteration launches every trial, and the ones after
    # an accepted trial are ~10 us no-ops each (the trial, its two reduction
    # launches); the FP32 phase's first trial is the accepted one as a rule
    mixed_bt_max: int = 2


class SolveResult(NamedTuple):
    """solver.py:289-298."""
    Q: torch.Tensor
    p: torch.Tensor
    mu: torch.Tensor
    density: torch.Tensor
    fock: torch.Tensor
    energy: torch.Tensor
    n_iter: int
    converged: bool
    history: dict


# ---------------------------------------------------------------------------
# shared helpers
# ---------------------------------------------------------------------------

class _Fock:
    """Fock build adaptor: the libhfx engine, or a user ``fock_build_fn`` (plugin seam)."""

    def __init__(self, kernel, cdt, fock_build_fn=None, block_sizes=None):
        self.kernel = kernel
        self.cdt = cdt
        self.fn = fock_build_fn
        self.block_specs = None if block_sizes is None else (("sizes", tuple(int(s) for s in block_sizes)),)
        self.eng = kernel.fock_engine(dtype=cdt)
        self.eng_nohs = None
        self._args = kernel.as_args() if (self.block_specs is not None or fock_build_fn is not None) else None

    def graph_safe(self, project_fn=None):
        """True when an iteration is device launches only (no host check, no
        callable that synchronises); callables marked ``hfx_graph_safe`` -- the
        projectors of ``jax_hf_amd.symmetry.make_project_fn`` and the build of
        ``jax_hf_amd.superlattice.make_superlattice_build_fock_fn`` -- qualify."""
        pf_ok = project_fn is None or bool(getattr(project_fn, "hfx_graph_safe", False))
        fn_ok = self.fn is None or bool(getattr(self.fn, "hfx_graph_safe", False))
        return fn_ok and self.block_specs is None and pf_ok

    def build(self, P, F, E_ptr=None, project_fn=None, refP="kernel", sub_h=False):
        eng = self.eng
        if self.fn is None and self.block_specs is None and (project_fn is None or not sub_h):
            ref = "default" if isinstance(refP, str) else refP
            eng.build(P, F, energy=E_ptr is not None, E_ptr=E_ptr, refP=ref, sub_h=sub_h)
            if project_fn is not None:      # fock.py:94-95: F = herm(project_fn(F)); E is unchanged
                X = to_tensor(project_fn(F), P.device, self.cdt)
                if X.data_ptr() == F.data_ptr():
                    X = X.clone()
                K.hermitize(X, F)
            return
        if self._args is None:
            self._args = self.kernel.as_args()
        a = self._args
        fn = self.fn
        if fn is None:
            from .fock import build_fock as fn
        ref = a["refP"] if isinstance(refP, str) else refP
        h = a["h"] if self.cdt == self.kernel.dtype else convert(a["h"], self.cdt)
        VR = a["VR"] if self.cdt == self.kernel.dtype else convert(a["VR"], self.cdt)
        S, H, Fo = fn(P, h=h, VR=VR, refP=ref, HH=a["HH"], w2d=self.kernel.w2d,
                      include_exchange=a["include_exchange"], include_hartree=a["include_hartree"],
                      exchange_hermitian_channel_packing=a["exchange_hermitian_channel_packing"],
                      contact_g=a["contact_g"], contact_Oi=a["contact_Oi"], contact_Oj=a["contact_Oj"],
                      exchange_block_specs=self.block_specs, project_fn=project_fn)
        Fo = to_tensor(Fo, P.device, self.cdt)
        if sub_h:
            L.call("hfx_lincomb", Fo.numel(), 2 if self.cdt == torch.complex128 else 3, 1.0, 0.0, L.ptr(Fo),
                   -1.0, 0.0, L.ptr(h), L.ptr(F), L.stream(P.device))
        else:
            F.copy_(Fo)
        if E_ptr is not None:
            from .fock import hf_energy
            E = hf_energy(P, h=h, Sigma=to_tensor(S, P.device, self.cdt), H=H, weights_b=self.kernel._w64)
            L.call("hfx_convert", L.ptr(E), 0, E_ptr, 0, 1, L.stream(P.device))


def _qr_retract(G, tau):
    """U = QR(I - tau G) with the R-diagonal phases normalised (solver.py:108-119).

    ``G`` (..., nb, nb) complex128 / complex64 device tensor (other array-likes
    are uploaded), ``tau`` a scalar or an array broadcastable to ``G.shape[:-2]``.
    Returns the unique Q whose R has a positive real diagonal, computed by one
    launch of ``hfx_qr_retract`` (Gram-Schmidt with re-orthogonalisation, one
    group of lanes per matrix).
    """
    Gt = to_tensor(G)
    if not Gt.is_complex():
        Gt = convert(Gt, torch.complex128 if Gt.dtype == torch.float64 else torch.complex64)
    if Gt.dim() < 2 or Gt.shape[-1] != Gt.shape[-2]:
        raise ValueError(f"G must have shape (..., nb, nb); got {tuple(Gt.shape)}")
    batch = tuple(Gt.shape[:-2])
    nb = int(Gt.shape[-1])
    tau_np = np.broadcast_to(np.asarray(host_array(tau), dtype=np.float64), batch)
    tt = torch.from_numpy(np.array(tau_np, dtype=np.float64).reshape(-1)).to(Gt.device)
    U = torch.empty_like(Gt)
    nk = int(np.prod(batch)) if batch else 1
    L.call("hfx_qr_retract", K.dt_code(Gt), nb, nk, L.ptr(Gt), L.ptr(tt), L.ptr(U), L.stream(Gt.device))
    return U


qr_retract = _qr_retract


def _project_herm(P, project_fn, tmp):
    """P <- herm(project_fn(P))  (solver.py:462)."""
    if project_fn is None:
        return P
    X = to_tensor(project_fn(P), P.device, P.dtype)
    K.hermitize(X, tmp)
    return tmp


# ---------------------------------------------------------------------------
# CG (solver.py:307-660)
# ---------------------------------------------------------------------------

class _CG:
    def __init__(self, kernel, cdt, cfg: SolverConfig, fock_build_fn=None):
        self.kernel = kernel
        self.cfg = cfg
        self.cdt = cdt
        dev = kernel.device
        self.dev = dev
        nk1, nk2, nb = kernel.h.shape[0], kernel.h.shape[1], kernel.h.shape[-1]
        self.shape4 = (nk1, nk2, nb, nb)
        self.nk, self.nb = nk1 * nk2, nb
        self.fock = _Fock(kernel, cdt, fock_build_fn, cfg.block_sizes)
        ws = max(kernel.weight_sum, 1e-30)
        self.w_norm = (kernel._w64 / ws).contiguous()
        self.T = max(kernel.T, 1e-12)
        self.ws = K.Workspace.get(dev, self.nk)
        c = dict(dtype=cdt, device=dev)
        r = dict(dtype=torch.float64, device=dev)
        z = lambda: torch.zeros(self.shape4, **c)  # noqa: E731
        self.Q, self.P, self.F, self.Ft, self.G, self.Gp = z(), z(), z(), z(), z(), z()
        self.H, self.D, self.V, self.Fe, self.tmp = z(), z(), z(), z(), z()
        v = lambda: torch.zeros((nk1, nk2, nb), **r)  # noqa: E731
        self.p, self.eps, self.gp, self.gpp, self.hp, self.dp, self.lam = v(), v(), v(), v(), v(), v(), v()
        self.st = torch.zeros(K.ST_LEN, **r)
        form = str(cfg.cayley).lower()
        if form not in ("gauss_jordan", "spectral"):
            raise ValueError(f"cayley must be 'gauss_jordan' or 'spectral'; got {cfg.cayley!r}")
        self.gauss_jordan = form == "gauss_jordan"
        self.graph = None        # captured iteration bodies {G buffer: graph}, kept with the runner

    # -- pieces ---------------------------------------------------------------
    def density(self):
        K.density(self.Q, self.P, p=self.p)
        return _project_herm(self.P, self.cfg.project_fn, self.tmp)

    # mu steps per mu solve inside the iteration after the bracket pass (which
    # takes the first step; a step after convergence returns at once but still
    # costs a dispatch): the fourth-order steps converge from the previous mu in
    # one or two, and a solve that needs more is finished on the readback
    # (96^2 x 16: no solve needed the readback completion with 1-4 steps here)
    MU_LOOP_STEPS = 3

    # mu steps of an eager (not captured) solve before its done flag is read:
    # the remaining steps of mu_maxiter are launched only when it is not set
    MU_EAGER_STEPS = 6

    def mu(self, eps, from_state, mu0=0.0, steps=None):
        """_solve_mu (solver.py:62-101): the safeguarded root of N(mu) = n inside the
        bracket, graph-capturable launches (fourth-order steps on the device).
        ``steps`` given: that many steps (inside a captured iteration); else an
        eager solve to convergence within mu_maxiter steps."""
        n = self.cfg.mu_maxiter if steps is None else min(steps, self.cfg.mu_maxiter)
        if steps is None and n > self.MU_EAGER_STEPS:
            K.solve_mu_steps(self.ws, eps, self.w_norm, self.nk, self.nb, self.n_norm, self.T, mu0=mu0,
                             mu0_from_state=from_state, maxiter=self.MU_EAGER_STEPS, st=self.st, p_out=self.p)
            if float(self.st[K.ST_MUDONE].item()) != 0.0:
                return
            from_state, n = True, n - self.MU_EAGER_STEPS
        K.solve_mu_steps(self.ws, eps, self.w_norm, self.nk, self.nb, self.n_norm, self.T, mu0=mu0,
                         mu0_from_state=from_state, maxiter=n, st=self.st, p_out=self.p)

    def set_target(self, n_electrons):
        self.n_norm = float(n_electrons) / max(self.kernel.weight_sum, 1e-30)

    def init_from(self, P0):
        """solver.py:380-394: eigh(F[herm(P0)]) -> (Q, eps), mu, p."""
        P0h = torch.empty(self.shape4, dtype=self.cdt, device=self.dev)
        if isinstance(P0, torch.Tensor) and P0.device == self.dev and P0.dtype in (torch.complex128, torch.complex64) \
                and P0.dtype != self.cdt and tuple(P0.shape) == self.shape4:
            P0c = P0.contiguous()      # hermitised and converted in one pass
            L.call("hfx_hermitize_convert", L.ptr(P0c), K.dt_code(P0c), L.ptr(P0h), K.dt_code(P0h),
                   P0c.numel() // (self.nb * self.nb), self.nb, L.stream(self.dev))
        else:
            K.hermitize(to_tensor(P0, self.dev, self.cdt), P0h)
        self.fock.build(P0h, self.F, project_fn=self.cfg.project_fn)
        K.eigh(self.F, self.eps, self.Q)
        self.mu(self.eps, False, 0.0)

    def warm_from(self, Q, p, mu):
        self.Q.copy_(convert(Q, self.cdt))
        self.p.copy_(p)
        self.st[K.ST_MU] = float(mu)

    def run(self, n_electrons, tol_E, tol_grad, max_iter, rel_tol=False, abs_window=False):
        cfg = self.cfg
        self.set_target(n_electrons)
        st = self.st
        self.G.zero_()
        self.Gp.zero_()
        self.D.zero_()
        self.gp.zero_()
        self.gpp.zero_()
        self.dp.zero_()
        hist_E, hist_G = [], []
        k = 0
        grad_norm, dE, E_prev = np.inf, np.inf, np.inf
        shrink_pow = cfg.bt_shrink ** cfg.bt_max
        tol_E_rel = tol_E
        st[K.ST_ITER] = 0.0                 # device iteration counter (CG restart schedule)
        use_graph = bool(cfg.use_graph) and self.fock.graph_safe(cfg.project_fn)
        while True:
            # stopping rule (solver.py:422-449); rel_tol scales tol_E by |E| (FP32 phase)
            if rel_tol and hist_E:
                tol_E = tol_E_rel * max(1.0, abs(hist_E[-1]))
            if cfg.plateau_window > 0:
                e_now = hist_E[max(k - 1, 0)] if hist_E else 0.0
                e_past = hist_E[max(k - 1 - cfg.plateau_window, 0)] if hist_E else 0.0
                # abs_window (refinement phase of the mixed form): |change| over the
                # window, since E may rise while the free energy falls when the
                # loop restarts from a nearby, not fully consistent, state
                win = abs(e_past - e_now) if abs_window else (e_past - e_now)
                enc = win > tol_E if k > cfg.plateau_window else dE > tol_E
            else:
                enc = dE > tol_E
            nc = grad_norm > tol_grad if tol_grad > 0.0 else enc
            if not (k < max_iter and nc):
                break
            if self.graph is None and use_graph and k >= 1:   # k = 0 ran eagerly: buffers exist
                self.graph = self._capture(shrink_pow)
            if use_graph and self.graph is not None:
                self.graph[self.G.data_ptr()].replay()
                self._swap()
            else:
                self._body(shrink_pow)
            sv = st.cpu().numpy()                                               # one readback
            if sv[K.ST_MUDONE] == 0.0:       # mu not converged in the loop's steps: finish it
                self.mu(self.eps, True)
            E = float(sv[K.ST_E])
            grad_norm = float(sv[K.ST_GRAD])
            hist_E.append(E)
            hist_G.append(grad_norm)
            dE = abs(E - E_prev)
            E_prev = E
            k += 1
        return k, grad_norm, hist_E, hist_G

    def _body(self, shrink_pow):
        """One CG iteration (solver.py:451-604) as device launches only: no host
        read, no allocation, so the same sequence can be captured in a graph."""
        cfg = self.cfg
        nk, nb = self.nk, self.nb
        dt = K.dt_code(self.Q)
        st, ws, part = self.st, self.ws, self.ws.part
        sm = L.stream(self.dev)
        P = self.density()                                                      # :462
        self.fock.build(P, self.F, E_ptr=K.st_ptr(st, K.ST_E))                  # :463-471
        L.call("hfx_cg_grad", dt, nb, nk, L.ptr(self.Q), L.ptr(self.F), L.ptr(self.p), L.ptr(self.Ft),
               L.ptr(self.eps), L.ptr(self.G), L.ptr(part), sm)                 # :472-478
        K.reduce(ws, part, 2, nk, w=self.w_norm, wmask=0b01, st=st, op=K.OP_GRAD,
                 params=[self.T, cfg.denom_scale, nk * nb])                     # :491-497
        L.call("hfx_cg_precond", dt, nb, nk, L.ptr(self.G), L.ptr(self.Gp), L.ptr(self.eps), L.ptr(self.p),
               L.ptr(self.gpp), L.ptr(st), self.T, L.ptr(self.H), L.ptr(self.gp), L.ptr(self.hp),
               L.ptr(part), 0, sm)                                              # :486, 499-504
        # restart when st[ITER] % cg_restart == 0 (k == 0 included), :522-526
        K.reduce(ws, part, 6, nk, w=self.w_norm, wmask=0b111111, st=st, op=K.OP_BETA,
                 params=[0, 0, 0, 0, 0, 0, 0, 0, 0, max(int(cfg.cg_restart), 1)])   # :517-526
        gj = self.gauss_jordan
        if gj:    # Cayley factor per trial by Gauss-Jordan: no eigh in the loop
            L.call("hfx_cg_direction_gj", dt, nb, nk, L.ptr(self.H), L.ptr(self.D), L.ptr(self.hp),
                   L.ptr(self.dp), L.ptr(st), L.ptr(part), 0, sm)
        else:     # spectral setup: eigh(i d) once, V^dagger Ft V  (:538-539)
            L.call("hfx_cg_direction", dt, nb, nk, L.ptr(self.H), L.ptr(self.D), L.ptr(self.hp), L.ptr(self.dp),
                   L.ptr(self.Ft), L.ptr(st), L.ptr(self.V), L.ptr(self.lam), L.ptr(self.Fe), L.ptr(part), sm)
        K.reduce(ws, part, 1, nk, w=self.w_norm, wmask=1, st=st, op=K.OP_TAU0,
                 params=[0, 0, 0, cfg.max_step, 0, shrink_pow])                 # :543-547
        for i in range(cfg.bt_max):                                             # :563-579
            if gj:
                L.call("hfx_cg_trial_gj", dt, nb, nk, L.ptr(self.D), L.ptr(self.Ft), L.ptr(self.p),
                       L.ptr(self.dp), L.ptr(self.eps), L.ptr(st), int(i == 0), L.ptr(part),
                       L.ptr(self.V), L.ptr(self.lam), 0, sm)
            else:
                L.call("hfx_cg_trial", dt, nb, nk, L.ptr(self.V), L.ptr(self.lam), L.ptr(self.Fe), L.ptr(self.p),
                       L.ptr(self.dp), L.ptr(self.eps), L.ptr(st), int(i == 0), L.ptr(part), sm)
            K.reduce(ws, part, 4 if i == 0 else 2, nk, w=self.w_norm, wmask=0b1111, st=st, op=K.OP_TRIAL,
                     params=[self.T, 0, 0, 0, cfg.bt_shrink, shrink_pow, 0, int(i == cfg.bt_max - 1),
                             int(i == 0)])
        if gj:
            L.call("hfx_cg_retract_gj", dt, nb, nk, L.ptr(self.Q), L.ptr(self.D), L.ptr(self.Ft), L.ptr(st),
                   L.ptr(self.eps), L.ptr(self.V), L.ptr(self.lam), 0, sm)      # :586-590
        else:
            L.call("hfx_cg_retract", dt, nb, nk, L.ptr(self.Q), L.ptr(self.V), L.ptr(self.lam), L.ptr(self.Fe),
                   L.ptr(st), 0.0, 0, L.ptr(self.eps), None, sm)                # :586-590
        self.mu(self.eps, True, steps=self.MU_LOOP_STEPS)                      # :591-592
        self._swap()     # this iteration's gradient / preconditioned norms become the previous ones

    def _swap(self):
        self.G, self.Gp = self.Gp, self.G
        self.gp, self.gpp = self.gpp, self.gp

    def _capture(self, shrink_pow):
        """HIP graphs of the iteration body (identical launches to the eager body),
        one per parity of the buffer swap it makes, keyed by the G buffer each was
        captured with (the two captures swap twice: the references end as they began)."""
        g0 = self.G.data_ptr()
        gs = capture_graphs(self.dev, lambda: self._body(shrink_pow), n=2)
        return {g0: gs[0], self.Gp.data_ptr(): gs[1]}

    def finalize(self):
        """solver.py:606-638: Fock, eigh canonicalisation, mu, p, Fock + E."""
        P = self.density()
        self.fock.build(P, self.F)
        # Q <- Q W, (eps, W) = eigh(Q^dagger F Q)  (:622-625): the register eigensolver
        # accumulates the rotations onto Q in place (the captured graph holds this buffer as Q)
        K.matmul(self.Q, self.F, self.Ft, C=self.Q, ops=0b001)              # Ft = Q^H F Q
        K.eigh_basis(self.Ft, self.eps, self.Q, herm_in=True)
        self.mu(self.eps, True)
        P = self.density()
        Pout = P.clone()
        F = torch.empty_like(Pout)
        self.fock.build(Pout, F, E_ptr=K.st_ptr(self.st, K.ST_E))
        return Pout, F


def _cg_runner(kernel, cdt, cfg, fock_build_fn, n_electrons):
    """The CG state for (kernel, precision, loop controls, electron count).

    Kept on the kernel between solves when the iteration is graph-captured:
    the buffers and the captured graph (which bakes in these controls and the
    electron count) are reused, so a repeated solve allocates and captures
    nothing.  At most four runners are kept per kernel.  A user Fock build or
    projector marked ``hfx_graph_safe`` is part of the key (by identity; the
    runner keeps a reference, so the identity cannot be reused while cached).
    """
    def safe(f):
        return f is None or bool(getattr(f, "hfx_graph_safe", False))

    if not (safe(fock_build_fn) and safe(cfg.project_fn)) or cfg.block_sizes is not None or not cfg.use_graph:
        return _CG(kernel, cdt, cfg, fock_build_fn)
    key = (cdt, float(cfg.denom_scale), float(cfg.max_step), int(cfg.cg_restart), float(cfg.bt_shrink),
           int(cfg.bt_max), int(cfg.mu_maxiter), str(cfg.cayley).lower(), float(n_electrons),
           id(fock_build_fn), id(cfg.project_fn))
    cache = kernel.__dict__.setdefault("_hfx_cg", {})
    r = cache.get(key)
    if r is None:
        while len(cache) >= 4:
            cache.pop(next(iter(cache)))
        r = _CG(kernel, cdt, cfg, fock_build_fn)
        r.kernel = r.fock.kernel = weakref.proxy(kernel)    # the kernel owns the runner, not both ways
        r.keep = (fock_build_fn, cfg.project_fn)
        cache[key] = r
    r.cfg = cfg
    return r


def _solve_cg(kernel, P0, n_electrons, cfg: SolverConfig, fock_build_fn):
    prec = _precision(kernel, cfg)
    dev = kernel.device
    polish = cfg.polish if cfg.polish is not None else ("diis" if prec == "mixed" else None)
    if polish not in (None, "diis"):
        raise ValueError(f"polish must be None or 'diis'; got {cfg.polish!r}")
    cdt = torch.complex128 if prec == "fp64" else torch.complex64
    cfg_cg = replace(cfg, bt_max=min(cfg.bt_max, max(int(cfg.mixed_bt_max), 1))) if prec == "mixed" else cfg
    runner = _cg_runner(kernel, cdt, cfg_cg, fock_build_fn, n_electrons)
    runner.set_target(n_electrons)
    runner.init_from(P0)
    if prec == "mixed":
        # FP32 phase: relative windowed stop, at most half of the iteration budget
        k, grad, hist_E, hist_G = runner.run(n_electrons, cfg.mixed_tol_E, 0.0, max(cfg.max_iter // 2, 1),
                                             rel_tol=True)
    else:
        k, grad, hist_E, hist_G = runner.run(n_electrons, cfg.tol_E, cfg.tol_grad, cfg.max_iter)
    if polish == "diis":
        return _polish_diis(kernel, runner, n_electrons, cfg, fock_build_fn, k, hist_E, hist_G)
    Pfin, Ffin = runner.finalize()
    converged = (grad <= cfg.tol_grad) if cfg.tol_grad > 0.0 else (k < cfg.max_iter)
    sv = runner.st.cpu()
    return _result(runner, Pfin, Ffin, sv, k, converged, hist_E, hist_G, cfg.max_iter, dev)


class _DIISPolish:
    """FP64 Pulay C-DIIS from the CG density to the fixed point [F, P] = 0.

    One eigh per iteration: F_i = F[P_i] and the DIIS error e_i = F_i P_i -
    P_i F_i (reference_scf.py:292-314) give P_{k+1} = density(herm(sum_i c_i
    F_i)).  The reference loop also forms density(F_k) for its density
    residual, a second eigh per iteration; here the step |P_k - P_{k-1}| the
    iteration itself takes stands in for it, and once the commutator residual
    is within tolerance the reference's |density(F_k) - P_k| is formed directly
    (_roothaan_check), so the stop is the reference's own test.  The density update
    (eigh of the Hermitised extrapolated Fock matrix, Newton mu warm-started
    at the previous mu, occupations, density) is one replayed HIP graph, and
    the object (buffers + graph) is kept on the kernel between solves.

    The eigh starts from the previous basis U: U^H F U is nearly diagonal once
    the iteration settles, so the Jacobi sweeps stop after one or two, and
    V = U W (the CG finish does the same, solver.py:622-625).  A run starts
    from the CG orbitals (re-orthonormalised in FP64) when they are given,
    else with a cold eigh; either way the rounding in the accumulated basis
    stays bounded by one run's updates.
    """

    def __init__(self, kernel, n_electrons, nh=6, cdt=torch.complex128):
        self.cdt = cdt
        self.kernel = kernel
        self.dev = dev = kernel.device
        nk1, nk2, nb = kernel.h.shape[0], kernel.h.shape[1], kernel.h.shape[-1]
        self.shape4 = (nk1, nk2, nb, nb)
        self.nk, self.nb = nk1 * nk2, nb
        self.n_e = float(n_electrons)
        self.fock = _Fock(kernel, cdt, None, None)
        self.w = kernel._w64.reshape(-1).contiguous()
        self.ws_sum = max(kernel.weight_sum, 1e-30)
        self.wsp = K.Workspace.get(dev, self.nk)
        self.nh = nh
        c = dict(dtype=cdt, device=dev)
        z = lambda: torch.zeros(self.shape4, **c)  # noqa: E731
        self.P_prev, self.P_cur, self.P_new, self.Fx, self.U, self.tmp = z(), z(), z(), z(), z(), z()
        self.Ft = z()
        self.F_hist = [z() for _ in range(nh)]
        # DIIS residual history: complex64 at nb = 16 (R and its norms are formed
        # in FP64; only the stored copies that feed the B matrix are rounded);
        # the complex64 stage keeps it in its own precision
        rdt = torch.complex64 if (self.HIST32 and nb == 16) or cdt == torch.complex64 else cdt
        self.R_hist = [torch.zeros(self.shape4, dtype=rdt, device=dev) for _ in range(nh)]
        self.eps = torch.zeros((nk1, nk2, nb), dtype=torch.float64, device=dev)
        self.occ = torch.zeros_like(self.eps)
        # one readback buffer: outs [0:2] residuals, [2:2+nh] B row, [16] E; then the state vector
        self.io = torch.zeros(64 + K.ST_LEN, dtype=torch.float64, device=dev)
        self.outs, self.st = self.io[:64], self.io[64:]
        self._eigh_tol = 0.0
        self.io_host = torch.empty(self.io.shape, dtype=torch.float64, pin_memory=True)
        self.ev = torch.cuda.Event()
        # the Roothaan check's own orbitals, occupations and mu state, and its readback
        self.U_chk = z()
        self.eps_chk, self.occ_chk = torch.zeros_like(self.eps), torch.zeros_like(self.eps)
        self.st_chk = torch.zeros(K.ST_LEN, dtype=torch.float64, device=dev)
        self.chk = torch.zeros(2, dtype=torch.float64, device=dev)
        self.chk_host = torch.empty(2, dtype=torch.float64, pin_memory=True)
        self.Bd = torch.zeros((nh, nh), dtype=torch.float64, device=dev)   # DIIS B matrix
        self.coef = torch.zeros(8, dtype=torch.float64, device=dev)        # DIIS coefficients
        import ctypes as C
        self._xa = (C.c_void_p * nh)(*[f.data_ptr() for f in self.F_hist])
        self.graphs = {}

    # Relative off-diagonal tolerance of the update's eigh, per unit of the
    # polish tolerance.  The update's density is exact for its (unitary) basis
    # either way and the fixed point is decided by the residual test, but the
    # neglected off-diagonals set a floor on the reachable residual (measured at
    # 96^2 x 16: ~250 x the eigh tolerance), so the eigh stays 1e5 times tighter
    # than the polish tolerance (2e-3 of it as the floor).
    EIGH_TOL_PER_TOL = 1e-5

    # the stored DIIS residuals in complex64 (nb = 16): measured at 96^2, 64^2 and
    # 48^2 x 16 for polish tolerances 1e-6 and 1e-8, the same iteration counts and
    # energies as FP64 storage, and the residual pass reads 1/3 fewer bytes
    HIST32 = True

    # mu steps per mu solve inside the update graph after the bracket pass: from
    # the previous mu one more converges; a solve that needs more is completed
    # on the next readback
    MU_GRAPH_STEPS = 2

    # Once the commutator residual is below the tolerance while the step residual
    # |P_k - P_{k-1}| (which lags one iteration) is not, test the reference's own
    # density residual |density(F_k) - P_k| (reference_scf.py:287-294) directly:
    # one warm eigh of F_k, whose orbitals are the returned state's when it passes,
    # instead of a further Fock build, update and final eigh.
    ROOTHAAN_CHECK = True
    MU_CHECK_STEPS = 4                # mu steps of the check before its one readback

    def _roothaan_check(self, Fk, Pk, T):
        """|density(F_k) - P_k| (weighted, as the residuals) with the orbitals,
        occupations and mu of F_k left in U_chk / occ_chk / st_chk."""
        K.extrap_rotate(self.Ft, self.U, [Fk], [1.0])
        K.eigh_basis(self.Ft, self.eps_chk, self.U, herm_in=True, out=self.U_chk)
        self.st_chk.copy_(self.st)
        mu_kw = dict(mu0_from_state=True, st=self.st_chk, p_out=self.occ_chk)
        for steps in (self.MU_CHECK_STEPS, 54):
            K.solve_mu_steps(self.wsp, self.eps_chk, self.w, self.nk, self.nb, self.n_e, T, maxiter=steps, **mu_kw)
            K.density_diff(self.U_chk, self.occ_chk, Pk, self.tmp, self.wsp.part)
            K.reduce(self.wsp, self.wsp.part, 1, self.nk, w=self.w, wmask=1, out=self.chk[0:1])
            self.chk[1:2].copy_(self.st_chk[K.ST_MUDONE:K.ST_MUDONE + 1])
            self.chk_host.copy_(self.chk, non_blocking=True)
            self.ev.record()
            self.ev.synchronize()
            if self.chk_host[1].item() != 0.0:    # else mu needed more steps: finish it, once
                break
        return float(np.sqrt(self.chk_host[0].item() / self.ws_sum))

    def _mu(self, maxiter=54):
        K.solve_mu_steps(self.wsp, self.eps, self.w, self.nk, self.nb, self.n_e, self.kernel.T,
                         mu0_from_state=True, maxiter=maxiter, st=self.st, p_out=self.occ)

    def _extrapolate(self, Xs, cs, warm):
        """herm(sum c X): in the basis U (Ft, warm) or as is (Fx)."""
        if warm:
            K.extrap_rotate(self.Ft, self.U, Xs, cs)
        else:
            K.lincomb_multi(self.Fx, Xs, cs, herm=True)

    def _update(self, out, warm=True):
        """-> out (the next iterate): launches only (the warm form is captured,
        once per output buffer).  Warm: eigh of Ft = U^H F U with the rotations
        accumulated onto U (U <- U W, eps), mu, then the density from (U, occ)."""
        if warm:
            K.eigh_basis(self.Ft, self.eps, self.U, herm_in=True, rel_tol=self._eigh_tol)
            self._mu(self.MU_GRAPH_STEPS)
            K.density(self.U, out, p=self.occ)
        else:
            K.eigh(self.Fx, self.eps, self.U, herm_in=True)
            self._mu(self.MU_GRAPH_STEPS)
            K.density(self.U, self.tmp, p=self.occ)
            K.hermitize(self.tmp, out)

    def _capture(self, out):
        return capture_graphs(self.dev, lambda: self._update(out))[0]

    def _eval(self, k):
        """F_k = F[P_k] with E_k, the residuals and the new B row of iterate k
        (P_cur), reduced into outs; then the readback, asynchronous."""
        nh, part = self.nh, self.wsp.part
        idx = k % nh
        Fk, Rk = self.F_hist[idx], self.R_hist[idx]
        self.fock.build(self.P_cur, Fk, E_ptr=self.outs.data_ptr() + 8 * 16)
        # residuals and the B row (reference_scf.py:110-111) in one pass; their
        # block partials are summed by the device DIIS solve (reference_scf.py:298-314)
        # of the update from iterate k, which writes the totals for the readback
        K.scf_resid_dots(Fk, self.P_cur, self.P_prev if k > 0 else None, Rk, self.R_hist, idx, part)
        K.reduce(self.wsp, part, 2 + nh, self.nk, w=self.w, wmask=(1 << (2 + nh)) - 1, op=K.OP_PARTIALS)
        L.call("hfx_diis_coefficients_reduce", L.ptr(self.wsp.scratch), self.nk, 2 + nh, L.ptr(self.outs), L.ptr(self.Bd),
               nh, idx, min(k + 1, nh), L.ptr(self.coef), L.stream(self.dev))
        self.io_host.copy_(self.io, non_blocking=True)
        self.ev.record()

    def _advance(self, k, warm):
        """Update from iterate k, launches only, with the DIIS coefficients
        (reference_scf.py:298-314) that _eval(k) solved on the device:
        extrapolation in the basis U, eigh, mu, density; then P_prev <- P_k,
        P_cur <- P_{k+1}."""
        nh = self.nh
        idx = k % nh
        if warm:
            count = min(k + 1, nh)           # active history slots 0 .. count-1 (coefficients: _eval)
            import ctypes as C
            L.call("hfx_extrap_rotate_batched", K.dt_code(self.U), self.nb, self.nk, self.nk, count,
                   C.cast(self._xa, C.c_void_p), L.ptr(self.coef), L.ptr(self.U), L.ptr(self.Ft),
                   L.stream(self.dev))
            key = self.P_new.data_ptr()   # the three iterate buffers rotate: one graph per output
            g = self.graphs.get(key)
            if g is None:
                self._update(self.P_new)
                self.graphs[key] = self._capture(self.P_new)
            else:
                g.replay()
        else:                             # first update of a cold run: the plain step, eigh of herm(F_0)
            self._extrapolate([self.F_hist[idx]], [1.0], False)
            self._update(self.P_new, warm=False)
        self.P_prev, self.P_cur, self.P_new = self.P_cur, self.P_new, self.P_prev

    def run(self, P_lo, mu0, tol, max_iter, Q_lo=None, stall=0.0, P_herm=False):
        """Q_lo (optional): the orbitals P_lo was built from.  Re-orthonormalised in
        FP64 (two Newton-Schulz steps), they are the first basis, so that the
        first update is a warm eigh as well.  P_herm: P_lo is exactly Hermitian
        already (a density of orbitals), so it is only converted.

        Pipelined: the update from iterate k is launched before iterate k's
        residuals are back on the host (the DIIS coefficients are computed on
        the device), so the host's stopping test overlaps the update instead of
        stalling the GPU.  When the test stops at k, that update is discarded
        (iterate k is kept in P_prev); when the residual trend predicts the
        stop, the update waits for the test instead.

        stall > 0: also stop (not converged) once the commutator residual has
        not fallen below stall x its smallest value for two iterations (the
        rounding floor of a complex64 stage)."""
        nh = self.nh
        max_iter = max(int(max_iter), 1)
        et = self.EIGH_TOL_PER_TOL * float(tol)
        if et != self._eigh_tol:          # the captured updates hold the eigh tolerance
            self._eigh_tol, self.graphs = et, {}
        if P_herm and P_lo.dtype != self.cdt:
            Pc = P_lo.contiguous()
            L.call("hfx_convert", L.ptr(Pc), KIND[Pc.dtype], L.ptr(self.P_cur), KIND[self.cdt], Pc.numel(),
                   L.stream(self.dev))
        else:
            K.hermitize(convert(P_lo, self.cdt) if P_lo.dtype != self.cdt else P_lo, self.P_cur)
        self.st[K.ST_MU] = mu0
        warm0 = Q_lo is not None
        if warm0:
            Qc = Q_lo.contiguous()
            # Newton-Schulz to this precision's unitarity: two steps from complex64
            # orbitals in FP64 (converted inside the first), one in complex64
            if self.cdt == torch.complex128 and Qc.dtype == torch.complex64:
                L.call("hfx_orthonormalize_from", self.nb, self.nk, L.ptr(Qc), L.ptr(self.U), 2, L.stream(self.dev))
            else:
                L.call("hfx_convert", L.ptr(Qc), KIND[Qc.dtype], L.ptr(self.U), KIND[self.cdt], Qc.numel(),
                       L.stream(self.dev))
                K.orthonormalize(self.U, 2 if self.cdt == torch.complex128 and Qc.dtype != torch.complex128 else 1)
        hE, hC, hD = [], [], []
        k, converged, E, idx, redo, checked = 0, False, float("nan"), 0, -1, False
        P_res = self.P_cur
        self._eval(0)
        while True:
            idx = k % nh
            Pk = self.P_cur
            # speculate unless the residual trend says iterate k may be the last
            spec = k + 1 < max_iter
            if spec and len(hC) >= 2 and hC[-2] > 0 and 0 < hD[-2] < float("inf"):
                pc = hC[-1] * hC[-1] / hC[-2]
                pd = hD[-1] * hD[-1] / hD[-2]
                # with the density-residual check, a predicted commutator stop
                # suffices: the check then decides before any further update
                spec = not (pc <= 2.0 * tol and (pd <= 2.0 * tol or self.ROOTHAAN_CHECK))
            if spec:
                self._advance(k, warm=k > 0 or warm0)
            self.ev.synchronize()
            sv_all = self.io_host.numpy()
            sv = sv_all[:64]
            c_res = float(np.sqrt(sv[0] / self.ws_sum))
            d_res = float(np.sqrt(sv[1] / self.ws_sum)) if k > 0 else float("inf")
            E = float(sv[16])
            if k > 0 and sv_all[64 + K.ST_MUDONE] == 0.0 and redo != k:
                # the update that made P_k needed more mu steps than its graph has:
                # P_k's electron count is off.  Finish mu on the newest eigenvalues
                # and rebuild that density (after the speculative update: P_{k+1};
                # else P_k, which is then evaluated again).
                self._mu()
                K.density(self.U, self.tmp, p=self.occ)
                K.hermitize(self.tmp, self.P_cur)
                if not spec:
                    redo = k                  # once per iterate
                    self._eval(k)
                    continue
                hE.append(E)
                hC.append(c_res)
                hD.append(d_res)
                k += 1
                self._eval(k)
                continue
            hE.append(E)
            hC.append(c_res)
            hD.append(d_res)
            if c_res <= tol and d_res <= tol:
                converged = True
                P_res = Pk
                break
            if c_res <= tol and k > 0 and self.ROOTHAAN_CHECK:
                d_ref = self._roothaan_check(self.F_hist[idx], Pk, self.kernel.T)
                if d_ref <= tol:
                    hD[-1] = d_ref
                    converged, checked = True, True
                    P_res = Pk
                    break
            if k + 1 >= max_iter:
                break
            if stall > 0.0 and len(hC) >= 4 and min(hC[-2:]) > stall * min(hC[:-2]):
                break
            if not spec:
                self._advance(k, warm=k > 0 or warm0)
            k += 1
            self._eval(k)
        n_it = len(hE)
        Ffin = self.F_hist[idx]           # F of the returned iterate (converged, or the last one)
        if not converged:
            P_res = self.P_cur
        if checked:                       # the check's orbitals are those of F_k
            # the check's buffers and F_k's history slot are in no captured graph:
            # hand them out and take fresh ones instead of copying
            Q, p = self.U_chk, self.occ_chk
            self.U_chk, self.occ_chk = torch.empty_like(Q), torch.empty_like(p)
            self.F_hist[idx] = torch.empty_like(Ffin)
            import ctypes as C
            self._xa = (C.c_void_p * self.nh)(*[f.data_ptr() for f in self.F_hist])
            return (P_res.clone(), Ffin, Q, p, self.st_chk[K.ST_MU].clone(), E, n_it, converged, hE, hC)
        # orbitals / occupations / mu of the returned state (solver.py:606-638 form)
        warm = k > 0 or warm0 or spec
        if warm:
            K.extrap_rotate(self.Ft, self.U, [Ffin], [1.0])
            K.eigh_basis(self.Ft, self.eps, self.U, herm_in=True)       # U <- U W
        else:
            K.eigh(Ffin, self.eps, self.U, herm_in=True)
        self._mu(self.MU_GRAPH_STEPS)
        if float(self.st[K.ST_MUDONE].item()) == 0.0:
            self._mu()
        Q = self.U.clone()
        return (P_res.clone(), Ffin.clone(), Q, self.occ.clone(), self.st[K.ST_MU].clone(), E, n_it,
                converged, hE, hC)


def _polish_state(kernel, n_electrons, cdt=torch.complex128):
    cache = kernel.__dict__.setdefault("_hfx_polish", {})
    key = (float(n_electrons), cdt)
    pol = cache.get(key)
    if pol is None:
        for k in [k for k in cache if k[1] == cdt]:
            del cache[k]              # one polish state per kernel and precision (the last electron count)
        pol = _DIISPolish(kernel, n_electrons, cdt=cdt)
        pol.kernel = pol.fock.kernel = weakref.proxy(kernel)
        cache[key] = pol
    return pol


def _polish_diis(kernel, runner, n_electrons, cfg, fock_build_fn, k_cg, hist_E, hist_G, P_lo=None, mu0=None):
    """FP64 Pulay-DIIS SCF from the CG density to the fixed point [F, P] = 0.

    The CG loop's windowed energy stop (solver.py:422-449) can end well before
    stationarity (E keeps drifting while the free energy still falls); the
    mixed form therefore finishes with the reference's own DIIS SCF
    (reference_scf.py:298-314) in FP64, to residuals <= polish_tol, so its
    energy is the FP64 fixed point's, independent of the FP32 trajectory.
    """
    from .reference_scf import SCFConfig, _SCF
    dev = kernel.device
    P_lo_given = P_lo is not None
    if P_lo is None:
        P_lo = runner.density()       # from the orbitals runner.Q, which then start the polish basis
    if fock_build_fn is None and cfg.project_fn is None and cfg.use_graph:
        mu_start = runner.st[K.ST_MU] if mu0 is None else mu0
        Q_lo = runner.Q if (runner is not None and P_lo_given is False) else None
        phases = [k_cg]
        hE, hG = list(hist_E), list(hist_G)
        stage32 = (_precision(kernel, cfg) == "mixed" and float(cfg.polish32_tol) > float(cfg.polish_tol)
                   and Q_lo is not None and Q_lo.dtype == torch.complex64)
        if stage32:
            # complex64 DIIS from the CG orbitals to its own floor: the FP64 polish
            # then starts a few refinement builds from the fixed point
            p32 = _polish_state(kernel, n_electrons, torch.complex64)
            P_lo, _, Q_lo, _, mu_start, _, n32, _, hE32, hC32 = p32.run(
                P_lo, mu_start, float(cfg.polish32_tol), cfg.polish_max_iter, Q_lo=Q_lo, stall=0.7)
            phases.append(n32)
            hE += list(hE32)
            hG += list(hC32)
        pol = _polish_state(kernel, n_electrons)
        density, F, Q, p, mu, E, n2, conv, hE2, hC2 = pol.run(P_lo, mu_start, cfg.polish_tol, cfg.polish_max_iter,
                                                              Q_lo=Q_lo, P_herm=not P_lo_given and cfg.project_fn is None)
        phases.append(n2)
        k = sum(phases)
        hE += list(hE2)
        hG += list(hC2)
        hEt, hGt = _histories_to_device([hE[:k], hG[:k]], max(cfg.max_iter, k), dev)
        hist = dict(E=hEt, grad_norm=hGt, phase_iters=torch.tensor(phases, dtype=torch.int64))
        return SolveResult(Q=Q, p=p, mu=mu, density=density, fock=F,
                           energy=torch.full((), E, dtype=torch.float64, device=dev), n_iter=int(k),
                           converged=bool(conv), history=hist)
    P64 = convert(P_lo, torch.complex128) if P_lo.dtype != torch.complex128 else P_lo.clone()
    scfg = SCFConfig(max_iter=cfg.polish_max_iter, acceleration="diis", density_tol=cfg.polish_tol,
                     comm_tol=cfg.polish_tol, project_fn=cfg.project_fn)
    scf = _SCF(kernel, scfg, fock_build_fn, dtype=torch.complex128, mu_newton=True)
    density, F, E, mu, n2, conv, hE2, hD2, hC2 = scf.run(P64, n_electrons)
    # orbitals / occupations of the returned state (solver.py:606-638 form)
    eps = torch.empty((F.shape[0], F.shape[1], F.shape[-1]), dtype=torch.float64, device=dev)
    Q = torch.empty_like(F)
    K.eigh(F, eps, Q, herm_in=True)
    p = torch.empty_like(eps)
    L.call("hfx_fermi", L.ptr(eps), L.ptr(mu), float(kernel.T), eps.numel(), L.ptr(p),
           L.stream(dev))
    k = k_cg + n2
    hE = list(hist_E) + list(hE2)
    hG = list(hist_G) + list(hC2)
    hEt = torch.zeros(max(cfg.max_iter, k), dtype=torch.float64)
    hGt = torch.zeros_like(hEt)
    hEt[:k] = torch.tensor(hE[:k], dtype=torch.float64)
    hGt[:k] = torch.tensor(hG[:k], dtype=torch.float64)
    hist = dict(E=hEt.to(dev), grad_norm=hGt.to(dev),
                phase_iters=torch.tensor([k_cg, n2], dtype=torch.int64))
    return SolveResult(Q=Q, p=p, mu=mu.to(dev), density=density, fock=F, energy=E, n_iter=int(k),
                       converged=bool(conv), history=hist)


def _histories_to_device(rows, m, dev):
    """Per-iteration histories as zero-padded length-m device vectors: one
    pinned host buffer and one asynchronous copy (no wait on the GPU)."""
    H = torch.zeros((len(rows), m), dtype=torch.float64, pin_memory=True)
    for j, r in enumerate(rows):
        n = min(len(r), m)
        if n:
            H[j, :n] = torch.tensor(r[:n], dtype=torch.float64)
    D = H.to(dev, non_blocking=True)
    return tuple(D[j] for j in range(len(rows)))


def _result(runner, Pfin, Ffin, sv, k, converged, hist_E, hist_G, max_iter, dev):
    hE = torch.zeros(max_iter, dtype=torch.float64)
    hG = torch.zeros(max_iter, dtype=torch.float64)
    n = min(len(hist_E), max_iter)
    if n:
        hE[:n] = torch.tensor(hist_E[:n], dtype=torch.float64)
        hG[:n] = torch.tensor(hist_G[:n], dtype=torch.float64)
    return SolveResult(Q=runner.Q.clone(), p=runner.p.clone(), mu=sv[K.ST_MU].to(dev), density=Pfin, fock=Ffin,
                       energy=sv[K.ST_E].to(dev), n_iter=int(k), converged=bool(converged),
                       history=dict(E=hE.to(dev), grad_norm=hG.to(dev)))


def _precision(kernel, cfg):
    p = str(cfg.precision).lower()
    if p == "auto":
        return "fp64" if kernel.dtype == torch.complex128 else "fp32"
    if p not in ("fp64", "mixed", "fp32"):
        raise ValueError(f"precision must be 'auto', 'fp64', 'mixed' or 'fp32'; got {cfg.precision!r}")
    return p


# ---------------------------------------------------------------------------
# public dispatcher (solver.py:1098-1251)
# ---------------------------------------------------------------------------

def solve_direct_minimization(kernel, P0, n_electrons: float, *, config: SolverConfig | None = None,
                              fock_build_fn=None) -> SolveResult:
    """Direct-minimisation HF solver (solver.py:1098-1247): CG or trust-region Newton."""
    from .utils import validate_kernel_electron_count
    if config is None:
        config = SolverConfig()
    validate_kernel_electron_count(kernel, n_electrons, context="n_electrons")
    optimizer = str(config.optimizer).lower()
    if optimizer not in ("cg", "newton"):
        raise ValueError(f"Unknown optimizer {config.optimizer!r}; expected 'cg' or 'newton'.")
    deflation_on = config.deflation_targets is not None and float(config.deflation_sigma) > 0.0
    if deflation_on and optimizer != "newton":
        raise ValueError("deflation is only supported with optimizer='newton' "
                         f"(got optimizer={config.optimizer!r}).")
    if kernel.h.shape[-1] > L.PERK_NB_MAX:
        raise ValueError(f"the GPU solvers support nb <= {L.PERK_NB_MAX} bands")
    if optimizer == "newton":
        from .newton import solve_newton
        if kernel.dtype != torch.complex128 and _precision(kernel, config) != "fp64":
            warnings.warn("optimizer='newton' needs float64 precision to converge reliably; the kernel is "
                          "float32. Build the kernel in complex128 (or pass precision='fp64'). The CG "
                          "optimizer works in float32.", stacklevel=2)
        return solve_newton(kernel, P0, n_electrons, config, fock_build_fn, deflation_on)
    return _solve_cg(kernel, P0, n_electrons, config, fock_build_fn)


solve = solve_direct_minimization
