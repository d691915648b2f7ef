"""Batched independent solves: a density / field scan in shared launches.

The reference runs a scan as a serial Python loop of ``solve`` calls
(``examples/multilayer_graphene_density_scan.py:117-137``); each call on a
small mesh leaves most of the GPU idle (a 36x36 mesh is 1296 per-k
wavefronts and 256 exchange workgroups).  ``solve_batch`` stacks the solves
that share a mesh, band count, temperature and interaction: one launch per
kernel serves all of them --

* the per-k CG kernels run over the stacked k-points, each k-point reading
  the state vector of its own solve (``seg_len``);
* the exchange runs all solves' channel planes in one launch with the shared
  VR (``hfx_planar_pack`` + ``hfx_selfenergy_planar``, or
  ``hfx_selfenergy_batched`` on the (k, i, j) layout), the epilogue writes
  one energy per solve;
* reductions, the line-search scalar steps and the mu solver are segmented
  per solve (``hfx_reduce_batched``, ``hfx_solve_mu_batched``);
* a solve that meets its stopping rule is frozen (its state vector's flag):
  its kernels skip their updates while the others continue;
* the iteration is captured once as HIP graphs and replayed, one readback
  of all state vectors per iteration.

The algorithm is ``solve``'s CG (optimizer "cg", Gauss-Jordan Cayley line
search) per solve; with ``precision="mixed"`` each solve is then polished by
the FP64 DIIS SCF (serially).  Solves that cannot be stacked (Hartree,
contact terms, a reference density, a user ``fock_build_fn``, block specs,
projection) run through ``solve`` one by one.
"""

from __future__ import annotations

from dataclasses import replace

import numpy as np
import torch

from . import _lib as L
from . import _ops as K
from . import fock as fock_mod
from ._util import KIND, capture_graphs, convert, to_tensor
from .solver import SolveResult, SolverConfig, _precision

__all__ = ["solve_batch"]


def _batchable(kernel, cfg):
    g = kernel._contact_host[0]
    contact = g is not None and bool(np.any(np.asarray(g) != 0.0))
    return (not kernel.include_hartree and kernel.include_exchange and kernel._refP is None and not contact
            and cfg.block_sizes is None and cfg.project_fn is None
            and str(cfg.optimizer).lower() == "cg" and str(cfg.cayley).lower() == "gauss_jordan"
            and kernel.h.shape[-1] <= L.PERK_NB_MAX)


def _group_key(kernel):
    return (tuple(kernel.h.shape), kernel.dtype, float(kernel.T), kernel._v_real_symmetric,
            tuple(kernel._VR_shifted.shape))


class _BatchFock:
    """Exchange-only Fock build + energy of S stacked solves sharing mesh and VR
    (fock.py:76-111 per solve); the energies go to st[s][HFX_ST_E]."""

    def __init__(self, kernels, cdt):
        k0 = kernels[0]
        self.dev = dev = k0.device
        nk1, nk2, nb = k0.h.shape[0], k0.h.shape[1], k0.h.shape[-1]
        self.S = S = len(kernels)
        self.plan = L.plan_for(nk1, nk2, nb, cdt, dev)
        self.herm = bool(k0._v_real_symmetric)
        VR = k0._VR_shifted
        self.VR = VR if VR.dtype == cdt else convert(VR, cdt)
        self.vfull = 0 if VR.shape[-1] == 1 else 1
        self.h = torch.stack([k.h if k.h.dtype == cdt else convert(k.h, cdt) for k in kernels]).contiguous()
        self.w_raw = torch.stack([k._w64 for k in kernels]).contiguous()
        c = dict(dtype=cdt, device=dev)
        self.planar = fock_mod.PLANAR
        if self.planar:
            nch = nb * (nb + 1) // 2 if self.herm else nb * nb
            self.Pt = torch.empty((S, nch, nk1, nk2), **c)
            self.St = torch.empty((S, nch, nk1, nk2), **c)
            self.vmap = None
            if self.vfull:
                self.VRc, self.vmap = fock_mod._channel_major_vr(self.plan, self.VR, (nk1, nk2, nb))
            else:
                self.VRc = self.VR.contiguous()
            if fock_mod.AUTOTUNE:
                self.plan.tune_once(self.VRc, bool(self.vfull), vr_map=self.vmap)
        else:
            self.Sig = torch.zeros((S, nk1, nk2, nb, nb), **c)
        self.epi = torch.zeros(256 * S, dtype=torch.float64, device=dev)   # energy partials, 256 per solve

    def __call__(self, P, F, st):
        sm = L.stream(self.dev)
        h = self.plan.handle
        if self.planar:
            herm = int(self.herm)
            L.call("hfx_planar_pack", h, self.S, L.ptr(P), None, None, 0, herm, L.ptr(self.Pt), sm)
            L.call("hfx_selfenergy_planar", h, self.S, L.ptr(self.Pt), L.ptr(self.VRc), self.vfull,
                   L.ptr(self.vmap), None, 0, herm, 0, None, L.ptr(self.St), None, sm)
            L.call("hfx_fock_epilogue_planar_batched", h, self.S, L.ptr(self.h), L.ptr(self.St), herm, L.ptr(P),
                   L.ptr(self.w_raw), L.ptr(F), L.ptr(st), L.ptr(self.epi), self.epi.numel(), sm)
            return
        L.call("hfx_selfenergy_batched", h, self.S, L.ptr(P), L.ptr(self.VR), self.vfull,
               int(self.herm), L.ptr(self.Sig), sm)
        L.call("hfx_fock_epilogue_batched", h, self.S, L.ptr(self.h), L.ptr(self.Sig), L.ptr(P),
               L.ptr(self.w_raw), L.ptr(F), L.ptr(st), L.ptr(self.epi), self.epi.numel(), sm)


class _BatchCG:
    """Stacked CG state of S solves on one (nk1, nk2, nb) mesh."""

    def __init__(self, kernels, cdt, cfg):
        k0 = kernels[0]
        self.kernels = kernels
        self.cfg = cfg
        self.cdt = cdt
        self.dev = dev = k0.device
        nk1, nk2, nb = k0.h.shape[0], k0.h.shape[1], k0.h.shape[-1]
        self.S = S = len(kernels)
        self.Nk, self.nb = nk1 * nk2, nb
        self.nk_tot = S * self.Nk
        self.shape4 = (S, nk1, nk2, nb, nb)
        self.fk = _BatchFock(kernels, cdt)
        self.w_norm = torch.stack([k._w64 / max(k.weight_sum, 1e-30) for k in kernels]).contiguous()
        self.T = max(float(k0.T), 1e-12)
        c = dict(dtype=cdt, device=dev)
        r = dict(dtype=torch.float64, device=dev)
        # every buffer is written before it is read (run() clears the CG history
        # terms), so none is filled at allocation: at 13 solves of 90^2 x 16 one
        # stacked matrix is 0.4 GB
        e = lambda: torch.empty(self.shape4, **c)  # noqa: E731
        self.Q, self.P, self.F, self.Ft, self.G, self.Gp = e(), e(), e(), e(), e(), e()
        self.H, self.D, self.V = e(), e(), e()
        v = lambda: torch.empty((S, nk1, nk2, nb), **r)  # noqa: E731
        self.p, self.eps, self.gp, self.gpp, self.hp, self.dp, self.lam = v(), v(), v(), v(), v(), v(), v()
        self.st = torch.zeros((S, K.ST_LEN), **r)
        self.part = torch.empty((8, self.nk_tot), **r)
        self.scratch = torch.zeros(int(L.lib().hfx_batch_scratch_doubles(S)), **r)
        self.targets = torch.zeros(S, **r)
        self.st_host = torch.empty((S, K.ST_LEN), dtype=torch.float64, pin_memory=True)
        self.ev = torch.cuda.Event()

    # -- pieces ---------------------------------------------------------------
    def set_targets(self, n_electrons):
        t = [float(n) / max(k.weight_sum, 1e-30) for n, k in zip(n_electrons, self.kernels)]
        self.targets.copy_(torch.tensor(t, dtype=torch.float64))

    def density(self):
        K.density(self.Q, self.P, p=self.p)
        return self.P

    def fock(self, P, F):
        self.fk(P, F, self.st)

    # mu steps per mu solve inside the captured iteration after the bracket pass
    # (which takes the first step); a solve that needs more is finished after
    # the iteration's readback (its done flag is in st)
    MU_LOOP_STEPS = 4

    # mu steps of an eager (not captured) solve before the done flags are read
    MU_EAGER_STEPS = 6

    def _mu_launch(self, n):
        L.call("hfx_solve_mu_batched", L.ptr(self.eps), L.ptr(self.w_norm), L.ptr(self.targets), self.nk_tot,
               self.Nk, self.nb, self.T, int(n), L.ptr(self.st), L.ptr(self.p),
               L.ptr(self.scratch), self.scratch.numel(), L.stream(self.dev))

    def mu(self, steps=None):
        """Per-solve mu: ``steps`` steps inside a captured iteration, else an eager
        solve to convergence within mu_maxiter steps (the remaining steps are
        launched only when a done flag is still clear after the first few)."""
        n = self.cfg.mu_maxiter if steps is None else min(steps, self.cfg.mu_maxiter)
        if steps is None and n > self.MU_EAGER_STEPS:
            self._mu_launch(self.MU_EAGER_STEPS)
            sv = self.st.view(-1, K.ST_LEN).cpu().numpy()
            if np.all((sv[:, K.ST_MUDONE] != 0.0) | (sv[:, K.ST_FROZEN] != 0.0)):
                return
            n -= self.MU_EAGER_STEPS
        self._mu_launch(n)

    def reduce(self, nval, wmask, op, params, out_slot=-1):
        vals = list(params) + [0.0] * (10 - len(params))
        import ctypes as C
        arr = (C.c_double * 10)(*[float(x) for x in vals])
        L.call("hfx_reduce_batched", L.ptr(self.part), int(nval), self.nk_tot, self.Nk, L.ptr(self.w_norm),
               int(wmask), L.ptr(self.st), int(op), C.cast(arr, C.c_void_p), int(out_slot), L.ptr(self.scratch),
               self.scratch.numel(), L.stream(self.dev))

    def init_from(self, P0s):
        """solver.py:380-394 per solve: eigh(F[herm(P0)]) -> (Q, eps), mu, p."""
        P0 = torch.stack([to_tensor(P, self.dev, self.cdt) for P in P0s]).contiguous()
        K.hermitize(P0, self.P)
        self.fock(self.P, self.F)
        K.eigh(self.F, self.eps, self.Q, herm_in=True)
        self.st.zero_()
        self.mu()

    def _swap(self):
        self.G, self.Gp = self.Gp, self.G
        self.gp, self.gpp = self.gpp, self.gp

    def _body(self, shrink_pow):
        cfg = self.cfg
        nk, nb, seg = self.nk_tot, self.nb, self.Nk
        dt = K.dt_code(self.Q)
        st, part = self.st, self.part
        sm = L.stream(self.dev)
        P = self.density()                                                      # solver.py:462
        self.fock(P, self.F)                                                    # :463-471
        L.call("hfx_cg_grad", dt, nb, nk, L.ptr(self.Q), L.ptr(self.F), L.ptr(self.p), L.ptr(self.Ft),
               L.ptr(self.eps), L.ptr(self.G), L.ptr(part), sm)                 # :472-478
        self.reduce(2, 0b01, K.OP_GRAD, [self.T, cfg.denom_scale, seg * nb])    # :491-497
        L.call("hfx_cg_precond", dt, nb, nk, L.ptr(self.G), L.ptr(self.Gp), L.ptr(self.eps), L.ptr(self.p),
               L.ptr(self.gpp), L.ptr(st), self.T, L.ptr(self.H), L.ptr(self.gp), L.ptr(self.hp),
               L.ptr(part), seg, sm)                                            # :486, 499-504
        self.reduce(6, 0b111111, K.OP_BETA, [0, 0, 0, 0, 0, 0, 0, 0, 0, max(int(cfg.cg_restart), 1)])
        L.call("hfx_cg_direction_gj", dt, nb, nk, L.ptr(self.H), L.ptr(self.D), L.ptr(self.hp),
               L.ptr(self.dp), L.ptr(st), L.ptr(part), seg, sm)
        self.reduce(1, 1, K.OP_TAU0, [0, 0, 0, cfg.max_step, 0, shrink_pow])     # :543-547
        for i in range(cfg.bt_max):                                             # :563-579
            L.call("hfx_cg_trial_gj", dt, nb, nk, L.ptr(self.D), L.ptr(self.Ft), L.ptr(self.p),
                   L.ptr(self.dp), L.ptr(self.eps), L.ptr(st), int(i == 0), L.ptr(part),
                   L.ptr(self.V), L.ptr(self.lam), seg, sm)
            self.reduce(4 if i == 0 else 2, 0b1111, K.OP_TRIAL,
                        [self.T, 0, 0, 0, cfg.bt_shrink, shrink_pow, 0, int(i == cfg.bt_max - 1), int(i == 0)])
        L.call("hfx_cg_retract_gj", dt, nb, nk, L.ptr(self.Q), L.ptr(self.D), L.ptr(self.Ft), L.ptr(st),
               L.ptr(self.eps), L.ptr(self.V), L.ptr(self.lam), seg, sm)      # :586-590
        self.mu(self.MU_LOOP_STEPS)                                             # :591-592
        self._swap()

    def _capture(self, shrink_pow):
        """Two graphs of the body, one per parity of the buffer swap it makes."""
        return capture_graphs(self.dev, lambda: self._body(shrink_pow), n=2)

    def run(self, tol_E, tol_grad, max_iter, rel_tol=False):
        """solver.py:422-449 per solve; a solve meeting its rule is frozen."""
        return _drive(self.run_steps(tol_E, tol_grad, max_iter, rel_tol))

    def run_steps(self, tol_E, tol_grad, max_iter, rel_tol=False):
        """run() as a generator that yields once per iteration, after queueing
        the iteration and the asynchronous readback of the state vectors and
        before waiting for them (the caller may queue other work meanwhile)."""
        cfg = self.cfg
        S = self.S
        self.G.zero_()
        self.Gp.zero_()
        self.D.zero_()
        self.gp.zero_()
        self.gpp.zero_()
        self.dp.zero_()
        self.st[:, K.ST_ITER] = 0.0
        self.st[:, K.ST_FROZEN] = 0.0
        shrink_pow = cfg.bt_shrink ** cfg.bt_max
        hist_E = [[] for _ in range(S)]
        hist_G = [[] for _ in range(S)]
        k_of = [0] * S
        grad = [np.inf] * S
        dE = [np.inf] * S
        E_prev = [np.inf] * S
        active = [True] * S
        graphs, g_base, k = None, 0, 0
        W = cfg.plateau_window
        while True:
            newly = []
            for s in range(S):
                if not active[s]:
                    continue
                hE = hist_E[s]
                tE = tol_E * max(1.0, abs(hE[-1])) if (rel_tol and hE) else tol_E
                ks = k_of[s]
                if W > 0:
                    e_now = hE[max(ks - 1, 0)] if hE else 0.0
                    e_past = hE[max(ks - 1 - W, 0)] if hE else 0.0
                    enc = (e_past - e_now) > tE if ks > W else dE[s] > tE
                else:
                    enc = dE[s] > tE
                nc = grad[s] > tol_grad if tol_grad > 0.0 else enc
                if not (ks < max_iter and nc):
                    active[s] = False
                    newly.append(s)
            if newly:
                self.st[newly, K.ST_FROZEN] = 1.0
            if not any(active):
                break
            if graphs is None and bool(cfg.use_graph) and k >= 1:
                graphs = self._capture(shrink_pow)
                g_base = k
            if graphs is not None:
                graphs[(k - g_base) % 2].replay()
                self._swap()
            else:
                self._body(shrink_pow)
            self.st_host.copy_(self.st, non_blocking=True)
            self.ev.record()
            yield
            self.ev.synchronize()
            sv = self.st_host.numpy()
            if any(active[s] and sv[s, K.ST_MUDONE] == 0.0 for s in range(S)):
                self.mu()                 # finish the mu solves the loop's steps left open
            for s in range(S):
                if not active[s]:
                    continue
                E = float(sv[s, K.ST_E])
                grad[s] = float(sv[s, K.ST_GRAD])
                hist_E[s].append(E)
                hist_G[s].append(grad[s])
                dE[s] = abs(E - E_prev[s])
                E_prev[s] = E
                k_of[s] += 1
            k += 1
        self.G, self.Gp = self.Gp, self.G
        self.st[:, K.ST_FROZEN] = 0.0
        return k_of, grad, hist_E, hist_G

    def finalize(self):
        """solver.py:606-638 for every solve: Fock, eigh canonicalisation, mu, p, Fock + E."""
        P = self.density()
        self.fock(P, self.F)
        K.matmul(self.Q, self.F, self.Ft, C=self.Q, ops=0b001)
        K.eigh(self.Ft, self.eps, self.V, herm_in=True)
        K.matmul(self.Q, self.V, self.H)
        self.Q, self.H = self.H, self.Q
        self.mu()
        P = self.density()
        Pout = P.clone()
        F = torch.empty_like(Pout)
        self.fock(Pout, F)
        return Pout, F


# state-vector slots the batched polish reduces into (unused by the CG layout)
_SL_RES = 20    # [20] sum w |FP - PF|^2, [21] sum w |P - P_prev|^2
_SL_B = 22      # [22 ...] the new row of the DIIS B matrix (= _SL_RES + 2: one reduction fills both)
_SL_CHK = 28    # the density residual |density(F_k) - P_k|^2 of the check (in the check's state copy)


class _BatchPolish:
    """FP64 C-DIIS polish of S stacked solves: the batched form of the mixed
    solve's polish (one eigh per iteration in a running basis,
    reference_scf.py:292-314 per solve).  Each solve keeps its own B matrix and
    coefficients (solved on the device, one thread per solve, and read per
    solve by the extrapolation kernel) and its own stopping test; a converged
    solve's iterate is copied out and it keeps iterating with the others.  As
    in the single polish, the update from iterate k is launched before the
    host has read iterate k's residuals."""

    MU_STEPS = 4     # mu steps per mu solve in an update after the bracket pass; more are run on the readback
    EIGH_TOL_PER_TOL = 1e-5   # update eigh tolerance per unit of the polish tolerance (as in solve)
    # the single polish's stopping test: once a solve's commutator residual is
    # within the tolerance, its density residual |density(F_k) - P_k| (the
    # reference's, reference_scf.py:287-294) is formed directly instead of
    # waiting one iteration for the step residual
    ROOTHAAN_CHECK = True
    MU_CHECK_STEPS = 6

    def __init__(self, kernels, n_electrons, nh=6):
        k0 = kernels[0]
        cdt = torch.complex128
        self.kernels = kernels
        self.dev = dev = k0.device
        nk1, nk2, nb = k0.h.shape[0], k0.h.shape[1], k0.h.shape[-1]
        self.S = S = len(kernels)
        self.Nk, self.nb, self.nk_tot = nk1 * nk2, nb, S * nk1 * nk2
        self.shape4 = (S, nk1, nk2, nb, nb)
        self.fk = _BatchFock(kernels, cdt)
        self.nh = nh
        c = dict(dtype=cdt, device=dev)
        r = dict(dtype=torch.float64, device=dev)
        # written before read (history slots are used only once filled), so not
        # filled at allocation: one stacked matrix is up to 0.4 GB here
        e = lambda: torch.empty(self.shape4, **c)  # noqa: E731
        self.P_prev, self.P_cur, self.P_new, self.Ft, self.U, self.tmp = e(), e(), e(), e(), e(), e()
        self.Pfin, self.Ffin = e(), e()
        self.F_hist = [e() for _ in range(nh)]
        # residual history in complex64 at nb = 16, as in the single polish
        rdt = torch.complex64 if nb == 16 else cdt
        self.R_hist = [torch.empty(self.shape4, dtype=rdt, device=dev) for _ in range(nh)]
        self.eps = torch.empty((S, nk1, nk2, nb), **r)
        self.occ = torch.empty_like(self.eps)
        self.st = torch.zeros((S, K.ST_LEN), **r)
        self.part = torch.empty((8, self.nk_tot), **r)
        self.scratch = torch.zeros(int(L.lib().hfx_batch_scratch_doubles(S)), **r)
        self.w_norm = torch.stack([k._w64 / max(k.weight_sum, 1e-30) for k in kernels]).contiguous()
        self.targets = torch.tensor([float(n) / max(k.weight_sum, 1e-30) for n, k in zip(n_electrons, kernels)],
                                    dtype=torch.float64, device=dev)
        self.T = max(float(k0.T), 1e-12)
        self.coef = torch.zeros((S, 8), **r)
        self.Bd = torch.zeros((S, nh, nh), **r)
        self.st_host = torch.empty((S, K.ST_LEN), dtype=torch.float64, pin_memory=True)
        self.ev = torch.cuda.Event()
        # the density-residual check's own orbitals, eigenvalues, occupations and state
        self.U_chk = e()
        self.eps_chk, self.occ_chk = torch.empty_like(self.eps), torch.empty_like(self.eps)
        self.st_chk = torch.zeros_like(self.st)
        self.chk_host = torch.empty((S, K.ST_LEN), dtype=torch.float64, pin_memory=True)
        import ctypes as C
        self._xa = (C.c_void_p * nh)(*[f.data_ptr() for f in self.F_hist])

    def _reduce(self, nval, slot, st=None):
        L.call("hfx_reduce_batched", L.ptr(self.part), int(nval), self.nk_tot, self.Nk, L.ptr(self.w_norm),
               (1 << nval) - 1, L.ptr(self.st if st is None else st), K.OP_SUM, None, int(slot),
               L.ptr(self.scratch), self.scratch.numel(), L.stream(self.dev))

    def _mu(self, maxiter=54, eps=None, st=None, occ=None):
        L.call("hfx_solve_mu_batched", L.ptr(self.eps if eps is None else eps), L.ptr(self.w_norm),
               L.ptr(self.targets), self.nk_tot, self.Nk, self.nb, self.T, int(maxiter),
               L.ptr(self.st if st is None else st), L.ptr(self.occ if occ is None else occ), L.ptr(self.scratch),
               self.scratch.numel(), L.stream(self.dev))

    def _check(self, idx, Pk):
        """Per solve |density(F_k) - P_k| (w-weighted as the residuals), one
        readback: the eigh of every solve's F_k in the running basis into U_chk,
        mu and occupations in a copy of the state (frozen solves skip), the
        density with its distance to P_k.  -> (squared residuals, mu done)."""
        K.extrap_rotate(self.Ft, self.U, [self.F_hist[idx]], [1.0])
        K.eigh_basis(self.Ft, self.eps_chk, self.U, herm_in=True, out=self.U_chk)
        self.st_chk.copy_(self.st)
        self._mu(self.MU_CHECK_STEPS, eps=self.eps_chk, st=self.st_chk, occ=self.occ_chk)
        K.density_diff(self.U_chk, self.occ_chk, Pk, self.tmp, self.part)
        self._reduce(1, _SL_CHK, st=self.st_chk)
        self.chk_host.copy_(self.st_chk, non_blocking=True)
        self.ev.record()
        self.ev.synchronize()
        sv = self.chk_host.numpy()
        return sv[:, _SL_CHK].copy(), sv[:, K.ST_MUDONE] != 0.0

    def _eval(self, k):
        """F, E, residuals and B row of every solve's iterate k; asynchronous readback."""
        nh = self.nh
        idx, count = k % nh, min(k + 1, nh)
        Fk, Rk = self.F_hist[idx], self.R_hist[idx]
        self.fk(self.P_cur, Fk, self.st)
        # residuals and the B row (reference_scf.py:110-111) in one pass: slots _SL_RES, _SL_B
        K.scf_resid_dots(Fk, self.P_cur, self.P_prev if k > 0 else None, Rk, self.R_hist[:count], idx, self.part)
        self._reduce(2 + count, _SL_RES)
        self.st_host.copy_(self.st, non_blocking=True)
        self.ev.record()

    def _advance(self, k, warm):
        """Update from iterate k for every solve (device DIIS coefficients), then
        P_prev <- P_k, P_cur <- P_{k+1}."""
        import ctypes as C
        nh = self.nh
        idx, count = k % nh, min(k + 1, nh)
        L.call("hfx_diis_coefficients_batched", K.st_ptr(self.st, _SL_B), K.ST_LEN, L.ptr(self.Bd), self.S, nh,
               idx, count, L.ptr(self.coef), L.stream(self.dev))
        if warm:
            L.call("hfx_extrap_rotate_batched", K.dt_code(self.U), self.nb, self.nk_tot, self.Nk, count,
                   C.cast(self._xa, C.c_void_p), L.ptr(self.coef), L.ptr(self.U), L.ptr(self.Ft),
                   L.stream(self.dev))
            K.eigh_basis(self.Ft, self.eps, self.U, herm_in=True, rel_tol=self._eigh_tol)   # U <- U W
            self._mu(self.MU_STEPS)
            K.density(self.U, self.P_new, p=self.occ)
        else:                             # cold first update: eigh of herm(F_0) (count = 1: plain step)
            K.lincomb_multi(self.Ft, [self.F_hist[idx]], [1.0], herm=True)
            K.eigh(self.Ft, self.eps, self.U, herm_in=True)
            self._mu(self.MU_STEPS)
            K.density(self.U, self.tmp, p=self.occ)
            K.hermitize(self.tmp, self.P_new)
        self.P_prev, self.P_cur = self.P_cur, self.P_prev
        K.lincomb(self.P_cur, self.P_new, 1.0)

    def run(self, P_lo, mu0, tol, max_iter, Q_lo=None):
        """Q_lo (optional): the stacked orbitals P_lo was built from; re-orthonormalised
        in FP64 they are the first basis (the first update is a warm eigh)."""
        return _drive(self.run_steps(P_lo, mu0, tol, max_iter, Q_lo))

    def run_steps(self, P_lo, mu0, tol, max_iter, Q_lo=None):
        """run() as a generator: yields before each wait for a readback."""
        S, nh = self.S, self.nh
        max_iter = max(int(max_iter), 1)
        self._eigh_tol = self.EIGH_TOL_PER_TOL * float(tol)
        K.hermitize(convert(P_lo, torch.complex128) if P_lo.dtype != torch.complex128 else P_lo, self.P_cur)
        self.st.zero_()
        self.st[:, K.ST_MU] = mu0
        warm0 = Q_lo is not None
        if warm0:
            Qc = Q_lo.contiguous()
            L.call("hfx_convert", L.ptr(Qc), KIND[Qc.dtype], L.ptr(self.U), KIND[torch.complex128], Qc.numel(),
                   L.stream(self.dev))
            K.orthonormalize(self.U, 2)
        done = [False] * S
        hE = [[] for _ in range(S)]
        hC = [[] for _ in range(S)]
        hD = [[] for _ in range(S)]
        E_out = [float("nan")] * S
        k, redo, spec = 0, -1, False
        self._eval(0)
        while True:
            idx = k % nh
            Pk = self.P_cur
            spec = k + 1 < max_iter
            if spec:                      # wait instead when every open solve is predicted to stop at k
                preds = []
                for s in range(S):
                    if done[s]:
                        continue
                    c, d = hC[s], hD[s]
                    if len(c) >= 2 and c[-2] > 0 and 0 < d[-2] < float("inf"):
                        # with the density-residual check a predicted commutator stop suffices
                        preds.append(c[-1] * c[-1] / c[-2] <= 2.0 * tol
                                     and (self.ROOTHAAN_CHECK or d[-1] * d[-1] / d[-2] <= 2.0 * tol))
                    else:
                        preds.append(False)
                spec = not (preds and all(preds))
            if spec:
                self._advance(k, warm=k > 0 or warm0)
            yield
            self.ev.synchronize()
            sv = self.st_host.numpy()
            bad_mu = k > 0 and redo != k and any(not done[s] and sv[s, K.ST_MUDONE] == 0.0 for s in range(S))
            if bad_mu:
                # an update's mu needed more steps than it had: finish mu on the newest
                # eigenvalues and rebuild those densities (P_{k+1} after the
                # speculative update, else P_k, evaluated again)
                self._mu()
                K.density(self.U, self.tmp, p=self.occ)
                K.hermitize(self.tmp, self.P_cur)
                if not spec:
                    redo = k
                    self._eval(k)
                    continue
            for s in range(S):
                if done[s]:
                    continue
                c_res = float(np.sqrt(max(sv[s, _SL_RES], 0.0)))
                d_res = float(np.sqrt(max(sv[s, _SL_RES + 1], 0.0))) if k > 0 else float("inf")
                E_out[s] = float(sv[s, K.ST_E])
                hE[s].append(E_out[s])
                hC[s].append(c_res)
                hD[s].append(d_res)
                if not bad_mu and c_res <= tol and d_res <= tol:     # this iterate is the result
                    done[s] = True
                    self.Pfin[s].copy_(Pk[s])
                    self.Ffin[s].copy_(self.F_hist[idx][s])
            ask = [s for s in range(S) if not done[s] and k > 0 and hC[s] and hC[s][-1] <= tol]
            if self.ROOTHAAN_CHECK and ask and not bad_mu:
                d2, mu_ok = self._check(idx, Pk)
                for s in ask:
                    d_ref = float(np.sqrt(max(d2[s], 0.0)))
                    if mu_ok[s] and d_ref <= tol:
                        hD[s][-1] = d_ref
                        done[s] = True
                        self.Pfin[s].copy_(Pk[s])
                        self.Ffin[s].copy_(self.F_hist[idx][s])
            if all(done) or k + 1 >= max_iter:
                break
            if not spec:
                self._advance(k, warm=k > 0 or warm0)
            k += 1
            self._eval(k)
        for s in range(S):
            if not done[s]:               # budget spent: the last evaluated iterate
                self.Pfin[s].copy_(Pk[s])
                self.Ffin[s].copy_(self.F_hist[idx][s])
        # orbitals / occupations / mu of the returned states (solver.py:606-638 form)
        if warm0 or k > 0 or spec:        # U holds a basis: warm eigh in it
            K.extrap_rotate(self.Ft, self.U, [self.Ffin], [1.0])
            K.eigh_basis(self.Ft, self.eps, self.U, herm_in=True)      # U <- U W
            self._mu()
        else:                             # every solve stopped before an update
            K.eigh(self.Ffin, self.eps, self.U, herm_in=True)
            self._mu()
            K.density(self.U, self.tmp, p=self.occ)
        return E_out, done, hE, hC


def _drive(gen):
    """Run a step generator to its end; its return value."""
    try:
        while True:
            next(gen)
    except StopIteration as e:
        return e.value


def _interleave(gens, dev, nstreams):
    """Run step generators side by side, each on its own stream (round robin
    over ``nstreams`` streams of default priority): each resume queues one
    iteration of that generator and its readback, so while the host waits for
    one mesh group the others' iterations are on the GPU."""
    cur = torch.cuda.current_stream(dev)
    if len(gens) == 1 or nstreams <= 1:
        for g in gens:
            _drive(g)
        return
    streams = [torch.cuda.Stream(dev) for _ in range(min(nstreams, len(gens)))]
    for sm in streams:
        sm.wait_stream(cur)             # the inputs were written on the caller's stream
    live = [(g, streams[i % len(streams)]) for i, g in enumerate(gens)]
    try:
        while live:
            nxt = []
            for g, sm in live:
                with torch.cuda.stream(sm):
                    try:
                        next(g)
                        nxt.append((g, sm))
                    except StopIteration:
                        pass
            live = nxt
    finally:
        for sm in streams:
            cur.wait_stream(sm)


def _same_vr(a, b):
    """Exact VR equality of two kernels, remembered on the kernels (a kernel's
    VR is fixed at construction) so that repeated scans compare on the host."""
    ta, tb = a.__dict__.get("_hfx_vr_token"), b.__dict__.get("_hfx_vr_token")
    if ta is not None and ta is tb:
        return True
    if a._VR_shifted.shape != b._VR_shifted.shape or a._VR_shifted.dtype != b._VR_shifted.dtype:
        return False
    if not torch.equal(a._VR_shifted, b._VR_shifted):
        return False
    tok = ta if ta is not None else (tb if tb is not None else object())
    a.__dict__["_hfx_vr_token"] = b.__dict__["_hfx_vr_token"] = tok
    return True


def _histories(rows, m, dev):
    """(S, m) device tensor of per-solve histories (zero-padded), one copy."""
    H = np.zeros((len(rows), m), np.float64)
    for j, r in enumerate(rows):
        H[j, :len(r)] = r[:m]
    return torch.from_numpy(H).to(dev)


def _keep_for(stream, *tensors):
    """The caller's stream uses these after the call: tell the allocator."""
    for t in tensors:
        t.record_stream(stream)


def _group_mixed(kernels, members, P0s, n_electrons, cfg, results, caller):
    """FP32 stacked CG then the FP64 stacked polish of one mesh group (steps)."""
    ks = [kernels[i] for i in members]
    ne = [n_electrons[i] for i in members]
    # the FP32 phase's line search as in solve: at most mixed_bt_max trials
    b = _BatchCG(ks, torch.complex64, replace(cfg, bt_max=min(cfg.bt_max, max(int(cfg.mixed_bt_max), 1))))
    b.set_targets(ne)
    b.init_from([P0s[i] for i in members])
    k_of, grad, hE, hG = yield from b.run_steps(cfg.mixed_tol_E, 0.0, max(cfg.max_iter // 2, 1), rel_tol=True)
    P_lo = b.density()
    Q_lo = b.Q
    mu0 = b.st[:, K.ST_MU].clone()
    del b                                            # free the FP32 stack before the FP64 one
    pol = _BatchPolish(ks, ne)
    E2, conv2, hE2, hC2 = yield from pol.run_steps(P_lo, mu0, cfg.polish_tol, cfg.polish_max_iter, Q_lo=Q_lo)
    del Q_lo, P_lo
    dev = pol.dev
    n_it = [k_of[j] + len(hE2[j]) for j in range(len(members))]
    m = max([cfg.max_iter] + n_it)
    hEt = _histories([list(hE[j]) + list(hE2[j]) for j in range(len(members))], m, dev)
    hGt = _histories([list(hG[j]) + list(hC2[j]) for j in range(len(members))], m, dev)
    Et = torch.tensor(E2, dtype=torch.float64, device=dev)
    _keep_for(caller, pol.U, pol.occ, pol.st, pol.Pfin, pol.Ffin, Et, hEt, hGt)
    for j, i in enumerate(members):
        results[i] = SolveResult(Q=pol.U[j], p=pol.occ[j], mu=pol.st[j, K.ST_MU], density=pol.Pfin[j],
                                 fock=pol.Ffin[j], energy=Et[j], n_iter=int(n_it[j]), converged=bool(conv2[j]),
                                 history=dict(E=hEt[j], grad_norm=hGt[j],
                                              phase_iters=torch.tensor([k_of[j], len(hE2[j])], dtype=torch.int64)))


def _group_cg(kernels, members, P0s, n_electrons, cfg, cdt, results, caller):
    """Stacked CG of one mesh group to the reference stopping rule (steps)."""
    ks = [kernels[i] for i in members]
    b = _BatchCG(ks, cdt, cfg)
    b.set_targets([n_electrons[i] for i in members])
    b.init_from([P0s[i] for i in members])
    k_of, grad, hE, hG = yield from b.run_steps(cfg.tol_E, cfg.tol_grad, cfg.max_iter)
    Pfin, Ffin = b.finalize()
    dev = b.dev
    m = max([cfg.max_iter] + list(k_of))
    hEt = _histories(hE, m, dev)
    hGt = _histories(hG, m, dev)
    _keep_for(caller, b.Q, b.p, b.st, Pfin, Ffin, hEt, hGt)
    for j, i in enumerate(members):
        kk = k_of[j]
        conv = (grad[j] <= cfg.tol_grad) if cfg.tol_grad > 0.0 else (kk < cfg.max_iter)
        results[i] = SolveResult(Q=b.Q[j], p=b.p[j], mu=b.st[j, K.ST_MU], density=Pfin[j], fock=Ffin[j],
                                 energy=b.st[j, K.ST_E], n_iter=int(kk), converged=bool(conv),
                                 history=dict(E=hEt[j], grad_norm=hGt[j]))


def solve_batch(kernels, P0s, n_electrons, *, config: SolverConfig | None = None, streams: int = 4):
    """Independent CG solves of a scan, stacked into shared launches.

    ``kernels``, ``P0s`` and ``n_electrons`` are sequences of equal length
    (``P0s`` / ``n_electrons`` may also be a single density / count used for
    all).  Returns one ``SolveResult`` per solve, in order; the result
    tensors of one mesh group are views of that group's stacked buffers.
    Solves are stacked per (mesh, bands, dtype, T, interaction); the mesh
    groups run side by side on up to ``streams`` streams, their host loops
    interleaved; the rest go through ``solve``.
    """
    from .solver import solve
    from .utils import validate_kernel_electron_count
    cfg = config or SolverConfig()
    n = len(kernels)
    if not isinstance(P0s, (list, tuple)):
        P0s = [P0s] * n
    if np.ndim(n_electrons) == 0:
        n_electrons = [float(n_electrons)] * n
    if len(P0s) != n or len(n_electrons) != n:
        raise ValueError("kernels, P0s and n_electrons must have the same length")
    for kern, ne in zip(kernels, n_electrons):
        validate_kernel_electron_count(kern, ne, context="n_electrons")
    results = [None] * n
    groups: dict = {}
    for i, kern in enumerate(kernels):
        if not _batchable(kern, cfg):
            results[i] = solve(kern, P0s[i], n_electrons[i], config=cfg)
            continue
        key = _group_key(kern)
        placed = False
        for gkey, members in groups.items():
            if gkey[0] == key and _same_vr(kernels[members[0]], kern):
                members.append(i)
                placed = True
                break
        if not placed:
            groups[(key, i)] = [i]
    if not groups:
        return results
    dev = kernels[next(iter(groups.values()))[0]].device
    caller = torch.cuda.current_stream(dev)
    gens = []
    for members in groups.values():
        prec = _precision(kernels[members[0]], cfg)
        if prec == "mixed":
            gens.append(_group_mixed(kernels, members, P0s, n_electrons, cfg, results, caller))
        else:
            cdt = torch.complex128 if prec == "fp64" else torch.complex64
            gens.append(_group_cg(kernels, members, P0s, n_electrons, cfg, cdt, results, caller))
    _interleave(gens, dev, int(streams))
    return results
