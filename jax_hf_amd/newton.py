"""Trust-region Newton direct minimisation on the GPU (solver.py:686-1078).

Same algorithm as ``_solve_newton_impl``: Steihaug truncated CG on the joint
(Q, p) Hessian with one Fock-response build per Hessian-vector product
(``F[dP] - h`` with refP = 0), block preconditioner, Cayley retraction with
tau = -1, trust-region ratio with the precision-scaled noise floor, the cold
start occupation relaxation and the optional deflation bias.  Every array op
is a libhfx kernel.  The Steihaug loop keeps its scalars on the device (the
last workgroup of each inner-product reduction forms alpha, the boundary
step, the stop test and the update coefficients -- the HFX_OP_NT* post-ops --
and ``hfx_lincomb_dev`` applies them), so an inner step
is launches only: it is captured as a HIP graph and replayed, and the host
reads one stop flag per step, one step behind (the step after the stop is a
no-op on the loop's state).  The outer trust-region logic runs on the host
from a few reduced scalars per outer iteration.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from . import _ops as K
from ._util import KIND, capture_graphs, convert, to_tensor

__all__ = ["solve_newton"]

# device scalars of the truncated CG (HFX_NT_* of include/hfx.h)
NT_RZ, NT_DELTA, NT_GRAD, NT_ALPHA, NT_TBD, NT_NEG, NT_DONE = 0, 1, 2, 3, 4, 5, 6
NT_CV, NT_CR, NT_CDA, NT_CDB, NT_NALPHA, NT_STEPS = 7, 8, 9, 10, 11, 12


class _Newton:
    def __init__(self, kernel, cfg, fock_build_fn, has_defl):
        from .solver import _Fock
        self.kernel = kernel
        self.cfg = cfg
        self.dev = kernel.device
        self.cdt = torch.complex128 if str(cfg.precision).lower() == "fp64" else kernel.dtype
        nk1, nk2, nb = kernel.h.shape[0], kernel.h.shape[1], kernel.h.shape[-1]
        self.shape4 = (nk1, nk2, nb, nb)
        self.shape3 = (nk1, nk2, nb)
        self.nk, self.nb = nk1 * nk2, nb
        self.fock = _Fock(kernel, self.cdt, fock_build_fn, cfg.block_sizes)
        # precision="mixed" on a complex128 kernel: the Fock-response builds of the
        # Hessian-vector products run in complex64 (FP32 FFT and products, the
        # bulk of the work); gradients, energies, mu, the Steihaug subproblem and
        # every other Fock build stay FP64
        self.hvp_fock = None
        if (str(cfg.precision).lower() == "mixed" and self.cdt == torch.complex128 and fock_build_fn is None
                and cfg.block_sizes is None):
            self.hvp_fock = _Fock(kernel, torch.complex64, None, None)
            z = lambda: torch.zeros((nk1, nk2, nb, nb), dtype=torch.complex64, device=self.dev)  # noqa: E731
            self.dP32, self.R32 = z(), z()
        self.w2d = kernel._w64.reshape(-1).contiguous()
        self.wsum = float(kernel.weight_sum)
        self.w_norm = (self.w2d / max(self.wsum, 1e-30)).contiguous()
        self.ones = torch.ones_like(self.w2d)
        self.T = max(kernel.T, 1e-12)
        self.ws = K.Workspace.get(self.dev, self.nk)
        self.kind = KIND[self.cdt]
        self.dt = K.dt_code(torch.empty(0, dtype=self.cdt))
        self.st = torch.zeros(K.ST_LEN, dtype=torch.float64, device=self.dev)
        self.out = torch.zeros(32, dtype=torch.float64, device=self.dev)
        self.nt = torch.zeros(16, dtype=torch.float64, device=self.dev)        # HFX_NT_* scalars
        self.nt_in = torch.zeros(2, dtype=torch.float64, pin_memory=True)      # Delta, grad
        self.flags = torch.zeros(max(64, int(cfg.tr_cg_max)), dtype=torch.float64,
                                 pin_memory=True)                               # stop flag per inner step
        self._tr_graphs = {}
        self.zero_ref = torch.zeros(self.shape4, dtype=self.cdt, device=self.dev)
        self.has_defl = has_defl
        if has_defl:
            t = to_tensor(cfg.deflation_targets, self.dev, self.cdt)
            if t.dim() == 4:
                t = t[None]
            if t.dim() != 5 or tuple(t.shape[1:]) != self.shape4:
                raise ValueError("deflation_targets must have shape (n_solutions, nk1, nk2, nb, nb) matching "
                                 f"the kernel; got {tuple(t.shape)}.")
            self.targets = [t[i].contiguous() for i in range(t.shape[0])]
            self.inv2L2 = 1.0 / (2.0 * max(float(cfg.deflation_length), 1e-12) ** 2)
            self.dsig = float(cfg.deflation_sigma)

    # ---- small helpers ----------------------------------------------------
    def cz(self):
        return torch.zeros(self.shape4, dtype=self.cdt, device=self.dev)

    def rz(self):
        return torch.zeros(self.shape3, dtype=torch.float64, device=self.dev)

    def sums(self, nval, w, wmask, off=0):
        K.reduce(self.ws, self.ws.part, nval, self.nk, w=w, wmask=wmask, out=self.out[off:off + nval])

    def jdots(self, pairs, wx=0, phase=None):
        """Reduce the inner products of n pairs of (X, p) vectors into out[0:2n],
        one fused pass and one fixed-order sum: X parts at out[i] (weighted by
        w2d for the bits of wx), p parts weighted by w2d at out[n + i]; with
        ``phase`` the sum's workgroup also takes that Steihaug scalar step;
        launches only."""
        import ctypes as C
        n = len(pairs)
        keep = [(C.c_void_p * n)(*[t.data_ptr() for t in ts])
                for ts in ([a[0] for a, _ in pairs], [b[0] for _, b in pairs],
                           [a[1] for a, _ in pairs], [b[1] for _, b in pairs])]
        wmask = int(wx) | (((1 << n) - 1) << n)
        op = K.OP_SUM if phase is None else K.OP_NT0 + int(phase)
        L.call("hfx_pair_dots", self.dt, self.nb, self.nk, n, *[C.cast(a, C.c_void_p) for a in keep],
               L.ptr(self.w2d), wmask, L.ptr(self.out), None if phase is None else L.ptr(self.nt), op,
               L.ptr(self.ws.scratch), L.stream(self.dev))

    def nt_ptr(self, idx):
        return None if idx is None else self.nt.data_ptr() + 8 * int(idx)

    def axpy(self, out, a, b, ca=1.0, ca_dev=None, cb=1.0, cb_dev=None):
        """(X, p) pair out = ca' a + cb' b, the coefficients scaled by the device
        scalars nt[ca_dev] / nt[cb_dev] when given; one launch."""
        (oX, op), (aX, ap), (bX, bp) = out, a, b
        L.call("hfx_lincomb_dev", oX.numel(), KIND[oX.dtype], L.ptr(aX), L.ptr(bX), L.ptr(oX),
               op.numel(), KIND[op.dtype], L.ptr(ap), L.ptr(bp), L.ptr(op),
               float(ca), self.nt_ptr(ca_dev), float(cb), self.nt_ptr(cb_dev), L.stream(self.dev))

    def axpys(self, ops):
        """Several axpy updates (out, a, b, ca, ca_dev, cb, cb_dev) of (X, p) pairs
        in one launch, applied in order at each element."""
        import ctypes as C
        n = len(ops)
        keep = [(C.c_void_p * n)(*[o[a][part].data_ptr() for o in ops])
                for part in (0, 1) for a in (1, 2, 0)]                  # X1, Y1, out1, X2, Y2, out2
        keep += [(C.c_double * n)(*[float(o[3]) for o in ops]),
                 (C.c_int * n)(*[-1 if o[4] is None else int(o[4]) for o in ops]),
                 (C.c_double * n)(*[float(o[5]) for o in ops]),
                 (C.c_int * n)(*[-1 if o[6] is None else int(o[6]) for o in ops])]
        oX, op_ = ops[0][0]
        L.call("hfx_lincomb_dev_batch", n, oX.numel(), KIND[oX.dtype], op_.numel(), KIND[op_.dtype],
               *[C.cast(a, C.c_void_p) for a in keep], L.ptr(self.nt), L.stream(self.dev))

    def tr_step(self, B, Q, p, Ft, eps, pen):
        """One step of the Steihaug truncated CG (solver.py:947-991), launches only:
        the Hessian-vector product, alpha and the boundary step, the trial point
        and residual, the preconditioned residual, the stop test and the updates
        of v, Hv, r and d with coefficients from the device."""
        dX, dp, HdX, Hdp = B["dX"], B["dp"], B["HdX"], B["Hdp"]
        v, Hv, r = (B["vX"], B["vp"]), (B["HvX"], B["Hvp"]), (B["rX"], B["rp"])
        va, ra, za = (B["vXa"], B["vpa"]), (B["rXa"], B["rpa"]), (B["zXa"], B["zpa"])
        d, Hd = (dX, dp), (HdX, Hdp)
        self.hvp(dX, dp, HdX, Hdp, Q, p, Ft, pen, B["dP"], B["R"])
        self.jdots([(d, Hd), (d, d), (v, d), (v, v)], phase=0)
        self.axpys([(va, v, d, 1.0, None, 1.0, NT_ALPHA),        # v + alpha d
                    (ra, r, Hd, 1.0, None, 1.0, NT_NALPHA)])     # r - alpha H d
        self.prec(ra[0], ra[1], za[0], za[1], eps, p)
        self.jdots([(va, va), (ra, ra), (ra, za)], phase=1)
        # in this order at each element: d is read by the first op and written by the last
        self.axpys([(v, v, d, 1.0, None, 1.0, NT_CV),            # boundary step or v + alpha d
                    (Hv, Hv, Hd, 1.0, None, 1.0, NT_CV),
                    (r, r, Hd, 1.0, None, 1.0, NT_CR),
                    (d, za, d, 1.0, NT_CDA, 1.0, NT_CDB)])       # d = z + beta d (unchanged once stopped)

    def project_occ(self, vec):
        """vec -= sum_k w_k sum_b vec / (sum w nb) (solver.py:224-233); part holds
        the per-k sums; one launch."""
        L.call("hfx_vec_shift_sum", L.ptr(vec), vec.numel(), L.ptr(self.ws.part), L.ptr(self.w2d), self.nk,
               1.0 / (self.wsum * self.nb), L.stream(self.dev))

    # ---- model pieces ------------------------------------------------------
    def eval_state(self, Q, p, Ft, G, eps, P, F):
        """solver.py:814-837 -> (E, grad, pen, lam); fills Ft, G, eps."""
        K.density(Q, P, p=p)
        if self.cfg.project_fn is not None:
            X = to_tensor(self.cfg.project_fn(P), self.dev, self.cdt)
            K.hermitize(X, P)
        self.fock.build(P, F, E_ptr=K.st_ptr(self.st, K.ST_E))
        Phi, pen = 0.0, 0.0
        if self.has_defl:
            Phi, pen = self.deflation(P, F)
        L.call("hfx_cg_grad", self.dt, self.nb, self.nk, L.ptr(Q), L.ptr(F), L.ptr(p), L.ptr(Ft), L.ptr(eps),
               L.ptr(G), L.ptr(self.ws.part), L.stream(self.dev))
        K.reduce(self.ws, self.ws.part, 2, self.nk, w=self.w_norm, wmask=0b01, st=self.st, op=K.OP_GRAD,
                 params=[self.T, self.cfg.denom_scale, self.nk * self.nb])
        sv = self.st.cpu().numpy()
        self.last_mu = float(sv[K.ST_MU])            # the mu this state's occupations were solved with
        return float(sv[K.ST_E]) + Phi, float(sv[K.ST_GRAD]), pen, float(sv[K.ST_LAM])

    def deflation(self, P, F):
        """Gaussian bias around previous solutions (solver.py:799-812); F += S_pen."""
        diffs = []
        for t in self.targets:
            d = self.cz()
            K.lincomb(d, P, 1.0, t, -1.0)
            diffs.append(d)
        d2 = []
        for j0 in range(0, len(diffs), 8):
            blk = diffs[j0:j0 + 8]
            K.dots(blk, blk, self.ws.part, mode=0)
            self.sums(len(blk), self.w2d, (1 << len(blk)) - 1, 0)
            d2 += list(self.out[:len(blk)].cpu().numpy())
        Phi, pen = 0.0, 0.0
        for d, dd in zip(diffs, d2):
            phi = self.dsig * np.exp(-dd * self.inv2L2)
            phip = -phi * self.inv2L2
            K.lincomb(F, F, 1.0, d, 2.0 * phip)
            Phi += phi
            pen += 2.0 * phip
        return float(Phi), float(pen)

    def prec(self, rX, rp, zX, zp, eps, p):
        L.call("hfx_newton_prec", self.dt, self.nb, self.nk, L.ptr(rX), L.ptr(rp), L.ptr(eps), L.ptr(p),
               L.ptr(self.st), self.T, L.ptr(zX), L.ptr(zp), L.ptr(self.ws.part), L.stream(self.dev))
        self.project_occ(zp)

    def hvp(self, X, dp, HX, Hp, Q, p, Ft, pen, dP, R):
        """solver.py:916-930."""
        L.call("hfx_newton_dP", self.dt, self.nb, self.nk, L.ptr(Q), L.ptr(X), L.ptr(p), L.ptr(dp), L.ptr(dP),
               L.stream(self.dev))
        # F[dP] - h with refP = 0 (solver.py:770-780); the engine takes None as "no reference"
        if self.hvp_fock is not None:
            n, st = dP.numel(), L.stream(self.dev)
            L.call("hfx_convert", L.ptr(dP), KIND[dP.dtype], L.ptr(self.dP32), KIND[torch.complex64], n, st)
            self.hvp_fock.build(self.dP32, self.R32, refP=None, sub_h=True)
            L.call("hfx_convert", L.ptr(self.R32), KIND[torch.complex64], L.ptr(R), KIND[R.dtype], n, st)
        else:
            self.fock.build(dP, R, refP=None if self.fock.fn is None else self.zero_ref, sub_h=True)
        if self.has_defl and pen != 0.0:
            K.lincomb(R, R, 1.0, dP, pen)
        L.call("hfx_newton_hvp", self.dt, self.nb, self.nk, L.ptr(Q), L.ptr(R), L.ptr(Ft), L.ptr(X), L.ptr(p),
               L.ptr(dp), self.T, L.ptr(HX), L.ptr(Hp), L.ptr(self.ws.part), L.stream(self.dev))
        self.project_occ(Hp)

    # ---- main loop ---------------------------------------------------------
    def run(self, P0, n_electrons):
        cfg = self.cfg
        tiny = 1e-30
        n_norm = float(n_electrons) / max(self.wsum, tiny)
        tol_g = float(cfg.tol_grad) if cfg.tol_grad > 0.0 else 1e-6
        delta_min = 1e-10
        feps = np.finfo(np.float64 if self.cdt == torch.complex128 else np.float32).eps
        noise_floor = max(1e-12, 4500.0 * float(feps))
        mu_it = cfg.mu_maxiter
        cz, rz = self.cz, self.rz
        Q, Ft, G, P, F = cz(), cz(), cz(), cz(), cz()
        p, eps = rz(), rz()
        Qt, Ftt, Gt = cz(), cz(), cz()
        pt, epst = rz(), rz()
        dP, R = cz(), cz()

        def mu_solve(e, pout, from_state, mu0=0.0):
            K.solve_mu(self.ws, e, self.w_norm, self.nk, self.nb, n_norm, self.T, mu0=mu0,
                       mu0_from_state=from_state, maxiter=mu_it, mode=K.MU_NEWTON_BRACKET, st=self.st, p_out=pout)

        P0h = cz()
        K.hermitize(to_tensor(P0, self.dev, self.cdt), P0h)                 # :840-848
        self.fock.build(P0h, F)
        K.eigh(F, eps, Q)
        mu_solve(eps, p, False, 0.0)
        mu = float(self.st[K.ST_MU])
        E, grad, pen, lam = self.eval_state(Q, p, Ft, G, eps, P, F)
        if grad <= tol_g:                                                    # :858-877
            for _ in range(30):
                K.density(Q, P, p=p)
                if cfg.project_fn is not None:
                    K.hermitize(to_tensor(cfg.project_fn(P), self.dev, self.cdt), P)
                self.fock.build(P, F)
                L.call("hfx_cg_grad", self.dt, self.nb, self.nk, L.ptr(Q), L.ptr(F), L.ptr(p), L.ptr(Ftt),
                       L.ptr(epst), L.ptr(Gt), L.ptr(self.ws.part), L.stream(self.dev))
                mu_solve(epst, p, True)
            mu = float(self.st[K.ST_MU])
            E, grad, pen, lam = self.eval_state(Q, p, Ft, G, eps, P, F)

        hE, hG = [], []
        k = 0
        Delta = float(cfg.tr_delta0)
        B = {k_: cz() for k_ in ("rX", "zX", "dX", "vX", "HvX", "HdX", "vXa", "rXa", "zXa")}
        B.update({k_: rz() for k_ in ("rp", "zp", "dp", "vp", "Hvp", "Hdp", "vpa", "rpa", "zpa")})
        B["dP"], B["R"] = dP, R
        rX, rp_, zX, zp = B["rX"], B["rp"], B["zX"], B["zp"]
        dX, dpv, vX, vp, HvX, Hvp = B["dX"], B["dp"], B["vX"], B["vp"], B["HvX"], B["Hvp"]
        use_graph = (bool(getattr(cfg, "use_graph", True)) and not self.has_defl
                     and self.fock.graph_safe() and (self.hvp_fock is None or self.hvp_fock.graph_safe()))
        evs = [torch.cuda.Event() for _ in range(cfg.tr_cg_max)]
        if use_graph and self.hvp_fock is not None:
            # first build of the complex64 engine outside the capture (its workspaces)
            self.hvp_fock.build(self.dP32, self.R32, refP=None, sub_h=True)
        while k < cfg.max_iter and grad > tol_g and Delta > delta_min:      # :886-891
            hE.append(E)
            hG.append(grad)
            self.st[K.ST_LAM] = lam                                         # preconditioner of this state
            K.lincomb(rX, G, -1.0)                                          # :933-943
            rp_.zero_()
            self.prec(rX, rp_, zX, zp, eps, p)
            dX.copy_(zX)
            dpv.copy_(zp)
            self.jdots([((rX, rp_), (zX, zp))], phase=2)                   # rz, not stopped
            self.nt_in[0], self.nt_in[1] = Delta, grad
            self.nt[NT_DELTA:NT_DELTA + 2].copy_(self.nt_in, non_blocking=True)
            vX.zero_(); vp.zero_(); HvX.zero_(); Hvp.zero_()                # noqa: E702
            graph = None
            if use_graph:
                key = Q.data_ptr()
                graph = self._tr_graphs.get(key)
                if graph is None:
                    graph = capture_graphs(self.dev, lambda: self.tr_step(B, Q, p, Ft, eps, pen))[0]
                    self._tr_graphs[key] = graph
            for cg in range(cfg.tr_cg_max):                                 # :947-991
                if graph is not None:
                    graph.replay()
                else:
                    self.tr_step(B, Q, p, Ft, eps, pen)
                self.flags[cg:cg + 1].copy_(self.nt[NT_DONE:NT_DONE + 1], non_blocking=True)
                evs[cg].record()
                # the stop flag of the previous step (this one is then a no-op)
                if cg >= 1:
                    evs[cg - 1].synchronize()
                    if float(self.flags[cg - 1]) != 0.0:
                        break

            # predicted reduction, all w2d-weighted (:998-1000), and the step norm: one
            # reduction, one readback (G.v needs no p part: the p part of the pair is v.v)
            self.jdots([((G, vp), (vX, vp)), ((vX, vp), (HvX, Hvp)), ((vX, vp), (vX, vp))], wx=0b011)
            o = self.out.cpu().numpy()
            gv, vHv = float(o[0]), float(o[1] + o[4])
            pred = -(gv + 0.5 * vHv)
            step_norm2 = float(o[2] + o[5])
            # retraction Q <- Q Cayley(tau=-1, vX) and frozen-F trial occupations (:1005-1011):
            # the Cayley factor by Gauss-Jordan elimination (the same unitary as the
            # spectral form of solver.py:140-197, without diagonalising i vX)
            Qt.copy_(Q)
            self.st[K.ST_TAUF] = -1.0
            L.call("hfx_cg_retract_gj", self.dt, self.nb, self.nk, L.ptr(Qt), L.ptr(vX), L.ptr(Ft), L.ptr(self.st),
                   L.ptr(epst), None, None, 0, L.stream(self.dev))
            self.st[K.ST_MU] = mu
            mu_solve(epst, pt, True)
            E_t, g_t, pen_t, lam_t = self.eval_state(Qt, pt, Ftt, Gt, epst, P, F)   # :1014
            mu_t = self.last_mu
            noise = noise_floor * max(abs(E), 1.0)                          # :1021-1046
            if pred <= noise:
                accept = (g_t <= grad) and (E_t <= E + noise)
                Delta_new = Delta if accept else 0.5 * Delta
            else:
                ratio = (E - E_t) / pred
                sn = float(np.sqrt(max(0.0, step_norm2)))
                if ratio < 0.25:
                    Delta_new = 0.25 * Delta
                elif ratio > 0.75 and sn > 0.9 * Delta:
                    Delta_new = min(2.0 * Delta, 5.0)
                else:
                    Delta_new = Delta
                accept = ratio > 0.1
            if accept:
                Q, Qt = Qt, Q
                p, pt = pt, p
                Ft, Ftt = Ftt, Ft
                G, Gt = Gt, G
                eps, epst = epst, eps
                E, mu, grad, pen, lam = E_t, mu_t, g_t, pen_t, lam_t
            Delta = Delta_new
            k += 1

        K.density(Q, P, p=p)                                                 # :1062-1066
        if cfg.project_fn is not None:
            K.hermitize(to_tensor(cfg.project_fn(P), self.dev, self.cdt), P)
        Pfin = P.clone()
        Ffin = cz()
        self.fock.build(Pfin, Ffin, E_ptr=K.st_ptr(self.st, K.ST_E))
        E_fin = float(self.st[K.ST_E])
        return Q, p, mu, Pfin, Ffin, E_fin, k, grad <= tol_g, hE, hG


def solve_newton(kernel, P0, n_electrons, cfg, fock_build_fn, has_defl):
    from .solver import SolveResult
    runner = _Newton(kernel, cfg, fock_build_fn, has_defl)
    Q, p, mu, P, F, E, k, conv, hE, hG = runner.run(P0, n_electrons)
    dev = kernel.device
    hEt = torch.zeros(cfg.max_iter, dtype=torch.float64)
    hGt = torch.zeros(cfg.max_iter, dtype=torch.float64)
    if hE:
        hEt[:len(hE)] = torch.tensor(hE, dtype=torch.float64)
        hGt[:len(hG)] = torch.tensor(hG, dtype=torch.float64)
    return SolveResult(Q=Q, p=p, mu=torch.tensor(mu, dtype=torch.float64, device=dev), density=P, fock=F,
                       energy=torch.tensor(E, dtype=torch.float64, device=dev), n_iter=int(k),
                       converged=bool(conv), history=dict(E=hEt.to(dev), grad_norm=hGt.to(dev)))
