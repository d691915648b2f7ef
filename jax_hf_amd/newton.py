"""Trust-region Newton direct minimisation on the GPU (solver.py:686-1078).

Same algorithm as ``_solve_newton_impl``: Steihaug truncated CG on the joint
(Q, p) Hessian with one Fock-response build per Hessian-vector product
(``F[dP] - h`` with refP = 0), block preconditioner, Cayley retraction with
tau = -1, trust-region ratio with the precision-scaled noise floor, the cold
start occupation relaxation and the optional deflation bias.  Every array op
is a libhfx kernel; the Steihaug control flow runs on the host from a few
reduced scalars per inner iteration.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from . import _ops as K
from ._util import KIND, convert, to_tensor

__all__ = ["solve_newton"]


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

    def jips(self, pairs):
        """[jip(aX, ap, bX, bp)] for pairs of (X, p) vectors: X unweighted, p weighted by w2d."""
        n = len(pairs)
        K.dots([a[0] for a, _ in pairs], [b[0] for _, b in pairs], self.ws.part, mode=0)
        self.sums(n, None, 0, 0)
        K.vdots([a[1] for a, _ in pairs], [b[1] for _, b in pairs], self.ws.part)
        self.sums(n, self.w2d, (1 << n) - 1, 8)
        o = self.out.cpu().numpy()
        return [float(o[i] + o[8 + i]) for i in range(n)]

    def project_occ(self, vec, part_written=True):
        """vec -= sum_k w_k sum_b vec / (sum w nb) (solver.py:224-233); part holds per-k sums."""
        K.reduce(self.ws, self.ws.part, 1, self.nk, w=self.w2d, wmask=1, out=self.out[16:17])
        L.call("hfx_vec_shift", L.ptr(vec), vec.numel(), self.out.data_ptr() + 16 * 8,
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
        dP, R, V, Fe, Dtmp = cz(), cz(), cz(), cz(), cz()
        lam_v, vtmp = rz(), rz()

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
        rX, rp_, zX, zp = cz(), rz(), cz(), rz()
        dX, dpv = cz(), rz()
        vX, vp, HvX, Hvp = cz(), rz(), cz(), rz()
        HdX, Hdp = cz(), rz()
        vXa, vpa, rXa, rpa = cz(), rz(), cz(), rz()
        zXa, zpa = cz(), rz()
        while k < cfg.max_iter and grad > tol_g and Delta > delta_min:      # :886-891
            hE.append(E)
            hG.append(grad)
            self.st[K.ST_LAM] = lam                                         # preconditioner of this state
            K.lincomb(rX, G, -1.0)                                          # :933-943
            rp_.zero_()
            self.prec(rX, rp_, zX, zp, eps, p)
            dX.copy_(zX)
            dpv.copy_(zp)
            (rzv,) = self.jips([((rX, rp_), (zX, zp))])
            vX.zero_(); vp.zero_(); HvX.zero_(); Hvp.zero_()                # noqa: E702
            cg = 0
            done = False
            while cg < cfg.tr_cg_max and not done:                          # :947-991
                self.hvp(dX, dpv, HdX, Hdp, Q, p, Ft, pen, dP, R)
                dHd, dd, vd, vv = self.jips([((dX, dpv), (HdX, Hdp)), ((dX, dpv), (dX, dpv)),
                                             ((vX, vp), (dX, dpv)), ((vX, vp), (vX, vp))])
                neg = dHd <= 1e-14 * dd
                b = 2.0 * vd
                cc = vv - Delta * Delta
                disc = max(b * b - 4.0 * dd * cc, 0.0)
                t_bd = (-b + np.sqrt(disc)) / (2.0 * dd + 1e-30)
                alpha = rzv / (1.0 if neg else dHd)
                K.lincomb(vXa, vX, 1.0, dX, alpha)
                K.lincomb(vpa, vp, 1.0, dpv, alpha)
                K.lincomb(rXa, rX, 1.0, HdX, -alpha)
                K.lincomb(rpa, rp_, 1.0, Hdp, -alpha)
                self.prec(rXa, rpa, zXa, zpa, eps, p)
                na2, rn2, rz_new = self.jips([((vXa, vpa), (vXa, vpa)), ((rXa, rpa), (rXa, rpa)),
                                              ((rXa, rpa), (zXa, zpa))])
                take_bd = neg or float(np.sqrt(max(0.0, na2))) >= Delta
                inner = float(np.sqrt(max(0.0, rn2))) < 0.05 * grad
                beta = rz_new / (rzv if abs(rzv) > tiny else 1.0)
                stop = take_bd or inner
                if take_bd:
                    K.lincomb(vX, vX, 1.0, dX, t_bd)
                    K.lincomb(vp, vp, 1.0, dpv, t_bd)
                    K.lincomb(HvX, HvX, 1.0, HdX, t_bd)
                    K.lincomb(Hvp, Hvp, 1.0, Hdp, t_bd)
                else:
                    vX, vXa = vXa, vX
                    vp, vpa = vpa, vp
                    K.lincomb(HvX, HvX, 1.0, HdX, alpha)
                    K.lincomb(Hvp, Hvp, 1.0, Hdp, alpha)
                    rX, rXa = rXa, rX
                    rp_, rpa = rpa, rp_
                if not stop:
                    K.lincomb(dX, zXa, 1.0, dX, beta)
                    K.lincomb(dpv, zpa, 1.0, dpv, beta)
                    rzv = rz_new
                done = stop
                cg += 1

            # predicted reduction, all w2d-weighted (:998-1000)
            K.dots([G, vX], [vX, HvX], self.ws.part, mode=0)
            self.sums(2, self.w2d, 0b11, 0)
            K.vdots([vp], [Hvp], self.ws.part)
            self.sums(1, self.w2d, 1, 8)
            o = self.out.cpu().numpy()
            gv, vHv = float(o[0]), float(o[1] + o[8])
            pred = -(gv + 0.5 * vHv)
            (step_norm2,) = self.jips([((vX, vp), (vX, vp))])
            # retraction Q <- Q Cayley(tau=-1, vX) and frozen-F trial occupations (:1005-1011)
            self.st[K.ST_BETA] = 0.0
            L.call("hfx_cg_direction", self.dt, self.nb, self.nk, L.ptr(vX), L.ptr(Dtmp), L.ptr(vtmp),
                   L.ptr(vtmp), L.ptr(Ft), L.ptr(self.st), L.ptr(V), L.ptr(lam_v), L.ptr(Fe),
                   L.ptr(self.ws.part), L.stream(self.dev))
            Qt.copy_(Q)
            L.call("hfx_cg_retract", self.dt, self.nb, self.nk, L.ptr(Qt), L.ptr(V), L.ptr(lam_v), L.ptr(Fe),
                   L.ptr(self.st), -1.0, 1, L.ptr(epst), None, L.stream(self.dev))
            self.st[K.ST_MU] = mu
            mu_solve(epst, pt, True)
            mu_t = float(self.st[K.ST_MU])
            E_t, g_t, pen_t, lam_t = self.eval_state(Qt, pt, Ftt, Gt, epst, P, F)   # :1014
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
