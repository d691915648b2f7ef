"""Hand-written adjoint of build_fock / hf_energy and converged-state gradients.

Bars: E is quadratic in P and linear in h and VR, so central differences are
exact up to rounding -- gradients must match them to 1e-9 relative (FP64);
they must also match torch autograd of a CPU float64 restatement of the
reference formulas (the stand-in for jax.grad of hf_energy, which cannot run
here).  Stationarity gradients dOmega/d(U0, q0, u) at a converged FP64 solve
must match central finite differences of re-converged free energies (step
1e-5) to 1e-6 relative.
"""

import numpy as np
import pytest

from jax_hf_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def _torch_energy_cpu(P, h, VR, w, refP=None):
    """CPU torch restatement of fock.py:51-64, 94, 109-111 (exchange only)."""
    import torch
    dP = P if refP is None else P - refP
    S = -torch.fft.ifftn(torch.fft.fftn(dP, dim=(0, 1)) * VR, dim=(0, 1))
    A = w[..., None, None] * (h + 0.5 * S)
    return torch.real(torch.einsum("klij,klji->", A, P))


def _rand_herm(rng, shape):
    X = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return 0.5 * (X + np.conj(np.swapaxes(X, -1, -2)))


@pytest.mark.parametrize("layer", [False, True])
def test_adjoint_matches_finite_differences_and_torch_autograd(dev, layer):
    import torch
    from jax_hf_amd.adjoint import hf_energy_autograd, vr_from_vq
    rng = np.random.default_rng(2)
    nk, nb = 12, 4
    w = rng.uniform(0.5, 1.5, (nk, nk)) / nk ** 2
    V = syn.coulomb_layer(nk, nk, nb, 2) if layer else syn.coulomb_scalar(nk, nk)
    P = syn.haar_density(nk, nk, nb, rng)
    h = syn.tight_binding(nk, nk, nb, 5)
    VRt = vr_from_vq(V, w, dev)
    Pt = torch.from_numpy(P).to(dev).requires_grad_(True)
    ht = torch.from_numpy(h).to(dev).requires_grad_(True)
    VRv = VRt.clone().requires_grad_(True)
    E = hf_energy_autograd(Pt, ht, VRv, w)
    gP, gh, gV = torch.autograd.grad(E, (Pt, ht, VRv))

    def Ev(Pm, hm, Vm):
        with torch.no_grad():
            return float(hf_energy_autograd(Pm, hm, Vm, w))

    t = 1e-4
    for name, base, g in (("P", Pt, gP), ("h", ht, gh), ("VR", VRv, gV)):
        D = torch.from_numpy(rng.standard_normal(tuple(base.shape)) + 1j * rng.standard_normal(tuple(base.shape))).to(dev)
        args_p = [Pt.detach(), ht.detach(), VRv.detach()]
        args_m = [Pt.detach(), ht.detach(), VRv.detach()]
        i = {"P": 0, "h": 1, "VR": 2}[name]
        args_p[i] = args_p[i] + t * D
        args_m[i] = args_m[i] - t * D
        fd = (Ev(*args_p) - Ev(*args_m)) / (2 * t)
        an = float(torch.sum(torch.real(torch.conj(g) * D)))
        assert abs(fd - an) <= 1e-9 * max(1.0, abs(an)), (name, fd, an)
    # a non-unit upstream gradient (read on the device by the backward) scales every gradient
    E2 = hf_energy_autograd(Pt, ht, VRv, w) * 2.5
    for a, b in zip(torch.autograd.grad(E2, (Pt, ht, VRv)), (gP, gh, gV)):
        assert torch.allclose(a, 2.5 * b, rtol=1e-14, atol=0)
    # torch autograd of the CPU restatement (jax.grad stand-in)
    Pc = torch.from_numpy(P).requires_grad_(True)
    hc = torch.from_numpy(h).requires_grad_(True)
    Vc = VRt.detach().cpu().clone().requires_grad_(True)
    Ec = _torch_energy_cpu(Pc, hc, Vc, torch.from_numpy(w))
    assert abs(float(Ec) - float(E)) < 1e-12 * max(1, abs(float(Ec)))
    rP, rh, rV = torch.autograd.grad(Ec, (Pc, hc, Vc))
    for a, b in ((gP, rP), (gh, rh), (gV, rV)):
        a = a.detach().cpu().numpy()
        b = b.detach().numpy()
        assert np.abs(a - b).max() <= 1e-11 * max(1.0, np.abs(b).max())


def test_adjoint_hartree_term(dev):
    import torch
    from jax_hf_amd.adjoint import hf_energy_autograd, vr_from_vq
    rng = np.random.default_rng(4)
    nk, nb = 6, 3
    w = rng.uniform(0.5, 1.5, (nk, nk))
    P = syn.haar_density(nk, nk, nb, rng)
    ref = syn.haar_density(nk, nk, nb, rng)
    h = syn.tight_binding(nk, nk, nb, 1)
    HH = rng.uniform(0.1, 0.5, (nb, nb))
    VR = vr_from_vq(syn.coulomb_scalar(nk, nk), w, dev)
    Pt = torch.from_numpy(P).to(dev).requires_grad_(True)
    reft = torch.from_numpy(ref).to(dev)
    E = hf_energy_autograd(Pt, torch.from_numpy(h).to(dev), VR, w, refP=reft, HH=HH)
    (gP,) = torch.autograd.grad(E, (Pt,))
    D = torch.from_numpy(_rand_herm(rng, P.shape)).to(dev)
    t = 1e-4

    def Ev(Pm):
        with torch.no_grad():
            return float(hf_energy_autograd(Pm, torch.from_numpy(h).to(dev), VR, w, refP=reft, HH=HH))

    fd = (Ev(Pt.detach() + t * D) - Ev(Pt.detach() - t * D)) / (2 * t)
    an = float(torch.sum(torch.real(torch.conj(gP) * D)))
    assert abs(fd - an) <= 1e-9 * max(1.0, abs(an))


def _layer_setup(nk, nb, U0, q0, d_, seed):
    h0 = syn.tight_binding(nk, nk, nb, seed)
    D = syn.displacement(nb, 1.0)
    w = syn.uniform_weights(nk, nk)
    q = syn.qnorm_min_image(nk, nk)[..., None, None]
    lay = syn.layer_of_band(nb, 2)
    dl = np.abs(lay[:, None] - lay[None, :])[None, None]

    def Vq(U, q0_):
        return U * np.exp(-q * d_ * dl) / (q + q0_)

    return h0, D, w, q, Vq


def test_explicit_parameter_gradients_at_fixed_density(dev):
    """d E(P; theta)/d theta at fixed P (U0, q0 through V(q); u through h) against
    central differences of the GPU energy: the adjoint's parameter part alone."""
    import jax_hf_amd as J
    from jax_hf_amd.adjoint import free_energy_gradients
    nk, nb, T = 12, 4, 0.1
    U0, q0, d_, u = 5.0, 0.3, 0.5, 0.2
    h0, D, w, q, Vq = _layer_setup(nk, nb, U0, q0, d_, 4)
    P = syn.haar_density(nk, nk, nb, np.random.default_rng(9))

    def energy(U, q0_, uu):
        G = J.HartreeFockKernel(w, h0 + uu * D, Vq(U, q0_), T)
        a = G.as_args()
        S, H, F = J.build_fock(P, h=a["h"], VR=a["VR"], refP=a["refP"], HH=a["HH"], w2d=G.w2d,
                               include_exchange=True, include_hartree=False,
                               exchange_hermitian_channel_packing=a["exchange_hermitian_channel_packing"])
        return float(J.hf_energy(P, h=G.h, Sigma=S, H=H, weights_b=G.weights_b)), G

    _, G = energy(U0, q0, u)
    grads = free_energy_gradients(G, P, dVq={"U0": Vq(U0, q0) / U0, "q0": -Vq(U0, q0) / (q + q0)},
                                  dh={"u": D})
    step = 1e-5
    fds = {"U0": (energy(U0 + step, q0, u)[0] - energy(U0 - step, q0, u)[0]) / (2 * step),
           "q0": (energy(U0, q0 + step, u)[0] - energy(U0, q0 - step, u)[0]) / (2 * step),
           "u": (energy(U0, q0, u + step)[0] - energy(U0, q0, u - step)[0]) / (2 * step)}
    for k, fd in fds.items():
        assert abs(grads[k] - fd) <= 1e-8 * max(abs(fd), 1e-3), (k, grads[k], fd)


def test_stationary_parameter_gradients_vs_finite_differences(dev):
    """dOmega/dtheta of the converged free energy (Hellmann-Feynman at an SCF
    fixed point, [F, P] ~ 1e-13) against central differences of re-converged
    solves, 1e-6 relative."""
    import jax_hf_amd as J
    from jax_hf_amd.adjoint import free_energy_gradients
    nk, nb, T = 12, 4, 0.1
    U0, q0, d_, u = 5.0, 0.3, 0.5, 0.2
    h0, D, w, q, Vq = _layer_setup(nk, nb, U0, q0, d_, 4)
    cfg = J.SCFConfig(max_iter=300, acceleration="diis", density_tol=1e-12, comm_tol=1e-11)

    def omega(U, q0_, uu):
        G = J.HartreeFockKernel(w, h0 + uu * D, Vq(U, q0_), T)
        r = J.solve_scf(G, np.zeros_like(h0), nb / 2.0, config=cfg)
        assert r.converged
        P = _np(r.density_matrix)
        p = np.linalg.eigvalsh(0.5 * (P + np.conj(np.swapaxes(P, -1, -2))))
        return float(J.free_energy(float(r.energy), p, w, T)), G, P

    _, G, P = omega(U0, q0, u)
    grads = free_energy_gradients(G, P, dVq={"U0": Vq(U0, q0) / U0, "q0": -Vq(U0, q0) / (q + q0)},
                                  dh={"u": D})
    step = 1e-5
    fds = {
        "U0": (omega(U0 + step, q0, u)[0] - omega(U0 - step, q0, u)[0]) / (2 * step),
        "q0": (omega(U0, q0 + step, u)[0] - omega(U0, q0 - step, u)[0]) / (2 * step),
        "u": (omega(U0, q0, u + step)[0] - omega(U0, q0, u - step)[0]) / (2 * step),
    }
    for k, fd in fds.items():
        assert abs(grads[k] - fd) <= 1e-6 * max(abs(fd), 1e-3), (k, grads[k], fd)
