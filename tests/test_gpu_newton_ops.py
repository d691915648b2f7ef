"""GPU checks of the fused inner-step operations of the trust-region Newton
solver through the C-ABI, against NumPy: hfx_pair_dots (weighted inner
products of (matrix, vector) pairs and the Steihaug scalar step) and
hfx_lincomb_dev_batch (device-coefficient updates applied in order).

Bars: sums to 1e-12 relative (FP64) / 1e-6 (complex64 matrices); updates
exact to FP64 rounding."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _arr(ts, ct=C.c_void_p):
    return (ct * len(ts))(*ts)


@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("nk,nb,npair", [(37, 8, 4), (64, 16, 3), (5, 3, 1)])
def test_pair_dots(dev, dt, nk, nb, npair):
    import torch
    from jax_hf_amd import _lib as L
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(nk + nb)
    cdt = torch.complex128 if dt == "c128" else torch.complex64
    X = [rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb)) for _ in range(npair)]
    Y = [rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb)) for _ in range(npair)]
    vx = [rng.standard_normal((nk, nb)) for _ in range(npair)]
    vy = [rng.standard_normal((nk, nb)) for _ in range(npair)]
    w = rng.uniform(0.1, 1.0, nk)
    wmask = 0b101 | (((1 << npair) - 1) << npair)
    t = lambda a, d=cdt: torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(d)  # noqa: E731
    Xt, Yt = [t(x) for x in X], [t(y) for y in Y]
    vxt, vyt = [t(v, torch.float64) for v in vx], [t(v, torch.float64) for v in vy]
    if dt == "c64":       # the reference from the rounded inputs
        X = [x.astype(np.complex64).astype(np.complex128) for x in X]
        Y = [y.astype(np.complex64).astype(np.complex128) for y in Y]
    wt = t(w, torch.float64)
    out = torch.zeros(16, dtype=torch.float64, device=dev)
    ws = K.Workspace.get(dev, nk)
    keep = [_arr([a.data_ptr() for a in ts]) for ts in (Xt, Yt, vxt, vyt)]
    L.call("hfx_pair_dots", K.dt_code(Xt[0]), nb, nk, npair, *[C.cast(a, C.c_void_p) for a in keep], L.ptr(wt),
           wmask, L.ptr(out), None, K.OP_SUM, L.ptr(ws.scratch), L.stream(dev))
    got = out.cpu().numpy()
    tol = 1e-12 if dt == "c128" else 1e-6
    for v in range(npair):
        per_k = np.real(np.sum(np.conj(X[v]) * Y[v], axis=(1, 2)))
        ref = np.sum((w if (wmask >> v) & 1 else 1.0) * per_k)
        assert abs(got[v] - ref) <= tol * np.sum(np.abs(per_k)) + 1e-14
        ref_p = np.sum(w * np.sum(vx[v] * vy[v], axis=1))
        assert abs(got[npair + v] - ref_p) <= 1e-12 * np.sum(np.abs(vx[v] * vy[v])) + 1e-14


def test_pair_dots_rejects_bad_arguments(dev):
    import torch
    from jax_hf_amd import _lib as L
    from jax_hf_amd import _ops as K
    x = torch.zeros((2, 2, 2), dtype=torch.complex128, device=dev)
    p = torch.zeros((2, 2), dtype=torch.float64, device=dev)
    ws = K.Workspace.get(dev, 2)
    keep = [_arr([a.data_ptr()] * 5) for a in (x, x, p, p)]
    with pytest.raises(ValueError, match="hfx_pair_dots"):   # five pairs: above the limit of four
        L.call("hfx_pair_dots", K.dt_code(x), 2, 2, 5, *[C.cast(a, C.c_void_p) for a in keep], None, 0,
               None, None, K.OP_SUM, L.ptr(ws.scratch), L.stream(dev))


def test_lincomb_dev_batch_in_order(dev):
    """The four updates of a truncated-CG inner step: v += cv d, Hv += cv Hd,
    r += cr Hd, d = cda z + cdb d, with the coefficients on the device; the
    first op reads d before the last op rewrites it (same element)."""
    import torch
    from jax_hf_amd import _lib as L
    from jax_hf_amd._util import KIND
    rng = np.random.default_rng(3)
    nk, nb = 41, 8
    mk = lambda: torch.from_numpy(rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb))).to(dev)  # noqa: E731
    vk = lambda: torch.from_numpy(rng.standard_normal((nk, nb))).to(dev)  # noqa: E731
    v, Hv, r, d, Hd, z = [(mk(), vk()) for _ in range(6)]
    before = {n: (a.cpu().numpy().copy(), b.cpu().numpy().copy())
              for n, (a, b) in dict(v=v, Hv=Hv, r=r, d=d, Hd=Hd, z=z).items()}
    sc = torch.tensor([0.0, 0.37, -1.25, 0.5, 2.0], dtype=torch.float64, device=dev)
    ops = [(v, v, d, 1.0, None, 1.0, 1), (Hv, Hv, Hd, 1.0, None, 1.0, 1), (r, r, Hd, 1.0, None, 1.0, 2),
           (d, z, d, 1.0, 3, 1.0, 4)]
    keep = [(C.c_void_p * 4)(*[o[a][part].data_ptr() for o in ops]) for part in (0, 1) for a in (1, 2, 0)]
    keep += [(C.c_double * 4)(*[o[3] for o in ops]), (C.c_int * 4)(*[-1 if o[4] is None else o[4] for o in ops]),
             (C.c_double * 4)(*[o[5] for o in ops]), (C.c_int * 4)(*[o[6] for o in ops])]
    L.call("hfx_lincomb_dev_batch", 4, v[0].numel(), KIND[v[0].dtype], v[1].numel(), KIND[v[1].dtype],
           *[C.cast(a, C.c_void_p) for a in keep], L.ptr(sc), L.stream(dev))
    cv, cr, cda, cdb = 0.37, -1.25, 0.5, 2.0
    for part in (0, 1):
        B = {n: x[part] for n, x in before.items()}
        ref = dict(v=B["v"] + cv * B["d"], Hv=B["Hv"] + cv * B["Hd"], r=B["r"] + cr * B["Hd"],
                   d=cda * B["z"] + cdb * B["d"])
        for n, got in dict(v=v, Hv=Hv, r=r, d=d).items():
            assert np.abs(got[part].cpu().numpy() - ref[n]).max() <= 1e-14 * (1 + np.abs(ref[n]).max())
