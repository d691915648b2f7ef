"""GPU checks of the polish helpers through the C-ABI, against NumPy:
hfx_density_diff (density of orbitals and its per-k distance to a reference
density), hfx_hermitize_convert (hermitise and change precision in one pass)
and hfx_orthonormalize_from (Newton-Schulz from complex64 orbitals into a
complex128 basis).  Bars: FP64 rounding (1e-13 relative), complex64 rounding
for the converted inputs."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("nb", [3, 16, 24])
def test_density_diff(dev, nb):
    import torch
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(nb)
    nk = 23
    Q = np.linalg.qr(rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb)))[0]
    p = rng.uniform(0, 1, (nk, nb))
    R = rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb))
    R = 0.5 * (R + np.conj(np.swapaxes(R, -1, -2)))
    t = lambda a, d=torch.complex128: torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(d)  # noqa: E731
    P = torch.empty((nk, nb, nb), dtype=torch.complex128, device=dev)
    part = torch.zeros(nk, dtype=torch.float64, device=dev)
    K.density_diff(t(Q), t(p, torch.float64), t(R), P, part)
    Pr = np.einsum("kin,kn,kjn->kij", Q, p, np.conj(Q))
    assert np.abs(_np(P) - Pr).max() < 1e-13
    Pg = _np(P)
    assert np.array_equal(Pg, np.conj(np.swapaxes(Pg, -1, -2)))         # exactly Hermitian
    d2 = np.sum(np.abs(Pr - R) ** 2, axis=(1, 2))
    assert np.abs(_np(part) - d2).max() < 1e-12 * d2.max()


@pytest.mark.parametrize("src,dst", [("c128", "c64"), ("c64", "c128"), ("c128", "c128")])
def test_hermitize_convert(dev, src, dst):
    import torch
    from jax_hf_amd import _lib as L
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(7)
    X = rng.standard_normal((9, 5, 5)) + 1j * rng.standard_normal((9, 5, 5))
    dt = {"c128": torch.complex128, "c64": torch.complex64}
    Xt = torch.from_numpy(X).to(dev).to(dt[src])
    Y = torch.empty(X.shape, dtype=dt[dst], device=dev)
    L.call("hfx_hermitize_convert", L.ptr(Xt), K.dt_code(Xt), L.ptr(Y), K.dt_code(Y), 9, 5, L.stream(dev))
    Xs = _np(Xt).astype(np.complex128)
    ref = 0.5 * (Xs + np.conj(np.swapaxes(Xs, -1, -2)))
    tol = 1e-15 if (src, dst) == ("c128", "c128") else 1e-6
    assert np.abs(_np(Y).astype(np.complex128) - ref).max() <= tol * np.abs(ref).max()
    Yn = _np(Y)
    assert np.array_equal(Yn, np.conj(np.swapaxes(Yn, -1, -2)))
    with pytest.raises(ValueError, match="hfx_hermitize_convert"):      # in place is refused
        L.call("hfx_hermitize_convert", L.ptr(Y), K.dt_code(Y), L.ptr(Y), K.dt_code(Y), 9, 5, L.stream(dev))


@pytest.mark.parametrize("nb", [8, 16, 24])
def test_orthonormalize_from_complex64(dev, nb):
    """The same basis as converting first and orthonormalising in place (bit
    for bit), unitary to FP64 rounding after two steps."""
    import torch
    from jax_hf_amd import _lib as L
    from jax_hf_amd import _ops as K
    from jax_hf_amd._util import KIND
    rng = np.random.default_rng(nb)
    nk = 17
    Q = np.linalg.qr(rng.standard_normal((nk, nb, nb)) + 1j * rng.standard_normal((nk, nb, nb)))[0]
    Q32 = torch.from_numpy(Q).to(dev).to(torch.complex64)
    U1 = torch.empty((nk, nb, nb), dtype=torch.complex128, device=dev)
    L.call("hfx_orthonormalize_from", nb, nk, L.ptr(Q32), L.ptr(U1), 2, L.stream(dev))
    U2 = torch.empty_like(U1)
    L.call("hfx_convert", L.ptr(Q32), KIND[Q32.dtype], L.ptr(U2), KIND[U2.dtype], Q32.numel(), L.stream(dev))
    K.orthonormalize(U2, 2)
    assert torch.equal(U1, U2)
    Un = _np(U1)
    assert np.abs(np.conj(np.swapaxes(Un, -1, -2)) @ Un - np.eye(nb)).max() < 1e-14
    assert np.abs(Un - Q).max() < 1e-6                                   # the polar factor of the rounded Q


@pytest.mark.parametrize("nk", [300, 9216])
def test_diis_coefficients_reduce_matches_two_launches(dev, nk):
    """hfx_diis_coefficients_reduce (the residual pass's block partials summed
    inside the device DIIS solve) gives the totals of hfx_reduce and the
    coefficients of hfx_diis_coefficients on them -- bit for bit above 2048
    k-points, where hfx_reduce itself takes the two-launch form (below it sums
    in one workgroup, in another order)."""
    import torch
    from jax_hf_amd import _lib as L
    from jax_hf_amd import _ops as K
    rng = np.random.default_rng(nk)
    nh = 6
    nval = 2 + nh
    part = torch.from_numpy(rng.uniform(0.1, 1.0, (nval, nk))).to(dev)
    w = torch.from_numpy(rng.uniform(0.5, 1.0, nk)).to(dev)
    ws = K.Workspace.get(dev, nk)
    out1 = torch.zeros(nval, dtype=torch.float64, device=dev)
    K.reduce(ws, part, nval, nk, w=w, wmask=(1 << nval) - 1, out=out1)
    B1 = torch.from_numpy(rng.uniform(0.1, 1.0, (nh, nh))).to(dev)
    B1 = 0.5 * (B1 + B1.T) + 5 * torch.eye(nh, dtype=torch.float64, device=dev)
    B2 = B1.clone()
    c1 = torch.zeros(8, dtype=torch.float64, device=dev)
    c2 = torch.zeros(8, dtype=torch.float64, device=dev)
    K.diis_coefficients(out1[2:], B1, 3, 5, c1)
    K.reduce(ws, part, nval, nk, w=w, wmask=(1 << nval) - 1, op=K.OP_PARTIALS)
    out2 = torch.zeros(nval, dtype=torch.float64, device=dev)
    L.call("hfx_diis_coefficients_reduce", L.ptr(ws.scratch), nk, nval, L.ptr(out2), L.ptr(B2), nh, 3, 5, L.ptr(c2),
           L.stream(dev))
    if nk > 2048:
        assert torch.equal(out1, out2)
        assert torch.equal(B1, B2)
        assert torch.equal(c1, c2)
    else:
        assert torch.allclose(out1, out2, rtol=1e-14, atol=0)
        assert torch.allclose(c1, c2, rtol=1e-10, atol=1e-12)
    ref = (_np(part) * _np(w)[None, :]).sum(axis=1)
    assert np.abs(_np(out2) - ref).max() < 1e-12 * ref.max()
