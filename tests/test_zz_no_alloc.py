"""hfx_solve_mu_noalloc (library-owned scratch, no allocation) against
hfx_solve_mu (caller scratch) through the C-ABI: the same kernels on the same
inputs give bit-identical mu and occupations, for every mode, on two calls in a
row with the second on a smaller input (the scratch keeps the first call's
contents), and sizes beyond the limit are refused before any launch."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from jax_hf_amd import _lib as L
    return L, L.lib()


def _case(dev, nk, nb, seed):
    import torch
    rng = np.random.default_rng(seed)
    eps = torch.from_numpy(np.sort(rng.standard_normal((nk, nb)), axis=-1).reshape(-1).copy()).to(dev)
    w = torch.from_numpy(rng.uniform(0.5, 1.5, nk)).to(dev)
    target = float(0.45 * nb * w.sum().item())
    return eps, w, target


@pytest.mark.parametrize("mode,shift", [(1, 0.0), (1, 0.7), (0, 0.0), (2, 0.0)])
def test_noalloc_matches_caller_scratch_entry(dev, mode, shift):
    import torch
    L, lib = _lib()
    assert lib.hfx_noalloc_scratch_doubles() >= lib.hfx_solver_scratch_doubles()
    s = L.stream(dev)
    scratch = torch.zeros(int(lib.hfx_solver_scratch_doubles()), dtype=torch.float64, device=dev)
    for nk, nb, seed in ((96 * 96, 16, 1), (24 * 24, 4, 2)):   # the second call is smaller
        eps, w, target = _case(dev, nk, nb, seed)
        out = {}
        for name in ("ref", "noalloc"):
            mu = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
            st = torch.zeros(32, dtype=torch.float64, device=dev)
            p = torch.full((nk * nb,), np.nan, dtype=torch.float64, device=dev)
            args = (L.ptr(eps), L.ptr(w), nk, nb, target, 0.05, 0.1, 0, 54 if mode == 1 else 25, mode, shift,
                    L.ptr(st), L.ptr(mu), L.ptr(p))
            if name == "ref":
                rc = lib.hfx_solve_mu(*args, L.ptr(scratch), s)
            else:
                rc = lib.hfx_solve_mu_noalloc(*args, s)
            assert rc == 0, L.last_error()
            torch.cuda.synchronize()
            out[name] = (mu.item(), st[0].item(), p.cpu().numpy())
        (m0, s0, p0), (m1, s1, p1) = out["ref"], out["noalloc"]
        assert np.isfinite(m0) and m0 == m1 and s0 == s1 == m0
        np.testing.assert_array_equal(p0, p1)
        # mu solves the count equation (bisection: to its bracket; Newton: to the rounding of the sum)
        wk = np.repeat(w.cpu().numpy(), nb)
        if shift == 0.0:
            x = np.clip((eps.cpu().numpy() - m0) / (0.05 + (1e-12 if mode == 1 else 0.0)), -700, 700)
            assert abs(np.sum(wk / (1 + np.exp(x))) - target) < 1e-9 * target


def test_noalloc_rejects_sizes_beyond_the_limit(dev):
    L, lib = _lib()
    d = C.c_void_p(4096)
    rc = lib.hfx_solve_mu_noalloc(d, d, 1 << 30, 4, 1.0, 0.1, 0.0, 0, 54, 1, 0.0, None, d, None, None)
    assert rc == -1 and "2^32" in L.last_error()
