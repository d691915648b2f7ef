"""ctypes binding of libhfx.so (the HIP kernels) and per-shape plan cache.

The library is loaded from this package directory only; a missing or stale
library raises immediately (there is no CPU fallback in the product path).
"""

from __future__ import annotations

import atexit
import ctypes as C
import gc
import os
import threading
import weakref
from pathlib import Path

import torch

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libhfx.so"

HFX_C128 = 0
HFX_C64 = 1
ABI_VERSION = 24
# Largest band count of the per-k solver kernels (nb <= 16: one wavefront per
# k-point, nb <= 32: a 256-thread workgroup per k-point).
PERK_NB_MAX = 32

_lib = None
_lock = threading.Lock()

vp = C.c_void_p
i32 = C.c_int
f64 = C.c_double
sz = C.c_size_t

# name -> argtypes (all functions return int status unless listed in _RESTYPES)
_SIGS = {
    "hfx_abi_version": [],
    "hfx_last_error": [C.c_char_p, sz],
    "hfx_plan_create": [i32, i32, i32, i32, C.POINTER(vp)],
    "hfx_plan_destroy": [vp],
    "hfx_energy_scratch_bytes": [vp],
    "hfx_plan_set_exchange_variant": [vp, i32],
    "hfx_plan_get_exchange_variant": [vp],
    "hfx_exchange_autotune": [vp, vp, vp, i32, vp, i32, vp, i32, vp, vp, vp],
    "hfx_selfenergy": [vp, vp, vp, vp, i32, vp, i32, i32, i32, vp, vp, vp, vp],
    "hfx_prepare_vr": [vp, vp, i32, vp, i32, vp, vp],
    "hfx_selfenergy_adjoint": [vp, vp, vp, i32, vp, vp],
    "hfx_fft2_weighted": [vp, vp, vp, vp, vp],
    "hfx_cmul": [i32, sz, i32, vp, vp, f64, f64, i32, vp, vp],
    "hfx_scale_k": [sz, i32, i32, vp, vp, f64, vp, vp],
    "hfx_scale_k_ex": [sz, i32, i32, vp, vp, f64, vp, i32, vp, vp],
    "hfx_density_moments": [vp, vp, vp, vp, vp, vp],
    "hfx_kterms": [vp, vp, vp, i32, vp, vp, vp, vp, vp, vp],
    "hfx_fock_epilogue": [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, f64, vp],
    "hfx_hf_energy": [vp, vp, vp, vp, vp, i32, vp, vp, vp, vp],
    "hfx_convert": [vp, i32, vp, i32, sz, vp],
    "hfx_hermitize": [vp, vp, sz, i32, i32, vp],
    "hfx_hermitize_convert": [vp, i32, vp, i32, sz, i32, vp],
    "hfx_lincomb": [sz, i32, f64, f64, vp, f64, f64, vp, vp, vp],
    "hfx_fermi": [vp, vp, f64, sz, vp, vp],
    "hfx_lincomb_multi": [sz, i32, i32, i32, vp, vp, i32, vp, vp],
    "hfx_absmax_masked": [vp, sz, i32, i32, vp, vp, vp, vp],
    "hfx_entropy": [vp, i32, vp, sz, i32, vp, f64, vp, vp, vp],
    "hfx_eigh_batched": [i32, i32, sz, vp, vp, vp, i32, vp],
    "hfx_eigh_batched_tol": [i32, i32, sz, vp, vp, vp, i32, f64, vp],
    "hfx_eigh_batched_basis": [i32, i32, sz, vp, vp, vp, vp, i32, f64, vp],
    "hfx_diis_coefficients_reduce": [vp, sz, i32, vp, vp, i32, i32, i32, vp, vp],
    "hfx_density_diff": [i32, i32, sz, vp, vp, vp, vp, vp, vp],
    "hfx_density_from_Qp": [i32, i32, sz, vp, vp, vp, vp, f64, vp, vp, vp],
    "hfx_cg_grad": [i32, i32, sz, vp, vp, vp, vp, vp, vp, vp, vp],
    "hfx_cg_precond": [i32, i32, sz, vp, vp, vp, vp, vp, vp, f64, vp, vp, vp, vp, i32, vp],
    "hfx_cg_direction": [i32, i32, sz, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
    "hfx_cg_trial": [i32, i32, sz, vp, vp, vp, vp, vp, vp, vp, i32, vp, vp],
    "hfx_cg_direction_gj": [i32, i32, sz, vp, vp, vp, vp, vp, vp, i32, vp],
    "hfx_cg_trial_gj": [i32, i32, sz, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, i32, vp],
    "hfx_cg_retract_gj": [i32, i32, sz, vp, vp, vp, vp, vp, vp, vp, i32, vp],
    "hfx_selfenergy_batched": [vp, i32, vp, vp, i32, i32, vp, vp],
    "hfx_fock_epilogue_batched": [vp, i32, vp, vp, vp, vp, vp, vp, vp, sz, vp],
    "hfx_batch_scratch_doubles": [i32],
    "hfx_planar_pack": [vp, i32, vp, vp, vp, i32, i32, vp, vp],
    "hfx_selfenergy_planar": [vp, i32, vp, vp, i32, vp, vp, i32, i32, i32, vp, vp, vp, vp],
    "hfx_fock_epilogue_planar": [vp, vp, vp, vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, f64, vp],
    "hfx_fock_epilogue_planar_batched": [vp, i32, vp, vp, i32, vp, vp, vp, vp, vp, sz, vp],
    "hfx_reduce_batched": [vp, i32, sz, i32, vp, C.c_uint, vp, i32, vp, i32, vp, sz, vp],
    "hfx_solve_mu_batched": [vp, vp, vp, sz, i32, i32, f64, i32, vp, vp, vp, sz, vp],
    "hfx_cg_retract": [i32, i32, sz, vp, vp, vp, vp, vp, f64, i32, vp, vp, vp],
    "hfx_rotate_eigh": [i32, i32, sz, vp, vp, vp, vp],
    "hfx_matmul_batched": [i32, i32, sz, vp, vp, vp, i32, vp, i32, vp, vp],
    "hfx_extrap_rotate": [i32, i32, sz, i32, vp, vp, vp, vp, vp],
    "hfx_extrap_rotate_batched": [i32, i32, sz, i32, i32, vp, vp, vp, vp, vp],
    "hfx_rotate_density": [i32, i32, sz, vp, vp, vp, vp, vp],
    "hfx_scf_resid": [i32, i32, sz, vp, vp, vp, vp, vp, vp],
    "hfx_scf_resid_dots": [i32, i32, sz, vp, vp, vp, vp, vp, i32, i32, i32, vp, vp],
    "hfx_orthonormalize": [i32, i32, sz, vp, i32, vp],
    "hfx_orthonormalize_from": [i32, sz, vp, vp, i32, vp],
    "hfx_symmetry_average": [i32, i32, sz, i32, vp, vp, vp, C.c_ulonglong, vp, f64, vp, vp],
    "hfx_resample_kgrid": [i32, i32, i32, sz, i32, vp, vp, vp],
    "hfx_superlattice_fock": [i32, i32, i32, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp,
                              vp, vp, i32, vp, vp, vp, vp, vp, f64, vp, vp, vp, vp, vp],
    "hfx_diis_coefficients": [vp, vp, i32, i32, i32, vp, vp],
    "hfx_diis_coefficients_batched": [vp, sz, vp, i32, i32, i32, i32, vp, vp],
    "hfx_batched_dots": [i32, i32, sz, vp, vp, i32, i32, vp, vp],
    "hfx_vec_dots": [vp, vp, i32, sz, i32, vp, vp],
    "hfx_vec_shift": [vp, sz, vp, f64, vp],
    "hfx_vec_shift_sum": [vp, sz, vp, vp, sz, f64, vp],
    "hfx_newton_dP": [i32, i32, sz, vp, vp, vp, vp, vp, vp],
    "hfx_newton_hvp": [i32, i32, sz, vp, vp, vp, vp, vp, vp, f64, vp, vp, vp, vp],
    "hfx_newton_prec": [i32, i32, sz, vp, vp, vp, vp, vp, f64, vp, vp, vp, vp],
    "hfx_solver_scratch_doubles": [],
    "hfx_reduce": [vp, i32, sz, vp, C.c_uint, vp, vp, i32, vp, vp, vp],
    "hfx_solve_mu_steps": [vp, vp, sz, i32, f64, f64, f64, i32, i32, vp, vp, vp, vp],
    "hfx_solve_mu": [vp, vp, sz, i32, f64, f64, f64, i32, i32, i32, f64, vp, vp, vp, vp, vp],
    "hfx_qr_retract": [i32, i32, sz, vp, vp, vp, vp],
    "hfx_lincomb_dev": [sz, i32, vp, vp, vp, sz, i32, vp, vp, vp, f64, vp, f64, vp, vp],
    "hfx_lincomb_dev_batch": [i32, sz, i32, sz, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
    "hfx_pair_dots": [i32, i32, sz, i32, vp, vp, vp, vp, vp, C.c_uint, vp, vp, i32, vp, vp],
    "hfx_noalloc_scratch_doubles": [],
    "hfx_solve_mu_noalloc": [vp, vp, sz, i32, f64, f64, f64, i32, i32, i32, f64, vp, vp, vp, vp],
}
_RESTYPES = {"hfx_energy_scratch_bytes": sz, "hfx_solver_scratch_doubles": sz, "hfx_batch_scratch_doubles": sz,
             "hfx_noalloc_scratch_doubles": sz}


class HfxError(RuntimeError):
    pass


def lib():
    """The loaded library (raises if it was never built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -m jax_hf_amd.build or __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, i32)
        v = L.hfx_abi_version()
        if v != ABI_VERSION:
            raise ImportError(f"libhfx ABI version {v} != expected {ABI_VERSION}; rebuild")
        _lib = L
        # registered after torch's own exit handlers, so it runs before them
        atexit.register(release_all)
        return L


def exported_symbols():
    return list(_SIGS)


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().hfx_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str):
    if rc == 0:
        return
    msg = last_error()
    if rc == -1:
        raise ValueError(f"{what}: {msg}")
    raise HfxError(f"{what} failed ({rc}): {msg}")


def call(name: str, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)


def ptr(t) -> int | None:
    """Device pointer of a contiguous tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("internal: non-contiguous tensor passed to libhfx")
    return t.data_ptr()


def stream(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class Plan:
    """Owns an hfx_plan (twiddles + channel lists) for one (nk1, nk2, nb, dtype)."""

    def __init__(self, nk1: int, nk2: int, nb: int, dtype: int, device):
        self.nk1, self.nk2, self.nb, self.dtype = nk1, nk2, nb, dtype
        self.device = device
        h = vp()
        with torch.cuda.device(device):
            call("hfx_plan_create", nk1, nk2, nb, dtype, C.byref(h))
        self.handle = h
        nbytes = lib().hfx_energy_scratch_bytes(h)
        self.scratch = torch.zeros(nbytes // 8, dtype=torch.float64, device=device)

    VARIANTS = {0: "auto", 1: "v1", 2: "v2_a", 3: "v2_b", 4: "v2_c", 5: "v3_a", 6: "v3_b", 7: "v3_c"}

    @property
    def exchange_variant(self) -> int:
        return int(lib().hfx_plan_get_exchange_variant(self.handle))

    @exchange_variant.setter
    def exchange_variant(self, v: int):
        call("hfx_plan_set_exchange_variant", self.handle, int(v))

    def autotune(self, P, VR, Sigma, reps: int = 5, planar: bool = False, vr_map=None):
        """Time every exchange-kernel variant on (P, VR) -> Sigma; keep the fastest.

        ``planar``: P / Sigma are nb^2 slot planes and VR the channel-major VR
        (or the scalar VR), read through ``vr_map`` when given.  Returns
        (best variant, {name: ms per launch}).
        """
        t = (C.c_double * 8)()
        best = C.c_int()
        if planar:
            vfull = 1 if VR.numel() > P.numel() // (self.nb * self.nb) else 0
        else:
            vfull = 0 if VR.shape[-1] == 1 else 1
        with torch.cuda.device(self.device):
            call("hfx_exchange_autotune", self.handle, ptr(P), ptr(VR), vfull, ptr(vr_map), int(planar), ptr(Sigma),
                 int(reps), t, C.byref(best), stream(self.device))
        return best.value, {self.VARIANTS[v]: t[v] for v in range(1, 8)}

    def tune_once(self, VR, vfull: bool, reps: int = 5, vr_map=None):
        """Autotune the exchange variant once per plan, on zero slot planes of
        all nb^2 channels with this VR (channel-major when vfull).  Skipped
        while the current stream is capturing; a mesh the slot-plane exchange
        does not support keeps the built-in default."""
        if getattr(self, "_tuned", False) or torch.cuda.is_current_stream_capturing():
            return
        self._tuned = True
        cdt = torch.complex128 if self.dtype == HFX_C128 else torch.complex64
        shape = (self.nb * self.nb if vfull else 1, self.nk1, self.nk2)
        if tuple(VR.shape) != shape and VR.numel() != shape[0] * shape[1] * shape[2]:
            raise ValueError(f"tune_once: VR of {tuple(VR.shape)} elements for a plan of {shape}")
        P = torch.zeros((self.nb * self.nb, self.nk1, self.nk2), dtype=cdt, device=self.device)
        S = torch.empty_like(P)
        try:
            self.autotune(P, VR, S, reps=reps, planar=True, vr_map=vr_map)
        except HfxError:
            self.exchange_variant = 0

    def destroy(self):
        """Free the plan's device tables now (idempotent)."""
        h, self.handle = self.handle, None
        self.scratch = None
        if h and _lib is not None:
            _lib.hfx_plan_destroy(h)

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # pragma: no cover
            pass


_plans: dict = {}


def plan_for(nk1: int, nk2: int, nb: int, cdtype: torch.dtype, device) -> Plan:
    dt = HFX_C128 if cdtype == torch.complex128 else HFX_C64
    dev = torch.device(device)
    key = (nk1, nk2, nb, dt, dev.index)
    p = _plans.get(key)
    if p is None:
        p = Plan(nk1, nk2, nb, dt, dev)
        _plans[key] = p
    return p


# Objects that keep device state between calls (kernels with cached solver
# states and captured graphs, Fock-build plug-ins) and module caches of device
# buffers.  release_all() drops all of it while the HIP runtime is still whole.
_owners: "weakref.WeakSet" = weakref.WeakSet()
_cache_clearers: list = []


def register_owner(obj):
    """obj._hfx_release() is called by release_all()."""
    _owners.add(obj)


def register_cache(clear_fn):
    """clear_fn() empties a module-level cache of device buffers."""
    _cache_clearers.append(clear_fn)


def release_all():
    """Release every cached graph, solver state, workspace, pinned buffer and
    plan, after the device has finished its work.

    Runs at interpreter exit (registered when the library is first loaded, so
    before torch's and the HIP runtime's own teardown) and may be called by a
    program at any point; later calls rebuild what they need.  Without it the
    plans and graphs are freed by destructors during interpreter teardown, in
    no fixed order relative to the runtime's and any loaded tool library's
    finalisation.
    """
    if _lib is None or not torch.cuda.is_available() or not torch.cuda.is_initialized():
        return
    try:
        torch.cuda.synchronize()
    except Exception:  # pragma: no cover - a faulted device: nothing to release safely
        return
    for o in list(_owners):
        try:
            o._hfx_release()
        except ReferenceError:  # pragma: no cover
            pass
    for clear in _cache_clearers:
        clear()
    gc.collect()
    torch.cuda.synchronize()
    for p in list(_plans.values()):
        p.destroy()
    _plans.clear()
    gc.collect()
    torch.cuda.synchronize()


def require_gpu(device=None):
    if not torch.cuda.is_available():
        raise HfxError("jax_hf_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    lib()
