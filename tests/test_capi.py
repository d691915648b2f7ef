"""The C-ABI library loads on the CPU host and exports every symbol include/hfx.h declares."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    txt = (ROOT / "include" / "hfx.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|size_t)\s+(hfx_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "hfx_selfenergy" in names and "hfx_plan_create" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    lib_path = ROOT / "jax_hf_amd" / "libhfx.so"
    if not lib_path.exists():
        pytest.skip("libhfx.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(str(lib_path))
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    L.hfx_abi_version.restype = ctypes.c_int
    from jax_hf_amd import _lib
    assert L.hfx_abi_version() == _lib.ABI_VERSION


def test_python_binding_covers_header():
    from jax_hf_amd import _lib
    assert sorted(_lib.exported_symbols()) == _declared()


def test_plan_create_rejects_bad_shapes_without_gpu():
    """Argument validation happens before any device call."""
    lib_path = ROOT / "jax_hf_amd" / "libhfx.so"
    if not lib_path.exists():
        pytest.skip("libhfx.so not built")
    from jax_hf_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.hfx_plan_create(0, 4, 2, 0, ctypes.byref(h)) == -1
    assert "bad shape" in _lib.last_error()
    assert L.hfx_plan_create(4, 4, 2, 7, ctypes.byref(h)) == -1
    assert L.hfx_plan_create(11, 11, 2, 0, ctypes.byref(h)) == -3
    assert "prime factor" in _lib.last_error()
    assert L.hfx_plan_create(128, 128, 2, 0, ctypes.byref(h)) == -3


def test_package_import_is_lazy():
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import jax_hf_amd, jax_hf_amd._lib as L; "
            "assert L._lib is None; print('ok')" % str(ROOT))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
