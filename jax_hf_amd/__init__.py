"""jax_hf_amd — MI355X-native Hartree-Fock Fock build and solvers.

Drop-in for the ``jax_hf`` hot path (``src/jax_hf/__init__.py:32-95``):
``HartreeFockKernel``, ``build_fock``, ``hf_energy``, ``free_energy``,
``occupation_entropy``, ``solve`` / ``solve_direct_minimization``,
``solve_scf``, the drivers ``solve_continuation`` / ``resample_kgrid`` and
``solve_deflated``, the superlattice Fock plug-in factories, the npz case
readers / writers (``cases``), and their config / result types (the export list of ``src/jax_hf/__init__.py:32-95``).  Arrays are torch tensors on a
ROCm device; all arithmetic runs in the hand-written HIP kernels of
``libhfx.so`` (``jax_hf_amd/csrc``), called through ctypes.

Importing the package touches no GPU and loads no library; the first call
that needs the kernels loads ``libhfx.so`` and raises if it is missing.
"""

from __future__ import annotations

__version__ = "0.1.0"

from .problem import HartreeFockKernel  # noqa: F401
from .fock import build_fock, hf_energy, free_energy, occupation_entropy  # noqa: F401

__all__ = [
    "HartreeFockKernel",
    "build_fock",
    "hf_energy",
    "free_energy",
    "occupation_entropy",
]


def _extend(mod_names):
    import importlib
    for modname, names in mod_names:
        try:
            mod = importlib.import_module(f".{modname}", __name__)
        except ModuleNotFoundError:
            continue
        for n in names:
            globals()[n] = getattr(mod, n)
            __all__.append(n)


_extend([
    ("solver", ["SolverConfig", "SolveResult", "solve", "solve_direct_minimization"]),
    ("reference_scf", ["SCFConfig", "SCFResult", "solve_scf"]),
    ("batch", ["solve_batch"]),
    ("utils", ["resample_kgrid"]),
    ("continuation", ["ContinuationResult", "solve_continuation"]),
    ("deflation", ["DeflatedResult", "solve_deflated"]),
    ("superlattice", ["ExtendedGridLayout", "build_extended_layout", "make_superlattice_fock_fn",
                      "make_superlattice_build_fock_fn"]),
    ("cases", ["RegressionCase", "load_case", "save_case", "case_from_scf", "save_scan", "load_scan"]),
])
