"""On-disk case formats of the reference: npz files read and written as plain
arrays (``numpy.load(allow_pickle=False)``; nothing in a file is executed).

* **Regression case** -- the schema of ``tests/data/meshgrid_regression_case_v1.npz``,
  consumed by ``examples/contimod_example_meshgrid_offline.py:43-62``: the
  kernel inputs (``weights``, ``hamiltonian``, ``coulomb_q``, ``T``,
  ``hartree_matrix``, ``reference_density``), the SCF start and controls
  (``P0``, ``electrondensity0``, ``max_iter``, ``comm_tol``, ``diis_size``,
  ``precond_mode``, ``precond_auto_nb``), the converged state it expects
  (``P_expected``, ``F_expected``, ``E_expected``, ``mu_expected``,
  ``k_expected``, ``hist_E_expected``, ``hist_dC_expected``) and metadata
  (``version``, ``nk``, ``kmax``).  :func:`load_case` / :func:`save_case`
  keep every field's dtype and shape, so a round trip reproduces the file's
  arrays exactly; :meth:`RegressionCase.kernel` builds the GPU kernel the
  case describes and :func:`case_from_scf` records a solve as a new case.
* **Scan results** -- the schema of ``tests/data/bilayer_reference.npz``
  written by ``tests/generate_bilayer_reference.py:175-194``: scalar run
  parameters plus one ``"{branch}_{density:+.2f}/{field}"`` entry per point
  and field (energy, n_iter, converged, n_electrons, mu).
  :func:`save_scan` / :func:`load_scan` write and read it;
  :func:`scan_key` forms the point key.
"""

from __future__ import annotations

from dataclasses import dataclass, field, fields
from pathlib import Path
from typing import Any

import numpy as np

__all__ = ["RegressionCase", "load_case", "save_case", "case_from_scf", "save_scan", "load_scan", "scan_key"]

_CASE_ORDER = ("version", "nk", "kmax", "T", "weights", "hamiltonian", "coulomb_q", "hartree_matrix",
               "reference_density", "P0", "electrondensity0", "max_iter", "comm_tol", "diis_size",
               "precond_mode", "precond_auto_nb", "P_expected", "F_expected", "E_expected", "mu_expected",
               "k_expected", "hist_E_expected", "hist_dC_expected")
_REQUIRED = ("T", "weights", "hamiltonian", "coulomb_q")


def _host(x):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(x)


@dataclass
class RegressionCase:
    """One saved Hartree-Fock case (every field a numpy array or None)."""
    T: Any
    weights: Any
    hamiltonian: Any
    coulomb_q: Any
    version: Any = None
    nk: Any = None
    kmax: Any = None
    hartree_matrix: Any = None
    reference_density: Any = None
    P0: Any = None
    electrondensity0: Any = None
    max_iter: Any = None
    comm_tol: Any = None
    diis_size: Any = None
    precond_mode: Any = None
    precond_auto_nb: Any = None
    P_expected: Any = None
    F_expected: Any = None
    E_expected: Any = None
    mu_expected: Any = None
    k_expected: Any = None
    hist_E_expected: Any = None
    hist_dC_expected: Any = None
    extra: dict = field(default_factory=dict)   # keys outside the schema, kept for the round trip

    def kernel(self, *, device=None, include_exchange: bool = True, precision: str = "file"):
        """The ``HartreeFockKernel`` of this case (Hartree on when the case has a
        Hartree matrix, as ``contimod_example_meshgrid_offline.py:45-54``).
        ``precision="file"`` keeps the file's dtypes (a complex64 case gives a
        complex64 kernel, as the reference's default float32 mode does);
        ``"fp64"`` promotes the arrays to float64 / complex128 first."""
        from .problem import HartreeFockKernel
        if precision not in ("file", "fp64"):
            raise ValueError(f"precision must be 'file' or 'fp64'; got {precision!r}")
        up = (lambda x: x) if precision == "file" else (
            lambda x: None if x is None else np.asarray(x).astype(
                np.complex128 if np.iscomplexobj(x) else np.float64))
        hartree = self.hartree_matrix is not None
        h = up(self.hamiltonian)
        ref = up(self.reference_density)
        if hartree and ref is None:
            ref = np.zeros_like(h)
        return HartreeFockKernel(up(self.weights), h, up(self.coulomb_q), float(self.T),
                                 include_hartree=hartree, include_exchange=include_exchange,
                                 reference_density=ref, hartree_matrix=up(self.hartree_matrix), device=device)

    @property
    def n_electrons(self) -> float:
        if self.electrondensity0 is None:
            raise ValueError("the case stores no electron count (electrondensity0)")
        return float(self.electrondensity0)

    def scf_config(self, **overrides):
        """An ``SCFConfig`` with the case's loop controls (max_iter, comm_tol,
        diis_size; DIIS acceleration when diis_size is given)."""
        from .reference_scf import SCFConfig
        kw = {}
        if self.max_iter is not None:
            kw["max_iter"] = int(self.max_iter)
        if self.comm_tol is not None:
            kw["comm_tol"] = float(self.comm_tol)
        if self.diis_size is not None:
            kw["diis_size"] = int(self.diis_size)
            kw["acceleration"] = "diis"
        kw.update(overrides)
        return SCFConfig(**kw)

    def to_dict(self) -> dict:
        out = {}
        for name in _CASE_ORDER:
            v = getattr(self, name)
            if v is not None:
                out[name] = np.asarray(_host(v))
        for k, v in self.extra.items():
            out[k] = np.asarray(_host(v))
        return out


def load_case(path) -> RegressionCase:
    """Read a regression case (arrays only; a pickled entry raises)."""
    with np.load(Path(path), allow_pickle=False) as d:
        data = {k: d[k] for k in d.files}
    missing = [k for k in _REQUIRED if k not in data]
    if missing:
        raise ValueError(f"{path}: not a regression case, missing {missing}")
    known = {f.name for f in fields(RegressionCase)} - {"extra"}
    kw = {k: v for k, v in data.items() if k in known}
    extra = {k: v for k, v in data.items() if k not in known}
    return RegressionCase(extra=extra, **kw)


def save_case(path, case: RegressionCase) -> None:
    """Write a regression case in the schema's key order (``numpy.savez``)."""
    h = np.asarray(_host(case.hamiltonian))
    if h.ndim != 4 or h.shape[-1] != h.shape[-2]:
        raise ValueError(f"hamiltonian must have shape (nk1, nk2, nb, nb); got {h.shape}")
    w = np.asarray(_host(case.weights))
    if w.shape != h.shape[:2]:
        raise ValueError(f"weights must have shape {h.shape[:2]}; got {w.shape}")
    np.savez(str(path), **case.to_dict())


def case_from_scf(kernel_inputs: dict, result, *, P0, n_electrons: float, scf_config=None,
                  version: str = "v1", nk=None, kmax=None) -> RegressionCase:
    """A regression case recording an SCF solve: ``kernel_inputs`` holds the
    kernel's arrays (weights, hamiltonian, coulomb_q, T and optionally
    hartree_matrix, reference_density); ``result`` an ``SCFResult``."""
    h = _host(kernel_inputs["hamiltonian"])
    hist = getattr(result, "history", {}) or {}
    k = int(result.iterations)
    case = RegressionCase(
        T=np.asarray(kernel_inputs["T"]), weights=_host(kernel_inputs["weights"]), hamiltonian=h,
        coulomb_q=_host(kernel_inputs["coulomb_q"]),
        hartree_matrix=None if kernel_inputs.get("hartree_matrix") is None else _host(kernel_inputs["hartree_matrix"]),
        reference_density=(None if kernel_inputs.get("reference_density") is None
                           else _host(kernel_inputs["reference_density"])),
        version=np.asarray(version), nk=None if nk is None else np.asarray(nk, np.int32),
        kmax=None if kmax is None else np.asarray(kmax),
        P0=_host(P0), electrondensity0=np.asarray(float(n_electrons)),
        P_expected=_host(result.density_matrix), F_expected=_host(result.fock_matrix),
        E_expected=np.asarray(float(result.energy)), mu_expected=np.asarray(float(result.chemical_potential)),
        k_expected=np.asarray(k, np.int32))
    if "E" in hist:
        case.hist_E_expected = np.asarray(_host(hist["E"]))[:k]
    if "commutator_residual" in hist:
        case.hist_dC_expected = np.asarray(_host(hist["commutator_residual"]))[:k]
    if scf_config is not None:
        case.max_iter = np.asarray(int(scf_config.max_iter), np.int32)
        case.comm_tol = np.asarray(float(scf_config.comm_tol))
        if scf_config.acceleration == "diis":
            case.diis_size = np.asarray(int(scf_config.diis_size), np.int32)
    return case


def scan_key(branch: str, density: float) -> str:
    """The point key of a scan file (``generate_bilayer_reference.py:163``)."""
    return f"{branch}_{float(density):+.2f}"


def save_scan(path, meta: dict, results: dict) -> None:
    """Write scan results: ``meta`` scalar / array run parameters, ``results``
    {point key: {field: value}} stored as ``"point/field"`` entries
    (``generate_bilayer_reference.py:176-194``)."""
    out = {}
    for k, v in meta.items():
        if "/" in k:
            raise ValueError(f"meta key {k!r} contains '/'")
        out[k] = np.asarray(v)
    for point, vals in results.items():
        if "/" in point:
            raise ValueError(f"point key {point!r} contains '/'")
        for fk, fv in vals.items():
            out[f"{point}/{fk}"] = np.asarray(fv)
    np.savez(str(path), **out)


def load_scan(path) -> tuple[dict, dict]:
    """Read scan results: (meta, {point key: {field: array}}), insertion order kept."""
    meta, results = {}, {}
    with np.load(Path(path), allow_pickle=False) as d:
        for k in d.files:
            if "/" in k:
                point, fk = k.split("/", 1)
                results.setdefault(point, {})[fk] = d[k]
            else:
                meta[k] = d[k]
    return meta, results
