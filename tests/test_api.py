"""The package exports the reference's public names (src/jax_hf/__init__.py:66-95),
so ``import jax_hf_amd as jax_hf`` finds every one.  Importing the package
touches no GPU and loads no library."""

import jax_hf_amd

REFERENCE_ALL = [
    "HartreeFockKernel",
    "SolverConfig", "SolveResult", "solve", "solve_direct_minimization",
    "SCFConfig", "SCFResult", "solve_scf",
    "build_fock", "hf_energy", "free_energy", "occupation_entropy",
    "resample_kgrid", "ContinuationResult", "solve_continuation",
    "DeflatedResult", "solve_deflated",
    "ExtendedGridLayout", "build_extended_layout", "make_superlattice_fock_fn", "make_superlattice_build_fock_fn",
]


def test_reference_public_names_are_exported():
    missing = [n for n in REFERENCE_ALL if not hasattr(jax_hf_amd, n) or n not in jax_hf_amd.__all__]
    assert not missing, missing


def test_config_defaults_match_reference_fields():
    """SolverConfig keeps the reference's fields and defaults (solver.py:240-286)."""
    c = jax_hf_amd.SolverConfig()
    assert (c.max_iter, c.tol_E, c.optimizer, c.deflation_sigma, c.deflation_length) == (200, 1e-7, "cg", 0.0, 1.0)
    assert c.project_fn is None and c.deflation_targets is None and c.block_sizes is None
