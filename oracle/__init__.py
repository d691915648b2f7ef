"""CPU FP64 oracle for the Hartree-Fock Fock-build / direct-minimisation path.

TEST INFRASTRUCTURE ONLY.  This package restates, in NumPy/SciPy float64 /
complex128 arithmetic, the algorithm of the reference package (jax_hf,
``src/jax_hf/{problem,fock,utils,solver,reference_scf}.py``) so that the HIP
implementation in ``jax_hf_amd`` can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker / CPU baseline: the product path never
calls into it and fails loudly when its HIP library is missing.

Pinning (what makes each function trustworthy, see ``tests/test_oracle.py``):

* ``HartreeFockKernel`` VR + ifftshift phase, ``selfenergy_fft``,
  ``build_fock``: the reference fixture ``meshgrid_regression_case_v1.npz``
  (complex64 precision), a brute-force O(Nk^2) direct convolution sum on
  non-uniform weights (1e-13), the single-k closed form, Hermitian packing ==
  full path, block specs == full path.
* ``hf_energy``: fixture ``E_expected`` plus the direct trace definition.
* ``occupation_entropy`` / ``free_energy``: closed form at p=1/2 and a central
  finite difference dOmega/dp = eps + T logit(p).
* ``solve_mu`` / ``find_chemical_potential``: count equation invariant,
  fixture ``mu_expected``, cold limit.
* spectral Cayley helpers: LU Cayley (``numpy.linalg.solve``) and unitarity.
* ``solve_cg`` / ``solve_newton`` / ``solve_scf``: the fixture's converged
  state (E, P, mu), commutator invariant, cross-solver agreement, Newton HVP vs
  central finite differences, DIIS coefficients vs a dense augmented solve.

* ``oracle/drivers.py`` (``resample_kgrid``, ``solve_continuation``,
  ``solve_deflated``) and ``deflation_bias_terms``: ``numpy.interp`` with a
  period, round trips and coincident samples; the reference tests' manual
  continuation driver and cold fixed point; the bistable model's analytic
  solutions; a central-difference gradient identity and closed forms of the
  bias (``tests/test_oracle_drivers.py``).

Parity unpinned (documented): iteration counts / history arrays of the
solvers vs JAX (depend on XLA round-off at the line-search accept boundary)
and the eigenvector gauge returned by ``jnp.linalg.eigh``.
"""

from .hf_oracle import *  # noqa: F401,F403
