"""A saved regression case drives the GPU solvers (jax_hf_amd.cases): the
fixture's inputs (promoted to FP64) and loop controls, through
RegressionCase.kernel and scf_config, reach its expected converged state (complex64 precision of the
file) and the oracle's (1e-10 per k-point); a GPU result recorded with
case_from_scf round-trips through save_case / load_case bit for bit."""

import numpy as np
import pytest

import oracle as O
from jax_hf_amd import cases as CS
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def test_saved_case_solves_to_its_expected_state(dev, tmp_path):
    import torch
    import jax_hf_amd as J
    case = CS.load_case(GOLDEN / "meshgrid_regression_case_v1.npz")
    # the file's own dtypes give a complex64 kernel
    assert case.kernel(device=dev).dtype == torch.complex64
    G = case.kernel(device=dev, precision="fp64")
    cfg = case.scf_config(density_tol=1e-10, comm_tol=1e-9, max_iter=300)
    r = J.solve_scf(G, case.P0.astype(np.complex128), case.n_electrons, config=cfg)
    assert r.converged
    np.testing.assert_allclose(_np(r.density_matrix), case.P_expected, atol=5e-6)
    np.testing.assert_allclose(_np(r.fock_matrix), case.F_expected, atol=2e-5)
    assert abs(float(r.energy) - float(case.E_expected)) < 2e-6 * abs(float(case.E_expected))
    assert abs(float(r.chemical_potential) - float(case.mu_expected)) < 5e-6
    K = O.OracleKernel(case.weights, case.hamiltonian, case.coulomb_q, float(case.T), include_hartree=True,
                       reference_density=case.reference_density, hartree_matrix=case.hartree_matrix)
    r0 = O.solve_scf(K, case.P0.astype(np.complex128), case.n_electrons, config=O.SCFConfig(
        max_iter=300, density_tol=1e-10, comm_tol=1e-9, acceleration="diis", diis_size=int(case.diis_size)))
    assert abs(float(r.energy) - r0.energy) < 1e-10
    # the GPU result as a new case file, read back unchanged
    inputs = dict(weights=case.weights, hamiltonian=case.hamiltonian, coulomb_q=case.coulomb_q, T=case.T,
                  hartree_matrix=case.hartree_matrix, reference_density=case.reference_density)
    new = CS.case_from_scf(inputs, r, P0=case.P0, n_electrons=case.n_electrons, scf_config=cfg)
    out = tmp_path / "gpu_case.npz"
    CS.save_case(out, new)
    back = CS.load_case(out)
    np.testing.assert_array_equal(back.P_expected, _np(r.density_matrix))
    np.testing.assert_array_equal(back.F_expected, _np(r.fock_matrix))
    assert float(back.E_expected) == float(r.energy) and int(back.k_expected) == int(r.iterations)
    np.testing.assert_array_equal(back.hist_E_expected, _np(r.history["E"])[:int(r.iterations)])
