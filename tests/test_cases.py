"""On-disk case formats (jax_hf_amd.cases) on the CPU: the reference's
regression case and scan-results files round-trip exactly (keys, dtypes,
shapes, values), the case's kernel inputs feed the oracle to the fixture's
converged state, and malformed inputs are refused."""

import numpy as np
import pytest

import oracle as O
from jax_hf_amd import cases as CS
from tests.conftest import GOLDEN


def _same(a, b):
    assert a.dtype == b.dtype and a.shape == b.shape
    if a.dtype.kind in "fc":
        np.testing.assert_array_equal(a, b)
    else:
        assert np.all(a == b)


def test_regression_case_round_trip(tmp_path):
    src = GOLDEN / "meshgrid_regression_case_v1.npz"
    case = CS.load_case(src)
    assert case.extra == {}
    out = tmp_path / "case.npz"
    CS.save_case(out, case)
    with np.load(src, allow_pickle=False) as a, np.load(out, allow_pickle=False) as b:
        assert list(a.files) == list(b.files)
        for k in a.files:
            _same(a[k], b[k])
    again = CS.load_case(out)
    assert str(again.precond_mode) == str(case.precond_mode) and int(again.k_expected) == int(case.k_expected)


def test_regression_case_drives_the_oracle_to_its_expected_state():
    """The loaded inputs and controls (as the reference's offline example uses
    them) reach P_expected / E_expected / mu_expected at the fixture's
    complex64 precision through the oracle's SCF."""
    case = CS.load_case(GOLDEN / "meshgrid_regression_case_v1.npz")
    K = O.OracleKernel(case.weights, case.hamiltonian, case.coulomb_q, float(case.T), include_hartree=True,
                       reference_density=case.reference_density, hartree_matrix=case.hartree_matrix)
    cfg = case.scf_config(density_tol=1e-10, comm_tol=1e-9, max_iter=300)
    assert cfg.acceleration == "diis" and cfg.diis_size == int(case.diis_size)
    r = O.solve_scf(K, case.P0.astype(np.complex128), case.n_electrons,
                    config=O.SCFConfig(**{f: getattr(cfg, f) for f in ("max_iter", "density_tol", "comm_tol",
                                                                        "acceleration", "diis_size")}))
    assert r.converged
    assert abs(r.energy - float(case.E_expected)) < 2e-6 * abs(float(case.E_expected))
    assert np.abs(r.density_matrix - case.P_expected).max() < 5e-6
    assert abs(r.chemical_potential - float(case.mu_expected)) < 5e-6


def test_case_from_scf_result_round_trip(tmp_path):
    case = CS.load_case(GOLDEN / "meshgrid_regression_case_v1.npz")
    K = O.OracleKernel(case.weights, case.hamiltonian, case.coulomb_q, float(case.T), include_hartree=True,
                       reference_density=case.reference_density, hartree_matrix=case.hartree_matrix)
    cfg = O.SCFConfig(max_iter=60, acceleration="diis", density_tol=1e-9, comm_tol=1e-8)
    r = O.solve_scf(K, case.P0.astype(np.complex128), case.n_electrons, config=cfg)
    inputs = dict(weights=case.weights, hamiltonian=case.hamiltonian, coulomb_q=case.coulomb_q, T=case.T,
                  hartree_matrix=case.hartree_matrix, reference_density=case.reference_density)
    new = CS.case_from_scf(inputs, r, P0=case.P0, n_electrons=case.n_electrons, scf_config=cfg, nk=12)
    out = tmp_path / "new.npz"
    CS.save_case(out, new)
    back = CS.load_case(out)
    assert int(back.k_expected) == r.iterations and float(back.E_expected) == r.energy
    np.testing.assert_array_equal(back.P_expected, r.density_matrix)
    np.testing.assert_array_equal(back.hist_E_expected, r.history["E"])
    np.testing.assert_array_equal(back.hist_dC_expected, r.history["commutator_residual"])
    assert int(back.diis_size) == cfg.diis_size and int(back.nk) == 12


def test_scan_results_round_trip(tmp_path):
    src = GOLDEN / "bilayer_reference.npz"
    meta, res = CS.load_scan(src)
    assert list(meta["branches"]) == ["PM", "SVP"] and int(meta["nk"]) == 49
    pts = [CS.scan_key(b, d) for b in meta["branches"] for d in meta["density_points_cm12"]]
    assert sorted(pts) == sorted(res)
    assert set(res[pts[0]]) == {"energy", "n_iter", "converged", "n_electrons", "mu"}
    out = tmp_path / "scan.npz"
    CS.save_scan(out, meta, res)
    with np.load(src, allow_pickle=False) as a, np.load(out, allow_pickle=False) as b:
        assert sorted(a.files) == sorted(b.files)
        for k in a.files:
            _same(a[k], b[k])


def test_malformed_inputs_are_refused(tmp_path):
    bad = tmp_path / "bad.npz"
    np.savez(bad, weights=np.ones((2, 2)))
    with pytest.raises(ValueError, match="missing"):
        CS.load_case(bad)
    case = CS.load_case(GOLDEN / "meshgrid_regression_case_v1.npz")
    case.weights = np.ones((3, 3))
    with pytest.raises(ValueError, match="weights"):
        CS.save_case(tmp_path / "x.npz", case)
    with pytest.raises(ValueError, match="contains"):
        CS.save_scan(tmp_path / "y.npz", {"a/b": 1}, {})
    obj = tmp_path / "obj.npz"
    np.savez(obj, T=np.array([{"x": 1}], dtype=object), weights=np.ones((2, 2)),
             hamiltonian=np.zeros((2, 2, 1, 1)), coulomb_q=np.zeros((2, 2, 1, 1)))
    with pytest.raises(ValueError):   # an object array needs pickle: refused
        CS.load_case(obj)
