"""The drop-in boundary on the CPU: `.prm` reader, output header, value
checks, C-ABI exports and the `phi` program (no GPU calls).

The oracle's reader/header (oracle/prm_oracle.py) is pinned by the
reference's own example output (example/one_qubit_100.txt); the product's
C++ reader and header (through the C ABI) must then agree with the oracle on
every input below, valid or not.
"""

import os
import re
import subprocess

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

import phi_amd as P
from oracle import prm_oracle as PO

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "tests", "golden", "reference_data")
CFG = os.path.join(ROOT, "phi_amd", "configs")
EXAMPLE = os.path.join(REF, "one_qubit.prm")


def test_oracle_header_matches_reference_example_output():
    """example/one_qubit.prm -> the header PHI wrote in one_qubit_100.txt."""
    p = PO.read_command_line(EXAMPLE, "rkf45")
    hdr = PO.full_header(p, 3)
    ref = open(os.path.join(REF, "one_qubit_100.txt")).read()
    assert ref.startswith(hdr)
    assert hdr.count("\n") == 48


SCALARS_D = ["hbar", "boltzmann_constant", "temperature", "integration_time", "timestep",
             "output_minimum_timestep", "rkf45_minimum_timestep", "rkf45_tolerance",
             "bicgstab_tolerance", "filter_tolerance", "minimum_field_change",
             "restart_output_time"]
SCALARS_I = ["verbose", "num_states", "hilbert_space_size", "num_bath_couplings",
             "num_matsubara_terms", "hierarchy_truncation_level", "matrix_count",
             "bicgstab_ell"]
FLAGS = ["use_time_local_truncation", "is_rho_normalized", "is_diagonal_bath_coupling",
         "is_full_bath_coupling", "is_truncated_system", "is_restarted",
         "stupid_hierarchy_partition", "are_parameters_valid"]
MATRICES = ["time_indepenent_hamiltonian", "initial_annealing_hamiltonian",
            "final_annealing_hamiltonian", "initial_density_matrix"]


def assert_same_params(path, method):
    o = PO.read_command_line(path, method)
    p = P.PhiParameters.read(path, method, check=False)
    assert bool(p.get_int("are_parameters_valid")) == o.are_parameters_valid
    assert p.run_method == o.run_method
    if not o.are_parameters_valid:
        return o, p
    for f in SCALARS_D:
        assert p.get_double(f) == getattr(o, f), f
    for f in SCALARS_I:
        assert p.get_int(f) == getattr(o, f), f
    for f in FLAGS:
        assert p.get_int(f) == int(getattr(o, f)), f
    for f, of in (("gamma", "gamma"), ("lambda", "lambda_"), ("kappa", "kappa"),
                  ("initial_hamiltonian_field", "initial_hamiltonian_field"),
                  ("final_hamiltonian_field", "final_hamiltonian_field")):
        ov = getattr(o, of)
        pv = p.get_array(f)
        assert np.array_equal(pv, np.asarray(ov if ov is not None else [], dtype=float)), f
    assert np.array_equal(p.get_array("multiple_independent_bath_indices"),
                          np.asarray(o.multiple_independent_bath_indices, dtype=float))
    for f in MATRICES:
        pm, om = p.get_matrix(f), getattr(o, f)
        assert (pm is None) == (om is None), f
        if om is not None:
            assert np.array_equal(pm, om), f
    assert p.get_string("output_filename") == o.output_filename
    assert p.header(o.matrix_count) == PO.full_header(o, o.matrix_count)
    return o, p


@pytest.mark.parametrize("path,method", [
    (EXAMPLE, "rkf45"), (EXAMPLE, "rk4"),
    (os.path.join(CFG, "c1.prm"), "rk4"), (os.path.join(CFG, "c2.prm"), "rk4"),
    (os.path.join(CFG, "c2.prm"), "bicgstab"), (os.path.join(CFG, "c3.prm"), "rkf45")])
def test_reader_matches_oracle_on_inputs(path, method):
    assert_same_params(path, method)


BASE = """NumStates=2
HierarchyTruncation=3
Temperature=300
Hamiltonian:
100,20
20,0
Lambda:
35,35
Gamma:
20,20
InitialDensityMatrix:
1,0
0,0
"""

EDGE_CASES = {
    "plain": BASE,
    "comments_between_rows": BASE.replace("100,20\n", "100,20\n# a comment, skipped\n"),
    "complex_rows": BASE.replace("100,20\n20,0", "(100,0),(20,-1)\n(20,1),(0,0)"),
    "unknown_keys_ignored": BASE + "RKF45Mindt=1e-06\nSomethingElse=3\n",
    "no_final_newline_drops_last_line": BASE + "Time=2",
    "final_newline_keeps_last_line": BASE + "Time=2\n",
    "aliases": BASE.replace("NumStates", "Ns").replace("HierarchyTruncation", "Lh") + "dt=0.002\nt=0.5\n",
    "matsubara_and_normalisation": BASE + "MatsubaraTerms=2\nrho_normalized=0\n",
    "blank_line_ends_matrix": BASE.replace("Lambda:", "\nLambda:"),
    "negative_states": BASE.replace("NumStates=2", "NumStates=-2"),
    "bad_float": BASE + "Temperature=3.0.1\n",
    "trailing_comma": BASE.replace("35,35", "35,35,"),
    "too_few_values": BASE.replace("35,35", "35"),
    "crlf": BASE.replace("\n", "\r\n"),
    "no_hamiltonian": "NumStates=2\nLambda:\n1,1\n",
    "spaces_around_values": BASE.replace("Temperature=300", "Temperature= 300 "),
    "hamiltonian_twice": BASE + "Hamiltonian:\n1,0\n0,1\n",
    "exponent_values": BASE + "RKF45tolerance=1e-08\nRKF45mindt=2.5E-7\n",
    "bool_values": BASE + "TimeLocal=1\nCyclicPartition=0\n",
    "bad_bool": BASE + "TimeLocal=2\n",
    "empty_value": BASE + "Temperature=\n",
    "multibath": BASE + "NumCouplingTerms=2\nDiagonalCouplingIndices=1,0\n",
}


@pytest.mark.parametrize("name", sorted(EDGE_CASES))
@pytest.mark.parametrize("method", ["rk4", "rkf45", "bicgstab"])
def test_reader_edge_cases(tmp_path, name, method):
    path = tmp_path / (name + ".prm")
    path.write_bytes(EDGE_CASES[name].encode())
    assert_same_params(str(path), method)


def test_last_line_without_newline_is_not_read(tmp_path):
    path = tmp_path / "x.prm"
    path.write_text(BASE + "Time=2")
    p = P.PhiParameters.read(str(path), "rk4")
    assert p.get_double("integration_time") == 0.0


def test_matrices_are_column_major(tmp_path):
    """Row r of a `.prm` matrix fills element (r, j) (phi_parameters.cc:356)."""
    path = tmp_path / "x.prm"
    path.write_text(BASE.replace("100,20\n20,0", "1,2\n3,4"))
    p = P.PhiParameters.read(str(path), "rk4")
    assert np.array_equal(p.get_matrix("time_indepenent_hamiltonian"), [[1, 2], [3, 4]])


ALPHABET = " 0123456789.eE+-,()x"


@settings(max_examples=600, deadline=None)
@given(st.text(alphabet=ALPHABET, min_size=1, max_size=14))
def test_value_checks_match_oracle(text):
    """ParameterInput's IsValid* decisions (parameter_input.cc:341-731)."""
    checks = [("int", PO.is_valid_integer), ("float", PO.is_valid_float),
              ("bool", PO.is_valid_bool), ("reals", PO.is_valid_array),
              ("complexes", PO.is_valid_complex_array)]
    for kind, fn in checks:
        try:
            want = bool(fn(text))
        except PO.ParseError:
            want = False
        assert P.validate(kind, text) == want, (kind, text)


def test_capi_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "phi_hip.h")).read()
    names = sorted(set(re.findall(r"\b(phi_hip_\w+)\s*\(", header)))
    assert len(names) >= 40
    lib = P.lib()
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", P.LIB_PATH], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (phi_hip_\w+)", out))
    assert set(names) <= exported


def test_params_fields_roundtrip():
    p = P.PhiParameters()
    assert p.get_double("hbar") == 5.3088354
    assert p.get_int("num_matsubara_terms") == 1
    p.set_double("temperature", 77.0)
    p.set_array("gamma", [1.0, 2.0])
    p.set_matrix("time_indepenent_hamiltonian", np.array([[1, 2 - 1j], [2 + 1j, 3]]))
    assert p.get_double("temperature") == 77.0
    assert np.array_equal(p.get_array("gamma"), [1.0, 2.0])
    assert np.array_equal(p.get_matrix("time_indepenent_hamiltonian"),
                          np.array([[1, 2 - 1j], [2 + 1j, 3]]))
    with pytest.raises(P.PhiError):
        p.get_double("no_such_field")


@pytest.mark.skipif(P.device_count() > 0, reason="checks the no-device error path")
def test_no_device_is_reported_not_crashed():
    p = P.PhiParameters.read(os.path.join(CFG, "c1.prm"), "rk4")
    with pytest.raises(P.PhiError) as e:
        P.Context(p)
    assert e.value.code == -7
    p.set_int("verbose", -2)
    it = P.HierarchyIntegrator(p)
    assert it.Launch(check=False) == -7


def test_cli_usage_and_memory_mode(tmp_path):
    r = P.run_cli([], timeout=60)
    assert "Usage" in r.stdout
    r = P.run_cli([str(tmp_path / "missing.prm"), "rk4"], timeout=60)
    assert "ERROR: No parameter file." in r.stdout
    r = P.run_cli([os.path.join(CFG, "c2.prm"), "memory"], timeout=60)
    assert "Total number of matrices: 6435" in r.stdout
    assert "rk4:" in r.stdout
