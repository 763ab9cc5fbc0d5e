"""Error paths of the C ABI on the CPU (no GPU calls reach a kernel):
over-size hierarchies are refused before anything is enumerated, run
parameters a loop cannot finish with are refused before anything is
allocated, short bath-index lists are refused, malformed `.prm` files
never crash the reader (its decisions equal the oracle's), and null contexts
are refused."""

import os
import random

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

import phi_amd as P
from oracle import prm_oracle as PO

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = os.path.join(ROOT, "phi_amd", "configs")
C1 = os.path.join(CFG, "c1.prm")
C2 = os.path.join(CFG, "c2.prm")


def launch_rc(p):
    p.set_int("verbose", -2)
    p.set_string("output_filename", "")
    it = P.HierarchyIntegrator(p)
    return it.Launch(check=False), P.lib().phi_hip_last_error().decode()


def test_oversize_hierarchy_is_refused_before_enumeration():
    """7 sites at HierarchyTruncation 200: C(206, 7) ~ 1.6e12 nodes."""
    p = P.PhiParameters.read(C2, "rk4")
    p.set_int("hierarchy_truncation_level", 200)
    with pytest.raises(P.PhiError) as e:
        P.Context(p)
    assert e.value.code == -2
    assert "hierarchy too large" in str(e.value)
    p.set_int("run_method", P.RUN_METHODS.index("PRINTHIERARCHY"))
    rc, msg = launch_rc(p)
    assert rc == -2 and "hierarchy too large" in msg


def test_largest_allowed_state_is_not_refused_by_the_size_check():
    # 2 states, 4 modes, HierarchyTruncation 40: C(43, 4) = 123410 nodes
    # pass the bound (the run then needs a device: without one it fails
    # for that reason only)
    p = P.PhiParameters.read(C1, "rk4")
    p.set_int("hierarchy_truncation_level", 40)
    with pytest.raises(P.PhiError) as e:
        P.Context(p)
    if P.device_count() == 0:
        assert e.value.code == -7


@pytest.mark.parametrize("field,value,method,what", [
    ("timestep", 0.0, "rk4", "timestep"),
    ("timestep", -1e-3, "rk4", "timestep"),
    ("timestep", float("nan"), "rk4", "timestep"),
    ("output_minimum_timestep", 0.0, "rk4", "OutputMinimumTimestep"),
    ("output_minimum_timestep", -1.0, "rk4", "OutputMinimumTimestep"),
    ("output_minimum_timestep", 1e30, "rk4", "OutputMinimumTimestep"),
    ("integration_time", float("inf"), "rk4", "finite"),
    ("integration_time", 1e300, "rk4", "1e15"),
    ("timestep", 0.0, "rkf45", "timestep"),
    ("rkf45_tolerance", 0.0, "rkf45", "RKF45tolerance"),
    ("rkf45_minimum_timestep", 0.0, "rkf45", "RKF45mindt"),
    ("bicgstab_tolerance", 0.0, "bicgstab", "BiCGSTAB"),
])
def test_run_parameters_a_loop_cannot_finish_with_are_refused(field, value, method, what):
    """Set through the C ABI (which skips the reader's checks): Launch
    returns PHI_HIP_EPARAMS before any device allocation."""
    p = P.PhiParameters.read(C1, method)
    p.set_double(field, value)
    rc, msg = launch_rc(p)
    assert rc == -2, msg
    assert what in msg


def test_short_bath_index_list_is_refused():
    p = P.PhiParameters.read(C2, "rk4")
    p.set_array("multiple_independent_bath_indices", [0, 1, 2])  # 7 baths
    with pytest.raises(P.PhiError) as e:
        P.Context(p)
    assert e.value.code == -2 and "one index per bath" in str(e.value)


def test_restart_from_missing_file_is_an_io_error(tmp_path):
    p = P.PhiParameters.read(C1, "rk4")
    p.set_string("restart_input_filename", str(tmp_path / "nope.txt"))
    p.set_int("is_restarted", 1)
    rc, msg = launch_rc(p)
    # no device here: the engine is created first; with one, the read fails
    assert rc in (-5, -7)


KEYS = ["NumStates", "HierarchyTruncation", "MatsubaraTerms", "Temperature", "Timestep",
        "Time", "OutputMinimumTimestep", "Lambda:", "Gamma:", "Hamiltonian:",
        "InitialDensityMatrix:", "NumCouplingTerms", "RKF45tolerance", "RKF45mindt",
        "TimeLocal", "FilterTolerance", "RestartOutputTimeSecs", "CorrelatedCouplingTerms:",
        "FullBathCouplingTerms:", "HilbertSpaceSize", "DiagonalCouplingIndices"]
VALUES = ["2", "-1", "0", "1e400", "nan", "1,2", "(1,0),(0,1)", "", "abc", "3.0.1", "1e-8",
          "2,2,2", "((1,2)", ",", "7", "255", "99999999999"]


@settings(max_examples=300, deadline=None)
@given(st.lists(st.tuples(st.sampled_from(KEYS), st.sampled_from(VALUES)), max_size=14),
       st.booleans())
def test_malformed_prm_files_never_crash(tmp_path_factory, entries, newline):
    """Random key/value files (keys out of order, repeated, truncated
    matrices, out-of-range numbers): the C++ reader returns, and its
    validity decision equals the oracle's."""
    lines = []
    for k, v in entries:
        if k.endswith(":"):
            lines += [k, v, v]
        else:
            lines.append("%s=%s" % (k, v))
    text = "\n".join(lines) + ("\n" if newline else "")
    path = tmp_path_factory.mktemp("prm") / "x.prm"
    path.write_text(text)
    for method in ("rk4", "rkf45"):
        p = P.PhiParameters.read(str(path), method, check=False)
        try:
            o = PO.read_command_line(str(path), method)
            want = o.are_parameters_valid
        except Exception:
            continue  # the oracle's restatement gave up; only "no crash" is checked
        assert bool(p.get_int("are_parameters_valid")) == want


def test_binary_junk_prm_never_crashes(tmp_path):
    rng = random.Random(5)
    for i in range(40):
        data = bytes(rng.randrange(256) for _ in range(rng.randrange(1, 400)))
        path = tmp_path / ("junk%d.prm" % i)
        path.write_bytes(data)
        p = P.PhiParameters.read(str(path), "rk4", check=False)
        assert p.get_int("are_parameters_valid") in (0, 1)


def test_node_count_estimate_of_huge_hierarchies_saturates(tmp_path):
    """PHI's double sum of factorials turns into NaN past 170! and its
    integer conversion is then undefined: the reader and the oracle both
    stop early and report INT_MAX (HierarchyTruncation 5000 with 255 * 256
    modes would otherwise take minutes in the oracle)."""
    import time
    text = open(C1).read().replace("HierarchyTruncation=9", "HierarchyTruncation=5000")
    text += "MatsubaraTerms=255\nNumCouplingTerms=255\n"
    path = tmp_path / "huge.prm"
    path.write_text(text)
    t0 = time.perf_counter()
    p = P.PhiParameters.read(str(path), "rk4", check=False)
    o = PO.read_command_line(str(path), "rk4")
    assert time.perf_counter() - t0 < 5
    assert p.get_int("matrix_count") == o.matrix_count == PO.INT_MAX
    assert PO.node_count_estimate(9, 4) == 495          # C(8+4, 4), the 2-site c1 case
    assert PO.node_count_estimate(5, 24) == 20475       # the 24-site ring at depth 4


def test_context_entry_points_refuse_a_null_context():
    """Entry points that take a context return PHI_HIP_EINVAL with a message
    for a null one (checked before any device call), the launch-shape calls
    among them."""
    import ctypes
    L = P.lib()
    ip = (ctypes.c_int * 4)()
    dp = ctypes.c_double()
    calls = {
        "phi_hip_set_launch": (None, 2, 0, 0, 0),
        "phi_hip_set_stage_launch": (None, 0, 0, 0),
        "phi_hip_stage_launch": (None, ip, ip),
        "phi_hip_autotune": (None, 3, 0.001, ctypes.byref(dp), ip),
        "phi_hip_get_index": (None, ip, 4),
        "phi_hip_set_rk4_form": (None, 2),
        "phi_hip_set_graphs": (None, 1),
        "phi_hip_steady_state": (None, 1e-10, 10, ip, ctypes.byref(dp)),
        "phi_hip_steady_state_l": (None, 2, 1e-10, 10, ip, ctypes.byref(dp)),
    }
    for name, args in calls.items():
        rc = getattr(L, name)(*args)
        assert rc == -1, (name, rc)
        assert L.phi_hip_last_error().decode(), name
