"""GPU runs of time-local truncation (PHI's TimeLocal option) through the
integrator (C ABI): PHI's nine TimeLocal known answers
(hierarchy_integrator_test.cc:505-552, 625-716; RK4 within 1e-6, RKF45
within 1e-3), the oracle's whole hierarchy (pinned in
test_oracle_timelocal.py) within 1e-10 with the same RKF45 attempt and
accepted-step counts, and restart files that carry the saved time-local
matrices (restart_helper.cc:30-84) and resume to the oracle's split run;
plus RK4 restarts that continue an uninterrupted run exactly."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import phi_amd as P  # noqa: E402
from oracle import td_oracle as TD  # noqa: E402

from test_oracle_timelocal import CASES, setup  # noqa: E402

pytestmark = pytest.mark.gpu


def params(name, method, precision="fp64", t_end=None, timelocal=True, restart_out=None,
           restart_in=None):
    n, full, M, bath, Hspec, model, ops, lh, T, _, _ = CASES[name]
    p = P.PhiParameters()
    p.set_int("verbose", -2)
    p.set_double("hbar", 1.0)
    p.set_double("boltzmann_constant", 1.0)
    p.set_int("num_states", n)
    p.set_int("hilbert_space_size", full)
    p.set_int("hilbert_space_num_elements", full * full)
    p.set_int("num_bath_couplings", M)
    p.set_int("num_matsubara_terms", 1)
    p.set_int("hierarchy_truncation_level", lh)
    p.set_double("temperature", 2.0)
    p.set_array("gamma", [bath["gamma"]] * M)
    p.set_array("lambda", [bath["lam"]] * M)
    if isinstance(Hspec, tuple):
        p.set_matrix("initial_annealing_hamiltonian", Hspec[0])
        p.set_matrix("final_annealing_hamiltonian", Hspec[1])
    else:
        p.set_matrix("time_indepenent_hamiltonian", Hspec)
    if model == "full":
        p.set_int("is_full_bath_coupling", 1)
        p.set_array("bath_coupling_op",
                    np.ascontiguousarray(np.asarray(ops, dtype=float).transpose(0, 2, 1)).reshape(-1))
    elif model == "diagonal":
        p.set_int("is_diagonal_bath_coupling", 1)
        p.set_array("bath_coupling_op", np.asarray(ops, dtype=float).reshape(-1))
    p.set_int("use_time_local_truncation", 1 if timelocal else 0)
    rho0 = np.zeros((full, full), dtype=complex)
    rho0[0, 0] = 1
    p.set_matrix("initial_density_matrix", rho0)
    p.set_double("timestep", 0.001)
    p.set_double("integration_time", T if t_end is None else t_end)
    p.set_double("output_minimum_timestep", 0.01)
    p.set_string("output_filename", "")
    p.set_int("bulk_precision", 1 if precision == "fp32" else 0)
    if restart_out:
        p.set_double("restart_output_time", 1.0)
        p.set_string("restart_output_filename", restart_out)
    if restart_in:
        p.set_int("is_restarted", 1)
        p.set_string("restart_input_filename", restart_in)
    assert p.fix(method) == 0
    return p


def launch(p):
    it = P.HierarchyIntegrator(p)
    it.Launch()
    return it


@pytest.mark.parametrize("name,method", [(k, m) for k, v in sorted(CASES.items()) for m in v[9]])
def test_timelocal_goldens_and_oracle(name, method):
    print("case", name, method, flush=True)
    it = launch(params(name, method))
    gold = CASES[name][10]
    r = it.density_matrix()
    assert r.shape == gold.shape
    assert np.abs(r - gold).max() < (1e-6 if method == "rk4" else 1e-3)
    orc, ham, tl, T = setup(name)
    if method == "rk4":
        TD.rk4(orc, ham, 0.0, T, 0.001, tl=tl)
    else:
        attempts, accepted, _ = TD.rkf45(orc, ham, 0.0, T, 0.001, 1e-6, 1e-6, tl=tl)
        info = it.info()
        assert (info["steps"], info["accepted"]) == (attempts, accepted)
    assert np.abs(it.state() - orc.state()).max() <= 1e-10


def test_timelocal_rejects_fp32():
    it = P.HierarchyIntegrator(params("medium_full", "rk4", precision="fp32"))
    assert it.Launch(check=False) != 0
    assert "FP64 only" in P.lib().phi_hip_last_error().decode()


def read_restart(path, nodes, ne):
    with open(path) as f:
        t = float(f.readline())
        dt = float(f.readline())
        rows = []
        for line in f:
            vals = [complex(*map(float, v.strip("()").split(","))) for v in line.split()]
            rows.append(vals)
    assert all(len(r) == ne for r in rows)
    return t, dt, np.array(rows[:nodes]), np.array(rows[nodes:])


@pytest.mark.parametrize("name", ["medium_full", "medium_independent"])
def test_timelocal_rkf45_restart(name, tmp_path):
    # first half with a restart file, second half resumed from it; the
    # oracle runs the same split with the dt and matrices the file holds
    f = str(tmp_path / "restart.txt")
    a = launch(params(name, "rkf45", t_end=0.5, restart_out=f))
    nodes = int(a.info()["num_nodes"])
    t, dt, rho, T = read_restart(f, nodes, 4)
    assert abs(t - 0.5) < 1e-12
    # the step size before the end clamp (the clamped one would be ~0)
    assert dt > 1e-4
    orc, ham, tl, _ = setup(name)
    TD.rkf45(orc, ham, 0.0, 0.5, 0.001, 1e-6, 1e-6, tl=tl)
    # M x K saved matrices, column-major lines, at 17 digits
    assert T.shape == (2, 4)
    saved = tl.saved.reshape(2, 2, 2)
    assert np.abs(T - saved.transpose(0, 2, 1).reshape(2, 4)).max() < 1e-15
    b = launch(params(name, "rkf45", restart_in=f))
    TD.rkf45(orc, ham, t, 1.0, dt, 1e-6, 1e-6, tl=tl)
    assert np.abs(b.state() - orc.state()).max() <= 1e-10


def test_rk4_restart_continues_exactly(tmp_path):
    # plain RK4 (no time-local term): a resumed run equals the uninterrupted
    # one (times are accumulated from the file's lossless start time)
    f = str(tmp_path / "restart.txt")
    launch(params("medium_independent", "rk4", t_end=0.5, timelocal=False, restart_out=f))
    b = launch(params("medium_independent", "rk4", timelocal=False, restart_in=f))
    c = launch(params("medium_independent", "rk4", timelocal=False))
    assert b.info()["final_time"] == c.info()["final_time"]
    assert np.abs(b.state() - c.state()).max() <= 1e-14
