"""GPU runs of PHI's DiagonalBath and FullBath coupling models through the
integrator (C ABI): PHI's known answers (hierarchy_integrator_test.cc:
388-505, 595-620; RK4 within 1e-6, RKF45 within 1e-3) and the oracle
(pinned in test_oracle_bath.py) within 1e-10 in FP64, with identical RKF45
step counts."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import phi_amd as P  # noqa: E402
from oracle import heom_oracle as HO  # noqa: E402
from oracle import td_oracle as TD  # noqa: E402

from test_oracle_bath import (K_WEAK_INDEPENDENT, K_WEAK_MATSUBARA, STATIC,  # noqa: E402
                              full_independent, diag_independent, problem)
from test_oracle_td import H2, HX, HZ, K_MEDIUM_ANNEAL, embed  # noqa: E402

pytestmark = pytest.mark.gpu


def params(model, ops, lh, gamma, lam, temperature=2.0, num_terms=1, H=None, Hi=None, Hf=None,
           full=2, method="rk4", precision="fp64"):
    p = P.PhiParameters()
    p.set_int("verbose", -2)
    p.set_double("hbar", 1.0)
    p.set_double("boltzmann_constant", 1.0)
    p.set_int("num_states", 2)
    p.set_int("hilbert_space_size", full)
    p.set_int("hilbert_space_num_elements", full * full)
    p.set_int("num_bath_couplings", len(gamma))
    p.set_int("num_matsubara_terms", num_terms)
    p.set_int("hierarchy_truncation_level", lh)
    p.set_double("temperature", temperature)
    p.set_array("gamma", gamma)
    p.set_array("lambda", lam)
    if H is not None:
        p.set_matrix("time_indepenent_hamiltonian", H)
    else:
        p.set_matrix("initial_annealing_hamiltonian", Hi)
        p.set_matrix("final_annealing_hamiltonian", Hf)
    ops = np.asarray(ops, dtype=float)
    if model == "full":
        p.set_int("is_full_bath_coupling", 1)
        p.set_array("bath_coupling_op", np.ascontiguousarray(ops.transpose(0, 2, 1)).reshape(-1))
    else:
        p.set_int("is_diagonal_bath_coupling", 1)
        p.set_array("bath_coupling_op", ops.reshape(-1))
    rho0 = np.zeros((full, full), dtype=complex)
    rho0[0, 0] = 1
    p.set_matrix("initial_density_matrix", rho0)
    p.set_double("timestep", 0.001)
    p.set_double("integration_time", 1.0)
    p.set_double("output_minimum_timestep", 0.01)
    p.set_string("output_filename", "")
    p.set_int("bulk_precision", 1 if precision == "fp32" else 0)
    assert p.fix(method) == 0
    return p


def launch(p):
    it = P.HierarchyIntegrator(p)
    it.Launch()
    return it


@pytest.mark.parametrize("name", sorted(STATIC))
def test_static_bath_models(name):
    (model, ops, lh), kw, gold = STATIC[name]
    print("case", name, flush=True)
    g = [1.0, 1.0]
    lam = [0.01, 0.01]
    # RK4: golden and the oracle's whole hierarchy
    it = launch(params(model, ops, lh, g, lam, H=H2, **kw))
    assert np.abs(it.density_matrix() - gold).max() < 1e-6
    orc = HO.Oracle(problem(model, ops, lh, **kw), threads=1)
    steps, _, _ = orc.rk4(1.0, 0.001, 0.01)
    assert it.info()["steps"] == steps
    assert np.abs(it.state() - orc.state()).max() <= 1e-10
    # RKF45: golden, the oracle's step counts and state
    it = launch(params(model, ops, lh, g, lam, H=H2, method="rkf45", **kw))
    assert np.abs(it.density_matrix() - gold).max() < 1e-3
    orc = HO.Oracle(problem(model, ops, lh, **kw), threads=1)
    res = orc.rkf45(1.0, 0.001, 1e-6, 1e-6, 0.01)
    info = it.info()
    assert info["steps"] == res["attempted"] and info["accepted"] == res["accepted"]
    assert np.abs(it.state() - orc.state()).max() <= 1e-10


@pytest.mark.parametrize("matsubara", [False, True])
def test_truncated_full_bath_rkf45(matsubara):
    kw = dict(temperature=0.5, num_terms=3) if matsubara else {}
    lh = 3 if matsubara else 2
    it = launch(params("full", full_independent(3), lh, [1.0] * 3, [0.01] * 3,
                       H=embed(H2, 3, 9.9), full=3, method="rkf45", **kw))
    gold = K_WEAK_MATSUBARA if matsubara else K_WEAK_INDEPENDENT
    assert np.abs(it.density_matrix() - embed(gold, 3)).max() < 1e-3
    ham = TD.Hamiltonian(2, H=embed(H2, 3, 9.9), bath_ops=full_independent(3))
    rho = np.zeros((3, 3), dtype=complex)
    rho[0, 0] = 1
    prob = HO.Problem(H=ham.H, rho0=ham.truncate(rho), gamma=[1.0] * 3, lambda_=[0.01] * 3,
                      lh=lh, hbar=1.0, kb=1.0, bath_model="full", bath_ops=ham.ops,
                      temperature=kw.get("temperature", 2.0), num_terms=kw.get("num_terms", 1))
    orc = HO.Oracle(prob, threads=1)
    res = orc.rkf45(1.0, 0.001, 1e-6, 1e-6, 0.01)
    assert it.info()["steps"] == res["attempted"]
    assert np.abs(it.state() - orc.state()).max() <= 1e-10


@pytest.mark.parametrize("method", ["rk4", "rkf45"])
@pytest.mark.parametrize("model,truncated", [("diagonal", False), ("full", False),
                                             ("full", True)])
def test_anneal_bath_models(model, truncated, method):
    print("case", model, truncated, method, flush=True)
    full = 3 if truncated else 2
    M = full
    if model == "diagonal":
        ops = diag_independent(M)
    else:
        ops = full_independent(M)
    Hi = embed(HX, 3, 9.9) if truncated else HX
    Hf = embed(HZ, 3, 9.9) if truncated else HZ
    it = launch(params(model, ops, 3, [0.5] * M, [0.5] * M, Hi=Hi, Hf=Hf, full=full,
                       method=method))
    gold = embed(K_MEDIUM_ANNEAL, 3) if truncated else K_MEDIUM_ANNEAL
    assert np.abs(it.density_matrix() - gold).max() < (1e-6 if method == "rk4" else 1e-3)
    ham = TD.Hamiltonian(2, H_initial=Hi, H_final=Hf, t_total=1.0,
                         bath_ops=ops if (truncated and model == "full") else None)
    rho = np.zeros((full, full), dtype=complex)
    rho[0, 0] = 1
    prob = HO.Problem(H=ham.H, rho0=ham.truncate(rho), gamma=[0.5] * M, lambda_=[0.5] * M,
                      temperature=2.0, lh=3, num_terms=1, hbar=1.0, kb=1.0, bath_model=model,
                      bath_ops=ham.ops if ham.ops is not None else ops)
    orc = HO.Oracle(prob, threads=1)
    if method == "rk4":
        TD.rk4(orc, ham, 0.0, 1.0, 0.001)
    else:
        attempts, accepted, _ = TD.rkf45(orc, ham, 0.0, 1.0, 0.001, 1e-6, 1e-6)
        info = it.info()
        assert (info["steps"], info["accepted"]) == (attempts, accepted)
    assert np.abs(it.state() - orc.state()).max() <= 1e-10


def test_bath_models_reject_fp32():
    (model, ops, lh), kw, gold = STATIC["weak_full"]
    p = params(model, ops, lh, [1.0, 1.0], [0.01, 0.01], H=H2, precision="fp32")
    it = P.HierarchyIntegrator(p)
    assert it.Launch(check=False) != 0
    assert "FP64 only" in P.lib().phi_hip_last_error().decode()
