"""GPU runs of time-dependent (annealing) and Hilbert-space-truncated
systems through the integrator (C ABI): PHI's known answers
(hierarchy_integrator_test.cc:350-373, 554-592) within 1e-6, and the whole
hierarchy against the oracle (tests/test_oracle_td.py pins it) within 1e-10
in FP64."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import phi_amd as P  # noqa: E402
from oracle import heom_oracle as HO  # noqa: E402
from oracle import td_oracle as TD  # noqa: E402

from test_oracle_td import (H2, HX, HZ, K_MEDIUM_ANNEAL, K_VACUUM,  # noqa: E402
                            K_VACUUM_ANNEAL, embed)

pytestmark = pytest.mark.gpu


def params(n, lh, gamma, lam, full=None, H=None, Hi=None, Hf=None, method="rk4",
           precision="fp64"):
    """PhiParameters set field by field as PHI's tests do
    (hierarchy_integrator_test.cc:27-44, 56-141)."""
    full = full or n
    p = P.PhiParameters()
    p.set_int("verbose", -2)
    p.set_double("hbar", 1.0)
    p.set_double("boltzmann_constant", 1.0)
    p.set_int("num_states", n)
    p.set_int("hilbert_space_size", full)
    p.set_int("hilbert_space_num_elements", full * full)
    p.set_int("num_bath_couplings", len(gamma))
    p.set_int("num_matsubara_terms", 1)
    p.set_int("hierarchy_truncation_level", lh)
    p.set_double("temperature", 2.0)
    p.set_array("gamma", gamma)
    p.set_array("lambda", lam)
    if H is not None:
        p.set_matrix("time_indepenent_hamiltonian", H)
    else:
        p.set_matrix("initial_annealing_hamiltonian", Hi)
        p.set_matrix("final_annealing_hamiltonian", Hf)
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


CASES = {
    "anneal_vacuum": (dict(n=2, lh=1, gamma=[0.5], lam=[0.5], Hi=HX, Hf=HZ), K_VACUUM_ANNEAL),
    "anneal_medium": (dict(n=2, lh=3, gamma=[0.5, 0.5], lam=[0.5, 0.5], Hi=HX, Hf=HZ),
                      K_MEDIUM_ANNEAL),
    "truncated_vacuum": (dict(n=2, lh=1, gamma=[1.0], lam=[0.01], full=3, H=embed(H2, 3, 9.9)),
                         embed(K_VACUUM, 3)),
    "truncated_anneal_vacuum": (dict(n=2, lh=1, gamma=[0.5], lam=[0.5], full=3,
                                     Hi=embed(HX, 3, 9.9), Hf=embed(HZ, 3, 9.9)),
                                embed(K_VACUUM_ANNEAL, 3)),
    # no published answer: a coupled 4-level system truncated to its 3 lowest
    # states, annealed, with independent baths on the eigenstates at depth 3
    # (the rotation of every live ADO buffer inside the step); oracle only
    "truncated_anneal_medium": (dict(n=3, lh=3, gamma=[0.5, 0.6, 0.7], lam=[0.5, 0.4, 0.3],
                                     full=4,
                                     Hi=np.array([[0, 1, 0.2, 0], [1, 0, 1, 0.3], [0.2, 1, 0, 1],
                                                  [0, 0.3, 1, 0]], dtype=complex),
                                     Hf=np.diag([1.0, -1.0, 0.5, 3.0]).astype(complex)),
                                None),
}


def oracle_run(kw, method="rk4"):
    kw = dict(kw)
    n, lh, gamma, lam = kw.pop("n"), kw.pop("lh"), kw.pop("gamma"), kw.pop("lam")
    full = kw.pop("full", None) or n
    ham = TD.Hamiltonian(n, H=kw.get("H"), H_initial=kw.get("Hi"), H_final=kw.get("Hf"),
                         t_total=1.0)
    rho_full = np.zeros((full, full), dtype=complex)
    rho_full[0, 0] = 1
    prob = HO.Problem(H=ham.H, rho0=ham.truncate(rho_full), gamma=gamma, lambda_=lam,
                      temperature=2.0, lh=lh, num_terms=1, hbar=1.0, kb=1.0)
    orc = HO.Oracle(prob, threads=1)
    if method == "rk4":
        TD.rk4(orc, ham, 0.0, 1.0, 0.001)
        return orc, ham
    # PHI's defaults: RKF45Tolerance 1e-6, RKF45MinimumTimestep 1e-6
    counts = TD.rkf45(orc, ham, 0.0, 1.0, 0.001, 1e-6, 1e-6)
    return orc, ham, counts


@pytest.mark.parametrize("name", sorted(CASES))
def test_td_and_truncated_goldens_and_oracle(name):
    kw, gold = CASES[name]
    print("case", name, flush=True)
    it = P.HierarchyIntegrator(params(**kw))
    it.Launch()
    assert it.info()["steps"] == 1000
    r = it.density_matrix()
    if gold is not None:
        assert r.shape == gold.shape
        assert np.abs(r - gold).max() < 1e-6
    orc, ham = oracle_run(kw)
    assert np.abs(r - ham.untruncate(orc.state()[0])).max() <= 1e-10
    # whole hierarchy (eigenbasis of the final H for a truncated system;
    # both use the same phase convention)
    assert np.abs(it.state() - orc.state()).max() <= 1e-10


def test_anneal_fp32_bulk_populations():
    kw, gold = CASES["anneal_medium"]
    it = P.HierarchyIntegrator(params(**kw, precision="fp32"))
    it.Launch()
    r = it.density_matrix()
    assert np.abs(np.diag(r).real - np.diag(gold).real).max() < 1e-5


TD_CASES = sorted(k for k in CASES if "anneal" in k)


@pytest.mark.parametrize("name", TD_CASES)
def test_anneal_rkf45_goldens_and_oracle(name):
    # PHI's RunRkf45Test (within 1e-3) and the oracle's RKF45 loop with H
    # updates at the stage times: the same attempts and accepted steps, the
    # same hierarchy within 1e-10
    kw, gold = CASES[name]
    print("case", name, flush=True)
    it = P.HierarchyIntegrator(params(**kw, method="rkf45"))
    it.Launch()
    r = it.density_matrix()
    if gold is not None:
        assert np.abs(r - gold).max() < 1e-3
    orc, ham, (attempts, accepted, _) = oracle_run(kw, "rkf45")
    info = it.info()
    assert (info["steps"], info["accepted"]) == (attempts, accepted)
    assert abs(info["final_time"] - 1.0) < 1e-12
    assert np.abs(r - ham.untruncate(orc.state()[0])).max() <= 1e-10
    assert np.abs(it.state() - orc.state()).max() <= 1e-10


def test_anneal_rkf45_fp32_bulk():
    # FP32 bulk with the rotations applied to FP32 stage inputs and k's:
    # the FP64 run's density matrix within 1e-5
    kw, _ = CASES["truncated_anneal_medium"]
    ref = P.HierarchyIntegrator(params(**kw, method="rkf45"))
    ref.Launch()
    it = P.HierarchyIntegrator(params(**kw, method="rkf45", precision="fp32"))
    it.Launch()
    assert np.abs(it.density_matrix() - ref.density_matrix()).max() < 1e-5
