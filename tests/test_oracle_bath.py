"""Pins of the oracle's DiagonalBath and FullBath coupling models against
PHI's own known answers (hierarchy_integrator_test.cc:388-505, 595-620):
RK4 within 1e-6 (RunRk4Test) and RKF45 within 1e-3 (RunRkf45Test), with
the fixtures' operators (:163-230).  Coupling operators that are site
projectors must reproduce the IndependentBath answers, which pins the
diagonal and full forms against the already pinned independent one."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import heom_oracle as HO  # noqa: E402
from oracle import td_oracle as TD  # noqa: E402

from test_oracle_td import H2, HX, HZ, K_MEDIUM_ANNEAL, colmajor, embed  # noqa: E402

K_WEAK_INDEPENDENT = colmajor([0.790012, 0.206755 - 0.346119j, 0.206755 + 0.346119j, 0.209988], 2)
K_WEAK_DIAGONAL = colmajor([0.793056, 0.194223 - 0.336944j, 0.194223 + 0.336944j, 0.206944], 2)
K_WEAK_FULL = colmajor([0.744586, 0.200681 - 0.309559j, 0.200681 + 0.309559j, 0.255414], 2)
K_WEAK_MATSUBARA = colmajor([0.789331, 0.208929 - 0.348365j, 0.208929 + 0.348365j, 0.210669], 2)


def diag_independent(n):
    """SetDiagonalIndependentBathCouplingTerms: bath i = diag(e_i)."""
    return np.eye(n)


def full_independent(n):
    """SetFullIndependentBathCouplingTerms: bath i = |i><i|."""
    ops = np.zeros((n, n, n))
    for i in range(n):
        ops[i, i, i] = 1
    return ops


def diag_alternating(m, n):
    """SetDiagonalBathCouplingTerms: bath i entry j = (-1)^(i+j)."""
    return np.array([[(-1.0) ** (i + j) for j in range(n)] for i in range(m)])


def full_diagonal(m, n):
    """SetFullDiagonalBathCouplingTerms: bath i = diag((-1)^(i+j))."""
    return np.array([np.diag(diag_alternating(m, n)[i]) for i in range(m)])


def full_alternating(m, n):
    """SetFullBathCouplingTerms: bath i element (j, k) = (-1)^(i+j+k)."""
    return np.array([[[(-1.0) ** (i + j + k) for k in range(n)] for j in range(n)]
                     for i in range(m)])


WEAK = dict(gamma=[1.0, 1.0], lambda_=[0.01, 0.01])


def problem(model, ops, lh, temperature=2.0, num_terms=1, H=H2, gamma=None, lam=None, n=2):
    rho0 = np.zeros((n, n), dtype=complex)
    rho0[0, 0] = 1
    return HO.Problem(H=H, rho0=rho0, gamma=gamma or WEAK["gamma"], lambda_=lam or WEAK["lambda_"],
                      temperature=temperature, lh=lh, num_terms=num_terms, hbar=1.0, kb=1.0,
                      bath_model=model, bath_ops=ops)


def run_rk4(prob):
    orc = HO.Oracle(prob, threads=1)
    steps, _, _ = orc.rk4(1.0, 0.001, 0.01)
    assert steps == 1000
    return orc.state()[0]


def run_rkf45(prob):
    orc = HO.Oracle(prob, threads=1)
    orc.rkf45(1.0, 0.001, 1e-6, 1e-6, 0.01)
    return orc.state()[0]


STATIC = {
    "weak_independent_diagonal": (("diagonal", diag_independent(2), 2), {}, K_WEAK_INDEPENDENT),
    "weak_independent_full": (("full", full_independent(2), 2), {}, K_WEAK_INDEPENDENT),
    "weak_diagonal": (("diagonal", diag_alternating(2, 2), 2), {}, K_WEAK_DIAGONAL),
    "weak_full_diagonal": (("full", full_diagonal(2, 2), 2), {}, K_WEAK_DIAGONAL),
    "weak_full": (("full", full_alternating(2, 2), 2), {}, K_WEAK_FULL),
    "weak_diagonal_matsubara": (("diagonal", diag_independent(2), 3),
                                dict(temperature=0.5, num_terms=3), K_WEAK_MATSUBARA),
    "weak_full_matsubara": (("full", full_independent(2), 3),
                            dict(temperature=0.5, num_terms=3), K_WEAK_MATSUBARA),
}


@pytest.mark.parametrize("name", sorted(STATIC))
def test_static_bath_goldens(name):
    (model, ops, lh), kw, gold = STATIC[name]
    r = run_rk4(problem(model, ops, lh, **kw))
    assert np.abs(r - gold).max() < 1e-6
    r = run_rkf45(problem(model, ops, lh, **kw))
    assert np.abs(r - gold).max() < 1e-3


def truncated_problem(H_full, ops_full, lh, gamma, lam, temperature=2.0, num_terms=1):
    ham = TD.Hamiltonian(2, H=H_full, bath_ops=ops_full)
    rho = np.zeros((3, 3), dtype=complex)
    rho[0, 0] = 1
    prob = HO.Problem(H=ham.H, rho0=ham.truncate(rho), gamma=gamma, lambda_=lam,
                      temperature=temperature, lh=lh, num_terms=num_terms, hbar=1.0, kb=1.0,
                      bath_model="full", bath_ops=ham.ops)
    return prob, ham


@pytest.mark.parametrize("matsubara", [False, True])
def test_truncated_full_bath_rkf45_goldens(matsubara):
    """TestTruncatedWeakIndependentFullBath(+Matsubara) (:406-416, 470-482)."""
    kw = dict(temperature=0.5, num_terms=3) if matsubara else {}
    prob, ham = truncated_problem(embed(H2, 3, 9.9), full_independent(3), 3 if matsubara else 2,
                                  [1.0] * 3, [0.01] * 3, **kw)
    r = ham.untruncate(run_rkf45(prob))
    gold = K_WEAK_MATSUBARA if matsubara else K_WEAK_INDEPENDENT
    assert np.abs(r - embed(gold, 3)).max() < 1e-3


def td_run(ham, orc, method):
    if method == "rk4":
        TD.rk4(orc, ham, 0.0, 1.0, 0.001)
        return 1e-6
    TD.rkf45(orc, ham, 0.0, 1.0, 0.001, 1e-6, 1e-6)
    return 1e-3


@pytest.mark.parametrize("method", ["rk4", "rkf45"])
@pytest.mark.parametrize("model", ["diagonal", "full"])
def test_anneal_medium_bath_models_golden(model, method):
    """TestAnnealMediumDiagonalIndependentBath / ...FullIndependentBath."""
    ops = diag_independent(2) if model == "diagonal" else full_independent(2)
    ham = TD.Hamiltonian(2, H_initial=HX, H_final=HZ, t_total=1.0)
    rho = np.zeros((2, 2), dtype=complex)
    rho[0, 0] = 1
    prob = HO.Problem(H=ham.H, rho0=rho, gamma=[0.5, 0.5], lambda_=[0.5, 0.5], temperature=2.0,
                      lh=3, num_terms=1, hbar=1.0, kb=1.0, bath_model=model, bath_ops=ops)
    orc = HO.Oracle(prob, threads=1)
    tol = td_run(ham, orc, method)
    assert np.abs(orc.state()[0] - K_MEDIUM_ANNEAL).max() < tol


@pytest.mark.parametrize("method", ["rk4", "rkf45"])
def test_truncated_anneal_full_bath_golden(method):
    """TestTruncatedAnnealMediumFullIndependentBath (:613-622)."""
    ham = TD.Hamiltonian(2, H_initial=embed(HX, 3, 9.9), H_final=embed(HZ, 3, 9.9), t_total=1.0,
                         bath_ops=full_independent(3))
    rho = np.zeros((3, 3), dtype=complex)
    rho[0, 0] = 1
    prob = HO.Problem(H=ham.H, rho0=ham.truncate(rho), gamma=[0.5] * 3, lambda_=[0.5] * 3,
                      temperature=2.0, lh=3, num_terms=1, hbar=1.0, kb=1.0, bath_model="full",
                      bath_ops=ham.ops)
    orc = HO.Oracle(prob, threads=1)
    tol = td_run(ham, orc, method)
    r = ham.untruncate(orc.state()[0])
    assert np.abs(r - embed(K_MEDIUM_ANNEAL, 3)).max() < tol
