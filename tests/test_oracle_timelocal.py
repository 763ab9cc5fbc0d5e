"""Pins of the oracle's time-local truncation (PHI's TimeLocal option)
against PHI's own known answers, hierarchy_integrator_test.cc:505-552 and
625-716: TestMedium{,Diagonal,Full}BathTimeLocal (RK4 within 1e-6 and RKF45
within 1e-3 of the deeper hierarchy's kMediumIndependentCouplingSoln),
TestTruncatedMediumIndependentFullBathTimeLocal and the five annealing
cases (RKF45 within 1e-3, integration time 10).  Plus what the algebra
fixes: with the bath switched off the matrices stay zero, the first step
equals the closed form, a time-local run differs from the plain truncation
at the same depth, and Q_m is what PHI's formula gives for a static H."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import heom_oracle as HO  # noqa: E402
from oracle import td_oracle as TD  # noqa: E402

from test_oracle_bath import (diag_alternating, diag_independent, full_alternating,  # noqa: E402
                              full_independent)
from test_oracle_td import H2, HX, HZ, colmajor, embed  # noqa: E402

K_MEDIUM_INDEPENDENT = colmajor([0.83145, 0.0804222 - 0.230756j, 0.0804222 + 0.230756j,
                                 0.16855], 2)
K_WEAK_TL_ANNEAL = colmajor([0.467201, 0.255327 + 0.207856j, 0.255327 - 0.207856j,
                             0.532799], 2)
K_WEAK_DIAGONAL_TL_ANNEAL = colmajor([0.423523, 0.0658185 + 0.0703083j,
                                      0.0658185 - 0.0703083j, 0.576477], 2)
K_WEAK_FULL_TL_ANNEAL = colmajor([0.428181, 0.0237366 + 0.0886596j, 0.0237366 - 0.0886596j,
                                  0.571819], 2)
K_TRUNCATED_WEAK_FULL_TL_ANNEAL = colmajor(
    [0.267895, -0.0281567 - 0.0246088j, -0.00247109 - 0.0184745j,
     -0.0281567 + 0.0246088j, 0.208783, 0.154777 + 0.00171474j,
     -0.00247109 + 0.0184745j, 0.154777 - 0.00171474j, 0.116023], 3)


def truncated_anneal_pair(full):
    """SetTruncatedAnnealingHamiltonian(full, 2) (:99-119)."""
    Hi = np.zeros((full, full), dtype=complex)
    Hf = np.zeros((full, full), dtype=complex)
    for i in range(full):
        Hf[i, i] = 1 + i
        Hi[i, i] = -1 - i
        for j in range(i):
            Hi[i, j] = Hi[j, i] = 1 + i
            Hf[i, j] = Hf[j, i] = -1 - j
    return Hi, Hf


MEDIUM = dict(gamma=0.5, lam=0.5)
WEAK = dict(gamma=1.0, lam=0.01)

# name: (n, full, M, bath, H | (Hi, Hf), model, ops, lh, t_total, methods, gold)
CASES = {
    "medium_independent": (2, 2, 2, MEDIUM, H2, "independent", None, 4, 1.0,
                           ("rk4", "rkf45"), K_MEDIUM_INDEPENDENT),
    "medium_diagonal": (2, 2, 2, MEDIUM, H2, "diagonal", diag_independent(2), 4, 1.0,
                        ("rk4", "rkf45"), K_MEDIUM_INDEPENDENT),
    "medium_full": (2, 2, 2, MEDIUM, H2, "full", full_independent(2), 4, 1.0,
                    ("rk4", "rkf45"), K_MEDIUM_INDEPENDENT),
    "truncated_medium_full": (2, 3, 3, MEDIUM, embed(H2, 3, 9.9), "full", full_independent(3), 4,
                              1.0, ("rkf45",), embed(K_MEDIUM_INDEPENDENT, 3)),
    "anneal_weak_independent": (2, 2, 2, WEAK, (HX, HZ), "independent", None, 3, 10.0,
                                ("rkf45",), K_WEAK_TL_ANNEAL),
    "truncated_anneal_weak_full_independent": (2, 3, 3, WEAK, (embed(HX, 3, 9.9),
                                                             embed(HZ, 3, 9.9)),
                                               "full", full_independent(3), 3, 10.0, ("rkf45",),
                                               embed(K_WEAK_TL_ANNEAL, 3)),
    "anneal_weak_diagonal": (2, 2, 2, WEAK, (HX, HZ), "diagonal", diag_alternating(2, 2), 3, 10.0,
                             ("rkf45",), K_WEAK_DIAGONAL_TL_ANNEAL),
    "anneal_weak_full": (2, 2, 2, WEAK, (HX, HZ), "full", full_alternating(2, 2), 3, 10.0,
                         ("rkf45",), K_WEAK_FULL_TL_ANNEAL),
    "truncated_anneal_weak_full": (2, 3, 2, WEAK, truncated_anneal_pair(3), "full",
                                   full_alternating(2, 3), 3, 10.0, ("rkf45",),
                                   K_TRUNCATED_WEAK_FULL_TL_ANNEAL),
}


def setup(name, timelocal=True):
    """(oracle, Hamiltonian, TimeLocal or None, t_total) for a case."""
    n, full, M, bath, Hspec, model, ops, lh, T, _, _ = CASES[name]
    if isinstance(Hspec, tuple):
        ham = TD.Hamiltonian(n, H_initial=Hspec[0], H_final=Hspec[1], t_total=T,
                             bath_ops=ops if (model == "full" and full > n) else None)
    else:
        ham = TD.Hamiltonian(n, H=Hspec, bath_ops=ops if (model == "full" and full > n) else None)
    rho = np.zeros((full, full), dtype=complex)
    rho[0, 0] = 1
    prob = HO.Problem(H=ham.H, rho0=ham.truncate(rho), gamma=[bath["gamma"]] * M,
                      lambda_=[bath["lam"]] * M, temperature=2.0, lh=lh, num_terms=1, hbar=1.0,
                      kb=1.0, bath_model=model,
                      bath_ops=ham.ops if ham.ops is not None else ops)
    orc = HO.Oracle(prob, threads=1)
    tl = TD.TimeLocal(orc) if timelocal else None
    return orc, ham, tl, T


def run(name, method, timelocal=True):
    orc, ham, tl, T = setup(name, timelocal)
    if method == "rk4":
        assert TD.rk4(orc, ham, 0.0, T, 0.001, tl=tl) == round(T / 0.001)
        counts = None
    else:
        counts = TD.rkf45(orc, ham, 0.0, T, 0.001, 1e-6, 1e-6, tl=tl)
    return ham.untruncate(orc.state()[0]), counts


@pytest.mark.parametrize("name,method", [(k, m) for k, v in sorted(CASES.items()) for m in v[9]])
def test_timelocal_goldens(name, method):
    r, _ = run(name, method)
    gold = CASES[name][10]
    assert np.abs(r - gold).max() < (1e-6 if method == "rk4" else 1e-3)


def test_timelocal_changes_the_answer():
    # the plain truncation at the same depth misses the golden by far more:
    # the time-local term is what brings depth 4 onto the converged answer
    r, _ = run("medium_independent", "rk4", timelocal=False)
    assert np.abs(r - K_MEDIUM_INDEPENDENT).max() > 1e-3


def test_timelocal_matrices_closed_forms():
    orc, ham, tl, _ = setup("medium_full")
    c = tl.c
    # first step from zero: T_mk = e^{-nu ts} P^dag (ts c_mk F_m) P
    ts = 0.01
    tl.step(ham.H, ts)
    P = np.eye(2) - 1j * ham.H * ts
    for m in range(2):
        F = full_independent(2)[m]
        want = np.exp(-tl.rate[m, 0] * ts) * (P.conj().T @ (ts * c[m, 0] * F) @ P)
        assert np.abs(tl.T[m, 0] - want).max() < 1e-16
    # a vanishing H and a single term: T(t) = c/gamma (1 - e^{-gamma t}) F exactly
    # in the limit of small steps; the recursion is its left Riemann sum
    tl2 = TD.TimeLocal(orc)
    Z = np.zeros((2, 2), dtype=complex)
    steps, ts = 2000, 0.0005
    for _ in range(steps):
        tl2.step(Z, ts)
    g = tl2.rate[0, 0]
    t = steps * ts
    # sum_{j=1..n} ts c e^{-g j ts} = ts c e^{-g ts} (1 - e^{-g t}) / (1 - e^{-g ts})
    geo = ts * c[0, 0] * np.exp(-g * ts) * (1 - np.exp(-g * t)) / (1 - np.exp(-g * ts))
    q = tl2.Q()[0][0, 0]
    assert abs(q - geo) < 1e-12      # 2000 rounded steps
    assert abs(geo - c[0, 0] / g * (1 - np.exp(-g * t))) < 1e-3 * abs(geo)
    # store / restore
    tl2.store()
    tl2.step(ham.H, 0.1)
    tl2.restore()
    assert tl2.Q()[0][0, 0] == q
    # steps below 1e-9 are skipped (hierarchy_integrator.cc:1678)
    before = tl2.T.copy()
    tl2.step(ham.H, 5e-10)
    assert np.array_equal(before, tl2.T)


def test_timelocal_term_structure():
    # the oracle's term for one node equals -sum_m [V_m, Q_m rho - rho Q_m^dag]
    # on the deepest tier and is absent elsewhere
    rng = np.random.default_rng(11)
    for name in ("medium_independent", "medium_diagonal", "medium_full"):
        orc, ham, tl, _ = setup(name)
        x = rng.standard_normal((orc.num_nodes, 2, 2)) + 1j * rng.standard_normal(
            (orc.num_nodes, 2, 2))
        base = orc.rhs(x)
        Q = rng.standard_normal((2, 2, 2)) + 1j * rng.standard_normal((2, 2, 2))
        orc.set_timelocal(Q)
        with_tl = orc.rhs(x)
        orc.set_timelocal(None)
        tiers = orc.index().sum(axis=1)
        model = CASES[name][5]
        ops = CASES[name][6]
        for i in range(orc.num_nodes):
            want = np.zeros((2, 2), dtype=complex)
            if tiers[i] == 3:
                for m in range(2):
                    X = Q[m] @ x[i] - x[i] @ Q[m].conj().T
                    if model == "independent":
                        V = np.zeros((2, 2))
                        V[m, m] = 1
                    elif model == "diagonal":
                        V = np.diag(ops[m])
                    else:
                        V = ops[m]
                    want -= V @ X - X @ V
            assert np.abs(with_tl[i] - base[i] - want).max() < 1e-13
