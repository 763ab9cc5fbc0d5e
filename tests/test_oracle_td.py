"""Pins of the oracle's time-dependent (annealing) and truncated-system
paths against PHI's own known answers (hierarchy_integrator_test.cc):
TestAnnealVacuum, TestAnnealMediumIndependentBath (kVacuumAnnealingSoln,
kMediumIndependentCouplingAnnealingSoln, :554-592), TestTruncatedVacuum
(kVacuumSoln embedded in 3 states, :350-373) and TestTruncatedAnnealVacuum
(:569-580), RK4 within 1e-6 as PHI's RunRk4Test and RKF45 within 1e-3 as
its RunRkf45Test; the RKF45 loop equals the C++ oracle's controller for a
static H and the RK4 loop for an H that follows every stage; plus closed
forms of the spline and the truncator."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import heom_oracle as HO  # noqa: E402
from oracle import td_oracle as TD  # noqa: E402


def colmajor(vals, n):
    return np.array(vals, dtype=complex).reshape(n, n).T


K_VACUUM = colmajor([0.788986, 0.211014 - 0.349228j, 0.211014 + 0.349228j, 0.211014], 2)
K_VACUUM_ANNEAL = colmajor([0.777864, 0.315941 - 0.270135j, 0.315941 + 0.270135j, 0.222136], 2)
K_MEDIUM_ANNEAL = colmajor([0.811767, 0.128111 - 0.144969j, 0.128111 + 0.144969j, 0.188233], 2)

HX = np.array([[0, 1], [1, 0]], dtype=complex)
HZ = np.array([[1, 0], [0, -1]], dtype=complex)
H2 = np.array([[0.5, 0.5], [0.5, -0.5]], dtype=complex)


def embed(small, full, corner=None):
    out = np.zeros((full, full), dtype=complex)
    n = small.shape[0]
    out[:n, :n] = small
    if corner is not None:
        out[full - 1, full - 1] = corner
    return out


def run(n, lh, gamma, lam, full=None, H=None, Hi=None, Hf=None):
    full = full or n
    ham = TD.Hamiltonian(n, H=H, H_initial=Hi, H_final=Hf, t_total=1.0)
    rho_full = np.zeros((full, full), dtype=complex)
    rho_full[0, 0] = 1
    prob = HO.Problem(H=ham.H, rho0=ham.truncate(rho_full), gamma=gamma, lambda_=lam,
                      temperature=2.0, lh=lh, num_terms=1, hbar=1.0, kb=1.0)
    orc = HO.Oracle(prob, threads=1)
    steps = TD.rk4(orc, ham, 0.0, 1.0, 0.001)
    assert steps == 1000
    return ham.untruncate(orc.state()[0]), orc, ham


def test_anneal_vacuum_golden():
    r, _, _ = run(2, 1, [0.5], [0.5], Hi=HX, Hf=HZ)
    assert np.abs(r - K_VACUUM_ANNEAL).max() < 1e-6


def test_anneal_medium_independent_bath_golden():
    r, _, _ = run(2, 3, [0.5, 0.5], [0.5, 0.5], Hi=HX, Hf=HZ)
    assert np.abs(r - K_MEDIUM_ANNEAL).max() < 1e-6


def test_truncated_vacuum_golden():
    r, _, ham = run(2, 1, [1.0], [0.01], full=3, H=embed(H2, 3, 9.9))
    assert ham.truncated
    assert np.abs(r - embed(K_VACUUM, 3)).max() < 1e-6


def test_truncated_anneal_vacuum_golden():
    r, _, ham = run(2, 1, [0.5], [0.5], full=3, Hi=embed(HX, 3, 9.9), Hf=embed(HZ, 3, 9.9))
    assert ham.truncated
    assert np.abs(r - embed(K_VACUUM_ANNEAL, 3)).max() < 1e-6


def run_rkf45(n, lh, gamma, lam, full=None, H=None, Hi=None, Hf=None, dt0=0.001, tol=1e-6):
    full = full or n
    ham = TD.Hamiltonian(n, H=H, H_initial=Hi, H_final=Hf, t_total=1.0)
    rho_full = np.zeros((full, full), dtype=complex)
    rho_full[0, 0] = 1
    prob = HO.Problem(H=ham.H, rho0=ham.truncate(rho_full), gamma=gamma, lambda_=lam,
                      temperature=2.0, lh=lh, num_terms=1, hbar=1.0, kb=1.0)
    orc = HO.Oracle(prob, threads=1)
    attempts, accepted, t = TD.rkf45(orc, ham, 0.0, 1.0, dt0, tol, 1e-6)
    assert abs(t - 1.0) < 1e-12
    return ham.untruncate(orc.state()[0]), (attempts, accepted), prob


def test_rkf45_static_matches_controller():
    # with a static H the loop is the C++ oracle's controller: same attempts,
    # same accepted steps, same state
    r, counts, prob = run_rkf45(2, 3, [0.5, 0.5], [0.5, 0.5], H=H2, dt0=0.2, tol=1e-7)
    orc = HO.Oracle(prob, threads=1)
    res = orc.rkf45(1.0, 0.2, 1e-7, 1e-6, 0.01)
    assert counts == (res["attempted"], res["accepted"])
    assert counts[0] > counts[1]          # some attempts were rejected
    assert np.abs(r - orc.state()[0]).max() < 1e-13


def test_rkf45_anneal_goldens():
    # PHI's RunRkf45Test of TestAnnealVacuum, TestAnnealMediumIndependentBath
    # and TestTruncatedAnnealVacuum (within 1e-3)
    r, _, _ = run_rkf45(2, 1, [0.5], [0.5], Hi=HX, Hf=HZ)
    assert np.abs(r - K_VACUUM_ANNEAL).max() < 1e-3
    r, _, _ = run_rkf45(2, 3, [0.5, 0.5], [0.5, 0.5], Hi=HX, Hf=HZ)
    assert np.abs(r - K_MEDIUM_ANNEAL).max() < 1e-3
    r, _, _ = run_rkf45(2, 1, [0.5], [0.5], full=3, Hi=embed(HX, 3, 9.9), Hf=embed(HZ, 3, 9.9))
    assert np.abs(r - embed(K_VACUUM_ANNEAL, 3)).max() < 1e-3


@pytest.mark.parametrize("full", [2, 3])
def test_rkf45_anneal_agrees_with_rk4(full):
    # With MinimumFieldChange 0 the Hamiltonian follows every stage time, and
    # the adaptive run at tolerance 1e-10 and the RK4 run at dt 1e-3 must then
    # reach one state to ~1e-11: a stage evaluated at a wrong time or a live
    # buffer left in the old eigenbasis would be off by O(dt).
    def go(method):
        Hi = embed(HX, 3, 9.9) if full == 3 else HX
        Hf = embed(HZ, 3, 9.9) if full == 3 else HZ
        ham = TD.Hamiltonian(2, H_initial=Hi, H_final=Hf, t_total=1.0, min_change=0.0)
        rho = np.zeros((full, full), dtype=complex)
        rho[0, 0] = 1
        prob = HO.Problem(H=ham.H, rho0=ham.truncate(rho), gamma=[0.5, 0.5],
                          lambda_=[0.5, 0.5], temperature=2.0, lh=3, num_terms=1, hbar=1.0,
                          kb=1.0)
        orc = HO.Oracle(prob, threads=1)
        if method == "rk4":
            TD.rk4(orc, ham, 0.0, 1.0, 0.001)
        else:
            TD.rkf45(orc, ham, 0.0, 1.0, 0.01, 1e-10, 1e-6)
        return ham.untruncate(orc.state()[0])

    assert np.abs(go("rkf45") - go("rk4")).max() < 1e-10


def test_spline_closed_forms():
    # two knots: the straight line between them
    s = TD.CubicSpline([1.0, 0.0])
    for x in np.linspace(0, 1, 11):
        assert abs(s(x) - (1 - x)) < 1e-15
    assert s(-0.5) == 1.0 and s(1.5) == 0.0
    # every knot is interpolated, the curve is C1 at interior knots
    y = [0.0, 0.3, 1.1, 0.7, 0.2]
    s = TD.CubicSpline(y)
    m = len(y)
    for i, v in enumerate(y):
        assert abs(s(i / (m - 1)) - v) < 1e-14
    h = 1e-7
    for i in range(1, m - 1):
        x = i / (m - 1)
        left = (s(x) - s(x - h)) / h
        right = (s(x + h) - s(x)) / h
        assert abs(left - right) < 1e-5


def test_truncator_closed_forms():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5))
    H = A + A.conj().T
    ham = TD.Hamiltonian(3, H=H)
    w, V = np.linalg.eigh(H)
    # the truncated H is diagonal with the 3 lowest eigenvalues
    assert np.abs(ham.H - np.diag(w[:3])).max() < 1e-12
    # projector identity: untruncate(truncate(P)) = P for P inside the space
    P = ham.V @ np.diag([0.5, 0.3, 0.2]) @ ham.V.conj().T
    assert np.abs(ham.untruncate(ham.truncate(P)) - P).max() < 1e-12
    # a rotation between two bases maps truncated matrices consistently
    ham2 = TD.Hamiltonian(3, H_initial=H, H_final=np.diag(np.arange(5.0)).astype(complex),
                          t_total=1.0, min_change=0.0)
    a_old = ham2.truncate(P)
    assert ham2.update(0.5)
    a_new = ham2.R @ a_old @ ham2.R.conj().T
    Pp = ham2.untruncate(a_new)
    # P projected onto the new space
    Q = ham2.V @ ham2.V.conj().T
    assert np.abs(Pp - Q @ P @ Q).max() < 1e-12
