"""Pins of the CPU oracle (CPU only).

Each oracle function is checked against something independent of its own
formulas: PHI's golden values, a numbering table from the reference source,
closed forms (vacuum propagator, pure-dephasing cumulant), the exact
polynomial form of RK4 for a linear operator, the steady state of a dense
operator, and physical invariants.  A dropped term, a flipped sign or a
transposed operand in the restatement fails at least one of them.
"""

import json
import os
from math import factorial

import numpy as np
import pytest
import scipy.linalg

from oracle import heom_oracle as HO

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H2 = np.array([[0.5, 0.5], [0.5, -0.5]])
RHO0 = np.array([[1, 0], [0, 0]], dtype=complex)
HBAR, KB = 5.3088354, 0.6950344


@pytest.fixture(scope="module", autouse=True)
def _lib(oracle_lib):
    return oracle_lib


def _golden(name):
    g = json.load(open(os.path.join(GOLD, "reference_goldens.json")))["cases"][name]
    rho = np.array([complex(a, b) for a, b in g["rho_colmajor"]]).reshape(2, 2).T
    return g, rho


@pytest.mark.parametrize("name", ["vacuum", "weak", "matsubara"])
def test_reference_goldens_rk4_rkf45(name):
    """hierarchy_integrator_test.cc:347-468: RK4 within 1e-6, RKF45 within 1e-3."""
    g, gold = _golden(name)
    prob = HO.Problem(H2, RHO0, g["gamma"], g["lambda"], g["temperature"], g["lh"],
                      num_terms=g["num_terms"], hbar=1.0, kb=1.0)
    o = HO.Oracle(prob)
    steps, _, _ = o.rk4(1.0, 0.001, 0.01)
    assert steps == 1000
    assert np.abs(o.state()[0] - gold).max() < 1e-6
    o2 = HO.Oracle(prob)
    o2.rkf45(1.0, 0.001, 1e-6, 1e-6, 0.01)
    assert np.abs(o2.state()[0] - gold).max() < 1e-3


def test_node_order_matches_reference_numbering():
    """ConstructI's id order (hierarchy.cc:45-155) for two modes and seven
    tiers against the numbering table written next to it in the reference."""
    tiers = json.load(open(os.path.join(GOLD, "hierarchy_order_d2.json")))["tiers"]
    prob = HO.Problem(np.diag([1.0, 0.0]), RHO0, [1.0, 1.0], [0.1, 0.1], 300.0,
                      lh=len(tiers), num_terms=1)
    idx = HO.Oracle(prob).index()
    assert idx.shape == (28, 2)
    for l, row in enumerate(tiers):
        for pos, one_based in enumerate(row):
            assert tuple(idx[one_based - 1]) == (l - pos, pos)


def test_node_count_and_tier_closure():
    """C(D+L, D) nodes and every node's +/- neighbours exist (c1 shape)."""
    prob = HO.Problem(np.array([[100.0, 100.0], [100.0, 0.0]]), RHO0, [20.0, 20.0],
                      [35.0, 35.0], 300.0, lh=9, num_terms=2)
    idx = HO.Oracle(prob).index()
    assert len(idx) == factorial(4 + 8) // (factorial(4) * factorial(8)) == 495
    assert len({tuple(r) for r in idx}) == 495
    assert idx.sum(axis=1).max() == 8
    # tier-major: tiers never decrease with the id
    assert np.all(np.diff(idx.sum(axis=1)) >= 0)


def test_vacuum_closed_form():
    """Lh = 1 is rho(t) = U rho0 U^dagger with U = exp(-i H t / hbar)."""
    prob = HO.Problem(H2, RHO0, [1.0], [0.01], 2.0, lh=1, hbar=1.0, kb=1.0)
    o = HO.Oracle(prob)
    _, ts, hist = o.rk4(1.0, 0.001, 0.1)
    for t, r in zip(ts, hist):
        U = scipy.linalg.expm(-1j * H2 * t)
        assert np.abs(r - U @ RHO0 @ U.conj().T).max() < 1e-12


def test_zero_coupling_is_vacuum():
    """lambda = 0 leaves c = Delta = 0: rho_0 follows the bare propagator
    even in a deep hierarchy.  (Unscaled ADOs: the scaled prefactors divide
    by sqrt|c| and are 0/0 there, in PHI as here.)"""
    H = np.array([[50.0, 20.0], [20.0, -30.0]])
    rho0 = np.array([[0.7, 0.2 - 0.1j], [0.2 + 0.1j, 0.3]])
    prob = HO.Problem(H, rho0, [20.0, 20.0], [0.0, 0.0], 300.0, lh=4, num_terms=2,
                      normalized=False)
    o = HO.Oracle(prob)
    _, ts, hist = o.rk4(0.2, 0.0002, 0.05)
    for t, r in zip(ts, hist):
        U = scipy.linalg.expm(-1j * H * t / HBAR)
        assert np.abs(r - U @ rho0 @ U.conj().T).max() < 1e-11


def _lineshape(c, nu, delta, t):
    g = np.zeros_like(t, dtype=complex)
    for ck, nk in zip(c, nu):
        g += ck / nk ** 2 * (np.exp(-nk * t) + nk * t - 1)
    return g + delta * t


@pytest.mark.parametrize("K1,lh,T,lam", [(1, 8, 300.0, 5.0), (2, 8, 100.0, 10.0)])
def test_pure_dephasing_cumulant(K1, lh, T, lam):
    """Diagonal H with different baths on the two sites: the hierarchy
    converges to rho_ab(t) = rho_ab(0) exp(-i w_ab t - g_a(t) - g_b(t)^*),
    g_j(t) = sum_k c_jk/nu_k^2 (e^{-nu_k t} + nu_k t - 1) + Delta_j t, the
    second-order cumulant of the Drude correlation function with its
    Markovian remainder.  The swapped conjugation (a wrong sign of Im c or a
    transposed coupling) misses by 5e-2 and more."""
    H = np.diag([100.0, 0.0])
    rho0 = np.full((2, 2), 0.5, dtype=complex)
    gams, lams = [20.0, 35.0], [lam, 2 * lam]
    o = HO.Oracle(HO.Problem(H, rho0, gams, lams, T, lh, num_terms=K1))
    c, _, delta = o.constants()
    _, ts, hist = o.rk4(0.3, 0.0002, 0.02)
    gs = []
    for j in range(2):
        nu = [gams[j]] + [2 * np.pi * KB * T / HBAR * k for k in range(1, K1)]
        gs.append(_lineshape(c[j * K1:(j + 1) * K1], nu, delta[j], ts))
    w = 100.0 / HBAR
    pred = 0.5 * np.exp(-1j * w * ts - gs[0] - np.conj(gs[1]))
    wrong = 0.5 * np.exp(-1j * w * ts - np.conj(gs[0]) - gs[1])
    assert np.abs(hist[:, 0, 1] - pred).max() < 1e-8
    assert np.abs(hist[:, 0, 1] - wrong).max() > 1e-2
    # populations of a pure-dephasing problem do not move
    assert np.abs(hist[:, 0, 0] - 0.5).max() < 1e-14


def _small_problem():
    H = np.array([[100.0, 60.0], [60.0, 0.0]])
    rho0 = np.array([[1, 0], [0, 0]], dtype=complex)
    return HO.Problem(H, rho0, [20.0, 30.0], [35.0, 20.0], 300.0, lh=4, num_terms=2)


def _dense_operator(o):
    n = o.num_nodes * o.N * o.N
    L = np.zeros((n, n), dtype=complex)
    for k in range(n):
        e = np.zeros(n, dtype=complex)
        e[k] = 1
        L[:, k] = o.rhs(e.reshape(o.num_nodes, o.N, o.N)).reshape(-1)
    return L


def test_rhs_is_complex_linear():
    o = HO.Oracle(_small_problem())
    rng = np.random.default_rng(3)
    shape = (o.num_nodes, 2, 2)
    x = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    y = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    a, b = 0.3 - 1.2j, -0.7 + 0.4j
    lhs = o.rhs(a * x + b * y)
    rhs = a * o.rhs(x) + b * o.rhs(y)
    assert np.abs(lhs - rhs).max() < 1e-9 * np.abs(rhs).max()


def test_rk4_step_is_taylor_polynomial():
    """For a linear time-independent operator PHI's RK4 step is
    x1 = sum_{k<=4} (dt L)^k / k! x0 (hierarchy_integrator.cc:2709-2825)."""
    o = HO.Oracle(_small_problem())
    L = _dense_operator(o)
    x0 = o.state().reshape(-1)
    dt = 0.001
    M = np.eye(len(x0), dtype=complex)
    P4 = np.zeros_like(M)
    term = np.eye(len(x0), dtype=complex)
    for k in range(5):
        P4 += term
        term = term @ (dt * L) / (k + 1)
    x1 = P4 @ x0
    o.rk4(dt * (1 - 1e-9), dt, 1.0)
    assert np.abs(o.state().reshape(-1) - x1).max() < 1e-14
    # and RK4 converges to the exact propagator at 4th order
    o2 = HO.Oracle(_small_problem())
    o2.rk4(0.05 * (1 - 1e-9), dt, 1.0)
    exact = scipy.linalg.expm(0.05 * L) @ x0
    err = np.abs(o2.state().reshape(-1) - exact).max()
    assert err < 5e-8


def test_trace_and_hermiticity_conserved():
    o = HO.Oracle(_small_problem())
    _, ts, hist = o.rk4(0.5, 0.001, 0.05)
    for r in hist:
        assert abs(np.trace(r) - 1) < 1e-12
        assert np.abs(r - r.conj().T).max() < 1e-13
    s = o.state()
    assert np.abs(s - s.conj().transpose(0, 2, 1)).max() < 1e-12


def test_steady_state_operator_and_long_time_limit():
    """The defined steady-state operator (L with the trace row) has a unique
    solution; it is the null vector of L normalised to unit trace and the
    long-time limit of the dynamics."""
    o = HO.Oracle(_small_problem())
    L = _dense_operator(o)
    w, v = np.linalg.eig(L)
    k = np.argmin(np.abs(w))
    ss = v[:, k].reshape(o.num_nodes, 2, 2)
    ss = ss / np.trace(ss[0])
    n = L.shape[0]
    A = np.zeros((n, n), dtype=complex)
    for c in range(n):
        e = np.zeros(n, dtype=complex)
        e[c] = 1
        A[:, c] = o.steady_matvec(e.reshape(o.num_nodes, 2, 2)).reshape(-1)
    b = np.zeros(n, dtype=complex)
    b[0] = 1
    x = np.linalg.solve(A, b).reshape(o.num_nodes, 2, 2)
    assert np.abs(x - ss).max() < 1e-10
    o.rk4(3.0, 0.001, 1.0)
    assert np.abs(o.state()[0] - ss[0]).max() < 1e-6


def test_rkf45_matches_fine_rk4():
    """The embedded 4(5) pair at tight tolerance agrees with a fine RK4."""
    o = HO.Oracle(_small_problem())
    res = o.rkf45(0.2, 0.001, 1e-11, 1e-8, 0.05)
    assert res["time"] == pytest.approx(0.2, abs=1e-12)
    assert res["attempted"] >= res["accepted"] > 0
    f = HO.Oracle(_small_problem())
    f.rk4(0.2 - 1e-9, 0.0002, 1.0)
    assert np.abs(o.state() - f.state()).max() < 1e-8


def test_thread_count_does_not_change_results():
    prob = _small_problem()
    a = HO.Oracle(prob, threads=1)
    b = HO.Oracle(prob, threads=3)
    a.rk4(0.05, 0.001, 1.0)
    b.rk4(0.05, 0.001, 1.0)
    assert np.array_equal(a.state(), b.state())


def _random_system(m, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    A = A / np.sqrt(m) + 3 * np.eye(m)
    b = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    return A, b


def test_krylov_oracle_bicgstab_l_with_ell_1_is_bicgstab():
    """BiCGSTAB(1) is BiCGSTAB (Sleijpen and Fokkema 1993, section 3): the
    dense restatements take the same residuals iteration by iteration."""
    from oracle import krylov_oracle as KO

    A, b = _random_system(40, 1)
    x0 = np.zeros(40, dtype=complex)
    r1, x1 = KO.bicgstab(A, b, x0, 10)
    rl, xl = KO.bicgstab_l(A, b, x0, 1, 10)
    assert np.allclose(r1, rl, rtol=1e-10, atol=1e-14 * r1[0])
    assert np.allclose(x1, xl, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("ell", [1, 2, 3, 4])
def test_krylov_oracle_converges_to_the_dense_solution(ell):
    """Every ell solves a well-conditioned complex system to numpy's
    solution, and the recursively updated residual it reports equals
    |b - A x| recomputed from its x."""
    from oracle import krylov_oracle as KO

    A, b = _random_system(60, 2 + ell)
    x0 = np.full(60, 0.1 + 0.2j)
    res, x = KO.bicgstab_l(A, b, x0, ell, 60 // ell)
    assert res[-1] < 1e-10 * res[0]
    assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-9
    assert abs(np.linalg.norm(b - A @ x) - res[-1]) < 1e-11 * res[0]
    r4, x4 = KO.bicgstab_l(A, b, x0, ell, 2)
    assert abs(np.linalg.norm(b - A @ x4) - r4[-1]) < 1e-11 * r4[0]
    rs, xs = KO.bicgstab(A, b, x0, 60)
    assert np.abs(xs - np.linalg.solve(A, b)).max() < 1e-9


@pytest.mark.parametrize("ell", [1, 2, 3, 5])
def test_krylov_oracle_mr_part_is_the_least_squares_polynomial(ell):
    """Sleijpen and Fokkema 1993, section 3: the minimal-residual part
    leaves r_0 - sum_j c_j r_j with c the least-squares coefficients of r_0
    against r_j = A^j r_0 (numpy.linalg.lstsq), applies the same polynomial
    to u_0 (u_j = A^j u_0), and moves x so that b - A x is the new r_0."""
    from oracle import krylov_oracle as KO

    m = 30
    A, b = _random_system(m, 20 + ell)
    rng = np.random.default_rng(ell)
    x = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    u0 = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    r = [b - A @ x]
    u = [u0]
    for _ in range(ell):
        r.append(A @ r[-1])
        u.append(A @ u[-1])
    R = np.stack(r[1:], axis=1)
    c = np.linalg.lstsq(R, r[0], rcond=None)[0]
    want_r = r[0] - R @ c
    want_u = u0 - np.stack(u[1:], axis=1) @ c
    rr = [v.copy() for v in r]
    uu = [v.copy() for v in u]
    x2, omega = KO.mr_part(rr, uu, x, ell)
    assert np.allclose(rr[0], want_r, atol=1e-10 * np.linalg.norm(r[0]))
    assert np.allclose(uu[0], want_u, atol=1e-9 * np.linalg.norm(u0))
    assert np.allclose(b - A @ x2, want_r, atol=1e-9 * np.linalg.norm(r[0]))
    assert omega == pytest.approx(c[-1], rel=1e-8)
