"""Pins of the oracle's adaptive hierarchy truncation (PHI's Shi filter).

PHI's HierarchyUpdater::Get is not virtual and the integrator calls it
through a HierarchyUpdater* (hierarchy_updater.h:36-52,
hierarchy_integrator.cc:1798, 2011), so with FilterTolerance > 0 the base
Get runs with IndependentBathAdaptiveTruncation's DoNext / DoPrevious
(hierarchy_updater.cc:256-304): every node -- active or not -- gets its own
terms (unitary, dephasing, Matsubara; DoUnitary overwrites drho), and the
terms of inactive neighbours are skipped.  The class's own Get
(hierarchy_updater.h:227-239, which would skip an inactive node's own
terms) is never called.  A node whose max |rho_n(i,j)| fell below the
tolerance after a step is inactive for the next step; flags change after
each step (hierarchy_integrator.cc:2828-2851): off when max < tol, on when
max > tol, unchanged on a tie.

What fixes these without retyping the oracle:
  * tolerance below every |rho| -> the plain (pinned) integrator, bit for bit;
  * RHS with flags == plain RHS of the masked input (inactive ADOs zeroed)
    on active nodes, and that plus the node's own block of the plain
    operator applied to its actual ADO on inactive nodes;
  * the flags after a run are the rule applied to the final state;
  * a tolerance above every |rho|: every node is inactive after step 1 and
    evolves alone, rho_n(m) = P4(dt L_nn)^(m-1) rho_n(1), with L_nn the
    node's diagonal block of the plain operator and P4 the RK4 polynomial.
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import heom_oracle as HO  # noqa: E402
from oracle import prm_oracle as PO  # noqa: E402

C1 = os.path.join(ROOT, "phi_amd", "configs", "c1.prm")


def _oracle(tol, lh=None):
    op = PO.read_command_line(C1, "rk4")
    prob = HO.Problem.from_prm(op)
    prob.filter_tolerance = tol
    if lh is not None:
        prob.lh = lh
    return op, HO.Oracle(prob, threads=4)


def _maxabs(state):
    return np.abs(state.reshape(state.shape[0], -1)).max(axis=1)


def _hermitian_state(o, seed, scale=1e-3):
    rng = np.random.default_rng(seed)
    N = o.N
    x = rng.standard_normal((o.num_nodes, N, N)) + 1j * rng.standard_normal((o.num_nodes, N, N))
    return scale * (x + np.conj(np.transpose(x, (0, 2, 1))))


def own_block(plain, n):
    """The plain operator's diagonal block at node n, as an N^2 x N^2
    matrix on column-major vec(rho_n) (from unit inputs)."""
    N = plain.N
    B = np.zeros((N * N, N * N), dtype=complex)
    for e in range(N * N):
        x = np.zeros((plain.num_nodes, N, N), dtype=complex)
        x[n, e % N, e // N] = 1
        y = plain.rhs(x)[n]
        B[:, e] = y.T.reshape(-1)
    return B


def test_tiny_tolerance_is_the_plain_integrator():
    """From a state with no zero ADO every node stays active, and the
    filtered run is the plain one bit for bit.  (From PHI's initial state
    the ADOs beyond the tiers reached so far are exactly zero, hence
    inactive for a step, so even a tiny tolerance changes that run.)"""
    op, o = _oracle(1e-300)
    _, plain = _oracle(0.0)
    x = _hermitian_state(o, 5)
    o.set_state(x)
    plain.set_state(x)
    dt = op.timestep
    o.rk4(40 * dt * (1 - 1e-9), dt, 1.0)
    plain.rk4(40 * dt * (1 - 1e-9), dt, 1.0)
    assert o.active().all()
    assert np.array_equal(o.state(), plain.state())
    # from PHI's initial state the same tolerance does deactivate nodes
    op, z = _oracle(1e-300)
    z.rk4(1 * dt * (1 - 1e-9), dt, 1.0)
    assert not z.active().all()


@pytest.mark.parametrize("tol", [1e-6, 1e-4, 1e-2])
def test_flags_follow_the_rule_and_rhs_structure(tol):
    op, o = _oracle(tol)
    dt = op.timestep
    o.rk4(50 * dt * (1 - 1e-9), dt, 1.0)
    act = o.active().astype(bool)
    mx = _maxabs(o.state())
    assert 0 < act.sum() < o.num_nodes
    tie = mx == tol
    assert np.array_equal(act[~tie], (mx > tol)[~tie])
    # RHS structure on an arbitrary Hermitian input
    rng = np.random.default_rng(3)
    N = o.N
    x = rng.standard_normal((o.num_nodes, N, N)) + 1j * rng.standard_normal((o.num_nodes, N, N))
    x = x + np.conj(np.transpose(x, (0, 2, 1)))
    got = o.rhs(x)
    _, plain = _oracle(0.0)
    want = plain.rhs(x * act[:, None, None])
    scale = np.abs(want).max()
    # active nodes: the plain operator on the masked input, exactly
    assert np.array_equal(got[act], want[act])
    # inactive nodes: plus their own terms on their actual ADO
    for n in np.flatnonzero(~act)[:40]:
        own = (own_block(plain, n) @ x[n].T.reshape(-1)).reshape(N, N).T
        assert np.abs(got[n] - want[n] - own).max() <= 1e-14 * scale
    # the skipped override (own terms dropped on inactive nodes) differs
    off = np.flatnonzero(~act)[0]
    assert np.abs(got[off] - want[off]).max() > 1e-6 * scale


def test_all_inactive_nodes_evolve_alone():
    op, o = _oracle(10.0)
    dt = op.timestep
    o.rk4(1 * dt * (1 - 1e-9), dt, 1.0)
    rho1 = o.state()
    assert not o.active().any()
    # step 1 itself is the plain step (every node active during it)
    _, plain = _oracle(0.0)
    plain.rk4(1 * dt * (1 - 1e-9), dt, 1.0)
    assert np.array_equal(rho1, plain.state())
    m = 30
    o.rk4((m - 1) * dt * (1 - 1e-9), dt, 1.0, t_start=0.0)
    assert not o.active().any()
    got = o.state()
    N = o.N
    checked = 0
    for n in range(o.num_nodes):
        if not np.abs(rho1[n]).max() > 0:
            assert np.abs(got[n]).max() == 0
            continue
        A = dt * own_block(plain, n)
        P4 = np.eye(N * N) + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24
        v = rho1[n].T.reshape(-1)
        for _ in range(m - 1):
            v = P4 @ v
        want = v.reshape(N, N).T
        assert np.abs(got[n] - want).max() <= 1e-12 * max(1.0, np.abs(want).max())
        checked += 1
    assert checked > 10
    # node 0 keeps evolving (its own unitary and Matsubara terms)
    assert np.abs(got[0] - rho1[0]).max() > 1e-3
