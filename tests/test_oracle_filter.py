"""Pins of the oracle's adaptive hierarchy truncation (PHI's Shi filter).

PHI (hierarchy_updater.h:217-245, hierarchy_updater.cc:256-304): a node whose
max |rho_n(i,j)| fell below FilterTolerance after a step is inactive for the
next step; an inactive node skips its own terms (unitary, dephasing,
Matsubara), inactive neighbours are skipped, and -- because the skipped
DoUnitary is the call that overwrites drho (beta = 0) -- an inactive node's
per-stage k buffer keeps its old contents and the neighbour terms are added
onto them.  Flags change after each step (hierarchy_integrator.cc:2828-2851):
off when max < tol, on when max > tol, unchanged on a tie.

What fixes these without retyping the oracle:
  * tolerance below every |rho| -> the plain (pinned) integrator, bit for bit;
  * RHS with flags == plain RHS of the masked input (inactive ADOs zeroed),
    exactly: own terms of an inactive node and terms of inactive neighbours
    are the plain operator's terms on zeros;
  * the flags after a run are the rule applied to the final state;
  * a tolerance above every |rho|: every node is inactive after step 1, no
    neighbour term survives, every k keeps step 1's value, so the state moves
    on a straight line: rho_m = rho_0 + m (rho_1 - rho_0).  Clearing k for
    inactive nodes would freeze it instead.
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
def test_flags_follow_the_rule_and_rhs_is_plain_rhs_of_masked_input(tol):
    op, o = _oracle(tol)
    dt = op.timestep
    o.rk4(50 * dt * (1 - 1e-9), dt, 1.0)
    act = o.active().astype(bool)
    mx = _maxabs(o.state())
    assert 0 < act.sum() < o.num_nodes
    tie = mx == tol
    assert np.array_equal(act[~tie], (mx > tol)[~tie])
    # RHS identity on an arbitrary Hermitian input
    rng = np.random.default_rng(3)
    N = o.N
    x = rng.standard_normal((o.num_nodes, N, N)) + 1j * rng.standard_normal((o.num_nodes, N, N))
    x = x + np.conj(np.transpose(x, (0, 2, 1)))
    got = o.rhs(x, out=np.zeros_like(x))
    _, plain = _oracle(0.0)
    want = plain.rhs(x * act[:, None, None])
    assert np.array_equal(got, want)


def test_all_inactive_state_moves_on_a_straight_line():
    op, o = _oracle(10.0)
    dt = op.timestep
    rho0 = o.state()
    o.rk4(1 * dt * (1 - 1e-9), dt, 1.0)
    rho1 = o.state()
    assert not o.active().any()
    # step 1 itself is the plain step (every node active during it)
    _, plain = _oracle(0.0)
    plain.rk4(1 * dt * (1 - 1e-9), dt, 1.0)
    assert np.array_equal(rho1, plain.state())
    m = 30
    o.rk4((m - 1) * dt * (1 - 1e-9), dt, 1.0, t_start=0.0)
    want = rho0 + m * (rho1 - rho0)
    assert not o.active().any()
    assert np.abs(o.state() - want).max() <= 1e-12 * max(1.0, np.abs(want).max())
    assert np.abs(o.state() - rho1).max() > 1e-6
