"""The benchmark configurations' legs at their stated integrator and size,
through the C ABI, against the oracle or against member-by-member runs.

  24-site ring, RKF45 abs tol 1e-8: at depth 2 (325 nodes of 24x24) the
      attempted and accepted step counts, the final time and the whole
      hierarchy equal the oracle's controller (1e-10); at full depth 4
      (20475 nodes) FP32-bulk vs FP64 to 1 ps, population error <= 1e-6.
  7-site depth 8 (6435 nodes), steady state: BiCGSTAB to 1e-10 from the
      1 ps state; the residual recomputed with the oracle's operator.
  static-disorder ensemble: members of depth 4..8 (PHI HierarchyTruncation
      5..9), 130 RK4 steps (hipGraph step blocks replay), autotuned shape,
      equal to each member integrated alone with eager launches.
"""

import os

import numpy as np
import pytest

import phi_amd as P
from oracle import heom_oracle as HO
from oracle import prm_oracle as PO

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = os.path.join(ROOT, "phi_amd", "configs")
C2 = os.path.join(CFG, "c2.prm")
C3 = os.path.join(CFG, "c3.prm")


def test_c3_rkf45_depth2_matches_oracle_controller():
    """24-site ring at HierarchyTruncation=3 (325 nodes), RKF45 tol 1e-8
    from dt 1 fs to 0.25 ps: PHI's controller on the GPU = the oracle's."""
    print("case: c3 rkf45 depth 2 vs oracle", flush=True)
    op = PO.read_command_line(C3, "rkf45")
    op.hierarchy_truncation_level = 3
    t_end = 0.25
    orc = HO.Oracle(HO.Problem.from_prm(op), threads=16)
    res = orc.rkf45(t_end, op.timestep, op.rkf45_tolerance, op.rkf45_minimum_timestep,
                    op.output_minimum_timestep)
    p = P.PhiParameters.read(C3, "rkf45")
    p.set_int("hierarchy_truncation_level", 3)
    p.set_double("integration_time", t_end)
    p.set_string("output_filename", "")
    p.set_int("verbose", -2)
    it = P.HierarchyIntegrator(p)
    it.Launch()
    info = it.info()
    print("attempts %d accepted %d (oracle %d %d)"
          % (info["steps"], info["accepted"], res["attempted"], res["accepted"]))
    assert info["num_nodes"] == orc.num_nodes == 325
    assert info["steps"] == res["attempted"]
    assert info["accepted"] == res["accepted"]
    assert info["final_time"] == res["time"]
    assert np.abs(it.state() - orc.state()).max() <= 1e-10


def test_c3_rkf45_full_fp32_bulk_vs_fp64():
    """24-site ring at depth 4 (20475 nodes), RKF45 tol 1e-8 to 1 ps:
    FP32-bulk populations within 1e-6 of FP64; coherence error reported."""
    out = {}
    for prec in ("fp64", "fp32"):
        print("case: c3 rkf45 full", prec, flush=True)
        p = P.PhiParameters.read(C3, "rkf45")
        p.set_string("output_filename", "")
        p.set_int("verbose", -2)
        p.set_int("bulk_precision", 1 if prec == "fp32" else 0)
        it = P.HierarchyIntegrator(p)
        it.Launch()
        out[prec] = (it.density_matrix(), it.info())
    r64, i64 = out["fp64"]
    r32, i32 = out["fp32"]
    d = r32 - r64
    pop = np.abs(np.real(np.diag(d))).max()
    coh = np.abs(d - np.diag(np.diag(d))).max()
    print("fp64 %d/%d fp32 %d/%d attempts/accepted; population error %.3e coherence error %.3e"
          % (i64["steps"], i64["accepted"], i32["steps"], i32["accepted"], pop, coh))
    assert i64["num_nodes"] == 20475
    assert i64["final_time"] == pytest.approx(1.0, abs=1e-12)
    assert abs(np.trace(r64) - 1) < 1e-10
    assert pop <= 1e-6
    assert coh <= 1e-5


def test_c2_steady_state_full_hierarchy():
    """7-site depth 8 (6435 nodes): BiCGSTAB from the 1 ps RK4 state to
    residual 1e-10; |b - A x| recomputed with the oracle's steady-state
    operator agrees, tr rho_0 = 1, and rho_0 is Hermitian."""
    print("case: c2 steady state 6435 nodes", flush=True)
    op = PO.read_command_line(C2, "bicgstab")
    p = P.PhiParameters.read(C2, "rk4")
    ctx = P.Context(p)
    assert ctx.num_nodes == 6435
    ctx.rk4_steps(1000, p.get_double("timestep"))
    rc, iters, res = ctx.steady_state(1e-10, 20000)
    x = ctx.state()
    ctx.close()
    print("iterations %d residual %.3e" % (iters, res))
    assert rc == 0 and res <= 1e-10
    orc = HO.Oracle(HO.Problem.from_prm(op), threads=16)
    b = np.zeros_like(x)
    b[0, 0, 0] = 1
    r = b - orc.steady_matvec(x)
    true_res = np.linalg.norm(r.reshape(-1))
    print("recomputed residual %.3e" % true_res)
    assert true_res <= 2e-10
    r0 = x[0]
    assert abs(np.trace(r0) - 1) < 1e-10
    assert np.abs(r0 - r0.conj().T).max() < 1e-9


def _members(depths):
    from phi_amd.configs.make_configs import c4_member
    Hs = [c4_member(r)[0] for r in range(len(depths))]
    return np.array(Hs)


def _member_params(H, lh):
    q = P.PhiParameters.read(C2, "rk4")
    q.set_matrix("time_indepenent_hamiltonian", H)
    q.set_int("hierarchy_truncation_level", lh)
    return q


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_ensemble_graph_blocks_deep_members(prec, tol):
    """Members of every depth 4..8, 130 steps (the autotuned shape, hipGraph
    step blocks replayed 3 times) = each member alone, eager launches."""
    depths = [5, 6, 7, 8, 9, 9, 7]
    Hs = _members(depths)
    p = P.PhiParameters.read(C2, "rk4")
    dt = p.get_double("timestep")
    print("case: ensemble graph blocks", prec, depths, flush=True)
    ens = P.Ensemble(p, Hs, depths, precision=prec)
    ens.set_graphs(True)
    ms, shape = ens.autotune(dt, steps=3)
    assert ens.graphs()
    print("autotuned shape", shape, "%.4f ms/step" % ms)
    ens.rk4_steps(130, dt)
    slab = ens.state()
    off = 0
    for m, lh in enumerate(depths):
        ctx = P.Context(_member_params(Hs[m], lh), precision=prec)
        ctx.set_graphs(False)
        assert ctx.num_nodes == ens.counts[m]
        ctx.rk4_steps(130, dt)
        d = np.abs(slab[off:off + ctx.num_nodes] - ctx.state()).max()
        print("member %d depth %d nodes %d max|d| %.3e" % (m, lh - 1, ctx.num_nodes, d))
        assert d <= tol
        off += ctx.num_nodes
        ctx.close()
    assert off == ens.num_nodes
    ens.close()
