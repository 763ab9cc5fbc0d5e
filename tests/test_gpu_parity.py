"""GPU parity of the HIP hot path against the CPU oracle and PHI's goldens.

Every call goes through the C ABI (libphi_hip.so).  Bars:
  * FP64 path vs oracle: max |delta rho| <= 1e-10 over the run (north star);
  * PHI's own integration goldens (hierarchy_integrator_test.cc:316-345):
    RK4 within 1e-6, RKF45 within 1e-3, plus Hermiticity / trace checks
    (hierarchy_integrator_test.cc:253-282);
  * FP32-bulk RK4: populations within 1e-6 of FP64.
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import phi_amd as P  # noqa: E402
from oracle import heom_oracle as HO  # noqa: E402
from oracle import prm_oracle as PO  # noqa: E402

pytestmark = pytest.mark.gpu

CFG = os.path.join(ROOT, "phi_amd", "configs")

H2 = np.array([[0.5, 0.5], [0.5, -0.5]])
RHO0_2 = np.array([[1, 0], [0, 0]], dtype=complex)


def colmajor_golden(vals):
    return np.array(vals, dtype=complex).reshape(2, 2).T


# hierarchy_integrator_test.cc:347-468 (IndependentBath goldens).
KATS = {
    "vacuum": (dict(gamma=[1.0], lambda_=[0.01], temperature=2.0, lh=1),
               colmajor_golden([0.788986, 0.211014 - 0.349228j, 0.211014 + 0.349228j, 0.211014])),
    "weak": (dict(gamma=[1.0, 1.0], lambda_=[0.01, 0.01], temperature=2.0, lh=2),
             colmajor_golden([0.790012, 0.206755 - 0.346119j, 0.206755 + 0.346119j, 0.209988])),
    "matsubara": (dict(gamma=[1.0, 1.0], lambda_=[0.01, 0.01], temperature=0.5, lh=3, num_terms=3),
                  colmajor_golden([0.789331, 0.208929 - 0.348365j, 0.208929 + 0.348365j,
                                   0.210669])),
}


def assert_valid_density_matrix(r):
    """hierarchy_integrator_test.cc:253-282."""
    n = r.shape[0]
    assert np.all(np.abs(np.imag(np.diag(r))) < 1e-9)
    assert np.all(np.real(np.diag(r)) >= 0)
    assert np.all(np.real(np.diag(r)) < 1.00001)
    assert np.allclose(r, r.conj().T, atol=1e-9)
    tr = np.real(np.trace(r))
    assert 0 <= tr < 1.00001


@pytest.mark.parametrize("name", sorted(KATS))
@pytest.mark.parametrize("method,tol", [("rk4", 1e-6), ("rkf45", 1e-3)])
def test_reference_goldens(name, method, tol):
    kw, gold = KATS[name]
    print("case", name, method, flush=True)
    p = P.params_from_problem(H2, RHO0_2, hbar=1.0, kb=1.0, t_end=1.0, method=method, **kw)
    it = P.HierarchyIntegrator(p)
    it.Launch()
    r = it.density_matrix()
    assert_valid_density_matrix(r)
    assert np.abs(r - gold).max() < tol


def _oracle_from_prm(path, threads=4):
    op = PO.read_command_line(path, "rk4")
    return op, HO.Oracle(HO.Problem.from_prm(op), threads=threads)


def test_rhs_matches_oracle_c1():
    """Batched Get on random Hermitian and non-Hermitian input."""
    path = os.path.join(CFG, "c1.prm")
    op, orc = _oracle_from_prm(path)
    p = P.PhiParameters.read(path, "rk4")
    ctx = P.Context(p)
    assert ctx.num_nodes == orc.num_nodes == 495
    assert np.array_equal(ctx.index(), orc.index())
    rng = np.random.default_rng(11)
    for herm in (True, False):
        x = rng.normal(size=(495, 2, 2)) + 1j * rng.normal(size=(495, 2, 2))
        if herm:
            x = x + x.conj().transpose(0, 2, 1)
        y_gpu = ctx.rhs(x)
        y_ref = orc.rhs(x)
        scale = np.abs(y_ref).max()
        assert np.abs(y_gpu - y_ref).max() <= 1e-12 * scale


def test_rhs_matches_oracle_c2_subset():
    """7-site system at depth 4 (330 nodes), general input."""
    path = os.path.join(CFG, "c2.prm")
    op = PO.read_command_line(path, "rk4")
    op.hierarchy_truncation_level = 5
    orc = HO.Oracle(HO.Problem.from_prm(op), threads=4)
    p = P.PhiParameters.read(path, "rk4")
    p.set_int("hierarchy_truncation_level", 5)
    ctx = P.Context(p)
    assert ctx.num_nodes == orc.num_nodes == 330
    assert np.array_equal(ctx.index(), orc.index())
    rng = np.random.default_rng(12)
    x = rng.normal(size=(330, 7, 7)) + 1j * rng.normal(size=(330, 7, 7))
    y_gpu = ctx.rhs(x)
    y_ref = orc.rhs(x)
    assert np.abs(y_gpu - y_ref).max() <= 1e-12 * np.abs(y_ref).max()


def test_rhs_nodes_is_per_node_get_on_live_state():
    """phi_hip_rhs_nodes (PHI's per-node Get) on the context's current state:
    equal to the batched RHS of that state at the listed nodes and to the
    oracle's Get of those nodes; ids outside the hierarchy are rejected."""
    path = os.path.join(CFG, "c2.prm")
    op = PO.read_command_line(path, "rk4")
    op.hierarchy_truncation_level = 5
    orc = HO.Oracle(HO.Problem.from_prm(op), threads=4)
    p = P.PhiParameters.read(path, "rk4")
    p.set_int("hierarchy_truncation_level", 5)
    ctx = P.Context(p)
    rng = np.random.default_rng(13)
    x = rng.normal(size=(330, 7, 7)) + 1j * rng.normal(size=(330, 7, 7))
    ctx.set_state(x)
    ids = [0, 1, 7, 36, 329, 36, 120]
    got = ctx.rhs_nodes(ids)
    assert got.shape == (len(ids), 7, 7)
    batched = ctx.rhs(x)
    assert np.array_equal(got, batched[ids])
    ref = orc.rhs(x)[ids]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # the live state: after RK4 steps the same call sees the new state
    ctx.rk4_steps(3, op.timestep)
    st = ctx.state()
    assert np.array_equal(ctx.rhs_nodes([5, 0]), ctx.rhs(st)[[5, 0]])
    for bad in ([330], [-1], [0, 400]):
        with pytest.raises(P.PhiError):
            ctx.rhs_nodes(bad)
    ctx.close()


def test_rk4_c1_full_run_fp64(tmp_path):
    """c1 (495 nodes, 1000 RK4 steps): every output time within 1e-10."""
    path = os.path.join(CFG, "c1.prm")
    op, orc = _oracle_from_prm(path, threads=1)
    steps, ts, hist = orc.rk4(op.integration_time, op.timestep, op.output_minimum_timestep)
    full = tmp_path / "c1.full"
    out = tmp_path / "c1.out"
    p = P.PhiParameters.read(path, "rk4")
    p.set_string("output_filename", str(out))
    p.set_string("full_precision_output", str(full))
    p.set_int("verbose", -2)
    it = P.HierarchyIntegrator(p)
    it.Launch()
    info = it.info()
    assert info["steps"] == steps == 1000
    rows = np.loadtxt(full)
    # initial line, 100 in-loop lines, final line
    assert rows.shape[0] == len(ts) + 2
    t_gpu = rows[1:-1, 0]
    assert np.array_equal(t_gpu, ts)
    r_gpu = rows[1:-1, 1:].view(complex).reshape(-1, 2, 2).transpose(0, 2, 1)
    assert np.abs(r_gpu - hist).max() <= 1e-10
    # Whole hierarchy at the end.
    assert np.abs(it.state() - orc.state()).max() <= 1e-10
    # PHI-format text output: header equals the oracle's, one line per output.
    text = out.read_text()
    hdr = PO.full_header(op, orc.num_nodes)
    assert text.startswith(hdr.replace("#Threads=1", "#Threads=%d" % op.num_threads))


def test_long_rk4_run_autotunes_and_matches_oracle():
    """An RK4 run of 10 000 steps (c1 over ten times its integration time)
    takes the autotuned launch shape before it starts; its final state still
    equals the oracle's within 1e-10."""
    path = os.path.join(CFG, "c1.prm")
    op, orc = _oracle_from_prm(path, threads=8)
    t_end = 10 * op.integration_time
    steps, _, _ = orc.rk4(t_end, op.timestep, op.output_minimum_timestep)
    assert steps >= 10000
    p = P.PhiParameters.read(path, "rk4")
    p.set_double("integration_time", t_end)
    p.set_string("output_filename", "")
    p.set_int("verbose", -2)
    print("case: c1 RK4 over %d steps" % steps, flush=True)
    it = P.HierarchyIntegrator(p)
    it.Launch()
    assert it.info()["steps"] == steps
    assert np.abs(it.state() - orc.state()).max() <= 1e-10


def test_rk4_c2_short_fp64_and_fp32():
    """7-site exciton, depth 8 (6435 nodes): 40 RK4 steps, FP64 within 1e-10;
    FP32-bulk populations within 1e-6 of FP64."""
    path = os.path.join(CFG, "c2.prm")
    op, orc = _oracle_from_prm(path, threads=8)
    op_steps = 40
    orc.rk4(op_steps * op.timestep * (1 - 1e-9), op.timestep, 1.0)
    ref = orc.state()
    p = P.PhiParameters.read(path, "rk4")
    for prec, tol in (("fp64", 1e-10), ("fp32", None)):
        ctx = P.Context(p, precision=prec)
        ctx.rk4_steps(op_steps, op.timestep)
        s = ctx.state()
        if tol is not None:
            assert np.abs(s - ref).max() <= tol
        else:
            pop_err = np.abs(np.real(np.diag(s[0])) - np.real(np.diag(ref[0]))).max()
            print("fp32 population error", pop_err)
            assert pop_err <= 1e-6
        ctx.close()


def test_rkf45_matches_oracle_controller():
    """RKF45 on the weak-bath golden problem and on c1 for 0.2 ps: same
    attempted/accepted counts and final state within 1e-10."""
    cases = []
    kw, _ = KATS["matsubara"]
    cases.append((dict(H=H2, rho0=RHO0_2, hbar=1.0, kb=1.0, **kw), 1.0, 1e-6, 1e-6))
    op = PO.read_command_line(os.path.join(CFG, "c1.prm"), "rkf45")
    cases.append((dict(H=op.time_indepenent_hamiltonian, rho0=op.initial_density_matrix,
                       gamma=op.gamma, lambda_=op.lambda_, temperature=op.temperature,
                       lh=op.hierarchy_truncation_level, num_terms=op.num_matsubara_terms),
                  0.2, 1e-8, 1e-6))
    for kw, t_end, tol, mindt in cases:
        prob = HO.Problem(**{k: v for k, v in kw.items()})
        orc = HO.Oracle(prob, threads=1)
        res = orc.rkf45(t_end, 0.001, tol, mindt, 0.01)
        pkw = dict(kw)
        H = pkw.pop("H")
        rho0 = pkw.pop("rho0")
        p = P.params_from_problem(H, rho0, t_end=t_end, method="rkf45", **pkw)
        p.set_double("rkf45_tolerance", tol)
        p.set_double("rkf45_minimum_timestep", mindt)
        it = P.HierarchyIntegrator(p)
        it.Launch()
        info = it.info()
        assert info["steps"] == res["attempted"]
        assert info["accepted"] == res["accepted"]
        assert info["final_time"] == res["time"]
        assert np.abs(it.state() - orc.state()).max() <= 1e-10


def test_steady_state_c1_small():
    """Steady state of a dimer at depth 3: BiCGSTAB residual below tol, the
    solution is a fixed point of the oracle's operator and has unit trace."""
    op = PO.read_command_line(os.path.join(CFG, "c1.prm"), "bicgstab")
    kw = dict(gamma=op.gamma, lambda_=op.lambda_, temperature=op.temperature, lh=4,
              num_terms=op.num_matsubara_terms)
    p = P.params_from_problem(op.time_indepenent_hamiltonian, op.initial_density_matrix,
                              method="bicgstab", **kw)
    ctx = P.Context(p)
    rc, iters, res = ctx.steady_state(1e-11, 5000)
    assert rc == 0, (iters, res)
    x = ctx.state()
    orc = HO.Oracle(HO.Problem(op.time_indepenent_hamiltonian, op.initial_density_matrix, **kw))
    Lx = orc.rhs(x)
    assert abs(np.trace(x[0]) - 1) < 1e-10
    assert np.abs(Lx).max() < 1e-8


def test_steady_state_graph_batches_equal_eager_launches():
    """The BiCGSTAB batches replayed from a captured hipGraph (one capture,
    many replays with the host's residual check between them) give the
    kernel-by-kernel solve bit for bit: iterations, residual and solution
    (7-site system, depth 4)."""
    _, p = _subset_params("c2", 5)
    dt = p.get_double("timestep")
    out = {}
    for graphs in (False, True):
        print("case: steady state graphs=%s" % graphs, flush=True)
        ctx = P.Context(p)
        ctx.set_graphs(graphs)
        ctx.rk4_steps(200, dt)
        rc, iters, res = ctx.steady_state(1e-10, 20000)
        out[graphs] = (rc, iters, res, ctx.state())
        ctx.close()
    print("iterations %d / %d" % (out[False][1], out[True][1]))
    assert out[False][0] == 0 and out[False][1] > 3 * 16
    assert out[True][:3] == out[False][:3]
    assert np.array_equal(out[True][3], out[False][3])


def test_steady_state_iterations_match_dense_bicgstab():
    """The device-scalar BiCGSTAB takes the textbook iterates (van der
    Vorst 1992): residuals after k = 1..12 iterations equal those of a dense
    numpy BiCGSTAB on the oracle's steady-state operator (A = L with the
    trace row, b = e_0, x_0 = the initial state)."""
    op = PO.read_command_line(os.path.join(CFG, "c1.prm"), "bicgstab")
    kw = dict(gamma=op.gamma, lambda_=op.lambda_, temperature=op.temperature, lh=3,
              num_terms=op.num_matsubara_terms)
    p = P.params_from_problem(op.time_indepenent_hamiltonian, op.initial_density_matrix,
                              method="bicgstab", **kw)
    orc = HO.Oracle(HO.Problem(op.time_indepenent_hamiltonian, op.initial_density_matrix, **kw))
    ctx = P.Context(p)
    x0 = ctx.state()
    ctx.close()
    shape = x0.shape
    m = x0.size
    A = np.zeros((m, m), dtype=complex)
    for c in range(m):
        e = np.zeros(m, dtype=complex)
        e[c] = 1
        A[:, c] = orc.steady_matvec(e.reshape(shape)).reshape(-1)
    b = np.zeros(m, dtype=complex)
    b[0] = 1
    x = x0.reshape(-1).copy()
    r = b - A @ x
    rh = r.copy()
    pv = np.zeros(m, dtype=complex)
    v = np.zeros(m, dtype=complex)
    rho_old = alpha = omega = 1.0 + 0j
    res = [np.linalg.norm(r)]
    for _ in range(12):
        rho_new = np.vdot(rh, r)
        beta = (rho_new / rho_old) * (alpha / omega)
        pv = r + beta * (pv - omega * v)
        v = A @ pv
        alpha = rho_new / np.vdot(rh, v)
        s = r - alpha * v
        t = A @ s
        omega = np.vdot(t, s) / np.vdot(t, t)
        x = x + alpha * pv + omega * s
        r = s - omega * t
        rho_old = rho_new
        res.append(np.linalg.norm(r))
    for k in (1, 2, 3, 5, 8, 12):
        ctx = P.Context(p)
        rc, it, got = ctx.steady_state(1e-30, k)
        ctx.close()
        assert it == k
        assert abs(got - res[k]) <= 1e-9 * res[0], (k, got, res[k])


def test_steady_state_l_cycles_match_dense_bicgstab_l():
    """BiCGSTAB(ell) on the device (ell = 1, 2, 4) takes the iterates of the
    dense restatement of Sleijpen and Fokkema's algorithm on the oracle's
    steady-state operator: residuals after 1, 2 and 4 cycles agree."""
    from oracle import krylov_oracle as KO

    op = PO.read_command_line(os.path.join(CFG, "c1.prm"), "bicgstab")
    kw = dict(gamma=op.gamma, lambda_=op.lambda_, temperature=op.temperature, lh=3,
              num_terms=op.num_matsubara_terms)
    p = P.params_from_problem(op.time_indepenent_hamiltonian, op.initial_density_matrix,
                              method="bicgstab", **kw)
    orc = HO.Oracle(HO.Problem(op.time_indepenent_hamiltonian, op.initial_density_matrix, **kw))
    ctx = P.Context(p)
    x0 = ctx.state()
    # ell outside 1..8 or a negative step cap: an error, the state untouched
    for ell, cap in ((0, 10), (9, 10), (2, -1)):
        rc, _, _ = ctx.steady_state_l(ell, 1e-10, cap)
        assert rc != 0, (ell, cap)
        assert P.lib().phi_hip_last_error().decode()
    assert np.array_equal(ctx.state(), x0)
    ctx.close()
    m = x0.size
    A = np.zeros((m, m), dtype=complex)
    for c in range(m):
        e = np.zeros(m, dtype=complex)
        e[c] = 1
        A[:, c] = orc.steady_matvec(e.reshape(x0.shape)).reshape(-1)
    b = np.zeros(m, dtype=complex)
    b[0] = 1
    for ell in (1, 2, 4):
        res, _ = KO.bicgstab_l(A, b, x0, ell, 4)
        for k in (1, 2, 4):
            print("case: bicgstab(%d) %d cycles" % (ell, k), flush=True)
            ctx = P.Context(p)
            rc, it, got = ctx.steady_state_l(ell, 1e-30, k * ell)
            ctx.close()
            assert it == k * ell
            # r_1..r_ell are A^j r_0, whose Gram-Schmidt loses digits as
            # |A|^ell grows: measured 1.4e-9 of |r_0| at ell = 4, 4 cycles.
            tol = 1e-9 if ell <= 2 else 1e-8
            assert abs(got - res[k]) <= tol * res[0], (ell, k, got, res[k])


def test_steady_state_l_converges_to_the_bicgstab_fixed_point():
    """BiCGSTAB(2), (4) and (8, the largest ell) on the 7-site system at depth
    4 reach the tolerance with fewer matrix-vector products than BiCGSTAB
    needs, and all land on the same fixed point of the oracle's operator; graph
    replays of the batches of cycles equal eager launches bit for bit."""
    op, p = _subset_params("c2", 5)
    dt = p.get_double("timestep")
    ctx = P.Context(p)
    ctx.rk4_steps(200, dt)
    start = ctx.state()
    rc, it_u, res = ctx.steady_state(1e-10, 20000)
    assert rc == 0, (it_u, res)
    ref = ctx.state()
    ctx.close()
    orc = HO.Oracle(HO.Problem(op.time_indepenent_hamiltonian, op.initial_density_matrix,
                               gamma=op.gamma, lambda_=op.lambda_, temperature=op.temperature,
                               lh=5, num_terms=op.num_matsubara_terms))
    for ell in (2, 4, 8):
        runs = []
        for graphs in (True, False):
            print("case: bicgstab(%d) c2 lh 5 graphs=%s" % (ell, graphs), flush=True)
            ctx = P.Context(p)
            ctx.set_graphs(graphs)
            ctx.set_state(start)
            runs.append(ctx.steady_state_l(ell, 1e-10, 20000) + (ctx.state(),))
            ctx.close()
        # batches replayed from a hipGraph = eager launches, bit for bit
        assert runs[0][:3] == runs[1][:3]
        assert np.array_equal(runs[0][3], runs[1][3])
        rc, it, res, x = runs[0]
        print("ell %d: %d BiCG steps (plain BiCGSTAB %d), residual %.3g" % (ell, it, it_u, res))
        assert rc == 0 and res <= 1e-10, (it, res)
        assert it <= it_u
        assert abs(np.trace(x[0]) - 1) < 1e-9
        assert np.abs(orc.rhs(x)).max() < 1e-7
        assert np.abs(x[0] - ref[0]).max() < 1e-8


def test_integrator_bicgstab_words_run_bicgstab_l_and_bicgstab():
    """PHI's bicgstab method runs BiCGSTAB(ell) with the configured ell
    (2 by default) and bicgstabu plain BiCGSTAB; both converge to the same
    steady-state density matrix of the dimer at depth 4."""
    op = PO.read_command_line(os.path.join(CFG, "c1.prm"), "bicgstab")
    kw = dict(gamma=op.gamma, lambda_=op.lambda_, temperature=op.temperature, lh=4,
              num_terms=op.num_matsubara_terms)
    out = {}
    for method in ("bicgstab", "bicgstabu"):
        print("case: integrator %s" % method, flush=True)
        p = P.params_from_problem(op.time_indepenent_hamiltonian, op.initial_density_matrix,
                                  method=method, **kw)
        p.set_double("bicgstab_tolerance", 1e-11)
        it = P.HierarchyIntegrator(p)
        it.Launch()
        out[method] = (it.info(), it.density_matrix())
    for method, (info, rho) in out.items():
        assert info["bicgstab_residual"] <= 1e-11, (method, info)
        assert abs(np.trace(rho) - 1) < 1e-9
    assert int(out["bicgstab"][0]["bicgstab_iterations"]) % 2 == 0
    assert np.abs(out["bicgstab"][1] - out["bicgstabu"][1]).max() < 1e-8


def _subset_params(cfg, lh):
    path = os.path.join(CFG, cfg + ".prm")
    op = PO.read_command_line(path, "rk4")
    op.hierarchy_truncation_level = lh
    p = P.PhiParameters.read(path, "rk4")
    p.set_int("hierarchy_truncation_level", lh)
    return op, p


@pytest.mark.parametrize("cfg,lh,steps", [("c2", 5, 30), ("c3", 3, 20)])
def test_rk4_forms_and_launch_shapes_match_oracle(cfg, lh, steps):
    """Every RK4 stage order (PHI's classic k0..k3, the nested Taylor form
    and its increment form), both GEMM forms (matrix cores / element) and
    several block shapes give the oracle's state within 1e-10 (FP64)."""
    op, p = _subset_params(cfg, lh)
    orc = HO.Oracle(HO.Problem.from_prm(op), threads=8)
    dt = op.timestep
    orc.rk4(steps * dt * (1 - 1e-9), dt, 1.0)
    ref = orc.state()
    shapes = ((2, 0, 0, 0), (2, 1, 0, 0), (2, 3, 0, 7), (2, 1, 64, 0), (2, 2, 256, 5),
              (1, 0, 0, 0), (0, 0, 0, 0), (1, 1, 64, 0), (1, 2, 128, 0), (0, 3, 192, 0))
    for form in ("increment", "nested", "classic"):
        for gemm, tile, threads, blocks in shapes:
            print("case", cfg, form, gemm, tile, threads, blocks, flush=True)
            ctx = P.Context(p)
            ctx.set_rk4_form(form)
            try:
                ctx.set_launch(gemm, tile, threads, blocks)
            except P.PhiError as e:
                # a shape this system cannot take (e.g. 4+ elements per thread)
                print("  skipped:", e)
                ctx.close()
                continue
            ctx.rk4_steps(steps, dt)
            err = np.abs(ctx.state() - ref).max()
            ctx.close()
            assert err <= 1e-10, (form, gemm, tile, threads, blocks, err)


@pytest.mark.parametrize("cfg,lh,tile,per_stage", [
    ("c2", 5, 4, ((64, 0), (0, 7), (128, 0), (192, 300))),
    ("c3", 3, 1, ((192, 0), (0, 512), (256, 0), (0, 300)))])
def test_stage_launch_shapes_match_oracle(cfg, lh, tile, per_stage):
    """Each RK4 stage with threads and grid of its own (one element per
    thread and 2..4 per thread, fixed and resident grids) gives the oracle's
    state within 1e-10 (FP64); set_launch resets the stages to its shape;
    shapes the tile cannot take and other kernel forms are refused."""
    op, p = _subset_params(cfg, lh)
    orc = HO.Oracle(HO.Problem.from_prm(op), threads=8)
    dt = op.timestep
    steps = 20
    orc.rk4(steps * dt * (1 - 1e-9), dt, 1.0)
    ctx = P.Context(p)
    ne = ctx.N * ctx.N
    ctx.set_launch(2, tile, 0, 0)
    for s, (th, b) in enumerate(per_stage):
        ctx.set_stage_launch(s, th, b)
    th, bl = ctx.stage_launch()
    one_per_thread = (tile * ne + 63) // 64 * 64
    assert th == tuple(t or one_per_thread for t, _ in per_stage)
    assert bl == tuple(b for _, b in per_stage)
    print("case", cfg, "tile", tile, "stages", th, bl, flush=True)
    ctx.rk4_steps(steps, dt)
    assert np.abs(ctx.state() - orc.state()).max() <= 1e-10
    with pytest.raises(P.PhiError):
        ctx.set_stage_launch(4, 0, 0)
    with pytest.raises(P.PhiError):
        ctx.set_stage_launch(0, 100, 0)
    with pytest.raises(P.PhiError):
        ctx.set_stage_launch(1, 0, -1)
    if tile * ne > 4 * 64:
        with pytest.raises(P.PhiError):  # more than 4 elements per thread
            ctx.set_stage_launch(2, 64, 0)
    ctx.set_launch(2, tile, 0, 0)
    assert ctx.stage_launch() == ((one_per_thread,) * 4, (0,) * 4)
    ctx.set_launch(1, 0, 0, 0)
    with pytest.raises(P.PhiError):
        ctx.set_stage_launch(0, 0, 0)
    ctx.close()


@pytest.mark.parametrize("cfg,steps", [("c2", 1000), ("c3", 1000)])
def test_fp32_bulk_population_error_vs_fp64(cfg, steps):
    """FP32-bulk RK4 (FP32 RHS, FP64 state) against the FP64 path over 1 ps:
    population error <= 1e-6; coherences reported."""
    p = P.PhiParameters.read(os.path.join(CFG, cfg + ".prm"), "rk4")
    dt = p.get_double("timestep")
    out = {}
    for prec in ("fp64", "fp32"):
        ctx = P.Context(p, precision=prec)
        ctx.rk4_steps(steps, dt)
        out[prec] = ctx.rho0()
        ctx.close()
    d = out["fp32"] - out["fp64"]
    pop = np.abs(np.real(np.diag(d))).max()
    coh = np.abs(d - np.diag(np.diag(d))).max()
    print("%s fp32 population error %.3e coherence error %.3e" % (cfg, pop, coh))
    assert abs(np.trace(out["fp64"]) - 1) < 1e-10
    assert pop <= 1e-6


@pytest.mark.parametrize("cfg,lh,prec", [("c2", 5, "fp64"), ("c2", 5, "fp32"), ("c1", 9, "fp64")])
def test_graph_replay_matches_eager(cfg, lh, prec):
    """RK4 blocks replayed from a captured hipGraph give the kernel-by-kernel
    result bit for bit, over several replays with host reads in between."""
    _, p = _subset_params(cfg, lh)
    dt = p.get_double("timestep")
    states = {}
    for graphs in (False, True):
        ctx = P.Context(p, precision=prec)
        ctx.set_graphs(graphs)
        got = []
        for chunk in (70, 130, 3, 97):
            ctx.rk4_steps(chunk, dt)
            got.append(ctx.state())
        states[graphs] = got
        ctx.close()
    for a, b in zip(states[False], states[True]):
        assert np.array_equal(a, b)


def test_autotune_keeps_state_and_parity():
    """The autotuner leaves the state untouched and its chosen shape still
    gives the oracle's result (7-site system, depth 4)."""
    op, p = _subset_params("c2", 5)
    dt = op.timestep
    orc = HO.Oracle(HO.Problem.from_prm(op), threads=8)
    orc.rk4(20 * dt * (1 - 1e-9), dt, 1.0)
    ctx = P.Context(p)
    before = ctx.state()
    ms, shape = ctx.autotune(dt, steps=3)
    print("autotune: %.4f ms/step, shape %s" % (ms, shape))
    assert ms > 0
    assert np.array_equal(ctx.state(), before)
    ctx.rk4_steps(20, dt)
    assert np.abs(ctx.state() - orc.state()).max() <= 1e-10
    # cached: a second context of the same problem gets the same shape
    ctx2 = P.Context(p)
    ms2, shape2 = ctx2.autotune(dt, steps=3)
    assert shape2 == shape
    print("stage shapes", ctx.stage_launch())
    assert ctx2.stage_launch() == ctx.stage_launch()
    ctx.close()
    ctx2.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stage_times_keep_state(prec):
    """Stage-kernel timing leaves the state bit-identical, so a run that is
    timed and one that is not give the same result."""
    op, p = _subset_params("c2", 5)
    dt = op.timestep
    ctx = P.Context(p, precision=prec)
    ref = P.Context(p, precision=prec)
    ctx.rk4_steps(5, dt)
    ref.rk4_steps(5, dt)
    before = ctx.state()
    ms = ctx.rk4_stage_times(7, dt)
    assert ms.shape == (4,) and np.all(ms > 0)
    assert np.array_equal(ctx.state(), before)
    ctx.rk4_steps(5, dt)
    ref.rk4_steps(5, dt)
    assert np.array_equal(ctx.state(), ref.state())
    # The event-timed loop runs the same steps.
    ms = ctx.rk4_steps_timed(9, dt)
    ref.rk4_steps(9, dt)
    assert ms > 0
    assert np.array_equal(ctx.state(), ref.state())
    ctx.close()
    ref.close()


def _ensemble_members(count, max_lh):
    from phi_amd.configs.make_configs import c4_member
    Hs, depths = [], []
    for r in range(count):
        H, lh = c4_member(r)
        Hs.append(H)
        depths.append(min(lh, max_lh))
    return np.array(Hs), depths


def _member_params(p, H, lh):
    q = P.PhiParameters.read(os.path.join(CFG, "c2.prm"), "rk4")
    q.set_matrix("time_indepenent_hamiltonian", H)
    q.set_int("hierarchy_truncation_level", lh)
    return q


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_ensemble_matches_member_runs(prec):
    """An ensemble of 7-site hierarchies of uneven depth (static disorder)
    advances every member exactly as integrating it alone; the recorded
    history is the member mean; one member checked against the oracle."""
    p = P.PhiParameters.read(os.path.join(CFG, "c2.prm"), "rk4")
    dt = p.get_double("timestep")
    Hs, depths = _ensemble_members(6, 6)
    print("depths", depths, flush=True)
    ens = P.Ensemble(p, Hs, depths, precision=prec)
    assert ens.members == 6
    steps = 30
    ens.rk4_steps(steps, dt)
    slab = ens.state()
    r0 = ens.member_rho0()
    off = 0
    tol = 1e-12 if prec == "fp64" else 1e-6
    for m in range(6):
        ctx = P.Context(_member_params(p, Hs[m], depths[m]), precision=prec)
        assert ctx.num_nodes == ens.counts[m]
        ctx.rk4_steps(steps, dt)
        single = ctx.state()
        d = np.abs(slab[off:off + ctx.num_nodes] - single).max()
        print("member %d nodes %d max|d| %.3e" % (m, ctx.num_nodes, d), flush=True)
        assert d <= tol
        assert np.abs(r0[m] - ctx.rho0()).max() <= tol
        off += ctx.num_nodes
        ctx.close()
    assert off == ens.num_nodes
    if prec == "fp64":
        op = PO.read_command_line(os.path.join(CFG, "c2.prm"), "rk4")
        op.time_indepenent_hamiltonian = Hs[2]
        op.hierarchy_truncation_level = depths[2]
        orc = HO.Oracle(HO.Problem.from_prm(op), threads=8)
        orc.rk4(steps * dt * (1 - 1e-9), dt, 1.0)
        assert np.abs(r0[2] - orc.state()[0]).max() <= 1e-10
    ens.close()
    # the recorded history is the mean of the members' rho_0
    ens2 = P.Ensemble(p, Hs, depths, precision=prec)
    hist = ens2.rk4_history(21, dt, 10)
    assert hist.shape == (3, 7, 7)
    ens3 = P.Ensemble(p, Hs, depths, precision=prec)
    ens3.rk4_steps(21, dt)
    assert np.abs(hist[-1] - ens3.member_rho0().mean(axis=0)).max() <= 1e-13
    ens2.close()
    ens3.close()
