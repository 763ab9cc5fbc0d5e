"""GPU tests of the learned truncation closure: the network kernels against
the CPU oracle (oracle/mlp_oracle.py), training on the device, sample
extraction from a deep hierarchy, and the closed shallow integration.

Bars: forward outputs and gradients equal the oracle's kernel-arithmetic
mode (bfloat16 GEMM operands, float32 sums) to float32 summation-order
differences; training lowers the loss; the closure's virtual neighbours
carry exactly what the network predicts (checked through the linear
response of the RHS)."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import phi_amd as P  # noqa: E402
from oracle import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu

CFG = os.path.join(ROOT, "phi_amd", "configs")


def _flat(params):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in params]).astype(np.float32)


def _unflat(flat, dims):
    out, o = [], 0
    for i in range(3):
        nin, nout = dims[i], dims[i + 1]
        W = flat[o:o + nin * nout].reshape(nout, nin)
        o += nin * nout
        b = flat[o:o + nout]
        o += nout
        out.append((W, b))
    return out


@pytest.mark.parametrize("dims,batch", [((105, 256, 686), 96), ((13, 40, 21), 33)])
def test_mlp_forward_and_gradients_match_oracle(dims, batch):
    n_in, n_hid, n_out = dims
    rng = np.random.default_rng(7)
    params = M.init_params(n_in, n_hid, n_out, seed=1)
    params = [(W, (rng.standard_normal(b.shape) * 0.1).astype(np.float32)) for W, b in params]
    x = rng.standard_normal((batch, n_in)).astype(np.float32)
    t = rng.standard_normal((batch, n_out)).astype(np.float32)
    net = P.Mlp(n_in, n_hid, n_out, max_batch=128)
    net.set_params(_flat(params))
    assert np.array_equal(net.get_params(), _flat(params))
    y = net.forward(x)
    yk, ck = M.forward(params, x, exact=False)
    scale = np.abs(yk).max()
    err = np.abs(y - yk).max() / scale
    print("forward rel err %.3e" % err)
    assert err < 2e-3
    loss, g = net.gradients(x, t)
    gk = M.backward(params, ck, yk, t, exact=False)
    assert abs(loss - M.mse(yk, t)) < 1e-3 * M.mse(yk, t)
    gd = _unflat(g, [n_in, n_hid, n_hid, n_out])
    for li, ((dW, db), (dWk, dbk)) in enumerate(zip(gd, gk)):
        ew = np.abs(dW - dWk).max() / np.abs(dWk).max()
        eb = np.abs(db - dbk).max() / np.abs(dbk).max()
        print("layer %d grad rel err W %.3e b %.3e" % (li, ew, eb))
        assert ew < 5e-3 and eb < 5e-3


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(3)
    n_in, n_hid, n_out = 24, 64, 16
    A = rng.standard_normal((n_out, n_in)) * 0.3
    X = rng.standard_normal((4096, n_in)).astype(np.float32)
    T = np.tanh(X @ A.T).astype(np.float32)
    net = P.Mlp(n_in, n_hid, n_out, max_batch=256)
    net.set_params(_flat(M.init_params(n_in, n_hid, n_out, seed=2)))
    losses = net.train(X, T, batch=128, epochs=10, lr=3e-3, seed=5)
    print("epoch losses", losses)
    assert np.all(np.diff(losses) < 0)
    assert losses[-1] < 0.3 * losses[0]
    # the trained network's forward pass agrees with the oracle on its parameters
    p = _unflat(net.get_params(), [n_in, n_hid, n_hid, n_out])
    y = net.forward(X[:128])
    yk, _ = M.forward(p, X[:128], exact=False)
    assert np.abs(y - yk).max() < 2e-3 * np.abs(yk).max()


def _c2_params():
    return P.PhiParameters.read(os.path.join(CFG, "c2.prm"), "rk4")


def test_closure_samples_are_the_deep_hierarchys_ados():
    """Samples for depth-5 hierarchies taken from a depth-7 ensemble: inputs
    are the tier-4 ADOs scaled by their largest entry, targets their tier-5
    neighbours with the same scale (checked against the full state)."""
    from phi_amd.configs.make_configs import c4_member
    p = _c2_params()
    Hs = np.array([c4_member(r)[0] for r in range(3)])
    ens = P.Ensemble(p, Hs, [7, 7, 7])
    ens.rk4_steps(60, p.get_double("timestep"))
    X, T = ens.closure_samples(5)
    idx = P.Context(_member_param(p, Hs[0], 7)).index()
    state = ens.state()
    tier = idx.sum(axis=1)
    b_nodes = np.nonzero(tier == 4)[0]
    assert X.shape == (3 * len(b_nodes), 105) and T.shape == (3 * len(b_nodes), 686)
    n_per = ens.counts[0]
    for m in (0, 2):
        for bi in (0, len(b_nodes) // 2, len(b_nodes) - 1):
            n = b_nodes[bi]
            rho = state[m * n_per + n]
            x, s = M.encode(rho, idx[n], 4)
            row = m * len(b_nodes) + bi
            assert np.allclose(X[row], x, rtol=1e-6, atol=1e-7)
            kids = M.decode_children(T[row], s, 7, 7)
            for j in range(7):
                child = np.nonzero((idx == idx[n] + np.eye(7, dtype=int)[j]).all(axis=1))[0][0]
                ref = state[m * n_per + child]
                assert np.abs(kids[j] - ref).max() <= 1e-6 * max(s, np.abs(ref).max())


def _member_param(p, H, lh):
    q = _c2_params()
    q.set_matrix("time_indepenent_hamiltonian", H)
    q.set_int("hierarchy_truncation_level", lh)
    return q


def test_closure_with_zero_network_is_plain_truncation():
    """A network whose output is identically zero adds nothing: the closed
    shallow ensemble equals the plain one (classic RK4 order both)."""
    from phi_amd.configs.make_configs import c4_member
    p = _c2_params()
    dt = p.get_double("timestep")
    Hs = np.array([c4_member(r)[0] for r in range(4)])
    net = P.Mlp(105, 256, 686, max_batch=4096)
    params = M.init_params(105, 256, 686, seed=0)
    params[2] = (np.zeros_like(params[2][0]), np.zeros_like(params[2][1]))
    net.set_params(_flat(params))
    closed = P.Ensemble(p, Hs, [5] * 4)
    closed.attach_closure(net)
    plain = P.Ensemble(p, Hs, [5] * 4)
    plain.set_rk4_form("classic")
    closed.rk4_steps(50, dt)
    plain.rk4_steps(50, dt)
    assert np.abs(closed.member_rho0() - plain.member_rho0()).max() <= 1e-13


def test_closure_injects_predicted_neighbours():
    """With a network that returns a fixed pattern c (last layer bias only),
    one RK4 step of tiny dt from a random Hermitian state differs from the
    plain truncation by dt times the next-neighbour coupling of the
    predicted ADOs s*c at every last-tier node n and mode j (PHI's DoNext,
    hierarchy_updater.cc:216-233, in the Hermitian form the stage kernels
    use: column j += i p rho_c(:, j), row j -= conj of it, p =
    sqrt((n_j + 1) |c_j|)); all other nodes agree to O(dt)."""
    from oracle import heom_oracle as HO
    from oracle import prm_oracle as PO
    from phi_amd.configs.make_configs import c4_member
    p = _c2_params()
    H = c4_member(0)[0]
    net = P.Mlp(105, 256, 686, max_batch=4096)
    params = M.init_params(105, 256, 686, seed=0)
    rng = np.random.default_rng(9)
    b3 = (rng.standard_normal(686) * 0.1).astype(np.float32)
    params = [params[0], params[1], (np.zeros_like(params[2][0]), b3)]
    net.set_params(_flat(params))
    lh = 3
    closed = P.Ensemble(p, np.array([H]), [lh])
    closed.attach_closure(net)
    plain = P.Ensemble(p, np.array([H]), [lh])
    plain.set_rk4_form("classic")
    n_real = plain.num_nodes
    idx = P.Context(_member_param(p, H, lh)).index()
    x0 = 0.1 * (rng.standard_normal((n_real, 7, 7)) + 1j * rng.standard_normal((n_real, 7, 7)))
    x0 = x0 + x0.conj().transpose(0, 2, 1)
    plain.set_state(x0)
    closed.set_state(np.concatenate([x0, np.zeros((closed.num_nodes - n_real, 7, 7))]))
    dt = 1e-8
    closed.rk4_steps(1, dt)
    plain.rk4_steps(1, dt)
    diff = (closed.state()[:n_real] - plain.state()[:n_real]) / dt
    op = PO.read_command_line(os.path.join(CFG, "c2.prm"), "rk4")
    op.hierarchy_truncation_level = lh
    _, abs_corr, _ = HO.Oracle(HO.Problem.from_prm(op)).constants()
    expect = np.zeros_like(diff)
    boundary = np.nonzero(idx.sum(axis=1) == lh - 1)[0]
    for n in boundary:
        s = M.scale(x0[n])
        kids = M.decode_children(b3, s, 7, 7)
        for j in range(7):
            pref = np.sqrt((idx[n, j] + 1) * abs_corr[j])
            col = 1j * pref * kids[j][:, j]
            expect[n][:, j] += col
            expect[n][j, :] += np.conj(col)
    scale = np.abs(expect).max()
    err = np.abs(diff - expect).max()
    print("closure coupling: max |expect| %.3e, max err %.3e" % (scale, err))
    assert scale > 0
    assert err <= 1e-3 * scale


def test_fp32_closure_run_tracks_fp64():
    """A closure run with FP32 bulk (PHI's classic RK4 order with the
    accumulator and a copy of rho in FP32; the stage inputs are FP32 values
    in either FP32 form) follows the FP64 closure run (FP64 accumulator and
    state): populations within 1e-6, coherences reported, after 200 steps
    of a network with small random outputs."""
    from phi_amd.configs.make_configs import c4_member
    p = _c2_params()
    dt = p.get_double("timestep")
    Hs = np.array([c4_member(r)[0] for r in range(4)])
    params = M.init_params(105, 256, 686, seed=4)
    params[2] = (params[2][0] * 0.05, params[2][1] * 0.05)
    net = P.Mlp(105, 256, 686, max_batch=4096)
    net.set_params(_flat(params))
    out = {}
    for prec in ("fp64", "fp32"):
        print("case: closure run", prec, flush=True)
        ens = P.Ensemble(p, Hs, [5] * 4, precision=prec)
        ens.attach_closure(net)
        ens.rk4_steps(200, dt)
        out[prec] = ens.member_rho0()
        ens.close()
    d = out["fp32"] - out["fp64"]
    pop = np.abs(np.real(np.einsum("mii->mi", d))).max()
    coh = np.abs(d - np.einsum("mii,ij->mij", d, np.eye(7))).max()
    print("closure fp32 vs fp64: population %.3e, coherence %.3e" % (pop, coh))
    assert np.isfinite(out["fp64"]).all()
    assert pop <= 1e-6


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fused_closure_pass_equals_unfused_and_network(prec):
    """The fused closure launch (inputs + 3 layers + scatter) gives the same
    states as the unfused kernels, and the virtual slots written at stage 0
    hold the network's prediction for the boundary ADOs."""
    from phi_amd.configs.make_configs import c4_member
    p = _c2_params()
    Hs = np.array([c4_member(r)[0] for r in range(3)])
    rng = np.random.default_rng(11)
    params = M.init_params(105, 256, 686, seed=4)
    params = [(W, (rng.standard_normal(b.shape) * 0.05).astype(np.float32)) for W, b in params]
    net = P.Mlp(105, 256, 686, max_batch=4096)
    net.set_params(_flat(params))
    lh = 4
    states = {}
    ens0 = None
    for fused in (True, False):
        ens = P.Ensemble(p, Hs, [lh] * 3, precision=prec)
        ens.attach_closure(net)
        ens.set_closure_fused(fused)
        n_real = ens.counts[0]
        span = ens.num_nodes // 3
        if ens0 is None:
            x0 = 0.05 * (rng.standard_normal((ens.num_nodes, 7, 7))
                         + 1j * rng.standard_normal((ens.num_nodes, 7, 7)))
            ens0 = x0 + x0.conj().transpose(0, 2, 1)
        ens.set_state(ens0)
        ens.rk4_steps(1, 1e-4)
        states[fused] = ens.state()
        ens.close()
    d = np.abs(states[True] - states[False]).max()
    print("fused vs unfused max |d| %.3e" % d)
    assert d <= 1e-12 * np.abs(states[False]).max()
    # virtual slots of member 1 after the step = network(boundary ADOs of x0)
    idx = P.Context(_member_param(p, Hs[0], lh)).index()
    bnodes = np.nonzero(idx.sum(axis=1) == lh - 1)[0]
    m = 1
    xs, ss = [], []
    for n in bnodes:
        x, s = M.encode(ens0[m * span + n], idx[n], lh - 1)
        xs.append(x)
        ss.append(s)
    y = net.forward(np.array(xs))
    for bi in (0, len(bnodes) // 2, len(bnodes) - 1):
        kids = M.decode_children(y[bi], ss[bi], 7, 7)
        for j in range(7):
            # only the column the stage kernels read (site j) is written; the
            # rest of the slot keeps the seeded values
            slot = states[True][m * span + n_real + bi * 7 + j]
            tol = (1e-5 if prec == "fp64" else 1e-5) * max(1e-30, np.abs(kids[j]).max())
            assert np.abs(slot[:, j] - kids[j][:, j]).max() <= tol


def test_failed_attach_leaves_the_ensemble_unchanged():
    """A network the fused pass cannot take (hidden width 48, not a multiple
    of 32) is refused, and the ensemble it was offered to still steps exactly
    as before: same slab, offsets and tiles, equal to its members run alone."""
    from phi_amd.configs.make_configs import c4_member
    p = _c2_params()
    dt = p.get_double("timestep")
    Hs = np.array([c4_member(r)[0] for r in range(3)])
    ens = P.Ensemble(p, Hs, [5] * 3)
    ens.rk4_steps(7, dt)
    before = ens.state()
    net = P.Mlp(105, 48, 686, max_batch=256)
    with pytest.raises(P.PhiError):
        ens.attach_closure(net)
    assert np.array_equal(ens.state(), before)
    ens.rk4_steps(40, dt)
    slab = ens.state()
    off = 0
    for m in range(3):
        ctx = P.Context(_member_param(p, Hs[m], 5))
        ctx.rk4_steps(47, dt)
        assert np.abs(slab[off:off + ctx.num_nodes] - ctx.state()).max() <= 1e-12
        off += ctx.num_nodes
        ctx.close()
    assert off == ens.num_nodes
    net.close()
    ens.close()


@pytest.mark.parametrize("fused", [True, False])
def test_fallback_to_plain_truncation(fused):
    """The closure's fallback: a network whose every output exceeds the
    threshold (bias 5, ymax 1) has all its predictions replaced by zeros,
    so the closed ensemble equals plain truncation (classic RK4 order both)
    and every boundary-node prediction of every stage is counted; below the
    threshold nothing falls back and the run differs from plain truncation."""
    from phi_amd.configs.make_configs import c4_member
    p = _c2_params()
    dt = p.get_double("timestep")
    Hs = np.array([c4_member(r)[0] for r in range(3)])
    params = M.init_params(105, 256, 686, seed=0)
    params[2] = (np.zeros_like(params[2][0]), np.full(686, 5.0, dtype=np.float32))
    net = P.Mlp(105, 256, 686, max_batch=4096)
    net.set_params(_flat(params))
    lh, steps = 5, 20
    plain = P.Ensemble(p, Hs, [lh] * 3)
    plain.set_rk4_form("classic")
    plain.rk4_steps(steps, dt)
    boundary = int((plain.index().sum(axis=1) == lh - 1).sum())
    for ymax, expect_fallen in ((1.0, 3 * boundary * 4 * steps), (10.0, 0)):
        print("case: fallback fused=%s ymax=%g" % (fused, ymax), flush=True)
        closed = P.Ensemble(p, Hs, [lh] * 3)
        closed.attach_closure(net)
        closed.set_closure_fused(fused)
        assert closed.closure_fallback(ymax) == 0
        closed.rk4_steps(steps, dt)
        fallen = closed.closure_fallback()
        d = np.abs(closed.member_rho0() - plain.member_rho0()).max()
        print("  fallen %d (expected %d), max |closed - plain| = %.3e" % (fallen, expect_fallen, d))
        assert fallen == expect_fallen
        if expect_fallen:
            assert d <= 1e-13
        else:
            assert d > 1e-10
        closed.close()
    plain.close()
