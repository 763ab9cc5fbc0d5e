"""Learned truncation closure: training data, training and evaluation.

A shallow hierarchy (PHI depth `shallow`, last tier L = shallow - 1) loses
the couplings of its last tier to tier L + 1.  The closure network predicts
those next-tier ADOs from each last-tier ADO (phi_amd.Mlp, 2 N^2 + D inputs,
2 N^2 D outputs); the engine feeds the predictions to the stage kernels as
virtual neighbours (Ensemble.attach_closure).

Pipeline (every step on the GPU through the C ABI):
  samples   deep ensembles (PHI depth `deep`) of random Hamiltonians are
            integrated; every `every` steps the last-tier ADOs of the shallow
            hierarchy and their next-tier neighbours are read out
            (Ensemble.closure_samples), a random fraction kept;
  train     Adam on the samples (Mlp.train, bf16 GEMM operands, float32
            accumulation); the fallback threshold is set from the targets:
            a prediction more than FALLBACK_FACTOR times the largest
            training target falls back to plain truncation;
  evaluate  held-out Hamiltonians integrated plain-deep, plain-shallow and
            shallow + closure; populations compared with the deep run, and
            the time to 1 ps of each.

There is no counterpart in PHI, which only truncates (hierarchy_integrator.cc
1204-1338).
"""

import os
import time

import numpy as np

import phi_amd as P

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIG = os.path.join(HERE, "configs", "c2.prm")


def base_params(precision="fp64"):
    """The 7-site bath model of the closure study (7 projector baths,
    lambda 35 cm^-1, gamma 20 ps^-1, 77 K, no Matsubara terms)."""
    p = P.PhiParameters.read(CONFIG, "rk4")
    return p


def hamiltonians(start, count):
    from phi_amd.configs.make_configs import c5_hamiltonian
    return np.array([c5_hamiltonian(i) for i in range(start, start + count)])


def generate_samples(Hs, deep, shallow, steps, dt, every, keep, seed=0, members=256,
                     precision="fp32", log=print):
    """(X, T) float32 rows from deep runs of the Hamiltonians Hs."""
    rng = np.random.default_rng(seed)
    p = base_params()
    xs, ts = [], []
    for b0 in range(0, len(Hs), members):
        H = Hs[b0:b0 + members]
        t0 = time.perf_counter()
        ens = P.Ensemble(p, H, [deep] * len(H), precision=precision)
        done = 0
        while done < steps:
            k = min(every, steps - done)
            ens.rk4_steps(k, dt)
            done += k
            X, T = ens.closure_samples(shallow)
            # ADOs that are still zero (the hierarchy fills tier by tier) teach nothing
            live = np.abs(X[:, :-H.shape[1]]).max(axis=1) > 0
            sel = np.nonzero(live & (rng.random(len(X)) < keep))[0]
            xs.append(X[sel])
            ts.append(T[sel])
        ens.close()
        log("closure samples: members %d..%d, %d rows so far, %.1f s"
            % (b0, b0 + len(H) - 1, sum(len(x) for x in xs), time.perf_counter() - t0))
    return np.concatenate(xs), np.concatenate(ts)


def init_flat(n_in, n_hidden, n_out, seed=0):
    """He-initialised parameters (last layer scaled by 0.1), flat float32."""
    rng = np.random.default_rng(seed)
    dims = [n_in, n_hidden, n_hidden, n_out]
    parts = []
    for i in range(3):
        W = rng.standard_normal((dims[i + 1], dims[i])) * np.sqrt(2.0 / dims[i])
        if i == 2:
            W *= 0.1
        parts += [W.reshape(-1), np.zeros(dims[i + 1])]
    return np.concatenate(parts).astype(np.float32)


FALLBACK_FACTOR = 4.0


def train(X, T, hidden=256, epochs=20, batch=512, lr=1e-3, lr_final=1e-5, seed=0, log=print):
    """Adam with the learning rate decaying geometrically from lr to
    lr_final over the epochs: against a constant 1e-3, the decay halved the
    closure's largest population error against depth 10 (512 training
    trajectories, 64 held-out Hamiltonians) at an equal training loss."""
    net = P.Mlp(X.shape[1], hidden, T.shape[1], max_batch=max(batch, 8192))
    net.ymax = FALLBACK_FACTOR * float(np.abs(T).max())
    net.set_params(init_flat(X.shape[1], hidden, T.shape[1], seed))
    t0 = time.perf_counter()
    losses = net.train(X, T, batch=batch, epochs=epochs, lr=lr, seed=seed, lr_final=lr_final)
    log("closure training: %d rows, %d epochs, loss %.3e -> %.3e, %.1f s"
        % (len(X), epochs, losses[0], losses[-1], time.perf_counter() - t0))
    return net, losses


def _run(ens, steps, dt, every):
    ens.synchronize()
    t0 = time.perf_counter()
    hist = []
    per_member = []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        ens.rk4_steps(k, dt)
        done += k
        per_member.append(np.real(np.einsum("mii->mi", ens.member_rho0())))
    ens.synchronize()
    el = time.perf_counter() - t0
    return np.array(per_member), el


def evaluate(Hs, net, deep, shallow, steps, dt, every, precision="fp64", deep_precision="fp64",
             log=print):
    """Population histories (time, member, site) of plain deep, plain
    shallow and shallow + closure runs of Hs, their errors against the deep
    run and the wall time of each (after autotuning).  The plain runs take
    the default RK4 order (the fastest for a linear right-hand side); the
    closure makes the right-hand side nonlinear, so its run takes PHI's
    classic order.  `precision` is the bulk precision of the shallow runs,
    `deep_precision` that of the deep reference run."""
    p = base_params()
    out = {}
    for name, lh, closed, prec in (("deep", deep, False, deep_precision),
                                   ("shallow", shallow, False, precision),
                                   ("closure", shallow, True, precision)):
        ens = P.Ensemble(p, Hs, [lh] * len(Hs), precision=prec)
        if closed:
            ens.attach_closure(net)
            if getattr(net, "ymax", None):
                ens.closure_fallback(net.ymax)
        # time a copy of the run after autotuning on a scratch ensemble
        ens.autotune(dt, steps=3)
        pops, el = _run(ens, steps, dt, every)
        out[name] = {"pops": pops, "seconds": el, "nodes": ens.num_nodes}
        if closed:
            out[name]["fallen"] = ens.closure_fallback()
        log("closure eval: %s (depth %d%s, %s): %.3f s for %d steps, %d slab nodes"
            % (name, lh, ", closure" if closed else "", prec, el, steps, ens.num_nodes))
        ens.close()
    ref = out["deep"]["pops"]
    for name in ("shallow", "closure"):
        out[name]["max_pop_err"] = float(np.abs(out[name]["pops"] - ref).max())
        out[name]["mean_pop_err"] = float(np.abs(out[name]["pops"] - ref).mean())
    return out


def timing_run(members=256, shallow=5, steps=10, dt=0.001, precision="fp32", seed=3,
               autotune=True, log=print):
    """A depth-`shallow` ensemble of `members` held-out Hamiltonians with a
    closure network of random weights attached, advanced `steps` RK4 steps
    (after autotuning, or with the default launch shape): the launches of a
    closure run (stage kernels and fused closure passes, one per stage) for
    timing and profiling, with no accuracy meaning."""
    p = base_params()
    n_in, n_out = 2 * 49 + 7, 2 * 49 * 7
    net = P.Mlp(n_in, 256, n_out, max_batch=8192)
    net.set_params(init_flat(n_in, 256, n_out, seed))
    ens = P.Ensemble(p, hamiltonians(100000, members), [shallow] * members, precision=precision)
    ens.attach_closure(net)
    if autotune:
        ens.autotune(dt, steps=3)
    ens.rk4_steps(2, dt)
    ens.synchronize()
    t0 = time.perf_counter()
    ens.rk4_steps(steps, dt)
    ens.synchronize()
    el = time.perf_counter() - t0
    log("closure timing run: %d members, depth %d, %s, %d slab nodes: %.4f ms/step"
        % (members, shallow, precision, ens.num_nodes, 1e3 * el / steps))
    ens.close()
    net.close()
    return el / steps


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(
        description="closure timing run (random network weights; launches only)")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--depth", type=int, default=5, help="PHI depth (HierarchyTruncation)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp64"))
    ap.add_argument("--no-autotune", action="store_true",
                    help="default launch shape (short runs under hardware counters)")
    a = ap.parse_args()
    timing_run(a.members, a.depth, a.steps, precision=a.precision, autotune=not a.no_autotune)
