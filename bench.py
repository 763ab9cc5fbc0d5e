"""Benchmark: HEOM RK4 on the 7-site exciton, one MI355X.

Workload (the reference's headline configuration): 7-site exciton with 7
projector baths, depth 8 (PHI HierarchyTruncation=9), 6435 nodes of 7x7
complex ADOs, RK4 with dt = 1 fs; one step = one RK4 step over every node
(4 fused RHS + stage kernels, nested form).  W warm-up steps, then K timed
steps (default 1000 = 1 ps) with the state resident in HBM, bracketed by
device synchronisation.  Prints one JSON line on stdout.

value    node-steps/s of the plain FP64 path (max |drho| <= 1e-10 against the
         oracle in the GPU tests); the FP32-bulk path (FP32 RHS, FP64 state)
         is reported beside it with its population error against FP64 after
         the same W + K steps.
roofline the dominant kernel, stream_stage (the four launches of each
         FP64 RK4 step; nothing else runs in the loop): algorithmic bytes
         per launch over the mean launch duration, i.e. a step's bytes over
         the timed region's HIP-event time per step (events on the engine's
         stream; launch gaps included).  Algorithmic bytes per node (nested
         RK4 form, s = bytes per bulk complex element): stage 0 N^2 (16 + s),
         stages 1-2 N^2 (16 + 2 s), stage 3 N^2 (32 + s), N^2 (80 + 6 s)
         per step; neighbour columns count as cache hits.
         isolated_launches: each stage launched back to back on the same
         buffers after the timed region (no producer-consumer hand-off).
         classic_accounting repeats the figures with the bytes of classic
         RK4 (N^2 (160 + 6 s) per node-step), which the nested form beats.
         traffic = mean HBM bytes per launch of the four stages from the
         committed rocprofv3 counter summary (profiles/), when present.
--shape-fp64/--shape-fp32/--shape-c3 pin a leg's launch shape instead of
autotuning it (profiles of the shape an earlier run chose).
cpu_baseline  the oracle's pthread restatement of PHI's RK4 loop (FP64) on
         the host cores, timed on a bounded sample of the same workload.
c3_fp32  the 24-site ring at depth 4 (20475 nodes of 24x24), FP32-bulk RK4:
         node-steps/s and the HBM fraction of its stage kernels.
c3_rkf45 the 24-site ring with RKF45 (abs tol 1e-8) to 1 ps, FP64 and FP32
         bulk: attempts, time to 1 ps, HBM fraction of N^2 (112 + 29 s) per
         attempted node-step, FP32-vs-FP64 population error.
c2_steady_state  BiCGSTAB to residual 1e-10 from the 7-site state at 1 ps.
c4_ensemble  512 static-disorder realisations of the 7-site system with
         depths U{4..8} in one slab (persistent kernels, autotuned shape,
         hipGraph step blocks): time to 1 ps, FP32-bulk and FP64.
c5_closure   the learned truncation closure: trained on depth-10 runs,
         depth 4 + closure vs plain depth 4 vs plain depth 10 on held-out
         Hamiltonians (populations and time to 1 ps).
"""

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CFG = os.path.join(ROOT, "phi_amd", "configs", "c2.prm")
CFG3 = os.path.join(ROOT, "phi_amd", "configs", "c3.prm")
PROFILE = os.path.join(ROOT, "profiles", "rk4_stage_7site_depth8_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(N, s):
    """Algorithmic bytes per node of the four nested-form stage kernels."""
    ne = N * N
    return [ne * (16 + s), ne * (16 + 2 * s), ne * (16 + 2 * s), ne * (32 + s)]


def classic_stage_bytes(N, s):
    """Bytes per node of the four stages of classic RK4 with an FP64 state,
    an FP64 accumulator and bulk stage vectors streamed once per stage
    (stage 1 reads rho, writes acc and Y; stages 2-3 read Y, rho, acc and
    write acc, Y; stage 4 reads Y, acc and writes rho): N^2 (160 + 6 s) per
    node-step in total.  The nested form moves fewer bytes; this accounting
    is reported beside it."""
    ne = N * N
    return [ne * (32 + s), ne * (48 + 2 * s), ne * (48 + 2 * s), ne * (32 + s)]


def cpu_baseline(seconds=12.0):
    from oracle import heom_oracle as HO
    from oracle import prm_oracle as PO
    op = PO.read_command_line(CFG, "rk4")
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 16))
    orc = HO.Oracle(HO.Problem.from_prm(op), threads=threads)
    dt = op.timestep
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        orc.rk4(10 * dt * (1 - 1e-9), dt, 1.0)
        steps += 10
    el = time.perf_counter() - t0
    return {"value": steps * orc.num_nodes / el, "unit": "node-steps/s", "cores": threads,
            "kind": "port",
            "sample": "%d RK4 steps of the 7-site depth-8 hierarchy (6435 nodes), FP64, "
                      "%d pthreads, %.1f s" % (steps, threads, el)}


def profile_traffic(kernel_stage, prec):
    """HBM bytes per launch of a stage kernel from the committed counter
    summary (lines 'stage=<s> prec=<p> fetch_bytes=<F> write_bytes=<W>')."""
    if not os.path.exists(PROFILE):
        return None
    for line in open(PROFILE):
        m = re.match(r"\s*stage=(\d) prec=(\w+) fetch_bytes=([\d.eE+]+) write_bytes=([\d.eE+]+)",
                     line)
        if m and int(m.group(1)) == kernel_stage and m.group(2) == prec:
            return float(m.group(3)) + float(m.group(4))
    return None


def run_rk4(P, p, prec, steps, warmup, shape=None):
    """RK4 timing of one configuration; `shape` = (form, tile, threads,
    blocks) pins the launch shape, None autotunes it."""
    dt = p.get_double("timestep")
    ctx = P.Context(p, precision=prec)
    if shape is None:
        _, shape = ctx.autotune(dt, steps=5)
    else:
        ctx.set_launch(*shape)
        shape = tuple(shape)
    ctx.rk4_steps(warmup, dt)
    ctx.synchronize()
    t0 = time.perf_counter()
    ev_ms = ctx.rk4_steps_timed(steps, dt)
    ctx.synchronize()
    el = time.perf_counter() - t0
    rho0 = ctx.rho0()
    ms = ctx.rk4_stage_times(min(200, max(steps, 20)), dt) / min(200, max(steps, 20))
    out = {"seconds": el, "node_steps_per_s": ctx.num_nodes * steps / el,
           "ms_per_step": 1e3 * el / steps, "event_ms_per_step": ev_ms / steps,
           "stage_ms": ms.tolist(), "rho0": rho0,
           "nodes": ctx.num_nodes, "shape": shape,
           "launch": "hipGraph blocks" if ctx.graphs() else "eager"}
    ctx.close()
    return out


def run_steady_state(P, np, p, steps, tol=1e-10, max_iter=20000):
    """BiCGSTAB steady state of the 7-site system from its state after
    `steps` RK4 steps (the reference configuration's second phase)."""
    dt = p.get_double("timestep")
    ctx = P.Context(p)
    ctx.rk4_steps(steps, dt)
    ctx.synchronize()
    t0 = time.perf_counter()
    rc, iters, resid = ctx.steady_state(tol, max_iter)
    el = time.perf_counter() - t0
    r0 = ctx.rho0()
    nodes = ctx.num_nodes
    N = r0.shape[0]
    ctx.close()
    # 24 FP64 vector passes per iteration: 384 N^2 B per node
    per_iter = nodes * N * N * 384
    return {"tolerance": tol, "status": rc, "iterations": iters, "residual": resid,
            "seconds": el, "trace": float(np.trace(r0).real),
            "hbm_frac_iteration": (per_iter * iters / el / (HBM_PEAK_GBS * 1e9)) if el > 0 else None}


def run_c3_rkf45(P, np, log):
    """The 24-site ring with RKF45 (abs tol 1e-8) to 1 ps, FP64 and FP32
    bulk: attempts, time to 1 ps, attempted node-steps/s and the HBM
    fraction of the attempt bytes N^2 (112 + 29 s)."""
    out = {}
    for prec in ("fp64", "fp32"):
        p3 = P.PhiParameters.read(CFG3, "rkf45")
        p3.set_string("output_filename", "")
        p3.set_int("verbose", -2)
        p3.set_int("bulk_precision", 1 if prec == "fp32" else 0)
        it = P.HierarchyIntegrator(p3)
        it.Launch()
        info = it.info()
        r = it.density_matrix()
        N = r.shape[0]
        s = 8 if prec == "fp32" else 16
        nodes = int(info["num_nodes"])
        bytes_attempt = nodes * N * N * (112 + 29 * s)
        out[prec] = {"attempts": int(info["steps"]), "accepted": int(info["accepted"]),
                     "time_to_1ps_s": info["seconds"],
                     "node_steps_per_s": nodes * info["steps"] / info["seconds"],
                     "hbm_frac_attempt": bytes_attempt * info["steps"] / info["seconds"]
                                         / (HBM_PEAK_GBS * 1e9),
                     "pops": np.real(np.diag(r))}
        log("bench: c3 rkf45 %s %d attempts %.3f s" % (prec, info["steps"], info["seconds"]))
    err = float(np.abs(out["fp32"]["pops"] - out["fp64"]["pops"]).max())
    for v in out.values():
        del v["pops"]
    out["population_error_fp32_vs_fp64"] = err
    out["workload"] = ("24-site ring, depth 4 (20475 nodes of 24x24), RKF45 abs tol 1e-8, "
                       "initial dt 1 fs, to 1 ps")
    return out


def run_ensemble(P, np, steps, dt, members, log):
    """Static-disorder ensemble: `members` realisations of the 7-site system
    (site energies + N(0, 100 cm^-1), depth U{4..8}) in one slab, FP32-bulk
    and FP64 RK4 to `steps` steps with the member-averaged rho_0 recorded
    every 10 steps."""
    from phi_amd.configs.make_configs import c4_member
    p = P.PhiParameters.read(CFG, "rk4")
    Hs, depths = [], []
    for r in range(members):
        H, lh = c4_member(r)
        Hs.append(H)
        depths.append(lh)
    out = {}
    for prec in ("fp32", "fp64"):
        ens = P.Ensemble(p, np.array(Hs), depths, precision=prec)
        _, shape = ens.autotune(dt, steps=3)
        ens.synchronize()
        t0 = time.perf_counter()
        hist = ens.rk4_history(steps, dt, 10)
        ens.synchronize()
        el = time.perf_counter() - t0
        out[prec] = {"seconds": el, "nodes": ens.num_nodes, "hist": hist, "shape": shape}
        log("bench: ensemble %s %d members %d nodes: %.3f s" % (prec, members, ens.num_nodes, el))
        ens.close()
    pops = {k: np.real(np.einsum("tii->ti", v["hist"])) for k, v in out.items()}
    return {
        "workload": "%d realisations of the 7-site exciton (site-energy disorder N(0,100), "
                    "depth U{4..8}), one slab, RK4 dt=1 fs, member-averaged populations every 10 fs"
                    % members,
        "members": members, "nodes": out["fp32"]["nodes"],
        "time_to_1ps_s_fp32_bulk": out["fp32"]["seconds"] * (1000.0 / steps),
        "time_to_1ps_s_fp64": out["fp64"]["seconds"] * (1000.0 / steps),
        "node_steps_per_s_fp32_bulk": out["fp32"]["nodes"] * steps / out["fp32"]["seconds"],
        "node_steps_per_s_fp64": out["fp64"]["nodes"] * steps / out["fp64"]["seconds"],
        "averaged_population_error_fp32_vs_fp64": float(np.abs(pops["fp32"] - pops["fp64"]).max()),
        "kernel_shape_fp32": out["fp32"]["shape"]}


def run_closure(P, np, steps, dt, train_members, test_members, log):
    """Learned truncation closure: train on deep (depth 10) trajectories of
    random 7-site Hamiltonians, then integrate held-out ones at depth 4 with
    and without the closure and at depth 10 (FP64), to 1 ps."""
    from phi_amd import closure as C
    t0 = time.perf_counter()
    X, T = C.generate_samples(C.hamiltonians(0, train_members), deep=11, shallow=5, steps=1000,
                              dt=dt, every=100, keep=0.25, seed=1, members=256, log=log)
    t_data = time.perf_counter() - t0
    net, losses = C.train(X, T, epochs=20, batch=512, lr=1e-3, seed=2, log=log)
    res = C.evaluate(C.hamiltonians(100000, test_members), net, deep=11, shallow=5, steps=steps,
                     dt=dt, every=10, log=log)
    scale = 1000.0 / steps
    return {
        "workload": "7-site exciton, random Hamiltonians; network trained on %d depth-10 "
                    "trajectories (%d rows), evaluated on %d held-out ones, FP64 RK4 dt=1 fs"
                    % (train_members, len(X), test_members),
        "train_seconds_data": t_data, "train_loss_first_last": [float(losses[0]), float(losses[-1])],
        "time_to_1ps_s_depth10": res["deep"]["seconds"] * scale,
        "time_to_1ps_s_depth4": res["shallow"]["seconds"] * scale,
        "time_to_1ps_s_depth4_closure": res["closure"]["seconds"] * scale,
        "speedup_closure_vs_depth10": res["deep"]["seconds"] / res["closure"]["seconds"],
        "max_population_error_vs_depth10": {"depth4": res["shallow"]["max_pop_err"],
                                            "depth4_closure": res["closure"]["max_pop_err"]},
        "mean_population_error_vs_depth10": {"depth4": res["shallow"]["mean_pop_err"],
                                             "depth4_closure": res["closure"]["mean_pop_err"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-closure", action="store_true")
    ap.add_argument("--no-steady-state", action="store_true")
    ap.add_argument("--no-c3-rkf45", action="store_true")
    ap.add_argument("--ensemble-members", type=int, default=512)
    ap.add_argument("--closure-train", type=int, default=512)
    ap.add_argument("--closure-test", type=int, default=256)
    ap.add_argument("--calibration-copy", action="store_true",
                    help="first copy 1 GiB device to device (a plain copy of known size, "
                         "the reference point of the HBM byte counters under rocprofv3)")
    shape = lambda s: tuple(int(v) for v in s.split(","))
    for leg in ("fp64", "fp32", "c3"):
        ap.add_argument("--shape-" + leg, type=shape, default=None,
                        help="launch shape form,tile,threads,blocks of the %s leg instead of "
                             "autotuning it (e.g. to profile the shape a previous run chose)" % leg)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this integrator runs on a single MI355X "
              "(the multi-GPU sharding is a design, not a build).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import phi_amd as P

    if args.calibration_copy:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("bench: calibration copy of %d bytes done" % (a.numel() * 4))
        del a, b

    p = P.PhiParameters.read(CFG, "rk4")
    N = p.get_int("num_states")
    dt = p.get_double("timestep")
    res = {}
    for prec in ("fp64", "fp32"):
        log("bench: 7-site depth 8, %s" % prec)
        res[prec] = run_rk4(P, p, prec, args.steps, args.warmup,
                            getattr(args, "shape_" + prec))
        log("bench: %s %.4f ms/step, shape %s, stage ms %s"
            % (prec, res[prec]["ms_per_step"], res[prec]["shape"], res[prec]["stage_ms"]))
    pop64 = np.real(np.diag(res["fp64"]["rho0"]))
    pop32 = np.real(np.diag(res["fp32"]["rho0"]))
    pop_err = float(np.abs(pop64 - pop32).max())

    nodes = res["fp64"]["nodes"]
    sb = stage_bytes(N, 16)
    st_ms = res["fp64"]["stage_ms"]
    dom = int(np.argmax(st_ms))
    # The four stage launches of a step, timed over the timed region.
    launch_ms = res["fp64"]["event_ms_per_step"] / 4
    per_launch = sum(sb) * nodes / 4
    achieved = per_launch / (launch_ms * 1e-3) / 1e9
    tr = [profile_traffic(s, "fp64") for s in range(4)]
    traffic = sum(tr) / 4 if None not in tr else None
    steps_per_ps = round(1.0 / dt)
    line = {
        "metric": "node-steps/s (RK4, 7-site exciton depth 8, FP64) + time-to-1ps",
        "value": res["fp64"]["node_steps_per_s"],
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["fp64"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 7-site exciton Hamiltonian, phi_amd/configs/c2.prm)",
        "config": {"workload": "7-site exciton HEOM, depth 8 (6435 nodes of 7x7), RK4 dt=1 fs",
                   "nodes": nodes, "num_states": N, "dt_ps": dt, "rk4_form": "nested",
                   "kernel_shape_fp64": res["fp64"]["shape"],
                   "kernel_shape_fp32": res["fp32"]["shape"],
                   "launch_fp64": res["fp64"]["launch"],
                   "launch_fp32": res["fp32"]["launch"],
                   "time_to_1ps_s_fp64": res["fp64"]["ms_per_step"] * steps_per_ps * 1e-3,
                   "time_to_1ps_s_fp32_bulk": res["fp32"]["ms_per_step"] * steps_per_ps * 1e-3},
        "fp32_bulk": {"value": res["fp32"]["node_steps_per_s"], "unit": "node-steps/s",
                      "ms_per_step": res["fp32"]["ms_per_step"],
                      "population_error_vs_fp64": pop_err,
                      "stage_ms": res["fp32"]["stage_ms"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "stream_stage (the 4 launches of a nested RK4 step, FP64)",
                     "bytes_per_launch": per_launch, "launch_ms": launch_ms,
                     "isolated_launches": {
                         "stage_ms_fp64": st_ms,
                         "dominant_stage": dom,
                         "achieved": sb[dom] * nodes / (st_ms[dom] * 1e-3) / 1e9},
                     "classic_accounting": {
                         "bytes_per_launch": sum(classic_stage_bytes(N, 16)) * nodes / 4,
                         "achieved": sum(classic_stage_bytes(N, 16)) * nodes / 4
                                     / (launch_ms * 1e-3) / 1e9,
                         "frac_step": nodes * N * N * (160 + 6 * 16)
                                      / (res["fp64"]["ms_per_step"] * 1e-3) / (HBM_PEAK_GBS * 1e9)}},
    }
    if not args.no_c3:
        log("bench: 24-site depth 4, fp32")
        p3 = P.PhiParameters.read(CFG3, "rk4")
        r3 = run_rk4(P, p3, "fp32", 50, 5, args.shape_c3)
        N3 = p3.get_int("num_states")
        b3 = stage_bytes(N3, 8)
        step_bytes = sum(b3) * r3["nodes"]
        line["c3_fp32"] = {
            "workload": "24-site ring, depth 4 (20475 nodes of 24x24), FP32-bulk RK4 dt=1 fs",
            "node_steps_per_s": r3["node_steps_per_s"], "ms_per_step": r3["ms_per_step"],
            "hbm_frac_step": step_bytes / (r3["ms_per_step"] * 1e-3) / (HBM_PEAK_GBS * 1e9),
            "hbm_frac_step_classic_accounting":
                r3["nodes"] * N3 * N3 * (160 + 6 * 8) / (r3["ms_per_step"] * 1e-3)
                / (HBM_PEAK_GBS * 1e9),
            "stage_ms": r3["stage_ms"], "kernel_shape": r3["shape"]}
    if not args.no_c3 and not args.no_c3_rkf45:
        log("bench: 24-site depth 4, rkf45")
        line["c3_rkf45"] = run_c3_rkf45(P, np, log)
    if not args.no_steady_state:
        log("bench: 7-site steady state")
        line["c2_steady_state"] = run_steady_state(P, np, p, 1000)
    if not args.no_ensemble:
        log("bench: static-disorder ensemble")
        line["c4_ensemble"] = run_ensemble(P, np, 1000, dt, args.ensemble_members, log)
    if not args.no_closure:
        log("bench: learned closure")
        line["c5_closure"] = run_closure(P, np, 1000, dt, args.closure_train, args.closure_test,
                                         log)
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
