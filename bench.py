"""Benchmark: HEOM RK4 over a static-disorder ensemble, one MI355X.

Workload (the largest single-GPU configuration): 512 realisations of the
7-site exciton (7 projector baths, site energies + N(0, 100 cm^-1) disorder,
depths drawn from U{4..8}, i.e. PHI HierarchyTruncation 5..9, 330-6435 nodes
each, 1.36 M nodes of 7x7 complex ADOs in one slab), RK4 with dt = 1 fs,
FP32-bulk right-hand side on the matrix cores with the FP64 state, the
member-averaged rho_0 recorded every 10 fs.  One step = one RK4 step over
every node of every member (4 persistent stage-kernel launches, increment
form of the nested step).  W warm-up steps, then K timed steps (default
1000 = 1 ps, so the time to 1 ps is measured, not extrapolated) with the
state resident in HBM, bracketed by device synchronisation.  Prints one JSON line on stdout.

value    node-steps/s of the FP32-bulk ensemble run (whole slab).  The same
         ensemble in plain FP64 is reported beside it ("fp64"), and the
         FP32-bulk result's population and coherence errors against it
         (every member's rho_0 and the member average).
roofline the dominant kernel, stream_stage (ensemble instantiation; the
         four launches of each RK4 step, nothing else but the snapshot of
         every 10th step runs in the loop): algorithmic bytes per launch over
         the mean launch duration, i.e. a step's bytes over the timed
         region's HIP-event time per step (events on the engine's stream).
         Algorithmic bytes per node (increment form of the nested RK4
         step, s = bytes per bulk complex element): stage 0 N^2 (16 + s),
         stage 1 N^2 2 s, stage 2 N^2 3 s, stage 3 N^2 (32 + 2 s),
         N^2 (48 + 8 s) per step; neighbour columns count as cache hits.  classic_accounting repeats the figures with
         the bytes of classic RK4, N^2 (160 + 6 s) per node-step.  traffic =
         calibrated HBM bytes per launch (FETCH_SIZE + WRITE_SIZE) of the
         four stages from the committed rocprofv3 counter summary.
cpu_baseline  the oracle's pthread restatement of PHI's RK4 loop (FP64,
         all host cores) on a bounded sample of the same ensemble: members
         in order, a fixed number of steps each, about 12 s.
c2_single  the 7-site depth-8 hierarchy alone (6435 nodes), FP64 and FP32
         bulk, with its own roofline figures.
c3_fp32  the 24-site ring at depth 4 (20475 nodes of 24x24), FP32-bulk RK4:
         node-steps/s, population / coherence error vs FP64, HBM fraction.
c3_rkf45 the 24-site ring with RKF45 (abs tol 1e-8) to 1 ps, FP64 and FP32
         bulk: attempts, time to 1 ps, HBM fraction of N^2 (112 + 29 s) per
         attempted node-step, FP32-vs-FP64 population / coherence error.
c2_steady_state  BiCGSTAB to residual 1e-10 from the 7-site state at 1 ps,
         and BiCGSTAB(2) from the same state.
c5_closure   the learned truncation closure: trained on depth-10 runs,
         depth 4 + closure vs plain depth 4 vs plain depth 10 on held-out
         Hamiltonians (populations and time to 1 ps).
--shape-ens/--shape-fp64/--shape-fp32/--shape-c3 pin a leg's launch shape
instead of autotuning it (profiles of the shape an earlier run chose):
form,tile,threads,blocks, optionally followed by threads,blocks of each of
the four RK4 stages (the 12 values a leg's kernel_shape records).
"""

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CFG = os.path.join(ROOT, "phi_amd", "configs", "c2.prm")
CFG3 = os.path.join(ROOT, "phi_amd", "configs", "c3.prm")
PROFILE_C2 = os.path.join(ROOT, "profiles", "rk4_stage_7site_depth8_traffic.txt")
PROFILE_ENS = os.path.join(ROOT, "profiles", "rk4_stage_ensemble_512x7site.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(N, s):
    """Algorithmic bytes per node of the four stage kernels of the increment
    form (the default RK4 order of the persistent kernels): stage 0 reads
    rho and writes K, stage 1 reads K and writes D, stage 2 reads D and K
    and writes D', stage 3 reads D', K and rho and writes rho:
    N^2 (48 + 8 s) per node-step."""
    ne = N * N
    return [ne * (16 + s), ne * (2 * s), ne * (3 * s), ne * (32 + 2 * s)]


def classic_stage_bytes(N, s):
    """Bytes per node of the four stages of classic RK4 with an FP64 state,
    an FP64 accumulator and bulk stage vectors streamed once per stage
    (stage 1 reads rho, writes acc and Y; stages 2-3 read Y, rho, acc and
    write acc, Y; stage 4 reads Y, acc and writes rho): N^2 (160 + 6 s) per
    node-step in total.  The nested form moves fewer bytes; this accounting
    is reported beside it."""
    ne = N * N
    return [ne * (32 + s), ne * (48 + 2 * s), ne * (48 + 2 * s), ne * (32 + s)]


def host_cores():
    """Threads this process may run on (its affinity set) and the host's
    online CPU count."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return aff, os.cpu_count()


def ensemble_members(members):
    from phi_amd.configs.make_configs import c4_member
    Hs, depths = [], []
    for r in range(members):
        H, lh = c4_member(r)
        Hs.append(H)
        depths.append(lh)
    return Hs, depths


def cpu_baseline(Hs, depths, seconds=12.0, steps_per_member=10):
    """The oracle (FP64, all host threads) on the ensemble's members in
    order, `steps_per_member` RK4 steps each, until `seconds` have passed."""
    from oracle import heom_oracle as HO
    from oracle import prm_oracle as PO
    threads, ncpu = host_cores()
    op = PO.read_command_line(CFG, "rk4")
    dt = op.timestep
    node_steps = 0
    m = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and m < len(Hs):
        op.time_indepenent_hamiltonian = Hs[m]
        op.hierarchy_truncation_level = depths[m]
        orc = HO.Oracle(HO.Problem.from_prm(op), threads=threads)
        orc.rk4(steps_per_member * dt * (1 - 1e-9), dt, 1.0)
        node_steps += orc.num_nodes * steps_per_member
        orc.close()
        m += 1
    el = time.perf_counter() - t0
    return {"value": node_steps / el, "unit": "node-steps/s", "cores": threads, "kind": "port",
            "sample": "members 0..%d of the 512-member ensemble (%d node-steps), %d RK4 steps "
                      "each, FP64, %d pthreads (affinity set of %d online CPUs), %.1f s"
                      % (m - 1, node_steps, steps_per_member, threads, ncpu, el)}


def profile_traffic(path, kernel_stage, prec):
    """HBM bytes per launch of a stage kernel from a committed counter
    summary (lines 'stage=<s> prec=<p> fetch_bytes=<F> write_bytes=<W>')."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*stage=(\d) prec=(\w+) fetch_bytes=([\d.eE+]+) write_bytes=([\d.eE+]+)",
                     line)
        if m and int(m.group(1)) == kernel_stage and m.group(2) == prec:
            return float(m.group(3)) + float(m.group(4))
    return None


def mean_traffic(path, prec):
    tr = [profile_traffic(path, s, prec) for s in range(4)]
    return sum(tr) / 4 if None not in tr else None


def errors(r, ref):
    """Population (diagonal) and coherence (off-diagonal) max |difference|
    of density matrices r and ref (..., N, N)."""
    d = np.asarray(r) - np.asarray(ref)
    diag = np.einsum("...ii->...i", d)
    off = d - np.einsum("...i,ij->...ij", diag, np.eye(d.shape[-1]))
    return float(np.abs(diag.real).max()), float(np.abs(off).max())


def launch_label(counts):
    g, e = counts
    if g == 0:
        return "eager launches (%d steps)" % e
    return "hipGraph blocks (%d steps) + eager head/tail (%d steps)" % (g, e)


def pin_shape(ctx, shape):
    """Pin a launch shape: (form, tile, threads, blocks), optionally
    followed by (threads, blocks) of each RK4 stage."""
    ctx.set_launch(*shape[:4])
    if len(shape) == 12:
        for s in range(4):
            ctx.set_stage_launch(s, shape[4 + 2 * s], shape[5 + 2 * s])


def full_shape(ctx, shape):
    """The common shape followed by (threads, blocks) of each RK4 stage."""
    th, bl = ctx.stage_launch()
    return tuple(shape[:4]) + tuple(v for s in range(4) for v in (th[s], bl[s]))


def copy_reference_gbs(nbytes=1 << 30, reps=10):
    """Device-to-device hipMemcpy rate of the HIP runtime the library
    loaded (nbytes per copy, read + write counted, HIP events on the null
    stream around `reps` copies after two untimed ones): the streaming rate
    this GPU reaches in practice, reported beside the published peak.  None
    when it cannot be measured."""
    import ctypes
    try:
        path = next(l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64" in l)
        hip = ctypes.CDLL(path)
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        ms = ctypes.c_float()
        ok = hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
        ok = ok and hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
        ok = ok and hip.hipMemset(a, 0, ctypes.c_size_t(nbytes)) == 0
        ok = ok and hip.hipEventCreate(ctypes.byref(e0)) == 0
        ok = ok and hip.hipEventCreate(ctypes.byref(e1)) == 0
        d2d = 3  # hipMemcpyDeviceToDevice
        for _ in range(2):
            ok = ok and hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), d2d) == 0
        ok = ok and hip.hipEventRecord(e0, None) == 0
        for _ in range(reps):
            ok = ok and hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), d2d) == 0
        ok = ok and hip.hipEventRecord(e1, None) == 0
        ok = ok and hip.hipEventSynchronize(e1) == 0
        ok = ok and hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        for e in (e0, e1):
            if e.value:
                hip.hipEventDestroy(e)
        for m in (a, b):
            if m.value:
                hip.hipFree(m)
        if not ok or ms.value <= 0:
            return None
        return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9
    except (OSError, StopIteration, AttributeError, ctypes.ArgumentError):
        return None


def run_rk4(P, p, prec, steps, warmup, shape=None):
    """RK4 timing of one configuration; `shape` (pin_shape's form) pins
    the launch shape, None autotunes it."""
    dt = p.get_double("timestep")
    ctx = P.Context(p, precision=prec)
    if shape is None:
        _, shape = ctx.autotune(dt, steps=5)
    else:
        pin_shape(ctx, shape)
    shape = full_shape(ctx, shape)
    ctx.rk4_steps(warmup, dt)
    ctx.synchronize()
    t0 = time.perf_counter()
    ev_ms = ctx.rk4_steps_timed(steps, dt)
    ctx.synchronize()
    el = time.perf_counter() - t0
    launch = launch_label(ctx.launch_counts())
    rho0 = ctx.rho0()
    ms = ctx.rk4_stage_times(min(200, max(steps, 20)), dt) / min(200, max(steps, 20))
    out = {"seconds": el, "node_steps_per_s": ctx.num_nodes * steps / el,
           "ms_per_step": 1e3 * el / steps, "event_ms_per_step": ev_ms / steps,
           "stage_ms": ms.tolist(), "rho0": rho0,
           "nodes": ctx.num_nodes, "shape": shape,
           "launch": launch}
    ctx.close()
    return out


def run_steady_state(P, np, p, steps, tol=1e-10, max_iter=20000):
    """BiCGSTAB steady state of the 7-site system from its state after
    `steps` RK4 steps (the reference configuration's second phase)."""
    dt = p.get_double("timestep")
    ctx = P.Context(p)
    ctx.rk4_steps(steps, dt)
    ctx.synchronize()
    start = ctx.state()
    t0 = time.perf_counter()
    rc, iters, resid = ctx.steady_state(tol, max_iter)
    el = time.perf_counter() - t0
    r0 = ctx.rho0()
    nodes = ctx.num_nodes
    N = r0.shape[0]
    # BiCGSTAB(2) (the reference's default steady-state solver) from the
    # same state
    ctx.set_state(start)
    ctx.synchronize()
    t0 = time.perf_counter()
    rc_l, steps_l, resid_l = ctx.steady_state_l(2, tol, max_iter)
    el_l = time.perf_counter() - t0
    trace_l = float(np.trace(ctx.rho0()).real)
    ctx.close()
    # 24 FP64 vector passes per iteration: 384 N^2 B per node
    per_iter = nodes * N * N * 384
    return {"tolerance": tol, "status": rc, "iterations": iters, "residual": resid,
            "seconds": el, "trace": float(np.trace(r0).real),
            "hbm_frac_iteration": (per_iter * iters / el / (HBM_PEAK_GBS * 1e9)) if el > 0 else None,
            "bicgstab_l2": {"status": rc_l, "bicg_steps": steps_l, "residual": resid_l,
                            "seconds": el_l, "trace": trace_l}}


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
                     "rho": r}
        log("bench: c3 rkf45 %s %d attempts %.3f s" % (prec, info["steps"], info["seconds"]))
    pe, ce = errors(out["fp32"]["rho"], out["fp64"]["rho"])
    for v in out.values():
        del v["rho"]
    out["population_error_fp32_vs_fp64"] = pe
    out["coherence_error_fp32_vs_fp64"] = ce
    out["workload"] = ("24-site ring, depth 4 (20475 nodes of 24x24), RKF45 abs tol 1e-8, "
                       "initial dt 1 fs, to 1 ps")
    return out


def run_ensemble(P, p, Hs, depths, prec, steps, warmup, dt, shape=None):
    """The ensemble in one slab: autotuned (or pinned) shape, W warm-up
    steps, then K steps with the member-averaged rho_0 every 10 steps,
    timed by the host clock (synchronised) and by HIP events on the
    engine's stream.  Also the stage kernels back to back."""
    ens = P.Ensemble(p, np.array(Hs), depths, precision=prec)
    if shape is None:
        _, shape = ens.autotune(dt, steps=3)
    else:
        pin_shape(ens, shape)
    shape = full_shape(ens, shape)
    ens.rk4_steps(warmup, dt)
    ens.synchronize()
    t0 = time.perf_counter()
    hist, ev_ms = ens.rk4_history(steps, dt, 10, timed=True)
    ens.synchronize()
    el = time.perf_counter() - t0
    launch = launch_label(ens.launch_counts())
    members_rho0 = ens.member_rho0()
    n_st = min(50, max(steps, 10))
    st_ms = ens.rk4_stage_times(n_st, dt) / n_st
    out = {"seconds": el, "event_ms": ev_ms, "nodes": ens.num_nodes, "hist": hist,
           "members_rho0": members_rho0, "shape": shape, "stage_ms": st_ms.tolist(),
           "launch": launch}
    ens.close()
    return out


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
                     dt=dt, every=10, precision="fp32", deep_precision="fp64", log=log)
    scale = 1000.0 / steps
    return {
        "workload": "7-site exciton, random Hamiltonians; network trained on %d depth-10 "
                    "trajectories (%d rows), evaluated on %d held-out ones, RK4 dt=1 fs: "
                    "depth 10 plain FP64 (increment form), depth 4 and depth 4 + closure "
                    "FP32 bulk (the closure run in PHI's classic order)"
                    % (train_members, len(X), test_members),
        "train_seconds_data": t_data, "train_loss_first_last": [float(losses[0]), float(losses[-1])],
        "time_to_1ps_s_depth10": res["deep"]["seconds"] * scale,
        "time_to_1ps_s_depth4": res["shallow"]["seconds"] * scale,
        "time_to_1ps_s_depth4_closure": res["closure"]["seconds"] * scale,
        "speedup_closure_vs_depth10": res["deep"]["seconds"] / res["closure"]["seconds"],
        "closure_vs_depth4_time": res["closure"]["seconds"] / res["shallow"]["seconds"],
        "fallback_threshold": getattr(net, "ymax", None),
        "fallback_predictions": res["closure"]["fallen"],
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
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-ensemble-fp64", action="store_true")
    ap.add_argument("--no-closure", action="store_true")
    ap.add_argument("--no-steady-state", action="store_true")
    ap.add_argument("--no-c3-rkf45", action="store_true")
    ap.add_argument("--ensemble-members", type=int, default=512)
    ap.add_argument("--closure-train", type=int, default=4096)
    ap.add_argument("--closure-test", type=int, default=256)
    ap.add_argument("--calibration-copy", action="store_true",
                    help="first copy 1 GiB device to device (a plain copy of known size, "
                         "the reference point of the HBM byte counters under rocprofv3)")
    def shape(s):
        v = tuple(int(x) for x in s.split(","))
        if len(v) not in (4, 12):
            raise argparse.ArgumentTypeError("a shape has 4 or 12 values")
        return v

    for leg in ("ens", "fp64", "fp32", "c3"):
        ap.add_argument("--shape-" + leg, type=shape, default=None,
                        help="launch shape form,tile,threads,blocks[,threads,blocks per RK4 "
                             "stage] of the %s leg instead of autotuning it (e.g. to profile "
                             "the shape a previous run chose)" % leg)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this integrator runs on a single MI355X "
              "(the multi-GPU sharding is a design, not a build).", file=sys.stderr)
        sys.exit(2)

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
    K, W = args.steps, args.warmup
    Hs, depths = ensemble_members(args.ensemble_members)

    # ---- headline: the ensemble, FP32 bulk --------------------------------------
    log("bench: ensemble fp32, %d members" % len(Hs))
    e32 = run_ensemble(P, p, Hs, depths, "fp32", K, W, dt, args.shape_ens)
    nodes = e32["nodes"]
    log("bench: ensemble fp32 %d nodes: %.4f ms/step (events %.4f), shape %s, %s, stage ms %s"
        % (nodes, 1e3 * e32["seconds"] / K, e32["event_ms"] / K, e32["shape"], e32["launch"],
           e32["stage_ms"]))
    sb = stage_bytes(N, 8)
    launch_ms = e32["event_ms"] / K / 4
    per_launch = sum(sb) * nodes / 4
    achieved = per_launch / (launch_ms * 1e-3) / 1e9
    classic = sum(classic_stage_bytes(N, 8)) * nodes / 4
    st_ms = e32["stage_ms"]
    dom = int(np.argmax(st_ms))
    ms_step = 1e3 * e32["seconds"] / K
    steps_per_ps = round(1.0 / dt)
    line = {
        "metric": "node-steps/s (RK4, 512-member static-disorder 7-site ensemble, FP32-bulk "
                  "RHS / FP64 state) + time-to-1ps",
        "value": nodes * K / e32["seconds"],
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded static-disorder realisations of the 7-site exciton, "
                "phi_amd/configs/make_configs.py c4_member)",
        "config": {"workload": "%d static-disorder realisations of the 7-site exciton (depth "
                               "U{4..8}, %d nodes of 7x7 in one slab), RK4 dt=1 fs, FP32-bulk "
                               "RHS on the matrix cores, FP64 state, member-averaged rho_0 "
                               "every 10 fs" % (len(Hs), nodes),
                   "members": len(Hs), "nodes": nodes, "num_states": N, "dt_ps": dt,
                   "rk4_form": "increment", "state_dtype": "f64",
                   "kernel_shape": e32["shape"], "launch": e32["launch"],
                   "time_to_1ps_s": ms_step * steps_per_ps * 1e-3,
                   "time_to_1ps": "measured" if K == steps_per_ps else
                                  "extrapolated from %d steps" % K},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": mean_traffic(PROFILE_ENS, "fp32"),
                     "kernel": "stream_stage, ensemble instantiation (the 4 launches of an "
                               "RK4 step, increment form, FP32 bulk)",
                     "bytes_per_launch": per_launch, "launch_ms": launch_ms,
                     "accounting": "increment-form RK4, N^2 (48 + 8 s) per node-step, s = 8",
                     "classic_accounting": {
                         "bytes_per_launch": classic,
                         "achieved": classic / (launch_ms * 1e-3) / 1e9,
                         "frac": classic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
                     "isolated_launches": {
                         "stage_ms": st_ms, "dominant_stage": dom,
                         "achieved": sb[dom] * nodes / (st_ms[dom] * 1e-3) / 1e9}},
    }
    cref = copy_reference_gbs()
    line["roofline"]["copy_reference"] = {
        "GBps": cref, "frac_of_copy": achieved / cref if cref else None,
        "what": "hipMemcpy device to device of 1 GiB, read + write counted: the streaming "
                "rate this GPU reaches in practice (peak stays the published 8 TB/s)"}
    log("bench: device copy reference %s GB/s" % ("%.0f" % cref if cref else "n/a"))
    if not args.no_ensemble_fp64:
        log("bench: ensemble fp64")
        e64 = run_ensemble(P, p, Hs, depths, "fp64", K, W, dt)
        pe, ce = errors(e32["members_rho0"], e64["members_rho0"])
        ape, ace = errors(e32["hist"], e64["hist"])
        ms64 = 1e3 * e64["seconds"] / K
        line["accuracy_vs_fp64"] = {
            "population_error": pe, "coherence_error": ce,
            "averaged_population_error": ape, "averaged_coherence_error": ace,
            "what": "max |FP32 bulk - FP64| over every member's rho_0 after W + K steps "
                    "(population: diagonal, coherence: off-diagonal), and over the "
                    "member-averaged rho_0 at every 10 fs snapshot"}
        line["fp64"] = {"node_steps_per_s": nodes * K / e64["seconds"], "ms_per_step": ms64,
                        "time_to_1ps_s": ms64 * steps_per_ps * 1e-3,
                        "kernel_shape": e64["shape"], "launch": e64["launch"],
                        "hbm_frac": sum(stage_bytes(N, 16)) * nodes / (e64["event_ms"] / K * 1e-3)
                                    / (HBM_PEAK_GBS * 1e9),
                        "stage_ms": e64["stage_ms"]}
        log("bench: ensemble fp64 %.4f ms/step; fp32 errors: population %.3e coherence %.3e"
            % (ms64, pe, ce))

    if not args.no_c2:
        c2 = {}
        for prec in ("fp64", "fp32"):
            log("bench: 7-site depth 8, %s" % prec)
            c2[prec] = run_rk4(P, p, prec, K, W, getattr(args, "shape_" + prec))
            log("bench: %s %.4f ms/step, shape %s, stage ms %s"
                % (prec, c2[prec]["ms_per_step"], c2[prec]["shape"], c2[prec]["stage_ms"]))
        pe, ce = errors(c2["fp32"]["rho0"], c2["fp64"]["rho0"])
        n2 = c2["fp64"]["nodes"]
        lm = c2["fp64"]["event_ms_per_step"] / 4
        pl = sum(stage_bytes(N, 16)) * n2 / 4
        line["c2_single"] = {
            "workload": "7-site exciton HEOM, depth 8 (6435 nodes of 7x7), RK4 dt=1 fs",
            "fp64": {"node_steps_per_s": c2["fp64"]["node_steps_per_s"],
                     "ms_per_step": c2["fp64"]["ms_per_step"],
                     "time_to_1ps_s": c2["fp64"]["ms_per_step"] * steps_per_ps * 1e-3,
                     "kernel_shape": c2["fp64"]["shape"], "launch": c2["fp64"]["launch"],
                     "stage_ms": c2["fp64"]["stage_ms"]},
            "fp32_bulk": {"node_steps_per_s": c2["fp32"]["node_steps_per_s"],
                          "ms_per_step": c2["fp32"]["ms_per_step"],
                          "kernel_shape": c2["fp32"]["shape"], "launch": c2["fp32"]["launch"],
                          "population_error_vs_fp64": pe, "coherence_error_vs_fp64": ce},
            "roofline_fp64": {"bound": "latency (state resident in MALL)",
                              "achieved": pl / (lm * 1e-3) / 1e9, "launch_ms": lm,
                              "frac_of_hbm": pl / (lm * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              "traffic": mean_traffic(PROFILE_C2, "fp64")}}
    if not args.no_c3:
        log("bench: 24-site depth 4, fp32 and fp64")
        p3 = P.PhiParameters.read(CFG3, "rk4")
        r3 = run_rk4(P, p3, "fp32", 50, 5, args.shape_c3)
        r3d = run_rk4(P, p3, "fp64", 50, 5)
        pe, ce = errors(r3["rho0"], r3d["rho0"])
        N3 = p3.get_int("num_states")
        b3 = stage_bytes(N3, 8)
        step_bytes = sum(b3) * r3["nodes"]
        line["c3_fp32"] = {
            "workload": "24-site ring, depth 4 (20475 nodes of 24x24), FP32-bulk RK4 dt=1 fs, "
                        "55 steps",
            "node_steps_per_s": r3["node_steps_per_s"], "ms_per_step": r3["ms_per_step"],
            "population_error_vs_fp64": pe, "coherence_error_vs_fp64": ce,
            "hbm_frac_step": step_bytes / (r3["event_ms_per_step"] * 1e-3) / (HBM_PEAK_GBS * 1e9),
            "hbm_frac_step_classic_accounting":
                r3["nodes"] * N3 * N3 * (160 + 6 * 8) / (r3["event_ms_per_step"] * 1e-3)
                / (HBM_PEAK_GBS * 1e9),
            "stage_ms": r3["stage_ms"], "kernel_shape": r3["shape"], "launch": r3["launch"],
            "fp64_ms_per_step": r3d["ms_per_step"]}
    if not args.no_c3 and not args.no_c3_rkf45:
        log("bench: 24-site depth 4, rkf45")
        line["c3_rkf45"] = run_c3_rkf45(P, np, log)
    if not args.no_steady_state:
        log("bench: 7-site steady state")
        line["c2_steady_state"] = run_steady_state(P, np, p, 1000)
    if not args.no_closure:
        log("bench: learned closure")
        line["c5_closure"] = run_closure(P, np, 1000, dt, args.closure_train, args.closure_test,
                                         log)
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(Hs, depths)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
