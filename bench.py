"""Benchmark: HEOM RK4 on the 7-site exciton (configs c2), one MI355X.

Workload: 7-site exciton, depth 8 (PHI HierarchyTruncation=9), 6435 nodes of
7x7 complex ADOs, RK4 with dt = 1 fs; one step = one RK4 step over every node
(4 fused RHS+stage kernels).  K timed steps (default 1000 = 1 ps) after W
warmup steps, with the state resident in HBM.  Prints one JSON line.

value = node-steps/s of the plain FP64 path; the FP32-bulk path (FP32 RHS,
FP64 state) is reported beside it with its population error against FP64.

roofline: the dominant kernel is the RK4 stage-1 kernel (RHS + acc += dt/3 k,
Y = rho + dt/2 k).  Algorithmic bytes per node for it are N^2 (48 + 2 s)
(reads Y, rho, acc; writes acc, Y'; s = 16 B for FP64) - stage 1 of the
per-step total N^2 (160 + 6 s); neighbour reads count as cache hits.  The
kernel's mean duration comes from HIP events on the engine's stream.

cpu_baseline: the oracle's pthread restatement of PHI's RK4 loop on the
host cores, timed on a bounded sample of the same workload.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CFG = os.path.join(ROOT, "phi_amd", "configs", "c2.prm")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(seconds=12.0):
    import numpy as np  # noqa: F401
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this integrator runs on a single MI355X "
              "(the multi-GPU sharding is a design, not a build).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import phi_amd as P

    p = P.PhiParameters.read(CFG, "rk4")
    N = p.get_int("num_states")
    dt = p.get_double("timestep")
    res = {}
    for prec in ("fp64", "fp32"):
        log("bench: %s context" % prec)
        ctx = P.Context(p, precision=prec)
        nodes = ctx.num_nodes
        ctx.rk4_steps(args.warmup, dt)
        ctx.synchronize()
        ctx2 = P.Context(p, precision=prec)
        ctx2.synchronize()
        t0 = time.perf_counter()
        ctx2.rk4_steps(args.steps, dt)
        ctx2.synchronize()
        el = time.perf_counter() - t0
        rho0 = ctx2.rho0()
        st_steps = min(args.steps, 200)
        ms = ctx.rk4_stage_times(st_steps, dt) / st_steps
        res[prec] = {"seconds": el, "node_steps_per_s": nodes * args.steps / el,
                     "ms_per_step": 1e3 * el / args.steps, "stage_ms": ms.tolist(),
                     "rho0": rho0}
        log("bench: %s %.3f ms/step, stage ms %s" % (prec, 1e3 * el / args.steps, ms))
        ctx.close()
        ctx2.close()
    pop64 = np.real(np.diag(res["fp64"]["rho0"]))
    pop32 = np.real(np.diag(res["fp32"]["rho0"]))
    pop_err = float(np.abs(pop64 - pop32).max())

    s = 16
    ne = N * N
    nodes = 6435
    stage1_bytes = ne * (48 + 2 * s) * nodes
    stage1_s = res["fp64"]["stage_ms"][1] * 1e-3
    achieved = stage1_bytes / stage1_s / 1e9
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
        "data": "synthetic (seeded 7-site exciton Hamiltonian, configs c2)",
        "config": {"workload": "7-site exciton HEOM, depth 8 (6435 nodes of 7x7), RK4 dt=1 fs",
                   "nodes": nodes, "num_states": N, "dt_ps": dt,
                   "time_to_1ps_s_fp64": res["fp64"]["ms_per_step"] * 1000 * 1e-3,
                   "time_to_1ps_s_fp32_bulk": res["fp32"]["ms_per_step"] * 1000 * 1e-3},
        "fp32_bulk": {"value": res["fp32"]["node_steps_per_s"], "unit": "node-steps/s",
                      "ms_per_step": res["fp32"]["ms_per_step"],
                      "population_error_vs_fp64": pop_err},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "rk4_stage<double,double,1>",
                     "stage_ms_fp64": res["fp64"]["stage_ms"],
                     "stage_ms_fp32": res["fp32"]["stage_ms"]},
    }
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
