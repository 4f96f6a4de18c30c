#!/usr/bin/env python3
"""Benchmark: LB2 bounded nodes/s and time-to-proven-optimum on one MI355X.

Default workload (one step = one solve): Taillard ta011, 20 jobs x 10
machines, int32 times in [1, 99], solved to proven optimum from an infinite
upper bound with the device-resident B&B using plain LB2, pools of 4096
children per sub-batch and ten sub-batches per round.  `value` is children
bounded per second over the K timed solves; ms_per_step is the
time-to-optimum.

Also reported ("batched"): the 50x20 batched bounding throughput, 2^20
synthetic nodes (prefix/suffix lengths uniform, random permutations) per
launch, inputs resident in HBM.

The CPU baseline is the oracle (the C restatement of the reference CPU path)
on this host: the same ta011 solve with the reference's DFS explorer, one
thread, and a 2^15-node sample of the batched workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_VALU_TOPS = 256 * 64 * 2.4e9 / 1e12  # int32 lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz
PEAK_HBM_GBS = 8000.0


def measured_traffic(workload):
    """HBM bytes per bound-kernel dispatch from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE passes scaled by a copy-kernel
    calibration), or None when absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    return rec["bytes_per_dispatch"] if rec else None


def node_work(n, m, l1, l2):
    """algorithmic int ops per node: P(n + 4u + 4) + 2m d1 + m(m+1) d2"""
    P = m * (m - 1) // 2
    d1 = l1.astype(np.int64) + 1
    d2 = n - l2.astype(np.int64)
    u = n - d1 - d2
    return P * (n + 4 * u + 4) + 2 * m * d1 + m * (m + 1) * d2


def bench_solve(wsbb_amd, steps, warmup):
    p = wsbb_amd.taillard(11)
    eng = wsbb_amd.Engine(p)
    eng.search_profile(1)
    for _ in range(warmup):
        eng.solve(pool_size=4096, max_batches=10)
    eng.synchronize()
    tot = dict(bounded=0, work=0, bound_ms=0.0, launches=0)
    best = None
    t0 = time.perf_counter()
    for _ in range(steps):
        res = eng.solve(pool_size=4096, max_batches=10)
        ms, launches = eng.search_profile(-1)
        tot["bounded"] += res["stats"]["bounded"]
        tot["work"] += res["stats"]["work_ops"]
        tot["bound_ms"] += ms
        tot["launches"] += launches
        best = res["best"]
        assert res["proven"]
    eng.synchronize()
    t1 = time.perf_counter()
    stats = res["stats"]
    eng.close()
    return dict(elapsed=t1 - t0, best=best, stats=stats, **tot)


def bench_batched(wsbb_amd, reps=20, autotune=True):
    p = wsbb_amd.taillard(51)
    m, n = p.shape
    N = 1 << 20
    permT, l1, l2 = wsbb_amd.random_nodes(n, N, 0x5EED)
    eng = wsbb_amd.Engine(p)
    tuned = eng.autotune(N, reps=3) if autotune else []
    dp, d1, d2, do = eng.malloc(permT.nbytes), eng.malloc(4 * N), eng.malloc(4 * N), eng.malloc(4 * N)
    eng.h2d(dp, permT)
    eng.h2d(d1, l1)
    eng.h2d(d2, l2)
    eng.bound_device(dp, d1, d2, N, do)
    eng.synchronize()
    eng.timer_start()
    for _ in range(reps):
        eng.bound_device(dp, d1, d2, N, do)
    ms = eng.timer_stop() / reps
    for d in (dp, d1, d2, do):
        eng.free(d)
    eng.close()
    work = int(node_work(n, m, l1, l2).sum())
    bytes_alg = N * (n + 12)  # uint8 permutation + two int32 limits in, int32 bound out
    return dict(N=N, ms=ms, work=work, bytes=bytes_alg, tuned=tuned, permT=permT, l1=l1, l2=l2,
                p=p)


def cpu_baselines(bt):
    import oracle
    p = oracle.taillard(11)[3]
    o = oracle.Oracle(p)
    t0 = time.perf_counter()
    best, st, _ = o.solve_dfs()
    dt = time.perf_counter() - t0
    solve = dict(value=st["bounded"] / dt, seconds=dt, bounded=st["bounded"], best=best)
    ob = oracle.Oracle(bt["p"])
    S = 1 << 15
    t0 = time.perf_counter()
    ref = ob.bound_batch(bt["permT"][:, :S], bt["l1"][:S], bt["l2"][:S], threads=1)
    dtb = time.perf_counter() - t0
    return solve, dict(value=S / dtb, seconds=dtb, sample=S, ref=ref)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("all", "solve", "batched", "copy"), default="all",
                    help="solve = ta011 time-to-optimum only; batched = 50x20 launches only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build targets a single MI355X "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import wsbb_amd
    if args.workload == "copy":  # calibration of the hardware byte counters
        eng = wsbb_amd.Engine(wsbb_amd.taillard(1))
        nbytes = 256 << 20
        src, dst = eng.malloc(nbytes), eng.malloc(nbytes)
        eng.timer_start()
        for _ in range(args.steps):
            eng.copy_device(dst, src, nbytes)
        ms = eng.timer_stop() / args.steps
        eng.free(src)
        eng.free(dst)
        print(json.dumps({"workload": "copy", "bytes_read": nbytes, "bytes_written": nbytes,
                          "ms": ms, "GBps": 2 * nbytes / ms / 1e6}), flush=True)
        return
    if args.workload == "batched":
        bt = bench_batched(wsbb_amd, reps=args.steps, autotune=not args.no_autotune)
        print(json.dumps({"workload": "batched", "ms_per_launch": bt["ms"],
                          "nodes_per_s": bt["N"] / (bt["ms"] / 1e3)}), flush=True)
        return
    sv = bench_solve(wsbb_amd, args.steps, args.warmup)
    bt = None if args.workload == "solve" else bench_batched(
        wsbb_amd, autotune=not args.no_autotune)
    cpu_solve = cpu_b = None
    if not args.no_cpu and bt:
        cpu_solve, cpu_b = cpu_baselines(bt)

    value = sv["bounded"] / sv["elapsed"]
    bound_s = sv["bound_ms"] / 1e3
    ach = sv["work"] / bound_s / 1e12 if bound_s > 0 else None
    line = {
        "metric": "LB2 bounded nodes/s + time-to-proven-optimum, 1 MI355X vs host CPU",
        "value": value,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sv["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (Taillard generator, seeds of the reference)",
        "config": {"workload": "ta011 20x10 solved to proven optimum, plain LB2, 4096-child "
                               "pools, 10 sub-batches per round, UB=+inf",
                   "parallelism": "single GPU"},
        "time_to_optimum_s": sv["elapsed"] / args.steps,
        "optimum": sv["best"],
        "nodes_bounded_per_solve": sv["stats"]["bounded"],
        "rounds_per_solve": sv["stats"]["rounds"],
        "roofline": {"bound": "valu", "kernel": "bound_kernel (children of a round)",
                     "achieved": ach, "peak": PEAK_VALU_TOPS, "unit": "Tint-op/s",
                     "frac": (ach / PEAK_VALU_TOPS) if ach else None,
                     "traffic": measured_traffic("solve"),
                     "algorithmic_bytes": sv["stats"]["bounded"] / max(sv["stats"]["rounds"], 1)
                     * (20 + 12),
                     "avg_launch_us": sv["bound_ms"] * 1e3 / max(sv["launches"], 1),
                     "launches": sv["launches"],
                     "share_of_solve": bound_s / sv["elapsed"]},
    }
    if cpu_solve:
        line["cpu_baseline"] = {"value": cpu_solve["value"], "unit": "nodes/s", "cores": 1,
                                "kind": "port",
                                "sample": "ta011 solved to proven optimum by the oracle's DFS "
                                          "explorer (reference CPU path), 1 thread",
                                "time_to_optimum_s": cpu_solve["seconds"],
                                "optimum": cpu_solve["best"]}
    if bt:
        ach_b = bt["work"] / (bt["ms"] / 1e3) / 1e12
        line["batched"] = {
            "workload": "ta051 50x20, 2^20 random nodes per launch, d1~U{0..49}, "
                        "d2~U{0..49-d1}, inputs resident in HBM",
            "nodes_per_s": bt["N"] / (bt["ms"] / 1e3),
            "ms_per_launch": bt["ms"],
            "roofline": {"bound": "valu", "achieved": ach_b, "peak": PEAK_VALU_TOPS,
                         "unit": "Tint-op/s", "frac": ach_b / PEAK_VALU_TOPS,
                         "hbm_gbs": bt["bytes"] / (bt["ms"] / 1e3) / 1e9,
                         "hbm_frac": bt["bytes"] / (bt["ms"] / 1e3) / 1e9 / PEAK_HBM_GBS,
                         "traffic": measured_traffic("batched"),
                         "algorithmic_bytes": bt["bytes"]},
            "autotune": bt["tuned"],
        }
        if cpu_b:
            line["batched"]["cpu_baseline"] = {
                "value": cpu_b["value"], "unit": "nodes/s", "cores": 1, "kind": "port",
                "sample": f"first {cpu_b['sample']} nodes of the same batch, oracle LB2, 1 thread"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
