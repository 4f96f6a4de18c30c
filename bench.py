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
launch, inputs resident in HBM; ("learned_20x20"): plain against learned
pair selection on ta021 solved to optimum, the selector trained on a
plain solve of the same instance; ("learned_online"): ta021-ta024, the
selector trained during the search on records of its first 2e6 bounded
children, every cost counted; ("mixed_batch"): 64 instances of
uneven size bounded together, plain and learned, over 1e8 nodes.

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


TIME_EVERY = 8  # the bound launch of every 8th round carries HIP events


def bench_solve(wsbb_amd, steps, warmup):
    p = wsbb_amd.taillard(11)
    eng = wsbb_amd.Engine(p)
    eng.search_profile(TIME_EVERY)
    for _ in range(warmup):
        eng.solve(pool_size=4096, max_batches=10)
    eng.synchronize()
    tot = dict(bounded=0, work=0, bound_ms=0.0, launches=0, rounds=0)
    best = None
    t0 = time.perf_counter()
    for _ in range(steps):
        res = eng.solve(pool_size=4096, max_batches=10)
        ms, launches = eng.search_profile(-1)
        tot["bounded"] += res["stats"]["bounded"]
        tot["work"] += res["stats"]["work_ops"]
        tot["bound_ms"] += ms
        tot["launches"] += launches
        tot["rounds"] += res["stats"]["rounds"]
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


def children_of(permT, l1, l2):
    """Every child of the given nodes in the search's order (per parent the
    BEGIN children, then the END children; decompose_beginend semantics):
    (child permT, l1, l2, children per parent)."""
    n, N = permT.shape
    cols, a, b, counts = [], [], [], []
    for k in range(N):
        u = l2[k] - l1[k] - 1
        cnt = 0
        for side in (0, 1):
            pos = l1[k] + 1 if side == 0 else l2[k] - 1
            for j in range(l1[k] + 1, l2[k]):
                c = permT[:, k].copy()
                c[pos], c[j] = c[j], c[pos]
                ca, cb = l1[k] + (side == 0), l2[k] - (side == 1)
                if cb - ca == 2:
                    ca += 1
                cols.append(c)
                a.append(ca)
                b.append(cb)
                cnt += 1
        counts.append(cnt)
        assert u >= 3
    return (np.stack(cols, axis=1), np.array(a, np.int32), np.array(b, np.int32),
            np.array(counts, np.int32))


def bench_learned(wsbb_amd):
    """20x20 (ta021): time-to-proven-optimum with plain LB2 against learned
    pair selection, no fallback.  The MLP scores the round's parents (the
    pairs that maximise LB2 over a parent's children); each 64-child tile
    bounds over the top-16 pairs of its parents' summed probabilities.
    Trained on the GPU on (parent features, child's LB2-argmax pair) records
    logged by a plain solve of the same instance.  Hit rate: share of the
    children of 4096 random parents whose tile-subset bound equals full LB2."""
    p = wsbb_amd.taillard(21)
    pool, nb, topk = 65536, 4, 16
    with wsbb_amd.Engine(p) as eng:
        eng.selector_init(seed=1)
        eng.selector_config(topk=topk, margin=-1, parents=True)
        eng.selector_log(stride=64, capacity=1 << 19)
        eng.solve(pool_size=pool, max_batches=nb)
        feat, label = eng.selector_records()
        eng.selector_log(stride=0, capacity=0)
        eng.selector_set_records(feat, label)
        t = time.perf_counter()
        losses = eng.selector_train(epochs=4, batch=1024, lr=2e-3, seed=2)
        train_s = time.perf_counter() - t
        eng.solve(pool_size=pool, max_batches=nb, budget=20_000)  # warm both paths
        t = time.perf_counter()
        plain = eng.solve(pool_size=pool, max_batches=nb)
        plain_s = time.perf_counter() - t
        eng.solve(kind=wsbb_amd.LEARNED, pool_size=pool, max_batches=nb, budget=20_000)
        t = time.perf_counter()
        learned = eng.solve(kind=wsbb_amd.LEARNED, pool_size=pool, max_batches=nb)
        learned_s = time.perf_counter() - t
        permT, l1, l2 = wsbb_amd.random_nodes(20, 4096, 0xC0FFEE)
        keep = (l2 - l1 - 1) >= 3
        permT, l1, l2 = permT[:, keep], l1[keep], l2[keep]
        cpermT, cl1, cl2, counts = children_of(permT, l1, l2)
        pfeat = eng.bound(permT, l1, l2, kind=wsbb_amd.LB1, feat=True)["feat"]
        _, tsel = eng.selector_infer_parents(pfeat, counts)
        full = eng.bound(cpermT, cl1, cl2)
        sub, _ = eng.bound_tiles(cpermT, cl1, cl2, tsel, margin=-1)
    ps, ls = plain["stats"], learned["stats"]
    return {
        "workload": "ta021 20x20 solved to proven optimum, 65536-child pools, 4 sub-batches",
        "selector": f"MLP 100-128-190 on each round's parents, BF16 weights / FP32 "
                    f"accumulation, top-{topk} pairs per 64-child tile from the parents' "
                    "probabilities weighted by their children in the tile, trained on records "
                    f"logged by a plain solve of ta021 ({feat.shape[1]} records, 4 epochs, "
                    f"{train_s:.2f} s on the GPU)",
        "train_loss": [float(x) for x in losses],
        "plain": {"optimum": plain["best"], "proven": plain["proven"], "seconds": plain_s,
                  "bounded": ps["bounded"], "pairs_evaluated": ps["pairs_evaluated"]},
        "learned": {"optimum": learned["best"], "proven": learned["proven"], "seconds": learned_s,
                    "bounded": ls["bounded"], "pairs_evaluated": ls["pairs_evaluated"]},
        "speedup_time_to_optimum": plain_s / learned_s,
        "node_inflation": ls["bounded"] / ps["bounded"],
        "hit_rate_random_children": float((sub == full).mean()),
        "hit_rate_sample": f"{int(counts.sum())} children of {len(counts)} random parents",
    }


def bench_learned_online(wsbb_amd, ids=(21, 22, 23, 24), warm=2_000_000, topk=24):
    """Online learning on one instance, all costs counted: a plain solve that
    logs (parent features, argmax pair) records stops after `warm` bounded
    children, the search is checkpointed, the selector is trained from
    scratch on those records (2 epochs, 2^17 records at most), and the search
    resumes with learned pair selection (top-`topk` pairs per tile, no
    fallback) to the proven optimum.  The time of the whole sequence is set
    against a plain solve of the same instance from scratch."""
    pool, nb, cap = 65536, 4, 1 << 17
    out = []
    for tid in ids:
        with wsbb_amd.Engine(wsbb_amd.taillard(tid)) as eng:
            eng.solve(pool_size=pool, max_batches=nb, budget=20_000)  # warm up
            t = time.perf_counter()
            plain = eng.solve(pool_size=pool, max_batches=nb)
            plain_s = time.perf_counter() - t
            eng.selector_init(seed=1)
            eng.selector_config(topk=topk, margin=-1, parents=True)
            eng.selector_log(stride=max(1, warm * eng.n // cap), capacity=cap)
            t = time.perf_counter()
            first = eng.solve(pool_size=pool, max_batches=nb, budget=warm)
            bounded = first["stats"]["bounded"]
            train_s = 0.0
            if first["proven"]:
                res = first
            else:
                ck = eng.checkpoint()
                t2 = time.perf_counter()
                eng.selector_train(epochs=2, batch=2048, lr=2e-3, seed=2)
                train_s = time.perf_counter() - t2
                eng.selector_log(stride=0, capacity=0)
                res = eng.resume(ck, kind=wsbb_amd.LEARNED, pool_size=pool, max_batches=nb)
                bounded = res["stats"]["bounded"]
            online_s = time.perf_counter() - t
        out.append({
            "instance": f"ta{tid:03d}",
            "plain": {"optimum": plain["best"], "proven": plain["proven"], "seconds": plain_s,
                      "bounded": plain["stats"]["bounded"]},
            "online": {"optimum": res["best"], "proven": res["proven"], "seconds": online_s,
                       "train_seconds": train_s, "bounded": bounded},
            "speedup_time_to_optimum": plain_s / online_s,
            "node_inflation": bounded / plain["stats"]["bounded"],
        })
    return {
        "protocol": f"plain solve with record logging for the first {warm} bounded children, "
                    "checkpoint, selector trained from scratch on the GPU (2 epochs, at most "
                    f"{cap} records, batch 2048), learned resume (top-{topk} pairs per 64-child "
                    "tile, no fallback) to the proven optimum; all of it timed",
        "instances": out,
    }


def bench_mixed(wsbb_amd, budget=100_000_000):
    """64 instances, n in {20, 50, 100} x m in {5, 10, 20}, job-correlated
    times, 16384 random nodes each, bounded together in one launch per pass
    until `budget` nodes are bounded: plain LB2 and learned pair selection
    (one MLP per machine count, trained on records of a plain solve of the
    first instance of that size, top-k = max(4, P/8) pairs per tile) with
    full LB2 for nodes within 30 of their instance's incumbent."""
    import itertools
    shapes = list(itertools.product((20, 50, 100), (5, 10, 20)))
    ps = [wsbb_amd.generate_correlated(*shapes[i % len(shapes)], 0xA11CE + i) for i in range(64)]
    per = 16384
    maxn = max(p.shape[1] for p in ps)
    N = per * len(ps)
    permT = np.zeros((maxn, N), np.uint8)
    l1 = np.zeros(N, np.int32)
    l2 = np.zeros(N, np.int32)
    offsets = np.arange(len(ps) + 1, dtype=np.int32) * per
    for i, p in enumerate(ps):
        pt, a, b = wsbb_amd.random_nodes(p.shape[1], per, 0x5EED + 977 * i)
        permT[:p.shape[1], i * per:(i + 1) * per] = pt
        l1[i * per:(i + 1) * per] = a
        l2[i * per:(i + 1) * per] = b
    with wsbb_amd.Multi(ps) as mb:
        inc = np.zeros(len(ps), np.int32)
        seen = {}
        t = time.perf_counter()
        for i, p in enumerate(ps):
            eng = mb.engines[i]
            m = p.shape[0]
            logging = m not in seen
            if logging:
                eng.selector_init(seed=3)
                eng.selector_log(stride=16, capacity=1 << 17)
            eng.set_pool_capacity(1 << 20)  # at most 5000 parents x 100 children
            # incumbent: a dive with one parent per round reaches leaves after
            # about n rounds
            inc[i] = eng.solve(pool_size=256, max_batches=1, budget=400)["best"]
            if logging:
                eng.solve(pool_size=16384, max_batches=2, budget=5_000)
                f, lab = eng.selector_records()
                eng.selector_log(stride=0, capacity=0)
                eng.selector_set_records(f, lab)
                eng.selector_train(epochs=3, batch=1024, lr=2e-3, seed=2)
                P = m * (m - 1) // 2
                eng.selector_config(topk=max(4, P // 8), margin=30)
                seen[m] = i
        prep_s = time.perf_counter() - t
        mb.load(permT, l1, l2, offsets)
        reps = -(-budget // N)
        full, st_p, ms_p = mb.run(kind=wsbb_amd.LB2, reps=reps)
        sub, st_l, ms_l = mb.run(kind=wsbb_amd.LEARNED, inc=inc, margin=30, reps=reps)
    return {
        "workload": f"64 instances (n x m over {{20,50,100}} x {{5,10,20}}, p_ij = a_j + U[0,20], "
                    f"a_j ~ U[1,80]), {per} random nodes each, {N} per launch, {reps} passes "
                    f"= {reps * N} nodes",
        "plain": {"nodes_per_s": N / (ms_p / 1e3), "ms_per_pass": ms_p,
                  "pairs_per_node": st_p["pairs"] / max(st_p["nodes"], 1)},
        "learned": {"nodes_per_s": N / (ms_l / 1e3), "ms_per_pass": ms_l,
                    "pairs_per_node": st_l["pairs"] / max(st_l["nodes"], 1),
                    "fallback_share": st_l["fallbacks"] / max(st_l["nodes"], 1),
                    "valid_bound": bool((sub <= full).all()),
                    "hit_rate": float((sub == full).mean())},
        "incumbents": "best leaf of a 400-parent depth-first dive per instance",
        "incumbent_found": int((inc < 2**31 - 1).sum()),
        "setup_s": prep_s,
    }


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
    cores = min(os.cpu_count() or 1, 16)
    S2 = 1 << 17
    t0 = time.perf_counter()
    ob.bound_batch(bt["permT"][:, :S2], bt["l1"][:S2], bt["l2"][:S2], threads=cores)
    dtc = time.perf_counter() - t0
    return solve, dict(value=S / dtb, seconds=dtb, sample=S, ref=ref,
                       all_cores=dict(value=S2 / dtc, cores=cores, sample=S2))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("all", "solve", "batched", "copy", "learned"),
                    default="all",
                    help="solve = ta011 time-to-optimum only; batched = 50x20 launches only; "
                         "learned = the ta021 plain-against-learned comparison only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the 20x20 learned-selector and mixed-batch measurements")
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
    if args.workload == "learned":
        print(json.dumps({"learned_20x20": bench_learned(wsbb_amd),
                          "learned_online": bench_learned_online(wsbb_amd)}), flush=True)
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
    # average bound-launch time from the timed rounds, times every round
    avg_ms = sv["bound_ms"] / max(sv["launches"], 1)
    bound_s = avg_ms * sv["rounds"] / 1e3
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
                     "avg_launch_us": avg_ms * 1e3,
                     "launches": sv["rounds"],
                     "timed_launches": sv["launches"],
                     "timing": f"HIP events on the bound launch of every {TIME_EVERY}th round "
                               "inside the timed solves (the launch records them itself)",
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
            ac = cpu_b["all_cores"]
            line["batched"]["cpu_baseline_all_cores"] = {
                "value": ac["value"], "unit": "nodes/s", "cores": ac["cores"], "kind": "port",
                "sample": f"first {ac['sample']} nodes of the same batch, oracle LB2, "
                          f"{ac['cores']} threads"}
    if args.workload == "all" and not args.no_learned:
        line["learned_20x20"] = bench_learned(wsbb_amd)
        line["learned_online"] = bench_learned_online(wsbb_amd)
        line["mixed_batch"] = bench_mixed(wsbb_amd)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
