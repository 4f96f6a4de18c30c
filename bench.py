#!/usr/bin/env python3
"""Benchmark: LB2 bounded nodes/s and time-to-proven-optimum on one MI355X.

Headline (one step = one bound launch): Taillard ta051, 50 jobs x 20
machines, int32 times in [1, 99]; a batch of 2^20 synthetic nodes (prefix
and suffix lengths uniform so the scheduled depth is in [0, 49], random job
permutations) resident in HBM, bounded with plain LB2 by the int32 kernel
with autotuned launch shape.  `value` is nodes bounded per second over the K
timed launches; `roofline` relates the algorithmic integer work of the batch
to the kernel's launch time (HIP events on the context stream) and the int32
VALU peak.

Sub-objects: "int16_two_node" (the same batch by the two-node int16 kernel,
against the packed 16-bit issue rate), "config1" (ta001-ta010, 20x5, solved
with LB2 and LB1 on the GPU and by the CPU explorer), "config2" (ta011, 20x10,
time-to-optimum with 4096-child pools: ten per round, and one per round as
the reference launches them), "learned_20x20" and "learned_online" (20x20
plain against learned pair selection), "mixed_batch" (64 instances of uneven
size in one launch), "solve_50x20" (50x20 instances to proven optimum, plain
and learned, against the CPU explorer).  CPU baselines are the oracle (the C
restatement of the reference CPU path) on this host, on one thread and on all
host cores.
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


def host_info():
    """CPU of this host: logical CPUs, usable CPUs, model and clock."""
    info = {"nproc": os.cpu_count(), "usable": len(os.sched_getaffinity(0))}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and "model" not in info:
                    info["model"] = line.split(":", 1)[1].strip()
                if line.startswith("cpu MHz") and "mhz" not in info:
                    info["mhz"] = float(line.split(":", 1)[1])
    except OSError:
        pass
    # the GPU box allots 16 host threads per GPU
    info["threads"] = max(1, min(info["usable"], 16))
    return info


def bench_solve(wsbb_amd, p, reps, pool=4096, nb=10, kind=0, profile=True):
    """`reps` solves of instance p to proven optimum after one untimed one;
    optional bound-kernel timing on every 8th round."""
    eng = wsbb_amd.Engine(p)
    if profile:
        eng.search_profile(TIME_EVERY)
    res = eng.solve(kind=kind, pool_size=pool, max_batches=nb)
    eng.synchronize()
    tot = dict(bounded=0, work=0, bound_ms=0.0, launches=0, rounds=0)
    t0 = time.perf_counter()
    for _ in range(reps):
        res = eng.solve(kind=kind, pool_size=pool, max_batches=nb)
        assert res["proven"]
        if profile:
            ms, launches = eng.search_profile(-1)
            tot["bound_ms"] += ms
            tot["launches"] += launches
        tot["bounded"] += res["stats"]["bounded"]
        tot["work"] += res["stats"]["work_ops"]
        tot["rounds"] += res["stats"]["rounds"]
    eng.synchronize()
    elapsed = time.perf_counter() - t0
    eng.close()
    return dict(elapsed=elapsed, seconds=elapsed / reps, best=res["best"], stats=res["stats"],
                **tot)


def cpu_solves(o, kind, host):
    """The oracle's explorer on one thread and on the host's threads."""
    import oracle
    t0 = time.perf_counter()
    b1, s1, _ = o.solve_dfs(kind=kind)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    bn, sn, _ = o.solve_parallel(kind=kind, threads=host["threads"])
    tn = time.perf_counter() - t0
    assert b1 == bn
    return {"optimum": b1, "one_thread": {"seconds": t1, "bounded": s1["bounded"]},
            "all_threads": {"seconds": tn, "bounded": sn["bounded"], "threads": host["threads"]}}


def bench_config1(wsbb_amd, host, cpu=True):
    """ta001-ta010 (20x5) solved to proven optimum with LB2 and with LB1:
    the GPU search with 4096-child pools (one per round) against the CPU
    explorer (DFS, begin-end branching, UB = +inf)."""
    import oracle
    rows = []
    for tid in range(1, 11):
        p = wsbb_amd.taillard(tid)
        row = {"instance": f"ta{tid:03d}"}
        for name, kind, okind in (("lb2", wsbb_amd.LB2, oracle.LB2),
                                  ("lb1", wsbb_amd.LB1, oracle.LB1)):
            g = bench_solve(wsbb_amd, p, 3, pool=4096, nb=1, kind=kind, profile=False)
            ent = {"gpu": {"optimum": g["best"], "seconds": g["seconds"],
                           "bounded": g["stats"]["bounded"], "rounds": g["stats"]["rounds"]}}
            if cpu:
                c = cpu_solves(oracle.Oracle(p), okind, host)
                ent["cpu"] = c
                ent["optimum_match"] = c["optimum"] == g["best"]
            row[name] = ent
        rows.append(row)
    tot = {}
    for name in ("lb2", "lb1"):
        tot[name] = {"gpu_seconds": sum(r[name]["gpu"]["seconds"] for r in rows),
                     "gpu_bounded": sum(r[name]["gpu"]["bounded"] for r in rows)}
        if cpu:
            tot[name]["cpu_one_thread_seconds"] = sum(r[name]["cpu"]["one_thread"]["seconds"]
                                                      for r in rows)
            tot[name]["cpu_all_threads_seconds"] = sum(r[name]["cpu"]["all_threads"]["seconds"]
                                                       for r in rows)
            tot[name]["all_optima_match"] = all(r[name]["optimum_match"] for r in rows)
    return {"workload": "ta001-ta010 20x5 solved to proven optimum, LB2 and LB1, UB=+inf; "
                        "GPU: 4096-child pools, one per round; CPU: the oracle's DFS explorer",
            "totals": tot, "instances": rows}


def bench_config2(wsbb_amd, host, reps=10, cpu=True):
    """ta011 (20x10) to proven optimum: pools of 4096 children, ten per round
    (the reference's ring of ten pools in flight) and one per round (one
    pool per launch); roofline of the bound kernel inside the solve."""
    import oracle
    p = wsbb_amd.taillard(11)
    ten = bench_solve(wsbb_amd, p, reps, pool=4096, nb=10)
    one = bench_solve(wsbb_amd, p, max(1, reps // 2), pool=4096, nb=1)
    out = {"workload": "ta011 20x10 solved to proven optimum, plain LB2, 4096-child pools, "
                       "UB=+inf"}
    for name, sv in (("ten_pools_per_round", ten), ("one_pool_per_round", one)):
        avg_ms = sv["bound_ms"] / max(sv["launches"], 1)
        bound_s = avg_ms * sv["rounds"] / 1e3
        ach = sv["work"] / bound_s / 1e12 if bound_s > 0 else None
        out[name] = {
            "time_to_optimum_s": sv["seconds"], "optimum": sv["best"],
            "nodes_per_s": sv["bounded"] / sv["elapsed"],
            "nodes_bounded_per_solve": sv["stats"]["bounded"],
            "rounds_per_solve": sv["stats"]["rounds"],
            "bound_kernel": {"avg_launch_us": avg_ms * 1e3, "share_of_solve": bound_s / sv["elapsed"],
                             "achieved": ach, "peak": PEAK_VALU_TOPS, "unit": "Tint-op/s",
                             "frac": ach / PEAK_VALU_TOPS if ach else None,
                             "timed_launches": sv["launches"]},
        }
    if cpu:
        c = cpu_solves(oracle.Oracle(p), oracle.LB2, host)
        c["note"] = ("the CPU explorer pops one node at a time (DFS) and bounds fewer "
                     "children than the batched GPU order")
        out["cpu"] = c
        out["optimum_match"] = c["optimum"] == ten["best"] == one["best"]
    return out


def bench_batched(wsbb_amd, steps=20, warmup=2, int16=False, autotune=True):
    """2^20 ta051 nodes bounded `steps` times by one launch each, after
    `warmup` untimed launches.  Returns the wall time and the HIP-event time
    (context stream) per launch and the batch's algorithmic work."""
    p = wsbb_amd.taillard(51)
    m, n = p.shape
    N = 1 << 20
    permT, l1, l2 = wsbb_amd.random_nodes(n, N, 0x5EED)
    eng = wsbb_amd.Engine(p)
    eng.set_int16(2 if int16 else 0)
    tuned = eng.autotune(N, reps=3) if autotune else []
    dp, d1, d2, do = eng.malloc(permT.nbytes), eng.malloc(4 * N), eng.malloc(4 * N), eng.malloc(4 * N)
    eng.h2d(dp, permT)
    eng.h2d(d1, l1)
    eng.h2d(d2, l2)
    for _ in range(max(warmup, 1)):
        eng.bound_device(dp, d1, d2, N, do)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.timer_start()
    for _ in range(steps):
        eng.bound_device(dp, d1, d2, N, do)
    ev_ms = eng.timer_stop() / steps
    eng.synchronize()
    wall_ms = (time.perf_counter() - t0) / steps * 1e3
    out = np.zeros(N, np.int32)
    eng.d2h(out, do)
    for d in (dp, d1, d2, do):
        eng.free(d)
    eng.close()
    work = int(node_work(n, m, l1, l2).sum())
    bytes_alg = N * (n + 12)  # uint8 permutation + two int32 limits in, int32 bound out
    return dict(N=N, ms=ev_ms, wall_ms=wall_ms, work=work, bytes=bytes_alg, tuned=tuned,
                permT=permT, l1=l1, l2=l2, p=p, out=out)


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
    pool, nb, topk = 65536, 8, 16  # the fastest plain shape measured on ta021-ta024
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
        "workload": "ta021 20x20 solved to proven optimum, 65536-child pools, 8 sub-batches",
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


def online_run(wsbb_amd, eng, pool, nb, warm, cap, topk):
    return online_run_from(wsbb_amd, eng, pool, nb, warm, cap, topk, dive_first=False)


def online_run_from(wsbb_amd, eng, pool, nb, warm, cap, topk, dive_first):
    """One online-learning solve on a context: a plain solve that logs
    records stops after `warm` bounded children, the search is checkpointed,
    the selector is trained from scratch on the records and the search
    resumes with learned selection.  Returns (result, seconds, training
    seconds, bounded children)."""
    eng.selector_init(seed=1)
    eng.selector_config(topk=topk, margin=-1, parents=True)
    eng.selector_log(stride=max(1, warm * eng.n // cap), capacity=cap)
    t = time.perf_counter()
    if dive_first:
        # a narrow dive for the first incumbent, then the warm-up from it
        # (budgets count decomposed parents: about 40 children each here)
        dive = eng.solve(pool_size=256, max_batches=1, budget=2000)
        first = eng.solve(pool_size=pool, max_batches=nb, ub0=dive["best"], budget=warm // 40)
    else:
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
        res = eng.resume(ck, kind=wsbb_amd.LEARNED, pool_size=pool, max_batches=nb,
                         budget=20_000_000 if dive_first else 0)
        bounded = res["stats"]["bounded"]
    return res, time.perf_counter() - t, train_s, bounded


def bench_learned_online(wsbb_amd, ids=tuple(range(21, 31)), runs=3, warm=2_000_000, topk=24):
    """Online learning on one instance, all costs counted (online_run), set
    against a plain solve of the same instance from scratch; `runs` of each
    per instance, alternating, with mean and spread of the times."""
    pool, nb, cap = 65536, 8, 1 << 17  # the fastest plain shape measured on ta021-ta024
    out = []
    for tid in ids:
        plain_t, online_t, train_t = [], [], []
        with wsbb_amd.Engine(wsbb_amd.taillard(tid)) as eng:
            eng.solve(pool_size=pool, max_batches=nb, budget=20_000)  # warm up
            for _ in range(runs):
                t = time.perf_counter()
                plain = eng.solve(pool_size=pool, max_batches=nb)
                plain_t.append(time.perf_counter() - t)
                res, sec, tr, bounded = online_run(wsbb_amd, eng, pool, nb, warm, cap, topk)
                online_t.append(sec)
                train_t.append(tr)
                assert res["proven"] and res["best"] == plain["best"]
        pm, om = float(np.mean(plain_t)), float(np.mean(online_t))
        out.append({
            "instance": f"ta{tid:03d}", "optimum": plain["best"],
            "plain_s": {"mean": pm, "min": min(plain_t), "max": max(plain_t)},
            "online_s": {"mean": om, "min": min(online_t), "max": max(online_t)},
            "train_s_mean": float(np.mean(train_t)),
            "plain_bounded": plain["stats"]["bounded"], "online_bounded": bounded,
            "speedup_mean": pm / om,
            "speedup_range": [min(plain_t) / max(online_t), max(plain_t) / min(online_t)],
            "node_inflation": bounded / plain["stats"]["bounded"],
        })
    sp = [r["speedup_mean"] for r in out]
    return {
        "protocol": f"plain solve with record logging for the first {warm} bounded children, "
                    "checkpoint, selector trained from scratch on the GPU (2 epochs, at most "
                    f"{cap} records, batch 2048), learned resume (top-{topk} pairs per 64-child "
                    "tile, no fallback) to the proven optimum; all of it timed, against a plain "
                    f"solve from scratch; {runs} runs of each per instance",
        "geomean_speedup": float(np.exp(np.mean(np.log(sp)))),
        "instances_faster": int(sum(x > 1.0 for x in sp)),
        "instances": out,
    }


def bench_50x20(wsbb_amd, host, cpu=True, seeds=(1, 11, 25, 27, 38, 49, 51, 65, 67, 69, 78)):
    """Machine-correlated 50x20 instances (p[k][j] = b_k + U[0,20], b_k ~
    U[1,80]; the seeds among 1-80 whose search does not end at the root)
    solved to proven optimum: a narrow dive (256-child pools) for the first
    incumbent, then plain LB2 with 4096-child pools, four per round, and
    with 65536-child pools, four per round; learned selection online from
    the same start at the large shape (warm-up of 300000 children, top-24,
    no fallback, 2e7-parent budget); the CPU explorer on the host's
    threads.  Job-correlated and Taillard 50x20 instances do not close within
    the budget."""
    import oracle
    rows = []
    for seed in seeds:
        p = wsbb_amd.generate_machine_correlated(50, 20, seed)
        row = {"seed": seed}
        with wsbb_amd.Engine(p) as eng:
            eng.solve(pool_size=256, max_batches=1, budget=2000)  # warm up
            for name, pool, nb in (("plain_4096x4", 4096, 4), ("plain_65536x4", 65536, 4)):
                t = time.perf_counter()
                dive = eng.solve(pool_size=256, max_batches=1, budget=2000)
                r = eng.solve(pool_size=pool, max_batches=nb, ub0=dive["best"], budget=20_000_000)
                row[name] = {"seconds": time.perf_counter() - t, "optimum": r["best"],
                             "proven": r["proven"], "bounded": r["stats"]["bounded"]}
            t = time.perf_counter()
            res, _, train_s, bounded = online_run_from(wsbb_amd, eng, 65536, 4, 300_000,
                                                       1 << 17, 24, dive_first=True)
            row["learned_online_65536x4"] = {"seconds": time.perf_counter() - t,
                                             "optimum": res["best"], "proven": res["proven"],
                                             "bounded": bounded, "train_seconds": train_s}
        if cpu:
            t = time.perf_counter()
            b, st, _ = oracle.Oracle(p).solve_parallel(threads=host["threads"])
            row["cpu_all_threads"] = {"seconds": time.perf_counter() - t, "optimum": b,
                                      "bounded": st["bounded"], "threads": host["threads"]}
            row["optimum_match"] = all(row[k]["optimum"] == b for k in
                                       ("plain_4096x4", "plain_65536x4") if row[k]["proven"]) and \
                (not row["learned_online_65536x4"]["proven"] or
                 row["learned_online_65536x4"]["optimum"] == b)
        rows.append(row)
    summ = {"instances": len(rows),
            "plain_proven": sum(r["plain_4096x4"]["proven"] for r in rows),
            "learned_proven": sum(r["learned_online_65536x4"]["proven"] for r in rows),
            "learned_faster_than_plain_65536x4": sum(
                r["learned_online_65536x4"]["proven"] and
                r["learned_online_65536x4"]["seconds"] < r["plain_65536x4"]["seconds"]
                for r in rows),
            "learned_faster_than_plain_4096x4": sum(
                r["learned_online_65536x4"]["proven"] and
                r["learned_online_65536x4"]["seconds"] < r["plain_4096x4"]["seconds"]
                for r in rows),
            "gpu_plain_4096x4_seconds": sum(r["plain_4096x4"]["seconds"] for r in rows)}
    if cpu:
        summ["cpu_all_threads_seconds"] = sum(r["cpu_all_threads"]["seconds"] for r in rows)
        summ["all_optima_match"] = all(r["optimum_match"] for r in rows)
    return {"workload": "machine-correlated 50x20 instances solved to proven optimum",
            "summary": summ, "instances": rows}


def bench_mixed(wsbb_amd, budget=100_000_000, host=None):
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
    cpu = None
    if host is not None:
        import oracle
        # the first 256 nodes of every instance, once on one thread and once
        # (the first 2048) on the host's threads
        t1 = tn = 0.0
        for i, p in enumerate(ps):
            o = oracle.Oracle(p)
            n = p.shape[1]
            sl = slice(i * per, i * per + 256)
            t = time.perf_counter()
            ref = o.bound_batch(permT[:n, sl], l1[sl], l2[sl], threads=1)
            t1 += time.perf_counter() - t
            assert np.array_equal(ref, full[sl])
            sl = slice(i * per, i * per + 2048)
            t = time.perf_counter()
            o.bound_batch(permT[:n, sl], l1[sl], l2[sl], threads=host["threads"])
            tn += time.perf_counter() - t
        cpu = {"one_thread": {"value": 64 * 256 / t1, "unit": "nodes/s", "cores": 1,
                              "sample": "first 256 nodes of each instance, oracle LB2"},
               "all_threads": {"value": 64 * 2048 / tn, "unit": "nodes/s",
                               "cores": host["threads"],
                               "sample": "first 2048 nodes of each instance, oracle LB2"},
               "kind": "port", "parity": "GPU plain bounds equal the oracle on the sample"}
    return {
        "cpu_baseline": cpu,
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


def cpu_batched(bt, host):
    """The oracle's LB2 on the head of the same batch: 2^17 nodes on one
    thread, 2^19 on the host's threads; the GPU's values are checked on the
    one-thread sample."""
    import oracle
    ob = oracle.Oracle(bt["p"])
    S1, S2 = 1 << 17, 1 << 19
    t0 = time.perf_counter()
    ref = ob.bound_batch(bt["permT"][:, :S1], bt["l1"][:S1], bt["l2"][:S1], threads=1)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    ob.bound_batch(bt["permT"][:, :S2], bt["l1"][:S2], bt["l2"][:S2], threads=host["threads"])
    tn = time.perf_counter() - t0
    return {"value": S2 / tn, "unit": "nodes/s", "cores": host["threads"], "kind": "port",
            "sample": f"first {S2} nodes of the same batch, oracle LB2, {host['threads']} threads "
                      f"({tn:.2f} s)",
            "one_thread": {"value": S1 / t1, "unit": "nodes/s", "cores": 1,
                           "sample": f"first {S1} nodes, 1 thread ({t1:.2f} s)"},
            "host": {k: host.get(k) for k in ("model", "mhz", "nproc", "usable")},
            "parity": bool(np.array_equal(ref, bt["out"][:S1]))}


def valu_peak(wsbb_amd, iters=20000):
    """Measured integer VALU throughput of the probe kernel (eight waves per
    SIMD, independent chains of one instruction), T lane-ops/s per kind."""
    out = {}
    with wsbb_amd.Engine(wsbb_amd.taillard(1)) as eng:
        nbytes = eng.num_cus * 8 * 256 * 4
        d = eng.malloc(nbytes)
        for kind, name in ((0, "v_add_u32"), (1, "v_mad_u32_u24"), (2, "v_max_i32")):
            eng.valu_probe(kind, 16, d, nbytes)
            eng.synchronize()
            eng.timer_start()
            ops = eng.valu_probe(kind, iters, d, nbytes)
            ms = eng.timer_stop()
            out[name] = ops / (ms / 1e3) / 1e12
        eng.free(d)
    return out


def roofline(bt, peak, note):
    ach = bt["work"] / (bt["ms"] / 1e3) / 1e12
    return {"bound": "valu", "kernel": note, "achieved": ach, "peak": peak, "unit": "Tint-op/s",
            "frac": ach / peak, "avg_launch_us": bt["ms"] * 1e3,
            "hbm_gbs": bt["bytes"] / (bt["ms"] / 1e3) / 1e9,
            "hbm_frac": bt["bytes"] / (bt["ms"] / 1e3) / 1e9 / PEAK_HBM_GBS,
            "algorithmic_bytes": bt["bytes"], "algorithmic_ops": bt["work"]}


_T0 = time.perf_counter()


def progress(what):
    """One line on stderr per finished section (the JSON line stays alone on stdout)."""
    print(f"bench: {what} done at {time.perf_counter() - _T0:.1f} s", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("all", "batched", "solve", "copy", "learned",
                                           "learned20", "50x20", "peak"),
                    default="all",
                    help="batched = the 50x20 launches only (--int16: two-node form); solve = "
                         "ta011 time-to-optimum only; learned = the 20x20 learned comparison "
                         "and the online protocol; learned20 = the comparison only")
    ap.add_argument("--int16", action="store_true", help="batched: the two-node int16 kernel")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the 20x20 learned-selector and mixed-batch measurements")
    ap.add_argument("--online-runs", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs on a single MI355X "
              "(sharding across GPUs is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import wsbb_amd
    host = host_info()
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
                          "learned_online": bench_learned_online(wsbb_amd,
                                                                 runs=args.online_runs)}),
              flush=True)
        return
    if args.workload == "peak":
        print(json.dumps({"workload": "peak", "valu_tops": valu_peak(wsbb_amd)}), flush=True)
        return
    if args.workload == "learned20":
        print(json.dumps({"learned_20x20": bench_learned(wsbb_amd)}), flush=True)
        return
    if args.workload == "50x20":
        print(json.dumps(bench_50x20(wsbb_amd, host, cpu=not args.no_cpu)), flush=True)
        return
    if args.workload == "solve":
        print(json.dumps(bench_config2(wsbb_amd, host, reps=args.steps, cpu=not args.no_cpu)),
              flush=True)
        return
    bt = bench_batched(wsbb_amd, args.steps, args.warmup, int16=args.int16 and
                       args.workload == "batched", autotune=not args.no_autotune)
    if args.workload == "batched":
        print(json.dumps({"workload": "batched", "int16": args.int16, "ms_per_launch": bt["ms"],
                          "nodes_per_s": bt["N"] / (bt["ms"] / 1e3)}), flush=True)
        return

    rf = roofline(bt, PEAK_VALU_TOPS, "bound_kernel<20, 2, 0, 256, BatchSource> (int32)")
    rf["traffic"] = measured_traffic("batched_int32")
    # the derived peak against the probe kernel's measured rates on this GPU
    rf["peak_check"] = {"measured_tlaneops": valu_peak(wsbb_amd),
                        "derived": "256 CU x 64 lanes x 2.4 GHz = 39.3 T int32 lane-ops/s",
                        "scan_instructions": "v_bfe_u32, v_mad_i32_i24, v_mad_u32_u24, v_max_i32"}
    line = {
        "metric": "LB2 bounded nodes/s + time-to-proven-optimum, 1 MI355X vs host CPU",
        "value": bt["N"] / (bt["wall_ms"] / 1e3),
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bt["wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (Taillard generator with the reference's seeds; splitmix64 nodes)",
        "config": {"workload": "ta051 50x20 batched LB2 bounding: 2^20 nodes per launch, "
                               "d1~U{0..49}, d2~U{0..49-d1}, random permutations, resident "
                               "in HBM; int32 kernel, autotuned launch shape",
                   "parallelism": "single GPU"},
        "roofline": rf,
        "autotune": bt["tuned"],
    }
    progress("ta051 batched int32")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_batched(bt, host)
        progress("CPU baseline sample")
    b16 = bench_batched(wsbb_amd, args.steps, args.warmup, int16=True,
                        autotune=not args.no_autotune)
    r16 = roofline(b16, 2 * PEAK_VALU_TOPS, "bound_kernel_pk<20, 2, BatchSource> (two nodes "
                   "per lane in int16 halves; peak = packed 16-bit issue rate, 2 x int32)")
    r16["traffic"] = measured_traffic("batched_int16")
    line["int16_two_node"] = {
        "dtype": "int16x2", "nodes_per_s": b16["N"] / (b16["wall_ms"] / 1e3),
        "ms_per_launch": b16["wall_ms"], "roofline": r16,
        "same_values_as_int32": bool(np.array_equal(b16["out"], bt["out"])),
        "applies": "instances whose path bound stays below 2^14 (every Taillard instance)"}
    progress("ta051 batched int16")
    line["config1"] = bench_config1(wsbb_amd, host, cpu=not args.no_cpu)
    progress("ta001-ta010 solves")
    line["config2"] = bench_config2(wsbb_amd, host, cpu=not args.no_cpu)
    progress("ta011 solves")
    if not args.no_learned:
        line["learned_20x20"] = bench_learned(wsbb_amd)
        progress("ta021 learned comparison")
        line["learned_online"] = bench_learned_online(wsbb_amd, runs=args.online_runs)
        progress("ta021-ta030 online learning")
        line["mixed_batch"] = bench_mixed(wsbb_amd, host=None if args.no_cpu else host)
        progress("mixed batch")
        line["solve_50x20"] = bench_50x20(wsbb_amd, host, cpu=not args.no_cpu)
        progress("50x20 solves")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
