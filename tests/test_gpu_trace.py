"""Device search rounds checked one by one through the search trace
(wsbb_search_trace): every child's bound against the oracle, the parents of
every round against the oracle's pop sequence, and the counters against
node conservation."""
import numpy as np
import pytest

import oracle
import wsbb_amd
from tracecheck import rebuild_children, taken

pytestmark = pytest.mark.gpu


def traced_solve(p, pool, nb, kind=wsbb_amd.LB2, int16=None, **kw):
    with wsbb_amd.Engine(p) as eng:
        if int16 is not None:
            eng.set_int16(int16)
        rounds = eng.search_trace()
        res = eng.solve(kind=kind, pool_size=pool, max_batches=nb, **kw)
        eng.search_trace(False)
    return res, rounds


def test_ta011_config_every_child_bound_matches_oracle():
    """20x10 solved with pools of 4096 children, one pool per round (the
    reference's launch shape): every child of every round, both sides,
    bit-exact against the oracle's bornes_calculer, and the proven optimum
    and counters equal the oracle's batched explorer."""
    p = oracle.taillard(11)[3]
    n = p.shape[1]
    o = oracle.Oracle(p)
    best, st, _ = o.solve_batched(pool_size=4096, max_batches=1)
    res, rounds = traced_solve(p, 4096, 1)
    assert res["proven"] and res["best"] == best == 1582
    for k in ("decomposed", "bounded", "leaves", "pruned_pop", "pruned_insert",
              "improvements", "max_pool", "rounds"):
        assert res["stats"][k] == st[k], k
    total = 0
    for rnd in rounds:
        permT, l1, l2 = rebuild_children(rnd, n)
        ref = o.bound_batch(permT, l1, l2, threads=8)
        bad = np.flatnonzero(ref != rnd["bound"])
        assert bad.size == 0, (rnd["round"], bad[:5], ref[bad[:5]], rnd["bound"][bad[:5]])
        total += l1.shape[0]
    assert total == st["bounded"]


@pytest.mark.parametrize("tid,pool,nb,budget,int16", [(11, 4096, 10, 0, 2), (21, 65536, 4, 30_000, 2),
                                                       (51, 4096, 10, 6_000, 2),
                                                       (81, 4096, 4, 1_500, 2),
                                                       (81, 4096, 4, 1_500, 0)])
def test_two_node_search_rounds_match_oracle(tid, pool, nb, budget, int16):
    """Search rounds forced through the two-node int16 form (int16 = 2), or
    through the int32 kernel (0): every child of every round against the
    oracle's bornes_calculer, on a 20x10 solve to the optimum (ten pools per
    round), a budgeted 20x20 solve with the 2^18-child rounds that select the
    two-node form by default, and budgeted 50x20 and 100x20 solves, where
    each 128-child tile of the two-node form scans only the union of its
    children's free jobs (two and four mask words per node)."""
    p = oracle.taillard(tid)[3]
    n = p.shape[1]
    o = oracle.Oracle(p)
    best, st, _ = o.solve_batched(pool_size=pool, max_batches=nb, budget=budget)
    res, rounds = traced_solve(p, pool, nb, int16=int16, budget=budget)
    # no leaf reached before the budget: the oracle says -1, the device INT32_MAX
    assert res["best"] == (best if best >= 0 else 2**31 - 1)
    assert res["stats"]["bounded"] == st["bounded"]
    assert res["stats"]["decomposed"] == st["decomposed"]
    total = 0
    for rnd in rounds:
        permT, l1, l2 = rebuild_children(rnd, n)
        ref = o.bound_batch(permT, l1, l2, threads=8)
        bad = np.flatnonzero(ref != rnd["bound"])
        assert bad.size == 0, (rnd["round"], bad[:5], ref[bad[:5]], rnd["bound"][bad[:5]])
        total += l1.shape[0]
    assert total == st["bounded"]


@pytest.mark.parametrize("tid,budget", [(11, 0), (51, 3_000)])
def test_lb1_search_rounds_match_oracle(tid, budget):
    """Search rounds with the one-machine bound: every child of every round
    against the oracle's LB1, on a 20x10 solve to the optimum and a budgeted
    50x20 solve (two mask words per node), with the oracle explorer's
    counters."""
    p = oracle.taillard(tid)[3]
    n = p.shape[1]
    o = oracle.Oracle(p)
    best, st, _ = o.solve_batched(kind=oracle.LB1, pool_size=4096, max_batches=10, budget=budget)
    res, rounds = traced_solve(p, 4096, 10, kind=wsbb_amd.LB1, budget=budget)
    assert res["best"] == (best if best >= 0 else 2**31 - 1)
    for k in ("decomposed", "bounded"):
        assert res["stats"][k] == st[k], k
    total = 0
    for rnd in rounds:
        permT, l1, l2 = rebuild_children(rnd, n)
        ref = o.bound_batch(permT, l1, l2, kind=oracle.LB1, threads=8)
        bad = np.flatnonzero(ref != rnd["bound"])
        assert bad.size == 0, (rnd["round"], bad[:5], ref[bad[:5]], rnd["bound"][bad[:5]])
        total += l1.shape[0]
    assert total == st["bounded"]


@pytest.mark.parametrize("tid,pool,nb,budget", [(1, 4096, 1, 0), (3, 300, 3, 0), (11, 4096, 10, 0),
                                                (21, 65536, 8, 60_000), (51, 65536, 4, 8_000)])
def test_parents_follow_the_oracle_pop_sequence(tid, pool, nb, budget):
    """The parents each device round takes, in pop order, are exactly the
    oracle explorer's (whose order is pinned to the reference's rank-string
    priority queue by the oracle's own tests); each parent's children are
    2u (or 2 at depth n-2), BEGIN side first.  The large pools (budgeted)
    take the chained multi-block select."""
    p = oracle.taillard(tid)[3]
    n = p.shape[1]
    o = oracle.Oracle(p)
    _, st, _, tr = o.solve_batched(pool_size=pool, max_batches=nb, budget=budget, trace=2_000_000)
    res, rounds = traced_solve(p, pool, nb, budget=budget)
    # a budget stop leaves the last round's pops in the oracle's trace only
    tk = {k: v[:st["decomposed"]] for k, v in taken(tr).items()}
    dev_perm = np.concatenate([r["permT"].T for r in rounds])
    dev_l1 = np.concatenate([r["l1"] for r in rounds])
    dev_l2 = np.concatenate([r["l2"] for r in rounds])
    dev_cost = np.concatenate([r["cost"] for r in rounds])
    assert dev_l1.shape[0] == tk["l1"].shape[0] == st["decomposed"]
    assert np.array_equal(dev_perm, tk["perm"].astype(np.uint8))
    assert np.array_equal(dev_l1, tk["l1"]) and np.array_equal(dev_l2, tk["l2"])
    assert np.array_equal(dev_cost, tk["cost"])
    for r in rounds:
        u = r["l2"] - r["l1"] - 1
        per = np.where(u == 2, 2, 2 * u)
        par = r["desc"] >> 9
        assert np.array_equal(np.bincount(par, minlength=u.shape[0]), per)
        assert np.all(np.diff(par) >= 0)
    assert res["stats"]["bounded"] == sum(r["desc"].shape[0] for r in rounds)
    # the work statistic (numerator of the search bound's roofline) sums,
    # over exactly the decomposed parents, the algorithmic ops of their
    # children: P(n + 4u + 4) + 2m d1 + m(m+1) d2 per child with prefix d1,
    # suffix d2 and u free jobs
    m = p.shape[0]
    P, mm1 = m * (m - 1) // 2, m * (m + 1)
    l1, l2 = tk["l1"].astype(np.int64), tk["l2"].astype(np.int64)
    d1, d2, u = l1 + 1, n - l2, l2 - l1 - 1
    base = P * (n + 4 * (u - 1) + 4)
    w = np.where(d1 + d2 == n - 2, 2 * (P * (n + 4) + 2 * m * (d1 + 2) + mm1 * d2),
                 u * (base + 2 * m * (d1 + 1) + mm1 * d2) + u * (base + 2 * m * d1 + mm1 * (d2 + 1)))
    assert res["stats"]["work_ops"] == int(w.sum())


def test_tied_leaves_keep_the_first_found_on_device():
    """Instances with many optimal leaves: the device keeps the first
    optimal leaf met in the batched order, like the oracle (whose choice is
    pinned against its own leaf trace)."""
    rng = np.random.default_rng(11)
    insts = [np.full((3, 7), 4, np.int32)]
    for _ in range(6):
        n, m = int(rng.integers(6, 9)), int(rng.integers(2, 4))
        insts.append(rng.integers(1, 4, size=(m, n)).astype(np.int32))
    for p in insts:
        o = oracle.Oracle(p)
        best, _, bp = o.solve_batched(pool_size=64, max_batches=2)
        with wsbb_amd.Engine(p) as eng:
            res = eng.solve(pool_size=64, max_batches=2)
        assert res["proven"] and res["best"] == best
        assert np.array_equal(res["perm"], bp)


def test_pool_overflow_stops_cleanly_and_resumes_exactly():
    """A device pool too small for the search: the solve stops with
    WSBB_EOVERFLOW before the round that would not fit, leaving the pool,
    incumbent and counters of the round before; the saved state resumed with
    the default capacity reaches the oracle's optimum with exactly the
    uninterrupted search's counters.  The same context then solves again
    normally."""
    p = oracle.taillard(1)[3]
    best, st, _ = oracle.Oracle(p).solve_batched(pool_size=4096, max_batches=1)
    cap = 1024
    assert st["max_pool"] > cap
    with wsbb_amd.Engine(p) as eng:
        eng.set_pool_capacity(cap)
        with pytest.raises(wsbb_amd.WsbbError) as ei:
            eng.solve(pool_size=4096, max_batches=1)
        assert ei.value.code == -4 and "overflow" in str(ei.value)
        ck = eng.checkpoint()
        assert 0 < ck["permT"].shape[1] <= cap
        assert ck["stats"][6] <= cap  # max pool never went past the capacity
        eng.set_pool_capacity(1 << 20)
        again = eng.solve(pool_size=4096, max_batches=1)
        assert again["proven"] and again["best"] == best
    with wsbb_amd.Engine(p) as eng:
        res = eng.resume(ck, pool_size=4096, max_batches=1)
    assert res["proven"] and res["best"] == best == 1278
    for k in ("decomposed", "bounded", "leaves", "pruned_pop", "pruned_insert",
              "improvements", "max_pool", "rounds"):
        assert res["stats"][k] == st[k], k


def test_resume_rejects_nodes_the_device_never_makes():
    """wsbb_solve_resume refuses a pool node with fewer than two free jobs,
    a negative bound, a column that is not a permutation, and a bad
    incumbent permutation; a bound beyond any incumbent still resumes."""
    p = oracle.taillard(1)[3]
    n = p.shape[1]
    base = dict(inc=2**31 - 1, stats=np.zeros(wsbb_amd.NSTATS, np.int64),
                best_perm=np.full(n, -1, np.int32),
                permT=np.arange(n, dtype=np.uint8).reshape(n, 1).copy(),
                l1=np.array([-1], np.int32), l2=np.array([n], np.int32),
                cost=np.array([0], np.int32))
    bad_leaf = dict(base, l1=np.array([n - 2], np.int32), l2=np.array([n - 1], np.int32))
    bad_one = dict(base, l1=np.array([n - 3], np.int32), l2=np.array([n - 1], np.int32))
    dup = base["permT"].copy()
    dup[1, 0] = dup[0, 0]
    bad_perm = dict(base, permT=dup)
    bad_inc = dict(base, best_perm=np.full(n, 0, np.int32))
    bad_cost = dict(base, cost=np.array([-5], np.int32))
    with wsbb_amd.Engine(p) as eng:
        for ck in (bad_leaf, bad_one, bad_perm, bad_inc, bad_cost):
            with pytest.raises(wsbb_amd.WsbbError) as ei:
                eng.resume(ck, pool_size=4096)
            assert ei.value.code == -1
        res = eng.resume(base, pool_size=4096)
        assert res["proven"] and res["best"] == 1278
        # a root carrying a huge bound is still taken without an incumbent
        res = eng.resume(dict(base, cost=np.array([2**30], np.int32)), pool_size=4096)
    assert res["proven"] and res["best"] == 1278
