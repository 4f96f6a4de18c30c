"""Device-resident branch-and-bound against the oracle's search loops.

The device engine follows the reference's batched GPU path, so besides the
proven optimum every counter (parents decomposed, children bounded, leaves,
prunes, incumbent updates, pool peak, rounds) must equal the oracle's
batched explorer on the same instance and pool shape."""
import numpy as np
import pytest

import oracle
import wsbb_amd

pytestmark = pytest.mark.gpu

STATS = ("decomposed", "bounded", "leaves", "pruned_pop", "pruned_insert", "improvements",
         "max_pool", "rounds")


def cmax_np(p, perm):
    t = np.zeros(p.shape[0], np.int64)
    for j in perm:
        t[0] += p[0, j]
        for k in range(1, p.shape[0]):
            t[k] = max(t[k], t[k - 1]) + p[k, j]
    return int(t[-1])


def check_solve(p, kind, pool, nb, ref_best=None):
    o = oracle.Oracle(p)
    okind = oracle.LB2 if kind == wsbb_amd.LB2 else oracle.LB1
    best, st, _ = o.solve_batched(kind=okind, pool_size=pool, max_batches=nb)
    with wsbb_amd.Engine(p) as eng:
        res = eng.solve(kind=kind, pool_size=pool, max_batches=nb)
    assert res["proven"]
    assert res["best"] == best
    if ref_best is not None:
        assert best == ref_best
    assert cmax_np(p, res["perm"]) == res["best"]
    assert sorted(res["perm"].tolist()) == list(range(p.shape[1]))
    for k in STATS:
        assert res["stats"][k] == st[k], (k, res["stats"][k], st[k])
    return res


@pytest.mark.parametrize("tid", range(1, 11))
def test_20x5_lb2_pool4096(tid):
    p = oracle.taillard(tid)[3]
    ref, _, _ = oracle.Oracle(p).solve_dfs()
    check_solve(p, wsbb_amd.LB2, 4096, 1, ref)


@pytest.mark.parametrize("pool,nb", [(4096, 10), (64, 1), (300, 3), (45, 7)])
def test_20x5_pool_shapes(pool, nb):
    for tid in (1, 5):
        check_solve(oracle.taillard(tid)[3], wsbb_amd.LB2, pool, nb)


def test_20x5_lb1():
    for tid in (1, 3):
        check_solve(oracle.taillard(tid)[3], wsbb_amd.LB1, 4096, 1)


def test_small_random_instances():
    rng = np.random.default_rng(5)
    for trial in range(12):
        n, m = int(rng.integers(3, 12)), int(rng.integers(2, 8))
        p = rng.integers(1, int(rng.choice([4, 99])) + 1, size=(m, n)).astype(np.int32)
        check_solve(p, wsbb_amd.LB2, 2 * n + 2 + int(rng.integers(0, 50)),
                    int(rng.integers(1, 4)))


def test_ta011_20x10_config():
    """20x10, pools of 4096 children, ten sub-batches per round."""
    check_solve(oracle.taillard(11)[3], wsbb_amd.LB2, 4096, 10, 1582)


def test_budget_stops_without_proof():
    p = oracle.taillard(11)[3]
    with wsbb_amd.Engine(p) as eng:
        res = eng.solve(pool_size=4096, max_batches=1, budget=500)
    assert not res["proven"]
    assert res["stats"]["decomposed"] <= 500
    if res["perm"][0] >= 0:
        assert cmax_np(p, res["perm"]) == res["best"]


@pytest.mark.parametrize("rounds", [1, 3, 16])
def test_graph_replay_matches_direct_launches(rounds):
    """Rounds replayed from a captured hipGraph (many replays per solve, host
    checks between them) give the same search as direct launches, which the
    tests above compare with the oracle.  ta021 at 65536 x 8 takes the
    chained select and the chained settle."""
    for tid, pool, nb in ((1, 4096, 1), (11, 4096, 10), (3, 300, 3), (21, 65536, 8)):
        p = oracle.taillard(tid)[3]
        with wsbb_amd.Engine(p) as eng:
            eng.search_options(use_graph=0, rounds_per_sync=rounds)
            direct = eng.solve(pool_size=pool, max_batches=nb)
            eng.search_options(use_graph=1)
            graph = eng.solve(pool_size=pool, max_batches=nb)
            again = eng.solve(pool_size=pool, max_batches=nb)
        for res in (graph, again):
            assert res["best"] == direct["best"] and res["proven"]
            for k in STATS:
                assert res["stats"][k] == direct["stats"][k], (tid, k)


@pytest.mark.parametrize("rounds,every", [(1, 8), (4, 8), (8, 8), (3, 5), (16, 4)])
def test_bound_timing_every_kth_round(rounds, every):
    """search_profile(k) times the bound launch of every k-th round of the
    solve, counted across the host's batches of rounds: ceil(R / k) launches
    for R rounds that did work, whatever the batch size; the timed solve
    proves the same optimum with the same counters."""
    p = oracle.taillard(11)[3]
    with wsbb_amd.Engine(p) as eng:
        eng.search_options(rounds_per_sync=rounds)
        plain = eng.solve(pool_size=4096, max_batches=10)
        eng.search_profile(every)
        res = eng.solve(pool_size=4096, max_batches=10)
        ms, timed = eng.search_profile(0)
    R = res["stats"]["rounds"]
    assert res["best"] == plain["best"] == 1582 and res["proven"]
    for k in STATS:
        assert res["stats"][k] == plain["stats"][k], k
    assert timed == -(-R // every)
    assert ms > 0


def test_checkpoint_resume_continues_exactly():
    """A search stopped by its node budget, saved (pool, incumbent, counters)
    and resumed in a fresh context, in segments, visits exactly the nodes of
    the uninterrupted search."""
    p = oracle.taillard(11)[3]
    with wsbb_amd.Engine(p) as eng:
        full = eng.solve(pool_size=4096, max_batches=10)
    ck = None
    res = None
    for seg_budget in (60_000, 150_000, 0):
        with wsbb_amd.Engine(p) as eng:
            if ck is None:
                res = eng.solve(pool_size=4096, max_batches=10, budget=seg_budget)
            else:
                res = eng.resume(ck, pool_size=4096, max_batches=10, budget=seg_budget)
            ck = eng.checkpoint()
        if seg_budget:
            assert not res["proven"] and ck["permT"].shape[1] > 0
            assert ck["stats"][0] <= seg_budget
    assert res["proven"] and res["best"] == full["best"] == 1582
    assert ck["permT"].shape[1] == 0
    assert cmax_np(p, res["perm"]) == res["best"]
    for k in STATS:
        assert res["stats"][k] == full["stats"][k], k
