"""The live node pool (wsbb_pool_*): a search driven round by round from
the host, nodes moved in and out between rounds, incumbents exchanged --
the GPU peer of the reference fed by host threads."""
import threading
import time

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


def test_rounds_in_segments_equal_one_solve():
    """Seeding the root and running the session a few rounds at a time
    visits exactly the nodes of one wsbb_solve; popping nodes off the top or
    the bottom and pushing them back leaves the search unchanged."""
    p = oracle.taillard(11)[3]
    with wsbb_amd.Engine(p) as eng:
        full = eng.solve(pool_size=4096, max_batches=1)
        eng.pool_open(pool_size=4096, max_batches=1)
        eng.pool_seed_root()
        assert eng.pool_size() == 1
        seg = 0
        while True:
            rc, size, st = eng.pool_run(rounds=7)
            assert rc == 0
            seg += 1
            if size == 0:
                break
            if seg % 5 == 0 and size >= 4:
                ck = eng.checkpoint()
                top = eng.pool_pop(3)
                bot = eng.pool_pop(2, from_bottom=True)
                assert eng.pool_size() == size - 5
                # bottom-first batches: the top three are the snapshot's last
                assert np.array_equal(top["permT"], ck["permT"][:, -3:])
                assert np.array_equal(bot["l1"], ck["l1"][:2])
                # put everything back where it was: the bottom two under the rest
                rest = eng.pool_pop(size - 5)
                eng.pool_push(bot)
                eng.pool_push(rest)
                eng.pool_push(top)
                again = eng.checkpoint()
                for k in ("permT", "l1", "l2", "cost"):
                    assert np.array_equal(again[k], ck[k]), k
        inc, perm = eng.incumbent()
    assert inc == full["best"] == 1582 and cmax_np(p, perm) == inc
    for k in STATS:
        assert st[k] == full["stats"][k], k


def test_incumbent_set_prunes_and_only_lowers():
    """An incumbent from elsewhere prunes like the device's own: with the
    optimum given before the root, the search ends with fewer decompositions
    and keeps that value; a larger value never replaces it."""
    p = oracle.taillard(11)[3]
    with wsbb_amd.Engine(p) as eng:
        plain = eng.solve(pool_size=4096, max_batches=10)
        eng.pool_open(pool_size=4096, max_batches=10)
        eng.set_incumbent(plain["best"], plain["perm"])
        eng.set_incumbent(plain["best"] + 5)
        eng.pool_seed_root()
        rc, size, st = eng.pool_run()
        inc, perm = eng.incumbent()
    assert rc == 0 and size == 0
    assert inc == plain["best"] and np.array_equal(perm, plain["perm"])
    assert st["decomposed"] < plain["stats"]["decomposed"]
    assert st["improvements"] == 0


def test_two_host_threads_feed_their_own_contexts():
    """Two host threads, one context each on the same GPU: the first seeds
    the root; a thread whose pool runs dry asks for work and the other
    gives it the bottom half of its pool (the shallowest nodes); between
    rounds each adopts the better incumbent.  Together they prove the
    oracle's optimum."""
    p = oracle.taillard(11)[3]
    ref, _, _ = oracle.Oracle(p).solve_dfs()
    lock = threading.Lock()
    shared = dict(inc=2**31 - 1, perm=None, mail=[[], []], want=[False, False],
                  idle=[False, False], done=False, given=0, errors=[])

    def worker(me):
        other = 1 - me
        try:
            with wsbb_amd.Engine(p) as eng:
                eng.pool_open(pool_size=4096, max_batches=2)
                if me == 0:
                    eng.pool_seed_root()
                while True:
                    with lock:
                        for batch in shared["mail"][me]:
                            eng.pool_push(batch)
                        shared["mail"][me] = []
                        if shared["inc"] < 2**31 - 1:
                            eng.set_incumbent(shared["inc"], shared["perm"])
                    size = eng.pool_size()
                    if size == 0:
                        with lock:
                            shared["idle"][me] = True
                            shared["want"][me] = True
                            if shared["idle"][other] and not shared["mail"][me] \
                                    and not shared["mail"][other]:
                                shared["done"] = True
                            if shared["done"]:
                                return
                        time.sleep(0.0005)
                        continue
                    with lock:
                        shared["idle"][me] = False
                    rc, size, _ = eng.pool_run(rounds=3)
                    assert rc == 0
                    inc, perm = eng.incumbent()
                    with lock:
                        if inc < shared["inc"]:
                            shared["inc"], shared["perm"] = inc, perm.copy()
                        if shared["want"][other] and size >= 2:
                            shared["mail"][other].append(eng.pool_pop(size // 2, from_bottom=True))
                            shared["want"][other] = False
                            shared["given"] += 1
        except Exception as e:  # reported by the main thread
            shared["errors"].append(repr(e))
            with lock:
                shared["done"] = True

    th = [threading.Thread(target=worker, args=(i,)) for i in (0, 1)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in th), "workers did not finish"
    assert not shared["errors"], shared["errors"]
    assert shared["given"] >= 1
    assert shared["inc"] == ref == 1582
    assert cmax_np(p, shared["perm"]) == ref
