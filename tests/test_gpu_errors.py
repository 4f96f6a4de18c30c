"""Error paths with a live context: bad kinds, sizes and limits at the
entry points are refused with WSBB_EINVAL and a message before any work is
queued, pool overflow and budget stops report their codes, and the same
context then solves correctly (the reference exits the process on any CUDA
error, gpu.cu:54-63; the library never does)."""
import ctypes

import numpy as np
import pytest

import oracle
import wsbb_amd

pytestmark = pytest.mark.gpu
L = wsbb_amd.lib
EINVAL = -1


def ptr(a, t=wsbb_amd._i32p):
    return a.ctypes.data_as(t)


def expect_einval(rc, what):
    assert rc == EINVAL, (what, rc)
    msg = L.wsbb_last_error().decode()
    assert msg, what
    return msg


def test_bad_arguments_leave_the_context_usable():
    p = oracle.taillard(1)[3]
    n = p.shape[1]
    with wsbb_amd.Engine(p) as eng:
        h = eng.h
        permT, l1, l2 = wsbb_amd.random_nodes(n, 64, 1)
        out = np.zeros(64, np.int32)
        u8 = wsbb_amd._u8p
        # batch bound: kind, size, limits, job ids
        expect_einval(L.wsbb_bound_batch(h, ptr(permT, u8), ptr(l1), ptr(l2), 64, 7, -1, ptr(out)),
                      "kind")
        expect_einval(L.wsbb_bound_batch(h, ptr(permT, u8), ptr(l1), ptr(l2), -3, 0, -1, ptr(out)),
                      "N < 0")
        for bad in ("l1", "l2", "order", "job"):
            pt, a, b = permT.copy(), l1.copy(), l2.copy()
            if bad == "l1":
                a[5] = -2
            elif bad == "l2":
                b[5] = n + 1
            elif bad == "order":
                a[5], b[5] = 7, 7
            else:
                pt[3, 9] = n
            msg = expect_einval(L.wsbb_bound_batch(h, ptr(pt, u8), ptr(a), ptr(b), 64, 0, -1,
                                                   ptr(out)), bad)
            assert "out of range" in msg
            expect_einval(L.wsbb_bound_batch_noalloc(h, ptr(pt, u8), ptr(a), ptr(b), 64, 0, -1,
                                                     ptr(out)), bad + " noalloc")
        # search shapes and kinds
        best, st, bp = ctypes.c_int32(), np.zeros(wsbb_amd.NSTATS, np.int64), np.zeros(n, np.int32)
        for kind, pool, nb in ((9, 4096, 1), (0, 2 * n + 1, 1), (0, 4096, 0), (0, 4096, 65),
                               (0, 1 << 20, 16), (wsbb_amd.LEARNED, 4096, 1)):
            expect_einval(L.wsbb_solve(h, kind, pool, nb, 2**31 - 1, 0, ctypes.byref(best),
                                       ptr(bp), ptr(st, wsbb_amd._i64p)), (kind, pool, nb))
        expect_einval(L.wsbb_set_pool_capacity(h, 100), "capacity")
        expect_einval(L.wsbb_search_options(h, -1, 65), "rounds per sync")
        expect_einval(L.wsbb_set_launch(h, 3, 0), "groups")
        expect_einval(L.wsbb_set_int16(h, 5), "int16 mode")
        # live pool: no session yet, then bad nodes
        expect_einval(L.wsbb_pool_seed_root(h), "no session")
        eng.pool_open(pool_size=4096, max_batches=1)
        expect_einval(L.wsbb_pool_pop(h, -1, 0, None, None, None, None, ctypes.byref(best)),
                      "max < 0")
        leaf = dict(permT=np.arange(n, dtype=np.uint8).reshape(n, 1),
                    l1=np.array([n - 2], np.int32), l2=np.array([n - 1], np.int32),
                    cost=np.array([0], np.int32))
        with pytest.raises(wsbb_amd.WsbbError) as ei:
            eng.pool_push(leaf)
        assert ei.value.code == EINVAL
        assert eng.pool_size() == 0
        # selector: bad top-k, no records
        expect_einval(L.wsbb_selector_config(h, 0, -1), "selector before init")
        eng.selector_init(seed=1)
        expect_einval(L.wsbb_selector_config(h, 0, -1), "top-k 0")
        # the context still bounds and solves correctly
        ref = oracle.Oracle(p).bound_batch(permT, l1, l2)
        assert np.array_equal(eng.bound(permT, l1, l2), ref)
        res = eng.solve(pool_size=4096, max_batches=1)
        assert res["proven"] and res["best"] == 1278


def test_budget_and_overflow_codes_then_a_full_solve():
    """A node budget returns WSBB_EBUDGET with the incumbent so far, a pool
    too small returns WSBB_EOVERFLOW with a message, and the same context
    then proves the optimum."""
    p = oracle.taillard(11)[3]
    with wsbb_amd.Engine(p) as eng:
        res = eng.solve(pool_size=4096, max_batches=10, budget=1000)
        assert not res["proven"] and res["stats"]["decomposed"] <= 1000
        eng.set_pool_capacity(1024)
        with pytest.raises(wsbb_amd.WsbbError) as ei:
            eng.solve(pool_size=4096, max_batches=10)
        assert ei.value.code == -4 and "overflow" in str(ei.value)
        eng.set_pool_capacity(1 << 22)
        res = eng.solve(pool_size=4096, max_batches=10)
        assert res["proven"] and res["best"] == 1582


def test_valu_probe_counts_its_lane_operations():
    """The integer VALU probe (the check of the derived int32 lane-op peak)
    counts 16 instructions per thread and iteration over eight 256-thread
    blocks per CU, refuses bad kinds and a short buffer, and a run of every
    kind finishes with the context still usable."""
    with wsbb_amd.Engine(oracle.taillard(1)[3]) as eng:
        nbytes = eng.num_cus * 8 * 256 * 4
        d = eng.malloc(nbytes)
        ops = ctypes.c_double()
        expect_einval(L.wsbb_valu_probe(eng.h, 3, 10, d, nbytes, ctypes.byref(ops)), "kind")
        expect_einval(L.wsbb_valu_probe(eng.h, 0, 10, d, nbytes - 4, ctypes.byref(ops)), "buffer")
        for kind in range(3):
            assert eng.valu_probe(kind, 10, d, nbytes) == 16 * 10 * eng.num_cus * 8 * 256
        eng.synchronize()
        eng.free(d)
        permT, l1, l2 = wsbb_amd.random_nodes(eng.n, 256, 3)
        assert np.array_equal(eng.bound(permT, l1, l2), oracle.Oracle(oracle.taillard(1)[3]).bound_batch(permT, l1, l2))
