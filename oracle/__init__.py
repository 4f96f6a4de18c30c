"""TEST INFRASTRUCTURE ONLY: ctypes loader for the C oracle (oracle.c).

The oracle is the CPU restatement of the reference's flowshop B&B path and is
used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker and the timed CPU baseline.  The product (wsbb_amd) never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

LB2, LB1 = 0, 1
# search trace record types (oracle.h)
TR_TAKEN, TR_PRUNED, TR_LEAF, TR_REPUSH = 0, 1, 2, 3
NSTATS = 8
STAT_NAMES = ("decomposed", "bounded", "leaves", "pruned_pop", "pruned_insert",
              "improvements", "max_pool", "rounds")

_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_unif.restype = ctypes.c_long
        L.orc_unif.argtypes = [ctypes.POINTER(ctypes.c_long), ctypes.c_long, ctypes.c_long]
        L.orc_taillard_dims.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.POINTER(ctypes.c_long)]
        L.orc_generate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, _i32p]
        L.orc_generate_correlated.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _i32p]
        L.orc_generate_machine_correlated.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                                      _i32p]
        L.orc_random_nodes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _u8p, _i32p, _i32p]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [_i32p, ctypes.c_int, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_pairs.argtypes = [ctypes.c_void_p]
        L.orc_tables.argtypes = [ctypes.c_void_p] + [_i32p] * 6
        L.orc_bound.argtypes = [ctypes.c_void_p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_lb1.argtypes = [ctypes.c_void_p, _i32p, ctypes.c_int, ctypes.c_int]
        L.orc_cmax.argtypes = [ctypes.c_void_p, _i32p]
        L.orc_heads_tails.argtypes = [ctypes.c_void_p, _i32p, ctypes.c_int, ctypes.c_int, _i32p, _i32p]
        L.orc_pair_values.argtypes = [ctypes.c_void_p, _i32p, ctypes.c_int, ctypes.c_int, _i32p]
        L.orc_bound_subset.argtypes = [ctypes.c_void_p, _i32p, ctypes.c_int, ctypes.c_int, _u8p]
        L.orc_bound_batch.argtypes = [ctypes.c_void_p, _u8p, _i32p, _i32p, ctypes.c_int,
                                      ctypes.c_int, _i32p, ctypes.c_int]
        L.orc_decompose.argtypes = [ctypes.c_void_p, _i32p, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int] + [_i32p] * 9
        L.orc_solve_dfs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int64, _i64p, _i32p]
        L.orc_solve_batched.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int64, _i64p, _i32p]
        L.orc_solve_batched_ex.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                           ctypes.c_int, _i64p, _i32p]
        L.orc_solve_parallel.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, _i64p, _i32p]
        L.orc_trace_begin.argtypes = [_i32p, ctypes.c_int64, ctypes.c_int]
        L.orc_trace_end.restype = ctypes.c_int64
        _lib = L
    return _lib


def _p(a, t=_i32p):
    return a.ctypes.data_as(t)


def unif_sequence(seed, count, low=1, high=99):
    s = ctypes.c_long(seed)
    return [lib().orc_unif(ctypes.byref(s), low, high) for _ in range(count)]


def taillard(tid):
    """(n, m, seed, p[m, n]) of Taillard instance `tid` (1-based)."""
    n, m, seed = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_long()
    if lib().orc_taillard_dims(tid, ctypes.byref(n), ctypes.byref(m), ctypes.byref(seed)):
        raise ValueError(tid)
    return n.value, m.value, seed.value, generate(n.value, m.value, seed.value)


def generate(n, m, seed):
    p = np.zeros((m, n), np.int32)
    lib().orc_generate(n, m, seed, _p(p))
    return p


def generate_correlated(n, m, seed):
    p = np.zeros((m, n), np.int32)
    lib().orc_generate_correlated(n, m, seed, _p(p))
    return p


def generate_machine_correlated(n, m, seed):
    p = np.zeros((m, n), np.int32)
    lib().orc_generate_machine_correlated(n, m, seed, _p(p))
    return p


def random_nodes(n, N, seed0=0x5EED):
    permT = np.zeros((n, N), np.uint8)
    l1 = np.zeros(N, np.int32)
    l2 = np.zeros(N, np.int32)
    lib().orc_random_nodes(n, N, seed0, _p(permT, _u8p), _p(l1), _p(l2))
    return permT, l1, l2


class Oracle:
    """One instance of the CPU restatement."""

    def __init__(self, p):
        self.p = np.ascontiguousarray(p, dtype=np.int32)
        self.m, self.n = self.p.shape
        self.h = lib().orc_create(_p(self.p), self.n, self.m)
        if not self.h:
            raise ValueError("bad instance shape")
        self.P = lib().orc_pairs(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def tables(self):
        P, n, m = self.P, self.n, self.m
        ma, mb = np.zeros(P, np.int32), np.zeros(P, np.int32)
        lag, john = np.zeros((P, n), np.int32), np.zeros((P, n), np.int32)
        R0, Q0 = np.zeros(m, np.int32), np.zeros(m, np.int32)
        lib().orc_tables(self.h, _p(ma), _p(mb), _p(lag), _p(john), _p(R0), _p(Q0))
        return dict(ma=ma, mb=mb, lag=lag, johnson=john, R0=R0, Q0=Q0)

    def bound(self, perm, l1, l2, best=999999):
        perm = np.ascontiguousarray(perm, np.int32)
        return lib().orc_bound(self.h, _p(perm), int(l1), int(l2), int(best))

    def lb1(self, perm, l1, l2):
        perm = np.ascontiguousarray(perm, np.int32)
        return lib().orc_lb1(self.h, _p(perm), int(l1), int(l2))

    def cmax(self, perm):
        perm = np.ascontiguousarray(perm, np.int32)
        return lib().orc_cmax(self.h, _p(perm))

    def heads_tails(self, perm, l1, l2):
        perm = np.ascontiguousarray(perm, np.int32)
        r, q = np.zeros(self.m, np.int32), np.zeros(self.m, np.int32)
        lib().orc_heads_tails(self.h, _p(perm), int(l1), int(l2), _p(r), _p(q))
        return r, q

    def pair_values(self, perm, l1, l2):
        perm = np.ascontiguousarray(perm, np.int32)
        v = np.zeros(max(self.P, 1), np.int32)
        lb = lib().orc_pair_values(self.h, _p(perm), int(l1), int(l2), _p(v))
        return lb, v[: self.P]

    def bound_subset(self, perm, l1, l2, sel):
        perm = np.ascontiguousarray(perm, np.int32)
        sel = np.ascontiguousarray(sel, np.uint8)
        return lib().orc_bound_subset(self.h, _p(perm), int(l1), int(l2), _p(sel, _u8p))

    def bound_batch(self, permT, l1, l2, kind=LB2, threads=1):
        permT = np.ascontiguousarray(permT, np.uint8)
        l1 = np.ascontiguousarray(l1, np.int32)
        l2 = np.ascontiguousarray(l2, np.int32)
        N = l1.shape[0]
        out = np.zeros(N, np.int32)
        lib().orc_bound_batch(self.h, _p(permT, _u8p), _p(l1), _p(l2), N, kind, _p(out), threads)
        return out

    def decompose(self, perm, l1, l2, depth, kind=LB2):
        n = self.n
        perm = np.ascontiguousarray(perm, np.int32)
        cperm = np.zeros((2 * n, n), np.int32)
        arrs = [np.zeros(2 * n, np.int32) for _ in range(6)]
        side, nb = np.zeros(1, np.int32), np.zeros(1, np.int32)
        cnt = lib().orc_decompose(self.h, _p(perm), l1, l2, depth, kind, _p(cperm),
                                  *[_p(a) for a in arrs], _p(side), _p(nb))
        cl1, cl2, cdepth, ccost, cjob, crank = [a[:cnt] for a in arrs]
        return dict(perm=cperm[:cnt], l1=cl1, l2=cl2, depth=cdepth, cost=ccost,
                    job=cjob, rank=crank, side=int(side[0]), nbounded=int(nb[0]))

    def _traced(self, run, trace):
        """Runs run() with the search trace on when trace is a record
        capacity; returns (run's result, trace dict or None)."""
        if not trace:
            return run(), None
        n = self.n
        buf = np.zeros((int(trace), 2 * n + 8), np.int32)
        lib().orc_trace_begin(_p(buf), int(trace), n)
        try:
            res = run()
        finally:
            cnt = lib().orc_trace_end()
        if cnt > trace:
            raise ValueError(f"trace needs {cnt} records, capacity {trace}")
        r = buf[:cnt]
        return res, dict(type=r[:, 0], round=r[:, 1], l1=r[:, 2], l2=r[:, 3], depth=r[:, 4],
                         cost=r[:, 5], perm=r[:, 7:7 + n], ranks=r[:, 7 + n:8 + 2 * n])

    def solve_dfs(self, kind=LB2, order=1, ub0=2**31 - 1, budget=0, trace=0):
        """CPU DFS explorer.  trace = record capacity: also return the
        trace of pops (oracle.h) as a fourth value."""
        st = np.zeros(NSTATS, np.int64)
        bp = np.zeros(self.n, np.int32)
        best, tr = self._traced(lambda: lib().orc_solve_dfs(
            self.h, kind, order, ub0, budget, _p(st, _i64p), _p(bp)), trace)
        res = (best, dict(zip(STAT_NAMES, st.tolist())), bp)
        return res + (tr,) if trace else res

    def solve_parallel(self, kind=LB2, threads=8, ub0=2**31 - 1):
        """All-cores CPU explorer over independent subtrees (same optimum
        as solve_dfs; counters depend on thread timing)."""
        st = np.zeros(NSTATS, np.int64)
        bp = np.zeros(self.n, np.int32)
        best = lib().orc_solve_parallel(self.h, kind, threads, ub0, _p(st, _i64p), _p(bp))
        return best, dict(zip(STAT_NAMES, st.tolist())), bp

    def solve_batched(self, kind=LB2, pool_size=4096, max_batches=1, ub0=2**31 - 1, budget=0,
                      order=1, trace=0):
        """Batched GPU-path explorer; order 1 = stack, 0 = ranks heap.
        trace = record capacity: also return the trace as a fourth value."""
        st = np.zeros(NSTATS, np.int64)
        bp = np.zeros(self.n, np.int32)
        best, tr = self._traced(lambda: lib().orc_solve_batched_ex(
            self.h, kind, pool_size, max_batches, ub0, budget, order, _p(st, _i64p), _p(bp)),
            trace)
        res = (best, dict(zip(STAT_NAMES, st.tolist())), bp)
        return res + (tr,) if trace else res
