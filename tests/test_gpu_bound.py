"""GPU parity of the bound kernels against the oracle, through the C ABI.

Integer work: every value must be bit-identical to the oracle's restatement
of bound_flowshop::bornes_calculer on the same node."""
import numpy as np
import pytest

import oracle
import wsbb_amd

pytestmark = pytest.mark.gpu


def edge_nodes(n, rng):
    """root, leaves, one free job, prefix-only and suffix-only nodes."""
    perms, l1s, l2s = [], [], []

    def add(perm, l1, l2):
        perms.append(perm)
        l1s.append(l1)
        l2s.append(l2)

    ident = np.arange(n)
    add(ident, -1, n)                           # root
    for _ in range(8):
        perm = rng.permutation(n)
        l1 = int(rng.integers(-1, n))
        add(perm, l1, l1 + 1)                   # leaf
        l1 = int(rng.integers(-1, n - 1))
        add(perm, l1, l1 + 2)                   # one free job
        add(perm, int(rng.integers(-1, n - 1)), n)  # no suffix
        add(perm, -1, int(rng.integers(1, n + 1)))  # no prefix
    permT = np.array(perms, np.uint8).T.copy()
    return permT, np.array(l1s, np.int32), np.array(l2s, np.int32)


def cases():
    rng = np.random.default_rng(0)
    out = [("ta001_20x5", oracle.taillard(1)[3]), ("ta011_20x10", oracle.taillard(11)[3]),
           ("ta021_20x20", oracle.taillard(21)[3]), ("ta051_50x20", oracle.taillard(51)[3]),
           ("ta081_100x20", oracle.taillard(81)[3]),
           ("corr_50x10", oracle.generate_correlated(50, 10, 5)),
           ("ties_30x6", rng.integers(1, 4, size=(6, 30)).astype(np.int32)),
           ("max_128x24", rng.integers(1, 100, size=(24, 128)).astype(np.int32)),
           # bound values beyond 2^14: the int16 two-node form does not apply
           ("big_p_20x10", rng.integers(1, 3000, size=(10, 20)).astype(np.int32)),
           ("tiny_3x2", rng.integers(1, 10, size=(2, 3)).astype(np.int32))]
    return out


def batch_for(n, N, seed):
    permT, l1, l2 = oracle.random_nodes(n, N, seed)
    e = edge_nodes(n, np.random.default_rng(seed))
    return (np.concatenate([permT, e[0]], axis=1), np.concatenate([l1, e[1]]),
            np.concatenate([l2, e[2]]))


@pytest.mark.parametrize("name,p", cases(), ids=[c[0] for c in cases()])
def test_lb2_lb1_parity(name, p):
    m, n = p.shape
    o = oracle.Oracle(p)
    N = 3000 if n <= 50 else 1000
    permT, l1, l2 = batch_for(n, N, 0x5EED)
    ref = o.bound_batch(permT, l1, l2, kind=oracle.LB2, threads=8)
    with wsbb_amd.Engine(p) as eng:
        # both the two-node int16 form (when the instance fits) and the int32 form
        for int16 in (2, 0):
            eng.set_int16(int16)
            for groups in (1, 2, 4):
                eng.set_launch(groups=groups)
                got = eng.bound(permT, l1, l2, kind=wsbb_amd.LB2)
                bad = np.nonzero(got != ref)[0]
                assert bad.size == 0, (int16, groups, bad[:5], got[bad[:5]], ref[bad[:5]])
        eng.set_int16(1)
        eng.set_launch(groups=0)
        got1 = eng.bound(permT, l1, l2, kind=wsbb_amd.LB1)
        ref1 = o.bound_batch(permT, l1, l2, kind=oracle.LB1, threads=8)
        assert np.array_equal(got1, ref1)
        leaf = (l2 - l1) == 1
        assert (got1[~leaf] <= ref[~leaf]).all() or m < 2


@pytest.mark.parametrize("tid", [11, 51])
def test_pair_values_argmax_best_subset(tid):
    p = oracle.taillard(tid)[3]
    m, n = p.shape
    o = oracle.Oracle(p)
    permT, l1, l2 = oracle.random_nodes(n, 700, 42)
    rng = np.random.default_rng(1)
    with wsbb_amd.Engine(p) as eng:
        P = eng.P
        res = eng.bound(permT, l1, l2, argmax=True, pairv=True, feat=True)
        for k in range(0, 700, 7):
            perm = permT[:, k].astype(np.int32)
            lb, v = o.pair_values(perm, l1[k], l2[k])
            assert res["bound"][k] == lb
            assert np.array_equal(res["pairv"][:, k], v)
            assert res["argmax"][k] == int(np.argmax(v))
            # features: head, tail, remaining work per machine (scaled)
            r, q = o.heads_tails(perm, l1[k], l2[k])
            t = o.tables()
            d1, d2 = l1[k] + 1, n - l2[k]
            head = r if d1 else t["R0"]
            tail = q if d2 else t["Q0"]
            rem = p[:, perm[l1[k] + 1:l2[k]]].sum(axis=1)
            f = res["feat"][:, k].reshape(5, m)
            scale = int(p.max(axis=0).sum() + p.max(axis=1).sum())  # path bound
            for row, ref in zip(f, [head, tail, rem, head + rem, rem + tail]):
                assert np.allclose(row * scale, ref, rtol=1e-5, atol=1e-3)
        # early exit against an incumbent
        best = rng.integers(500, 5000, size=700).astype(np.int32)
        for k in range(0, 700, 13):
            got = eng.bound(permT[:, k:k + 1], l1[k:k + 1], l2[k:k + 1], best=int(best[k]))[0]
            assert got == o.bound(permT[:, k].astype(np.int32), l1[k], l2[k], best=int(best[k]))
        # pair subsets
        nsw = (P + 31) // 32
        selb = rng.random((P, 700)) < 0.1
        sel = np.zeros((nsw, 700), np.uint32)
        for s in range(P):
            sel[s // 32] |= (selb[s].astype(np.uint32) << np.uint32(s % 32))
        got = eng.bound(permT, l1, l2, sel=sel)
        for k in range(0, 700, 5):
            ref = o.bound_subset(permT[:, k].astype(np.int32), l1[k], l2[k],
                                 selb[:, k].astype(np.uint8))
            assert got[k] == ref


def test_full_size_batch_50x20():
    """2^20 nodes of the 50x20 batched workload: full parity on a 2^15
    sample (oracle), properties on all nodes, and run-to-run determinism."""
    p = oracle.taillard(51)[3]
    n = p.shape[1]
    N = 1 << 20
    permT, l1, l2 = wsbb_amd.random_nodes(n, N, 0x5EED)
    o = oracle.Oracle(p)
    with wsbb_amd.Engine(p) as eng:
        lb2 = eng.bound(permT, l1, l2)
        lb2b = eng.bound(permT, l1, l2)
        lb1 = eng.bound(permT, l1, l2, kind=wsbb_amd.LB1)
    assert np.array_equal(lb2, lb2b)
    assert (lb1 <= lb2).all() and (lb2 > 0).all()
    idx = np.random.default_rng(3).choice(N, 1 << 15, replace=False)
    ref = o.bound_batch(permT[:, idx], l1[idx], l2[idx], threads=16)
    assert np.array_equal(lb2[idx], ref)


def test_device_resident_api_matches_host_api():
    p = oracle.taillard(21)[3]
    n = p.shape[1]
    N = 5000
    permT, l1, l2 = oracle.random_nodes(n, N, 7)
    with wsbb_amd.Engine(p) as eng:
        host = eng.bound(permT, l1, l2)
        dp, d1, d2, do = (eng.malloc(permT.nbytes), eng.malloc(4 * N), eng.malloc(4 * N),
                          eng.malloc(4 * N))
        try:
            eng.h2d(dp, permT)
            eng.h2d(d1, l1)
            eng.h2d(d2, l2)
            eng.bound_device(dp, d1, d2, N, do)
            eng.synchronize()
            dev = eng.d2h(np.zeros(N, np.int32), do)
        finally:
            for d in (dp, d1, d2, do):
                eng.free(d)
    assert np.array_equal(host, dev)


def test_invalid_nodes_are_rejected_on_the_host():
    p = oracle.taillard(1)[3]
    with wsbb_amd.Engine(p) as eng:
        permT, l1, l2 = oracle.random_nodes(20, 10, 1)
        bad = permT.copy()
        bad[0, 0] = 20
        with pytest.raises(wsbb_amd.WsbbError):
            eng.bound(bad, l1, l2)
        bl = l2.copy()
        bl[3] = l1[3]
        with pytest.raises(wsbb_amd.WsbbError):
            eng.bound(permT, l1, bl)
