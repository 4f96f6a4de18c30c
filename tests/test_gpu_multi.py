"""Mixed batches of several instances bounded in one launch (wsbb_multi_*):
every node must get exactly the value the oracle gives it for its own
instance (plain LB2), or its LB2 over its tile's pair subset."""
import numpy as np
import pytest

import oracle
import wsbb_amd

pytestmark = pytest.mark.gpu


def mixed_batch(specs, seed=3):
    """specs: list of (p, count) -> (ps, permT [maxn][N], l1, l2, offsets)"""
    ps = [p for p, _ in specs]
    maxn = max(p.shape[1] for p in ps)
    counts = [c for _, c in specs]
    N = sum(counts)
    permT = np.zeros((maxn, N), np.uint8)
    l1 = np.zeros(N, np.int32)
    l2 = np.zeros(N, np.int32)
    offsets = np.zeros(len(specs) + 1, np.int32)
    at = 0
    for i, (p, c) in enumerate(specs):
        offsets[i] = at
        if c:
            pt, a, b = oracle.random_nodes(p.shape[1], c, seed + 17 * i)
            permT[:p.shape[1], at:at + c] = pt
            l1[at:at + c] = a
            l2[at:at + c] = b
        at += c
    offsets[-1] = at
    return ps, permT, l1, l2, offsets


def specs_uneven():
    c = oracle.generate_correlated
    return [(c(20, 5, 11), 100), (c(50, 10, 12), 64), (c(100, 20, 13), 1), (oracle.taillard(21)[3], 130),
            (c(50, 5, 14), 0), (c(20, 10, 15), 257), (oracle.taillard(1)[3], 70), (c(100, 10, 16), 65)]


def test_mixed_batch_plain_matches_oracle():
    ps, permT, l1, l2, off = mixed_batch(specs_uneven())
    with wsbb_amd.Multi(ps) as mb:
        mb.load(permT, l1, l2, off)
        out, st, _ = mb.run(kind=wsbb_amd.LB2)
    assert st["nodes"] == len(l1)
    for i, p in enumerate(ps):
        a, b = off[i], off[i + 1]
        if a == b:
            continue
        n = p.shape[1]
        ref = oracle.Oracle(p).bound_batch(np.ascontiguousarray(permT[:n, a:b]), l1[a:b], l2[a:b])
        assert np.array_equal(out[a:b], ref), i


def test_mixed_batch_tile_subsets_match_oracle():
    ps, permT, l1, l2, off = mixed_batch(specs_uneven(), seed=9)
    Pmax = max(p.shape[0] * (p.shape[0] - 1) // 2 for p in ps)
    nsw = (Pmax + 31) // 32
    rng = np.random.default_rng(4)
    # tiles in instance order, 64 nodes each
    tiles = [(i, t) for i in range(len(ps)) for t in range(off[i], off[i + 1], 64)]
    masks = np.zeros((len(tiles), nsw), np.uint32)
    subsets = []
    for ti, (i, t0) in enumerate(tiles):
        P = ps[i].shape[0] * (ps[i].shape[0] - 1) // 2
        sel = rng.random(P) < 0.2
        sel[rng.integers(P)] = True
        subsets.append(sel)
        for s in np.nonzero(sel)[0]:
            masks[ti, s // 32] |= np.uint32(1) << np.uint32(s % 32)
    with wsbb_amd.Multi(ps) as mb:
        mb.load(permT, l1, l2, off)
        out, _, _ = mb.run(kind=wsbb_amd.LEARNED, tile_sel=masks)
        full, _, _ = mb.run(kind=wsbb_amd.LB2)
        inc = np.array([int(np.median(full[off[i]:off[i + 1]])) if off[i + 1] > off[i] else 0
                        for i in range(len(ps))], np.int32)
        fb, stf, _ = mb.run(kind=wsbb_amd.LEARNED, tile_sel=masks, inc=inc, margin=40)
    for ti, (i, t0) in enumerate(tiles):
        o = oracle.Oracle(ps[i])
        n = ps[i].shape[1]
        for k in range(t0, min(t0 + 64, off[i + 1])):
            perm = permT[:n, k].astype(np.int32)
            ref = o.bound_subset(perm, int(l1[k]), int(l2[k]), subsets[ti].astype(np.uint8))
            assert out[k] == ref, (i, k)
            leaf = l2[k] - l1[k] == 1
            near = (not leaf) and inc[i] - 40 <= ref < inc[i]
            assert fb[k] == (full[k] if near else ref), (i, k)
    assert stf["fallbacks"] == sum(
        1 for i in range(len(ps)) for k in range(off[i], off[i + 1])
        if l2[k] - l1[k] != 1 and inc[i] - 40 <= out[k] < inc[i])


def test_mixed_batch_learned_is_a_valid_bound():
    """MLP-chosen subsets (untrained selectors) never exceed full LB2; with
    the fallback over every node below the incumbent they reproduce it."""
    ps, permT, l1, l2, off = mixed_batch(specs_uneven(), seed=21)
    with wsbb_amd.Multi(ps) as mb:
        seen = set()
        for i, p in enumerate(ps):
            if p.shape[0] not in seen:
                seen.add(p.shape[0])
                mb.engines[i].selector_init(seed=5 + i)
                mb.engines[i].selector_config(topk=4, margin=0)
        mb.load(permT, l1, l2, off)
        full, _, _ = mb.run(kind=wsbb_amd.LB2)
        sub, st, _ = mb.run(kind=wsbb_amd.LEARNED)
        big = np.full(len(ps), 2**30, np.int32)
        fb, _, _ = mb.run(kind=wsbb_amd.LEARNED, inc=big, margin=2**30)
    assert (sub <= full).all()
    assert np.array_equal(fb, full)
    assert st["pairs"] < len(full) * 190
