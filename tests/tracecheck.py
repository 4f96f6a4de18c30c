"""Helpers for checks on search traces (oracle and device).

A rank string is the reference's ranks[] (tree.cpp:37-46 orders the pool by
them lexicographically).  Children of a traced device round are rebuilt from
their descriptors the way the reference builds them (create_raw_problem3,
bb_problem_calculator.cpp:25-33: swap, limites_set 40-48)."""
import numpy as np


def lex_nondecreasing(ranks, strict=False):
    """True if the rows of ranks are (strictly) increasing lexicographically."""
    r = np.asarray(ranks)
    if r.shape[0] < 2:
        return True
    a, b = r[:-1].astype(np.int64), r[1:].astype(np.int64)
    diff = a != b
    anyd = diff.any(axis=1)
    first = np.argmax(diff, axis=1)
    rows = np.arange(a.shape[0])
    up = b[rows, first] > a[rows, first]
    ok = np.where(anyd, up, not strict)
    return bool(ok.all())


def rebuild_children(rnd, n):
    """(permT [n, nc], l1, l2) of a traced device round's children."""
    desc = rnd["desc"]
    p, side, j = desc >> 9, (desc >> 8) & 1, desc & 255
    pl1, pl2 = rnd["l1"][p], rnd["l2"][p]
    permT = rnd["permT"][:, p].copy()
    t = np.where(side == 0, pl1 + 1, pl2 - 1)
    cols = np.arange(desc.shape[0])
    a, b = permT[t, cols].copy(), permT[j, cols].copy()
    permT[t, cols] = b
    permT[j, cols] = a
    l1 = (pl1 + (side == 0)).astype(np.int32)
    l2 = (pl2 - (side == 1)).astype(np.int32)
    l1 = np.where(l2 - l1 == 2, l1 + 1, l1).astype(np.int32)
    return permT, l1, l2


def taken(tr, taken_type=0):
    """Records of parents taken for decomposition, as arrays."""
    sel = tr["type"] == taken_type
    return {k: v[sel] for k, v in tr.items()}
