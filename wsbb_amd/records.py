"""The reference's binary node record, as its work-stealing messages carry
nodes between peers: int32 fields in host (little-endian) byte order,
permutation[MAX_NBJOBS], limite1, limite2, couts_somme, depth, begin_end,
ranks[MAX_NBJOBS + 1] (src/work_stealing/node.cpp:524-556 writes them,
1240-1283 reads them back with couts[0] = couts_somme).  MAX_NBJOBS is 20 in
the reference (headers/tree.h:31): a record is 184 bytes.  The same codec is
in include/wsbb_plugins.hpp (serialize_nodes / deserialize_nodes)."""
import numpy as np

MAX_NBJOBS = 20


def record_dtype(max_jobs=MAX_NBJOBS):
    return np.dtype([("permutation", "<i4", (max_jobs,)), ("limite1", "<i4"),
                     ("limite2", "<i4"), ("couts_somme", "<i4"), ("depth", "<i4"),
                     ("begin_end", "<i4"), ("ranks", "<i4", (max_jobs + 1,))])


def encode(recs, max_jobs=MAX_NBJOBS):
    """A structured array of record_dtype(max_jobs) as message bytes."""
    return np.ascontiguousarray(recs, dtype=record_dtype(max_jobs)).tobytes()


def decode(data, max_jobs=MAX_NBJOBS):
    dt = record_dtype(max_jobs)
    if len(data) % dt.itemsize:
        raise ValueError("partial node record")
    return np.frombuffer(data, dtype=dt).copy()


def from_pool(ckpt, max_jobs=MAX_NBJOBS):
    """Device pool nodes (a checkpoint dict: permT [n, c] bottom-first, l1,
    l2, cost) as records whose rank strings give the reference's priority
    queue the device's pop order: ranks[0] = nodes above it on the stack."""
    permT = np.asarray(ckpt["permT"])
    n, c = permT.shape
    if n > max_jobs:
        raise ValueError("more jobs than the record holds")
    r = np.zeros(c, dtype=record_dtype(max_jobs))
    r["permutation"][:, :n] = permT.T
    r["limite1"] = ckpt["l1"]
    r["limite2"] = ckpt["l2"]
    r["couts_somme"] = ckpt["cost"]
    r["depth"] = (np.asarray(ckpt["l1"]) + 1) + (n - np.asarray(ckpt["l2"]))
    r["begin_end"] = -1
    r["ranks"][:, 0] = c - 1 - np.arange(c)
    return r


def to_pool(recs, n):
    """Records (any order) as pool nodes, bottom-first: the smallest rank
    string, which the reference's queue pops first, ends on top."""
    keys = [recs["ranks"][:, i] for i in range(recs["ranks"].shape[1] - 1, -1, -1)]
    order = np.lexsort(keys)[::-1]  # descending rank strings
    r = recs[order]
    return dict(permT=np.ascontiguousarray(r["permutation"][:, :n].T).astype(np.uint8),
                l1=r["limite1"].astype(np.int32), l2=r["limite2"].astype(np.int32),
                cost=r["couts_somme"].astype(np.int32))
