"""The reference's DFS explorer written against the C++ plugin face
(include/wsbb_plugins.hpp: bound_abstract -> bound_hip, decompose_beginend,
problems, select_worst) must reproduce the oracle's CPU explorer exactly:
same optimum, decompositions, bounded children, leaves and prunes."""
import os
import subprocess

import pytest

import oracle

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "plugin_explorer")


@pytest.mark.parametrize("tid", [1, 2, 3, 4, 7])
def test_plugin_explorer_matches_oracle_dfs(tid):
    assert os.path.exists(BIN), "build the library first (wsbb_amd/build.py)"
    out = subprocess.run([BIN, str(tid)], capture_output=True, text=True, timeout=240,
                         check=True).stdout.split()
    best, decomposed, bounded, leaves, pruned = map(int, out)
    p = oracle.taillard(tid)[3]
    ref, st, _ = oracle.Oracle(p).solve_dfs()
    assert best == ref
    assert (decomposed, bounded, leaves, pruned) == (
        st["decomposed"], st["bounded"], st["leaves"], st["pruned_pop"])


def test_pool_checkpoint_file_resume(tmp_path):
    """A device search saved to a pool checkpoint file mid-way and resumed
    from the file in a fresh context ends with the uninterrupted search's
    optimum and counters."""
    import wsbb_amd
    assert os.path.exists(BIN), "build the library first (wsbb_amd/build.py)"
    path = str(tmp_path / "ta011.pool")
    out = subprocess.run([BIN, "--resume", "11", "100000", path], capture_output=True, text=True,
                         timeout=240, check=True).stdout.split()
    best, proven, st = int(out[0]), int(out[1]), list(map(int, out[2:]))
    with wsbb_amd.Engine(oracle.taillard(11)[3]) as eng:
        full = eng.solve(pool_size=4096, max_batches=10)
    assert proven == 1 and best == full["best"] == 1582
    for i, k in enumerate(wsbb_amd.STAT_NAMES[:8]):
        assert st[i] == full["stats"][k], k
    with open(path) as f:
        assert f.readline().startswith("wsbb-pool 20 ")


def test_binary_records_resume(tmp_path):
    """A device search stopped mid-way, its pool written as the reference's
    184-byte binary node records, read back and pushed into a fresh
    context's live pool with the incumbent (gpu_pool), ends at the optimum
    having decomposed exactly the parents the uninterrupted search does."""
    import wsbb_amd
    assert os.path.exists(BIN), "build the library first (wsbb_amd/build.py)"
    path = str(tmp_path / "ta011.rec")
    out = subprocess.run([BIN, "--records", "11", "100000", path], capture_output=True,
                         text=True, timeout=240, check=True).stdout.split()
    best, nrec, first, rest = map(int, out)
    with wsbb_amd.Engine(oracle.taillard(11)[3]) as eng:
        full = eng.solve(pool_size=4096, max_batches=10)
    assert best == full["best"] == 1582
    assert nrec > 0 and os.path.getsize(path) == 184 * nrec
    assert first + rest == full["stats"]["decomposed"]


def test_pool_through_binary_records_python():
    """The same through the Python codec and wsbb_solve_resume."""
    import wsbb_amd
    from wsbb_amd import records
    p = oracle.taillard(11)[3]
    with wsbb_amd.Engine(p) as eng:
        full = eng.solve(pool_size=4096, max_batches=10)
        part = eng.solve(pool_size=4096, max_batches=10, budget=80_000)
        ck = eng.checkpoint()
    assert not part["proven"]
    msg = records.encode(records.from_pool(ck))
    nodes = records.to_pool(records.decode(msg)[::-1], p.shape[1])  # any message order
    ck2 = dict(ck, **nodes)
    with wsbb_amd.Engine(p) as eng:
        res = eng.resume(ck2, pool_size=4096, max_batches=10)
    assert res["proven"] and res["best"] == full["best"]
    for k in ("decomposed", "bounded", "leaves", "pruned_pop", "pruned_insert", "rounds"):
        assert res["stats"][k] == full["stats"][k], k
