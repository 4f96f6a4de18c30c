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
