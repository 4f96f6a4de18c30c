"""50x20 instances solved to proven optimum on the GPU, plain LB2 and
learned pair selection, against the oracle's CPU explorer.  The instances
are machine-correlated (p[k][j] = b_k + U[0,20], b_k ~ U[1,80]); Taillard's
50x20 class and job-correlated 50x20 instances do not close within minutes
of B&B with this bound.  A narrow dive (256-child pools) gives the first
incumbent; the search proper then starts from it."""
import numpy as np
import pytest

import oracle
import wsbb_amd

pytestmark = pytest.mark.gpu


def cmax_np(p, perm):
    t = np.zeros(p.shape[0], np.int64)
    for j in perm:
        t[0] += p[0, j]
        for k in range(1, p.shape[0]):
            t[k] = max(t[k], t[k - 1]) + p[k, j]
    return int(t[-1])


def gpu_plain(eng, pool=4096, nb=4):
    dive = eng.solve(pool_size=256, max_batches=1, budget=2000)
    return eng.solve(pool_size=pool, max_batches=nb, ub0=dive["best"], budget=20_000_000)


def gpu_online(eng, pool=65536, nb=4, warm=300_000, topk=24):
    eng.selector_init(seed=1)
    eng.selector_config(topk=topk, margin=-1, parents=True)
    cap = 1 << 17
    eng.selector_log(stride=max(1, warm * eng.n // cap), capacity=cap)
    dive = eng.solve(pool_size=256, max_batches=1, budget=2000)
    first = eng.solve(pool_size=pool, max_batches=nb, ub0=dive["best"], budget=warm // 40)
    if first["proven"]:
        return first
    ck = eng.checkpoint()
    eng.selector_train(epochs=2, batch=2048, lr=2e-3, seed=2)
    eng.selector_log(stride=0, capacity=0)
    return eng.resume(ck, kind=wsbb_amd.LEARNED, pool_size=pool, max_batches=nb,
                      budget=20_000_000)


@pytest.mark.parametrize("seed", [38, 65])
def test_50x20_plain_and_learned_prove_the_cpu_optimum(seed):
    p = wsbb_amd.generate_machine_correlated(50, 20, seed)
    assert np.array_equal(p, oracle.generate_machine_correlated(50, 20, seed))
    ref, _, ref_perm = oracle.Oracle(p).solve_parallel(threads=16)
    assert cmax_np(p, ref_perm) == ref
    with wsbb_amd.Engine(p) as eng:
        plain = gpu_plain(eng)
        assert plain["proven"] and plain["best"] == ref
        assert cmax_np(p, plain["perm"]) == ref
        learned = gpu_online(eng)
        assert learned["proven"] and learned["best"] == ref
        assert cmax_np(p, learned["perm"]) == ref
