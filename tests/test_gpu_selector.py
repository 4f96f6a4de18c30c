"""Learned pair selector: kernels against torch (CPU, fp32) and the oracle.

- training gradients (hand-written forward/backward kernels) against torch
  autograd of the same network with the same BF16 roundings;
- MFMA inference logits against the torch forward, tile subsets against a
  numpy top-k of the summed softmax;
- the subset bound of every node against the oracle's LB2 over that subset,
  with the full-LB2 fallback inside the margin;
- learned solves prove the plain optimum; training beats random weights;
  on 20x20 the trained network cuts pair evaluations at a small node
  inflation (the gain over the search without it).
Floating-point tolerances are written next to each comparison."""
import numpy as np
import pytest
import torch

import oracle
import wsbb_amd

pytestmark = pytest.mark.gpu


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def torch_grads(params, feat, label, idx):
    X = bf(torch.from_numpy(feat[:, idx].T.copy()))
    y = torch.from_numpy(label[idx].astype(np.int64))
    W1 = bf(torch.from_numpy(params["W1"])).requires_grad_()
    W2 = bf(torch.from_numpy(params["W2"])).requires_grad_()
    b1 = torch.from_numpy(params["b1"]).clone().requires_grad_()
    b2 = torch.from_numpy(params["b2"]).clone().requires_grad_()
    h = torch.relu(X @ W1 + b1)
    hb = h + (bf(h) - h).detach()  # rounding with identity gradient
    z = hb @ W2 + b2
    ok = y >= 0
    loss = torch.nn.functional.cross_entropy(z[ok], y[ok])
    loss.backward()
    return loss.item(), dict(W1=W1.grad.numpy(), b1=b1.grad.numpy(), W2=W2.grad.numpy(),
                             b2=b2.grad.numpy())


def torch_logits(params, feat):
    X = bf(torch.from_numpy(feat.T.copy()))
    h = bf(torch.relu(X @ bf(torch.from_numpy(params["W1"])) + torch.from_numpy(params["b1"])))
    return (h @ bf(torch.from_numpy(params["W2"])) + torch.from_numpy(params["b2"])).numpy()


def random_records(eng, R, seed):
    rng = np.random.default_rng(seed)
    feat = rng.random((5 * eng.m, R), dtype=np.float32)
    label = rng.integers(0, eng.P, size=R).astype(np.int32)
    label[rng.random(R) < 0.1] = -1
    return feat, label


@pytest.mark.parametrize("tid", [1, 21])
def test_gradients_match_torch(tid):
    p = oracle.taillard(tid)[3]
    with wsbb_amd.Engine(p) as eng:
        eng.selector_init(seed=3)
        params = eng.selector_params()
        feat, label = random_records(eng, 300, 5)
        eng.selector_set_records(feat, label)
        idx = np.random.default_rng(6).choice(300, 128, replace=False).astype(np.int32)
        loss, g = eng.selector_grad(idx)
    tloss, tg = torch_grads(params, feat, label, idx)
    # fp32 sums in a different order can flip a BF16 rounding of a hidden
    # activation: tolerances are 1e-3 relative on the loss and 2% of each
    # gradient tensor's largest entry elementwise; a wrong index, sign or
    # transposed operand gives O(1) errors.
    assert abs(loss - tloss) <= 1e-3 * abs(tloss)
    for k in ("W1", "b1", "W2", "b2"):
        scale = np.abs(tg[k]).max()
        assert np.allclose(g[k], tg[k], rtol=0, atol=2e-2 * scale), k


@pytest.mark.parametrize("tid", [1, 11, 21, 51])
def test_inference_matches_torch_and_tile_topk(tid):
    p = oracle.taillard(tid)[3]
    with wsbb_amd.Engine(p) as eng:
        eng.selector_init(seed=7)
        params = eng.selector_params()
        rng = np.random.default_rng(8)
        params["b2"] = rng.normal(0, 0.5, eng.P).astype(np.float32)
        eng.selector_params(params)
        topk = min(16, eng.P)
        eng.selector_config(topk=topk, margin=0)
        N = 1000
        feat = rng.random((5 * eng.m, N), dtype=np.float32)
        scores, tsel = eng.selector_infer(feat)
    ref = torch_logits(params, feat).T  # [P][N]
    # MFMA accumulates in fp32 in its own order: 2e-3 absolute on logits
    assert np.allclose(scores, ref, rtol=0, atol=2e-3)
    e = np.exp(ref - ref.max(axis=0))
    prob = e / e.sum(axis=0)
    for t in range((N + 63) // 64):
        agg = prob[:, 64 * t:64 * t + 64].sum(axis=1)
        chosen = [s for s in range(eng.P) if (tsel[t, s // 32] >> (s % 32)) & 1]
        assert len(chosen) == topk
        kth = np.sort(agg)[::-1][topk - 1]
        # every chosen pair is within rounding of the k-th largest sum
        assert agg[chosen].min() >= kth - 1e-4
        assert (agg[[s for s in range(eng.P) if s not in chosen]] <= kth + 1e-4).all()


@pytest.mark.parametrize("tid", [1, 11, 21, 51])
def test_parent_inference_and_tile_sums(tid):
    """Per-parent inference: BF16 softmax probabilities against the torch
    forward; the mask of every 64-child tile against a numpy top-k of its
    parents' probabilities weighted by their children in the tile (parents
    without children, two-leaf parents and tiles spanning many parents
    included)."""
    p = oracle.taillard(tid)[3]
    with wsbb_amd.Engine(p) as eng:
        eng.selector_init(seed=7)
        params = eng.selector_params()
        rng = np.random.default_rng(9)
        params["b2"] = rng.normal(0, 0.5, eng.P).astype(np.float32)
        eng.selector_params(params)
        topk = min(16, eng.P)
        eng.selector_config(topk=topk, margin=0)
        Np = 700
        feat = rng.random((5 * eng.m, Np), dtype=np.float32)
        counts = rng.integers(0, 41, Np).astype(np.int32)
        counts[rng.random(Np) < 0.2] = 2
        counts[rng.random(Np) < 0.05] = 0
        counts[-1] = 0
        probs, tsel = eng.selector_infer_parents(feat, counts)
    ref = torch_logits(params, feat)  # [Np][P]
    e = np.exp(ref - ref.max(axis=1, keepdims=True))
    pr = e / e.sum(axis=1, keepdims=True)
    # BF16 rounding (2^-9 relative) on top of the MFMA logits' 2e-3 absolute
    # error: 1% relative, 1e-6 absolute for underflowing entries
    assert np.allclose(probs, pr, rtol=1e-2, atol=1e-6)
    off = np.concatenate([[0], np.cumsum(counts)])
    Nc = int(off[-1])
    assert tsel.shape[0] == (Nc + 63) // 64
    for t in range(tsel.shape[0]):
        c0, c1 = 64 * t, min(64 * t + 64, Nc)
        w = np.clip(np.minimum(off[1:], c1) - np.maximum(off[:-1], c0), 0, None)
        agg = (w[:, None].astype(np.float64) * probs).sum(axis=0)
        chosen = [s for s in range(eng.P) if (tsel[t, s // 32] >> (s % 32)) & 1]
        assert len(chosen) == topk
        kth = np.sort(agg)[::-1][topk - 1]
        # fp32 sums in the kernel's order: 1e-4 absolute (sums reach 64)
        assert agg[chosen].min() >= kth - 1e-4
        assert (agg[[s for s in range(eng.P) if s not in chosen]] <= kth + 1e-4).all()


def test_parent_selector_logging_and_solves():
    """Per-parent mode on 20x10 ta011: records hold the parent's features
    and the child's label (the first round's records all carry the root's
    features); learned solves over parent-scored tiles prove the plain
    optimum with and without the full-LB2 fallback."""
    p = oracle.taillard(11)[3]
    with wsbb_amd.Engine(p) as eng:
        eng.selector_init(seed=1)
        eng.selector_config(topk=8, margin=30, parents=True)
        eng.selector_log(stride=4, capacity=100000)
        plain = eng.solve(pool_size=4096, max_batches=10)
        feat, label = eng.selector_records()
        eng.selector_log(stride=0, capacity=0)
        eng.selector_set_records(feat, label)
        assert plain["best"] == 1582 and feat.shape[1] > 1000
        assert ((label >= -1) & (label < eng.P)).all()
        # the root round has 2n = 40 children: records 0..9 (stride 4) carry
        # the root's features, head R0, tail Q0, all the work
        m, n = p.shape
        t = oracle.Oracle(p).tables()
        scale = int(p.max(axis=0).sum() + p.max(axis=1).sum())  # path bound
        rem = p.sum(axis=1)
        root = np.concatenate([t["R0"], t["Q0"], rem, t["R0"] + rem, rem + t["Q0"]])
        for r in range(2 * n // 4):
            assert np.allclose(feat[:, r] * scale, root, rtol=1e-5, atol=1e-3)
        # every record: head + work and work + tail blocks are sums of the
        # first three
        f = feat.reshape(5, m, -1)
        assert np.allclose(f[3], f[0] + f[2], atol=1e-5)
        assert np.allclose(f[4], f[2] + f[1], atol=1e-5)
        losses = eng.selector_train(epochs=4, batch=512, lr=2e-3, seed=2)
        assert losses[-1] < losses[0]
        fb = eng.solve(kind=wsbb_amd.LEARNED, pool_size=4096, max_batches=10)
        # the host keeps a batch of rounds queued behind the one it checks;
        # rounds after the end must leave every counter alone, whatever the
        # batch size (the fallback count included)
        eng.search_options(rounds_per_sync=1)
        fb1 = eng.solve(kind=wsbb_amd.LEARNED, pool_size=4096, max_batches=10)
        eng.search_options(rounds_per_sync=32)
        fb32 = eng.solve(kind=wsbb_amd.LEARNED, pool_size=4096, max_batches=10)
        eng.search_options(rounds_per_sync=8)
        eng.selector_config(topk=8, margin=-1)
        nofb = eng.solve(kind=wsbb_amd.LEARNED, pool_size=4096, max_batches=10)
    for r in (fb, nofb):
        assert r["proven"] and r["best"] == plain["best"]
        assert r["stats"]["subset_bounded"] == r["stats"]["bounded"]
    assert fb1["stats"] == fb["stats"] == fb32["stats"]
    assert fb["stats"]["fallbacks"] > 0
    assert nofb["stats"]["fallbacks"] == 0
    print(f"parents: nodes plain {plain['stats']['bounded']} learned+fallback "
          f"{fb['stats']['bounded']} ({fb['stats']['fallbacks']} fallbacks) learned "
          f"{nofb['stats']['bounded']}")


@pytest.mark.parametrize("tid,N,int16", [(11, 640, 0), (51, 640, 0), (11, 566, 2), (21, 566, 2),
                                         (51, 640, 2)])
def test_tile_subset_bounds_match_oracle(tid, N, int16):
    """Subset bounds per 64-node tile against the oracle, with the two-node
    form allowed (int16=2: it stays with plain LB2) and without; N = 566
    leaves a partial last tile."""
    p = oracle.taillard(tid)[3]
    m, n = p.shape
    o = oracle.Oracle(p)
    permT, l1, l2 = oracle.random_nodes(n, N, 11)
    rng = np.random.default_rng(12)
    with wsbb_amd.Engine(p) as eng:
        eng.set_int16(int16)
        P, nsw = eng.P, (eng.P + 31) // 32
        selb = rng.random(((N + 63) // 64, P)) < 0.15
        selb[:, 0] = True
        tsel = np.zeros(((N + 63) // 64, nsw), np.uint32)
        for s in range(P):
            tsel[:, s // 32] |= selb[:, s].astype(np.uint32) << np.uint32(s % 32)
        full = eng.bound(permT, l1, l2)
        inc = int(np.median(full))
        sub, st0 = eng.bound_tiles(permT, l1, l2, tsel, margin=-1)
        mixed, st1 = eng.bound_tiles(permT, l1, l2, tsel, margin=60, inc=inc)
    for k in range(N):
        perm = permT[:, k].astype(np.int32)
        ref = o.bound_subset(perm, l1[k], l2[k], selb[k // 64].astype(np.uint8))
        assert sub[k] == ref
        flagged = ref < inc and ref >= inc - 60
        assert mixed[k] == (full[k] if flagged else ref)
    assert (sub <= full).all()
    assert st0["nodes"] == N and st0["fallbacks"] == 0
    assert st1["fallbacks"] == int(sum(1 for k in range(N) if inc - 60 <= sub[k] < inc))


def test_logging_training_and_learned_solves():
    """20x10 ta011: records from a plain solve train the selector; learned
    solves (with the full-LB2 fallback) prove the same optimum, and the
    trained network hits the LB2 value more often than a random network."""
    p = oracle.taillard(11)[3]
    with wsbb_amd.Engine(p) as eng:
        eng.selector_init(seed=1)
        eng.selector_log(stride=4, capacity=100000)
        plain = eng.solve(pool_size=4096, max_batches=10)
        feat, label = eng.selector_records()
        eng.selector_log(stride=0, capacity=0)  # stop logging, drop the store
        eng.selector_set_records(feat, label)
        assert plain["best"] == 1582 and feat.shape[1] > 1000
        assert ((label >= -1) & (label < eng.P)).all()
        valid = label >= 0
        # held-out batch for hit rates
        permT, l1, l2 = oracle.random_nodes(20, 4096, 21)
        full = eng.bound(permT, l1, l2)

        def hit_rate():
            fe = eng.bound(permT, l1, l2, kind=wsbb_amd.LB1, feat=True)["feat"]
            _, tsel = eng.selector_infer(fe)
            sub, _ = eng.bound_tiles(permT, l1, l2, tsel, margin=-1)
            return float((sub == full).mean())

        eng.selector_config(topk=8, margin=30)
        h_random = hit_rate()
        losses = eng.selector_train(epochs=4, batch=512, lr=2e-3, seed=2)
        assert losses[-1] < losses[0]
        h_trained = hit_rate()
        assert h_trained > h_random
        learned = eng.solve(kind=wsbb_amd.LEARNED, pool_size=4096, max_batches=10)
    assert learned["proven"] and learned["best"] == plain["best"]
    st = learned["stats"]
    assert st["subset_bounded"] == st["bounded"]
    # with 45 pairs and top-8 tiles the subset saving does not cover the
    # extra nodes and the fallbacks; only report the counts here
    print(f"hit rate random {h_random:.3f} trained {h_trained:.3f}; "
          f"pair evaluations learned {st['pairs_evaluated']} plain "
          f"{plain['stats']['pairs_evaluated']}; nodes learned {st['bounded']} plain "
          f"{plain['stats']['bounded']}; fallbacks {st['fallbacks']}")


@pytest.mark.parametrize("parents", [False, True])
def test_trained_selector_gains_on_20x20(parents):
    """What the trained network gains over the method without it, on ta021
    (20 jobs x 20 machines, 190 pairs): trained on records of a plain solve
    of the instance, top-16 pairs per tile, the learned search proves the
    same optimum with a small node inflation while evaluating a fraction of
    the pairs; the untrained network (same search, random weights) needs far
    more nodes.  Per child and per parent (the network scores the round's
    parents)."""
    p = oracle.taillard(21)[3]
    pool, nb = 65536, 4
    with wsbb_amd.Engine(p) as eng:
        eng.selector_init(seed=1)
        eng.selector_config(topk=16, margin=-1, parents=parents)
        untrained = eng.solve(kind=wsbb_amd.LEARNED, pool_size=pool, max_batches=nb,
                              budget=40_000_000)
        eng.selector_log(stride=64, capacity=1 << 19)
        plain = eng.solve(pool_size=pool, max_batches=nb)
        feat, label = eng.selector_records()
        eng.selector_log(stride=0, capacity=0)
        eng.selector_set_records(feat, label)
        eng.selector_train(epochs=4, batch=1024, lr=2e-3, seed=2)
        learned = eng.solve(kind=wsbb_amd.LEARNED, pool_size=pool, max_batches=nb)
    assert plain["proven"] and learned["proven"]
    assert learned["best"] == plain["best"] == 2297
    ps, ls, us = plain["stats"], learned["stats"], untrained["stats"]
    assert ls["bounded"] < 1.25 * ps["bounded"]
    assert ls["pairs_evaluated"] < 0.2 * ps["pairs_evaluated"]
    assert us["bounded"] > 1.5 * ls["bounded"] or not untrained["proven"]
    print(f"plain {ps['bounded']} nodes {ps['pairs_evaluated']} pairs; learned {ls['bounded']} "
          f"nodes {ls['pairs_evaluated']} pairs; untrained {us['bounded']} nodes")


def test_online_training_resumes_to_the_optimum():
    """Online learning on ta024 (20x20): a plain solve logs records for its
    first 2e6 bounded children and stops, the search is checkpointed, the
    network is trained from scratch on those records and the search resumes
    with learned selection (top-24 pairs per tile, no fallback).  The resumed
    search proves the optimum plain LB2 proves, with a small node
    inflation, and evaluates a fraction of the pairs after the switch."""
    p = oracle.taillard(24)[3]
    pool, nb, cap = 65536, 4, 1 << 17
    with wsbb_amd.Engine(p) as eng:
        plain = eng.solve(pool_size=pool, max_batches=nb)
        eng.selector_init(seed=1)
        eng.selector_config(topk=24, margin=-1, parents=True)
        eng.selector_log(stride=max(1, 2_000_000 * 20 // cap), capacity=cap)
        first = eng.solve(pool_size=pool, max_batches=nb, budget=2_000_000)
        assert not first["proven"]
        ck = eng.checkpoint()
        feat, label = eng.selector_records()
        assert 0 < feat.shape[1] <= cap
        assert ((label >= -1) & (label < eng.P)).all() and (label >= 0).any()
        eng.selector_set_records(feat, label)
        losses = eng.selector_train(epochs=2, batch=2048, lr=2e-3, seed=2)
        assert losses[-1] < losses[0]
        eng.selector_log(0, 0)
        res = eng.resume(ck, kind=wsbb_amd.LEARNED, pool_size=pool, max_batches=nb)
    assert plain["proven"] and res["proven"]
    assert res["best"] == plain["best"]
    assert oracle.Oracle(p).cmax(res["perm"]) == res["best"]
    ps, rs, fs = plain["stats"], res["stats"], first["stats"]
    assert rs["bounded"] < 1.3 * ps["bounded"]
    after = rs["pairs_evaluated"] - fs["pairs_evaluated"]
    assert after < 0.3 * 190 * (rs["bounded"] - fs["bounded"])
    print(f"plain {ps['bounded']} nodes; online {rs['bounded']} nodes, "
          f"{after / max(rs['bounded'] - fs['bounded'], 1):.1f} pairs per node after the switch")
