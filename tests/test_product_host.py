"""Host-side product checks that need no GPU: the library loads and exports
every symbol declared in include/wsbb_amd.h, its instance precompute and
generators match the oracle bit for bit, and GPU entry points fail cleanly
without a device."""
import os
import re

import numpy as np
import pytest

import oracle
import wsbb_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "wsbb_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(wsbb_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(wsbb_amd.lib, name), name


def instances():
    rng = np.random.default_rng(0)
    out = [oracle.taillard(t)[3] for t in (1, 11, 21, 51, 81)]
    out.append(oracle.generate_correlated(50, 10, 1234))
    out.append(rng.integers(1, 4, size=(6, 30)).astype(np.int32))   # heavy ties
    out.append(rng.integers(0, 2, size=(4, 17)).astype(np.int32))   # zeros and ties
    out.append(rng.integers(1, 100, size=(24, 128)).astype(np.int32))
    return out


@pytest.mark.parametrize("idx", range(9))
def test_precompute_matches_oracle(idx):
    p = instances()[idx]
    mine = wsbb_amd.precompute(p)
    ref = oracle.Oracle(p).tables()
    for k in ("ma", "mb", "lag", "johnson", "R0", "Q0"):
        assert np.array_equal(mine[k], ref[k]), k


def test_taillard_generator_matches_oracle():
    for tid in range(1, 121):
        assert np.array_equal(wsbb_amd.taillard(tid), oracle.taillard(tid)[3]), tid


def test_other_generators_match_oracle():
    for n, m, seed in [(20, 5, 1), (50, 20, 77), (100, 10, 2**40 + 3)]:
        assert np.array_equal(wsbb_amd.generate_correlated(n, m, seed),
                              oracle.generate_correlated(n, m, seed))
        mc = wsbb_amd.generate_machine_correlated(n, m, seed)
        assert np.array_equal(mc, oracle.generate_machine_correlated(n, m, seed))
        # a machine's times lie within 20 of each other, in [1, 100]
        assert (mc.max(axis=1) - mc.min(axis=1) <= 20).all() and mc.min() >= 1 and mc.max() <= 100
    for n, N, s0 in [(20, 1000, 0x5EED), (50, 4096, 0x5EED), (100, 777, 9)]:
        a = wsbb_amd.random_nodes(n, N, s0)
        b = oracle.random_nodes(n, N, s0)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        permT, l1, l2 = a
        assert (np.sort(permT, axis=0) == np.arange(n)[:, None]).all()
        d1, d2 = l1 + 1, n - l2
        assert (d1 >= 0).all() and (d2 >= 0).all() and (d1 + d2 <= n - 1).all()


def test_precompute_rejects_out_of_range_instances():
    with pytest.raises(wsbb_amd.WsbbError):
        wsbb_amd.precompute(np.ones((5, 129), np.int32))
    with pytest.raises(wsbb_amd.WsbbError):
        wsbb_amd.precompute(np.ones((25, 20), np.int32))
    with pytest.raises(wsbb_amd.WsbbError):
        wsbb_amd.precompute(np.full((5, 20), 70000, np.int32))


@pytest.mark.skipif(wsbb_amd.device_count() > 0, reason="a GPU is present")
def test_create_without_gpu_fails_cleanly():
    with pytest.raises(wsbb_amd.WsbbError) as e:
        wsbb_amd.Engine(oracle.taillard(1)[3])
    assert e.value.code == -6


# entry points that take no context, or accept a null one by design
_NO_CTX = {"wsbb_version", "wsbb_last_error", "wsbb_device_count", "wsbb_taillard",
           "wsbb_generate", "wsbb_generate_correlated", "wsbb_generate_machine_correlated",
           "wsbb_random_nodes", "wsbb_precompute", "wsbb_create", "wsbb_destroy",
           "wsbb_multi_create", "wsbb_multi_destroy"}


def test_every_context_entry_point_refuses_a_null_context():
    """Called with a null context (and null or zero arguments otherwise),
    every entry point that needs one returns WSBB_EINVAL with a message and
    touches nothing; no GPU is needed to reach that check."""
    import ctypes
    names = [n for n in header_functions() if n not in _NO_CTX]
    assert len(names) >= 40
    for name in names:
        fn = getattr(wsbb_amd.lib, name)
        args = []
        for t in fn.argtypes:
            if t in (ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64,
                     ctypes.c_size_t, ctypes.c_long, ctypes.c_float):
                args.append(t(0))
            elif t is wsbb_amd._TRACE_FN:
                args.append(wsbb_amd._TRACE_FN())
            else:
                args.append(None)
        assert fn(*args) == -1, name
        assert wsbb_amd.lib.wsbb_last_error(), name


def test_create_rejects_bad_instances_before_any_device():
    """Shapes and times outside the documented limits are refused with
    WSBB_EINVAL and the reason, whether or not a GPU is present."""
    import ctypes
    h = ctypes.c_void_p()

    def create(p, n, m):
        p = np.ascontiguousarray(p, np.int32)
        return wsbb_amd.lib.wsbb_create(p.ctypes.data_as(wsbb_amd._i32p), n, m, 0, ctypes.byref(h))

    ok = np.ones((5, 20), np.int32)
    for p, n, m in ((np.ones((5, 1), np.int32), 1, 5), (np.ones((5, 129), np.int32), 129, 5),
                    (np.ones((1, 20), np.int32), 20, 1), (np.ones((25, 20), np.int32), 20, 25),
                    (np.where(np.arange(100).reshape(5, 20) == 7, -1, ok), 20, 5),
                    (np.where(np.arange(100).reshape(5, 20) == 7, 65536, ok), 20, 5),
                    (np.full((24, 128), 5000, np.int32), 128, 24)):  # path bound > 2^19
        assert create(p, n, m) == -1
        assert wsbb_amd.lib.wsbb_last_error()
        assert not h.value
    assert wsbb_amd.lib.wsbb_create(None, 20, 5, 0, ctypes.byref(h)) == -1
    for bad in (0, 121):
        assert wsbb_amd.lib.wsbb_taillard(bad, None, None, None) == -1
    z = np.zeros(8, np.int32)
    zp = z.ctypes.data_as(wsbb_amd._i32p)
    assert wsbb_amd.lib.wsbb_generate(0, 5, 1, zp) == -1
    assert wsbb_amd.lib.wsbb_generate_correlated(5, 0, 1, zp) == -1
    assert wsbb_amd.lib.wsbb_generate_machine_correlated(5, 5, 1, None) == -1
    assert wsbb_amd.lib.wsbb_random_nodes(1, 4, 0, None, zp, zp) == -1
    assert wsbb_amd.lib.wsbb_precompute(None, 20, 5, None, None, None, None, None, None) == -1
