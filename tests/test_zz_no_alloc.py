"""wsbb_bound_batch_noalloc against wsbb_bound_batch through the C ABI: two
calls in a row on the library's fixed staging buffer, the second smaller,
each bit-identical to the allocating entry point, and a batch too large for
the buffer refused with the limit in the message."""
import numpy as np
import pytest

import wsbb_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tid", [11, 51])
def test_noalloc_matches_allocating_entry_point(tid):
    p = wsbb_amd.taillard(tid)
    n = p.shape[1]
    with wsbb_amd.Engine(p) as eng:
        for N, seed, kind in ((40000, 0x5EED, wsbb_amd.LB2), (777, 0xBEEF, wsbb_amd.LB2),
                              (3001, 0xCAFE, wsbb_amd.LB1)):
            permT, l1, l2 = wsbb_amd.random_nodes(n, N, seed)
            ref = eng.bound(permT, l1, l2, kind=kind)
            got = eng.bound_noalloc(permT, l1, l2, kind=kind)
            assert np.array_equal(got, ref), (N, np.flatnonzero(got != ref)[:5])
        # the `best` early exit passes through unchanged
        permT, l1, l2 = wsbb_amd.random_nodes(n, 5000, 7)
        ref = eng.bound(permT, l1, l2, best=1000)
        assert np.array_equal(eng.bound_noalloc(permT, l1, l2, best=1000), ref)


def test_noalloc_refuses_batches_beyond_its_buffer():
    p = wsbb_amd.taillard(51)
    n = p.shape[1]
    N = (8 << 20) // (n + 12) + 1000
    permT, l1, l2 = wsbb_amd.random_nodes(n, N, 3)
    with wsbb_amd.Engine(p) as eng:
        with pytest.raises(wsbb_amd.WsbbError) as ei:
            eng.bound_noalloc(permT, l1, l2)
        assert ei.value.code == -1 and "fixed buffer" in str(ei.value)
        small = eng.bound_noalloc(permT[:, :100], l1[:100], l2[:100])
        assert np.array_equal(small, eng.bound(permT[:, :100], l1[:100], l2[:100]))
