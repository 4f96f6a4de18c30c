"""Text formats of the C++ plugin face (include/wsbb_plugins.hpp): a node
record, a bab.solutions block and the formater CSV row must read exactly as
the reference writes them (raw_bb_problem operator<<,
bb_problem_calculator.cpp:124-144; solutions operator<<, solutions.cpp:88-95;
formater.cpp:94-116), and every written form must read back unchanged."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_text_formats(tmp_path):
    exe = str(tmp_path / "format_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "format_check.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.split("\n")
    perm = " ".join(str((j * 7) % 20) for j in range(20))
    rec = f"1278 0 1278 {perm} 3 17 7 14 1 "
    assert lines[0] == rec
    assert lines[1] == "2 "
    assert lines[2] == rec
    assert lines[3] == f"1290 0 1290 {perm} 3 17 7 3 1 "
    assert lines[4] == "11;1582;4295624;0.5;"
    # the first binary record, as the Python codec writes it
    from wsbb_amd import records
    r = np.zeros(1, dtype=records.record_dtype())
    r["permutation"][0] = [(j * 7) % 20 for j in range(20)]
    r["limite1"], r["limite2"], r["couts_somme"], r["depth"], r["begin_end"] = 3, 17, 1278, 7, 1
    r["ranks"][0] = [(i * 3) % 7 for i in range(21)]
    assert len(records.encode(r)) == 184
    assert lines[5] == records.encode(r).hex()
    assert lines[6] == "ok"


def test_binary_records_round_trip_a_pool():
    """A device pool written as the reference's binary node records and
    read back gives the same nodes in the same order, for a shuffled
    message too (the rank strings carry the pop order)."""
    from wsbb_amd import records
    rng = np.random.default_rng(3)
    n, c = 20, 57
    ck = dict(permT=np.stack([rng.permutation(n) for _ in range(c)], axis=1).astype(np.uint8),
              l1=rng.integers(-1, 8, c).astype(np.int32), l2=rng.integers(12, 21, c).astype(np.int32),
              cost=rng.integers(1000, 2000, c).astype(np.int32))
    recs = records.from_pool(ck)
    data = records.encode(recs)
    assert len(data) == 184 * c
    back = records.decode(data)
    shuffled = back[rng.permutation(c)]
    for got in (records.to_pool(back, n), records.to_pool(shuffled, n)):
        for k in ("permT", "l1", "l2", "cost"):
            assert np.array_equal(got[k], ck[k]), k
    assert (back["depth"] == (ck["l1"] + 1) + (n - ck["l2"])).all()
