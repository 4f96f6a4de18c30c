"""Text formats of the C++ plugin face (include/wsbb_plugins.hpp): a node
record, a bab.solutions block and the formater CSV row must read exactly as
the reference writes them (raw_bb_problem operator<<,
bb_problem_calculator.cpp:124-144; solutions operator<<, solutions.cpp:88-95;
formater.cpp:94-116), and every written form must read back unchanged."""
import os
import subprocess

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
    assert lines[5] == "ok"
