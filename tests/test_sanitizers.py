"""CPU builds under AddressSanitizer and UndefinedBehaviorSanitizer (any
report is fatal) of the oracle, the product's host-side instance code and
the C++ face's format harness, each driven through its entry points."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]


def build_and_run(tmp_path, name, cmd):
    exe = str(tmp_path / name)
    r = subprocess.run(cmd + ["-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert out.stdout.strip().endswith("ok"), out.stdout
    assert "runtime error" not in out.stderr and "ERROR: AddressSanitizer" not in out.stderr


def test_oracle_under_sanitizers(tmp_path):
    build_and_run(tmp_path, "san_oracle",
                  ["gcc", "-std=c11", "-pthread"] + SAN +
                  [os.path.join(ROOT, "oracle", "oracle.c"),
                   os.path.join(ROOT, "tests", "cpp", "sanitize_oracle.c")])


def test_instance_code_under_sanitizers(tmp_path):
    build_and_run(tmp_path, "san_instance",
                  ["g++", "-std=c++17"] + SAN +
                  [os.path.join(ROOT, "wsbb_amd", "csrc", "instance.cpp"),
                   os.path.join(ROOT, "tests", "cpp", "sanitize_instance.cpp")])


def test_plugin_formats_under_sanitizers(tmp_path):
    build_and_run(tmp_path, "san_formats",
                  ["g++", "-std=c++17", "-I", os.path.join(ROOT, "include")] + SAN +
                  [os.path.join(ROOT, "tests", "cpp", "format_check.cpp")])
