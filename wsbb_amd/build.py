"""Builds wsbb_amd/libwsbb_amd.so (HIP for gfx950) with hipcc, in-tree."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwsbb_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
SOURCES = ["instance.cpp", "bound.hip", "search.hip", "selector.hip", "multi.hip", "capi.hip"]
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Wno-unused-result",
    # no packed-FP32 VALU ops anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + (["-x", "hip"] if src.endswith(".hip") else []) + [
        "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msgs}")
    return obj, msgs


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    for obj, msgs in results:
        if verbose and msgs.strip():
            print(msgs)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_plugin_example()
    return LIB


def build_plugin_example():
    """tests/cpp/plugin_explorer: the reference's DFS loop over the C++
    plugin face (include/wsbb_plugins.hpp), linked against the library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "plugin_explorer.cpp")
    out = os.path.join(root, "tests", "cpp", "plugin_explorer")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(root, "include"), src,
           "-o", out, "-L", HERE, "-lwsbb_amd", "-Wl,-rpath,$ORIGIN/../../wsbb_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("plugin example build failed:\n" + r.stdout + r.stderr)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
