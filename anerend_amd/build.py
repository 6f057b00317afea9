"""Builds libane_ddgi.so (gfx950) in-tree with hipcc.

Usage: python -m anerend_amd.build [--quiet]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libane_ddgi.so")
SOURCES = ["api.hip", "bvh.hip", "trace.hip", "conv.hip", "translate.hip", "serve.hip", "sungrid.hip"]

# Float contract: no contraction, correctly rounded divide/sqrt, denormals kept.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def build(quiet=False, out=None, defines=()):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    lib = out or LIB
    tmp = lib + ".tmp"
    cmd = [hipcc()] + FLAGS + [f"-D{d}" for d in defines] + ["-o", tmp] + srcs
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = res.stdout + res.stderr
    # the host pass does not know the packed-fp32 feature; that warning is expected
    lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
    if res.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError("hipcc failed building libane_ddgi.so")
    if lines and not quiet:
        sys.stderr.write("\n".join(lines) + "\n")
    os.replace(tmp, lib)
    return lib


HOST = os.path.join(HERE, "host")
HOST_BIN = os.path.join(HOST, "ddgi_frames")
HOST_SOURCES = ["ddgi_passes.cpp", "ddgi_frames.cpp"]


def build_host(quiet=False):
    """C++ frame-graph driver over the C-ABI (plain g++, links libane_ddgi.so)."""
    cxx = os.environ.get("CXX", "g++")
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES]
    tmp = HOST_BIN + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-o", tmp] + srcs + [
        "-L" + HERE, "-lane_ddgi", "-Wl,-rpath,$ORIGIN/.."]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building ddgi_frames")
    if (res.stdout or res.stderr) and not quiet:
        sys.stderr.write(res.stdout + res.stderr)
    os.replace(tmp, HOST_BIN)
    return HOST_BIN


if __name__ == "__main__":
    print(build(quiet="--quiet" in sys.argv))
    print(build_host(quiet="--quiet" in sys.argv))
