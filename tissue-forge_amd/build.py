"""Builds libtfg.so (HIP device module + C ABI + host engine) in-tree for gfx950.

Usage: python tissue-forge_amd/build.py
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libtfg.so")

SOURCES = [
    os.path.join(HERE, "csrc", "device", "ensemble.hip"),
    os.path.join(HERE, "csrc", "device", "bin.hip"),
    os.path.join(HERE, "csrc", "device", "nonbond.hip"),
    os.path.join(HERE, "csrc", "device", "advance.hip"),
    os.path.join(HERE, "csrc", "device", "noalloc.hip"),
    os.path.join(HERE, "csrc", "host", "capi.cpp"),
    os.path.join(HERE, "csrc", "host", "engine.cpp"),
    os.path.join(HERE, "csrc", "host", "potential.cpp"),
    os.path.join(HERE, "csrc", "host", "checkpoint.cpp"),
]
OBJDIR = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Packed FP32 VALU forms stay off on this target.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]

# hipcc's host pass reports the device-only target feature as unknown; that
# line is expected and carries no information.
_NOISE = "is not a recognized feature for this target"


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (res.stdout + res.stderr).splitlines() if _NOISE not in l]
    return res.returncode, msgs


def build(verbose=False, extra_flags=()):
    """Compiles every source (in parallel) and links tissue-forge_amd/libtfg.so;
    raises on failure."""
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
        jobs.append((src, obj, [HIPCC] + FLAGS + list(extra_flags) + ["-c", src, "-o", obj]))
    if verbose:
        for _, _, cmd in jobs:
            print(" ".join(cmd), flush=True)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as pool:
        results = list(pool.map(lambda j: _run(j[2]), jobs))
    for (src, _, _), (rc, msgs) in zip(jobs, results):
        if rc != 0:
            raise RuntimeError("hipcc failed on %s:\n%s" % (src, "\n".join(msgs)))
        if verbose and msgs:
            print("\n".join(msgs))
    tmp = LIB + ".tmp"
    rc, msgs = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + [o for _, o, _ in jobs] + ["-o", tmp])
    if rc != 0:
        raise RuntimeError("link failed:\n" + "\n".join(msgs))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
