"""Builds libtfg.so (HIP device module + C ABI + host engine) in-tree for gfx950.

Usage: python tissue-forge_amd/build.py [--sanitize address|undefined]

--sanitize builds a host-instrumented copy (device code unchanged) into
tissue-forge_amd/_build_<name>/libtfg.so, for the CPU test suite:

  TFG_LIB=tissue-forge_amd/_build_address/libtfg.so ASAN_OPTIONS=detect_leaks=0 \
  LD_PRELOAD=$(/opt/rocm/llvm/bin/clang -print-file-name=libclang_rt.asan-x86_64.so) \
  python -m pytest tests -m "not gpu"

(the sanitizer runtime is clang's, preloaded because Python itself is not
instrumented; with undefined, libclang_rt.ubsan_standalone-x86_64.so).
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


def build(verbose=False, extra_flags=(), sanitize=None):
    """Compiles every source (in parallel) and links tissue-forge_amd/libtfg.so
    (with sanitize, a host-instrumented copy in _build_<sanitize>/); raises on
    failure."""
    objdir, lib, link_flags = OBJDIR, LIB, []
    flags = FLAGS + list(extra_flags)
    if sanitize:
        objdir = os.path.join(HERE, "_build_" + sanitize)
        lib = os.path.join(objdir, "libtfg.so")
        # host code only: the GPU code of this pool is never built instrumented
        san = ["-Xarch_host", "-fsanitize=" + sanitize]
        flags = [f if f != "-O3" else "-O1" for f in flags] + ["-g"] + san + ["-Xarch_host", "-fno-omit-frame-pointer"]
        link_flags = san
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        jobs.append((src, obj, [HIPCC] + flags + ["-c", src, "-o", obj]))
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
    tmp = lib + ".tmp"
    rc, msgs = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + link_flags + [o for _, o, _ in jobs] +
                    ["-o", tmp])
    if rc != 0:
        raise RuntimeError("link failed:\n" + "\n".join(msgs))
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    san = None
    if "--sanitize" in sys.argv:
        san = sys.argv[sys.argv.index("--sanitize") + 1]
        if san not in ("address", "undefined"):
            sys.exit("--sanitize takes address or undefined")
    print(build(verbose=True, sanitize=san))
    sys.exit(0)
