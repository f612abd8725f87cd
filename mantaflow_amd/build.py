"""Build libmfx.so (the HIP kernels + C ABI) for gfx950 in-tree.

Usage: python -m mantaflow_amd.build [--jobs N]

Every .hip file under csrc/ is compiled to an object with hipcc, then linked
into mantaflow_amd/libmfx.so.  Flags:
  -ffp-contract=off      no fused multiply-add: the path reproduces the
                         reference's float operation order bit for bit
  -target-feature -packed-fp32-ops
                         keep packed FP32 instructions out of the code
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libmfx.so")
ARCH = os.environ.get("MFX_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-value", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OBJ, os.path.basename(s)[:-4] + ".o") for s in srcs]

    def compile_one(pair):
        s, o = pair
        _run([HIPCC] + CFLAGS + ["-c", s, "-o", o])
        if verbose:
            print("compiled", os.path.basename(s), flush=True)
        return o

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("linked", LIB, flush=True)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    build(a.jobs, a.verbose)


if __name__ == "__main__":
    sys.exit(main())
