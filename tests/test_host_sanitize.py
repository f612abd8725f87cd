"""The library's host code under AddressSanitizer and
UndefinedBehaviorSanitizer: every .hip source is built with the sanitizers
on its host side only (the device code is compiled as usual, at -O1), and
the CPU tests of the host logic (argument parsing, defaults, error paths
without a device, the serving engine's threads, queues and refusals) run
in a child process against that build with the sanitizer runtime
preloaded; any report fails the run.  CPU only."""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mantaflow_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O1", "-std=c++17", "-fPIC", "-fno-omit-frame-pointer",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]


def _runtime():
    libs = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return libs[-1] if libs else None


def test_host_code_under_asan_and_ubsan(tmp_path):
    rt = _runtime()
    if not os.path.exists(HIPCC) or rt is None:
        pytest.skip("hipcc or the clang sanitizer runtime is not installed")
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))

    def compile_one(src):
        obj = str(tmp_path / (os.path.basename(src)[:-4] + ".o"))
        r = subprocess.run([HIPCC] + FLAGS + ["-c", src, "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    lib = str(tmp_path / "libmfx_asan.so")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fsanitize=address", "-fsanitize=undefined",
                        "-o", lib] + objs, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, MFX_LIB=lib, LD_PRELOAD=rt,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    tests = ["tests/test_host.py::" + t for t in (
        "test_library_exports_every_declared_symbol", "test_pressure_defaults_match_reference",
        "test_errors_surface_as_messages_without_a_device", "test_serve_engine_host_logic")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] + tests,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "4 passed" in out, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
