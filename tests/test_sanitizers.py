"""The host code and the oracle under AddressSanitizer + UndefinedBehaviorSanitizer.

Two sanitizer builds (no GPU needed): libcellspe_fx_asan.so -- the library's
host code (C-ABI, the task runtime's unit threads, the ingest and frame-header
decoders, the spectrometer and minicorrelator front ends) instrumented on the
host pass only, with the normal GPU code -- and liboracle_fx_asan.so, the
oracle.  Each runs a share of the CPU suite in a child process with the shared
clang sanitizer runtime preloaded (Python itself is not instrumented), UB
fatal and leak checking off (the interpreter's own allocations are not ours):
the host-ABI and task-runtime tests against the sanitized library, the
oracle's pinning and frame-decoder tests against the sanitized oracle.  A
sanitizer report ends the child with a non-zero status.
"""
import glob
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cellspe-tasklib_amd")
LIB = os.path.join(PKG, "libcellspe_fx_asan.so")
ORACLE = os.path.join(ROOT, "oracle", "liboracle_fx_asan.so")


def _runtime():
    found = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return found[-1] if found else None


def _run(env_extra, args):
    rt = _runtime()
    if rt is None:
        pytest.skip("no shared clang sanitizer runtime in this image")
    env = dict(os.environ)
    env.update(env_extra)
    env["LD_PRELOAD"] = rt
    env["LD_LIBRARY_PATH"] = "/opt/rocm/llvm/lib:" + env.get("LD_LIBRARY_PATH", "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["PYTHONDONTWRITEBYTECODE"] = "1"
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-p", "no:xdist",
           "-m", "not gpu"] + args
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in tail and "runtime error:" not in tail, tail
    return tail


def _passed(tail):
    m = re.search(r"(\d+) passed", tail)
    return int(m.group(1)) if m and " failed" not in tail else 0


def _build(target, directory, path):
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", directory, target], check=True, capture_output=True)
    return path


def test_host_code_under_sanitizers():
    lib = _build("asan", PKG, LIB)
    # the runtime and error-path tests (the ones that compile C with gcc run
    # in the normal suite: the preloaded runtime would also enter the compiler)
    tail = _run({"CELLSPE_FX_LIB": lib},
                [os.path.join("tests", "test_host_abi.py"), "-k",
                 "not compiles and not layouts and not exports and not parser and not reference_declarations"])
    assert _passed(tail) >= 10, tail
    tail = _run({"CELLSPE_FX_LIB": lib}, [os.path.join("tests", "test_frames_oracle.py"), "-k", "library_decoders"])
    assert _passed(tail) == 1, tail


def test_oracle_under_sanitizers():
    orc = _build("asan", os.path.join(ROOT, "oracle"), ORACLE)
    tail = _run({"FX_ORACLE_LIB": orc}, [os.path.join("tests", "test_oracle.py"),
                                         os.path.join("tests", "test_frames_oracle.py")])
    assert _passed(tail) >= 40, tail


def test_sanitizer_harness_catches_an_overflow():
    """The harness itself: the sanitized oracle asked to unpack more samples
    than its output buffer holds is stopped by AddressSanitizer (the child
    exits non-zero with a heap-buffer-overflow report)."""
    orc = _build("asan", os.path.join(ROOT, "oracle"), ORACLE)
    rt = _runtime()
    if rt is None:
        pytest.skip("no shared clang sanitizer runtime in this image")
    code = ("import ctypes as C, numpy as np\n"
            f"lib = C.CDLL({orc!r})\n"
            "w = np.zeros(64, np.uint32)\n"
            "out = np.zeros(16, np.float32)\n"   # 1024 samples wanted
            "lib.fxo_unpack2b_1ch(w.ctypes.data_as(C.c_void_p), C.c_size_t(1024), out.ctypes.data_as(C.c_void_p))\n")
    env = dict(os.environ, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0",
               LD_LIBRARY_PATH="/opt/rocm/llvm/lib:" + os.environ.get("LD_LIBRARY_PATH", ""))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr, r.stderr[-2000:]
