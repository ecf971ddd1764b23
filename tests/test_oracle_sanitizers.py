"""The CPU oracle under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer (oracle/sanitize.mk, driver oracle/sanitize_main.c).

The driver runs every oracle routine over boundary inputs (clamped voxel
boxes, out-of-row solid runs clipped at the table end, degenerate triangles,
every writer into a buffer of exactly its queried size) and checks that one
thread, four threads and the lock-free mode give the same tables: the
reference merges bits with commutative OR/XOR under a critical section
(cpu_voxelizer.cpp:7-15,186-194), so any difference would be a race.  The
ThreadSanitizer build uses LLVM's libomp with its Archer tool loaded, which
gives TSan OpenMP's barrier and critical-section ordering."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MK = os.path.join(ROOT, "oracle", "sanitize.mk")
OUT = os.path.join(ROOT, "oracle", "_san")
LLVM = os.environ.get("LLVM", "/opt/rocm/llvm")


def _build(target):
    r = subprocess.run(["make", "-s", "-f", MK, target], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def _run(exe, env_extra):
    env = dict(os.environ, **env_extra)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    return r.returncode, r.stdout, r.stderr


def test_oracle_under_asan_ubsan():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    _build("asan")
    rc, out, err = _run(os.path.join(OUT, "oracle_asan"),
                        {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1:abort_on_error=0",
                         "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1", "OMP_NUM_THREADS": "4"})
    assert rc == 0 and ", 0 failed" in out, out + err[-4000:]
    assert "Sanitizer" not in err and "runtime error" not in err, err[-4000:]


def test_oracle_under_tsan_with_archer():
    archer = os.path.join(LLVM, "lib", "libarcher.so")
    clang = os.path.join(LLVM, "bin", "clang")
    if not (os.path.exists(archer) and os.path.exists(clang)):
        pytest.skip("LLVM clang with libomp/Archer not available")
    _build("tsan")
    rc, out, err = _run(os.path.join(OUT, "oracle_tsan"),
                        {"OMP_TOOL_LIBRARIES": archer, "ARCHER_OPTIONS": "verbose=1",
                         "TSAN_OPTIONS": "halt_on_error=1:exitcode=66:ignore_noninstrumented_modules=1"})
    assert "Archer detected OpenMP application with TSan" in out + err, err[-2000:]  # the tool is active
    assert rc == 0 and ", 0 failed" in out, out + err[-4000:]
    assert "ThreadSanitizer" not in err, err[-4000:]
