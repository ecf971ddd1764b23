"""The host mesh loader (vox_load_mesh, cuda_voxelizer_amd/csrc/mesh_io.cpp)
under AddressSanitizer + UndefinedBehaviorSanitizer, on seeded mutations of
small OBJ, OFF, STL and PLY files (tests/cpp/loader_fuzz.cpp).  CPU only: the
loader has no device code, and the driver links it alone with g++."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_loader_fuzz_under_asan_ubsan(tmp_path):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the HIP headers not available")
    exe = str(tmp_path / "loader_fuzz")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined,float-cast-overflow",
           "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           os.path.join(ROOT, "tests", "cpp", "loader_fuzz.cpp"),
           os.path.join(ROOT, "cuda_voxelizer_amd", "csrc", "mesh_io.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe, "40000", str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and "0 failed" in r.stdout, r.stdout + r.stderr[-4000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
