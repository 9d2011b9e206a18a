"""Host-side C++ under AddressSanitizer and UBSan, built with g++ (host code
only, no GPU), finishing without a sanitizer report:
  * the engine's host routines (em_host.cpp) through tests/cpp/host_sanitized.cpp;
  * the C++ simulator surface's header code (atom helpers, text / raw I/O,
    parameter accessors) through the CPU modes of tests/cpp/test_simulators.cpp."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_under_asan_and_ubsan(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "host_sanitized"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "cudaem_amd", "csrc"),
           os.path.join(ROOT, "tests", "cpp", "host_sanitized.cpp"),
           os.path.join(ROOT, "cudaem_amd", "csrc", "em_host.cpp"), "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True)
    if b.returncode != 0 and ("asan" in b.stderr.lower() or "ubsan" in b.stderr.lower()):
        pytest.skip("sanitizer runtime not available: " + b.stderr[-200:])
    assert b.returncode == 0, b.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    assert r.returncode == 0, r.stderr[-3000:]
    assert "done" in r.stdout


def test_simulator_surface_cpu_modes_under_asan_and_ubsan(tmp_path):
    gxx = shutil.which("g++")
    lib = os.path.join(ROOT, "cudaem_amd", "libcudaem_amd.so")
    if gxx is None or not os.path.exists(lib):
        pytest.skip("g++ or the built library not available")
    exe = tmp_path / "test_simulators_san"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "test_simulators.cpp"),
           "-L", os.path.dirname(lib), "-lcudaem_amd", "-Wl,-rpath," + os.path.dirname(lib), "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True)
    if b.returncode != 0 and ("asan" in b.stderr.lower() or "ubsan" in b.stderr.lower()):
        pytest.skip("sanitizer runtime not available: " + b.stderr[-200:])
    assert b.returncode == 0, b.stderr[-2000:]
    (tmp_path / "atoms.txt").write_text("# c\n38 0 0 0\n22 1.9525 1.9525 1.9525 0.5\n8 1.9525 1.9525 0\n")
    # the HIP runtime the library links keeps allocations to process exit
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    for mode in ("io", "cut", "params", "vibrate"):
        r = subprocess.run([str(exe), mode, str(tmp_path)], capture_output=True, text=True, timeout=120, env=env)
        assert r.returncode == 0, (mode, r.stderr[-3000:])
