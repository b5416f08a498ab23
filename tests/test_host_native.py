"""Native host tests on the CPU (no GPU): the C++ tests under tests/cpp/,
each built with g++ and run three ways: plainly, under ThreadSanitizer, and
under AddressSanitizer with UndefinedBehaviorSanitizer.

  * test_workspace_cache.cpp  WorkspaceCache with fake workspaces (the
    reference's tests/cpp/test_workspace_cache.cpp:36-131 behaviours);
  * test_device.cpp           device selection, device guard and error
    mapping against a fake HIP runtime (tests/cpp/test_cuda_device.cpp:9-101);
  * test_engine.cpp           the host engine (thread pool, contour stages,
    two-workspace pipeline, failure handling) against a CPU stand-in of the
    backend's C ABI (fake_backend.cpp);
  * test_host_copy.cpp        the backend's parallel staging copy (its thread
    pool, concurrent callers, finiteness flags).
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CPP = ROOT / "tests" / "cpp"
CSRC = ROOT / "harvest-cuda_amd" / "csrc"
INCLUDE = ROOT / "include"
SOURCES = {
    "test_workspace_cache": [CPP / "test_workspace_cache.cpp"],
    "test_device": [CPP / "test_device.cpp"],
    "test_engine": [CPP / "test_engine.cpp", CPP / "fake_backend.cpp", CSRC / "engine.cpp"],
    "test_host_copy": [CPP / "test_host_copy.cpp", CSRC / "host_copy.cpp"],
}
MODES = {
    "plain": ["-O2"],
    "tsan": ["-O1", "-fsanitize=thread"],
    "asan": ["-O1", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
             "-fno-sanitize-recover=undefined"],
}


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("name", sorted(SOURCES))
def test_native(tmp_path, name, mode):
    exe = tmp_path / name
    cmd = ["g++", "-std=c++17", "-g", "-pthread", "-ffp-contract=off", *MODES[mode],
           f"-I{CSRC}", f"-I{INCLUDE}", "-o", str(exe), *map(str, SOURCES[name])]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = {"TSAN_OPTIONS": "halt_on_error=1 second_deadlock_stack=1",
           "ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1", "PATH": "/usr/bin:/bin"}
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "passed" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr
