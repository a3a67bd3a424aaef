"""The library's host code (p2z-tests_amd/host: generator, report, workloads
and their files, descriptor checks) compiled on its own with
AddressSanitizer + UndefinedBehaviorSanitizer and driven through valid calls,
argument errors and damaged files (tests/host_sanitize_main.cpp).  CPU only;
the GPU code is not part of this build."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HOST = os.path.join(REPO, "p2z-tests_amd", "host")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_sanitize")
    srcs = [os.path.join(HOST, f) for f in ("p2z_host.cpp", "p2z_presets.cpp", "p2z_workload.cpp")]
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-fopenmp", "-I" + os.path.join(REPO, "include"),
           "-I" + os.path.join(REPO, "p2z-tests_amd", "csrc"), *srcs, os.path.join(HERE, "host_sanitize_main.cpp"),
           "-o", exe]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    run = subprocess.run([exe, str(tmp_path / "workload.bin")], capture_output=True, text=True, timeout=300, env=env)
    if "runtime does not come first" in run.stderr:
        pytest.skip("the sanitizer runtime cannot be first in this process's library list here")
    assert run.returncode == 0, (run.returncode, run.stdout[-2000:], run.stderr[-4000:])
    assert "clean" in run.stdout
