"""The library's host code (graph generator, relabelings, partitioner, plan
builder, file writers and loaders, MatrixMarket reader) built with g++ under
AddressSanitizer and UBSan and driven over small graphs of every order,
several partition counts and plan options, plus malformed inputs that must be
rejected (tests/host_sanitize/driver.cpp).  No GPU code is involved."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cola_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    exe = str(tmp_path / "driver")
    srcs = [os.path.join(ROOT, "tests", "host_sanitize", "driver.cpp")] + [
        os.path.join(CSRC, f) for f in ("host.cpp", "mtx.cpp", "community.cpp")]
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fopenmp", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-I" + os.path.join(ROOT, "include"),
           *srcs, "-o", exe]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("the sanitizer runtime is not installed: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-3000:]
    scratch = tmp_path / "files"
    scratch.mkdir()
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe, str(scratch)], capture_output=True, text=True, timeout=600, env=env)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert run.stdout.strip().splitlines()[-1].startswith("ok:"), run.stdout[-2000:]
