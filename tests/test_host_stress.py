"""The host logic of the search driver and the request loop -- plans, batches
in flight, hit slots, the pinned-buffer pool and its process-wide cache, the
insertion thread, plan eviction, deadlines, CU splits, hit-record chunk
sizes and graph settings changed between one plan's searches -- built for the CPU against a
stand-in of the HIP runtime whose streams run asynchronously on threads of
their own (tests/native/hipstub), under AddressSanitizer and under
ThreadSanitizer, with three requests in flight.  Every served request must
return its plan's alone result (tests/native/server_stress.cpp)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "prestozl_amd", "csrc")
SOURCES = [os.path.join(HERE, "native", "hipstub", "hipstub.cpp"), os.path.join(HERE, "native", "server_stress.cpp")] + \
          [os.path.join(CSRC, f) for f in ("pzl_engine.cpp", "pzl_server.cpp", "pzl_geometry.cpp", "pzl_stats.cpp")]


def _build_and_run(tmp_path, compiler, flags, env, rounds):
    exe = str(tmp_path / "server_stress")
    cmd = [compiler, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread",
           "-I" + os.path.join(HERE, "native", "hipstub"), "-I" + os.path.join(ROOT, "include"), "-I" + CSRC] + \
        flags + SOURCES + ["-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        pytest.fail("build failed:\n" + r.stderr[-3000:])
    r = subprocess.run([exe, str(rounds)], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, **env))
    print(r.stdout[-3000:])
    print(r.stderr[-6000:])
    assert r.returncode == 0
    assert r.stdout.strip().endswith("OK")


def test_request_loop_under_asan(tmp_path):
    # the process-wide pinned-buffer cache holds its buffers until exit by design
    _build_and_run(tmp_path, "g++", ["-fsanitize=address"], {"ASAN_OPTIONS": "detect_leaks=0"}, 4)


def test_request_loop_under_tsan(tmp_path):
    # clang's runtime intercepts pthread_cond_clockwait (condition-variable
    # waits with a timeout); gcc 11's does not and reports false races
    clang = "/opt/rocm/llvm/bin/clang++"
    if not os.path.exists(clang):
        pytest.skip("clang++ with ThreadSanitizer not present")
    _build_and_run(tmp_path, clang, ["-fsanitize=thread"], {"TSAN_OPTIONS": "halt_on_error=1"}, 2)
