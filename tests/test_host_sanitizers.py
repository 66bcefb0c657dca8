"""The host C++ under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C gpurify_amd/csrc sanitize` rebuilds the library with its host code
(mesher, .qui parser, tree builder, CPQR, resistance, API and orchestration)
instrumented (libgpurify_san.so) and builds tests/host/tree_check, the
octree builder against brute-force constructions of its lists.  The CPU
tests that drive host code then run in a child process with the clang ASan
runtime preloaded; any report ends the child with a failure."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang"
HOST_TESTS = ["tests/test_fastcap.py", "tests/test_resistance.py", "tests/test_workloads.py",
              "tests/test_m2l_cpqr.py", "tests/test_m2l_low_rank.py", "tests/test_abi.py"]
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}


@pytest.fixture(scope="module")
def sanitized():
    if not os.path.exists(CLANG):
        pytest.skip("no ROCm clang for the sanitizer runtime")
    jobs = str(min(8, os.cpu_count() or 4))
    r = subprocess.run(["make", "-j", jobs, "-C", os.path.join(ROOT, "gpurify_amd", "csrc"),
                        "sanitize"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    rt = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"],
                        capture_output=True, text=True, check=True).stdout.strip()
    assert os.path.exists(rt), rt
    return rt


def test_tree_builder_under_sanitizers(sanitized):
    exe = os.path.join(ROOT, "gpurify_amd", "csrc", "_build_san", "tree_check")
    r = subprocess.run([exe], capture_output=True, text=True, env=dict(os.environ, **SAN_ENV),
                       timeout=600)
    assert r.returncode == 0 and "0 failures" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]


def test_host_code_tests_under_sanitizers(sanitized):
    env = dict(os.environ, LD_PRELOAD=sanitized, GPURIFY_LIB_VARIANT="san", **SAN_ENV)
    # -s: a sanitizer report is written to the child's stderr as it happens
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-s", "-q", "-m", "not gpu",
                        "-p", "no:cacheprovider", *HOST_TESTS], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=1200)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    assert " passed" in out
