"""CPU checks of the host-only code under AddressSanitizer and
UndefinedBehaviorSanitizer: the oracle (oracle/cpu_join.c), the reference
file readers and writers, the autotuner's table file, the planner and the
request loop's admission policy (deadlines, cancellation, budget), built with
g++ from the sources and run as one program (tests/sanitize/host_checks.cpp).
No GPU and no HIP runtime are involved."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "eating_one_amd", "csrc")
HOST_SOURCES = ["api_io.cpp", "plan.cpp", "tune_table.cpp", "status.cpp", "serve_queue.cpp"]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def _compile(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}"


def test_host_code_under_asan_ubsan(tmp_path):
    gxx, gcc = shutil.which("g++"), shutil.which("gcc")
    if not gxx or not gcc:
        pytest.skip("g++/gcc not available")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
    objs = []
    oracle_obj = str(tmp_path / "cpu_join.o")
    _compile([gcc, "-std=c11", *SAN, "-pthread", "-ffp-contract=off", "-c",
              os.path.join(ROOT, "oracle", "cpu_join.c"), "-o", oracle_obj])
    objs.append(oracle_obj)
    for src in HOST_SOURCES + [os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp")]:
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        path = src if os.path.isabs(src) else os.path.join(CSRC, src)
        _compile([gxx, "-std=c++17", *SAN, "-Wall", *inc, "-c", path, "-o", obj])
        objs.append(obj)
    exe = str(tmp_path / "host_checks")
    # the sanitizer runtimes linked statically: nothing depends on the order of preloaded libraries
    _compile([gxx, *SAN, "-static-libasan", "-static-libubsan", "-pthread", *objs, "-o", exe])
    env = dict(os.environ, EO_CHECK_TMP=str(tmp_path),
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:strict_string_checks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host checks ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
