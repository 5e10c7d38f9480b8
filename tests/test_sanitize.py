"""The host-only C and C++ code built with AddressSanitizer and
UndefinedBehaviorSanitizer and run on the CPU: the serving loop's region
allocator (dida-thhpca_amd/csrc/letkf_pool.h) and the CPU oracle with the CPU
baseline (oracle/letkf_oracle.c, oracle/letkf_baseline.c).  A sanitizer report
aborts the program (-fno-sanitize-recover=all), so a clean exit is the pass."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)
    return r.returncode, r.stdout + r.stderr


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return env


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_region_allocator_under_sanitizers(tmp_path):
    exe = tmp_path / "pool_check"
    rc, out = _run(["g++", "-std=c++17", *SAN, "-o", str(exe),
                    os.path.join(ROOT, "tests", "native", "pool_check.cpp")])
    assert rc == 0, out
    rc, out = _run([str(exe)], env=_env())
    print(out)
    assert rc == 0 and "pool_check ok" in out, out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no gcc")
def test_oracle_and_baseline_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_check"
    src = [os.path.join(ROOT, "tests", "native", "oracle_check.c"),
           os.path.join(ROOT, "oracle", "letkf_oracle.c"),
           os.path.join(ROOT, "oracle", "letkf_baseline.c")]
    rc, out = _run(["gcc", "-std=c11", "-ffp-contract=off", *SAN, "-o", str(exe), *src,
                    "-lm", "-lpthread", "-ldl"])
    assert rc == 0, out
    rc, out = _run([str(exe)], env=_env())
    print(out)
    assert rc == 0 and "oracle_check ok" in out, out
