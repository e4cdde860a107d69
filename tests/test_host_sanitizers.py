"""The host plan (stir_amd/csrc/pmrt_plan.cpp: geometry, symmetries, subsets,
families -- the index arithmetic a device fault would trace back to) and the
C oracle, built with g++/gcc under AddressSanitizer and UndefinedBehavior-
Sanitizer (halting on the first finding), run on the workload geometries:
the 8-ring scanner, mMR-size with and without TOF, the reference's RPT
fixture geometry and six of the uneven batch problems. CPU only."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def _runtime(name):
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.mark.slow
def test_plan_and_oracle_under_sanitizers(tmp_path):
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("the AddressSanitizer runtime is not installed")
    csrc = os.path.join(ROOT, "stir_amd", "csrc")
    plan_so = tmp_path / "libplan_san.so"
    subprocess.run(["g++", "-std=c++17", "-fPIC", "-shared", *SAN, "-ffp-contract=off", "-I", csrc, "-I",
                    os.path.join(ROOT, "include"), "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    os.path.join(csrc, "pmrt_plan.cpp"), os.path.join(ROOT, "tests", "sanitize", "plan_probe.cpp"),
                    "-o", str(plan_so)], check=True)
    oracle_so = tmp_path / "liborpmrt_san.so"
    subprocess.run(["gcc", "-std=gnu11", "-fPIC", "-shared", *SAN, "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    os.path.join(ROOT, "oracle", "pmrt_oracle.c"), "-lm", "-o", str(oracle_so)], check=True)
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan + (":" + _runtime("libubsan.so") if _runtime("libubsan.so") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["OMP_NUM_THREADS"] = "4"
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "sanitize", "run_sanitized.py"), str(plan_so),
                        str(oracle_so), ROOT], env=env, capture_output=True, text=True, timeout=900)
    print(r.stdout[-4000:])
    print(r.stderr[-4000:])
    assert r.returncode == 0
    assert "oracle:" in r.stdout
