"""Host code under AddressSanitizer / UndefinedBehaviorSanitizer: the readers
on seeded mutations of the golden inputs, the writers, the QUBO builder and
the generators at their edges (tests/host_fuzz.cpp).  CPU only: the host
source is built with g++, no GPU code is involved."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_is_clean_under_sanitizers(tmp_path):
    exe = tmp_path / "host_fuzz"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined",
                    "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "sparse-ising-solver_amd", "csrc"),
                    os.path.join(ROOT, "sparse-ising-solver_amd", "csrc", "sis_host.cpp"),
                    os.path.join(ROOT, "tests", "host_fuzz.cpp"), "-o", str(exe)], check=True)
    inputs = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "inputs", "*")))
    env = dict(os.environ, SIS_FUZZ_TRIALS="150", SIS_FUZZ_DIR=str(tmp_path),
               ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)] + inputs, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert "parsed ok" in r.stdout
