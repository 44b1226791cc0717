"""SHAFFT's C API exercised from C: tests/c/c_api_onerank.c follows the call
sequences of the reference's C integration tests (tests/integration/c/
test_c_api.c, test_c_roundtrip.c, test_c_FFT1D.c) at one rank, compiled
with gcc as C99 against include/shafft_amd.h and linked to the product
library."""
import os
import subprocess

import pytest

from shafft_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_program(out_dir):
    lib_dir = os.path.dirname(api.library_path())
    exe = os.path.join(str(out_dir), "c_api_onerank")
    cmd = ["gcc", "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c", "c_api_onerank.c"),
           "-L", lib_dir, "-lshafft_amd", "-Wl,-rpath," + lib_dir, "-lm", "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def run_program(exe, *args, timeout):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr)
    return r


def test_c_program_host_checks(tmp_path):
    """Builds as strict C99 and passes every check that needs no device."""
    exe = build_program(tmp_path)
    r = run_program(exe, "host", timeout=60)
    assert r.returncode == 0, r.stdout
    assert "FAIL" not in r.stdout and r.stdout.count("PASS") >= 15


@pytest.mark.gpu
def test_c_program_full(tmp_path):
    """Round trips within 1e-5 (complex64) / 1e-12 (complex128), the delta
    and plane-wave closed forms, buffer bookkeeping and layout queries."""
    exe = build_program(tmp_path)
    r = run_program(exe, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "FAIL" not in r.stdout and r.stdout.count("PASS") >= 30
