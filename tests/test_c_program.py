"""SHAFFT's C and C++ APIs exercised from C and C++: tests/c/c_api_onerank.c
follows the call sequences of the reference's C integration tests
(tests/integration/c/test_c_api.c, test_c_roundtrip.c, test_c_FFT1D.c) at
one rank, compiled with gcc as C99 against include/shafft_amd.h;
tests/c/cpp_api_onerank.cpp does the same for the class API of
include/shafft_amd.hpp (tests/integration/cpp/test_roundtrip.cpp,
test_FFT1D_roundtrip.cpp), compiled with g++ as C++17.  Both link to the
product library."""
import os
import subprocess

import pytest

from shafft_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_program(out_dir, cpp=False):
    lib_dir = os.path.dirname(api.library_path())
    name = "cpp_api_onerank" if cpp else "c_api_onerank"
    exe = os.path.join(str(out_dir), name)
    lang = ["g++", "-std=c++17"] if cpp else ["gcc", "-std=c99"]
    cmd = lang + ["-O1", "-Wall", "-Wextra", "-Werror",
                  "-I", os.path.join(ROOT, "include"),
                  os.path.join(ROOT, "tests", "c", name + (".cpp" if cpp else ".c")),
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


def test_cpp_program_host_checks(tmp_path):
    """The class API builds as strict C++17 and passes every check that
    needs no device (init, queries, moves, configuration)."""
    exe = build_program(tmp_path, cpp=True)
    r = run_program(exe, "host", timeout=60)
    assert r.returncode == 0, r.stdout
    assert "FAIL" not in r.stdout and r.stdout.count("PASS") >= 12


@pytest.mark.gpu
def test_cpp_program_full(tmp_path):
    """Class-API round trips within 1e-5 / 1e-12 (one through a moved plan)
    and FFT1D deltas."""
    exe = build_program(tmp_path, cpp=True)
    r = run_program(exe, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "FAIL" not in r.stdout and r.stdout.count("PASS") >= 16
