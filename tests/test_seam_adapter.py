"""The gfx950 seam adapter (shafft_amd/csrc/seam/gfx950_method) compiled as
SHAFFT's core would compile a backend directory -- its own fftnd_method.hpp,
fftnd_handle.hpp, fft1d_method.hpp, fft1d_handle.hpp and common/normalize.hpp
on the include path, MPI from the image, no hipFFT -- and driven by
tests/seam/seam_program.cpp through the core's call sequences
(src/cpp/shafft_internal.cpp): by-value N-D handle, void*& buffer swap,
streams (tests/integration/cpp/test_backend.cpp:251-450), normalize, FFT1D.

On the CPU the adapter links the emulation build of the library (the same
host code with the passes replayed on the host), also under a two-rank
mpiexec; on the GPU it links libshafft_amd.so."""
import os
import subprocess

import pytest

from shafft_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEAM = os.path.join(ROOT, "shafft_amd", "csrc", "seam", "gfx950_method")
MPI_ROOT = "/opt/conda"
SYS_LIB = "/usr/lib/x86_64-linux-gnu"

need_mpi = pytest.mark.skipif(not os.path.exists(os.path.join(MPI_ROOT, "include", "mpi.h")),
                              reason="no MPI headers in this image")


def build(out_dir, lib_path, extra=()):
    exe = os.path.join(str(out_dir), "seam_program")
    lib_dir = os.path.dirname(lib_path)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "tests", "seam", "include"), "-I", SEAM,
           "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
           "-I", os.path.join(MPI_ROOT, "include"),
           os.path.join(ROOT, "tests", "seam", "seam_program.cpp"),
           os.path.join(SEAM, "fftnd_method.cpp"), os.path.join(SEAM, "fft1d_method.cpp"),
           os.path.join(SEAM, "common", "normalize.cpp"),
           lib_path, os.path.join(MPI_ROOT, "lib", "libmpi.so"), *extra,
           # the image's own libstdc++ ahead of the older one beside MPI
           f"-Wl,-rpath,{lib_dir}:{SYS_LIB}:{os.path.join(MPI_ROOT, 'lib')}", "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def run(cmd, timeout):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr)
    return r


@need_mpi
def test_seam_adapter_on_emulator(emu_lib, tmp_path):
    exe = build(tmp_path, os.path.join(ROOT, "_build", "emu", "libshafft_emu.so"))
    r = run([exe], timeout=120)
    assert r.returncode == 0 and "FAIL" not in r.stdout and r.stdout.count("PASS") >= 20


@need_mpi
def test_seam_adapter_two_ranks_on_emulator(emu_lib, tmp_path):
    exe = build(tmp_path, os.path.join(ROOT, "_build", "emu", "libshafft_emu.so"))
    r = run([os.path.join(MPI_ROOT, "bin", "mpiexec"), "-n", "2", exe, "multirank"], timeout=120)
    assert r.returncode == 0 and "FAIL" not in r.stdout
    assert r.stdout.count("PASS FFT1D on a communicator of more than one rank") == 2


@need_mpi
@pytest.mark.gpu
def test_seam_adapter_on_gpu(gpu_lib, tmp_path):
    exe = build(tmp_path, api.library_path(), extra=["-L/opt/rocm/lib", "-lamdhip64"])
    r = run([exe], timeout=120)
    assert r.returncode == 0 and "FAIL" not in r.stdout and r.stdout.count("PASS") >= 20
