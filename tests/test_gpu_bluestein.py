"""GPU parity of the three-pass Bluestein convolution over every padded
length it covers (M = 2^13 .. 2^22), on the compile-time-length kernels
where they apply (32..512-point chirped column passes, 1024..8192-point
filter row passes) and on the generic pass elsewhere, for FFT1D and for a
prime contiguous axis of an N-D array (outer lines).  Every element against
numpy's FP64 FFT (pinned against the direct DFT at primes in
tests/test_oracle.py), both directions, at TOL_F = 1e-5 / TOL_D = 1e-12 as
tests/include/test_utils.hpp:84-122; the generic-kernel plan of the same
length is run beside it and must agree to the same tolerance.
"""
import os
import subprocess
import tempfile
import sys

import numpy as np
import pytest

import oracle
from shafft_amd import api

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = {api.C2C: 1e-5, api.Z2Z: 1e-12}
DT = {api.C2C: np.complex64, api.Z2Z: np.complex128}

# prime n -> padded length M = next_pow2(2n - 1)
PRIMES = [3001, 5003, 12007, 20011, 40009, 70001, 150001, 300007, 600011, 1500007]


def _run_1d(n, prec, x):
    p = api.OneDBackendPlan(n, prec)
    desc = p.describe()
    ext = max(max(p.pass_extents(i)) for i in range(p.num_passes()))
    a, b = api.DeviceBuffer(max(ext, n), prec), api.DeviceBuffer(max(ext, n), prec)
    a.upload(x)
    out = {}
    for direction in (api.FORWARD, api.BACKWARD):
        p.execute(a.ptr, b.ptr, direction)
        p.synchronize()
        out[direction] = api.download_ptr(b.ptr, n, prec)
    assert np.array_equal(api.download_ptr(a.ptr, n, prec), x)  # input preserved
    return desc, out


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
@pytest.mark.parametrize("n", PRIMES)
def test_bluestein_lengths_vs_numpy(gpu_lib, n, prec):
    x = oracle.gaussian_input((n,), 44, DT[prec])
    desc, out = _run_1d(n, prec, x)
    print(f"n={n} prec={prec}\n{desc}", flush=True)
    x128 = x.astype(np.complex128)
    for direction, y in out.items():
        ref = np.fft.fft(x128) if direction == api.FORWARD else np.fft.ifft(x128) * n
        e = oracle.rel_l2(y, ref)
        print(f"  dir {direction}: relL2 {e:.3e}", flush=True)
        assert oracle.check_rel_error(y, ref, TOL[prec]), (n, prec, direction, e)


def test_bluestein_kernels_selected(gpu_lib):
    """The 1,000,003-point plans run all three passes on the fixed-length
    Bluestein kernels; the 4096 columns go 16 to a tile in both precisions
    (a tile for each of the 256 CUs: 128-byte segments for complex float,
    256-byte for complex double)."""
    for prec, cols in ((api.C2C, 16), (api.Z2Z, 16)):
        desc = api.OneDBackendPlan(1000003, prec).describe()
        print(desc, flush=True)
        lines = desc.strip().splitlines()
        assert len(lines) == 3
        for line, kind in zip(lines, ("blue1", "blue2", "blue3")):
            assert kind in line, desc
        assert f"C={cols} " in lines[0] and f"C={cols} " in lines[2], desc


@pytest.mark.parametrize("dims,prec", [((3, 40009), api.C2C), ((2, 3, 20011), api.Z2Z), ((5, 70001), api.Z2Z)])
def test_bluestein_prime_axis_with_outer_lines(gpu_lib, dims, prec):
    x = oracle.uniform_input(dims, 45, DT[prec])
    p = api.NDBackendPlan(list(dims), [len(dims) - 1], prec)
    print(p.describe(), flush=True)
    a, b = api.DeviceBuffer(x.size, prec), api.DeviceBuffer(x.size, prec)
    x128 = x.astype(np.complex128)
    for direction in (api.FORWARD, api.BACKWARD):
        a.upload(x.ravel())
        d, _ = p.execute(a.ptr, b.ptr, direction)
        y = api.download_ptr(d, x.size, prec).reshape(dims)
        ref = np.fft.fft(x128, axis=-1) if direction == api.FORWARD else np.fft.ifft(x128, axis=-1) * dims[-1]
        e = oracle.rel_l2(y, ref)
        print(f"  dims {dims} dir {direction}: relL2 {e:.3e}", flush=True)
        assert oracle.check_rel_error(y, ref, TOL[prec]), (dims, direction, e)


def test_bluestein_fixed_matches_generic(gpu_lib):
    """The same 1,000,003-point transforms with the fixed-length Bluestein
    kernels switched off (SHAFFT_AMD_BLUE_FX=0, the generic pass) agree with
    the default plan within the tolerance (a second process: the knob is
    read at plan time)."""
    code = r"""
import sys, numpy as np
sys.path.insert(0, '.')
import oracle
from shafft_amd import api
api.lib()
for prec in (api.C2C, api.Z2Z):
    n = 1000003
    x = oracle.gaussian_input((n,), 44, np.complex64 if prec == api.C2C else np.complex128)
    p = api.OneDBackendPlan(n, prec)
    assert not any(f' blue{k} ' in p.describe() for k in (1, 2, 3))
    ext = max(max(p.pass_extents(i)) for i in range(p.num_passes()))
    a, b = api.DeviceBuffer(ext, prec), api.DeviceBuffer(ext, prec)
    a.upload(x)
    p.execute(a.ptr, b.ptr, api.FORWARD)
    p.synchronize()
    np.save(sys.argv[1] + f'_{prec}.npy', api.download_ptr(b.ptr, n, prec))
"""
    tmpdir = tempfile.mkdtemp(prefix="blue_generic_")
    stem = os.path.join(tmpdir, "out")
    env = dict(os.environ, SHAFFT_AMD_BLUE_FX="0")
    r = subprocess.run([sys.executable, "-c", code, stem], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    for prec in (api.C2C, api.Z2Z):
        n = 1000003
        x = oracle.gaussian_input((n,), 44, DT[prec])
        desc, out = _run_1d(n, prec, x)
        g = np.load(stem + f"_{prec}.npy")
        os.remove(stem + f"_{prec}.npy")
        e = oracle.rel_l2(out[api.FORWARD], g.astype(np.complex128))
        print(f"prec {prec}: fixed vs generic relL2 {e:.3e}", flush=True)
        assert e <= TOL[prec]
    os.rmdir(tmpdir)


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
@pytest.mark.parametrize("n", [5003, 12007, 40009, 1000003])
def test_bluestein_row_forms(gpu_lib, monkeypatch, n, prec):
    """Both forms of the filter row pass, forced (16 elements per thread,
    radix-16 stages; 8 elements per thread, radix-8/4 stages, twice the
    threads): each against numpy's FFT, and the two within the tolerance of
    each other."""
    x = oracle.gaussian_input((n,), 45, DT[prec])
    x128 = x.astype(np.complex128)
    outs = {}
    for e in ("16", "8"):
        monkeypatch.setenv("SHAFFT_AMD_BLUE_ROW_E", e)
        desc, out = _run_1d(n, prec, x)
        assert "blue2" in desc, desc
        outs[e] = out
        for direction, y in out.items():
            ref = np.fft.fft(x128) if direction == api.FORWARD else np.fft.ifft(x128) * n
            assert oracle.check_rel_error(y, ref, TOL[prec]), (n, prec, e, direction)
    for direction in outs["16"]:
        assert oracle.check_rel_error(outs["8"][direction], outs["16"][direction], TOL[prec])
