"""GPU parity of the FFT1D path (one rank) against the FP64 oracle.

Sizes follow tests/integration/cpp/test_FFT1D_roundtrip.cpp (delta at 256,
round trips at 128, sizes {32,33,64,100,128}, non-divisible {10,17,31,100,
127}); the headline lengths 2^27 (four-step, 3 LDS passes) and the prime
1,000,003 (Bluestein through a 2^21 convolution, 3 passes), and the prime
4,194,319 (a 2^24 convolution, past the three-pass form: two M-point
four-step chains) are checked by direct-summation spot bins, Parseval and
round trips; 9 * 2^28 (element indices past 2^31) by an impulse.
"""
import numpy as np
import pytest

import oracle
from shafft_amd import api

pytestmark = pytest.mark.gpu
TOL = {api.C2C: 1e-5, api.Z2Z: 1e-12}
DT = {api.C2C: np.complex64, api.Z2Z: np.complex128}


def run_api(n, prec, x):
    p = api.FFT1D()
    p.init(n, n, 0, prec)
    p.plan()
    assert p.alloc_size() == n
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x)
    p.set_buffers(a.ptr, b.ptr)
    p.execute(api.FORWARD)
    d, w = p.get_buffers()
    assert (d, w) == (b.ptr, a.ptr)
    spec = api.download_ptr(d, n, prec)
    assert np.array_equal(api.download_ptr(w, n, prec), x.astype(DT[prec]))  # input kept
    p.normalize()
    p.execute(api.BACKWARD)
    p.normalize()
    d, _ = p.get_buffers()
    back = api.download_ptr(d, n, prec)
    return spec, back


@pytest.mark.parametrize("prec", [api.C2C, api.Z2Z])
@pytest.mark.parametrize("n", [10, 17, 31, 32, 33, 64, 100, 127, 128, 256, 1000, 4096, 65536, 40009])
def test_fft1d_vs_oracle(gpu_lib, prec, n):
    x = oracle.uniform_input((n,), 44, DT[prec])
    spec, back = run_api(n, prec, x)
    ref = oracle.dftn(x.astype(np.complex128), None, -1) if n <= 4096 else np.fft.fft(x.astype(np.complex128))
    e1, e2 = oracle.rel_l2(spec, ref), oracle.rel_l2(back, x)
    print(f"FFT1D n={n} prec={prec}: forward relL2 {e1:.3e} roundtrip relL2 {e2:.3e}")
    assert oracle.check_rel_error(spec, ref, TOL[prec])
    assert oracle.check_rel_error(back, x, TOL[prec])


def test_fft1d_delta(gpu_lib):
    for prec in (api.C2C, api.Z2Z):
        n = 256
        x = np.zeros(n, DT[prec])
        x[0] = 1
        spec, _ = run_api(n, prec, x)
        assert np.max(np.abs(spec - 1)) < (1e-6 if prec == api.C2C else 1e-14)


def _bins_direct(x, ks, chunk=1 << 24):
    n = x.size
    out = []
    for k in ks:
        acc = 0j
        for s in range(0, n, chunk):
            idx = np.arange(s, min(n, s + chunk), dtype=np.int64)
            ph = np.mod(idx * k, n).astype(np.float64) / n
            acc += np.sum(x[s:s + chunk].astype(np.complex128) * np.exp(-2j * np.pi * ph))
        out.append(acc)
    return np.array(out)


def _energy(a, chunk=1 << 24):
    return sum(float(np.vdot(a[s:s + chunk].astype(np.complex128), a[s:s + chunk].astype(np.complex128)).real)
               for s in range(0, a.size, chunk))


@pytest.mark.parametrize("n,prec", [(1 << 27, api.C2C), (1 << 27, api.Z2Z), (1000003, api.C2C), (1000003, api.Z2Z),
                                    (4194319, api.C2C), (4194319, api.Z2Z)])
def test_fft1d_headline_lengths(gpu_lib, n, prec):
    x = oracle.gaussian_input((n,), 44, DT[prec])  # C3 input recipe (Box-Muller)
    seam = api.OneDBackendPlan(n, prec)
    print(seam.describe())
    spec, back = run_api(n, prec, x)
    ex = _energy(x)
    e_par = abs(_energy(spec) / n - ex) / ex
    ks = [0, 1, 12345, n // 2 + 7, n - 1]
    ref = _bins_direct(x, ks)
    e_bins = np.abs(spec[ks] - ref) / np.sqrt(ex)
    e_rt = oracle.rel_l2(back, x)
    print(f"FFT1D n={n} prec={prec}: bins {e_bins.max():.3e} Parseval {e_par:.3e} roundtrip {e_rt:.3e}")
    assert e_bins.max() < TOL[prec] * 10
    assert e_par < TOL[prec]
    assert e_rt < TOL[prec] * np.sqrt(np.log(n))


def test_fft1d_past_int32_indices(gpu_lib):
    """N = 9 * 2^28 = 2,415,919,104 complex64 elements (18 GB per buffer):
    element indices past 2^31 - 1, beyond the reference's INT_MAX guard on
    its 1-D lengths (src/cpp/fft/hipfft_method/1d/fft1d_method.hip:416-581).
    A unit impulse at p = N - 5 must transform to the plane wave
    W_N^{k p} (spot bins with the phase k*p mod N in exact integers, and
    Parseval), and forward -> normalize -> backward -> normalize must put it
    back at p: an index that wrapped anywhere in the passes would move or
    lose it."""
    n, prec = 9 << 28, api.C2C
    p = n - 5
    x = np.zeros(n, np.complex64)
    x[p] = 1
    print("N", n, api.OneDBackendPlan(n, prec).describe(), flush=True)
    fp = api.FFT1D()
    fp.init(n, n, 0, prec)
    fp.plan()
    a, b = api.DeviceBuffer(n, prec), api.DeviceBuffer(n, prec)
    a.upload(x)
    fp.set_buffers(a.ptr, b.ptr)
    fp.execute(api.FORWARD)
    d, _ = fp.get_buffers()
    spec = api.download_ptr(d, n, prec)
    ks = np.array([0, 1, 12345, (1 << 31) + 17, n // 2 + 7, n - 1], dtype=np.int64)
    ph = np.array([(int(k) * p) % n for k in ks], dtype=np.float64) / n
    ref = np.exp(-2j * np.pi * ph)
    e_bins = np.abs(spec[ks] - ref).max()
    chunk = 1 << 26
    e_par = abs(_energy(spec) / n - 1)  # FP64 accumulation
    del spec
    fp.normalize()
    fp.execute(api.BACKWARD)
    fp.normalize()
    d, _ = fp.get_buffers()
    back = api.download_ptr(d, n, prec)
    peak = complex(back[p])
    back[p] = 0
    rest = max(float(np.abs(back[s:s + chunk]).max()) for s in range(0, n, chunk))
    print(f"N={n}: bins {e_bins:.3e} Parseval {e_par:.3e} back[p] {peak:.7f} max elsewhere {rest:.3e}",
          flush=True)
    assert e_bins < 1e-4
    assert e_par < 1e-5
    assert abs(peak - 1) < 1e-5
    assert rest < 1e-5
